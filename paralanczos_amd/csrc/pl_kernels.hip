// HIP kernels for gfx950 (MI355X): the batched principal-submatrix Lanczos
// eigenvalue path and the synthetic-input generators.
//
// One lane group of GROUP lanes (8, 16, 32 or 64; a 64-lane wavefront holds
// 64/GROUP groups) solves one subset of size k <= GROUP. Lane i owns row i of
// the gathered submatrix G_S, entry i of every Lanczos vector, row i of the
// Lanczos basis V (vrow) and the whole Lanczos vector v_i (vcol). Rows and
// columns k..GROUP-1 are zero padding; they stay exactly zero through the
// recurrence, and every sum runs in the same order for any GROUP >= k, so a
// subset gives bit-identical Lanczos arithmetic in any lane-group size.
//
// Stages per subset (reference: POS_SYM_SUBMATRIX_KER2, ParaLanczos.cu:1157-1480):
//  (a) gather (ParaLanczos.cu:1237-1242 / :471-490): the upper triangle of G_S
//      in global index order, G[min(i,j)][max(i,j)], is loaded one subset row
//      per instruction (lanes = columns), mirrored into a padded LDS tile and
//      each lane then reads its row into VGPRs.
//  (b) k Lanczos steps (ParaLanczos.cu:1285-1374) in FP32: matvec from the
//      VGPR row; alpha = v^T G v by a compensated (TwoProduct/TwoSum) group
//      reduction; three-term update; one full classical Gram-Schmidt pass
//      against all previous vectors, repeated when it removed more than half
//      of w's norm (Daniel, Gragg, Kaufman, Stewart, Math. Comp. 1976);
//      beta = ||w|| compensated; deterministic restart on breakdown. The all-to-all broadcasts of v, w and the
//      Gram-Schmidt coefficients use DPP row broadcasts fused into v_fmac for
//      groups of 16/32/64 lanes (v_permlane16/32_swap bring the other rows'
//      values into each row), and LDS for groups of 8.
//  (c) FP64 Sturm-sequence multisection on T for lambda_min and lambda_max
//      (replaces the serial TQLRAT of ParaLanczos.cu:32-189, :1454-1462), or
//      per-lane bisection for every eigenvalue (legacy all-eigenvalue mode).
//      The count uses the characteristic-polynomial recurrence (Barth, Martin,
//      Wilkinson, Numer. Math. 9, 1967; the reference's unused BM_POLY,
//      ParaLanczos.cu:504-552, without its off-by-one defects), which needs no
//      division, rescaled by a power of two every 8 steps.
// No kernel uses a block-level barrier: all cross-lane traffic stays inside
// one wavefront, so groups finish independently.
#include <float.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <math.h>
#include <stddef.h>

#include "pl_common.hpp"
#include "pl_internal.hpp"

namespace pl {

constexpr int BLOCK = 256;
constexpr int BIS_MAXIT = 64;        // iteration cap (the floor below bounds it by ~54)
constexpr int BIS_TOLBITS_EXT = 27;  // extremes (FP32 outputs): width <= 2^-27 max(|lo|,|hi|)
constexpr int BIS_TOLBITS_ALL = 44;  // all eigenvalues (FP64 legacy outputs)

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---- generic group reductions through ds_bpermute (setup paths, GROUP 8) ----
template <int W>
__device__ __forceinline__ uint32_t g_or(uint32_t x) {
#pragma unroll
  for (int m = 1; m < W; m <<= 1) x |= (uint32_t)__shfl_xor((int)x, m, W);
  return x;
}
template <int W>
__device__ __forceinline__ double g_min(double x) {
#pragma unroll
  for (int m = 1; m < W; m <<= 1) x = fmin(x, __shfl_xor(x, m, W));
  return x;
}
template <int W>
__device__ __forceinline__ double g_max(double x) {
#pragma unroll
  for (int m = 1; m < W; m <<= 1) x = fmax(x, __shfl_xor(x, m, W));
  return x;
}
template <int W>
__device__ __forceinline__ double g_prod(double x) {
#pragma unroll
  for (int m = 1; m < W; m <<= 1) x *= __shfl_xor(x, m, W);
  return x;
}
template <int W>
__device__ __forceinline__ float g_sum(float x) {
#pragma unroll
  for (int m = 1; m < W; m <<= 1) x += __shfl_xor(x, m, W);
  return x;
}

// true when pred holds on every lane of the caller's group
template <int W>
__device__ __forceinline__ bool g_all(bool pred) {
  const uint64_t bad = __ballot(!pred);
  if constexpr (W == 64) {
    return bad == 0ull;
  } else {
    const int lane = __lane_id();
    const int base = lane - (lane % W);
    return ((bad >> base) & ((1ull << W) - 1ull)) == 0ull;
  }
}

__device__ __forceinline__ void two_sum(float a, float b, float &s, float &e) {
  s = a + b;
  const float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// ---- DPP / permlane building blocks (groups of 16, 32, 64 lanes) ----
constexpr int DPP_QUAD_1032 = 0xB1;     // quad_perm [1,0,3,2]  (partner lane ^ 1)
constexpr int DPP_QUAD_2301 = 0x4E;     // quad_perm [2,3,0,1]  (partner lane ^ 2)
constexpr int DPP_ROW_MIRROR = 0x140;   // lane i <- 15 - i within a row of 16
constexpr int DPP_ROW_HMIRROR = 0x141;  // lane i <- 7 - i within each half row

template <int CTRL>
__device__ __forceinline__ float dpp(float x) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), CTRL, 0xf, 0xf, true));
}

__device__ __forceinline__ float swap_lo16(float x, float &hi) {
  // v_permlane16_swap with both operands x: result 0 holds row 2h's values and
  // result 1 row 2h+1's values in every lane of the 32-lane half h
  const auto r =
      __builtin_amdgcn_permlane16_swap(__builtin_bit_cast(unsigned, x), __builtin_bit_cast(unsigned, x), false, false);
  hi = __builtin_bit_cast(float, (unsigned)r[1]);
  return __builtin_bit_cast(float, (unsigned)r[0]);
}

__device__ __forceinline__ float swap_lo32(float x, float &hi) {
  // v_permlane32_swap with both operands x: result 0 holds lanes 0..31's
  // values, result 1 lanes 32..63's, in both halves
  const auto r =
      __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(unsigned, x), __builtin_bit_cast(unsigned, x), false, false);
  hi = __builtin_bit_cast(float, (unsigned)r[1]);
  return __builtin_bit_cast(float, (unsigned)r[0]);
}

// x of every lane of the group, split in rows of 16: part p holds, in every
// lane, the values of row p of the group (lane 16p + n at row position n)
template <int GROUP>
__device__ __forceinline__ void row_parts(float x, float (&p)[GROUP / 16]) {
  if constexpr (GROUP == 16) {
    p[0] = x;
  } else if constexpr (GROUP == 32) {
    p[0] = swap_lo16(x, p[1]);
  } else {
    float hi32;
    const float lo32 = swap_lo32(x, hi32);
    p[0] = swap_lo16(lo32, p[1]);
    p[2] = swap_lo16(hi32, p[3]);
  }
}

// Fused row broadcast + FMA: a0 += sum_{n even} bcast_n(v) * g[n] and
// a1 += sum_{n odd} bcast_n(v) * g[n] over one row of 16 lanes (v_fmac_f32
// with DPP row_newbcast:n on its first source). The s_nop covers the DPP
// read-after-VALU-write and VALU-exec-write hazards (at most 5 wait states).
#define PL_BC(N, ACC, G)                                                               \
  "v_fmac_f32_dpp %" #ACC ", %2, %" #G " row_newbcast:" #N " row_mask:0xf bank_mask:0xf\n\t"
__device__ __forceinline__ void row_dot16(float &a0, float &a1, float v, const float *g) {
  asm volatile("s_nop 4\n\t" PL_BC(0, 0, 3) PL_BC(1, 1, 4) PL_BC(2, 0, 5) PL_BC(3, 1, 6)
                   PL_BC(4, 0, 7) PL_BC(5, 1, 8) PL_BC(6, 0, 9) PL_BC(7, 1, 10) PL_BC(8, 0, 11)
                       PL_BC(9, 1, 12) PL_BC(10, 0, 13) PL_BC(11, 1, 14) PL_BC(12, 0, 15)
                           PL_BC(13, 1, 16) PL_BC(14, 0, 17) PL_BC(15, 1, 18)
               : "+v"(a0), "+v"(a1)
               : "v"(v), "v"(g[0]), "v"(g[1]), "v"(g[2]), "v"(g[3]), "v"(g[4]), "v"(g[5]),
                 "v"(g[6]), "v"(g[7]), "v"(g[8]), "v"(g[9]), "v"(g[10]), "v"(g[11]), "v"(g[12]),
                 "v"(g[13]), "v"(g[14]), "v"(g[15]));
}
#undef PL_BC

// sum_c x[c] * (value of lane c of the group) for a group of 16/32/64 lanes
template <int GROUP>
__device__ __forceinline__ float group_dot_dpp(float v, const float (&x)[GROUP]) {
  float parts[GROUP / 16];
  row_parts<GROUP>(v, parts);
  float a0 = 0.f, a1 = 0.f;
#pragma unroll
  for (int p = 0; p < GROUP / 16; ++p) row_dot16(a0, a1, parts[p], x + 16 * p);
  return a0 + a1;
}

// Group butterfly over DPP/permlane partners; every lane ends with identical bits.
template <int GROUP>
__device__ __forceinline__ float g_sum_dpp(float x) {
  x += dpp<DPP_QUAD_1032>(x);
  x += dpp<DPP_QUAD_2301>(x);
  x += dpp<DPP_ROW_HMIRROR>(x);
  x += dpp<DPP_ROW_MIRROR>(x);
  if constexpr (GROUP >= 32) {
    float b;
    const float a = swap_lo16(x, b);
    x = a + b;
  }
  if constexpr (GROUP == 64) {
    float b;
    const float a = swap_lo32(x, b);
    x = a + b;
  }
  return x;
}

// compensated dot product over the group (Ogita, Rump, Oishi, "Accurate sum
// and dot product", SIAM J. Sci. Comput. 2005, Dot2 with a pairwise tree)
__device__ __forceinline__ void comp_step(float &p, float &e, float po, float eo) {
  float s, t;
  two_sum(p, po, s, t);
  p = s;
  e = (e + eo) + t;
}

template <int GROUP>
__device__ __forceinline__ void g_dot2_pe(float x, float y, float &p, float &e) {
  p = x * y;
  e = __builtin_fmaf(x, y, -p);
  if constexpr (GROUP < 16) {
#pragma unroll
    for (int m = 1; m < GROUP; m <<= 1) comp_step(p, e, __shfl_xor(p, m, GROUP), __shfl_xor(e, m, GROUP));
  } else {
    comp_step(p, e, dpp<DPP_QUAD_1032>(p), dpp<DPP_QUAD_1032>(e));
    comp_step(p, e, dpp<DPP_QUAD_2301>(p), dpp<DPP_QUAD_2301>(e));
    comp_step(p, e, dpp<DPP_ROW_HMIRROR>(p), dpp<DPP_ROW_HMIRROR>(e));
    comp_step(p, e, dpp<DPP_ROW_MIRROR>(p), dpp<DPP_ROW_MIRROR>(e));
    if constexpr (GROUP >= 32) {
      float pb, eb;
      const float pa = swap_lo16(p, pb);
      const float ea = swap_lo16(e, eb);
      p = pa;
      e = ea;
      comp_step(p, e, pb, eb);
    }
    if constexpr (GROUP == 64) {
      float pb, eb;
      const float pa = swap_lo32(p, pb);
      const float ea = swap_lo32(e, eb);
      p = pa;
      e = ea;
      comp_step(p, e, pb, eb);
    }
  }
}

template <int GROUP>
__device__ __forceinline__ float g_dot2(float x, float y) {
  float p, e;
  g_dot2_pe<GROUP>(x, y, p, e);
  return p + e;
}

// ---- LDS broadcast path (groups of 8) ----
// sum_c x[c] * L[c] with L a 16-byte aligned LDS vector read by broadcast;
// two accumulators split by index parity (the order of row_dot16)
template <int GROUP>
__device__ __forceinline__ float bcast_dot(const float *L, const float (&x)[GROUP]) {
  float acc0 = 0.f, acc1 = 0.f;
#pragma unroll
  for (int q = 0; q < GROUP / 4; ++q) {
    const float4 t = reinterpret_cast<const float4 *>(L)[q];
    acc0 = __builtin_fmaf(x[4 * q + 0], t.x, acc0);
    acc1 = __builtin_fmaf(x[4 * q + 1], t.y, acc1);
    acc0 = __builtin_fmaf(x[4 * q + 2], t.z, acc0);
    acc1 = __builtin_fmaf(x[4 * q + 3], t.w, acc1);
  }
  return acc0 + acc1;
}

// Stage (b) for groups of 8 (LDS broadcasts): one classical Gram-Schmidt pass
// of w against the basis, h_l = v_l . w by lane l from its vcol, then
// w -= sum_l V[i][l] h_l with lane i's basis row in vrow
template <int GROUP>
__device__ __forceinline__ float cgs_pass_lds(float w, const float (&vrow)[GROUP],
                                              const float (&vcol)[GROUP], float *Lw, float *Lh,
                                              int gl) {
  Lw[gl] = w;
  wave_sync();
  const float h = bcast_dot<GROUP>(Lw, vcol);
  Lh[gl] = h;
  wave_sync();
  return w - bcast_dot<GROUP>(Lh, vrow);
}

// index of the first lane of largest residual 1 - ||V[i][:]||^2 (restart rule)
template <int GROUP>
__device__ __forceinline__ int restart_lane(float res, int gl, int k) {
  float bv = (gl < k) ? res : -1.f;
  int bi = gl;
#pragma unroll
  for (int m = 1; m < GROUP; m <<= 1) {
    const float ov = __shfl_xor(bv, m, GROUP);
    const int oi = __shfl_xor(bi, m, GROUP);
    if (ov > bv || (ov == bv && oi < bi)) {
      bv = ov;
      bi = oi;
    }
  }
  return bi;
}

// Lanczos recurrence shared by both broadcast paths, ParaLanczos.cu:1285-1374
// with the build's substitutions (ones/sqrt(k) start, full reorthogonalisation,
// restart). Lane i ends with my_a = alpha_i and my_b = beta_i (beta_0 = 0).
template <int GROUP>
__device__ uint32_t lanczos_lds(const float (&g)[GROUP], int k, int gl, float *R, float &my_a,
                                float &my_b) {
  float *Lv = R;
  float *Lw = R + GROUP;
  float *Lh = R + 2 * GROUP;
  float vrow[GROUP], vcol[GROUP];
#pragma unroll
  for (int l = 0; l < GROUP; ++l) {
    vrow[l] = 0.f;
    vcol[l] = 0.f;
  }
  float vi = (gl < k) ? (1.0f / sqrtf((float)k)) : 0.f;
  vrow[0] = vi;
  float vp = 0.f, bj = 0.f, amax = 0.f, bmax = 0.f;
  uint32_t st = 0;
  my_a = 0.f;
  my_b = 0.f;
  for (int j = 0; j < k; ++j) {
    Lv[gl] = vi;
    wave_sync();
    const bool mine = (gl == j);
    float acc0 = 0.f, acc1 = 0.f;
#pragma unroll
    for (int q = 0; q < GROUP / 4; ++q) {
      const float4 t = reinterpret_cast<const float4 *>(Lv)[q];
      acc0 = __builtin_fmaf(g[4 * q + 0], t.x, acc0);
      acc1 = __builtin_fmaf(g[4 * q + 1], t.y, acc1);
      acc0 = __builtin_fmaf(g[4 * q + 2], t.z, acc0);
      acc1 = __builtin_fmaf(g[4 * q + 3], t.w, acc1);
      vcol[4 * q + 0] = mine ? t.x : vcol[4 * q + 0];
      vcol[4 * q + 1] = mine ? t.y : vcol[4 * q + 1];
      vcol[4 * q + 2] = mine ? t.z : vcol[4 * q + 2];
      vcol[4 * q + 3] = mine ? t.w : vcol[4 * q + 3];
    }
    float w = acc0 + acc1;
    const float a = g_dot2<GROUP>(vi, w);
    my_a = mine ? a : my_a;
    amax = fmaxf(amax, fabsf(a));
    w = __builtin_fmaf(-a, vi, __builtin_fmaf(-bj, vp, w));
    const float pre = g_sum<GROUP>(w * w);
    w = cgs_pass_lds<GROUP>(w, vrow, vcol, Lw, Lh, gl);
    if (j == k - 1) break;
    float b2 = g_dot2<GROUP>(w, w);
    if (b2 < 0.5f * pre) {  // DGKS: w was mostly in span(V); orthogonalise again
      w = cgs_pass_lds<GROUP>(w, vrow, vcol, Lw, Lh, gl);
      b2 = g_dot2<GROUP>(w, w);
    }
    const float b = sqrtf(b2);
    vp = vi;
    float scale;
    if (b > TAU * (amax + bmax)) {
      bj = b;
      bmax = fmaxf(bmax, b);
      scale = 1.0f / b;
    } else {
      // restart: e_i for the first i of largest residual 1 - ||V[i][:]||^2,
      // orthogonalised twice against the basis; beta_{j+1} = 0
      st |= ST_BREAKDOWN;
      bj = 0.f;
      float res = 1.f;
#pragma unroll
      for (int l = 0; l < GROUP; ++l) res = __builtin_fmaf(-vrow[l], vrow[l], res);
      const int bi = restart_lane<GROUP>(res, gl, k);
      w = (gl == bi) ? 1.f : 0.f;
      w = cgs_pass_lds<GROUP>(w, vrow, vcol, Lw, Lh, gl);
      w = cgs_pass_lds<GROUP>(w, vrow, vcol, Lw, Lh, gl);
      scale = 1.0f / sqrtf(g_dot2<GROUP>(w, w));
    }
    vi = w * scale;
#pragma unroll
    for (int l = 1; l < GROUP; ++l) vrow[l] = (l == j + 1) ? vi : vrow[l];
    my_b = (gl == j + 1) ? bj : my_b;
  }
  return st;
}

// row gl of the basis (LDS, row stride STRIDE floats) into registers
template <int GROUP, int STRIDE>
__device__ __forceinline__ void load_basis_row(const float *V, int gl, float (&x)[GROUP]) {
  if constexpr (STRIDE % 4 == 0) {
#pragma unroll
    for (int q = 0; q < GROUP / 4; ++q) {
      const float4 t = reinterpret_cast<const float4 *>(V + gl * STRIDE)[q];
      x[4 * q + 0] = t.x;
      x[4 * q + 1] = t.y;
      x[4 * q + 2] = t.z;
      x[4 * q + 3] = t.w;
    }
  } else {
#pragma unroll
    for (int c = 0; c < GROUP; ++c) x[c] = V[gl * STRIDE + c];
  }
}

// column gl of the basis: entries 0..GROUP-1 of Lanczos vector v_gl
// (consecutive lanes read consecutive LDS words: no bank conflicts)
template <int GROUP, int STRIDE>
__device__ __forceinline__ void load_basis_col(const float *V, int gl, float (&x)[GROUP]) {
#pragma unroll
  for (int r = 0; r < GROUP; ++r) x[r] = V[r * STRIDE + gl];
}

// Gram-Schmidt pass for groups of 16/32/64 (DPP broadcasts) against the LDS
// basis V (V[i*STRIDE + l] = entry i of v_l, zero for vectors not yet formed):
// lane l forms h_l = v_l . w from column l, lane i subtracts sum_l V[i][l] h_l
template <int GROUP, int STRIDE>
__device__ __forceinline__ float cgs_pass_dpp(float w, const float *V, int gl) {
  float x[GROUP];
  load_basis_col<GROUP, STRIDE>(V, gl, x);
  const float h = group_dot_dpp<GROUP>(w, x);
  load_basis_row<GROUP, STRIDE>(V, gl, x);
  return w - group_dot_dpp<GROUP>(h, x);
}

template <int GROUP, int STRIDE>
__device__ uint32_t lanczos_dpp(const float (&g)[GROUP], int k, int gl, float *R, float &my_a,
                                float &my_b) {
  float *V = R;  // the gather tile is dead: reuse it as the basis
  float vi = (gl < k) ? (1.0f / sqrtf((float)k)) : 0.f;
  wave_sync();
  if constexpr (STRIDE % 4 == 0) {
#pragma unroll
    for (int q = 0; q < GROUP / 4; ++q)
      reinterpret_cast<float4 *>(V + gl * STRIDE)[q] = make_float4(q == 0 ? vi : 0.f, 0.f, 0.f, 0.f);
  } else {
#pragma unroll
    for (int c = 0; c < GROUP; ++c) V[gl * STRIDE + c] = (c == 0) ? vi : 0.f;
  }
  float vp = 0.f, bj = 0.f, amax = 0.f, bmax = 0.f;
  uint32_t st = 0;
  my_a = 0.f;
  my_b = 0.f;
  for (int j = 0; j < k; ++j) {
    // matvec w_i = sum_c G_S[i][c] v_c
    float w = group_dot_dpp<GROUP>(vi, g);
    const float a = g_dot2<GROUP>(vi, w);
    my_a = (gl == j) ? a : my_a;
    amax = fmaxf(amax, fabsf(a));
    w = __builtin_fmaf(-a, vi, __builtin_fmaf(-bj, vp, w));
    const float pre = g_sum_dpp<GROUP>(w * w);
    wave_sync();
    w = cgs_pass_dpp<GROUP, STRIDE>(w, V, gl);
    if (j == k - 1) break;
    float b2 = g_dot2<GROUP>(w, w);
    if (b2 < 0.5f * pre) {  // DGKS: w was mostly in span(V); orthogonalise again
      w = cgs_pass_dpp<GROUP, STRIDE>(w, V, gl);
      b2 = g_dot2<GROUP>(w, w);
    }
    const float b = sqrtf(b2);
    vp = vi;
    float scale;
    if (b > TAU * (amax + bmax)) {
      bj = b;
      bmax = fmaxf(bmax, b);
      scale = 1.0f / b;
    } else {
      // restart: e_i for the first i of largest residual 1 - ||V[i][:]||^2,
      // orthogonalised twice against the basis; beta_{j+1} = 0
      st |= ST_BREAKDOWN;
      bj = 0.f;
      float vr[GROUP];
      load_basis_row<GROUP, STRIDE>(V, gl, vr);
      float res = 1.f;
#pragma unroll
      for (int l = 0; l < GROUP; ++l) res = __builtin_fmaf(-vr[l], vr[l], res);
      const int bi = restart_lane<GROUP>(res, gl, k);
      w = (gl == bi) ? 1.f : 0.f;
      w = cgs_pass_dpp<GROUP, STRIDE>(w, V, gl);
      w = cgs_pass_dpp<GROUP, STRIDE>(w, V, gl);
      scale = 1.0f / sqrtf(g_dot2<GROUP>(w, w));
    }
    vi = w * scale;
    V[gl * STRIDE + j + 1] = vi;
    wave_sync();
    my_b = (gl == j + 1) ? bj : my_b;
  }
  return st;
}

template <int GROUP, int STRIDE>
__device__ __forceinline__ uint32_t lanczos_core(const float (&g)[GROUP], int k, int gl, float *R,
                                                 float &my_a, float &my_b) {
  uint32_t st;
  if constexpr (GROUP >= 16) st = lanczos_dpp<GROUP, STRIDE>(g, k, gl, R, my_a, my_b);
  else st = lanczos_lds<GROUP>(g, k, gl, R, my_a, my_b);
  const bool finite = (gl >= k) || (isfinite(my_a) && isfinite(my_b));
  if (!g_all<GROUP>(finite)) st |= ST_NONFINITE;
  return st;
}

// Stage (b) epilogue: T to the workspace, tri[i] = (alpha_i, beta_i) in FP32
// (beta_0 = 0), and the subset's bracket data in FP64: the Gershgorin interval
// of T and the range of its diagonal
template <int GROUP>
__device__ __forceinline__ void store_tridiag(float2 *tri, double4 *bounds, int k, int gl,
                                              float my_a, float my_b) {
  const double al = my_a;
  const double b = my_b;
  const double bn = __shfl_down(b, 1, GROUP);  // beta_{i+1}
  double glo = INFINITY, ghi = -INFINITY, amin = INFINITY, amax = -INFINITY;
  if (gl < k) {
    tri[gl] = make_float2(my_a, gl > 0 ? my_b : 0.f);
    const double r = (gl > 0 ? fabs(b) : 0.0) + (gl + 1 < k ? fabs(bn) : 0.0);
    glo = al - r;
    ghi = al + r;
    amin = amax = al;
  }
  glo = g_min<GROUP>(glo);
  ghi = g_max<GROUP>(ghi);
  amin = g_min<GROUP>(amin);
  amax = g_max<GROUP>(amax);
  if (gl == 0) *bounds = make_double4(glo, ghi, amin, amax);
}

// waves per SIMD each lane-group size is register-budgeted for (VGPRs <= 512/n)
template <int GROUP>
constexpr int waves_per_simd() {
  return GROUP == 8 ? 7 : GROUP == 16 ? 6 : GROUP == 32 ? 4 : 2;
}

template <int GROUP, int PAD, int MODE>
__device__ __forceinline__ void solve_subset(const SolveArgs &a, float *R, int gl, int64_t s) {
  constexpr int STRIDE = GROUP + PAD;
  int k;
  int64_t off;
  if (a.offsets) {
    off = a.offsets[s];
    k = (int)(a.offsets[s + 1] - off);
  } else {
    k = a.kuni;
    off = s * (int64_t)k;
  }
  uint32_t st = (k < 1 || k > GROUP || k > a.kcap) ? ST_BADK : 0u;
  int my = 0;
  if (!st && gl < k) {
    my = a.idx[off + gl];
    if (my < 0 || (int64_t)my >= a.n) st = ST_BADINDEX;
  }
  st = g_or<GROUP>(st);

  // ---- (a) gather: row r of the subset per step, lanes = columns c >= r ----
  float vals[GROUP];
#pragma unroll
  for (int r = 0; r < GROUP; ++r) {
    const int ir = __shfl(my, r, GROUP);
    float x = 0.f;
    if (st == 0 && r < k && gl < k && gl >= r) {
      const int64_t lo = min(ir, my), hi = max(ir, my);
      x = a.G[lo * a.ld + hi];
    }
    vals[r] = x;
  }
  bool fin = true;
#pragma unroll
  for (int r = 0; r < GROUP; ++r) fin = fin && isfinite(vals[r]);
  if (!g_all<GROUP>(fin)) st |= ST_NONFINITE;
#pragma unroll
  for (int r = 0; r < GROUP; ++r) {
    if (gl >= r) {
      R[r * STRIDE + gl] = vals[r];
      R[gl * STRIDE + r] = vals[r];
    }
  }
  wave_sync();
  float g[GROUP];
  if constexpr (STRIDE % 4 == 0) {
#pragma unroll
    for (int q = 0; q < GROUP / 4; ++q) {
      const float4 t = reinterpret_cast<const float4 *>(R + gl * STRIDE)[q];
      g[4 * q + 0] = t.x;
      g[4 * q + 1] = t.y;
      g[4 * q + 2] = t.z;
      g[4 * q + 3] = t.w;
    }
  } else {
#pragma unroll
    for (int c = 0; c < GROUP; ++c) g[c] = R[gl * STRIDE + c];
  }
  wave_sync();

  if constexpr (MODE == MODE_DUMP) {
    if (gl < k && st == 0) {
      float *o = a.gs_out + s * (int64_t)k * k + (int64_t)gl * k;
#pragma unroll
      for (int c = 0; c < GROUP; ++c)
        if (c < k) o[c] = g[c];
    }
    return;
  } else if constexpr (MODE == MODE_CHECKSUM) {
    float acc = 0.f;
#pragma unroll
    for (int c = 0; c < GROUP; ++c) acc += g[c];
    acc = g_sum<GROUP>(acc);
    if (gl == 0) a.gs_out[s] = st ? NAN : acc;
    return;
  } else {
    constexpr uint32_t FATAL = ST_BADK | ST_BADINDEX | ST_NONFINITE;
    if (!(st & FATAL)) {
      float my_a, my_b;
      st |= lanczos_core<GROUP, STRIDE>(g, k, gl, R, my_a, my_b);
      if (!(st & FATAL))
        store_tridiag<GROUP>(a.tri + s * (int64_t)a.tstride, a.bounds + s, k, gl, my_a, my_b);
    }
    if (gl == 0) a.status[s] = st;
  }
}

template <int GROUP, int PAD, int MODE>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(waves_per_simd<GROUP>())))
void solve_kernel(SolveArgs a) {
  constexpr int SPB = BLOCK / GROUP;
  constexpr int STRIDE = GROUP + PAD;
  constexpr int REGION = GROUP * STRIDE;
  static_assert(REGION >= 5 * GROUP, "LDS region too small for the vector buffers");
  __shared__ __attribute__((aligned(16))) float lds[SPB * REGION];

  const int tid = threadIdx.x;
  const int gl = tid % GROUP;
  const int gib = tid / GROUP;
  float *R = lds + gib * REGION;

  // slots of this launch: all of them, or the run of one bin (persistent grid);
  // lane groups advance independently, nothing below syncs the block
  int64_t cnt = a.nslots;
  const int32_t *ord = a.order;
  if (a.bins) {
    cnt = a.bins[BIN_COUNT + a.bin];
    ord = a.order + a.bins[BIN_START + a.bin];
  }
  for (int64_t slot = (int64_t)blockIdx.x * SPB + gib; slot < cnt;
       slot += (int64_t)gridDim.x * SPB) {
    solve_subset<GROUP, PAD, MODE>(a, R, gl, ord ? (int64_t)ord[slot] : slot);
    wave_sync();
  }
}

// ---------------------------------------------------------------------------
// Wide subsets (64 < k <= 128): one 128-lane workgroup (two wavefronts) per
// subset. Lane i owns row i of G_S in VGPRs; the padded LDS tile that receives
// the gather becomes the Lanczos basis V (V[i*S + l] = entry i of v_l), read
// by rows (b128) and by columns (conflict-free b32); the broadcasts of v, w
// and the Gram-Schmidt coefficients and the cross-wave halves of every
// reduction go through LDS with workgroup barriers.
// ---------------------------------------------------------------------------
constexpr int WIDE = 128;

__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename T>
struct WideShared {
  float lv[WIDE];  // next Lanczos vector, unnormalised: the matvec input
  float lw[WIDE];  // w broadcast
  T lwT[sizeof(T) == 8 ? WIDE : 4];  // w in T (FP64 coefficients only)
  float lh[WIDE];  // Gram-Schmidt coefficients (accumulated in T, applied in FP32)
  float ra[4];     // alpha partials (p, e) of the two wavefronts
  float rb[4];     // beta^2 partials
  float rp[2];     // ||w||^2 partials before a Gram-Schmidt pass
  float red[8];
  double redd[8];
  int ired[4];
};

template <typename T>
__device__ __forceinline__ double wide_min(double x, WideShared<T> &sh) {
  x = g_min<64>(x);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh.redd[threadIdx.x >> 6] = x;
  __syncthreads();
  return fmin(sh.redd[0], sh.redd[1]);
}

template <typename T>
__device__ __forceinline__ double wide_max(double x, WideShared<T> &sh) {
  x = g_max<64>(x);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh.redd[2 + (threadIdx.x >> 6)] = x;
  __syncthreads();
  return fmax(sh.redd[2], sh.redd[3]);
}

// sum_c x[c] * L[c] over a 16-byte aligned LDS vector read by broadcast;
// four accumulators by index mod 4
template <int N>
__device__ __forceinline__ float lds_dot4(const float *L, const float (&x)[N]) {
  float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
#pragma unroll
  for (int q = 0; q < N / 4; ++q) {
    const float4 t = reinterpret_cast<const float4 *>(L)[q];
    a0 = __builtin_fmaf(x[4 * q + 0], t.x, a0);
    a1 = __builtin_fmaf(x[4 * q + 1], t.y, a1);
    a2 = __builtin_fmaf(x[4 * q + 2], t.z, a2);
    a3 = __builtin_fmaf(x[4 * q + 3], t.w, a3);
  }
  return (a0 + a1) + (a2 + a3);
}

// beta^2 = w.w (compensated: per-wave Dot2, the two partials combined in
// wave order) and the next matvec input lv = w, behind one barrier
template <typename T>
__device__ __forceinline__ float wide_beta2(float w, WideShared<T> &sh) {
  const int gl = threadIdx.x;
  float p, e;
  g_dot2_pe<64>(w, w, p, e);
  sh.lv[gl] = w;
  if ((gl & 63) == 0) {
    sh.rb[2 * (gl >> 6)] = p;
    sh.rb[2 * (gl >> 6) + 1] = e;
  }
  __syncthreads();
  float p0 = sh.rb[0], e0 = sh.rb[1];
  comp_step(p0, e0, sh.rb[2], sh.rb[3]);
  return p0 + e0;
}

// Gram-Schmidt pass against basis columns 0..j in LDS: lane l forms
// h_l = v_l . w from column l over the first k rows (accumulated in T), lane
// i subtracts sum_l V[i][l] h_l from row i (FP32). With WANT_PRE, ||w||^2
// before the pass comes back in pre (its partials ride with the broadcast).
template <int S, typename T, bool WANT_PRE>
__device__ __forceinline__ float wide_cgs(float w, int j, int k, const float *V,
                                          WideShared<T> &sh, float &pre) {
  const int gl = threadIdx.x;
  float ws = 0.f;
  if constexpr (WANT_PRE) ws = g_sum_dpp<64>(w * w);
  sh.lw[gl] = w;
  if constexpr (sizeof(T) == 8) sh.lwT[gl] = (T)w;
  if (WANT_PRE && (gl & 63) == 0) sh.rp[gl >> 6] = ws;
  __syncthreads();
  if constexpr (WANT_PRE) pre = sh.rp[0] + sh.rp[1];
  const int ncol = j + 1;
  T hc = 0;
  if (gl < ncol) {
    // 16 rows per iteration, loaded before the FMAs consume them; rows >= k
    // are zero (k <= 128, so the rounded count stays inside the tile)
    const int rn = (k + 15) & ~15;
    T a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    for (int r = 0; r < rn; r += 16) {
      float v[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) v[u] = V[(r + u) * S + gl];
      if constexpr (sizeof(T) == 8) {
        const double2 *lw2 = reinterpret_cast<const double2 *>(sh.lwT + r);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const double2 l0 = lw2[2 * u], l1 = lw2[2 * u + 1];
          a0 = fmaT((T)v[4 * u + 0], (T)l0.x, a0);
          a1 = fmaT((T)v[4 * u + 1], (T)l0.y, a1);
          a2 = fmaT((T)v[4 * u + 2], (T)l1.x, a2);
          a3 = fmaT((T)v[4 * u + 3], (T)l1.y, a3);
        }
      } else {
        const float4 *lw4 = reinterpret_cast<const float4 *>(sh.lw + r);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const float4 l = lw4[u];
          a0 = fmaT((T)v[4 * u + 0], (T)l.x, a0);
          a1 = fmaT((T)v[4 * u + 1], (T)l.y, a1);
          a2 = fmaT((T)v[4 * u + 2], (T)l.z, a2);
          a3 = fmaT((T)v[4 * u + 3], (T)l.w, a3);
        }
      }
    }
    hc = (a0 + a1) + (a2 + a3);
  }
  sh.lh[gl] = (float)hc;  // zero for the columns not yet formed
  __syncthreads();
  // row update over columns 0..j rounded up to 4 (the rest hold zeros)
  const int nc4 = (ncol + 3) & ~3;
  float u0 = 0.f, u1 = 0.f, u2 = 0.f, u3 = 0.f;
  for (int c = 0; c < nc4; c += 4) {
    const float4 t = *reinterpret_cast<const float4 *>(sh.lh + c);
    float4 r;
    if constexpr (S % 4 == 0) {
      r = *reinterpret_cast<const float4 *>(V + gl * S + c);
    } else {
      r = make_float4(V[gl * S + c], V[gl * S + c + 1], V[gl * S + c + 2], V[gl * S + c + 3]);
    }
    u0 = __builtin_fmaf(r.x, t.x, u0);
    u1 = __builtin_fmaf(r.y, t.y, u1);
    u2 = __builtin_fmaf(r.z, t.z, u2);
    u3 = __builtin_fmaf(r.w, t.w, u3);
  }
  return w - ((u0 + u1) + (u2 + u3));
}

template <int PAD, int MODE, typename T>
__device__ __forceinline__ void solve_wide_subset(const SolveArgs &a, float *V, WideShared<T> &sh,
                                                  int64_t s) {
  constexpr int S = WIDE + PAD;
  const int gl = threadIdx.x;
  int k;
  int64_t off;
  if (a.offsets) {
    off = a.offsets[s];
    k = (int)(a.offsets[s + 1] - off);
  } else {
    k = a.kuni;
    off = s * (int64_t)k;
  }
  const bool badk = k < 1 || k > WIDE || k > a.kcap;
  uint32_t st = badk ? ST_BADK : 0u;
  int my = 0;
  if (!st && gl < k) {
    my = a.idx[off + gl];
    if (my < 0 || (int64_t)my >= a.n) st = ST_BADINDEX;
  }
  if (__syncthreads_or(st != 0)) st |= badk ? ST_BADK : ST_BADINDEX;
  int *lidx = reinterpret_cast<int *>(sh.lv);
  lidx[gl] = my;
  __syncthreads();

  // ---- (a) gather: row r per step, lanes = columns c >= r, 8 loads in flight ----
  bool fin = true;
  for (int r0 = 0; r0 < WIDE; r0 += 8) {
    float x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int r = r0 + u;
      x[u] = 0.f;
      if (st == 0 && r < k && gl < k && gl >= r) {
        const int ir = lidx[r];
        const int64_t lo = min(ir, my), hi = max(ir, my);
        x[u] = a.G[lo * a.ld + hi];
      }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int r = r0 + u;
      fin = fin && isfinite(x[u]);
      if (gl >= r) {
        V[r * S + gl] = x[u];
        V[gl * S + r] = x[u];
      }
    }
  }
  if (__syncthreads_or(!fin)) st |= ST_NONFINITE;
  float g[WIDE];
#pragma unroll
  for (int c = 0; c < WIDE; ++c) g[c] = V[gl * S + c];
  __syncthreads();

  if constexpr (MODE == MODE_DUMP) {
    if (gl < k && st == 0) {
      float *o = a.gs_out + s * (int64_t)k * k + (int64_t)gl * k;
#pragma unroll
      for (int c = 0; c < WIDE; ++c)
        if (c < k) o[c] = g[c];
    }
    return;
  } else if constexpr (MODE == MODE_CHECKSUM) {
    float acc = 0.f;
#pragma unroll
    for (int c = 0; c < WIDE; ++c) acc += g[c];
    acc = g_sum<64>(acc);
    __syncthreads();
    if ((gl & 63) == 0) sh.red[gl >> 6] = acc;
    __syncthreads();
    if (gl == 0) a.gs_out[s] = st ? NAN : sh.red[0] + sh.red[1];
    return;
  } else {
    constexpr uint32_t FATAL = ST_BADK | ST_BADINDEX | ST_NONFINITE;
    if (!(st & FATAL)) {
      // ---- (b) Lanczos recurrence, ParaLanczos.cu:1285-1374 with the build's
      // substitutions (ones/sqrt(k) start, full reorthogonalisation, restart).
      // Each step: matvec of the unnormalised vector in lv, scaled by 1/beta,
      // alpha's compensated partials exchanged behind the first barrier;
      // three-term update; Gram-Schmidt pass (||w||^2 rides along, two
      // barriers); beta^2 published with the next vector (one barrier). ----
      float scale = 1.0f / sqrtf((float)k);
      float vi = (gl < k) ? scale : 0.f;
#pragma unroll
      for (int c = 0; c < WIDE; ++c) V[gl * S + c] = (c == 0) ? vi : 0.f;
      sh.lv[gl] = (gl < k) ? 1.f : 0.f;
      __syncthreads();
      float vp = 0.f, bj = 0.f, amax = 0.f, bmax = 0.f, my_a = 0.f, my_b = 0.f;
      for (int j = 0; j < k; ++j) {
        const float mv = lds_dot4<WIDE>(sh.lv, g);
        float p, e;
        g_dot2_pe<64>(vi, mv, p, e);
        if ((gl & 63) == 0) {
          sh.ra[2 * (gl >> 6)] = p;
          sh.ra[2 * (gl >> 6) + 1] = e;
        }
        __syncthreads();
        float p0 = sh.ra[0], e0 = sh.ra[1];
        comp_step(p0, e0, sh.ra[2], sh.ra[3]);
        const float al = (p0 + e0) * scale;
        float w = mv * scale;
        my_a = (gl == j) ? al : my_a;
        amax = fmaxf(amax, fabsf(al));
        if (j == k - 1) break;
        w = __builtin_fmaf(-al, vi, __builtin_fmaf(-bj, vp, w));
        float pre = 0.f, unused;
        w = wide_cgs<S, T, true>(w, j, k, V, sh, pre);
        float b2 = wide_beta2(w, sh);
        if (b2 < 0.5f * pre) {  // DGKS: w was mostly in span(V); orthogonalise again
          w = wide_cgs<S, T, false>(w, j, k, V, sh, unused);
          b2 = wide_beta2(w, sh);
        }
        const float b = sqrtf(b2);
        vp = vi;
        if (b > TAU * (amax + bmax)) {
          bj = b;
          bmax = fmaxf(bmax, b);
          scale = 1.0f / b;
        } else {
          // restart: e_i for the first i of largest residual 1 - ||V[i][:]||^2,
          // orthogonalised twice against the basis; beta_{j+1} = 0
          st |= ST_BREAKDOWN;
          bj = 0.f;
          float res = 1.f;
          for (int l = 0; l < WIDE; ++l) res = __builtin_fmaf(-V[gl * S + l], V[gl * S + l], res);
          const int wb = restart_lane<64>(res, gl, k);
          const float wv = __shfl(gl < k ? res : -1.f, wb & 63, 64);
          __syncthreads();
          if ((gl & 63) == 0) {
            sh.red[gl >> 6] = wv;
            sh.ired[gl >> 6] = wb;
          }
          __syncthreads();
          const int bi = (sh.red[1] > sh.red[0]) ? sh.ired[1] : sh.ired[0];
          w = (gl == bi) ? 1.f : 0.f;
          w = wide_cgs<S, T, false>(w, j, k, V, sh, unused);
          w = wide_cgs<S, T, false>(w, j, k, V, sh, unused);
          scale = 1.0f / sqrtf(wide_beta2(w, sh));
        }
        vi = w * scale;
        V[gl * S + j + 1] = vi;
        my_b = (gl == j + 1) ? bj : my_b;
      }
      const bool finite = (gl >= k) || (isfinite(my_a) && isfinite(my_b));
      if (__syncthreads_or(!finite)) st |= ST_NONFINITE;
      if (!(st & FATAL)) {
        // T and its bracket data to the workspace
        float2 *tri = a.tri + s * (int64_t)a.tstride;
        sh.lw[gl] = my_b;
        __syncthreads();
        const double al = my_a, b = my_b, bn = (gl + 1 < WIDE) ? (double)sh.lw[gl + 1] : 0.0;
        double glo = INFINITY, ghi = -INFINITY, amin = INFINITY, amx = -INFINITY;
        if (gl < k) {
          tri[gl] = make_float2(my_a, gl > 0 ? my_b : 0.f);
          const double r = (gl > 0 ? fabs(b) : 0.0) + (gl + 1 < k ? fabs(bn) : 0.0);
          glo = al - r;
          ghi = al + r;
          amin = amx = al;
        }
        glo = wide_min(glo, sh);
        ghi = wide_max(ghi, sh);
        amin = wide_min(amin, sh);
        amx = wide_max(amx, sh);
        if (gl == 0) a.bounds[s] = make_double4(glo, ghi, amin, amx);
      }
    }
    if (gl == 0) a.status[s] = st;
  }
}

template <int PAD, int MODE, typename T = float>
__global__ __launch_bounds__(WIDE) void solve_wide_kernel(SolveArgs a) {
  __shared__ __attribute__((aligned(16))) float V[WIDE * (WIDE + PAD)];
  __shared__ __attribute__((aligned(16))) WideShared<T> sh;
  int64_t cnt = a.nslots;
  const int32_t *ord = a.order;
  if (a.bins) {
    cnt = a.bins[BIN_COUNT + a.bin];
    ord = a.order + a.bins[BIN_START + a.bin];
  }
  // the slot is uniform over the workgroup, so every barrier inside is too
  for (int64_t slot = blockIdx.x; slot < cnt; slot += gridDim.x) {
    solve_wide_subset<PAD, MODE, T>(a, V, sh, ord ? (int64_t)ord[slot] : slot);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Row-part kernels, subset capacity KC = 128 or 256, Q = 2 or 4 parts per
// row: one workgroup of Q KC lanes per subset. Lane t = (row i = t % KC,
// part h = t / KC) holds G_S[i][KC/Q h .. KC/Q (h + 1) - 1] in VGPRs. The
// Lanczos basis V keeps columns 0..L-1 in LDS (row-major, stride S). At
// KC = 256 an FP32 G_S and an FP32 basis do not both fit one CU's LDS, and at
// KC = 128 three subsets share a CU, so columns L..KC-1 live in VGPRs: the
// columns L + Q m + h (m = 0, 1, ...) of part h, row i, on lane (i, h), as a
// shift register: the newest in slot vr[0], column m in slot nm_h - 1 - m
// once nm_h of them are formed. Adding a column shifts the part's slots up by
// one (a wave-uniform branch: the part index is uniform per wavefront).
// Gram-Schmidt coefficients of the LDS columns are column dot products (the
// lanes split into columns x row groups); those of the register columns come
// from a wavefront reduce-scatter of the lanes' products (wave_rs) and a sum
// over the wavefronts of each part in LDS. The coefficients are accumulated
// in T (float, or double: FP64 reorthogonalisation coefficients) and applied
// in an FP32 update.
//   KC = 128, Q = 2: 256 lanes, L = 92 (three workgroups per CU: a third
//   wavefront per SIMD hides more LDS latency than the 18 register columns
//   cost; L = 128 with two workgroups ran k = 128 at 171 us per subset and CU
//   in FP32, 184 in FP64, against 131 and 169)
//   KC = 256, Q = 2: 512 lanes, L = 148 (the widest LDS share that fits: each
//   column moved out of the VGPRs relieves the step loop's register pressure)
//   Q = 4 (1024 lanes, four wavefronts per SIMD) agreed with Q = 2 at L = 140
//   but spilled at 128 VGPRs per lane and ran at half the speed; its LDS
//   share does not fit beside L = 148, so it is not instantiated.
// ---------------------------------------------------------------------------
template <int KC, int Q>
struct HR {
  static_assert((KC == 128 && Q == 2) || (KC == 256 && (Q == 2 || Q == 4)), "row-part layout");
  static constexpr int NT = Q * KC;             // threads per workgroup
  static constexpr int HW = KC / Q;             // G_S columns per lane
  static constexpr int WPH = KC / 64;           // wavefronts per row part
  static constexpr int NW = Q * WPH;            // wavefronts
  static constexpr int Q_ = Q;
  static constexpr int LOGK = KC == 256 ? 8 : 7;
  static constexpr int L = KC == 256 ? 148 : 92;  // basis columns in LDS
  static constexpr int S = KC == 256 ? 148 : 92;  // LDS row stride: 16-byte rows, stride/4 odd
  static constexpr int R = (KC - L) / Q;        // register basis columns per part
  // KC = 256: the register columns of a part are held in blocks of eight
  // rows. Lane e of the block of rows 8b..8b+7 holds the part's columns m
  // with m mod 8 = e (slot s = m / 8), entry p of a slot being row
  // 8b + (p xor e). A column's dot then sums eight rows inside the lane, so
  // the wavefront reduce-scatter moves one value per slot instead of one per
  // column, and the xor order lets the update's per-row partials meet by
  // three DPP adds with no selects. KC = 128 keeps one row per lane.
  static constexpr bool OCT = KC == 256;
  static constexpr int RB = OCT ? (R + 7) / 8 : R;  // register slots per lane
  static constexpr int RA = OCT ? 8 * RB : (R > 0 ? R : 1);  // their array extent
  static_assert(S % 4 == 0 && (S / 4) % 2 == 1 && L % 4 == 0 && L <= S, "row layout");
  static_assert(!OCT || (Q == 2 && NW == 8 && RB <= 8), "8-row blocks: slots fit one reduce-scatter");
};
constexpr int DPP_ROW_ROR8 = 0x128;  // lane i <- lane (i + 8) mod 16 within a row
constexpr int DPP_QUAD_3210 = 0x1B;  // quad_perm [3,2,1,0]  (partner lane ^ 3)

// reduce-scatter chunk: 16 FP32 or 8 FP64 values (register budget)
template <typename T>
constexpr int k2_chunk() {
  return sizeof(T) == 8 ? 8 : 16;
}

// LDS work buffers. Members of a union are live in disjoint parts of a
// Lanczos step, separated by barriers: pw (matvec partials) from the matvec
// to the barrier after it, rpart from the reduce-scatter to the coefficient
// sums, part from the update to the sum after the next barrier; lw (the
// broadcast of w, and the index staging of the gather) until the column dots
// end, lh from the coefficient sums to the update; lvn from its store after
// a Gram-Schmidt pass to the next matvec, cp from the column dots to the
// coefficient sums.
template <int KC, int Q, typename T>
struct HRShared {
  using C = HR<KC, Q>;
  static constexpr int RP = C::OCT ? 8 * C::RB
                          : C::R > 0 ? (C::R + k2_chunk<T>() - 1) / k2_chunk<T>() * k2_chunk<T>()
                                     : 1;  // register partial slots per (part, wavefront)
  static constexpr int RPN = C::Q_ * C::WPH * RP;
  // LDS-column coefficient partials: four per dot unit (8 NW units); at
  // KC = 256 also the tail groups' four row parts (256 + TST p + c - 128)
  static constexpr int TST = (C::L - 128 + 3) / 4 * 4;  // tail columns, rounded to groups
  static constexpr int CPN = KC == 256 ? 256 + 4 * TST : 32 * C::NW;
  // every buffer read 16 bytes at a time starts on a 16-byte boundary
  union alignas(16) {
    float lw[KC];    // w broadcast
    float lh[KC];    // Gram-Schmidt coefficients (accumulated in T, applied in FP32)
  };
  // the register columns' coefficients per part (16-byte reads): in slot
  // order, or at KC = 256 by [lane of the block][slot]
  alignas(16) float lhr[Q][C::OCT ? 64 : C::R > 0 ? (C::R + 3) / 4 * 4 : 4];
  union alignas(16) {
    float pw[Q][KC];    // matvec partials of the row parts
    float part[Q][KC];  // Gram-Schmidt update partials of the row parts
    // register-column partials per part and wavefront [h][wv][m] (KC = 128:
    // written after the broadcast of w, when pw is dead)
    T rpart_u[C::OCT ? 1 : RPN];
  };
  // KC = 256: with FP64 coefficients the register-column partials are
  // written before the barrier that publishes w, while other wavefronts may
  // still read pw
  alignas(16) T rpart_s[C::OCT ? RPN : 1];
  union alignas(16) {
    float lvn[KC];      // next Lanczos vector, unnormalised: the matvec input
    T cp[CPN];          // LDS-column coefficient partials
    double redd[16];    // bounds reduction, after the recurrence
  };
  __device__ T *rpart(int h, int wv) { return (C::OCT ? rpart_s : rpart_u) + (h * C::WPH + wv) * RP; }
  // FP64 coefficients: w widened once to FP64 for the column dots, so the
  // dot loops convert only V's entries (at KC = 256 this takes 2 KiB of the
  // 2.2 KiB of LDS left beside the basis; at KC = 128, 1 KiB per workgroup
  // keeps three workgroups per CU)
  static constexpr bool WD = sizeof(T) == 8;
  alignas(16) double lwd[WD ? KC : 2];
  double ra[C::NW];     // alpha partials (FP64) of all wavefronts
  double rb[C::WPH];    // beta^2 partials (FP64) of the wavefronts of part 0
  float rp[C::WPH];     // ||w||^2 partials before a Gram-Schmidt pass
  float red[16];
  int ired[8];
};
// The workgroup's LDS: the work buffers first, at addresses that the 16-bit
// offset fields of LDS instructions reach (no address arithmetic for the
// broadcasts of w and the coefficients), then V.
template <int KC, int Q, typename T>
struct HRLds {
  HRShared<KC, Q, T> sh;
  alignas(16) float VL[KC * HR<KC, Q>::S];
};
template <int KC, int Q, typename T>
constexpr bool hr_aligned() {
  using SH = HRShared<KC, Q, T>;
  using LD = HRLds<KC, Q, T>;
  return offsetof(SH, lvn) % 16 == 0 && offsetof(SH, lw) % 16 == 0 && offsetof(SH, lhr) % 16 == 0 &&
         offsetof(SH, pw) % 16 == 0 && offsetof(LD, VL) % 16 == 0;
}
static_assert(hr_aligned<256, 2, float>() && hr_aligned<256, 2, double>() &&
                  hr_aligned<128, 2, float>() && hr_aligned<128, 2, double>(),
              "16-byte reads of the work buffers");
template <int KC, int Q>
constexpr size_t hr_lds_bytes() {
  return sizeof(HRLds<KC, Q, double>);
}
static_assert(hr_lds_bytes<256, 2>() <= 163840, "LDS budget, 512 lanes");
static_assert(3 * hr_lds_bytes<128, 2>() <= 163840, "LDS budget, 256 lanes: three workgroups per CU");

template <typename T>
__device__ __forceinline__ T swap_add(T a, T b, bool wide32);
template <>
__device__ __forceinline__ float swap_add<float>(float a, float b, bool wide32) {
  // v_permlane32_swap (lanes 32..63 of a <-> lanes 0..31 of b) or
  // v_permlane16_swap (odd rows of a <-> even rows of b): afterwards the lane
  // holds its own and its partner's value of the element its half keeps
  const unsigned ua = __builtin_bit_cast(unsigned, a), ub = __builtin_bit_cast(unsigned, b);
  const auto r = wide32 ? __builtin_amdgcn_permlane32_swap(ua, ub, false, false)
                        : __builtin_amdgcn_permlane16_swap(ua, ub, false, false);
  return __builtin_bit_cast(float, (unsigned)r[0]) + __builtin_bit_cast(float, (unsigned)r[1]);
}
template <>
__device__ __forceinline__ double swap_add<double>(double a, double b, bool wide32) {
  const unsigned long long ua = __builtin_bit_cast(unsigned long long, a);
  const unsigned long long ub = __builtin_bit_cast(unsigned long long, b);
  const unsigned alo = (unsigned)ua, ahi = (unsigned)(ua >> 32);
  const unsigned blo = (unsigned)ub, bhi = (unsigned)(ub >> 32);
  const auto lo = wide32 ? __builtin_amdgcn_permlane32_swap(alo, blo, false, false)
                         : __builtin_amdgcn_permlane16_swap(alo, blo, false, false);
  const auto hi = wide32 ? __builtin_amdgcn_permlane32_swap(ahi, bhi, false, false)
                         : __builtin_amdgcn_permlane16_swap(ahi, bhi, false, false);
  const double x0 = __builtin_bit_cast(
      double, ((unsigned long long)(unsigned)hi[0] << 32) | (unsigned)lo[0]);
  const double x1 = __builtin_bit_cast(
      double, ((unsigned long long)(unsigned)hi[1] << 32) | (unsigned)lo[1]);
  return x0 + x1;
}

template <int CTRL>
__device__ __forceinline__ double dpp(double x) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
  const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)u, CTRL, 0xf, 0xf, true);
  const unsigned hi =
      (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), CTRL, 0xf, 0xf, true);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Reduce-scatter of N (16 or 8) values over the 64 lanes of a wavefront:
// lane q returns the sum over all lanes of x[q >> S] (S = 2 for N = 16,
// 3 for N = 8; 64/N lanes hold each element). Partners by stage: lane ^ 32
// and ^ 16 (permlane swaps), ^ 8 (row_ror:8), ^ 7 (row_half_mirror), then an
// all-reduce over ^ 2 and ^ 1 (quad_perm) - the partner sets span all 64 lanes.
template <int N, typename T>
__device__ __forceinline__ T wave_rs(const T (&x)[N], int q) {
  T y[N / 2], z[N / 4];
#pragma unroll
  for (int u = 0; u < N / 2; ++u) y[u] = swap_add<T>(x[u], x[u + N / 2], true);
#pragma unroll
  for (int u = 0; u < N / 4; ++u) z[u] = swap_add<T>(y[u], y[u + N / 4], false);
  const bool b3 = (q >> 3) & 1;
  T r;
  if constexpr (N == 16) {
    const bool b2 = (q >> 2) & 1;
    T v[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const T send = b3 ? z[u] : z[u + 2];
      const T keep = b3 ? z[u + 2] : z[u];
      v[u] = keep + dpp<DPP_ROW_ROR8>(send);
    }
    const T send = b2 ? v[0] : v[1];
    r = (b2 ? v[1] : v[0]) + dpp<DPP_ROW_HMIRROR>(send);
  } else {
    const T send = b3 ? z[0] : z[1];
    r = (b3 ? z[1] : z[0]) + dpp<DPP_ROW_ROR8>(send);
    r += dpp<DPP_ROW_HMIRROR>(r);
  }
  r += dpp<DPP_QUAD_2301>(r);
  r += dpp<DPP_QUAD_1032>(r);
  return r;
}

// Reduce-scatter of 8 values over the eight 8-lane blocks of a wavefront
// (lanes that agree in bits 0..2): lane q returns the sum over its eight
// block partners of x[(q >> 3) & 7]. Partners by stage: ^32 and ^16
// (permlane swaps: the lower half keeps the first operand), then ^8
// (row_ror:8).
template <typename T>
__device__ __forceinline__ T blk_rs(const T (&x)[8], int q) {
  T y[4], z[2];
#pragma unroll
  for (int u = 0; u < 4; ++u) y[u] = swap_add<T>(x[u], x[u + 4], true);
#pragma unroll
  for (int u = 0; u < 2; ++u) z[u] = swap_add<T>(y[u], y[u + 2], false);
  const bool b3 = (q >> 3) & 1;
  const T send = b3 ? z[0] : z[1];
  const T keep = b3 ? z[1] : z[0];
  return keep + dpp<DPP_ROW_ROR8>(send);
}

// Sum over the eight lanes of a block of their x[p xor (lane mod 8)]: lane e
// holds in x[p] its partial for row p xor e of the block and gets the total
// of row e. Partners ^7 (row_half_mirror), ^2 and ^1 (quad perms) split the
// rows by bits 2, 1, 0; under the xor order the value a lane keeps and the
// one its partner sends sit in fixed registers, so no selects are needed.
__device__ __forceinline__ float xor_rs8(const float (&x)[8]) {
  float y[4], z[2];
#pragma unroll
  for (int p = 0; p < 4; ++p) y[p] = x[p] + dpp<DPP_ROW_HMIRROR>(x[7 - p]);
#pragma unroll
  for (int p = 0; p < 2; ++p) z[p] = y[p] + dpp<DPP_QUAD_2301>(y[p ^ 2]);
  return z[0] + dpp<DPP_QUAD_1032>(z[1]);
}

// Sum over the 16 lanes of a row of their x[p xor (lane mod 16) / 2]: lane
// c holds in x[p] its partial for row p xor (c >> 1) of an 8-row block and
// gets the total of row c >> 1 (lanes 2 r and 2 r + 1 both). Partners ^15
// (row_mirror), ^7 (row_half_mirror) and ^3 (quad perm) split the rows by
// bits 2, 1, 0 of (c >> 1), ^1 adds the two lanes of each row; under the xor
// order the kept and the sent values sit in fixed registers (no selects).
__device__ __forceinline__ float xor_rs16(const float (&x)[8]) {
  float y[4], z[2];
#pragma unroll
  for (int p = 0; p < 4; ++p) y[p] = x[p] + dpp<DPP_ROW_MIRROR>(x[p ^ 7]);
#pragma unroll
  for (int p = 0; p < 2; ++p) z[p] = y[p] + dpp<DPP_ROW_HMIRROR>(y[p ^ 3]);
  const float r = z[0] + dpp<DPP_QUAD_3210>(z[1]);
  return r + dpp<DPP_QUAD_1032>(r);
}

// x[p] = v of lane (lane xor p) within the lane's 8-lane block
__device__ __forceinline__ void xor_gather8(float v, float (&x)[8]) {
  x[0] = v;
  x[1] = dpp<DPP_QUAD_1032>(v);
  x[2] = dpp<DPP_QUAD_2301>(v);
  x[3] = dpp<DPP_QUAD_3210>(v);
  x[7] = dpp<DPP_ROW_HMIRROR>(v);
  x[4] = dpp<DPP_QUAD_3210>(x[7]);
  x[5] = dpp<DPP_QUAD_2301>(x[7]);
  x[6] = dpp<DPP_QUAD_1032>(x[7]);
}

// min or max of x over the rows of part 0 (all lanes return it)
template <bool MAX, int KC, int Q, typename T>
__device__ __forceinline__ double hr_minmax(double x, int slot, HRShared<KC, Q, T> &sh) {
  constexpr int WPH = HR<KC, Q>::WPH;
  x = MAX ? g_max<64>(x) : g_min<64>(x);
  const int t = threadIdx.x;
  __syncthreads();
  if (t < KC && (t & 63) == 0) sh.redd[WPH * slot + (t >> 6)] = x;
  __syncthreads();
  const double *r = sh.redd + WPH * slot;
  double y = r[0];
#pragma unroll
  for (int u = 1; u < WPH; ++u) y = MAX ? fmax(y, r[u]) : fmin(y, r[u]);
  return y;
}

// The matvec of a lane's G_S tile: rows 8 rb + (p xor f) (f = cb / 2) and
// the CW = KC / 16 columns CW cb .. CW cb + CW - 1, held as g[CW p + u],
// where u runs over the lane's columns in 4-column blocks rotated by
// cb / (16 / NB) (NB = CW / 4 blocks): then the 16 lanes of a DPP row read
// x = lvn by 16-byte reads from 16 distinct bank quads at every u, where
// unrotated the lanes cb, cb + 4, .. hit the same banks (16 CW words apart).
// Each row's CW products are summed in u order, the row partials of the 16
// lanes of a DPP row (one 8-row block) meet by xor_rs16, and the lane returns
// the sum of row 8 rb + f (lanes 2 f and 2 f + 1 both). pa: alpha's FP64 term
// of that row, the exact product v_r (row sum) with v_r = lvn_r scale as the
// row's owner forms it, on both lanes of the pair (wave_sum_pairs_f64 counts
// each row once; computing it on both lanes avoids a divergent branch).
template <int KC>
__device__ __forceinline__ float hr_tile_matvec(const float (&g)[KC / 2], const float *lvn, int cb,
                                                int rb, int f, float scale, double &pa) {
  constexpr int CW = KC / 16, NB = CW / 4;
  static_assert(CW % 4 == 0, "16-byte reads of the vector");
  const int rot = cb / (16 / NB);
  float x[CW];
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    const float4 t4 = reinterpret_cast<const float4 *>(lvn + CW * cb)[(u + rot) % NB];
    x[4 * u] = t4.x;
    x[4 * u + 1] = t4.y;
    x[4 * u + 2] = t4.z;
    x[4 * u + 3] = t4.w;
  }
  float y[8];
#pragma unroll
  for (int p = 0; p < 8; ++p) {
    float acc = 0.f;
#pragma unroll
    for (int c = 0; c < CW; ++c) acc = __builtin_fmaf(g[CW * p + c], x[c], acc);
    y[p] = acc;
  }
  const float ys = xor_rs16(y);
  // alpha: the exact product of the row's v and its sum, on both lanes of
  // the row's pair (the caller sums over the even lanes only)
  const float v = lvn[8 * rb + f] * scale;
  pa = (double)v * (double)ys;
  return ys;
}

// sum_c g[c] * L[c] over N entries, L a 16-byte aligned LDS vector read by
// broadcast; four accumulators by index mod 4
template <int N>
__device__ __forceinline__ float hr_rowdot(const float (&g)[N], const float *L) {
  static_assert(N % 16 == 0, "blocks of 16 entries");
  float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
  // four 16-byte reads (16 entries) issued together per block
#pragma unroll
  for (int b = 0; b < N / 16; ++b) {
    float4 t[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) t[u] = reinterpret_cast<const float4 *>(L)[4 * b + u];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int q = 4 * b + u;
      a0 = __builtin_fmaf(g[4 * q + 0], t[u].x, a0);
      a1 = __builtin_fmaf(g[4 * q + 1], t[u].y, a1);
      a2 = __builtin_fmaf(g[4 * q + 2], t[u].z, a2);
      a3 = __builtin_fmaf(g[4 * q + 3], t[u].w, a3);
    }
  }
  return (a0 + a1) + (a2 + a3);
}

// FP64 sum over the 64 lanes of a wavefront (every lane gets it). Fed with
// exact products of FP32 operands, it gives the extended-precision dot
// product that compensated FP32 summation (Dot2) emulates, in fewer
// instructions: FP64 is native here.
// the same sum when lanes 2l and 2l + 1 hold the same value: the total over
// the even lanes, on every lane
__device__ __forceinline__ double wave_sum_pairs_f64(double x) {
  x += dpp<DPP_QUAD_2301>(x);
  x += dpp<DPP_ROW_HMIRROR>(x);
  x += dpp<DPP_ROW_MIRROR>(x);
  x = swap_add<double>(x, x, false);
  x = swap_add<double>(x, x, true);
  return x;
}
__device__ __forceinline__ double wave_sum_f64(double x) {
  x += dpp<DPP_QUAD_1032>(x);
  return wave_sum_pairs_f64(x);
}

// The total over the wavefront on lane 63 only (other lanes: unspecified),
// for x equal on the lanes of each row of 16 (a row total): rows 1 and 3 add
// lane 15 of the row before (row_bcast:15), then rows 2 and 3 add lane 31
// (row_bcast:31): two 64-bit DPP moves and an add per step, against two
// copies, two permlane swaps and an add for the all-lane sum.
// rows outside ROWS receive an unspecified value (only lane 63 is used)
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_rows_f64(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)u, CTRL, ROWS, 0xf, false);
  const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), CTRL, ROWS, 0xf, false);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ double rows_to_lane63_f64(double x) {
  x += dpp_rows_f64<0x142, 0xa>(x);  // row_bcast:15 into rows 1 and 3
  x += dpp_rows_f64<0x143, 0xc>(x);  // row_bcast:31 into rows 2 and 3
  return x;
}
// g_sum_dpp<64>'s total on lane 63 only (other lanes: unspecified)
__device__ __forceinline__ float wave_total63_f32(float x) {
  x += dpp<DPP_QUAD_1032>(x);
  x += dpp<DPP_QUAD_2301>(x);
  x += dpp<DPP_ROW_HMIRROR>(x);
  x += dpp<DPP_ROW_MIRROR>(x);
  x += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), 0x142, 0xa, 0xf, false));
  x += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), 0x143, 0xc, 0xf, false));
  return x;
}
// wave_sum_pairs_f64's total on lane 63
__device__ __forceinline__ double wave_total63_pairs_f64(double x) {
  x += dpp<DPP_QUAD_2301>(x);
  x += dpp<DPP_ROW_HMIRROR>(x);
  x += dpp<DPP_ROW_MIRROR>(x);
  return rows_to_lane63_f64(x);
}
// wave_sum_f64's total on lane 63
__device__ __forceinline__ double wave_total63_f64(double x) {
  x += dpp<DPP_QUAD_1032>(x);
  return wave_total63_pairs_f64(x);
}

// beta^2 = w.w (exact squares summed in FP64, over the rows of part 0) and
// the next matvec input lvn = w, published together behind one barrier
template <int KC, int Q, typename T>
__device__ __forceinline__ void hr_publish(float w, HRShared<KC, Q, T> &sh) {
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  const int i = t & (KC - 1);
  const double sd = wave_total63_f64((double)w * (double)w);
  if (t < KC) {
    sh.lvn[i] = w;
    if ((t & 63) == 63) sh.rb[t >> 6] = sd;
  }
  __syncthreads();
}
// beta^2 published by hr_publish (valid until the next barrier that is
// followed by a new publication)
template <int KC, int Q, typename T>
__device__ __forceinline__ float hr_b2(HRShared<KC, Q, T> &sh) {
  double tot = sh.rb[0];
#pragma unroll
  for (int u = 1; u < HR<KC, Q>::WPH; ++u) tot += sh.rb[u];
  return (float)tot;
}
template <int KC, int Q, typename T>
__device__ __forceinline__ float hr_beta2(float w, HRShared<KC, Q, T> &sh) {
  hr_publish(w, sh);
  return hr_b2(sh);
}

// beta = sqrt(beta^2) and 1 / beta from one reciprocal square root (1 ulp):
// a short chain on the step's critical path, where sqrt and a correctly
// rounded division took about 25 dependent instructions. beta^2 below the
// normal range (where the instruction flushes its input and returns inf) or
// NaN gives beta = 0, which the breakdown test always catches.
__device__ __forceinline__ void hr_beta_scale(float b2, float &b, float &sc) {
  sc = __builtin_amdgcn_rsqf(b2);
  b = b2 >= 1.17549435e-38f ? b2 * sc : 0.f;
}

// true on every lane when p holds on any lane of the workgroup (a barrier,
// like __syncthreads_or, without the LDS block that one reserves):
// per-wavefront flags in ired
template <int KC, int Q, typename T>
__device__ __forceinline__ bool hr_any(bool p, HRShared<KC, Q, T> &sh) {
  const bool w = __ballot(p) != 0;
  if ((threadIdx.x & 63) == 0) sh.ired[threadIdx.x >> 6] = w ? 1 : 0;
  __syncthreads();
  int any = 0;
#pragma unroll
  for (int u = 0; u < HR<KC, Q>::NW; ++u) any |= sh.ired[u];
  return any != 0;
}

// One column-dot unit: LDS columns 4 cg .. 4 cg + 3 over the 16-row blocks
// [b0, b1) (b1 - b0 even), taken by the eight lanes g = 0..7 of a slot over
// the row pairs 16 b + 2 g, 16 b + 2 g + 1; the unit's first lane writes the
// four partials at cp[wo..wo+3].
template <int KC, int Q, typename T>
__device__ __forceinline__ void hr_dot_unit(const float *VL, HRShared<KC, Q, T> &sh, int g, int cg,
                                            int b0, int b1, int wo) {
  using C = HR<KC, Q>;
  const float *col = VL + 2 * g * C::S + 4 * cg;
  T a0 = 0, a1 = 0, a2 = 0, a3 = 0;
  for (int b = b0; b < b1; b += 2) {
    float4 v[4];
    T l[4];
#pragma unroll
    for (int u2 = 0; u2 < 2; ++u2) {
      const int r = 16 * (b + u2);
      v[2 * u2] = *reinterpret_cast<const float4 *>(col + r * C::S);
      v[2 * u2 + 1] = *reinterpret_cast<const float4 *>(col + (r + 1) * C::S);
      if constexpr (HRShared<KC, Q, T>::WD) {
        const double2 x = *reinterpret_cast<const double2 *>(sh.lwd + r + 2 * g);
        l[2 * u2] = x.x;
        l[2 * u2 + 1] = x.y;
      } else {
        // FP32 coefficients: w read in FP32
        const float2 x = *reinterpret_cast<const float2 *>(sh.lw + r + 2 * g);
        l[2 * u2] = (T)x.x;
        l[2 * u2 + 1] = (T)x.y;
      }
    }
#pragma unroll
    for (int u2 = 0; u2 < 4; ++u2) {
      a0 = fmaT((T)v[u2].x, l[u2], a0);
      a1 = fmaT((T)v[u2].y, l[u2], a1);
      a2 = fmaT((T)v[u2].z, l[u2], a2);
      a3 = fmaT((T)v[u2].w, l[u2], a3);
    }
  }
  if constexpr (sizeof(T) == 8) {
    // reduce-scatter over the eight lanes (a 64-bit DPP move is two moves,
    // so halving the values each level beats summing all four everywhere):
    // partner 7 - g splits the columns {0, 1} / {2, 3}, partner g xor 2
    // splits each pair, partner g xor 1 completes; lanes 2c and 2c + 1 end
    // with column c and the even one writes it
    const bool lo = g < 4, pb = (g & 2) == 0;
    T k0 = lo ? a0 : a2, k1 = lo ? a1 : a3;
    const T s0 = lo ? a2 : a0, s1 = lo ? a3 : a1;
    k0 += dpp<DPP_ROW_HMIRROR>(s0);
    k1 += dpp<DPP_ROW_HMIRROR>(s1);
    T kk = pb ? k0 : k1;
    kk += dpp<DPP_QUAD_2301>(pb ? k1 : k0);
    kk += dpp<DPP_QUAD_1032>(kk);
    if ((g & 1) == 0) sh.cp[wo + (g >> 1)] = kk;
  } else {
    a0 += dpp<DPP_QUAD_1032>(a0);
    a1 += dpp<DPP_QUAD_1032>(a1);
    a2 += dpp<DPP_QUAD_1032>(a2);
    a3 += dpp<DPP_QUAD_1032>(a3);
    a0 += dpp<DPP_QUAD_2301>(a0);
    a1 += dpp<DPP_QUAD_2301>(a1);
    a2 += dpp<DPP_QUAD_2301>(a2);
    a3 += dpp<DPP_QUAD_2301>(a3);
    a0 += dpp<DPP_ROW_HMIRROR>(a0);
    a1 += dpp<DPP_ROW_HMIRROR>(a1);
    a2 += dpp<DPP_ROW_HMIRROR>(a2);
    a3 += dpp<DPP_ROW_HMIRROR>(a3);
    if (g == 0) {
      T *cq = sh.cp + wo;
      cq[0] = a0;
      cq[1] = a1;
      cq[2] = a2;
      cq[3] = a3;
    }
  }
}

// One classical Gram-Schmidt pass of w (row i's entry, identical on all row
// parts) against basis columns 0..j; returns the new entry. With WANT_PRE,
// ||w||^2 before the pass is returned in pre (its partials ride along with
// the broadcast of w).
template <int KC, int Q, typename T, bool WANT_PRE>
__device__ __forceinline__ float hr_cgs(float w, int j, int k, const float *VL,
                                        const float (&vr)[HR<KC, Q>::RA],
                                        HRShared<KC, Q, T> &sh, float &pre) {
  using C = HR<KC, Q>;
  constexpr int NC = k2_chunk<T>();
  constexpr int SH = NC == 16 ? 2 : 3;
  // lane coordinates derived here rather than kept live across the step
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  const int i = t & (KC - 1), h = t >> C::LOGK, wv = (t >> 6) & (C::WPH - 1), q = t & 63;
  float ws = 0.f;
  if constexpr (WANT_PRE) ws = wave_total63_f32(w * w);  // on lane 63
  if (h == 0) {
    sh.lw[i] = w;
    if constexpr (HRShared<KC, Q, T>::WD) sh.lwd[i] = (double)w;
    if (WANT_PRE && q == 63) sh.rp[wv] = ws;
  }
  // (0) KC = 256, register columns L + Q m + h: eight-row blocks, each slot
  // summed over its eight rows in the lane and reduce-scattered over the
  // blocks of the wavefront (slots past the part's formed columns skipped,
  // unformed columns inside a slot hold zeros). The rows' w come from the
  // block's lanes by DPP, so they need no LDS. With FP64 coefficients they
  // run before w is published; with FP32 coefficients next to the LDS-column
  // dots, in the order that lets the two wavefronts of a SIMD (one of each
  // part) overlap the two (measured: FP32 983 -> 971 us per subset and CU at
  // k = 256, FP64 1,094 -> 1,108, so FP64 keeps the first placement).
  constexpr bool IL = C::OCT && sizeof(T) == 4;
  auto reg_dots = [&] {
    if constexpr (C::OCT) {
      const int hu = __builtin_amdgcn_readfirstlane(h);
      const int nmh = j >= C::L + hu ? (j - C::L - hu) / Q + 1 : 0;
      const int smax = (nmh + 7) >> 3;
      if (smax > 0) {
        float wf[8];
        xor_gather8(w, wf);
        T acc[8];
#pragma unroll
        for (int s = 0; s < 8; ++s) {
          T a = 0;
          if (s < C::RB && s < smax) {
#pragma unroll
            for (int p = 0; p < 8; ++p)
              a = fmaT((T)vr[8 * (s < C::RB ? s : 0) + p], (T)wf[p], a);
          }
          acc[s] = a;
        }
        const T r = blk_rs<T>(acc, q);
        const int m = (t & 7) + 8 * ((q >> 3) & 7);
        if (m < C::R) sh.rpart(h, wv)[m] = r;
      }
    }
  };
  if constexpr (C::OCT && !IL) reg_dots();
  __syncthreads();
  if constexpr (WANT_PRE) {
    float s = sh.rp[0];
#pragma unroll
    for (int u = 1; u < C::WPH; ++u) s += sh.rp[u];
    pre = s;
  }
  // (1) LDS columns 0..ncl-1, in groups of four adjacent columns. A unit
  // is a group over a part of its rows, taken by eight lanes of one slot
  // (slot = t / 8 mod 8) of a wavefront; wavefront w takes row part
  // w mod P of the eight consecutive groups 8 (w / P) + slot. At KC = 256,
  // P is the largest power of two <= 8 with ncg P <= 64 (a part holds at
  // least two 16-row blocks), so a step with few columns spreads its dots
  // over every SIMD and both wavefronts of each; the part is uniform per
  // wavefront and adjacent slots take adjacent groups. At KC = 256 with two
  // row parts (16 < ncg <= 32) the two wavefronts of a SIMD (w, w + 4) take
  // the same eight groups, the first over 5/8 of the rows and the second
  // over the rest, since the SIMD issues the first one first; past 32 groups
  // the first takes 3/4 of the rows of its eight of groups 0..31, and the
  // second the rest plus one quarter of the rows of the tail group 32 +
  // slot. At KC = 128 (three
  // subsets per CU) P = 1. A unit's eight lanes take its four columns over
  // the row pairs 16 b + 2 g, 16 b + 2 g + 1 (g = t % 8) of its blocks: one
  // 16-byte read of V per row feeds four FMAs, and one read of w serves both
  // rows of a pair. With the row stride S (S / 4 odd), the even rows of the
  // eight lanes fall in distinct even bank quads and the odd rows in odd
  // ones, so the two units of a quarter wavefront read disjoint banks. The
  // eight lanes' partial sums meet by DPP, and each unit's first lane writes
  // its four partials, part-major (part p of column c at p (256 / P) + c;
  // the four tail parts at 256 + TST p + c - 128), so the coefficient sums read
  // each part with consecutive lanes on consecutive words and add the parts
  // by a fixed tree. Rows >= k and columns >= ncl are zero.
  const int ncl = min(j + 1, C::L);
  const int ncg = (ncl + 3) >> 2;
  constexpr int SL = 8 * C::NW, PMAX = KC == 256 ? 8 : 1;
  const bool split_tail = C::NW == 8 && ncg > 32;  // wave-uniform
  int lp = 0;  // log2 P, wave-uniform
  while ((2 << lp) <= PMAX && ncg * (2 << lp) <= SL) ++lp;
  lp = __builtin_amdgcn_readfirstlane(lp);
  auto lds_dots = [&] {
    static_assert((C::S / 4) % 2 == 1, "bank layout of the dot reads");
    static_assert(PMAX <= C::NW, "a wavefront takes one row part");
    static_assert(KC != 256 || C::L <= 4 * 40, "tail groups fit the eight slots");
    const int wv0 = t >> 6, sl = (t >> 3) & 7, g = t & 7;
    // 16-row blocks rounded up to even (rows beyond k hold zeros, rows < KC
    // exist); a part's blocks [b0, b1), nbp = blocks per part, even
    const int nbe = (((k + 15) >> 4) + 1) & ~1;
    int cg, wo, b0, b1;  // group, partial offset of its column 4 cg, blocks
    if (!split_tail && !(KC == 256 && lp == 1)) {  // P parts, NW / P wavefronts each
      const int pp = wv0 & ((1 << lp) - 1);
      const int nbp = (((nbe + (1 << lp) - 1) >> lp) + 1) & ~1;
      cg = 8 * (wv0 >> lp) + sl;
      wo = pp * ((4 * SL) >> lp) + 4 * cg;
      b0 = pp * nbp;
      b1 = min(b0 + nbp, nbe);
    } else if (!split_tail) {
      // KC = 256, two row parts (16 < ncg <= 32): the two wavefronts of a
      // SIMD (w, w + 4) on the same eight groups, the one issued first over
      // 5/8 of the blocks (measured with FP64 coefficients 1,077 -> 1,073 us
      // per subset and CU at k = 256; 3/4: 1,083)
      const int pp = wv0 >> 2;
      const int nb0 = min(nbe, ((nbe * 5 + 8 - 1) / 8 + 1) & ~1);
      cg = 8 * (wv0 & 3) + sl;
      wo = pp * ((4 * SL) >> 1) + 4 * cg;
      b0 = pp ? nb0 : 0;
      b1 = pp ? nbe : nb0;
    } else {
      // past 32 groups: the same pairing, the first wavefront over 7/8 of
      // the blocks and the second over the rest plus the tail groups below
      // (3/4 measured 1,087 -> 1,079 us per subset and CU at k = 256 with
      // FP64 coefficients, 972 -> 964 with FP32, against halves on the pairs
      // w, w + 1 and the tail over all eight wavefronts; since the FP64 dot
      // changes 7/8 beats 3/4 by 0.4% in both forms, 5/8 and 8/8 lose)
      const int pp = wv0 >> 2;
      const int nb0 = min(nbe, ((nbe * 13 + 15) / 16 + 1) & ~1);
      cg = 8 * (wv0 & 3) + sl;
      wo = pp * 128 + 4 * cg;
      b0 = pp ? nb0 : 0;
      b1 = pp ? nbe : nb0;
    }
    b0 = __builtin_amdgcn_readfirstlane(b0);
    b1 = __builtin_amdgcn_readfirstlane(b1);
    if (cg < ncg) hr_dot_unit<KC, Q, T>(VL, sh, g, cg, b0, b1, wo);
    if (split_tail) {
      // tail group 32 + slot, row part wv0 - 4 of four, on the second
      // wavefront of each SIMD (it holds 1/8 of the main rows)
      constexpr int TP = 4;
      const int tw = wv0 - 4;
      if (tw >= 0 && tw < TP) {
        const int nbp = (((nbe + TP - 1) / TP) + 1) & ~1;
        int tb0 = tw * nbp, tb1 = min(tb0 + nbp, nbe);
        tb0 = __builtin_amdgcn_readfirstlane(tb0);
        tb1 = __builtin_amdgcn_readfirstlane(tb1);
        if (32 + sl < ncg)
          hr_dot_unit<KC, Q, T>(VL, sh, g, 32 + sl, tb0, tb1, 256 + HRShared<KC, Q, T>::TST * tw + 4 * sl);
      }
    }
  };
  // FP32 coefficients: part 0 (the first wavefront of each SIMD) starts
  // with its LDS-column dots, part 1 with its register-column dots
  if constexpr (IL) {
    if (__builtin_amdgcn_readfirstlane(h) == 0) {
      lds_dots();
      reg_dots();
    } else {
      reg_dots();
      lds_dots();
    }
  } else {
    lds_dots();
  }
  // (2) register columns L + Q m + h (m < nm), KC = 128: products
  // reduce-scattered per chunk, the chunks holding formed columns only
  // (uniform branch per chunk; unformed columns inside a chunk hold zeros)
  int nm = 0;
  if constexpr (!C::OCT && C::R > 0) {
    nm = (max(0, j + 1 - C::L) + Q - 1) / Q;
    if (nm > 0) {
#pragma unroll
      for (int c = 0; c < (C::R + NC - 1) / NC; ++c) {
        if (NC * c >= nm) break;
        T x[NC];
#pragma unroll
        for (int u = 0; u < NC; ++u) {
          const int m = NC * c + u;
          x[u] = (m < C::R) ? (T)vr[m < C::R ? m : 0] * (T)w : (T)0;
        }
        const T r = wave_rs<NC, T>(x, q);
        if ((q & ((1 << SH) - 1)) == 0) sh.rpart(h, wv)[NC * c + (q >> SH)] = r;
      }
    }
  }
  __syncthreads();
  // (3) the coefficients: LDS column t on lane t < KC of part 0 (zero
  // beyond j), the sum of its group's row parts in part order; register
  // column L + r on lane r of part 1, summed over the owning part's
  // wavefronts in wavefront order (the two kinds on different wavefronts,
  // so neither waits behind the other's branch)
  if (h == 0) {
    T hc = 0;
    if (i < ncl) {
      // the column's parts, by a pairwise tree: P of them (a wave-uniform
      // count) at i + p 4 SL / P; past 32 groups two at i, i + 128, or the
      // four tail parts at 128 + i + TST p
      if (split_tail) {
        if (i < 128) {
          hc = sh.cp[i] + sh.cp[i + 128];
        } else {
          constexpr int TS = HRShared<KC, Q, T>::TST;
          const T *cq = sh.cp + 128 + i;
          hc = (cq[0] + cq[TS]) + (cq[2 * TS] + cq[3 * TS]);
        }
      } else if (lp == 0) {
        hc = sh.cp[i];
      } else if (lp == 1) {
        hc = sh.cp[i] + sh.cp[i + 2 * SL];
      } else if (lp == 2) {
        const T *cq = sh.cp + i;
        hc = (cq[0] + cq[SL]) + (cq[2 * SL] + cq[3 * SL]);
      } else {
        const T *cq = sh.cp + i;
        const int st = SL / 2;
        hc = ((cq[0] + cq[st]) + (cq[2 * st] + cq[3 * st])) +
             ((cq[4 * st] + cq[5 * st]) + (cq[6 * st] + cq[7 * st]));
      }
    }
    sh.lh[i] = (float)hc;
  } else if constexpr (C::R > 0) {
    if (h == 1 && i < KC - C::L && C::L + i <= j) {
      // column L + i is the m-th register column of part hh: at KC = 256
      // its partials sit at m, its coefficient goes to [m mod 8][m / 8];
      // else in slot nm_hh - 1 - m with nm_hh = (j - L - hh) / Q + 1 formed
      // in that part
      const int m = i / Q, hh = i % Q;
      const int sl = C::OCT ? m : (j - C::L - hh) / Q - m;
      T s = sh.rpart(hh, 0)[sl];
#pragma unroll
      for (int u = 1; u < C::WPH; ++u) s += sh.rpart(hh, u)[sl];
      sh.lhr[hh][C::OCT ? 8 * (m & 7) + (m >> 3) : sl] = (float)s;
    }
  }
  __syncthreads();
  // (4) update partials (FP32): part h takes LDS columns [h cs, (h + 1) cs)
  // below ncl, cs = ceil(ncl / 4Q) * 4; register columns go to their owner
  float u0 = 0.f, u1 = 0.f, u2 = 0.f, u3 = 0.f;
  {
    const int cs = ((ncl + 4 * Q - 1) / (4 * Q)) * 4;
    const int c0 = h * cs;
    const int cn = __builtin_amdgcn_readfirstlane(max(0, min(cs, ncl - c0)));  // per part
    const float *row = VL + i * C::S + c0;
    const float *hl = sh.lh + c0;
    // 16 (FP32) or 8 columns per iteration, then at most one group of 4; reads stay in
    // [c0, c0 + cn rounded up to 4), inside this part's share of the row and
    // below L (columns past j hold zeros and zero coefficients)
    // the coefficients are read 16 bytes at a time too (c0 is a multiple of
    // 4; the part is uniform per wavefront, so these reads are broadcasts)
    const int cn4 = (cn + 3) & ~3;
    int c = 0;
    // 16 columns per round trip
    for (; c + 16 <= cn4; c += 16) {
      float4 ra[4], ha[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        ra[q] = *reinterpret_cast<const float4 *>(row + c + 4 * q);
        ha[q] = *reinterpret_cast<const float4 *>(hl + c + 4 * q);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        u0 = __builtin_fmaf(ra[q].x, ha[q].x, u0);
        u1 = __builtin_fmaf(ra[q].y, ha[q].y, u1);
        u2 = __builtin_fmaf(ra[q].z, ha[q].z, u2);
        u3 = __builtin_fmaf(ra[q].w, ha[q].w, u3);
      }
    }
    for (; c + 8 <= cn4; c += 8) {
      const float4 ra = *reinterpret_cast<const float4 *>(row + c);
      const float4 rb = *reinterpret_cast<const float4 *>(row + c + 4);
      const float4 ha = *reinterpret_cast<const float4 *>(hl + c);
      const float4 hb = *reinterpret_cast<const float4 *>(hl + c + 4);
      u0 = __builtin_fmaf(ra.x, ha.x, u0);
      u1 = __builtin_fmaf(ra.y, ha.y, u1);
      u2 = __builtin_fmaf(ra.z, ha.z, u2);
      u3 = __builtin_fmaf(ra.w, ha.w, u3);
      u0 = __builtin_fmaf(rb.x, hb.x, u0);
      u1 = __builtin_fmaf(rb.y, hb.y, u1);
      u2 = __builtin_fmaf(rb.z, hb.z, u2);
      u3 = __builtin_fmaf(rb.w, hb.w, u3);
    }
    if (c < cn4) {
      const float4 ra = *reinterpret_cast<const float4 *>(row + c);
      const float4 ha = *reinterpret_cast<const float4 *>(hl + c);
      u0 = __builtin_fmaf(ra.x, ha.x, u0);
      u1 = __builtin_fmaf(ra.y, ha.y, u1);
      u2 = __builtin_fmaf(ra.z, ha.z, u2);
      u3 = __builtin_fmaf(ra.w, ha.w, u3);
    }
  }
  if constexpr (C::OCT) {
    // the lane's slots (eight rows each) times their coefficients, read as
    // two 16-byte words per lane (slots past the formed columns hold zeros,
    // as do their coefficients); the block's lanes then sum each row's
    // partials
    const int hu = __builtin_amdgcn_readfirstlane(h);
    const int nmh = j >= C::L + hu ? (j - C::L - hu) / Q + 1 : 0;
    const int smax = (nmh + 7) >> 3;
    if (smax > 0) {
      const float4 *lr = reinterpret_cast<const float4 *>(sh.lhr[hu] + 8 * (t & 7));
      const float4 c0 = lr[0], c1 = lr[1];
      const float cf[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
      float x[8];
#pragma unroll
      for (int p = 0; p < 8; ++p) x[p] = 0.f;
#pragma unroll
      for (int s = 0; s < C::RB; ++s) {
        if (s < smax) {
#pragma unroll
          for (int p = 0; p < 8; ++p) x[p] = __builtin_fmaf(vr[8 * s + p], cf[s], x[p]);
        }
      }
      u0 += xor_rs8(x);
    }
  } else if constexpr (C::R > 0) {
    // slots 0..nm_h-1 of this part, coefficients read 4 at a time in slot
    // order (a broadcast: the part is uniform per wavefront), the four reads
    // of a block of 16 slots issued together (slots past nm_h hold zeros)
    constexpr int NG = (C::R + 3) / 4;  // groups of 4 slots
    const int hu = __builtin_amdgcn_readfirstlane(h);
    const int nmh = j >= C::L + hu ? (j - C::L - hu) / Q + 1 : 0;
    const float4 *lr = reinterpret_cast<const float4 *>(sh.lhr[hu]);
#pragma unroll
    for (int g0 = 0; g0 < NG; g0 += 4) {
      if (4 * g0 < nmh) {  // uniform
        float4 cf[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (g0 + q < NG) cf[q] = lr[g0 + q];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int m0 = 4 * (g0 + q);
          if (g0 + q < NG && m0 < nmh) {
            u0 = __builtin_fmaf(vr[m0], cf[q].x, u0);
            if (m0 + 1 < C::R) u1 = __builtin_fmaf(vr[m0 + 1], cf[q].y, u1);
            if (m0 + 2 < C::R) u2 = __builtin_fmaf(vr[m0 + 2], cf[q].z, u2);
            if (m0 + 3 < C::R) u3 = __builtin_fmaf(vr[m0 + 3], cf[q].w, u3);
          }
        }
      }
    }
  }
  sh.part[h][i] = (u0 + u1) + (u2 + u3);
  __syncthreads();
  float tot = sh.part[0][i];
#pragma unroll
  for (int hh = 1; hh < Q; ++hh) tot += sh.part[hh][i];
  return w - tot;
}

template <int KC, int Q, int MODE, typename T>
__device__ __forceinline__ void hr_subset(const SolveArgs &a, float *VL, HRShared<KC, Q, T> &sh,
                                          int64_t s) {
  using C = HR<KC, Q>;
  // lane coordinates kept opaque per subset, so that nothing derived from
  // them is hoisted out of the persistent loop and held live across it
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  const int i = t & (KC - 1), h = t >> C::LOGK, q = t & 63;
  int k;
  int64_t off;
  if (a.offsets) {
    off = a.offsets[s];
    k = (int)(a.offsets[s + 1] - off);
  } else {
    k = a.kuni;
    off = s * (int64_t)k;
  }
  const bool badk = k < 1 || k > KC || k > a.kcap;
  uint32_t st = badk ? ST_BADK : 0u;
  int my = 0;
  if (!st && i < k) {
    my = a.idx[off + i];
    if (my < 0 || (int64_t)my >= a.n) st = ST_BADINDEX;
  }
  if (hr_any(st != 0, sh)) st |= badk ? ST_BADK : ST_BADINDEX;
  int *lidx = reinterpret_cast<int *>(sh.lw);
  if (h == 0) lidx[i] = my;
  __syncthreads();

  // ---- (a) gather: the upper triangle of G_S (r <= c, KC (KC + 1) / 2
  // words) is fetched once, each word G[min(idx_r, idx_c)][max(..)] by one
  // lane: lane t takes the entries t + NT m of the triangle folded into a
  // KC/2 x (KC + 1) rectangle (row pr holds row pr of the triangle, then row
  // KC - 1 - pr), so every lane has about KC (KC + 1) / (2 NT) independent
  // loads in flight and each word of G is requested once. The words go to
  // LDS in the basis area (free until the recurrence starts), rectangle row
  // stride TS = KC + 3, and each lane then reads its tile from there (rows
  // 8 rb + (p xor f), columns CW cb + c: lane t has cb = t mod 16, rb = t /
  // 16, f = cb / 2, CW = KC / 16, the 4-column blocks of a lane's columns
  // rotated by cb / (16 / NB), NB = CW / 4, as hr_tile_matvec reads them; a
  // DPP row of 16 lanes holds one 8-row block): G_S[r][c] is triangle entry
  // (min(r, c), max(r, c)). ----
  float g[C::HW];
  bool fin = true;
  {
    constexpr int TS = KC + 3, NP = (KC / 2) * (KC + 1), MI = (NP + C::NT - 1) / C::NT;
    static_assert((KC / 2) * TS <= KC * C::S, "folded triangle fits the basis area");
    float x[MI];
#pragma unroll
    for (int m = 0; m < MI; ++m) {
      const int p = t + C::NT * m;
      const int pr = p / (KC + 1), qc = p - pr * (KC + 1);
      const bool first = qc < KC - pr;
      const int r = first ? pr : KC - 1 - pr, c = first ? pr + qc : qc - 1;
      float y = 0.f;
      if (st == 0 && p < NP && c < k) {  // r <= c
        const int ir = lidx[r], ic = lidx[c];
        const int64_t lo = min(ir, ic), hi = max(ir, ic);
        y = a.G[lo * a.ld + hi];
      }
      x[m] = y;
    }
#pragma unroll
    for (int m = 0; m < MI; ++m) {
      const int p = t + C::NT * m;
      if (p < NP) {
        const int pr = p / (KC + 1), qc = p - pr * (KC + 1);
        VL[pr * TS + qc] = x[m];
      }
    }
    __syncthreads();
    constexpr int CW = KC / 16, NB = CW / 4;
    static_assert(8 * CW == C::HW && C::NT == 16 * (KC / 8), "one 8-row block per DPP row");
    const int cb = t & 15, rb = t >> 4, f = cb >> 1, rot = cb / (16 / NB);
#pragma unroll
    for (int p = 0; p < 8; ++p) {
#pragma unroll
      for (int u = 0; u < CW; ++u) {
        const int r = 8 * rb + (p ^ f), c = CW * cb + 4 * (((u >> 2) + rot) % NB) + (u & 3);
        float y = 0.f;
        if (st == 0 && r < k && c < k) {
          const int r0 = min(r, c), c0 = max(r, c);
          const bool first = r0 < KC / 2;
          const int pr = first ? r0 : KC - 1 - r0;
          const int qc = first ? c0 - r0 : c0 + 1;
          y = VL[pr * TS + qc];
        }
        g[CW * p + u] = y;
        fin = fin && isfinite(y);
      }
    }
  }
  if (hr_any(!fin, sh)) st |= ST_NONFINITE;

  if constexpr (MODE == MODE_DUMP) {
    if (st == 0) {
      constexpr int CW = KC / 16, NB = CW / 4;
      const int cb = t & 15, rb = t >> 4, f = cb >> 1, rot = cb / (16 / NB);
      float *o = a.gs_out + s * (int64_t)k * k;
#pragma unroll
      for (int p = 0; p < 8; ++p) {
        const int r = 8 * rb + (p ^ f);
#pragma unroll
        for (int u = 0; u < CW; ++u) {
          const int c = CW * cb + 4 * (((u >> 2) + rot) % NB) + (u & 3);
          if (r < k && c < k) o[(int64_t)r * k + c] = g[CW * p + u];
        }
      }
    }
    return;
  } else if constexpr (MODE == MODE_CHECKSUM) {
    float acc = 0.f;
#pragma unroll
    for (int c = 0; c < C::HW; ++c) acc += g[c];
    acc = g_sum<64>(acc);
    if (q == 0) sh.red[t >> 6] = acc;
    __syncthreads();
    if (t == 0) {
      float sum = 0.f;
      for (int w = 0; w < C::NW; ++w) sum += sh.red[w];
      a.gs_out[s] = st ? NAN : sum;
    }
    return;
  } else {
    constexpr uint32_t FATAL = ST_BADK | ST_BADINDEX | ST_NONFINITE;
    if (!(st & FATAL)) {
      // ---- (b) Lanczos recurrence, ParaLanczos.cu:1285-1374 with the build's
      // substitutions (ones/sqrt(k) start, full reorthogonalisation, restart).
      // Each step: matvec of the unnormalised vector published in lvn, scaled
      // by 1/beta, with alpha's FP64 partials exchanged alongside; three-term
      // update; Gram-Schmidt pass (||w||^2 rides along); beta^2 published
      // with the next vector. Six barriers per step. ----
      float vr[C::RA];
#pragma unroll
      for (int m = 0; m < C::RA; ++m) vr[m] = 0.f;
      for (int e = t; e < KC * C::S; e += C::NT) VL[e] = 0.f;
      // register-column coefficients: the update reads whole groups of 4
      // slots, so slots not yet formed must hold zeros, not stale LDS
      if constexpr (C::R > 0)
        for (int e = t; e < (int)(sizeof(sh.lhr) / sizeof(float)); e += C::NT)
          (&sh.lhr[0][0])[e] = 0.f;
      float scale = 1.0f / sqrtf((float)k);
      float vi = (i < k) ? scale : 0.f;
      __syncthreads();
      if (h == 0) {
        VL[i * C::S] = vi;
        sh.lvn[i] = (i < k) ? 1.0f : 0.f;
      }
      __syncthreads();
      float vp = 0.f, bj = 0.f, amax = 0.f, bmax = 0.f, my_a = 0.f, my_b = 0.f;
      for (int j = 0; j < k; ++j) {
        // lane coordinates derived here rather than kept live across the step
        int tl = threadIdx.x;
        asm volatile("" : "+v"(tl));
        const int il = tl & (KC - 1), hl = tl >> C::LOGK;
        // matvec of the lane's tile, row sums through LDS; alpha's terms
        // (exact products v_r times the row sum, once per row, summed in FP64
        // over the wavefront and then the wavefronts) exchanged alongside
        const int cbl = tl & 15, rbl = tl >> 4, fl = cbl >> 1;
        double pl;
        const float mv = hr_tile_matvec<KC>(g, sh.lvn, cbl, rbl, fl, scale, pl);
        const double pa = wave_total63_pairs_f64(pl);
        if ((cbl & 1) == 0) sh.pw[0][8 * rbl + fl] = mv;
        if ((tl & 63) == 63) sh.ra[tl >> 6] = pa;
        __syncthreads();
        float w = sh.pw[0][il] * scale;
        double asum = sh.ra[0];
#pragma unroll
        for (int w2 = 1; w2 < C::NW; ++w2) asum += sh.ra[w2];
        const float al = (float)(asum * (double)scale);
        my_a = (i == j) ? al : my_a;
        amax = fmaxf(amax, fabsf(al));
        if (j == k - 1) break;
        w = __builtin_fmaf(-al, vi, __builtin_fmaf(-bj, vp, w));
        float pre = 0.f, unused;
        w = hr_cgs<KC, Q, T, true>(w, j, k, VL, vr, sh, pre);
        float b2 = hr_beta2(w, sh);
        float sc, b;
        hr_beta_scale(b2, b, sc);
        bool brk = false;
        if (b2 < 0.5f * pre || !(b > TAU * (amax + bmax))) {  // uniform, rare
          if (b2 < 0.5f * pre) {  // DGKS: w was mostly in span(V); orthogonalise again
            w = hr_cgs<KC, Q, T, false>(w, j, k, VL, vr, sh, unused);
            b2 = hr_beta2(w, sh);
            hr_beta_scale(b2, b, sc);
          }
          if (!(b > TAU * (amax + bmax))) {
            // restart: e_i for the first i of largest residual 1 - ||V[i][:]||^2,
            // orthogonalised twice against the basis; beta_{j+1} = 0
            brk = true;
            st |= ST_BREAKDOWN;
            float rs = 0.f;
            if (h == 0)
              for (int l = 0; l < C::L; ++l)
                rs = __builtin_fmaf(VL[i * C::S + l], VL[i * C::S + l], rs);
            if constexpr (C::OCT) {
              // squares per row of the block, summed over the block's lanes
              float x[8];
#pragma unroll
              for (int p = 0; p < 8; ++p) x[p] = 0.f;
#pragma unroll
              for (int m = 0; m < C::RA; ++m) x[m & 7] = __builtin_fmaf(vr[m], vr[m], x[m & 7]);
              rs += xor_rs8(x);
            } else if constexpr (C::R > 0) {
#pragma unroll
              for (int m = 0; m < C::R; ++m) rs = __builtin_fmaf(vr[m], vr[m], rs);
            }
            __syncthreads();
            sh.pw[h][i] = rs;
            __syncthreads();
            float rsum = sh.pw[0][i];
#pragma unroll
            for (int hh = 1; hh < Q; ++hh) rsum += sh.pw[hh][i];
            const float res = 1.f - rsum;
            const int wb = restart_lane<64>(res, i, k);
            const float wvv = __shfl(i < k ? res : -1.f, wb & 63, 64);
            if (t < KC && q == 0) {
              sh.red[t >> 6] = wvv;
              sh.ired[t >> 6] = wb;
            }
            __syncthreads();
            int bi = sh.ired[0];
            float bv = sh.red[0];
            for (int w2 = 1; w2 < C::WPH; ++w2)
              if (sh.red[w2] > bv) {
                bv = sh.red[w2];
                bi = sh.ired[w2];
              }
            w = (i == bi) ? 1.f : 0.f;
            w = hr_cgs<KC, Q, T, false>(w, j, k, VL, vr, sh, unused);
            w = hr_cgs<KC, Q, T, false>(w, j, k, VL, vr, sh, unused);
            sc = 1.0f / sqrtf(hr_beta2(w, sh));
          }
        }
        vp = vi;
        bj = brk ? 0.f : b;
        bmax = brk ? bmax : fmaxf(bmax, b);
        scale = sc;
        vi = w * sc;
        // basis column j + 1: LDS, or at KC = 256 slot m / 8 of lane m mod 8
        // of each 8-row block of its part (the m-th column there), or else
        // slot 0 of the register shift of its part (the part is uniform per
        // wavefront)
        const int c = j + 1;
        if (c < C::L && h == 0) VL[i * C::S + c] = vi;
        if constexpr (C::OCT) {
          const int hu = __builtin_amdgcn_readfirstlane(h);
          const int cc = c - C::L;
          if (cc >= 0 && cc % Q == hu) {
            const int m = cc / Q, sm = m >> 3;
            float x[8];
            xor_gather8(vi, x);
            const bool mine = (t & 7) == (m & 7);
#pragma unroll
            for (int s = 0; s < C::RB; ++s) {
              if (s == sm) {
#pragma unroll
                for (int p = 0; p < 8; ++p) vr[8 * s + p] = mine ? x[p] : vr[8 * s + p];
              }
            }
          }
        } else if constexpr (C::R > 0) {
          const int hu = __builtin_amdgcn_readfirstlane(h);
          const int cc = c - C::L;
          if (cc >= 0 && cc % Q == hu) {
#pragma unroll
            for (int r = C::R - 1; r > 0; --r) vr[r] = vr[r - 1];
            vr[0] = vi;
          }
        }
        my_b = (i == j + 1) ? bj : my_b;
      }
      const bool finite = (i >= k) || (isfinite(my_a) && isfinite(my_b));
      if (hr_any(!finite, sh)) st |= ST_NONFINITE;
      if (!(st & FATAL)) {
        float2 *tri = a.tri + s * (int64_t)a.tstride;
        sh.pw[0][i] = my_b;  // every part writes the same value
        __syncthreads();
        const double al = my_a, b = my_b, bn = (i + 1 < KC) ? (double)sh.pw[0][i + 1] : 0.0;
        double glo = INFINITY, ghi = -INFINITY, amin = INFINITY, amx = -INFINITY;
        if (i < k) {
          if (h == 0) tri[i] = make_float2(my_a, i > 0 ? my_b : 0.f);
          const double r = (i > 0 ? fabs(b) : 0.0) + (i + 1 < k ? fabs(bn) : 0.0);
          glo = al - r;
          ghi = al + r;
          amin = amx = al;
        }
        glo = hr_minmax<false>(glo, 0, sh);
        ghi = hr_minmax<true>(ghi, 1, sh);
        amin = hr_minmax<false>(amin, 2, sh);
        amx = hr_minmax<true>(amx, 3, sh);
        if (t == 0) a.bounds[s] = make_double4(glo, ghi, amin, amx);
      }
    }
    if (t == 0) a.status[s] = st;
  }
}

template <int KC, int Q, int MODE, typename T>
__global__ __launch_bounds__(Q * KC, KC == 128 ? 3 : 1) void solve_rowpart_kernel(SolveArgs a) {
  __shared__ __attribute__((aligned(16))) HRLds<KC, Q, T> lds;
  int64_t cnt = a.nslots;
  const int32_t *ord = a.order;
  if (a.bins) {
    cnt = a.bins[BIN_COUNT + a.bin];
    ord = a.order + a.bins[BIN_START + a.bin];
  }
  for (int64_t slot = blockIdx.x; slot < cnt; slot += gridDim.x) {
    hr_subset<KC, Q, MODE, T>(a, lds.VL, lds.sh, ord ? (int64_t)ord[slot] : slot);
    __syncthreads();
  }
}

using SolveFn = void (*)(SolveArgs);

template <int GROUP, int PAD>
struct VariantFns {
  static constexpr SolveFn fns[3] = {solve_kernel<GROUP, PAD, 0>, solve_kernel<GROUP, PAD, 1>,
                                     solve_kernel<GROUP, PAD, 2>};
};

template <int KC, int Q>
static SolveFn pick_rowpart(int mode, bool f64) {
  static constexpr SolveFn f32[3] = {solve_rowpart_kernel<KC, Q, 0, float>,
                                     solve_rowpart_kernel<KC, Q, 1, float>,
                                     solve_rowpart_kernel<KC, Q, 2, float>};
  return (f64 && mode == MODE_TRIDIAG) ? solve_rowpart_kernel<KC, Q, 0, double> : f32[mode];
}

// group = lanes per subset: 8..64 (lane groups), 128 (one row per lane),
// 256 and 512 (half rows: capacities 128 and 256)
static SolveFn pick(int group, int pad, int mode, bool f64) {
  if (group >= 256 && pad != 4) return nullptr;
  if (group == 512) return pick_rowpart<256, 2>(mode, f64);
  if (group == 256) return pick_rowpart<128, 2>(mode, f64);
  if (group == WIDE) {
    static constexpr SolveFn w1[3] = {solve_wide_kernel<1, 0>, solve_wide_kernel<1, 1>,
                                      solve_wide_kernel<1, 2>};
    static constexpr SolveFn w4[3] = {solve_wide_kernel<4, 0>, solve_wide_kernel<4, 1>,
                                      solve_wide_kernel<4, 2>};
    if (f64 && mode == MODE_TRIDIAG) {
      if (pad == 1) return solve_wide_kernel<1, 0, double>;
      if (pad == 4) return solve_wide_kernel<4, 0, double>;
    }
    if (pad == 1) return w1[mode];
    if (pad == 4) return w4[mode];
    return nullptr;
  }
#define PL_CASE(G, P) \
  if (group == G && pad == P) return VariantFns<G, P>::fns[mode];
  PL_CASE(8, 0) PL_CASE(8, 1) PL_CASE(8, 4)
  PL_CASE(16, 0) PL_CASE(16, 1) PL_CASE(16, 4)
  PL_CASE(32, 0) PL_CASE(32, 1) PL_CASE(32, 4)
  PL_CASE(64, 0) PL_CASE(64, 1) PL_CASE(64, 4)
#undef PL_CASE
  return nullptr;
}

bool solve_variant_exists(int group, int pad) { return pick(group, pad, 0, false) != nullptr; }

int variant_capacity(int group) { return group >= 256 ? group / 2 : group; }

// workgroups of a variant that fit on the device at once (cached per variant)
static int resident_blocks(SolveFn fn, int threads) {
  static std::mutex mu;
  static std::map<SolveFn, int> cache;
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find(fn);
  if (it != cache.end()) return it->second;
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)fn, threads, 0) !=
          hipSuccess ||
      per_cu < 1)
    return 0;
  const int r = cus * per_cu;
  cache[fn] = r;
  return r;
}

hipError_t launch_solve(int group, int pad, int mode, const SolveArgs &a, hipStream_t s,
                        int64_t max_slots, bool f64) {
  if (mode < 0 || mode > 2) return hipErrorInvalidValue;
  SolveFn fn = pick(group, pad, mode, f64);
  if (!fn) return hipErrorInvalidValue;
  const int threads = group >= WIDE ? group : BLOCK;
  const int spb = group >= WIDE ? 1 : BLOCK / group;
  int64_t blocks;
  if (a.bins) {
    // persistent: as many workgroups as are resident, never more than the
    // bin can use
    const int res = resident_blocks(fn, threads);
    if (res <= 0) return hipErrorInvalidConfiguration;
    blocks = std::min<int64_t>(res, (max_slots + spb - 1) / spb);
  } else {
    blocks = (a.nslots + spb - 1) / spb;
  }
  if (blocks <= 0) return hipSuccess;
  if (blocks > 0x7fffffffLL) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL(fn, dim3((unsigned)blocks), dim3(threads), 0, s, a);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// PL_FORM_PLAIN_WAVE: one 64-lane wavefront per subset at every k <= 256.
// Beyond k = 64, G_S and the Lanczos basis do not fit a wavefront's share of
// the registers and LDS, so they live in a global scratch slot of the
// resident wavefront (2 x 256 x 256 FP32): column c of G_S (equal to row c)
// and basis vector c are each a run of 256 floats, read coalesced (lane l
// takes entries l + 64 m, m < 4) and streamed every step. FP32 matvec and
// Gram-Schmidt (coefficients by wavefront reduce-scatters); alpha and beta^2 as FP64
// sums of exact products; one classical Gram-Schmidt pass, a second one by
// the DGKS test; the breakdown rule of the other kernels (restart with the
// e_i of largest residual, orthogonalised twice). T and the bounds go to
// the workspace like every stage-1 kernel's.
// ---------------------------------------------------------------------------
constexpr int WV_K = 256, WV_M = WV_K / 64;

// w -= V[:, 0..j] (V[:, 0..j]^T w), one pass; coefficients through lh. The
// lanes' partial dots of 16 columns at a time are summed by one wavefront
// reduce-scatter (4 lanes end with each column's sum).
__device__ __forceinline__ void wv_cgs(float (&w)[WV_M], int j, const float *Vs, float *lh) {
  const int l = threadIdx.x;
  for (int c0 = 0; c0 <= j; c0 += 16) {
    float x[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const float *vc = Vs + (size_t)(c0 + u) * WV_K + l;
      float p = 0.f;
      if (c0 + u <= j) {
#pragma unroll
        for (int m = 0; m < WV_M; ++m) p = __builtin_fmaf(vc[64 * m], w[m], p);
      }
      x[u] = p;
    }
    const float r = wave_rs<16, float>(x, l);
    if ((l & 3) == 0) lh[c0 + (l >> 2)] = r;
  }
  __syncthreads();
  for (int c = 0; c <= j; ++c) {
    const float *vc = Vs + (size_t)c * WV_K + l;
    const float hc = lh[c];
#pragma unroll
    for (int m = 0; m < WV_M; ++m) w[m] = __builtin_fmaf(-vc[64 * m], hc, w[m]);
  }
  __syncthreads();
}

__device__ __forceinline__ double wv_sum_sq(const float (&w)[WV_M]) {
  double x = 0.0;
#pragma unroll
  for (int m = 0; m < WV_M; ++m) x += (double)w[m] * (double)w[m];
  return wave_sum_f64(x);
}

__global__ __launch_bounds__(64) void solve_wave_kernel(SolveArgs a, float *scratch) {
  __shared__ float lv[WV_K];   // matvec input
  __shared__ float lh[WV_K];   // Gram-Schmidt coefficients
  __shared__ int lidx[WV_K];   // the subset's indices
  __shared__ float la[WV_K], lb[WV_K];
  float *Gs = scratch + (size_t)blockIdx.x * 2 * WV_K * WV_K;
  float *Vs = Gs + (size_t)WV_K * WV_K;
  const int l = threadIdx.x;
  int64_t cnt = a.nslots;
  const int32_t *ord = a.order;
  if (a.bins) {
    cnt = a.bins[BIN_COUNT + a.bin];
    ord = a.order + a.bins[BIN_START + a.bin];
  }
  for (int64_t slot = blockIdx.x; slot < cnt; slot += gridDim.x) {
    const int64_t s = ord ? (int64_t)ord[slot] : slot;
    int k;
    int64_t off;
    if (a.offsets) {
      off = a.offsets[s];
      k = (int)(a.offsets[s + 1] - off);
    } else {
      k = a.kuni;
      off = s * (int64_t)k;
    }
    const bool badk = k < 1 || k > WV_K || k > a.kcap;
    uint32_t st = badk ? ST_BADK : 0u;
    bool bad = false;
#pragma unroll
    for (int m = 0; m < WV_M; ++m) {
      const int r = l + 64 * m;
      int x = 0;
      if (!badk && r < k) {
        x = a.idx[off + r];
        bad = bad || x < 0 || (int64_t)x >= a.n;
      }
      lidx[r] = x;
    }
    if (__ballot(bad) != 0) st |= ST_BADINDEX;
    __syncthreads();
    // ---- (a) gather: G_S column c (= row c) into scratch row c, zeros in
    // rows k..255 (the matvec reads whole rows) ----
    bool fin = true;
    if (!st) {
      for (int c = 0; c < k; ++c) {
        const int ic = lidx[c];
#pragma unroll
        for (int m = 0; m < WV_M; ++m) {
          const int r = l + 64 * m;
          float x = 0.f;
          if (r < k) {
            const int ir = lidx[r];
            const int64_t lo = min(ic, ir), hi = max(ic, ir);
            x = a.G[lo * a.ld + hi];
            fin = fin && isfinite(x);
          }
          Gs[(size_t)c * WV_K + r] = x;
        }
      }
    }
    if (__ballot(!fin) != 0) st |= ST_NONFINITE;
    __syncthreads();
    constexpr uint32_t FATAL = ST_BADK | ST_BADINDEX | ST_NONFINITE;
    if (!(st & FATAL)) {
      // ---- (b) Lanczos, full reorthogonalisation ----
      float v[WV_M], vp[WV_M], w[WV_M];
      const float sc0 = 1.0f / sqrtf((float)k);
#pragma unroll
      for (int m = 0; m < WV_M; ++m) {
        const int r = l + 64 * m;
        v[m] = r < k ? sc0 : 0.f;
        vp[m] = 0.f;
        Vs[r] = v[m];
        lv[r] = v[m];
      }
      __syncthreads();
      float bj = 0.f, amax = 0.f, bmax = 0.f;
      for (int j = 0; j < k; ++j) {
        float y[WV_M] = {0.f, 0.f, 0.f, 0.f};
        for (int c = 0; c < k; ++c) {
          const float *gc = Gs + (size_t)c * WV_K + l;
          const float vc = lv[c];
#pragma unroll
          for (int m = 0; m < WV_M; ++m) y[m] = __builtin_fmaf(gc[64 * m], vc, y[m]);
        }
        double pa = 0.0;
#pragma unroll
        for (int m = 0; m < WV_M; ++m) pa += (double)v[m] * (double)y[m];
        const float al = (float)wave_sum_f64(pa);
        if (l == 0) la[j] = al;
        amax = fmaxf(amax, fabsf(al));
        if (j == k - 1) break;
#pragma unroll
        for (int m = 0; m < WV_M; ++m) w[m] = __builtin_fmaf(-al, v[m], __builtin_fmaf(-bj, vp[m], y[m]));
        const float pre = (float)wv_sum_sq(w);
        wv_cgs(w, j, Vs, lh);
        float b2 = (float)wv_sum_sq(w);
        if (b2 < 0.5f * pre) {
          wv_cgs(w, j, Vs, lh);
          b2 = (float)wv_sum_sq(w);
        }
        float b = sqrtf(b2);
        float scale;
        if (b > TAU * (amax + bmax)) {
          bj = b;
          bmax = fmaxf(bmax, b);
          scale = 1.0f / b;
        } else {
          // restart: e_i for the first i of largest residual 1 - ||V[i][:]||^2
          st |= ST_BREAKDOWN;
          bj = 0.f;
          float best = -1.f;
          int bi = 0;
#pragma unroll
          for (int m = 0; m < WV_M; ++m) {
            const int r = l + 64 * m;
            float rs = 0.f;
            for (int c = 0; c <= j; ++c) {
              const float x = Vs[(size_t)c * WV_K + r];
              rs = __builtin_fmaf(x, x, rs);
            }
            const float res = r < k ? 1.f - rs : -1.f;
            if (res > best) {
              best = res;
              bi = r;
            }
          }
          bi = restart_lane<64>(best, bi, WV_K);
#pragma unroll
          for (int m = 0; m < WV_M; ++m) w[m] = (l + 64 * m == bi) ? 1.f : 0.f;
          wv_cgs(w, j, Vs, lh);
          wv_cgs(w, j, Vs, lh);
          scale = 1.0f / sqrtf((float)wv_sum_sq(w));
        }
        if (l == 0) lb[j + 1] = bj;
#pragma unroll
        for (int m = 0; m < WV_M; ++m) {
          const int r = l + 64 * m;
          vp[m] = v[m];
          v[m] = w[m] * scale;
          Vs[(size_t)(j + 1) * WV_K + r] = v[m];
          lv[r] = v[m];
        }
        __syncthreads();
      }
      if (l == 0) lb[0] = 0.f;
      __syncthreads();
      // ---- T, bounds ----
      bool finite = true;
      double glo = INFINITY, ghi = -INFINITY, amin = INFINITY, amx = -INFINITY;
      float2 *tri = a.tri + s * (int64_t)a.tstride;
#pragma unroll
      for (int m = 0; m < WV_M; ++m) {
        const int r = l + 64 * m;
        if (r < k) {
          const float ar = la[r], br = lb[r], bn = r + 1 < k ? lb[r + 1] : 0.f;
          finite = finite && isfinite(ar) && isfinite(br);
          tri[r] = make_float2(ar, br);
          const double rad = (double)fabsf(br) + (double)fabsf(bn);
          glo = fmin(glo, (double)ar - rad);
          ghi = fmax(ghi, (double)ar + rad);
          amin = fmin(amin, (double)ar);
          amx = fmax(amx, (double)ar);
        }
      }
      if (__ballot(!finite) != 0) st |= ST_NONFINITE;
      glo = g_min<64>(glo);
      ghi = g_max<64>(ghi);
      amin = g_min<64>(amin);
      amx = g_max<64>(amx);
      if (l == 0 && !(st & ST_NONFINITE)) a.bounds[s] = make_double4(glo, ghi, amin, amx);
    }
    if (l == 0) a.status[s] = st;
    __syncthreads();
  }
}

hipError_t launch_wave(const SolveArgs &a, float *scratch, int64_t slots, hipStream_t s) {
  int64_t blocks = a.bins ? slots : std::min<int64_t>(slots, a.nslots);
  if (blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(solve_wave_kernel, dim3((unsigned)blocks), dim3(64), 0, s, a, scratch);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Power-iteration estimate of lambda_max per subset: the counterpart of
// UPPER_KER (ParaLanczos.cu:1484-1626), which runs 10 steps of w = A v,
// b = ||w||, v = w / b on one full n <= 128 matrix held in per-lane 4 x 4
// register tiles and returns the last b. Here: any principal submatrix with
// k <= KC (64, 128 or 256), ones/sqrt(k) as the start vector, `iters` steps.
// One workgroup per subset of Q = KC / 64 row parts (KC Q lanes): lane (i, h)
// holds G_S[i][64 h + c], c < 64, in VGPRs; v is broadcast from LDS; the row
// parts' partial dots meet in LDS; ||w||^2 sums the exact FP32 squares in
// FP64 over the wavefronts of part 0. Three barriers per step.
// ---------------------------------------------------------------------------
template <int KC>
__global__ __launch_bounds__(KC * (KC / 64)) void upper_kernel(SolveArgs a, int iters,
                                                               float *upper) {
  constexpr int Q = KC / 64, NT = KC * Q, WP = KC / 64;  // parts, lanes, waves of a part
  __shared__ __attribute__((aligned(16))) float lv[KC];
  __shared__ float part[Q][KC];
  __shared__ double rn[WP];
  __shared__ int lidx[KC];
  __shared__ int flag[NT / 64];
  const int t = threadIdx.x, i = t % KC, h = t / KC, q = t & 63;
  for (int64_t s = blockIdx.x; s < a.nslots; s += gridDim.x) {
    int k;
    int64_t off;
    if (a.offsets) {
      off = a.offsets[s];
      k = (int)(a.offsets[s + 1] - off);
    } else {
      k = a.kuni;
      off = s * (int64_t)k;
    }
    const bool badk = k < 1 || k > KC || k > a.kcap;
    uint32_t st = badk ? ST_BADK : 0u;
    int my = 0;
    if (!st && i < k) {
      my = a.idx[off + i];
      if (my < 0 || (int64_t)my >= a.n) st = ST_BADINDEX;
    }
    if (h == 0) lidx[i] = my;
    {
      const bool any = __ballot(st != 0) != 0;
      if (q == 0) flag[t >> 6] = any ? 1 : 0;
    }
    __syncthreads();
    int anyb = 0;
#pragma unroll
    for (int u = 0; u < NT / 64; ++u) anyb |= flag[u];
    if (anyb) st = badk ? ST_BADK : ST_BADINDEX;
    float g[64];
    bool fin = true;
#pragma unroll
    for (int c = 0; c < 64; ++c) {
      const int col = 64 * h + c;
      float x = 0.f;
      if (st == 0 && i < k && col < k) {
        const int ic = lidx[col];
        const int64_t lo = min(ic, my), hi = max(ic, my);
        x = a.G[lo * a.ld + hi];
      }
      g[c] = x;
      fin = fin && isfinite(x);
    }
    __syncthreads();  // lidx is reused by the next subset
    {
      const bool bad = __ballot(!fin) != 0;
      if (q == 0) flag[t >> 6] = bad ? 1 : 0;
    }
    if (h == 0) lv[i] = (i < k) ? 1.0f / sqrtf((float)k) : 0.f;
    __syncthreads();
    anyb = 0;
#pragma unroll
    for (int u = 0; u < NT / 64; ++u) anyb |= flag[u];
    if (anyb && st == 0) st = ST_NONFINITE;
    float b = 0.f;
    if (st == 0) {
      for (int l = 0; l < iters; ++l) {
        float d0 = 0.f, d1 = 0.f, d2 = 0.f, d3 = 0.f;
        const float4 *v4 = reinterpret_cast<const float4 *>(lv + 64 * h);
#pragma unroll
        for (int c = 0; c < 16; ++c) {
          const float4 x = v4[c];
          d0 = __builtin_fmaf(g[4 * c + 0], x.x, d0);
          d1 = __builtin_fmaf(g[4 * c + 1], x.y, d1);
          d2 = __builtin_fmaf(g[4 * c + 2], x.z, d2);
          d3 = __builtin_fmaf(g[4 * c + 3], x.w, d3);
        }
        part[h][i] = (d0 + d1) + (d2 + d3);
        __syncthreads();
        float w = part[0][i];
#pragma unroll
        for (int u = 1; u < Q; ++u) w += part[u][i];
        const double sq = wave_sum_f64((double)w * (double)w);
        if (h == 0 && q == 0) rn[t >> 6] = sq;
        __syncthreads();
        double tot = rn[0];
#pragma unroll
        for (int u = 1; u < WP; ++u) tot += rn[u];
        b = (float)sqrt(tot);
        if (h == 0) lv[i] = b > 0.f ? w / b : 0.f;
        __syncthreads();
      }
      if (!isfinite(b)) st = ST_NONFINITE;
    }
    if (t == 0) {
      upper[s] = (st & (ST_BADK | ST_BADINDEX | ST_NONFINITE)) ? NAN : b;
      a.status[s] = st;
    }
    __syncthreads();
  }
}

hipError_t launch_upper(const SolveArgs &a, int iters, float *upper, hipStream_t s) {
  if (a.nslots <= 0) return hipSuccess;
  const unsigned blocks = (unsigned)std::min<int64_t>(a.nslots, 1 << 16);
  if (a.kcap <= 64)
    hipLaunchKernelGGL(upper_kernel<64>, dim3(blocks), dim3(64), 0, s, a, iters, upper);
  else if (a.kcap <= 128)
    hipLaunchKernelGGL(upper_kernel<128>, dim3(blocks), dim3(256), 0, s, a, iters, upper);
  else
    hipLaunchKernelGGL(upper_kernel<256>, dim3(blocks), dim3(1024), 0, s, a, iters, upper);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Binning of subsets by size: count per bin (LDS histogram per block, one
// global atomic per bin and block), exclusive scan, scatter of subset ids into
// the order array (LDS atomics give each subset its place within the block's
// share of the bin, one global atomic per bin and block reserves that share).
// Subsets keep their own results whatever slot they land in, so the order
// inside a bin does not matter.
// ---------------------------------------------------------------------------
constexpr int BIN_THREADS = 256;
constexpr int BIN_ITEMS = 8;

__device__ __forceinline__ int bin_of(const int64_t *offsets, int kuni, const BinLut &lut,
                                      int64_t s) {
  const int64_t k = offsets ? offsets[s + 1] - offsets[s] : (int64_t)kuni;
  return lut.bin[lut_index(k)];
}

__global__ void bin_clear_kernel(int32_t *table) {
  for (int i = threadIdx.x; i < 3 * MAXBIN; i += blockDim.x) table[i] = 0;
}

__global__ __launch_bounds__(BIN_THREADS) void bin_count_kernel(const int64_t *offsets, int kuni,
                                                                int64_t nsub, BinLut lut,
                                                                int nbins, int32_t *table) {
  __shared__ int h[MAXBIN];
  for (int i = threadIdx.x; i < nbins; i += BIN_THREADS) h[i] = 0;
  __syncthreads();
  for (int64_t s = (int64_t)blockIdx.x * BIN_THREADS + threadIdx.x; s < nsub;
       s += (int64_t)gridDim.x * BIN_THREADS)
    atomicAdd(&h[bin_of(offsets, kuni, lut, s)], 1);
  __syncthreads();
  for (int i = threadIdx.x; i < nbins; i += BIN_THREADS)
    if (h[i]) atomicAdd(&table[BIN_COUNT + i], h[i]);
}

__global__ void bin_scan_kernel(int32_t *table, int nbins) {
  if (threadIdx.x != 0) return;
  int acc = 0;
  for (int b = 0; b < nbins; ++b) {
    table[BIN_START + b] = acc;
    table[BIN_FILL + b] = 0;
    acc += table[BIN_COUNT + b];
  }
}

__global__ __launch_bounds__(BIN_THREADS) void bin_scatter_kernel(const int64_t *offsets,
                                                                  int kuni, int64_t nsub,
                                                                  BinLut lut, int nbins,
                                                                  int32_t *table,
                                                                  int32_t *order) {
  __shared__ int h[MAXBIN];
  __shared__ int base[MAXBIN];
  for (int i = threadIdx.x; i < nbins; i += BIN_THREADS) h[i] = 0;
  __syncthreads();
  const int64_t t0 = (int64_t)blockIdx.x * BIN_THREADS * BIN_ITEMS;
  int b[BIN_ITEMS], loc[BIN_ITEMS];
#pragma unroll
  for (int u = 0; u < BIN_ITEMS; ++u) {
    const int64_t s = t0 + u * BIN_THREADS + threadIdx.x;
    b[u] = 0;
    loc[u] = 0;
    if (s < nsub) {
      b[u] = bin_of(offsets, kuni, lut, s);
      loc[u] = atomicAdd(&h[b[u]], 1);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nbins; i += BIN_THREADS)
    if (h[i]) base[i] = table[BIN_START + i] + atomicAdd(&table[BIN_FILL + i], h[i]);
  __syncthreads();
#pragma unroll
  for (int u = 0; u < BIN_ITEMS; ++u) {
    const int64_t s = t0 + u * BIN_THREADS + threadIdx.x;
    if (s < nsub) order[base[b[u]] + loc[u]] = (int32_t)s;
  }
}

hipError_t launch_binning(const int64_t *offsets, int kuni, int64_t nsub, const BinLut &lut,
                          int nbins, int32_t *table, int32_t *order, hipStream_t s) {
  if (nbins < 1 || nbins > MAXBIN || nsub < 0 || nsub > 0x7fffffffLL)
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(bin_clear_kernel, dim3(1), dim3(256), 0, s, table);
  if (nsub == 0) return hipGetLastError();
  const int64_t cblocks = std::min<int64_t>((nsub + BIN_THREADS - 1) / BIN_THREADS, 1024);
  hipLaunchKernelGGL(bin_count_kernel, dim3((unsigned)cblocks), dim3(BIN_THREADS), 0, s, offsets,
                     kuni, nsub, lut, nbins, table);
  hipLaunchKernelGGL(bin_scan_kernel, dim3(1), dim3(64), 0, s, table, nbins);
  const int64_t tile = (int64_t)BIN_THREADS * BIN_ITEMS;
  const int64_t sblocks = (nsub + tile - 1) / tile;
  hipLaunchKernelGGL(bin_scatter_kernel, dim3((unsigned)sblocks), dim3(BIN_THREADS), 0, s,
                     offsets, kuni, nsub, lut, nbins, table, order);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Stage (c): FP64 Sturm-sequence bisection. A block of 64 lanes stages the T
// of its subsets into LDS once (row j of the stage holds subset j's
// (alpha_i, beta_i), rows kstride+1 float2 apart so the staging writes and
// the per-step reads hit distinct banks); the extremes are then found by
// multisection, 8 lanes per eigenvalue, and every eigenvalue of a subset
// (the legacy path) by bisection, one lane each.
// ---------------------------------------------------------------------------
constexpr int BIS_BLOCK = 64;

// Number of eigenvalues of T below sigma: sign changes of the Sturm sequence
// p_0 = 1, p_1 = alpha_0 - sigma, p_i = (alpha_{i-1} - sigma) p_{i-1} -
// beta_{i-1}^2 p_{i-2} (Barth, Martin, Wilkinson, Numer. Math. 9, 1967; the
// reference's unused BM_POLY, ParaLanczos.cu:504-552, without its off-by-one
// defects), evaluated in FP64 from the FP32 entries (exact conversions). It
// needs no division. An exactly zero p_i is replaced by -2^-60 p_{i-1}, so it
// takes the sign opposite to p_{i-1} (Wilkinson's convention) and the
// recurrence continues through a vanishing beta; the pair (p_{i-1}, p_i) is
// rescaled by a power of two every 8 steps so that nothing overflows.
__device__ __forceinline__ int sturm_count(const float2 *tri, int k, double sigma) {
  double p0 = 1.0;
  double p1 = (double)tri[0].x - sigma;
  if (p1 == 0.0) p1 = -0x1p-60;
  bool neg1 = p1 < 0.0;
  int c = neg1;
  float2 t = tri[k > 1 ? 1 : 0];
  for (int i = 1; i < k; ++i) {
    const float2 tn = tri[i + 1 < k ? i + 1 : i];  // next entry, read ahead of the math
    const double b = (double)t.y;
    double p2 = __builtin_fma((double)t.x - sigma, p1, -(b * b) * p0);
    p2 = (p2 == 0.0) ? -p1 * 0x1p-60 : p2;
    t = tn;
    const bool neg2 = p2 < 0.0;
    c += neg2 != neg1;
    neg1 = neg2;
    p0 = p1;
    p1 = p2;
    if ((i & 7) == 0) {
      int e;
      (void)frexp(fmax(fabs(p0), fabs(p1)), &e);
      p0 = ldexp(p0, -e);
      p1 = ldexp(p1, -e);
    }
  }
  return c;
}

// eigenvalue `target` (0-based, ascending) of T by bisection of [lo, hi];
// stops at width <= max(2^-tolbits max(|lo|,|hi|), 2^-52 tnorm)
__device__ __forceinline__ double bisect(const float2 *tri, int k, int target, double lo,
                                         double hi, double tnorm, int tolbits, bool &conv) {
  const double floor_tol = ldexp(tnorm, -52) + 0x1p-1000;
  int it = 0;
  for (; it < BIS_MAXIT; ++it) {
    const double tol = fmax(ldexp(fmax(fabs(lo), fabs(hi)), -tolbits), floor_tol);
    if (hi - lo <= tol) break;
    const double mid = lo + 0.5 * (hi - lo);
    if (sturm_count(tri, k, mid) > target) hi = mid;
    else lo = mid;
  }
  conv = it < BIS_MAXIT;
  return lo + 0.5 * (hi - lo);
}

// Gershgorin bounds widened as LAPACK dstebz does
__device__ __forceinline__ void widened(const double4 &b, int k, double &glo, double &ghi,
                                        double &tnorm) {
  tnorm = fmax(fabs(b.x), fabs(b.y));
  const double fudge = 2.0 * DBL_EPSILON * tnorm * k + 0x1p-1000;
  glo = b.x - fudge;
  ghi = b.y + fudge;
}

__device__ __forceinline__ int subset_k(const BisectArgs &a, int64_t s) {
  return a.offsets ? (int)(a.offsets[s + 1] - a.offsets[s]) : a.kuni;
}

__device__ __forceinline__ int64_t subset_at(const BisectArgs &a, int64_t slot) {
  return a.order ? (int64_t)a.order[slot] : slot;
}

// copy the T of the subsets in slots s0 .. s0+cnt-1 into the LDS stage (one
// row per subset)
__device__ __forceinline__ void stage_tridiag(const BisectArgs &a, int64_t s0, int cnt,
                                              float2 *stage) {
  const int rs = a.tstride + 1;
  for (int j = 0; j < cnt; ++j) {
    const float2 *src = a.tri + subset_at(a, s0 + j) * (int64_t)a.tstride;
    for (int i = threadIdx.x; i < a.tstride; i += BIS_BLOCK) stage[j * rs + i] = src[i];
  }
  __syncthreads();
}

// lambda_min / lambda_max by multisection: MS lanes per extreme (1 or 8),
// BIS_BLOCK / (2 MS) subsets per 64-lane block (lanes 2 MS j .. 2 MS j +
// MS - 1 seek lambda_min of the block's subset j, the next MS lanes its
// lambda_max). Each iteration the MS lanes of an extreme evaluate the Sturm
// count at the interior points lo + (r + 1) d, d = (hi - lo) / (MS + 1),
// r = 0..MS-1, and all move to the part of [lo, hi] that holds the target:
// the first point whose count exceeds it (a ballot) is the new hi, the point
// before it (or lo) the new lo, each the same double the evaluating lane
// formed. MS = 1 is plain bisection. Brackets [Gershgorin lo, min alpha] and
// [max alpha, Gershgorin hi] hold since lambda_min(T) <= T_ii <=
// lambda_max(T), and the stopping width is the same for both. With MS = 8
// an iteration gains log2(9) = 3.2 bits instead of 1: the serial chain of
// Sturm sweeps a subset waits for is a third as long, for 2.9x the sweeps,
// which pays while the launch has few subsets (launch_bisect chooses).
constexpr int BIS_MS_WIDE = 8;
constexpr int BIS_WIDE_MAX_SUBSETS = 16384;  // multisection up to this many subsets
template <int MS>
__global__ __launch_bounds__(BIS_BLOCK) void bisect_extreme_kernel(BisectArgs a) {
  constexpr int BIS_MS = MS;
  constexpr int BIS_SUBS = BIS_BLOCK / (2 * MS);  // subsets per block
  extern __shared__ float2 stage[];
  const int64_t s0 = (int64_t)blockIdx.x * BIS_SUBS;
  const int cnt = (int)min<int64_t>(BIS_SUBS, a.nsub - s0);
  stage_tridiag(a, s0, cnt, stage);
  const int grp = threadIdx.x / BIS_MS, r = threadIdx.x % BIS_MS;
  const int j = grp >> 1, upper = grp & 1;
  if (j >= cnt) return;  // both groups of a subset leave together
  const int64_t s = subset_at(a, s0 + j);
  const int k = subset_k(a, s);
  const uint32_t st = a.status[s];
  double lam = NAN;
  int noconv = 0;
  if (!(st & (ST_BADK | ST_BADINDEX | ST_NONFINITE))) {
    const float2 *tri = stage + j * (a.tstride + 1);
    const double4 bd = a.bounds[s];
    double glo, ghi, tnorm;
    widened(bd, k, glo, ghi, tnorm);
    const double floor_tol = ldexp(tnorm, -52) + 0x1p-1000;
    double lo = upper ? bd.w : glo, hi = upper ? ghi : bd.z;
    const int target = upper ? k - 1 : 0;
    int it = 0;
    for (; it < BIS_MAXIT; ++it) {
      const double tol = fmax(ldexp(fmax(fabs(lo), fabs(hi)), -BIS_TOLBITS_EXT), floor_tol);
      if (hi - lo <= tol) break;  // lo, hi are the same on the group's lanes
      const double d = (hi - lo) * (1.0 / (BIS_MS + 1));
      const bool above = sturm_count(tri, k, lo + (double)(r + 1) * d) > target;
      const unsigned m = (unsigned)(__ballot(above) >> (grp * BIS_MS)) & ((1u << BIS_MS) - 1);
      const int f = m ? __builtin_ctz(m) : BIS_MS;  // first point above the target
      const double nlo = lo + (double)f * d;
      hi = f < BIS_MS ? lo + (double)(f + 1) * d : hi;
      lo = nlo;
    }
    noconv = it == BIS_MAXIT;
    lam = lo + 0.5 * (hi - lo);
  }
  const double lmax = __shfl_xor(lam, BIS_MS, 64);
  noconv |= __shfl_xor(noconv, BIS_MS, 64);
  if (!upper && r == 0) {
    a.lmin[s] = (float)lam;
    a.lmax[s] = (float)lmax;
    a.delta[s] = (float)fmax(lmax - 1.0, 1.0 - lam);
    a.status[s] = st | (noconv ? ST_NOCONV : 0u);
  }
}

// every eigenvalue (uniform k): lane t of the block's targets bisects
// eigenvalue t % k of subset t / k
__global__ __launch_bounds__(BIS_BLOCK) void bisect_all_kernel(BisectArgs a) {
  extern __shared__ float2 stage[];
  const int k = a.kuni;
  const int64_t t0 = (int64_t)blockIdx.x * BIS_BLOCK;
  const int64_t s0 = t0 / k;
  const int64_t s_last = min<int64_t>((t0 + BIS_BLOCK - 1) / k, a.nsub - 1);
  stage_tridiag(a, s0, (int)(s_last - s0 + 1), stage);
  const int64_t gid = t0 + threadIdx.x;
  const int64_t s = gid / k;
  if (s >= a.nsub) return;
  const int t = (int)(gid - s * k);
  const uint32_t st = a.status[s];
  double lam = NAN;
  if (!(st & (ST_BADK | ST_BADINDEX | ST_NONFINITE))) {
    const float2 *tri = stage + (s - s0) * (a.tstride + 1);
    double glo, ghi, tnorm;
    widened(a.bounds[s], k, glo, ghi, tnorm);
    bool conv;
    lam = bisect(tri, k, t, glo, ghi, tnorm, BIS_TOLBITS_ALL, conv);
    if (!conv) atomicOr(a.status + s, ST_NOCONV);
  }
  a.eig[gid] = lam;
}

// det = product of the eigenvalues (after bisect_all_kernel)
__global__ __launch_bounds__(256) void det_kernel(BisectArgs a) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.nsub) return;
  double d = 1.0;
  for (int t = 0; t < a.kuni; ++t) d *= a.eig[s * a.kuni + t];
  a.det[s] = d;
}

hipError_t launch_bisect(int mode, const BisectArgs &a, hipStream_t s) {
  if (a.nsub <= 0) return hipSuccess;
  if (mode == 0) {
    const bool wide = a.nsub <= BIS_WIDE_MAX_SUBSETS;
    const int subs = BIS_BLOCK / (2 * (wide ? BIS_MS_WIDE : 1));
    const int64_t blocks = (a.nsub + subs - 1) / subs;
    const size_t lds = (size_t)subs * (a.tstride + 1) * sizeof(float2);
    const void *fn = wide ? (const void *)bisect_extreme_kernel<BIS_MS_WIDE>
                          : (const void *)bisect_extreme_kernel<1>;
    if (lds > 65536) {
      hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      if (e != hipSuccess) return e;
    }
    if (wide)
      hipLaunchKernelGGL((bisect_extreme_kernel<BIS_MS_WIDE>), dim3((unsigned)blocks),
                         dim3(BIS_BLOCK), lds, s, a);
    else
      hipLaunchKernelGGL((bisect_extreme_kernel<1>), dim3((unsigned)blocks), dim3(BIS_BLOCK),
                         lds, s, a);
  } else {
    const int64_t threads = a.nsub * a.kuni;
    const int64_t blocks = (threads + BIS_BLOCK - 1) / BIS_BLOCK;
    // a block's 64 targets span at most 64/k + 2 subsets
    const int rows = BIS_BLOCK / a.kuni + 2;
    const size_t lds = (size_t)rows * (a.tstride + 1) * sizeof(float2);
    if (lds > 65536) {
      hipError_t e = hipFuncSetAttribute((const void *)bisect_all_kernel,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(bisect_all_kernel, dim3((unsigned)blocks), dim3(BIS_BLOCK), lds, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(det_kernel, dim3((unsigned)((a.nsub + 255) / 256)), dim3(256), 0, s, a);
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// RIP statistics of a solved batch, per subset size (the per-index-set
// lower/upper bound means and deviations the reference's SubSeq record holds,
// ParaLanczos0.h:82-90, and the largest deviation delta). Block-local FP64
// sums in LDS, one global atomic per field, size and block, then a finalize
// pass. Row layout (RIP_FIELDS doubles per size): count, sum/mean lmin,
// sumsq/std lmin, sum/mean lmax, sumsq/std lmax, max delta, min lmin, max lmax.
// Row 0 counts the subsets skipped (fatal status or size outside [1, kmax]).
// ---------------------------------------------------------------------------
constexpr int RIP_FIELDS = 8;
constexpr int RIP_THREADS = 256;

// order-preserving map of a float onto unsigned (for atomic min/max)
__device__ __forceinline__ unsigned f2ord(float f) {
  const unsigned u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(unsigned o) {
  return __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}

__global__ void rip_clear_kernel(double *out, unsigned *ext, int kmax) {
  for (int i = threadIdx.x + blockIdx.x * blockDim.x; i < (kmax + 1) * RIP_FIELDS;
       i += blockDim.x * gridDim.x)
    out[i] = 0.0;
  for (int i = threadIdx.x + blockIdx.x * blockDim.x; i < (kmax + 1) * 3;
       i += blockDim.x * gridDim.x)
    ext[i] = (i % 3 == 1) ? f2ord(INFINITY) : f2ord(-INFINITY);  // max delta, min lmin, max lmax
}

__global__ __launch_bounds__(RIP_THREADS) void rip_accumulate_kernel(
    const float *lmin, const float *lmax, const float *delta, const uint32_t *status,
    const int64_t *offsets, int kuni, int64_t nsub, int kmax, double *out, unsigned *ext) {
  extern __shared__ double acc[];  // (kmax+1) x 5 sums, then (kmax+1) x 3 ordered extremes
  unsigned *lext = reinterpret_cast<unsigned *>(acc + (kmax + 1) * 5);
  for (int i = threadIdx.x; i < (kmax + 1) * 5; i += RIP_THREADS) acc[i] = 0.0;
  for (int i = threadIdx.x; i < (kmax + 1) * 3; i += RIP_THREADS)
    lext[i] = (i % 3 == 1) ? f2ord(INFINITY) : f2ord(-INFINITY);
  __syncthreads();
  constexpr uint32_t FATAL = ST_BADK | ST_BADINDEX | ST_NONFINITE;
  for (int64_t s = (int64_t)blockIdx.x * RIP_THREADS + threadIdx.x; s < nsub;
       s += (int64_t)gridDim.x * RIP_THREADS) {
    const int64_t k = offsets ? offsets[s + 1] - offsets[s] : (int64_t)kuni;
    if (k < 1 || k > kmax || (status[s] & FATAL)) {
      atomicAdd(&acc[0], 1.0);
      continue;
    }
    const double lo = lmin[s], hi = lmax[s];
    double *a = acc + k * 5;
    atomicAdd(&a[0], 1.0);
    atomicAdd(&a[1], lo);
    atomicAdd(&a[2], lo * lo);
    atomicAdd(&a[3], hi);
    atomicAdd(&a[4], hi * hi);
    atomicMax(&lext[k * 3 + 0], f2ord(delta[s]));
    atomicMin(&lext[k * 3 + 1], f2ord(lmin[s]));
    atomicMax(&lext[k * 3 + 2], f2ord(lmax[s]));
  }
  __syncthreads();
  for (int i = threadIdx.x; i < (kmax + 1) * 5; i += RIP_THREADS) {
    if (acc[i] != 0.0) atomicAdd(&out[(i / 5) * RIP_FIELDS + (i % 5)], acc[i]);
  }
  for (int i = threadIdx.x; i < (kmax + 1) * 3; i += RIP_THREADS) {
    if ((i % 3) == 1) atomicMin(&ext[i], lext[i]);
    else atomicMax(&ext[i], lext[i]);
  }
}

__global__ void rip_finalize_kernel(double *out, const unsigned *ext, int kmax) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > kmax) return;
  double *r = out + (size_t)k * RIP_FIELDS;
  if (k == 0) {
    for (int f = 1; f < RIP_FIELDS; ++f) r[f] = 0.0;
    return;
  }
  const double n = r[0];
  if (n == 0.0) {
    for (int f = 1; f < RIP_FIELDS; ++f) r[f] = NAN;
    return;
  }
  const double m1 = r[1] / n, m2 = r[3] / n;
  const double v1 = fmax(r[2] / n - m1 * m1, 0.0), v2 = fmax(r[4] / n - m2 * m2, 0.0);
  r[1] = m1;
  r[2] = sqrt(v1);
  r[3] = m2;
  r[4] = sqrt(v2);
  r[5] = ord2f(ext[k * 3 + 0]);
  r[6] = ord2f(ext[k * 3 + 1]);
  r[7] = ord2f(ext[k * 3 + 2]);
}

hipError_t launch_rip_stats(const float *lmin, const float *lmax, const float *delta,
                            const uint32_t *status, const int64_t *offsets, int kuni,
                            int64_t nsub, int kmax, double *out, unsigned *ext, hipStream_t s) {
  hipLaunchKernelGGL(rip_clear_kernel, dim3(4), dim3(256), 0, s, out, ext, kmax);
  if (nsub > 0) {
    const size_t lds = (size_t)(kmax + 1) * (5 * sizeof(double) + 3 * sizeof(unsigned));
    const int64_t blocks = std::min<int64_t>((nsub + RIP_THREADS - 1) / RIP_THREADS, 512);
    hipLaunchKernelGGL(rip_accumulate_kernel, dim3((unsigned)blocks), dim3(RIP_THREADS), lds, s,
                       lmin, lmax, delta, status, offsets, kuni, nsub, kmax, out, ext);
  }
  hipLaunchKernelGGL(rip_finalize_kernel, dim3((kmax + 1 + 255) / 256), dim3(256), 0, s, out, ext,
                     kmax);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Generators
// ---------------------------------------------------------------------------

// Floyd sample per subset, one thread per subset; membership test against the
// thread's own earlier writes
__global__ void gen_subsets_kernel(uint64_t seed, int64_t s0, int64_t nsub,
                                   const int64_t *offsets, int kuni, int n, int32_t *idx) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsub) return;
  const int64_t off = offsets ? offsets[s] : s * (int64_t)kuni;
  const int k = offsets ? (int)(offsets[s + 1] - off) : kuni;
  if (k < 1 || k > n) return;
  Stream st(seed, TAG_IDX, (uint64_t)(s0 + s));
  int32_t *out = idx + off;
  int cnt = 0;
  for (int j = n - k; j < n; ++j) {
    const int32_t t = (int32_t)st.bounded((uint32_t)j + 1u);
    bool seen = false;
    for (int q = 0; q < cnt; ++q) seen = seen || (out[q] == t);
    out[cnt++] = seen ? j : t;
  }
}

hipError_t launch_gen_subsets(uint64_t seed, int64_t s0, int64_t nsub, const int64_t *offsets,
                              int kuni, int n, int32_t *idx, hipStream_t s) {
  if (nsub <= 0) return hipSuccess;
  const int64_t blocks = (nsub + 255) / 256;
  hipLaunchKernelGGL(gen_subsets_kernel, dim3((unsigned)blocks), dim3(256), 0, s, seed, s0, nsub,
                     offsets, kuni, n, idx);
  return hipGetLastError();
}

struct KList {
  int32_t v[32];
};

__global__ void gen_ks_kernel(uint64_t seed, int64_t s0, int64_t nsub, KList kl, int nk,
                              int32_t *ks) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsub) return;
  Stream st(seed, TAG_K, (uint64_t)(s0 + s));
  ks[s] = kl.v[st.bounded((uint32_t)nk)];
}

hipError_t launch_gen_ks(uint64_t seed, int64_t s0, int64_t nsub, const int32_t *klist, int nk,
                         int32_t *ks, hipStream_t s) {
  if (nk < 1 || nk > 32) return hipErrorInvalidValue;
  if (nsub <= 0) return hipSuccess;
  KList kl;
  for (int i = 0; i < 32; ++i) kl.v[i] = i < nk ? klist[i] : 0;
  const int64_t blocks = (nsub + 255) / 256;
  hipLaunchKernelGGL(gen_ks_kernel, dim3((unsigned)blocks), dim3(256), 0, s, seed, s0, nsub, kl,
                     nk, ks);
  return hipGetLastError();
}

// Gaussian entries: Box-Muller on pair p = e/2 of the column-major element
// index e = j*m + r, 53-bit uniforms from stream (seed, TAG_PHI, p) word 0..3
__global__ void gen_gauss_kernel(uint64_t seed, int64_t total, double scale, float *phi) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (2 * p >= total) return;
  uint32_t r[4];
  stream_block(seed, TAG_PHI, (uint64_t)p, 0u, r);
  const double u1 = uniform53(r[0], r[1]), u2 = uniform53(r[2], r[3]);
  const double rad = sqrt(-2.0 * log(1.0 - u1));
  const double ang = 6.283185307179586476925286766559 * u2;
  phi[2 * p] = (float)(rad * cos(ang) * scale);
  if (2 * p + 1 < total) phi[2 * p + 1] = (float)(rad * sin(ang) * scale);
}

// divide each column by its FP64 L2 norm (one block per column)
__global__ void normalize_cols_kernel(int m, float *phi) {
  __shared__ double part[256];
  float *col = phi + (int64_t)blockIdx.x * m;
  double acc = 0.0;
  for (int r = threadIdx.x; r < m; r += blockDim.x) acc += (double)col[r] * (double)col[r];
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
    __syncthreads();
  }
  const double inv = 1.0 / sqrt(part[0]);
  for (int r = threadIdx.x; r < m; r += blockDim.x) col[r] = (float)((double)col[r] * inv);
}

hipError_t launch_gen_phi_gauss(uint64_t seed, int m, int n, float *phi, hipStream_t s) {
  const int64_t total = (int64_t)m * n;
  const int64_t pairs = (total + 1) / 2;
  hipLaunchKernelGGL(gen_gauss_kernel, dim3((unsigned)((pairs + 255) / 256)), dim3(256), 0, s,
                     seed, total, 1.0 / sqrt((double)m), phi);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(normalize_cols_kernel, dim3((unsigned)n), dim3(256), 0, s, m, phi);
  return hipGetLastError();
}

// +-1/sqrt(m), bit e&31 of word (e>>5)&3 of stream (seed, TAG_PHI, e>>7)
__global__ void gen_bernoulli_kernel(uint64_t seed, int64_t total, float v, float *phi) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  uint32_t r[4];
  stream_block(seed, TAG_PHI, (uint64_t)e >> 7, 0u, r);
  const uint32_t w = r[(e >> 5) & 3];
  phi[e] = ((w >> (e & 31)) & 1u) ? v : -v;
}

hipError_t launch_gen_phi_bernoulli(uint64_t seed, int m, int n, float *phi, hipStream_t s) {
  const int64_t total = (int64_t)m * n;
  hipLaunchKernelGGL(gen_bernoulli_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                     seed, total, (float)(1.0 / sqrt((double)m)), phi);
  return hipGetLastError();
}

// rows of the orthonormal DCT-II scaled by sqrt(n/m); angle reduced exactly
__global__ void gen_dct_kernel(const int32_t *rows, int m, int n, float *phi) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)m * n) return;
  const int64_t j = e / m;
  const int i = (int)(e - j * m);
  const int64_t r = rows[i];
  const int64_t a = (r * (2 * j + 1)) % (4 * (int64_t)n);
  const double c = cos(3.14159265358979323846264338327950288 * (double)a / (2.0 * n));
  const double amp = sqrt(2.0 / n) * sqrt((double)n / m);
  phi[e] = (float)(amp * (r == 0 ? sqrt(0.5) : 1.0) * c);
}

hipError_t launch_gen_phi_dct(const int32_t *rows_dev, int m, int n, float *phi, hipStream_t s) {
  const int64_t total = (int64_t)m * n;
  hipLaunchKernelGGL(gen_dct_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                     rows_dev, m, n, phi);
  return hipGetLastError();
}

}  // namespace pl
