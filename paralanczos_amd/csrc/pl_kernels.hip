// HIP kernels for gfx950 (MI355X): the batched principal-submatrix Lanczos
// eigenvalue path and the synthetic-input generators.
//
// One lane group of GROUP lanes (8, 16, 32 or 64; a 64-lane wavefront holds
// 64/GROUP groups) solves one subset of size k <= GROUP. Lane i owns row i of
// the gathered submatrix G_S, entry i of every Lanczos vector, row i of the
// Lanczos basis V (vrow) and the whole Lanczos vector v_i (vcol). Rows and
// columns k..GROUP-1 are zero padding; they stay exactly zero through the
// recurrence, so a subset gives the same arithmetic in any GROUP >= k.
//
// Stages per subset (reference: POS_SYM_SUBMATRIX_KER2, ParaLanczos.cu:1157-1480):
//  (a) gather (ParaLanczos.cu:1237-1242 / :471-490): the upper triangle of G_S
//      in global index order, G[min(i,j)][max(i,j)], is loaded one subset row
//      per instruction (lanes = columns), mirrored into a padded LDS tile and
//      each lane then reads its row into VGPRs.
//  (b) k Lanczos steps (ParaLanczos.cu:1285-1374) in FP32: LDS all-gather of
//      v and matvec from the VGPR row; alpha = v^T G v by a compensated
//      (TwoProduct/TwoSum) group reduction; three-term update; one full
//      classical Gram-Schmidt pass against all previous vectors; beta = ||w||
//      compensated; deterministic restart on breakdown.
//  (c) FP64 Sturm-count multisection on T for lambda_min and lambda_max
//      (replaces the serial TQLRAT of ParaLanczos.cu:32-189, :1454-1462), or
//      per-lane bisection for every eigenvalue (legacy all-eigenvalue mode).
// No kernel uses a block-level barrier: all cross-lane traffic stays inside
// one wavefront, so groups finish independently.
#include <float.h>
#include <hip/hip_runtime.h>
#include <math.h>

#include "pl_common.hpp"
#include "pl_internal.hpp"

namespace pl {

constexpr int BLOCK = 256;
constexpr int MS_MAXIT = 96;      // multisection iteration cap
constexpr int MS_TOLBITS = 34;    // stop at width <= 2^-34 * max(|lo|,|hi|)
constexpr int BIS_MAXIT = 160;    // all-eigenvalue bisection cap
constexpr int BIS_TOLBITS = 44;

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---- group reductions (xor butterflies: every lane ends with identical bits) ----
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

__device__ __forceinline__ void two_sum(float a, float b, float &s, float &e) {
  s = a + b;
  const float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// compensated dot product over the group (Ogita, Rump, Oishi, "Accurate sum
// and dot product", SIAM J. Sci. Comput. 2005, Dot2 with a pairwise tree)
template <int W>
__device__ __forceinline__ float g_dot2(float x, float y) {
  float p = x * y;
  float e = __builtin_fmaf(x, y, -p);
#pragma unroll
  for (int m = 1; m < W; m <<= 1) {
    const float po = __shfl_xor(p, m, W);
    const float eo = __shfl_xor(e, m, W);
    float s, t;
    two_sum(p, po, s, t);
    p = s;
    e = (e + eo) + t;
  }
  return p + e;
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

// sum_c x[c] * L[c] with L a 16-byte aligned LDS vector read by broadcast;
// two accumulators split by index parity (fixed order for every GROUP)
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

// one classical Gram-Schmidt pass of w (entry per lane) against the basis:
// h_l = v_l . w computed by lane l from its vcol, then w -= sum_l V[i][l] h_l
template <int GROUP>
__device__ __forceinline__ float cgs_pass(float w, const float (&vrow)[GROUP],
                                          const float (&vcol)[GROUP], float *Lw, float *Lh,
                                          int gl) {
  Lw[gl] = w;
  wave_sync();
  const float h = bcast_dot<GROUP>(Lw, vcol);
  Lh[gl] = h;
  wave_sync();
  return w - bcast_dot<GROUP>(Lh, vrow);
}

// Stage (b): k Lanczos steps. T goes to Lt as interleaved (alpha_i, beta_i),
// beta_0 = 0 and beta_i coupling rows i-1 and i.
template <int GROUP>
__device__ uint32_t lanczos_core(const float (&g)[GROUP], int k, int gl, float *R) {
  float *Lv = R;
  float *Lw = R + GROUP;
  float *Lh = R + 2 * GROUP;
  float *Lt = R + 3 * GROUP;
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
  if (gl == 0) Lt[1] = 0.f;
  for (int j = 0; j < k; ++j) {
    Lv[gl] = vi;
    wave_sync();
    // matvec w_i = sum_c G_S[i][c] v_c; lane j keeps v_j as its basis vector
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
    amax = fmaxf(amax, fabsf(a));
    w = __builtin_fmaf(-a, vi, __builtin_fmaf(-bj, vp, w));
    w = cgs_pass<GROUP>(w, vrow, vcol, Lw, Lh, gl);
    if (gl == 0) Lt[2 * j] = a;
    if (j == k - 1) break;
    const float b = sqrtf(g_dot2<GROUP>(w, w));
    vp = vi;
    if (b > TAU * (amax + bmax)) {
      bj = b;
      bmax = fmaxf(bmax, b);
      vi = w * (1.0f / b);
    } else {
      // restart: e_i for the first i of largest residual 1 - ||V[i][:]||^2,
      // orthogonalised twice against the basis; beta_{j+1} = 0
      st |= ST_BREAKDOWN;
      bj = 0.f;
      float res = 1.f;
#pragma unroll
      for (int l = 0; l < GROUP; ++l) res = __builtin_fmaf(-vrow[l], vrow[l], res);
      if (gl >= k) res = -1.f;
      float bv = res;
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
      float r = (gl == bi) ? 1.f : 0.f;
      r = cgs_pass<GROUP>(r, vrow, vcol, Lw, Lh, gl);
      r = cgs_pass<GROUP>(r, vrow, vcol, Lw, Lh, gl);
      const float nb = sqrtf(g_dot2<GROUP>(r, r));
      vi = r * (1.0f / nb);
    }
    if (gl == 0) Lt[2 * (j + 1) + 1] = bj;
#pragma unroll
    for (int l = 1; l < GROUP; ++l) vrow[l] = (l == j + 1) ? vi : vrow[l];
  }
  wave_sync();
  bool finite = true;
  if (gl < k) finite = isfinite(Lt[2 * gl]) && isfinite(Lt[2 * gl + 1]);
  if (!g_all<GROUP>(finite)) st |= ST_NONFINITE;
  return st;
}

// number of eigenvalues of T below sigma (LAPACK dstebz-style pivot guard)
__device__ __forceinline__ int negcount(const float *Lt, int k, double sigma, double pivmin) {
  double d = (double)Lt[0] - sigma;
  if (fabs(d) < pivmin) d = -pivmin;
  int c = d < 0.0;
  for (int i = 1; i < k; ++i) {
    const float2 ab = reinterpret_cast<const float2 *>(Lt)[i];
    const double b = (double)ab.y;
    d = ((double)ab.x - sigma) - (b * b) / d;
    if (fabs(d) < pivmin) d = -pivmin;
    c += d < 0.0;
  }
  return c;
}

// Gershgorin interval of T (widened as LAPACK dstebz does) and pivmin
template <int GROUP>
__device__ __forceinline__ void t_bounds(const float *Lt, int k, int gl, double &lo, double &hi,
                                         double &pivmin) {
  double glo = INFINITY, ghi = -INFINITY, b2 = 0.0;
  if (gl < k) {
    const double al = Lt[2 * gl];
    const double bl = (gl > 0) ? fabs((double)Lt[2 * gl + 1]) : 0.0;
    const double br = (gl + 1 < k) ? fabs((double)Lt[2 * (gl + 1) + 1]) : 0.0;
    glo = al - bl - br;
    ghi = al + bl + br;
    b2 = bl * bl;
  }
  glo = g_min<GROUP>(glo);
  ghi = g_max<GROUP>(ghi);
  b2 = g_max<GROUP>(b2);
  pivmin = DBL_MIN * fmax(1.0, b2);
  const double tnorm = fmax(fabs(glo), fabs(ghi));
  const double fudge = 2.0 * DBL_EPSILON * tnorm * k + 4.0 * pivmin;
  lo = glo - fudge;
  hi = ghi + fudge;
}

// Stage (c), extreme eigenvalues: the group's lower half brackets lambda_min
// (index 0), the upper half lambda_max (index k-1). Each iteration every lane
// counts at one of HALF equispaced interior points; by monotonicity of the
// count the new bracket follows from a ballot popcount.
template <int GROUP>
__device__ __forceinline__ double extreme_eig(const float *Lt, int k, int gl, uint32_t &st) {
  constexpr int HALF = GROUP / 2;
  double lo, hi, pivmin;
  t_bounds<GROUP>(Lt, k, gl, lo, hi, pivmin);
  const int target = (gl < HALF) ? 0 : k - 1;
  const int p = gl % HALF;
  const int base = __lane_id() - p;
  int it = 0;
  for (; it < MS_MAXIT; ++it) {
    const double tol = ldexp(fmax(fabs(lo), fabs(hi)), -MS_TOLBITS) + 4.0 * pivmin;
    const bool done = (hi - lo) <= tol;
    if (g_all<GROUP>(done)) break;
    const double step = (hi - lo) / (double)(HALF + 1);
    const double sigma = lo + step * (double)(p + 1);
    const int c = negcount(Lt, k, sigma, pivmin);
    const uint64_t m = __ballot(c <= target);
    const int cnt = __popcll((m >> base) & ((1ull << HALF) - 1ull));
    if (!done) {
      const double nlo = (cnt > 0) ? lo + step * (double)cnt : lo;
      const double nhi = (cnt < HALF) ? lo + step * (double)(cnt + 1) : hi;
      lo = nlo;
      hi = nhi;
    }
  }
  if (it == MS_MAXIT) st |= ST_NOCONV;
  return lo + 0.5 * (hi - lo);
}

// Stage (c), all eigenvalues: lane t bisects eigenvalue t (ascending order,
// as TQLRAT returns them, ParaLanczos.cu:168-185)
template <int GROUP>
__device__ __forceinline__ double lane_eig(const float *Lt, int k, int gl, uint32_t &st) {
  double lo, hi, pivmin;
  t_bounds<GROUP>(Lt, k, gl, lo, hi, pivmin);
  bool conv = true;
  if (gl < k) {
    int it = 0;
    for (; it < BIS_MAXIT; ++it) {
      const double tol = ldexp(fmax(fabs(lo), fabs(hi)), -BIS_TOLBITS) + 4.0 * pivmin;
      if (hi - lo <= tol) break;
      const double mid = lo + 0.5 * (hi - lo);
      if (negcount(Lt, k, mid, pivmin) > gl) hi = mid;
      else lo = mid;
    }
    conv = it < BIS_MAXIT;
  }
  if (!g_all<GROUP>(conv)) st |= ST_NOCONV;
  return lo + 0.5 * (hi - lo);
}

template <int GROUP, int PAD, int MODE>
__global__ __launch_bounds__(BLOCK) void solve_kernel(SolveArgs a) {
  constexpr int SPB = BLOCK / GROUP;
  constexpr int STRIDE = GROUP + PAD;
  constexpr int REGION = GROUP * STRIDE;
  static_assert(REGION >= 5 * GROUP, "LDS region too small for the vector buffers");
  __shared__ __attribute__((aligned(16))) float lds[SPB * REGION];

  const int tid = threadIdx.x;
  const int gl = tid % GROUP;
  const int gib = tid / GROUP;
  float *R = lds + gib * REGION;

  const int64_t slot = (int64_t)blockIdx.x * SPB + gib;
  if (slot >= a.nslots) return;  // whole group leaves; nothing below syncs the block
  const int64_t s = a.order ? (int64_t)a.order[slot] : slot;
  int k;
  int64_t off;
  if (a.offsets) {
    off = a.offsets[s];
    k = (int)(a.offsets[s + 1] - off);
  } else {
    k = a.kuni;
    off = s * (int64_t)k;
  }
  uint32_t st = (k < 1 || k > GROUP) ? ST_BADK : 0u;
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
    const bool bad = (st & (ST_BADK | ST_BADINDEX | ST_NONFINITE)) != 0u;
    if (!bad) st |= lanczos_core<GROUP>(g, k, gl, R);
    const float *Lt = R + 3 * GROUP;
    if constexpr (MODE == MODE_EXTREME) {
      double lam = NAN;
      if (!(st & (ST_BADK | ST_BADINDEX | ST_NONFINITE))) lam = extreme_eig<GROUP>(Lt, k, gl, st);
      const double lmax = __shfl(lam, GROUP / 2, GROUP);
      if (gl == 0) {
        a.lmin[s] = (float)lam;
        a.lmax[s] = (float)lmax;
        a.delta[s] = (float)fmax(lmax - 1.0, 1.0 - lam);
        a.status[s] = st;
      }
    } else {  // MODE_ALLEIG
      double lam = NAN;
      if (!(st & (ST_BADK | ST_BADINDEX | ST_NONFINITE))) lam = lane_eig<GROUP>(Lt, k, gl, st);
      const double det = g_prod<GROUP>(gl < k ? lam : 1.0);
      if (gl < k) a.eig[off + gl] = lam;
      if (gl == 0) {
        a.det[s] = det;
        a.err[s] = (int32_t)st;
      }
    }
  }
}

using SolveFn = void (*)(SolveArgs);

template <int GROUP, int PAD>
struct Variant {
  static constexpr SolveFn fns[4] = {solve_kernel<GROUP, PAD, 0>, solve_kernel<GROUP, PAD, 1>,
                                     solve_kernel<GROUP, PAD, 2>, solve_kernel<GROUP, PAD, 3>};
};

static SolveFn pick(int group, int pad, int mode) {
#define PL_CASE(G, P) \
  if (group == G && pad == P) return Variant<G, P>::fns[mode];
  PL_CASE(8, 0) PL_CASE(8, 1) PL_CASE(8, 4)
  PL_CASE(16, 0) PL_CASE(16, 1) PL_CASE(16, 4)
  PL_CASE(32, 0) PL_CASE(32, 1) PL_CASE(32, 4)
  PL_CASE(64, 0) PL_CASE(64, 1) PL_CASE(64, 4)
#undef PL_CASE
  return nullptr;
}

bool solve_variant_exists(int group, int pad) { return pick(group, pad, 0) != nullptr; }

hipError_t launch_solve(int group, int pad, int mode, const SolveArgs &a, hipStream_t s) {
  if (mode < 0 || mode > 3) return hipErrorInvalidValue;
  SolveFn fn = pick(group, pad, mode);
  if (!fn) return hipErrorInvalidValue;
  if (a.nslots <= 0) return hipSuccess;
  const int spb = BLOCK / group;
  const int64_t blocks = (a.nslots + spb - 1) / spb;
  if (blocks > 0x7fffffffLL) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL(fn, dim3((unsigned)blocks), dim3(BLOCK), 0, s, a);
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
