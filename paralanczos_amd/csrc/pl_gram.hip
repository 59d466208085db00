// Gram matrix construction on the device: G = Phi^T Phi with FP32 matrix
// cores (pl_gram). The reference builds Gram matrices element by element
// (DSG_B, DSG_B1, SG_B1, ParaLanczos.cu:364-438: one dot product over the
// frame's M rows per output entry); here the whole G is one symmetric rank-m
// update, tiled over the upper triangle and run on v_mfma_f32_32x32x2f32.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "../../include/paralanczos.h"
#include "pl_internal.hpp"

namespace pl {
namespace {

constexpr int GT = 128;        // G tile edge (rows and columns)
constexpr int GK = 32;         // rows of Phi per stage
constexpr int GSTR = GT + 32;  // LDS row stride of a staged Phi slice: the two
                               // rows one MFMA operand read spans start 32 banks apart
constexpr int GTHREADS = 256;  // four wavefronts, each a 64 x 64 quarter of the tile
constexpr int TSTR = 64 + 1;   // LDS row stride of a transposed 64 x 64 quarter

typedef float v16f __attribute__((ext_vector_type(16)));

struct GramArgs {
  const float *phi;  // m x n: entry (r, c) at phi[r ld + c] (row layout) or phi[c ld + r] (column layout)
  int64_t m, n, ld;
  float *G;          // n x n, row-major, leading dimension ldg
  int64_t ldg;
  int nt;            // tiles per edge, ceil(n / GT)
  int64_t ntri;      // tiles of the upper triangle, nt (nt + 1) / 2
};

// workgroup b -> upper-triangle tile (ti <= tj). Consecutive workgroups go to
// different XCDs (round robin: b on XCD b % 8), so the linear tile order is
// regrouped: XCD x takes the contiguous run of q + (x < nb % 8) tiles that
// starts at x q + min(x, nb % 8), q = nb / 8 - a bijection of [0, nb) - and
// its tiles share their Phi row slices in that XCD's L2.
__device__ __forceinline__ void gram_tile(const GramArgs &a, int &ti, int &tj) {
  const int64_t b = blockIdx.x, nb = gridDim.x;
  const int64_t q = nb / 8, rr = nb % 8, x = b % 8;
  const int64_t r = x * q + (x < rr ? x : rr) + b / 8;
  // row ti of the upper triangle starts at ti nt - ti (ti - 1) / 2
  int lo = 0, hi = a.nt - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    const int64_t start = (int64_t)mid * a.nt - (int64_t)mid * (mid - 1) / 2;
    if (start <= r) lo = mid; else hi = mid - 1;
  }
  ti = lo;
  tj = ti + (int)(r - ((int64_t)ti * a.nt - (int64_t)ti * (ti - 1) / 2));
}

// rows k0..k0+GK-1 of Phi, columns c0..c0+GT-1, into 16 registers per thread
// (zero outside the matrix), with consecutive lanes on consecutive addresses.
// Row layout: thread t holds columns c0 + 4 (t % 32) + 0..3 of rows
// k0 + t / 32 + 8u; column layout: column c0 + t / 2, rows k0 + 16 (t % 2) +
// 4u + 0..3. VEC: the four entries of each group are one 16-byte load (the
// caller checked alignment and that a group is all in or all out).
template <bool COL, bool VEC>
__device__ __forceinline__ void gram_fetch(const GramArgs &a, int64_t k0, int64_t c0, float (&r)[16]) {
  const int t = threadIdx.x;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    int64_t k, c;
    if constexpr (!COL) {
      k = k0 + t / 32 + 8 * u;
      c = c0 + 4 * (t % 32);
    } else {
      k = k0 + 16 * (t & 1) + 4 * u;
      c = c0 + (t >> 1);
    }
    if constexpr (VEC) {
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (k < a.m && c < a.n)
        v = *reinterpret_cast<const float4 *>(a.phi + (COL ? c * a.ld + k : k * a.ld + c));
      r[4 * u + 0] = v.x;
      r[4 * u + 1] = v.y;
      r[4 * u + 2] = v.z;
      r[4 * u + 3] = v.w;
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t kk = COL ? k + e : k, cc = COL ? c : c + e;
        r[4 * u + e] = (kk < a.m && cc < a.n) ? a.phi[COL ? cc * a.ld + kk : kk * a.ld + cc] : 0.f;
      }
    }
  }
}

// the fetched values into an LDS slice S[row k][column c] (stride GSTR)
template <bool COL>
__device__ __forceinline__ void gram_stage(float *S, const float (&r)[16]) {
  const int t = threadIdx.x;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    if constexpr (!COL) {
      *reinterpret_cast<float4 *>(S + (t / 32 + 8 * u) * GSTR + 4 * (t % 32)) =
          make_float4(r[4 * u], r[4 * u + 1], r[4 * u + 2], r[4 * u + 3]);
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) S[(16 * (t & 1) + 4 * u + e) * GSTR + (t >> 1)] = r[4 * u + e];
    }
  }
}

template <bool COL, bool VEC>
__global__ __launch_bounds__(GTHREADS) void gram_kernel(GramArgs a) {
  // two buffers of two Phi slices (the tile's row and column blocks); after
  // the K loop the same LDS holds the transposed quarters for the mirror
  __shared__ __attribute__((aligned(16))) float lds[4 * 64 * TSTR > 4 * GK * GSTR ? 4 * 64 * TSTR
                                                                                   : 4 * GK * GSTR];
  int ti, tj;
  gram_tile(a, ti, tj);
  const bool diag = ti == tj;
  const int64_t i0 = (int64_t)ti * GT, j0 = (int64_t)tj * GT;
  const int t = threadIdx.x, l = t & 63, w = t >> 6, wm = w >> 1, wn = w & 1;
  // buffer b: row-block slice at lds + 2b GK GSTR, column-block slice after it
  auto A = [&](int b) { return lds + 2 * b * GK * GSTR; };
  auto B = [&](int b) { return lds + (2 * b + 1) * GK * GSTR; };

  v16f acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[x][y][e] = 0.f;

  float ra[16], rb[16];
  gram_fetch<COL, VEC>(a, 0, i0, ra);
  if (!diag) gram_fetch<COL, VEC>(a, 0, j0, rb);
  gram_stage<COL>(A(0), ra);
  if (!diag) gram_stage<COL>(B(0), rb);
  __syncthreads();
  int buf = 0;
  for (int64_t k0 = 0; k0 < a.m; k0 += GK) {
    const bool more = k0 + GK < a.m;
    if (more) {
      gram_fetch<COL, VEC>(a, k0 + GK, i0, ra);
      if (!diag) gram_fetch<COL, VEC>(a, k0 + GK, j0, rb);
    }
    const float *As = A(buf);
    const float *Bs = diag ? A(buf) : B(buf);
    // v_mfma_f32_32x32x2f32: lane l supplies A[l % 32][l / 32] and B[l / 32][l % 32]
#pragma unroll
    for (int kk = 0; kk < GK; kk += 2) {
      const int row = (kk + (l >> 5)) * GSTR;
      float av[2], bv[2];
#pragma unroll
      for (int x = 0; x < 2; ++x) av[x] = As[row + 64 * wm + 32 * x + (l & 31)];
#pragma unroll
      for (int y = 0; y < 2; ++y) bv[y] = Bs[row + 64 * wn + 32 * y + (l & 31)];
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
          acc[x][y] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[x], bv[y], acc[x][y], 0, 0, 0);
    }
    if (more) {
      gram_stage<COL>(A(buf ^ 1), ra);
      if (!diag) gram_stage<COL>(B(buf ^ 1), rb);
    }
    __syncthreads();
    buf ^= 1;
  }

  // epilogue. Accumulator e of block (x, y) on lane l is entry
  // (32x + 8(e / 4) + 4(l / 32) + e % 4, 32y + l % 32) of this wavefront's
  // 64 x 64 quarter. Direct stores are row segments (coalesced); the mirror
  // G[j][i] goes through LDS so that its stores are row segments too. On a
  // diagonal tile only entries with i <= j are stored directly and i < j
  // mirrored, so G is exactly symmetric.
  const int64_t qi = i0 + 64 * wm, qj = j0 + 64 * wn;  // quarter origin
  const bool skip = diag && wm > wn;                  // below the diagonal: the mirror covers it
  float *T = lds + w * 64 * TSTR;
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int r = 32 * x + 8 * (e >> 2) + 4 * (l >> 5) + (e & 3), c = 32 * y + (l & 31);
        const int64_t gi = qi + r, gj = qj + c;
        const float v = acc[x][y][e];
        if (!skip && gi < a.n && gj < a.n && (!diag || gi <= gj)) a.G[gi * a.ldg + gj] = v;
        T[c * TSTR + r] = v;  // transposed quarter: row c holds column c of the block
      }
  __syncthreads();
  if (skip) return;
  // mirror: G[qj + c][qi + r] = quarter[r][c], one row c per step, lanes over r
  for (int c = 0; c < 64; ++c) {
    const int64_t gi = qj + c;
    if (gi >= a.n) break;
    const int r = l;
    const int64_t gj = qi + r;
    if (gj < a.n && (!diag || gj < gi)) a.G[gi * a.ldg + gj] = T[c * TSTR + r];
  }
}

}  // namespace
}  // namespace pl

extern "C" {

pl_status pl_gram(const float *phi, int64_t m, int64_t n, int64_t ld, int layout, float *G,
                  int64_t ldg, void *stream) {
  if (!phi || !G) return pl::set_error(PL_ERR_ARG, "phi and G must be non-null");
  if (m < 1 || n < 1) return pl::set_error(PL_ERR_ARG, "m and n must be >= 1");
  if (layout != PL_LAYOUT_ROW && layout != PL_LAYOUT_COL)
    return pl::set_error(PL_ERR_ARG, "layout must be PL_LAYOUT_ROW or PL_LAYOUT_COL");
  if (ld < (layout == PL_LAYOUT_ROW ? n : m))
    return pl::set_error(PL_ERR_ARG, "ld below the row (row layout) or column (column layout) length");
  if (ldg < n) return pl::set_error(PL_ERR_ARG, "ldg < n");
  pl::GramArgs a{phi, m, n, ld, G, ldg, 0, 0};
  a.nt = (int)std::min<int64_t>((n + pl::GT - 1) / pl::GT, 0x7fffffff);
  a.ntri = (int64_t)a.nt * (a.nt + 1) / 2;
  if (a.ntri > 0x7fffffffLL) return pl::set_error(PL_ERR_UNSUPPORTED, "n too large");
  // 16-byte loads when every group of four entries is aligned and lies all
  // inside or all outside the frame
  const bool vec = ((uintptr_t)phi % 16 == 0) && ld % 4 == 0 &&
                   (layout == PL_LAYOUT_ROW ? n % 4 == 0 : m % 4 == 0);
  const dim3 grid((unsigned)a.ntri), block(pl::GTHREADS);
  hipStream_t s = (hipStream_t)stream;
  if (layout == PL_LAYOUT_ROW) {
    if (vec) hipLaunchKernelGGL((pl::gram_kernel<false, true>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((pl::gram_kernel<false, false>), grid, block, 0, s, a);
  } else {
    if (vec) hipLaunchKernelGGL((pl::gram_kernel<true, true>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((pl::gram_kernel<true, false>), grid, block, 0, s, a);
  }
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? PL_OK : pl::set_hip_error(e, "gram launch");
}

}  // extern "C"
