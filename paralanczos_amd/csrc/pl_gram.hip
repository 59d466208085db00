// Gram matrices on the device with FP32 matrix cores:
//   pl_gram:    G = Phi^T Phi, the whole n x n Gram matrix of a frame;
//   pl_subgram: G_S = Phi_S^T Phi_S for a batch of index sets S (the frame's
//               columns S), straight from the frame.
// The reference builds Gram matrices entry by entry (DSG_B, DSG_B1, SG_B1,
// ParaLanczos.cu:364-438: one dot product over the frame's M rows per output
// entry, DSG_B for the columns of one subset). Here each is a symmetric
// rank-m update, tiled over the upper triangle and run on
// v_mfma_f32_32x32x2f32; the sub-Gram form reads the subset's columns
// through its index set.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "../../include/paralanczos.h"
#include "pl_common.hpp"
#include "pl_internal.hpp"

namespace pl {
namespace {

constexpr int GK = 32;         // rows of Phi per stage
constexpr int GTHREADS = 256;  // four wavefronts, each one quarter of the tile

typedef float v16f __attribute__((ext_vector_type(16)));

struct GramArgs {
  const float *phi;      // frame: entry (r, c) at phi[r ld + c] (row layout) or phi[c ld + r] (column layout)
  int64_t m, ncols, ld;  // frame rows, frame columns, leading dimension
  const int32_t *cols;   // logical column c of output s is frame column cols[s n + c]; nullptr: c itself
  int64_t n;             // edge of each output
  float *G;              // output s at G + s gstride: n x n, row-major, leading dimension ldg
  int64_t ldg, gstride;
  uint32_t *status;      // per-output status (index sets only), or nullptr
  int64_t s0;            // first output of this launch (grid y indexes s0 + y)
  int nt;                // tiles per edge, ceil(n / GT)
  int64_t ntri;          // tiles of the upper triangle, nt (nt + 1) / 2
};

// layout constants of a GT x GT tile
template <int GT>
struct GramTile {
  static constexpr int QS = GT / 2;      // quarter edge: one wavefront's part
  static constexpr int NB = QS / 32;     // 32 x 32 MFMA blocks per quarter edge
  static constexpr int STR = GT + 32;    // LDS row stride of a staged slice: the two rows one
                                         // MFMA operand read spans start 32 banks apart
  static constexpr int NG = GT / 32;     // groups of four entries each thread fetches per slice
  static constexpr int TPR = GT / 4;     // row layout: threads per slice row (four columns each)
  static constexpr int RPP = GTHREADS / TPR;  // row layout: slice rows per pass
  static constexpr int TPC = GTHREADS / GT;   // column layout: threads per column
  static constexpr int RPT = GK / TPC;   // column layout: consecutive rows per thread
  static constexpr int TSTR = QS + 1;    // LDS row stride of a transposed quarter
  static constexpr int LDS = (4 * QS * TSTR > 4 * GK * STR) ? 4 * QS * TSTR : 4 * GK * STR;
  static_assert(NG * 4 * GTHREADS == GK * GT, "one slice per fetch");
};

// workgroup x -> upper-triangle tile (ti <= tj). Consecutive workgroups go to
// different XCDs (round robin: x on XCD x % 8), so the linear tile order is
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

// rows k0..k0+GK-1 of Phi, logical columns c0..c0+GT-1 of output s, into
// registers (zero outside the frame and past n), with consecutive lanes on
// consecutive addresses. Row layout: thread t holds logical columns
// c0 + 4 (t % TPR) + 0..3 of rows k0 + t / TPR + RPP u; column layout:
// logical column c0 + t / TPC, rows k0 + RPT (t % TPC) + 4u + 0..3. VEC: the
// four entries of a group are one 16-byte load (the caller checked alignment
// and that a group is all in or all out; in the row layout only without an
// index set, whose columns need not be adjacent).
template <int GT, bool COL, bool VEC, bool IND>
__device__ __forceinline__ void gram_fetch(const GramArgs &a, int64_t s, int64_t k0, int64_t c0,
                                           float (&r)[4 * GramTile<GT>::NG]) {
  using P = GramTile<GT>;
  const int t = threadIdx.x;
#pragma unroll
  for (int u = 0; u < P::NG; ++u) {
    int64_t k, c;
    if constexpr (!COL) {
      k = k0 + t / P::TPR + P::RPP * u;
      c = c0 + 4 * (t % P::TPR);
    } else {
      k = k0 + P::RPT * (t % P::TPC) + 4 * u;
      c = c0 + t / P::TPC;
    }
    if constexpr (VEC) {
      static_assert(COL || !IND, "row-layout 16-byte loads need adjacent columns");
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (k < a.m && c < a.n) {
        const int64_t fc = IND ? (int64_t)a.cols[s * a.n + c] : c;
        v = *reinterpret_cast<const float4 *>(a.phi + (COL ? fc * a.ld + k : k * a.ld + fc));
      }
      r[4 * u + 0] = v.x;
      r[4 * u + 1] = v.y;
      r[4 * u + 2] = v.z;
      r[4 * u + 3] = v.w;
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t kk = COL ? k + e : k, cc = COL ? c : c + e;
        float x = 0.f;
        if (kk < a.m && cc < a.n) {
          const int64_t fc = IND ? (int64_t)a.cols[s * a.n + cc] : cc;
          x = a.phi[COL ? fc * a.ld + kk : kk * a.ld + fc];
        }
        r[4 * u + e] = x;
      }
    }
  }
}

// the fetched values into an LDS slice S[row k][column c] (stride STR)
template <int GT, bool COL>
__device__ __forceinline__ void gram_stage(float *S, const float (&r)[4 * GramTile<GT>::NG]) {
  using P = GramTile<GT>;
  const int t = threadIdx.x;
#pragma unroll
  for (int u = 0; u < P::NG; ++u) {
    if constexpr (!COL) {
      *reinterpret_cast<float4 *>(S + (t / P::TPR + P::RPP * u) * P::STR + 4 * (t % P::TPR)) =
          make_float4(r[4 * u], r[4 * u + 1], r[4 * u + 2], r[4 * u + 3]);
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e)
        S[(P::RPT * (t % P::TPC) + 4 * u + e) * P::STR + t / P::TPC] = r[4 * u + e];
    }
  }
}

template <int GT, bool COL, bool VEC, bool IND>
__global__ __launch_bounds__(GTHREADS) void gram_kernel(GramArgs a) {
  using P = GramTile<GT>;
  // two buffers of two Phi slices (the tile's row and column blocks); after
  // the K loop the same LDS holds the transposed quarters for the mirror
  __shared__ __attribute__((aligned(16))) float lds[P::LDS];
  const int t = threadIdx.x, l = t & 63, w = t >> 6, wm = w >> 1, wn = w & 1;
  const int64_t s = a.s0 + blockIdx.y;
  if constexpr (IND) {
    // an index outside the frame flags the output and leaves it unwritten;
    // the block-wide flag goes through the first word of the (not yet staged) LDS
    bool bad = false;
    for (int64_t c = t; c < a.n; c += GTHREADS) {
      const int32_t v = a.cols[s * a.n + c];
      bad = bad || v < 0 || (int64_t)v >= a.ncols;
    }
    int *flag = reinterpret_cast<int *>(lds);
    if (t == 0) *flag = 0;
    __syncthreads();
    if (bad) *flag = 1;
    __syncthreads();
    const bool any = *flag != 0;
    __syncthreads();
    if (t == 0 && blockIdx.x == 0 && a.status) a.status[s] = any ? ST_BADINDEX : 0u;
    if (any) return;
  }
  int ti, tj;
  gram_tile(a, ti, tj);
  const bool diag = ti == tj;
  const int64_t i0 = (int64_t)ti * GT, j0 = (int64_t)tj * GT;
  float *G = a.G + s * a.gstride;
  // buffer b: row-block slice at lds + 2b GK STR, column-block slice after it
  auto A = [&](int b) { return lds + 2 * b * GK * P::STR; };
  auto B = [&](int b) { return lds + (2 * b + 1) * GK * P::STR; };

  v16f acc[P::NB][P::NB];
#pragma unroll
  for (int x = 0; x < P::NB; ++x)
#pragma unroll
    for (int y = 0; y < P::NB; ++y)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[x][y][e] = 0.f;

  float ra[4 * P::NG], rb[4 * P::NG];
  gram_fetch<GT, COL, VEC, IND>(a, s, 0, i0, ra);
  if (!diag) gram_fetch<GT, COL, VEC, IND>(a, s, 0, j0, rb);
  gram_stage<GT, COL>(A(0), ra);
  if (!diag) gram_stage<GT, COL>(B(0), rb);
  __syncthreads();
  int buf = 0;
  for (int64_t k0 = 0; k0 < a.m; k0 += GK) {
    const bool more = k0 + GK < a.m;
    if (more) {
      gram_fetch<GT, COL, VEC, IND>(a, s, k0 + GK, i0, ra);
      if (!diag) gram_fetch<GT, COL, VEC, IND>(a, s, k0 + GK, j0, rb);
    }
    const float *As = A(buf);
    const float *Bs = diag ? A(buf) : B(buf);
    // v_mfma_f32_32x32x2f32: lane l supplies A[l % 32][l / 32] and B[l / 32][l % 32]
#pragma unroll
    for (int kk = 0; kk < GK; kk += 2) {
      const int row = (kk + (l >> 5)) * P::STR;
      float av[P::NB], bv[P::NB];
#pragma unroll
      for (int x = 0; x < P::NB; ++x) av[x] = As[row + P::QS * wm + 32 * x + (l & 31)];
#pragma unroll
      for (int y = 0; y < P::NB; ++y) bv[y] = Bs[row + P::QS * wn + 32 * y + (l & 31)];
#pragma unroll
      for (int x = 0; x < P::NB; ++x)
#pragma unroll
        for (int y = 0; y < P::NB; ++y)
          acc[x][y] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[x], bv[y], acc[x][y], 0, 0, 0);
    }
    if (more) {
      gram_stage<GT, COL>(A(buf ^ 1), ra);
      if (!diag) gram_stage<GT, COL>(B(buf ^ 1), rb);
    }
    __syncthreads();
    buf ^= 1;
  }

  // epilogue. Accumulator e of block (x, y) on lane l is entry
  // (32x + 8(e / 4) + 4(l / 32) + e % 4, 32y + l % 32) of this wavefront's
  // QS x QS quarter. Direct stores are row segments (coalesced); the mirror
  // G[j][i] goes through LDS so that its stores are row segments too. On a
  // diagonal tile only entries with i <= j are stored directly and i < j
  // mirrored, so G is exactly symmetric.
  const int64_t qi = i0 + P::QS * wm, qj = j0 + P::QS * wn;  // quarter origin
  const bool skip = diag && wm > wn;  // below the diagonal: the mirror covers it
  float *T = lds + w * P::QS * P::TSTR;
#pragma unroll
  for (int x = 0; x < P::NB; ++x)
#pragma unroll
    for (int y = 0; y < P::NB; ++y)
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int r = 32 * x + 8 * (e >> 2) + 4 * (l >> 5) + (e & 3), c = 32 * y + (l & 31);
        const int64_t gi = qi + r, gj = qj + c;
        const float v = acc[x][y][e];
        if (!skip && gi < a.n && gj < a.n && (!diag || gi <= gj)) G[gi * a.ldg + gj] = v;
        T[c * P::TSTR + r] = v;  // transposed quarter: row c holds column c of the block
      }
  __syncthreads();
  if (skip || l >= P::QS) return;
  // mirror: G[qj + c][qi + r] = quarter[r][c], one row c per step, lanes over r
  for (int c = 0; c < P::QS; ++c) {
    const int64_t gi = qj + c;
    if (gi >= a.n) break;
    const int r = l;
    const int64_t gj = qi + r;
    if (gj < a.n && (!diag || gj < gi)) G[gi * a.ldg + gj] = T[c * P::TSTR + r];
  }
}

template <int GT, bool COL, bool IND>
hipError_t gram_launch(GramArgs a, bool vec, int64_t nout, hipStream_t s) {
  a.nt = (int)((a.n + GT - 1) / GT);
  a.ntri = (int64_t)a.nt * (a.nt + 1) / 2;
  // the grid's y extent is kept within 65535 outputs per launch
  for (int64_t s0 = 0; s0 < nout; s0 += 65535) {
    a.s0 = s0;
    const dim3 grid((unsigned)a.ntri, (unsigned)std::min<int64_t>(65535, nout - s0));
    if (vec && (COL || !IND))
      hipLaunchKernelGGL((gram_kernel<GT, COL, COL || !IND, IND>), grid, dim3(GTHREADS), 0, s, a);
    else
      hipLaunchKernelGGL((gram_kernel<GT, COL, false, IND>), grid, dim3(GTHREADS), 0, s, a);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// 16-byte loads when every group of four entries is aligned and lies all
// inside or all outside the frame
bool gram_vec(const float *phi, int64_t m, int64_t n, int64_t ld, int layout) {
  return ((uintptr_t)phi % 16 == 0) && ld % 4 == 0 &&
         (layout == PL_LAYOUT_ROW ? n % 4 == 0 : m % 4 == 0);
}

pl_status check_frame(const float *phi, int64_t m, int64_t n, int64_t ld, int layout) {
  if (!phi) return set_error(PL_ERR_ARG, "phi must be non-null");
  if (m < 1 || n < 1) return set_error(PL_ERR_ARG, "m and n must be >= 1");
  if (n > 0x7fffffffLL) return set_error(PL_ERR_UNSUPPORTED, "n >= 2^31");
  if (layout != PL_LAYOUT_ROW && layout != PL_LAYOUT_COL)
    return set_error(PL_ERR_ARG, "layout must be PL_LAYOUT_ROW or PL_LAYOUT_COL");
  if (ld < (layout == PL_LAYOUT_ROW ? n : m))
    return set_error(PL_ERR_ARG, "ld below the row (row layout) or column (column layout) length");
  return PL_OK;
}

}  // namespace
}  // namespace pl

extern "C" {

pl_status pl_gram(const float *phi, int64_t m, int64_t n, int64_t ld, int layout, float *G,
                  int64_t ldg, void *stream) {
  pl_status st = pl::check_frame(phi, m, n, ld, layout);
  if (st != PL_OK) return st;
  if (!G) return pl::set_error(PL_ERR_ARG, "G must be non-null");
  if (ldg < n) return pl::set_error(PL_ERR_ARG, "ldg < n");
  const int64_t nt = (n + 127) / 128;
  if (nt * (nt + 1) / 2 > 0x7fffffffLL) return pl::set_error(PL_ERR_UNSUPPORTED, "n too large");
  pl::GramArgs a{};
  a.phi = phi;
  a.m = m;
  a.ncols = n;
  a.ld = ld;
  a.n = n;
  a.G = G;
  a.ldg = ldg;
  const bool vec = pl::gram_vec(phi, m, n, ld, layout);
  hipStream_t s = (hipStream_t)stream;
  const hipError_t e = layout == PL_LAYOUT_ROW ? pl::gram_launch<128, false, false>(a, vec, 1, s)
                                               : pl::gram_launch<128, true, false>(a, vec, 1, s);
  return e == hipSuccess ? PL_OK : pl::set_hip_error(e, "gram launch");
}

pl_status pl_subgram(const float *phi, int64_t m, int64_t n, int64_t ld, int layout,
                     const int32_t *idx, int k, int64_t nsub, float *out, uint32_t *status,
                     void *stream) {
  pl_status st = pl::check_frame(phi, m, n, ld, layout);
  if (st != PL_OK) return st;
  if (!idx || !out) return pl::set_error(PL_ERR_ARG, "idx and out must be non-null");
  if (k < 1 || k > pl::KMAX_ALL) return pl::set_error(PL_ERR_UNSUPPORTED, "k outside [1, 256]");
  if (nsub < 0) return pl::set_error(PL_ERR_ARG, "nsub < 0");
  if (nsub == 0) return PL_OK;
  pl::GramArgs a{};
  a.phi = phi;
  a.m = m;
  a.ncols = n;
  a.ld = ld;
  a.cols = idx;
  a.n = k;
  a.G = out;
  a.ldg = k;
  a.gstride = (int64_t)k * k;
  a.status = status;
  const bool vec = pl::gram_vec(phi, m, n, ld, layout);
  hipStream_t s = (hipStream_t)stream;
  hipError_t e;
  if (layout == PL_LAYOUT_ROW)
    e = k <= 64 ? pl::gram_launch<64, false, true>(a, vec, nsub, s)
                : pl::gram_launch<128, false, true>(a, vec, nsub, s);
  else
    e = k <= 64 ? pl::gram_launch<64, true, true>(a, vec, nsub, s)
                : pl::gram_launch<128, true, true>(a, vec, nsub, s);
  return e == hipSuccess ? PL_OK : pl::set_hip_error(e, "subgram launch");
}

}  // extern "C"
