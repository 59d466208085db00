// Internal interface between the C-ABI layer (pl_capi.hip) and the HIP
// kernels (pl_kernels.hip). Nothing here is exported from the library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pl {

// what a solve launch writes per subset
enum Mode : int {
  MODE_TRIDIAG = 0,   // stages (a)+(b): T (FP32 alpha, beta), its bounds, status
  MODE_DUMP = 1,      // stage (a) only: the gathered k x k submatrix (FP32), uniform k
  MODE_CHECKSUM = 2,  // stage (a) only: sum of the G_S entries (FP32)
};

struct SolveArgs {
  const float *G;          // n x n, row-major, leading dimension ld (floats)
  int64_t n, ld;
  const int32_t *idx;      // concatenated index sets
  const int64_t *offsets;  // nsub+1 offsets into idx, or nullptr (uniform k)
  int kuni;                // subset size when offsets == nullptr
  int64_t nslots;          // work items in this launch
  const int32_t *order;    // slot -> subset id, or nullptr (identity)
  uint32_t *status;        // per-subset status word
  float2 *tri;             // T of subset s at tri + s*tstride, (alpha_i, beta_i)
  int tstride;
  double4 *bounds;         // (Gershgorin lo, hi, min alpha, max alpha) per subset
  float *gs_out;           // MODE_DUMP / MODE_CHECKSUM output
};

struct BisectArgs {
  const float2 *tri;
  int tstride;
  const double4 *bounds;
  const int64_t *offsets;  // as in SolveArgs
  int kuni;
  int64_t nsub;
  uint32_t *status;        // read (fatal bits) and updated (NOCONV)
  float *lmin, *lmax, *delta;
  double *eig, *det;
};

// workspace of a solve: T entries, then (256-byte aligned) per-subset bounds
inline size_t tridiag_bytes(int64_t nsub, int kstride) {
  return ((size_t)nsub * (size_t)kstride * sizeof(float2) + 255) & ~(size_t)255;
}
inline size_t workspace_bytes(int64_t nsub, int kstride) {
  return tridiag_bytes(nsub, kstride) + (size_t)nsub * sizeof(double4);
}

// lanes per subset (8, 16, 32, 64) x LDS row padding (0, 1, 4) x mode
hipError_t launch_solve(int group, int pad, int mode, const SolveArgs &a, hipStream_t s);
bool solve_variant_exists(int group, int pad);
// mode 0: lambda_min/lambda_max/delta; mode 1: every eigenvalue + det (uniform k)
hipError_t launch_bisect(int mode, const BisectArgs &a, hipStream_t s);

hipError_t launch_gen_subsets(uint64_t seed, int64_t s0, int64_t nsub, const int64_t *offsets,
                              int kuni, int n, int32_t *idx, hipStream_t s);
hipError_t launch_gen_ks(uint64_t seed, int64_t s0, int64_t nsub, const int32_t *klist,
                         int nk, int32_t *ks, hipStream_t s);
hipError_t launch_gen_phi_gauss(uint64_t seed, int m, int n, float *phi, hipStream_t s);
hipError_t launch_gen_phi_bernoulli(uint64_t seed, int m, int n, float *phi, hipStream_t s);
hipError_t launch_gen_phi_dct(const int32_t *rows_dev, int m, int n, float *phi, hipStream_t s);

}  // namespace pl
