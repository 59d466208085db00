// Internal interface between the C-ABI layer (pl_capi.hip) and the HIP
// kernels (pl_kernels.hip). Nothing here is exported from the library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pl {

// what a solve launch writes per subset
enum Mode : int {
  MODE_EXTREME = 0,   // lambda_min, lambda_max, delta, status (FP32 + u32)
  MODE_ALLEIG = 1,    // all k eigenvalues ascending (FP64), det, err
  MODE_DUMP = 2,      // the gathered k x k submatrix (FP32), uniform k only
  MODE_CHECKSUM = 3,  // gather stage only: sum of G_S entries (FP32)
};

struct SolveArgs {
  const float *G;          // n x n, row-major, leading dimension ld (floats)
  int64_t n, ld;
  const int32_t *idx;      // concatenated index sets
  const int64_t *offsets;  // nsub+1 offsets into idx, or nullptr (uniform k)
  int kuni;                // subset size when offsets == nullptr
  int64_t nslots;          // work items in this launch
  const int32_t *order;    // slot -> subset id, or nullptr (identity)
  float *lmin, *lmax, *delta;
  uint32_t *status;
  double *eig, *det;
  int32_t *err;
  float *gs_out;           // MODE_DUMP / MODE_CHECKSUM output
};

// lanes per subset (8, 16, 32, 64) x LDS row padding (0, 1, 4) x mode
hipError_t launch_solve(int group, int pad, int mode, const SolveArgs &a, hipStream_t s);
bool solve_variant_exists(int group, int pad);

hipError_t launch_gen_subsets(uint64_t seed, int64_t s0, int64_t nsub, const int64_t *offsets,
                              int kuni, int n, int32_t *idx, hipStream_t s);
hipError_t launch_gen_ks(uint64_t seed, int64_t s0, int64_t nsub, const int32_t *klist,
                         int nk, int32_t *ks, hipStream_t s);
hipError_t launch_gen_phi_gauss(uint64_t seed, int m, int n, float *phi, hipStream_t s);
hipError_t launch_gen_phi_bernoulli(uint64_t seed, int m, int n, float *phi, hipStream_t s);
hipError_t launch_gen_phi_dct(const int32_t *rows_dev, int m, int n, float *phi, hipStream_t s);

}  // namespace pl
