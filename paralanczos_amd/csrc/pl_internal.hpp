// Internal interface between the C-ABI layer (pl_capi.hip) and the HIP
// kernels (pl_kernels.hip). Nothing here is exported from the library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/paralanczos.h"

namespace pl {

// what a solve launch writes per subset
enum Mode : int {
  MODE_TRIDIAG = 0,   // stages (a)+(b): T (FP32 alpha, beta), its bounds, status
  MODE_DUMP = 1,      // stage (a) only: the gathered k x k submatrix (FP32), uniform k
  MODE_CHECKSUM = 2,  // stage (a) only: sum of the G_S entries (FP32)
};

// Subset binning. Subsets are grouped by size k into bins (one bin per solve
// variant, or one per exact k); the bin table in device memory holds, per bin,
// its count, the start of its run in the order array and a fill cursor.
constexpr int KMAX_ALL = 256;        // largest k any kernel accepts
constexpr int KLUT = KMAX_ALL + 2;   // lut index: 0 = k < 1, 1..256 = k, 257 = k > 256
constexpr int MAXBIN = 272;
constexpr int BIN_COUNT = 0, BIN_START = MAXBIN, BIN_FILL = 2 * MAXBIN;
struct BinLut {
  uint16_t bin[KLUT];
};
__host__ __device__ inline int lut_index(int64_t k) { return k < 1 ? 0 : k > KMAX_ALL ? KMAX_ALL + 1 : (int)k; }

struct SolveArgs {
  const float *G;          // n x n, row-major, leading dimension ld (floats)
  int64_t n, ld;
  const int32_t *idx;      // concatenated index sets
  const int64_t *offsets;  // nsub+1 offsets into idx, or nullptr (uniform k)
  int kuni;                // subset size when offsets == nullptr
  int kcap;                // largest valid k (the T stride); larger subsets get BADK
  int64_t nslots;          // work items in this launch (ignored when bins != nullptr)
  const int32_t *order;    // slot -> subset id, or nullptr (identity)
  const int32_t *bins;     // device bin table: when set, the launch drains bin `bin`
  int bin;                 //   (slots 0..count-1 of order + start) with a grid-stride loop
  uint32_t *status;        // per-subset status word
  float2 *tri;             // T of subset s at tri + s*tstride, (alpha_i, beta_i)
  int tstride;
  double4 *bounds;         // (Gershgorin lo, hi, min alpha, max alpha) per subset
  float *gs_out;           // MODE_DUMP / MODE_CHECKSUM output
};

struct BisectArgs {
  const int32_t *order;    // slot -> subset id, or nullptr (identity)
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

// workspace of a solve, each part 256-byte aligned: T entries, per-subset
// bounds, the bin table and the order array
inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
inline size_t tridiag_bytes(int64_t nsub, int kstride) {
  return align256((size_t)nsub * (size_t)kstride * sizeof(float2));
}
inline size_t bounds_bytes(int64_t nsub) { return align256((size_t)nsub * sizeof(double4)); }
inline size_t bintable_bytes() { return align256(3 * MAXBIN * sizeof(int32_t)); }
inline size_t workspace_bytes(int64_t nsub, int kstride) {
  return tridiag_bytes(nsub, kstride) + bounds_bytes(nsub) + bintable_bytes() +
         align256((size_t)nsub * sizeof(int32_t));
}
struct WorkspaceParts {
  float2 *tri;
  double4 *bounds;
  int32_t *table;
  int32_t *order;
};
inline WorkspaceParts split_workspace(void *p, int64_t nsub, int kstride) {
  char *c = (char *)p;
  WorkspaceParts w;
  w.tri = (float2 *)c;
  c += tridiag_bytes(nsub, kstride);
  w.bounds = (double4 *)c;
  c += bounds_bytes(nsub);
  w.table = (int32_t *)c;
  c += bintable_bytes();
  w.order = (int32_t *)c;
  return w;
}

// lanes per subset (8, 16, 32, 64, 128; 256 and 512 for the half-row kernels
// of capacity 128 and 256) x LDS row padding (0, 1, 4) x mode.
// With a.bins == nullptr the grid covers a.nslots slots; otherwise the launch
// is persistent (grid = resident workgroups, capped by `max_slots`) and drains
// bin a.bin of the table.
// f64 selects FP64 Gram-Schmidt coefficients (128 lanes and more).
hipError_t launch_solve(int group, int pad, int mode, const SolveArgs &a, hipStream_t s,
                        int64_t max_slots = 0, bool f64 = false);
bool solve_variant_exists(int group, int pad);
// one 64-lane wavefront per subset at any k <= 256, G_S and the basis in
// `scratch` (wave_scratch_bytes(slots)); `slots` resident wavefronts
constexpr int64_t kWaveSlots = 2048;
inline size_t wave_scratch_bytes(int64_t slots) {
  return (size_t)slots * 2 * 256 * 256 * sizeof(float);
}
hipError_t launch_wave(const SolveArgs &a, float *scratch, int64_t slots, hipStream_t s);
// power-iteration estimate of lambda_max per subset (iters steps from
// ones/sqrt(k)) into upper[s], status into a.status; kernel capacity by a.kcap
hipError_t launch_upper(const SolveArgs &a, int iters, float *upper, hipStream_t s);
// largest k a variant with this many lanes per subset takes
int variant_capacity(int group);

// bin the nsub subsets (sizes from offsets, or kuni) by lut into table/order:
// clear, count, scan, scatter - four stream-ordered launches, no host sync
hipError_t launch_binning(const int64_t *offsets, int kuni, int64_t nsub, const BinLut &lut,
                          int nbins, int32_t *table, int32_t *order, hipStream_t s);
// mode 0: lambda_min/lambda_max/delta; mode 1: every eigenvalue + det (uniform k)
hipError_t launch_bisect(int mode, const BisectArgs &a, hipStream_t s);

// per-size RIP statistics of a solved batch into out[(kmax+1) x 8] (FP64);
// ext is device scratch of (kmax+1) x 3 unsigned
hipError_t launch_rip_stats(const float *lmin, const float *lmax, const float *delta,
                            const uint32_t *status, const int64_t *offsets, int kuni,
                            int64_t nsub, int kmax, double *out, unsigned *ext, hipStream_t s);

hipError_t launch_gen_subsets(uint64_t seed, int64_t s0, int64_t nsub, const int64_t *offsets,
                              int kuni, int n, int32_t *idx, hipStream_t s);
hipError_t launch_gen_ks(uint64_t seed, int64_t s0, int64_t nsub, const int32_t *klist,
                         int nk, int32_t *ks, hipStream_t s);
hipError_t launch_gen_phi_gauss(uint64_t seed, int m, int n, float *phi, hipStream_t s);
hipError_t launch_gen_phi_bernoulli(uint64_t seed, int m, int n, float *phi, hipStream_t s);
hipError_t launch_gen_phi_dct(const int32_t *rows_dev, int m, int n, float *phi, hipStream_t s);

// ---- host-side solve jobs (pl_capi.hip), shared with the serving engine ----
struct Variant {
  int lanes, pad;
  bool f64 = false;   // FP64 Gram-Schmidt coefficients
  bool wave = false;  // one 64-lane wavefront beyond k = 64 (global scratch)
};

// everything one solve call needs, bound to device buffers
struct SolveJob {
  const float *G;
  int64_t n, ld;
  const int32_t *idx;
  const int64_t *offsets;
  int k;  // uniform size, or the largest size of a mixed batch
  int64_t nsub;
  float *lmin, *lmax, *delta;
  uint32_t *status;
  int form;           // PL_FORM_TUNED / PL_FORM_PLAIN (| precision flags), or -1 for an
                      // explicit variant
  Variant explicit_v;  // used when form == -1
  // tuned mixed batches: the bins other than the widest on a helper stream,
  // so that they fill the wide kernel's tail
  bool overlap_bins = true;
  // one-wavefront plain form: a scratch of wave_scratch_bytes(wave_slots)
  // owned by the caller (a plan); nullptr = allocated on the stream per launch
  float *wave_scratch = nullptr;
  int64_t wave_slots = 0;
};

// the stream-ordered launches of one solve on a workspace of
// workspace_bytes(nsub, k) bytes (no host synchronisation except for the
// plain form of a mixed batch)
pl_status solve_job(const SolveJob &j, void *ws, hipStream_t s);
pl_status validate_job(const SolveJob &j);
pl_status set_error(pl_status st, const char *msg);
pl_status set_hip_error(hipError_t e, const char *what);
// LDS bytes the variant for size k occupies per subset
int lds_bytes_of(int form, int k);

}  // namespace pl
