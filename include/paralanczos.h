/*
 * paralanczos.h - device-pointer C ABI of the MI355X batched-submatrix
 * Lanczos eigenvalue library (libparalanczos.so).
 *
 * All array arguments are device pointers allocated by the caller (any HIP
 * allocator); `stream` is a hipStream_t (NULL = default stream). Every call
 * is asynchronous on `stream` unless stated otherwise; argument errors are
 * reported synchronously through the returned pl_status and pl_last_error().
 *
 * Reference interfaces these replace (btuomanen/ParaLanczos):
 *   pl_solve          - POS_SYM_SUBMATRIX_KER / POS_SYM_SUBMATRIX_KER2
 *                       (ParaLanczos.h:32-34, launched at ParaLanczos.cu:1754,
 *                       :1852 and by Matlab/gpu_multi_subsvd.m:54-60), reduced
 *                       to lambda_min, lambda_max and the RIP deviation.
 *   pl_solve_alleig   - the same kernels' full outputs Eigenvalues[s*k + j]
 *                       ascending, DetMat[s], Eigerror[s]
 *                       (ParaLanczos.cu:1465-1478).
 *   pl_upper          - UPPER_KER (ParaLanczos.h:36, ParaLanczos.cu:1484-1626),
 *                       per subset instead of one full matrix.
 *   pl_gen_*          - no reference counterpart: seeded index sets and
 *                       synthetic sensing matrices (the reference takes a
 *                       caller-supplied subIndex, ParaLanczos.cu:1811).
 */
#ifndef PARALANCZOS_AMD_H
#define PARALANCZOS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  PL_OK = 0,
  PL_ERR_ARG = 1,         /* invalid argument (see pl_last_error) */
  PL_ERR_HIP = 2,         /* a HIP runtime call failed */
  PL_ERR_NOMEM = 3,       /* device allocation failed */
  PL_ERR_UNSUPPORTED = 4  /* valid request this build does not support */
} pl_status;

/* per-subset status bits written by pl_solve / pl_solve_alleig */
#define PL_SUB_BREAKDOWN 1u /* Lanczos restarted (invariant subspace found) */
#define PL_SUB_NONFINITE 2u /* non-finite entry in G_S or in T            */
#define PL_SUB_NOCONV 4u    /* bisection hit its iteration cap            */
#define PL_SUB_BADINDEX 8u  /* an index outside [0, n)                    */
#define PL_SUB_BADK 16u     /* subset size outside the supported range    */

/* solve forms */
#define PL_FORM_TUNED 0 /* variant (lanes, LDS pad) from the tuning table: lane
                           groups up to k = 64, half-row kernels of 256 / 512
                           lanes beyond; FP64 Gram-Schmidt coefficients from
                           k = 128 on                                           */
#define PL_FORM_PLAIN 1 /* one 64-lane wavefront per subset (128 / 512 lanes
                           beyond k = 64 / 128), FP32 throughout, one launch per k */
#define PL_FORM_PLAIN_WAVE 2 /* one 64-lane wavefront per subset at every k: beyond
                                k = 64 the wavefront streams G_S and the basis
                                from a global scratch (512 KiB per resident
                                wavefront, allocated per call, or once by a plan
                                of this form); FP32 throughout,
                                one launch per k; no precision flags            */
/* precision flags, OR-ed into the form: FP32 or FP64 Gram-Schmidt
   coefficients at every k > 64 (subsets of k <= 64 always use FP32) */
#define PL_FORM_REORTH_F32 0x100
#define PL_FORM_REORTH_F64 0x200
/* tuned mixed batches: every bin on the caller's stream. Without it, outside
   graph capture, the bins other than the 512-lane one run on a helper stream
   (a fixed pool of two per device) between a fork after the binning and a
   join before the bisection, and fill the wide kernel's tail. Results are
   the same either way. */
#define PL_FORM_NO_OVERLAP 0x400

const char *pl_version(void);
const char *pl_status_string(pl_status st);
/* message of the last failing call on this thread ("" if none) */
const char *pl_last_error(void);

/* select the device for subsequent calls on this thread */
pl_status pl_init(int device);

/* largest subset size the solve kernels accept */
int pl_max_k(void);

/* diagnostic: mixed tuned batches whose bins were split over the caller's
   stream and a helper stream since the library was loaded */
int64_t pl_overlap_forks(void);

/*
 * lambda_min, lambda_max and delta = max(lambda_max - 1, 1 - lambda_min) of
 * the principal submatrices G[S_s, S_s], s = 0..nsub-1, of the symmetric
 * n x n FP32 matrix G (row-major, leading dimension ld >= n; only the upper
 * triangle G[min(i,j)][max(i,j)] is read).
 *   offsets == NULL: S_s = idx[s*k .. s*k+k-1] (all subsets of size k);
 *   offsets != NULL: S_s = idx[offsets[s] .. offsets[s+1]-1] (int64 offsets,
 *                    device memory) and k is the largest subset size; the
 *                    tuned form bins the subsets by size on the device and
 *                    drains each bin with one persistent launch, the plain
 *                    form makes one launch per distinct size (and waits for
 *                    the bin counts on the host).
 * Outputs: lmin, lmax, delta (FP32) and status (PL_SUB_* bits) per subset.
 * nsub = 0 is an empty batch: nothing is launched, and idx and the output
 * pointers may then be NULL (so may those of pl_upper and pl_solve_alleig).
 * Subsets with status BADINDEX/BADK/NONFINITE get NaN eigenvalues.
 */
pl_status pl_solve(const float *G, int64_t n, int64_t ld, const int32_t *idx,
                   const int64_t *offsets, int k, int64_t nsub, float *lmin, float *lmax,
                   float *delta, uint32_t *status, int form, void *stream);

/* Bytes of device workspace pl_solve_ws needs for nsub subsets of size <= k
   (the tridiagonal matrices and their bounds between the two stages). */
size_t pl_solve_workspace_size(int64_t nsub, int k);

/* pl_solve with a caller-provided device workspace of workspace_bytes >=
   pl_solve_workspace_size(nsub, k); with workspace == NULL it allocates one,
   stream-ordered, as pl_solve does. Calls that share a workspace on
   different streams race. */
pl_status pl_solve_ws(const float *G, int64_t n, int64_t ld, const int32_t *idx,
                      const int64_t *offsets, int k, int64_t nsub, float *lmin, float *lmax,
                      float *delta, uint32_t *status, int form, void *workspace,
                      size_t workspace_bytes, void *stream);

/* pl_solve without any device allocation. No counterpart in the reference:
   it is for callers that cannot allocate (code under HIP graph capture, tight
   serving loops, harnesses that forbid hipMalloc). The workspace is a
   library-owned device array of PL_NOALLOC_SCRATCH_BYTES per GPU, held from
   library load until the process ends; calls with
   pl_solve_workspace_size(nsub, k) above that are rejected (PL_ERR_ARG).
   Every call on a GPU uses that GPU's one array, so calls that overlap on one
   GPU (other streams, other threads) race: order them. Tuned mixed batches
   run every bin on `stream` (as with PL_FORM_NO_OVERLAP); PL_FORM_PLAIN_WAVE
   is not accepted (PL_ERR_UNSUPPORTED); PL_FORM_PLAIN with offsets waits on
   the host for its bin counts. Results equal pl_solve's. */
#define PL_NOALLOC_SCRATCH_BYTES (32u << 20)
pl_status pl_solve_noalloc(const float *G, int64_t n, int64_t ld, const int32_t *idx,
                           const int64_t *offsets, int k, int64_t nsub, float *lmin,
                           float *lmax, float *delta, uint32_t *status, int form,
                           void *stream);

/* One stage of pl_solve_ws on a caller workspace: stage 1 gathers each G_S
   and runs the Lanczos recurrence (the tridiagonal T of every subset and, for
   a mixed batch, its bins stay in the workspace); stage 2 bisects T for
   lambda_min, lambda_max and delta. Stage 2 must follow stage 1 with the same
   arguments and workspace (on one stream, or ordered by the caller). */
pl_status pl_solve_stage(const float *G, int64_t n, int64_t ld, const int32_t *idx,
                         const int64_t *offsets, int k, int64_t nsub, float *lmin, float *lmax,
                         float *delta, uint32_t *status, int form, int stage, void *workspace,
                         size_t workspace_bytes, void *stream);

/* same, with an explicit variant and FP32 coefficients: lanes per subset and
   LDS row padding (0, 1 or 4 floats) - used by the autotuner. Variants: 8, 16,
   32 or 64 lanes (k <= lanes; pad 0/1/4), 128 lanes (k <= 128; pad 1/4),
   256 and 512 lanes (half rows: k <= 128 and k <= 256; pad 4) */
pl_status pl_solve_variant(const float *G, int64_t n, int64_t ld, const int32_t *idx,
                           const int64_t *offsets, int k, int64_t nsub, float *lmin,
                           float *lmax, float *delta, uint32_t *status, int lanes, int pad,
                           void *stream);

/*
 * Plans (no reference counterpart: the reference re-uploads G and relaunches
 * per call, ParaLanczos.cu:1837-1858). A plan binds one solve to fixed device
 * buffers and owns its workspace (allocated by pl_plan_create, nothing is
 * allocated afterwards). The first pl_plan_execute of a plan whose work needs
 * no host synchronisation (every form except PL_FORM_PLAIN with offsets)
 * records the launch sequence into a HIP graph; later executions replay it.
 * The contents of idx, offsets and G may change between executions; subset
 * sizes are re-binned on the device each time. Executions of one plan must
 * not overlap; pl_plan_destroy waits for the device.
 */
typedef struct pl_plan pl_plan;
pl_status pl_plan_create(pl_plan **plan, const float *G, int64_t n, int64_t ld,
                         const int32_t *idx, const int64_t *offsets, int k, int64_t nsub,
                         float *lmin, float *lmax, float *delta, uint32_t *status, int form);
pl_status pl_plan_execute(pl_plan *plan, void *stream);
int pl_plan_is_graph(const pl_plan *plan); /* 1 once the plan replays a graph */
pl_status pl_plan_destroy(pl_plan *plan);

/*
 * Tuning table of PL_FORM_TUNED: the variant (lanes per subset, LDS row
 * padding) used for each k; mixed batches are binned by runs of equal
 * variants. pl_autotune times every variant that takes k of the family of
 * k's default variant (lane groups of 8..64 lanes up to k = 64, the 256-lane
 * half-row kernel up to 128, the 512-lane one beyond; a family gives
 * bit-identical results whatever its lanes and padding) and pad 0/1/4 on
 * `sample` seeded subsets of each size in klist (nk sizes, each <= n), keeps
 * the fastest, and gives every unlisted size the winner of the next listed
 * size above it when that winner is of the size's own family (else the
 * default). best_ms (nk entries, may be NULL) receives the winning
 * times. Synchronous; the table is shared by all threads of the process.
 */
pl_status pl_autotune(const float *G, int64_t n, int64_t ld, const int32_t *klist, int nk,
                      int64_t sample, uint64_t seed, double *best_ms, void *stream);
/* pl_autotune with the bin edges tuned too: at each listed size every variant
   that takes it competes, across kernel families, and every unlisted size
   takes the winner of the next listed size above it, so the boundaries
   between the families (the bin edges of mixed batches) move to the measured
   optimum. Different families round differently (all within the stated
   tolerance); while the table stays fixed, every subset's result is the same
   alone and in any mix. sweep_ms (nk x PL_TUNE_VARIANTS, may be NULL): the
   time of variant v = 3 * (index of its lanes in 8, 16, 32, 64, 128, 256,
   512) + (index of its pad in 0, 1, 4) for each listed size, -1 where that
   variant does not exist or does not take the size. */
#define PL_TUNE_VARIANTS 21
pl_status pl_autotune_edges(const float *G, int64_t n, int64_t ld, const int32_t *klist, int nk,
                            int64_t sample, uint64_t seed, double *best_ms, double *sweep_ms,
                            void *stream);
pl_status pl_tuning_get(int k, int *lanes, int *pad);
pl_status pl_tuning_set(int k, int lanes, int pad);
pl_status pl_tuning_reset(void);

/*
 * Serving engine (no reference counterpart): continuous batching of requests
 * against a device-resident G. A request is nsub subsets whose index sets
 * (idx, nidx entries), CSR offsets (request-local, offsets[0] = 0 and
 * offsets[nsub] = nidx; or NULL with uniform size k) and outputs (lmin, lmax, delta, status; nsub entries
 * each) are caller-owned DEVICE buffers that must stay valid until the
 * request is waited for; kmax is the largest subset size in it. Requests are
 * packed in arrival order into steps bounded by the per-step arena
 * (max_subsets, max_indices, max_requests) and by work_budget (sum over the
 * requests of nsub * LDS bytes per subset of the kernel for kmax * kmax; 0 =
 * no bound); a request alone in a step is always admitted. Up to `depth`
 * steps run at once, each on its own stream. Results equal those of pl_solve
 * with the same form, whatever the mix. A request whose offsets do not run
 * from 0 to nidx gets status PL_SUB_BADK and NaN outputs for every subset;
 * its indices never reach another request's subsets. Submission and waiting
 * are thread-safe; pl_server_destroy finishes every queued request first.
 */
typedef struct pl_server pl_server;
typedef struct {
  int64_t max_subsets;  /* subsets per step (device arena)          */
  int64_t max_indices;  /* index entries per step                   */
  int32_t max_requests; /* requests per step                        */
  int32_t max_k;        /* largest subset size accepted (<= pl_max_k) */
  int32_t depth;        /* steps in flight, 1..8                    */
  int32_t form;         /* PL_FORM_TUNED or PL_FORM_PLAIN (| precision flag) */
  double work_budget;   /* per-step LDS-occupancy budget, 0 = none  */
} pl_server_config;
typedef struct {
  int64_t requests, subsets, steps; /* completed since the last reset */
  double span_s;                    /* first submission to last completion */
  double p50_s, p99_s, max_s;       /* request latency, submission to completion */
} pl_server_stats;
pl_status pl_server_create(const float *G, int64_t n, int64_t ld, const pl_server_config *cfg,
                           pl_server **server);
/* `stream` is the stream whose work writes the request's inputs (NULL = the
   default stream): the request's step starts after that work, as if ordered
   on it. Outputs are complete when pl_server_wait returns. */
pl_status pl_server_submit(pl_server *server, const int32_t *idx, const int64_t *offsets, int k,
                           int64_t nsub, int64_t nidx, int kmax, float *lmin, float *lmax,
                           float *delta, uint32_t *status, void *stream, int64_t *ticket);
/* block until the request is done; latency_s = submission to completion,
   queued_s = submission to the start of its step (either may be NULL) */
pl_status pl_server_wait(pl_server *server, int64_t ticket, double *latency_s, double *queued_s);
pl_status pl_server_poll(pl_server *server, int64_t ticket, int *done);
pl_status pl_server_query(pl_server *server, pl_server_stats *stats, int reset);
pl_status pl_server_destroy(pl_server *server);

/* Power-iteration estimate of lambda_max per subset (the reference's
   UPPER_KER, ParaLanczos.h:36, ParaLanczos.cu:1484-1626, which runs 10 steps
   on one full FP64 matrix of n <= 128): from v = ones/sqrt(k), `iters` times
   w = G_S v, b = ||w||, v = w / b; upper[s] = the last b (FP32; FP32 products,
   ||w||^2 summed in FP64), a lower bound of lambda_max that approaches it.
   Subsets as in pl_solve (uniform k, or CSR offsets with k the largest size);
   status gets PL_SUB_BADINDEX / BADK / NONFINITE with NaN in upper. A full
   matrix is the identity index set 0..n-1. */
pl_status pl_upper(const float *G, int64_t n, int64_t ld, const int32_t *idx,
                   const int64_t *offsets, int k, int64_t nsub, int iters, float *upper,
                   uint32_t *status, void *stream);

/* all k eigenvalues ascending eig[s*k + j] (FP64), det[s] = prod_j eig, and
   err[s] = per-subset status, for nsub subsets of size k (uniform) */
pl_status pl_solve_alleig(const float *G, int64_t n, int64_t ld, const int32_t *idx, int k,
                          int64_t nsub, double *eig, double *det, int32_t *err, void *stream);

/*
 * RIP statistics of a solved batch (the outputs of pl_solve*), per subset
 * size, for the reference's consumer of the bounds (the SubSeq record,
 * ParaLanczos0.h:82-90: per index-set size the lower and upper bound as mean
 * and standard deviation). out is a device array of (kmax+1) x 8 doubles;
 * row k (1 <= k <= kmax) holds {count, mean lambda_min, std lambda_min,
 * mean lambda_max, std lambda_max, max delta, min lambda_min, max lambda_max}
 * over the subsets of size k without a fatal status bit (NaN when count is 0);
 * row 0 holds the number of subsets skipped. Sizes come from offsets (CSR) or
 * are all k. The sums are FP64 atomics: the last bits of the means and
 * deviations may vary from run to run; counts and extremes do not.
 */
pl_status pl_rip_stats(const float *lmin, const float *lmax, const float *delta,
                       const uint32_t *status, const int64_t *offsets, int k, int64_t nsub,
                       int kmax, double *out, void *stream);

/* gather stage only: out[s*k*k + r*k + c] = G_S[r][c] (mode 0), or
   out[s] = sum of G_S (mode 1); uniform k */
pl_status pl_gather(const float *G, int64_t n, int64_t ld, const int32_t *idx, int k,
                    int64_t nsub, float *out, int mode, int lanes, int pad, void *stream);

/* seeded index sets: Floyd sample of k distinct indices of [0,n) per subset
   s0+s, Philox4x32-10 stream (seed, tag 1, s0+s), kept in draw order */
pl_status pl_gen_subsets(uint64_t seed, int64_t s0, int64_t nsub, int k, int n, int32_t *idx,
                         void *stream);
/* same for variable sizes: subset s occupies idx[offsets[s] .. offsets[s+1]) */
pl_status pl_gen_subsets_csr(uint64_t seed, int64_t s0, int64_t nsub, const int64_t *offsets,
                             int n, int32_t *idx, void *stream);
/* ks[s] = klist[u], u uniform in [0,nk) from stream (seed, tag 2, s0+s);
   klist is a HOST array of at most 32 sizes */
pl_status pl_gen_ks(uint64_t seed, int64_t s0, int64_t nsub, const int32_t *klist, int nk,
                    int32_t *ks, void *stream);

/* Gram matrix G = Phi^T Phi (n x n FP32, row-major, leading dimension ldg >= n)
   of an m x n FP32 frame Phi, on the matrix cores (FP32 products, FP32
   accumulation). Replaces the reference's entry-by-entry Gram builders DSG_B /
   DSG_B1 / SG_B1 (ParaLanczos.cu:364-438), which take the frame row-major,
   Frame[r*N + c]. Layouts: PL_LAYOUT_ROW, entry (r, c) at phi[r*ld + c]
   (ld >= n; the reference's), or PL_LAYOUT_COL, entry (r, c) at phi[c*ld + r]
   (ld >= m; the layout pl_gen_phi writes). Only upper-triangle tiles are
   computed; G is written whole and exactly symmetric. Asynchronous on stream. */
#define PL_LAYOUT_ROW 0
#define PL_LAYOUT_COL 1
pl_status pl_gram(const float *phi, int64_t m, int64_t n, int64_t ld, int layout, float *G,
                  int64_t ldg, void *stream);

/* Sub-Gram matrices straight from the frame, the reference's DSG_B form
   (ParaLanczos.cu:364-382, subgrams[j + i*n] = sum_k Frame[subset[i] + k*N]
   * Frame[subset[j] + k*N]) over a batch: out[s*k*k + r*k + c] =
   sum_t Phi(t, idx[s*k + r]) Phi(t, idx[s*k + c]), uniform k in [1, 256],
   on the matrix cores with the accumulation order of pl_gram (so each G_S
   equals the one pl_gather takes from pl_gram's G, bit for bit). An
   alternative to gathering from a materialised G. Frame layouts as pl_gram.
   status (nsub entries, may be NULL): 0, or PL_SUB_BADINDEX for a subset with
   an index outside [0, n), whose output is left unwritten. */
pl_status pl_subgram(const float *phi, int64_t m, int64_t n, int64_t ld, int layout,
                     const int32_t *idx, int k, int64_t nsub, float *out, uint32_t *status,
                     void *stream);

/* synthetic sensing matrix, column-major m x n (phi[j*m + r]):
   kind 0: N(0,1/m) entries, columns normalised to unit L2 norm;
   kind 1: +-1/sqrt(m) equiprobable;
   kind 2: m seeded rows of the orthonormal n-point DCT-II, scaled sqrt(n/m) */
pl_status pl_gen_phi(int kind, uint64_t seed, int m, int n, float *phi, void *stream);

#ifdef __cplusplus
}
#endif
#endif
