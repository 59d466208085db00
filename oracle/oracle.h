/*
 * CPU oracle for the batched-submatrix Lanczos eigenvalue path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP product
 * path is compared against; it is linked only by tests/, by the smoke check
 * and by the CPU-baseline leg of bench.py. Product code never calls it.
 *
 * It restates, in plain C99 + FP64, the per-subset algorithm of
 * ParaLanczos.cu (POS_SYM_SUBMATRIX_KER2, ParaLanczos.cu:1157-1480) with the
 * substitutions the MI355X build makes (deterministic start vector, full
 * reorthogonalisation, Sturm bisection for the extreme eigenvalues), plus a
 * reference-faithful variant (no reorthogonalisation, rational QL on T) that
 * follows ParaLanczos.cu:1285-1374 and :32-189 literally apart from the start
 * vector.
 *
 * Pins (what each function is checked against) are listed in oracle.c.
 */
#ifndef PL_ORACLE_H
#define PL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status bits, shared meaning with the product's per-subset status word */
#define OR_ST_BREAKDOWN 1u  /* Lanczos restarted after beta <= tau*scale       */
#define OR_ST_NONFINITE 2u  /* non-finite entry in the gathered submatrix / T  */
#define OR_ST_NOCONV    4u  /* bisection hit its iteration cap                 */
#define OR_ST_BADINDEX  8u  /* an index outside [0, n)                         */
#define OR_ST_BADK      16u /* subset size outside [1, 256]                    */

/* breakdown threshold: beta_{j+1} <= OR_TAU * (max|alpha| + max beta) */
#define OR_TAU (1.0 / 1073741824.0)

/* ---- counter-based PRNG (Philox4x32-10) and the derived samplers ---- */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* word w of the stream (seed, tag, sid): block w/4 of Philox, lane w%4 */
uint32_t or_stream_word(uint64_t seed, uint32_t tag, uint64_t sid, uint32_t w);
/* Floyd sample of k distinct indices of [0,n) for subset sid, in draw order */
void or_floyd(uint64_t seed, uint64_t sid, int n, int k, int32_t *out);
void or_gen_subsets_uniform(uint64_t seed, int64_t s0, int64_t nsub, int k, int n,
                            int32_t *idx);
void or_gen_ks(uint64_t seed, int64_t s0, int64_t nsub, const int32_t *klist, int nk,
               int32_t *ks);
void or_gen_subsets_csr(uint64_t seed, int64_t s0, int64_t nsub, const int64_t *offsets,
                        int n, int32_t *idx);

/* ---- synthetic sensing matrices, column-major m x n (phi[j*m + r]) ---- */
void or_gen_phi_gauss(uint64_t seed, int m, int n, float *phi);
void or_gen_phi_bernoulli(uint64_t seed, int m, int n, float *phi);
void or_dct_rows(uint64_t seed, int m, int n, int32_t *rows);
void or_gen_phi_dct(uint64_t seed, int m, int n, float *phi);

/* ---- per-subset pieces ---- */
/* A[r*k+c] = G[min(i,j)*ld + max(i,j)], i = idx[r], j = idx[c]  (FP64 copy) */
uint32_t or_gather(const float *G, int64_t n, int64_t ld, const int32_t *idx, int k,
                   double *A);
/* FP64 Lanczos with full reorthogonalisation. alpha[0..k-1], beta[0..k-1]
   (beta[0] = 0, beta[j] couples rows j-1 and j). Returns status bits. */
uint32_t or_lanczos(const double *A, int k, double *alpha, double *beta);
/* Sturm count: number of eigenvalues of T strictly below sigma */
int or_sturm_count(const double *alpha, const double *beta, int k, double sigma);
/* t-th smallest eigenvalue of T (0-based) by bisection; *nit = iterations */
double or_sturm_eig(const double *alpha, const double *beta, int k, int t, int *nit);
/* EISPACK rational QL (TQLRAT) on d[0..n-1], e2[0..n-1] (e2[i] = beta_i^2,
   e2[0] unused); ascending eigenvalues in d; returns ierr. */
int or_tqlrat(int n, double *d, double *e2);
/* reference-faithful Lanczos: no reorthogonalisation, ones/sqrt(k) start */
void or_lanczos_noreorth(const double *A, int k, double *alpha, double *beta);

/* ---- batched drivers (OpenMP over subsets when nthreads > 1) ---- */
/* offsets == NULL -> subset s is idx[s*kuni .. s*kuni+kuni-1] */
void or_solve_batch(const float *G, int64_t n, int64_t ld, const int32_t *idx,
                    const int64_t *offsets, int kuni, int64_t nsub, double *lmin,
                    double *lmax, double *delta, uint32_t *status, int nthreads);
/* all eigenvalues ascending (eig[s*k + t]) and det = prod(eig) */
void or_solve_batch_alleig(const float *G, int64_t n, int64_t ld, const int32_t *idx,
                           int k, int64_t nsub, double *eig, double *det,
                           uint32_t *status, int nthreads);
/* reference-faithful: FP64 G, no reorth, TQLRAT, all eigenvalues + det */
void or_solve_batch_faithful(const double *G, int64_t n, const int32_t *idx, int k,
                             int64_t nsub, double *eig, double *det, int32_t *ierr,
                             int nthreads);

/* reference-faithful path (FP64 gather of the FP32 G, no reorth, TQLRAT)
   over CSR or uniform index sets; the ends of the ascending eigenvalues */
void or_faithful_extremes(const float *G, int64_t n, int64_t ld, const int32_t *idx,
                          const int64_t *offsets, int kuni, int64_t nsub, double *lmin,
                          double *lmax, int32_t *ierr, int nthreads);
/* power-iteration estimate of lambda_max (UPPER_KER): ones/sqrt(k) start,
   `iters` steps of w = A v, b = ||w||, v = w / b; returns the last b */
double or_upper(const double *A, int k, int iters);
void or_upper_batch(const float *G, int64_t n, int64_t ld, const int32_t *idx,
                    const int64_t *offsets, int kuni, int64_t nsub, int iters, double *upper,
                    uint32_t *status, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
