/*
 * CPU oracle for the batched-submatrix Lanczos eigenvalue path.
 *
 * TEST INFRASTRUCTURE ONLY: linked by tests/, the smoke check and the
 * CPU-baseline leg of bench.py, never by the product library.
 *
 * What pins each part (tests/test_oracle.py):
 *   or_philox4x32_10  - the published Random123 known-answer vectors
 *                       (Salmon et al., SC'11) for philox4x32-10.
 *   or_floyd & co.    - invariants (distinct, in range, count), chi-square
 *                       uniformity, committed index-stream digests.
 *   or_gen_phi_*      - moments / unit column norms / DCT orthonormality.
 *   or_gather         - direct indexing on small matrices.
 *   or_lanczos        - V^T A V = T and orthonormality invariants, closed
 *                       forms (diagonal, 2x2, identity breakdown), numpy
 *                       dsyevd of the gathered matrix, the reference's own
 *                       example matrix (example1.cu:19-26, example3.cu).
 *   or_sturm_*        - tridiagonal Toeplitz closed form
 *                       a + 2b cos(j pi/(k+1)), scipy eigvalsh_tridiagonal
 *                       fixtures, brute-force counts on tiny T.
 *   or_tqlrat         - the same closed form and fixtures.
 *   or_upper          - closed forms: a diagonal matrix (||D^L 1|| /
 *                       ||D^(L-1) 1||), a rank-one matrix (exact after one
 *                       step), a multiple of the identity; bounded above by
 *                       numpy's largest eigenvalue and converging to it.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as   */
/* easy as 1, 2, 3", SC'11). The reference uses curand XORWOW seeded from    */
/* clock() (ParaLanczos.cu:1196-1206); the build replaces it with a counter- */
/* based generator so host and device draw identical streams.                */
/* ------------------------------------------------------------------------ */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int round = 0; round < 10; ++round) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

uint32_t or_stream_word(uint64_t seed, uint32_t tag, uint64_t sid, uint32_t w) {
  uint32_t ctr[4] = {w >> 2, (uint32_t)sid, (uint32_t)(sid >> 32), tag};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t out[4];
  or_philox4x32_10(ctr, key, out);
  return out[w & 3u];
}

/* sequential reader over one (seed, tag, sid) stream */
typedef struct {
  uint64_t seed, sid;
  uint32_t tag, w;
  uint32_t buf[4];
} or_strm;

static uint32_t strm_next(or_strm *s) {
  if ((s->w & 3u) == 0) {
    uint32_t ctr[4] = {s->w >> 2, (uint32_t)s->sid, (uint32_t)(s->sid >> 32), s->tag};
    uint32_t key[2] = {(uint32_t)s->seed, (uint32_t)(s->seed >> 32)};
    or_philox4x32_10(ctr, key, s->buf);
  }
  return s->buf[s->w++ & 3u];
}

/* uniform integer in [0, range), range >= 1: Lemire's multiply-shift with
   rejection (Lemire, "Fast random integer generation in an interval",
   ACM TOMACS 2019). */
static uint32_t strm_bounded(or_strm *s, uint32_t range) {
  uint64_t m = (uint64_t)strm_next(s) * range;
  uint32_t l = (uint32_t)m;
  if (l < range) {
    uint32_t t = (uint32_t)(0u - range) % range;
    while (l < t) {
      m = (uint64_t)strm_next(s) * range;
      l = (uint32_t)m;
    }
  }
  return (uint32_t)(m >> 32);
}

enum { TAG_IDX = 1, TAG_K = 2, TAG_PHI = 3, TAG_DCT = 4 };

/* Floyd's algorithm (Bentley & Floyd, CACM 1987) for k distinct values of
   [0, n), kept in draw order. */
static void floyd_tag(uint64_t seed, uint32_t tag, uint64_t sid, int n, int k,
                      int32_t *out) {
  or_strm s = {seed, sid, tag, 0u, {0, 0, 0, 0}};
  int cnt = 0;
  for (int j = n - k; j < n; ++j) {
    int32_t t = (int32_t)strm_bounded(&s, (uint32_t)j + 1u);
    int seen = 0;
    for (int q = 0; q < cnt; ++q)
      if (out[q] == t) { seen = 1; break; }
    out[cnt++] = seen ? (int32_t)j : t;
  }
}

void or_floyd(uint64_t seed, uint64_t sid, int n, int k, int32_t *out) {
  floyd_tag(seed, TAG_IDX, sid, n, k, out);
}

void or_gen_subsets_uniform(uint64_t seed, int64_t s0, int64_t nsub, int k, int n,
                            int32_t *idx) {
  for (int64_t s = 0; s < nsub; ++s)
    floyd_tag(seed, TAG_IDX, (uint64_t)(s0 + s), n, k, idx + s * (int64_t)k);
}

void or_gen_ks(uint64_t seed, int64_t s0, int64_t nsub, const int32_t *klist, int nk,
               int32_t *ks) {
  for (int64_t s = 0; s < nsub; ++s) {
    or_strm st = {seed, (uint64_t)(s0 + s), TAG_K, 0u, {0, 0, 0, 0}};
    ks[s] = klist[strm_bounded(&st, (uint32_t)nk)];
  }
}

void or_gen_subsets_csr(uint64_t seed, int64_t s0, int64_t nsub, const int64_t *offsets,
                        int n, int32_t *idx) {
  for (int64_t s = 0; s < nsub; ++s) {
    int k = (int)(offsets[s + 1] - offsets[s]);
    floyd_tag(seed, TAG_IDX, (uint64_t)(s0 + s), n, k, idx + offsets[s]);
  }
}

/* ------------------------------------------------------------------------ */
/* Synthetic sensing matrices (column-major, phi[j*m + r]).                  */
/* ------------------------------------------------------------------------ */
static double u53(uint32_t hi, uint32_t lo) {
  return ((double)(hi >> 5) * 67108864.0 + (double)(lo >> 6)) * (1.0 / 9007199254740992.0);
}

void or_gen_phi_gauss(uint64_t seed, int m, int n, float *phi) {
  const double two_pi = 6.283185307179586476925286766559;
  const double scale = 1.0 / sqrt((double)m);
  int64_t total = (int64_t)m * n;
  for (int64_t p = 0; 2 * p < total; ++p) {
    uint32_t ctr[4] = {0u, (uint32_t)p, (uint32_t)((uint64_t)p >> 32), TAG_PHI};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t r[4];
    or_philox4x32_10(ctr, key, r);
    double u1 = u53(r[0], r[1]), u2 = u53(r[2], r[3]);
    double rad = sqrt(-2.0 * log(1.0 - u1));
    double ang = two_pi * u2;
    phi[2 * p] = (float)(rad * cos(ang) * scale);
    if (2 * p + 1 < total) phi[2 * p + 1] = (float)(rad * sin(ang) * scale);
  }
  for (int j = 0; j < n; ++j) {
    float *col = phi + (int64_t)j * m;
    double s = 0.0;
    for (int r = 0; r < m; ++r) s += (double)col[r] * (double)col[r];
    double inv = 1.0 / sqrt(s);
    for (int r = 0; r < m; ++r) col[r] = (float)((double)col[r] * inv);
  }
}

void or_gen_phi_bernoulli(uint64_t seed, int m, int n, float *phi) {
  const float v = (float)(1.0 / sqrt((double)m));
  int64_t total = (int64_t)m * n;
  for (int64_t e = 0; e < total; ++e) {
    uint32_t w = or_stream_word(seed, TAG_PHI, (uint64_t)e >> 7, (uint32_t)(e >> 5) & 3u);
    phi[e] = ((w >> (e & 31)) & 1u) ? v : -v;
  }
}

void or_dct_rows(uint64_t seed, int m, int n, int32_t *rows) {
  floyd_tag(seed, TAG_DCT, 0u, n, m, rows);
}

/* rows R of the orthonormal DCT-II C[r][j] = sqrt(2/n) e_r cos(pi r (2j+1) / (2n)),
   e_0 = 1/sqrt(2), scaled by sqrt(n/m); the angle is reduced exactly in integers */
void or_gen_phi_dct(uint64_t seed, int m, int n, float *phi) {
  const double pi = 3.14159265358979323846264338327950288;
  int32_t *rows = (int32_t *)malloc(sizeof(int32_t) * (size_t)m);
  or_dct_rows(seed, m, n, rows);
  const double amp = sqrt(2.0 / n) * sqrt((double)n / m);
  const int64_t period = 4 * (int64_t)n;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) {
      int64_t r = rows[i];
      int64_t a = (r * (2 * (int64_t)j + 1)) % period;
      double c = cos(pi * (double)a / (2.0 * n));
      double e = (r == 0) ? sqrt(0.5) : 1.0;
      phi[(int64_t)j * m + i] = (float)(amp * e * c);
    }
  free(rows);
}

/* ------------------------------------------------------------------------ */
/* Per-subset pieces.                                                        */
/* ------------------------------------------------------------------------ */

/* Gather of the k x k principal submatrix, ParaLanczos.cu:471-490
   (LOAD_SMS4HF: mat[idx_r * n + idx_c]). The build reads the upper triangle
   of G in global index order, G[min(i,j)][max(i,j)], so G_S is exactly
   symmetric and independent of the order of the index set. */
uint32_t or_gather(const float *G, int64_t n, int64_t ld, const int32_t *idx, int k,
                   double *A) {
  uint32_t st = 0;
  for (int r = 0; r < k; ++r)
    if (idx[r] < 0 || idx[r] >= n) st |= OR_ST_BADINDEX;
  if (st) return st;
  for (int r = 0; r < k; ++r)
    for (int c = 0; c < k; ++c) {
      int64_t i = idx[r], j = idx[c];
      int64_t lo = i < j ? i : j, hi = i < j ? j : i;
      double x = (double)G[lo * ld + hi];
      if (!isfinite(x)) st |= OR_ST_NONFINITE;
      A[r * k + c] = x;
    }
  return st;
}

static double dotd(const double *a, const double *b, int k) {
  double s = 0.0;
  for (int i = 0; i < k; ++i) s += a[i] * b[i];
  return s;
}

/* one classical Gram-Schmidt pass of w against V[:, 0..nv-1] (columns at V + l*k) */
static void cgs_pass(const double *V, int nv, int k, double *w, double *h) {
  for (int l = 0; l < nv; ++l) h[l] = dotd(V + (int64_t)l * k, w, k);
  for (int l = 0; l < nv; ++l) {
    const double *vl = V + (int64_t)l * k;
    for (int i = 0; i < k; ++i) w[i] -= h[l] * vl[i];
  }
}

/* Lanczos recurrence of ParaLanczos.cu:1285-1374 (alpha from v^T G v before
   the three-term subtraction; beta_k discarded as in :1377-1378) with the
   build's substitutions: start vector ones/sqrt(k) instead of curand normals
   (:1196-1233), one full classical Gram-Schmidt pass against all previous
   Lanczos vectors after the three-term step (a second one when the first
   removed more than half of the vector's norm: the DGKS criterion of Daniel,
   Gragg, Kaufman, Stewart, Math. Comp. 1976), and a deterministic restart when
   beta_{j+1} <= OR_TAU * (max|alpha| + max beta): the new vector is e_i for
   the first i of largest residual 1 - sum_l V[i][l]^2, orthogonalised twice,
   and beta_{j+1} := 0. */
uint32_t or_lanczos(const double *A, int k, double *alpha, double *beta) {
  uint32_t st = 0;
  double *V = (double *)malloc(sizeof(double) * (size_t)k * k);
  double *v = (double *)malloc(sizeof(double) * (size_t)k);
  double *vp = (double *)calloc((size_t)k, sizeof(double));
  double *w = (double *)malloc(sizeof(double) * (size_t)k);
  double *h = (double *)malloc(sizeof(double) * (size_t)k);
  const double v0 = 1.0 / sqrt((double)k);
  for (int i = 0; i < k; ++i) v[i] = v0;
  double amax = 0.0, bmax = 0.0;
  beta[0] = 0.0;
  for (int j = 0; j < k; ++j) {
    memcpy(V + (int64_t)j * k, v, sizeof(double) * (size_t)k);
    for (int i = 0; i < k; ++i) w[i] = dotd(A + (int64_t)i * k, v, k);
    double a = dotd(v, w, k);
    alpha[j] = a;
    if (fabs(a) > amax) amax = fabs(a);
    for (int i = 0; i < k; ++i) w[i] -= a * v[i] + beta[j] * vp[i];
    const double pre = dotd(w, w, k);
    cgs_pass(V, j + 1, k, w, h);
    if (j == k - 1) break;
    double b2 = dotd(w, w, k);
    if (b2 < 0.5 * pre) {  /* DGKS criterion: orthogonalise a second time */
      cgs_pass(V, j + 1, k, w, h);
      b2 = dotd(w, w, k);
    }
    double b = sqrt(b2);
    memcpy(vp, v, sizeof(double) * (size_t)k);
    if (b > OR_TAU * (amax + bmax)) {
      beta[j + 1] = b;
      if (b > bmax) bmax = b;
      for (int i = 0; i < k; ++i) v[i] = w[i] / b;
    } else {
      st |= OR_ST_BREAKDOWN;
      beta[j + 1] = 0.0;
      int best = 0;
      double bres = -1.0;
      for (int i = 0; i < k; ++i) {
        double r = 1.0;
        for (int l = 0; l <= j; ++l) r -= V[(int64_t)l * k + i] * V[(int64_t)l * k + i];
        if (r > bres) { bres = r; best = i; }
      }
      for (int i = 0; i < k; ++i) w[i] = (i == best) ? 1.0 : 0.0;
      cgs_pass(V, j + 1, k, w, h);
      cgs_pass(V, j + 1, k, w, h);
      double nb = sqrt(dotd(w, w, k));
      for (int i = 0; i < k; ++i) v[i] = w[i] / nb;
    }
  }
  for (int j = 0; j < k; ++j)
    if (!isfinite(alpha[j]) || !isfinite(beta[j])) st |= OR_ST_NONFINITE;
  free(V); free(v); free(vp); free(w); free(h);
  return st;
}

/* Reference-faithful recurrence, ParaLanczos.cu:1285-1374: no
   reorthogonalisation, no breakdown guard; start vector ones/sqrt(k). */
void or_lanczos_noreorth(const double *A, int k, double *alpha, double *beta) {
  double *v = (double *)malloc(sizeof(double) * (size_t)k);
  double *vp = (double *)calloc((size_t)k, sizeof(double));
  double *w = (double *)malloc(sizeof(double) * (size_t)k);
  const double v0 = 1.0 / sqrt((double)k);
  for (int i = 0; i < k; ++i) v[i] = v0;
  beta[0] = 0.0;
  for (int j = 0; j < k; ++j) {
    for (int i = 0; i < k; ++i) w[i] = dotd(A + (int64_t)i * k, v, k);
    double a = dotd(v, w, k);
    for (int i = 0; i < k; ++i) w[i] -= a * v[i] + beta[j] * vp[i];
    double b = sqrt(dotd(w, w, k));
    memcpy(vp, v, sizeof(double) * (size_t)k);
    for (int i = 0; i < k; ++i) v[i] = w[i] / b;
    alpha[j] = a;
    if (j + 1 < k) beta[j + 1] = b;
  }
  free(v); free(vp); free(w);
}

static double pivmin_of(const double *beta, int k) {
  double m = 1.0;
  for (int i = 1; i < k; ++i)
    if (beta[i] * beta[i] > m) m = beta[i] * beta[i];
  return DBL_MIN * m;
}

/* Sturm sequence count with the LAPACK dstebz pivot guard (replaces the
   unused and defective EIGVALS_BELOW of ParaLanczos.cu:504-552) */
static int sturm_count_p(const double *alpha, const double *beta, int k, double sigma,
                         double pivmin) {
  double d = alpha[0] - sigma;
  if (fabs(d) < pivmin) d = -pivmin;
  int c = d < 0.0;
  for (int i = 1; i < k; ++i) {
    d = (alpha[i] - sigma) - beta[i] * beta[i] / d;
    if (fabs(d) < pivmin) d = -pivmin;
    c += d < 0.0;
  }
  return c;
}

int or_sturm_count(const double *alpha, const double *beta, int k, double sigma) {
  return sturm_count_p(alpha, beta, k, sigma, pivmin_of(beta, k));
}

/* bisection for the t-th smallest eigenvalue of T (replaces TQLRAT,
   ParaLanczos.cu:32-189, for the extreme eigenvalues) */
double or_sturm_eig(const double *alpha, const double *beta, int k, int t, int *nit) {
  double pivmin = pivmin_of(beta, k);
  double gl = INFINITY, gu = -INFINITY;
  for (int i = 0; i < k; ++i) {
    double r = fabs(beta[i] * (i > 0)) + (i + 1 < k ? fabs(beta[i + 1]) : 0.0);
    if (alpha[i] - r < gl) gl = alpha[i] - r;
    if (alpha[i] + r > gu) gu = alpha[i] + r;
  }
  double tnorm = fmax(fabs(gl), fabs(gu));
  gl -= 2.0 * DBL_EPSILON * tnorm * k + 4.0 * pivmin;
  gu += 2.0 * DBL_EPSILON * tnorm * k + 4.0 * pivmin;
  double lo = gl, hi = gu;
  int it = 0;
  for (; it < 256; ++it) {
    double tol = ldexp(fmax(fabs(lo), fabs(hi)), -46) + 4.0 * pivmin;
    if (hi - lo <= tol) break;
    double mid = lo + 0.5 * (hi - lo);
    if (sturm_count_p(alpha, beta, k, mid, pivmin) > t) hi = mid;
    else lo = mid;
  }
  if (nit) *nit = it;
  return lo + 0.5 * (hi - lo);
}

/* EISPACK TQLRAT (rational QL, Reinsch 1973), as used by the reference at
   ParaLanczos.cu:32-189 and called at :1461: e2[i] = beta_i^2 is shifted
   down one slot on entry (:59-62); eigenvalues come out ascending (:168-185);
   ierr = l+1 after 30 sweeps on eigenvalue l (:94-98). */
int or_tqlrat(int n, double *d, double *e2) {
  if (n == 1) return 0;
  for (int i = 1; i < n; ++i) e2[i - 1] = e2[i];
  e2[n - 1] = 0.0;
  double f = 0.0, t = 0.0, b = 0.0, c = 0.0;
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    double h = fabs(d[l]) + sqrt(e2[l]);
    if (t <= h) {
      t = h;
      b = fabs(t) * DBL_EPSILON;
      c = b * b;
    }
    int m = l;
    while (m < n && !(e2[m] <= c)) ++m;
    if (m == n) m = n - 1;
    if (m != l) {
      for (;;) {
        if (iter >= 30) return l + 1;
        ++iter;
        int l1 = l + 1;
        double s = sqrt(e2[l]);
        double g = d[l];
        double p = (d[l1] - g) / (2.0 * s);
        double r = sqrt(p * p + 1.0);
        d[l] = s / (p + (p < 0.0 ? -fabs(r) : fabs(r)));
        double hh = g - d[l];
        for (int i = l1; i < n; ++i) d[i] -= hh;
        f += hh;
        g = d[m];
        if (g == 0.0) g = b;
        hh = g;
        s = 0.0;
        for (int i = m - 1; i >= l; --i) {
          p = g * hh;
          r = p + e2[i];
          e2[i + 1] = s * r;
          s = e2[i] / r;
          d[i + 1] = hh + s * (hh + d[i]);
          g = d[i] - e2[i] / g;
          if (g == 0.0) g = b;
          hh = g * p / r;
        }
        e2[l] = s * g;
        d[l] = hh;
        if (hh == 0.0) break;
        if (fabs(e2[l]) <= fabs(c / hh)) break;
        e2[l] = hh * e2[l];
        if (e2[l] == 0.0) break;
      }
    }
    double p = d[l] + f;
    int i = l;
    while (i > 0 && d[i - 1] > p) {
      d[i] = d[i - 1];
      --i;
    }
    d[i] = p;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Batched drivers.                                                          */
/* ------------------------------------------------------------------------ */
static void solve_one(const float *G, int64_t n, int64_t ld, const int32_t *idx, int k,
                      double *A, double *al, double *be, double *lmin, double *lmax,
                      double *delta, uint32_t *status) {
  uint32_t st = 0;
  if (k < 1 || k > 256) st |= OR_ST_BADK;
  if (!st) st |= or_gather(G, n, ld, idx, k, A);
  if (st) {
    *lmin = *lmax = *delta = NAN;
    *status = st;
    return;
  }
  st |= or_lanczos(A, k, al, be);
  if (st & OR_ST_NONFINITE) {
    *lmin = *lmax = *delta = NAN;
    *status = st;
    return;
  }
  int it0 = 0, it1 = 0;
  double lo = or_sturm_eig(al, be, k, 0, &it0);
  double hi = or_sturm_eig(al, be, k, k - 1, &it1);
  if (it0 >= 256 || it1 >= 256) st |= OR_ST_NOCONV;
  *lmin = lo;
  *lmax = hi;
  *delta = fmax(hi - 1.0, 1.0 - lo);
  *status = st;
}

void or_solve_batch(const float *G, int64_t n, int64_t ld, const int32_t *idx,
                    const int64_t *offsets, int kuni, int64_t nsub, double *lmin,
                    double *lmax, double *delta, uint32_t *status, int nthreads) {
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
  {
    double *A = (double *)malloc(sizeof(double) * 256 * 256);
    double *al = (double *)malloc(sizeof(double) * 256);
    double *be = (double *)malloc(sizeof(double) * 256);
#pragma omp for schedule(dynamic, 16)
    for (int64_t s = 0; s < nsub; ++s) {
      int64_t o = offsets ? offsets[s] : s * (int64_t)kuni;
      int k = offsets ? (int)(offsets[s + 1] - offsets[s]) : kuni;
      solve_one(G, n, ld, idx + o, k, A, al, be, lmin + s, lmax + s, delta + s, status + s);
    }
    free(A); free(al); free(be);
  }
}

void or_solve_batch_alleig(const float *G, int64_t n, int64_t ld, const int32_t *idx,
                           int k, int64_t nsub, double *eig, double *det,
                           uint32_t *status, int nthreads) {
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
  {
    double *A = (double *)malloc(sizeof(double) * 256 * 256);
    double *al = (double *)malloc(sizeof(double) * 256);
    double *be = (double *)malloc(sizeof(double) * 256);
#pragma omp for schedule(dynamic, 16)
    for (int64_t s = 0; s < nsub; ++s) {
      uint32_t st = (k < 1 || k > 256) ? OR_ST_BADK : 0u;
      if (!st) st |= or_gather(G, n, ld, idx + s * (int64_t)k, k, A);
      if (!st) st |= or_lanczos(A, k, al, be);
      double prod = 1.0;
      for (int t = 0; t < k; ++t) {
        int it = 0;
        double e = (st & (OR_ST_BADINDEX | OR_ST_NONFINITE | OR_ST_BADK))
                       ? NAN
                       : or_sturm_eig(al, be, k, t, &it);
        if (it >= 256) st |= OR_ST_NOCONV;
        eig[s * (int64_t)k + t] = e;
        prod *= e;
      }
      det[s] = prod;
      status[s] = st;
    }
    free(A); free(al); free(be);
  }
}

void or_solve_batch_faithful(const double *G, int64_t n, const int32_t *idx, int k,
                             int64_t nsub, double *eig, double *det, int32_t *ierr,
                             int nthreads) {
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
  {
    double *A = (double *)malloc(sizeof(double) * (size_t)k * k);
    double *al = (double *)malloc(sizeof(double) * (size_t)k);
    double *be = (double *)malloc(sizeof(double) * (size_t)k);
#pragma omp for schedule(dynamic, 16)
    for (int64_t s = 0; s < nsub; ++s) {
      const int32_t *ix = idx + s * (int64_t)k;
      for (int r = 0; r < k; ++r)
        for (int c = 0; c < k; ++c) A[r * k + c] = G[(int64_t)ix[r] * n + ix[c]];
      or_lanczos_noreorth(A, k, al, be);
      for (int i = 0; i < k; ++i) be[i] = be[i] * be[i];
      double *d = eig + s * (int64_t)k;
      memcpy(d, al, sizeof(double) * (size_t)k);
      ierr[s] = or_tqlrat(k, d, be);
      double prod = 1.0;
      for (int i = 0; i < k; ++i) prod *= d[i];
      det[s] = prod;
    }
    free(A); free(al); free(be);
  }
}

/* ------------------------------------------------------------------------ */
/* Power-iteration estimate of lambda_max, UPPER_KER (ParaLanczos.cu:       */
/* 1484-1626): v normalised, then `iters` times w = A v, b = ||w||,         */
/* v = w / b; the estimate is the last b. Substitutions, as in the Lanczos  */
/* path: start vector ones/sqrt(k) instead of clock()-seeded curand normals */
/* (:1486-1512), any k <= 256 and principal submatrices instead of one full */
/* n <= 128 matrix. Not reproduced: the reference's norm sums every lane's  */
/* copy of w[0..3] (:1591-1599) and its next vector repeats w[0..3] in      */
/* every lane's chunk (:1605-1616), so it does not compute ||A v||.        */
/* ------------------------------------------------------------------------ */
double or_upper(const double *A, int k, int iters) {
  double *v = (double *)malloc(sizeof(double) * (size_t)k);
  double *w = (double *)malloc(sizeof(double) * (size_t)k);
  const double v0 = 1.0 / sqrt((double)k);
  for (int i = 0; i < k; ++i) v[i] = v0;
  double b = 0.0;
  for (int l = 0; l < iters; ++l) {
    for (int i = 0; i < k; ++i) w[i] = dotd(A + (int64_t)i * k, v, k);
    b = sqrt(dotd(w, w, k));
    for (int i = 0; i < k; ++i) v[i] = b > 0.0 ? w[i] / b : 0.0;
  }
  free(v); free(w);
  return b;
}

void or_upper_batch(const float *G, int64_t n, int64_t ld, const int32_t *idx,
                    const int64_t *offsets, int kuni, int64_t nsub, int iters, double *upper,
                    uint32_t *status, int nthreads) {
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
  {
    double *A = (double *)malloc(sizeof(double) * 256 * 256);
#pragma omp for schedule(dynamic, 16)
    for (int64_t s = 0; s < nsub; ++s) {
      int64_t o = offsets ? offsets[s] : s * (int64_t)kuni;
      int k = offsets ? (int)(offsets[s + 1] - offsets[s]) : kuni;
      uint32_t st = (k < 1 || k > 256) ? OR_ST_BADK : 0u;
      if (!st) st |= or_gather(G, n, ld, idx + o, k, A);
      upper[s] = st ? NAN : or_upper(A, k, iters);
      status[s] = st;
    }
    free(A);
  }
}

/* Reference-faithful extremes over an FP32 G and CSR index sets: each G_S is
   gathered into FP64 and run through the reference's path as it stands,
   ParaLanczos.cu:1285-1374 (no reorthogonalisation) then TQLRAT (:32-189)
   for every eigenvalue (:1454-1462); lmin/lmax are the ends of the ascending
   list, ierr TQLRAT's error code. The baseline the GPU path is timed against
   in its reference form (FP64 throughout, all eigenvalues). */
void or_faithful_extremes(const float *G, int64_t n, int64_t ld, const int32_t *idx,
                          const int64_t *offsets, int kuni, int64_t nsub, double *lmin,
                          double *lmax, int32_t *ierr, int nthreads) {
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
  {
    double *A = (double *)malloc(sizeof(double) * 256 * 256);
    double *al = (double *)malloc(sizeof(double) * 256);
    double *be = (double *)malloc(sizeof(double) * 256);
#pragma omp for schedule(dynamic, 16)
    for (int64_t s = 0; s < nsub; ++s) {
      const int64_t o = offsets ? offsets[s] : s * (int64_t)kuni;
      const int k = offsets ? (int)(offsets[s + 1] - offsets[s]) : kuni;
      if (k < 1 || k > 256 || or_gather(G, n, ld, idx + o, k, A) != 0) {
        lmin[s] = lmax[s] = NAN;
        ierr[s] = -1;
        continue;
      }
      or_lanczos_noreorth(A, k, al, be);
      for (int i = 0; i < k; ++i) be[i] = be[i] * be[i];
      ierr[s] = or_tqlrat(k, al, be);
      lmin[s] = al[0];
      lmax[s] = al[k - 1];
    }
    free(A); free(al); free(be);
  }
}
