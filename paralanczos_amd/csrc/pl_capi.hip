// C ABI of libparalanczos.so: the device-pointer API of paralanczos.h and the
// drop-in host entry points of ParaLanczos.h. Host code only; kernels live in
// pl_kernels.hip.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/ParaLanczos.h"
#include "../../include/paralanczos.h"
#include "pl_common.hpp"
#include "pl_internal.hpp"

namespace {

thread_local std::string g_err;

pl_status fail(pl_status st, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return st;
}

pl_status hip_fail(hipError_t e, const char *what) {
  return fail(e == hipErrorOutOfMemory ? PL_ERR_NOMEM : PL_ERR_HIP, "%s: %s", what,
              hipGetErrorString(e));
}

constexpr int kMaxK = 128;

int group_for(int k) {
  int g = 8;
  while (g < k) g <<= 1;
  return g;
}

// tuned LDS row padding per lane-group size (index log2(group) - 3)
int g_tuned_pad[5] = {4, 4, 4, 4, 4};

int tuned_pad(int group) {
  const int i = group == 8 ? 0 : group == 16 ? 1 : group == 32 ? 2 : group == 64 ? 3 : 4;
  return g_tuned_pad[i];
}

pl_status check_solve_args(const float *G, int64_t n, int64_t ld, const int32_t *idx, int k,
                           int64_t nsub) {
  if (!G || !idx) return fail(PL_ERR_ARG, "G and idx must be non-null");
  if (n < 1 || n > 0x7fffffffLL) return fail(PL_ERR_ARG, "n = %lld out of range", (long long)n);
  if (ld < n) return fail(PL_ERR_ARG, "ld = %lld < n = %lld", (long long)ld, (long long)n);
  if (nsub < 0) return fail(PL_ERR_ARG, "nsub = %lld < 0", (long long)nsub);
  if (k < 1) return fail(PL_ERR_ARG, "k = %d < 1", k);
  if (k > kMaxK) return fail(PL_ERR_UNSUPPORTED, "k = %d > %d is not supported", k, kMaxK);
  return PL_OK;
}

// stream-ordered workspace: the caller's, or one allocated for this call
struct Workspace {
  void *p = nullptr;
  bool owned = false;
  hipStream_t stream = nullptr;
  pl_status acquire(void *given, size_t given_bytes, size_t need, hipStream_t s) {
    stream = s;
    if (given) {
      if (given_bytes < need)
        return fail(PL_ERR_ARG, "workspace of %zu bytes < %zu needed", given_bytes, need);
      p = given;
      return PL_OK;
    }
    hipError_t e = hipMallocAsync(&p, need ? need : 16, s);
    if (e != hipSuccess) return hip_fail(e, "hipMallocAsync(workspace)");
    owned = true;
    return PL_OK;
  }
  ~Workspace() {
    if (owned) (void)hipFreeAsync(p, stream);
  }
};

pl_status run_solve(const float *G, int64_t n, int64_t ld, const int32_t *idx,
                    const int64_t *offsets, int k, int64_t nsub, float *lmin, float *lmax,
                    float *delta, uint32_t *status, int group, int pad, void *ws,
                    size_t ws_bytes, void *stream) {
  if (!lmin || !lmax || !delta || !status)
    return fail(PL_ERR_ARG, "output pointers must be non-null");
  if (group < k) return fail(PL_ERR_ARG, "lanes per subset %d < k = %d", group, k);
  if (!pl::solve_variant_exists(group, pad))
    return fail(PL_ERR_ARG, "no variant with %d lanes and pad %d", group, pad);
  if (nsub == 0) return PL_OK;
  hipStream_t s = (hipStream_t)stream;
  Workspace w;
  pl_status st = w.acquire(ws, ws_bytes, pl::workspace_bytes(nsub, k), s);
  if (st != PL_OK) return st;
  float2 *tri = (float2 *)w.p;
  double4 *bounds = (double4 *)((char *)w.p + pl::tridiag_bytes(nsub, k));
  pl::SolveArgs a{};
  a.G = G;
  a.n = n;
  a.ld = ld;
  a.idx = idx;
  a.offsets = offsets;
  a.kuni = offsets ? 0 : k;
  a.nslots = nsub;
  a.status = status;
  a.tri = tri;
  a.tstride = k;
  a.bounds = bounds;
  hipError_t e = pl::launch_solve(group, pad, pl::MODE_TRIDIAG, a, s);
  if (e != hipSuccess) return hip_fail(e, "solve launch");
  pl::BisectArgs b{};
  b.tri = tri;
  b.tstride = k;
  b.bounds = bounds;
  b.offsets = offsets;
  b.kuni = offsets ? 0 : k;
  b.nsub = nsub;
  b.status = status;
  b.lmin = lmin;
  b.lmax = lmax;
  b.delta = delta;
  e = pl::launch_bisect(0, b, s);
  if (e != hipSuccess) return hip_fail(e, "bisection launch");
  return PL_OK;
}

pl_status pick_form(int form, int k, int &group, int &pad) {
  if (form == PL_FORM_PLAIN) {
    // one 64-lane wavefront per subset; a 128-lane workgroup beyond k = 64
    group = k <= 64 ? 64 : 128;
    pad = 4;
  } else if (form == PL_FORM_TUNED) {
    group = group_for(k);
    pad = tuned_pad(group);
  } else {
    return fail(PL_ERR_ARG, "unknown form %d", form);
  }
  return PL_OK;
}

}  // namespace

extern "C" {

const char *pl_version(void) { return "paralanczos-amd 0.1 (gfx950)"; }

const char *pl_status_string(pl_status st) {
  switch (st) {
    case PL_OK: return "ok";
    case PL_ERR_ARG: return "invalid argument";
    case PL_ERR_HIP: return "HIP runtime error";
    case PL_ERR_NOMEM: return "out of device memory";
    case PL_ERR_UNSUPPORTED: return "unsupported";
  }
  return "unknown status";
}

const char *pl_last_error(void) { return g_err.c_str(); }

pl_status pl_init(int device) {
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  return PL_OK;
}

int pl_max_k(void) { return kMaxK; }

size_t pl_solve_workspace_size(int64_t nsub, int k) {
  if (nsub < 0 || k < 1) return 0;
  return pl::workspace_bytes(nsub, k);
}

pl_status pl_solve(const float *G, int64_t n, int64_t ld, const int32_t *idx,
                   const int64_t *offsets, int k, int64_t nsub, float *lmin, float *lmax,
                   float *delta, uint32_t *status, int form, void *stream) {
  return pl_solve_ws(G, n, ld, idx, offsets, k, nsub, lmin, lmax, delta, status, form, nullptr, 0,
                     stream);
}

pl_status pl_solve_ws(const float *G, int64_t n, int64_t ld, const int32_t *idx,
                      const int64_t *offsets, int k, int64_t nsub, float *lmin, float *lmax,
                      float *delta, uint32_t *status, int form, void *workspace,
                      size_t workspace_bytes, void *stream) {
  pl_status st = check_solve_args(G, n, ld, idx, k, nsub);
  if (st != PL_OK) return st;
  int group, pad;
  st = pick_form(form, k, group, pad);
  if (st != PL_OK) return st;
  return run_solve(G, n, ld, idx, offsets, k, nsub, lmin, lmax, delta, status, group, pad,
                   workspace, workspace_bytes, stream);
}

pl_status pl_solve_variant(const float *G, int64_t n, int64_t ld, const int32_t *idx,
                           const int64_t *offsets, int k, int64_t nsub, float *lmin,
                           float *lmax, float *delta, uint32_t *status, int lanes, int pad,
                           void *stream) {
  pl_status st = check_solve_args(G, n, ld, idx, k, nsub);
  if (st != PL_OK) return st;
  return run_solve(G, n, ld, idx, offsets, k, nsub, lmin, lmax, delta, status, lanes, pad,
                   nullptr, 0, stream);
}

pl_status pl_solve_alleig(const float *G, int64_t n, int64_t ld, const int32_t *idx, int k,
                          int64_t nsub, double *eig, double *det, int32_t *err, void *stream) {
  pl_status st = check_solve_args(G, n, ld, idx, k, nsub);
  if (st != PL_OK) return st;
  if (!eig || !det || !err) return fail(PL_ERR_ARG, "output pointers must be non-null");
  if (nsub == 0) return PL_OK;
  hipStream_t s = (hipStream_t)stream;
  Workspace w;
  st = w.acquire(nullptr, 0, pl::workspace_bytes(nsub, k), s);
  if (st != PL_OK) return st;
  float2 *tri = (float2 *)w.p;
  double4 *bounds = (double4 *)((char *)w.p + pl::tridiag_bytes(nsub, k));
  const int group = group_for(k);
  pl::SolveArgs a{};
  a.G = G;
  a.n = n;
  a.ld = ld;
  a.idx = idx;
  a.kuni = k;
  a.nslots = nsub;
  a.status = (uint32_t *)err;
  a.tri = tri;
  a.tstride = k;
  a.bounds = bounds;
  hipError_t e = pl::launch_solve(group, tuned_pad(group), pl::MODE_TRIDIAG, a, s);
  if (e != hipSuccess) return hip_fail(e, "all-eigenvalue solve launch");
  pl::BisectArgs b{};
  b.tri = tri;
  b.tstride = k;
  b.bounds = bounds;
  b.kuni = k;
  b.nsub = nsub;
  b.status = (uint32_t *)err;
  b.eig = eig;
  b.det = det;
  e = pl::launch_bisect(1, b, s);
  if (e != hipSuccess) return hip_fail(e, "all-eigenvalue bisection launch");
  return PL_OK;
}

pl_status pl_gather(const float *G, int64_t n, int64_t ld, const int32_t *idx, int k,
                    int64_t nsub, float *out, int mode, int lanes, int pad, void *stream) {
  pl_status st = check_solve_args(G, n, ld, idx, k, nsub);
  if (st != PL_OK) return st;
  if (!out) return fail(PL_ERR_ARG, "out must be non-null");
  if (mode != 0 && mode != 1) return fail(PL_ERR_ARG, "gather mode must be 0 or 1");
  if (lanes == 0) lanes = group_for(k);
  if (lanes < k || !pl::solve_variant_exists(lanes, pad))
    return fail(PL_ERR_ARG, "no variant with %d lanes and pad %d for k = %d", lanes, pad, k);
  pl::SolveArgs a{};
  a.G = G;
  a.n = n;
  a.ld = ld;
  a.idx = idx;
  a.kuni = k;
  a.nslots = nsub;
  a.gs_out = out;
  hipError_t e = pl::launch_solve(lanes, pad, mode == 0 ? pl::MODE_DUMP : pl::MODE_CHECKSUM, a,
                                  (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "gather launch");
  return PL_OK;
}

pl_status pl_gen_subsets(uint64_t seed, int64_t s0, int64_t nsub, int k, int n, int32_t *idx,
                         void *stream) {
  if (!idx) return fail(PL_ERR_ARG, "idx must be non-null");
  if (n < 1 || k < 1 || k > n) return fail(PL_ERR_ARG, "need 1 <= k <= n (k=%d, n=%d)", k, n);
  if (nsub < 0 || s0 < 0) return fail(PL_ERR_ARG, "negative subset range");
  hipError_t e = pl::launch_gen_subsets(seed, s0, nsub, nullptr, k, n, idx, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "gen_subsets launch");
  return PL_OK;
}

pl_status pl_gen_subsets_csr(uint64_t seed, int64_t s0, int64_t nsub, const int64_t *offsets,
                             int n, int32_t *idx, void *stream) {
  if (!idx || !offsets) return fail(PL_ERR_ARG, "idx and offsets must be non-null");
  if (n < 1 || nsub < 0 || s0 < 0) return fail(PL_ERR_ARG, "bad n or subset range");
  hipError_t e = pl::launch_gen_subsets(seed, s0, nsub, offsets, 0, n, idx, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "gen_subsets_csr launch");
  return PL_OK;
}

pl_status pl_gen_ks(uint64_t seed, int64_t s0, int64_t nsub, const int32_t *klist, int nk,
                    int32_t *ks, void *stream) {
  if (!klist || !ks) return fail(PL_ERR_ARG, "klist and ks must be non-null");
  if (nk < 1 || nk > 32) return fail(PL_ERR_ARG, "nk = %d outside [1, 32]", nk);
  if (nsub < 0 || s0 < 0) return fail(PL_ERR_ARG, "negative subset range");
  hipError_t e = pl::launch_gen_ks(seed, s0, nsub, klist, nk, ks, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "gen_ks launch");
  return PL_OK;
}

pl_status pl_gen_phi(int kind, uint64_t seed, int m, int n, float *phi, void *stream) {
  if (!phi) return fail(PL_ERR_ARG, "phi must be non-null");
  if (m < 1 || n < 1) return fail(PL_ERR_ARG, "m and n must be positive");
  hipStream_t s = (hipStream_t)stream;
  hipError_t e;
  if (kind == 0) {
    e = pl::launch_gen_phi_gauss(seed, m, n, phi, s);
  } else if (kind == 1) {
    e = pl::launch_gen_phi_bernoulli(seed, m, n, phi, s);
  } else if (kind == 2) {
    if (m > n) return fail(PL_ERR_ARG, "DCT rows: m = %d > n = %d", m, n);
    // row set: Floyd sample of m rows of [0,n) from stream (seed, TAG_DCT, 0)
    std::vector<int32_t> rows((size_t)m);
    pl::Stream st(seed, pl::TAG_DCT, 0u);
    int cnt = 0;
    for (int j = n - m; j < n; ++j) {
      const int32_t t = (int32_t)st.bounded((uint32_t)j + 1u);
      bool seen = false;
      for (int q = 0; q < cnt && !seen; ++q) seen = rows[(size_t)q] == t;
      rows[(size_t)cnt++] = seen ? j : t;
    }
    int32_t *drows = nullptr;
    e = hipMalloc(&drows, sizeof(int32_t) * (size_t)m);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(dct rows)");
    e = hipMemcpyAsync(drows, rows.data(), sizeof(int32_t) * (size_t)m, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = pl::launch_gen_phi_dct(drows, m, n, phi, s);
    hipError_t e2 = hipStreamSynchronize(s);
    if (e == hipSuccess) e = e2;
    (void)hipFree(drows);
  } else {
    return fail(PL_ERR_ARG, "unknown phi kind %d", kind);
  }
  if (e != hipSuccess) return hip_fail(e, "gen_phi");
  return PL_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Drop-in host entry points (ParaLanczos.h)
// ---------------------------------------------------------------------------
namespace {

std::mutex g_legacy_mu;

bool quiet() {
  const char *q = getenv("PARALANCZOS_QUIET");
  return q && *q && strcmp(q, "0") != 0;
}

template <class T>
struct DevBuf {
  T *p = nullptr;
  hipError_t alloc(size_t count) { return hipMalloc(&p, sizeof(T) * (count ? count : 1)); }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

// Uploads mat (FP64 -> FP32) and the index sets, runs the all-eigenvalue solve
// and copies back n_eig eigenvalues, n_det determinants and n_err status words.
// Returns 0 on success, -1 (with a message on stderr) on failure.
int legacy_solve(const double *mat, int matsize, const int *subIndex, int L, int subn,
                 double *eig_out, int64_t n_eig, double *det_out, int64_t n_det, int *err_out,
                 int64_t n_err) {
  if (subn < 1 || subn > kMaxK) {
    fprintf(stderr, " subset size %d is not supported by this build (1..%d) \n", subn, kMaxK);
    return -1;
  }
  if (L < subn || L % subn != 0 || matsize < 1) {
    fprintf(stderr, " index length %d is not a positive multiple of the subset size %d \n", L,
            subn);
    return -1;
  }
  for (int i = 0; i < L; ++i)
    if (subIndex[i] < 0 || subIndex[i] >= matsize) {
      fprintf(stderr, " subIndex[%d] = %d is outside [0, %d) \n", i, subIndex[i], matsize);
      return -1;
    }
  const int64_t nsub = L / subn;
  const size_t nn = (size_t)matsize * (size_t)matsize;
  std::vector<float> g32(nn);
  for (size_t i = 0; i < nn; ++i) g32[i] = (float)mat[i];
  DevBuf<float> dG;
  DevBuf<int32_t> dIdx, dErr;
  DevBuf<double> dEig, dDet;
  hipError_t e = dG.alloc(nn);
  if (e == hipSuccess) e = dIdx.alloc((size_t)L);
  if (e == hipSuccess) e = dEig.alloc((size_t)L);
  if (e == hipSuccess) e = dDet.alloc((size_t)nsub);
  if (e == hipSuccess) e = dErr.alloc((size_t)nsub);
  if (e == hipSuccess) e = hipMemcpy(dG.p, g32.data(), sizeof(float) * nn, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(dIdx.p, subIndex, sizeof(int32_t) * (size_t)L, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    fprintf(stderr, " ParaLanczos: %s \n", hipGetErrorString(e));
    return -1;
  }
  pl_status st =
      pl_solve_alleig(dG.p, matsize, matsize, dIdx.p, subn, nsub, dEig.p, dDet.p, dErr.p, nullptr);
  if (st != PL_OK) {
    fprintf(stderr, " ParaLanczos: %s \n", pl_last_error());
    return -1;
  }
  e = hipStreamSynchronize(nullptr);
  if (e == hipSuccess && n_eig > 0)
    e = hipMemcpy(eig_out, dEig.p, sizeof(double) * (size_t)n_eig, hipMemcpyDeviceToHost);
  if (e == hipSuccess && n_det > 0)
    e = hipMemcpy(det_out, dDet.p, sizeof(double) * (size_t)n_det, hipMemcpyDeviceToHost);
  if (e == hipSuccess && n_err > 0)
    e = hipMemcpy(err_out, dErr.p, sizeof(int32_t) * (size_t)n_err, hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    fprintf(stderr, " ParaLanczos: %s \n", hipGetErrorString(e));
    return -1;
  }
  return 0;
}

void print_eigs(const double *eig, int count, const char *word) {
  if (quiet()) return;
  for (int i = 0; i < count; ++i) printf("%d %s: %g \n", i, word, eig[i]);
  fflush(stdout);
}

}  // namespace

extern "C" {

void print_tridiagonal(double *alpha, double *beta, int n) {
  printf(" [ ");
  for (int r = 0; r < n; ++r) {
    for (int c = 0; c < n; ++c) {
      double x = 0.0;
      if (r == c) x = alpha[c];
      else if (c == r - 1) x = beta[c + 1];
      else if (c == r + 1) x = beta[c];
      printf(" %g ", x);
      if (c != n - 1) printf(", ");
    }
    printf(r != n - 1 ? "; \n" : "] \n");
  }
  fflush(stdout);
}

void calculate_eigs(double *mat, int n, double *Eigenvalues, double *Det) {
  if (n % 4 != 0 || n > 128 || n < 1) {
    fprintf(stderr, " 'int n' must be divisible by 4, and less than or equal to 128! \n");
    return;
  }
  std::lock_guard<std::mutex> lock(g_legacy_mu);
  std::vector<int> ident((size_t)n);
  for (int i = 0; i < n; ++i) ident[(size_t)i] = i;
  if (legacy_solve(mat, n, ident.data(), n, n, Eigenvalues, n, Det, 1, nullptr, 0) != 0) return;
  print_eigs(Eigenvalues, n, "eigevalue");
}

void calculate_subeigs(double *mat, int n, double *Eigenvalues, double *Det, int *subIndex,
                       int subn) {
  if (subn % 4 != 0 || n > 128 || subn < 1) {
    fprintf(stderr, " 'int n' must be divisible by 4, and less than or equal to 128! \n");
    return;
  }
  std::lock_guard<std::mutex> lock(g_legacy_mu);
  if (legacy_solve(mat, n, subIndex, subn, subn, Eigenvalues, subn, Det, 1, nullptr, 0) != 0)
    return;
  print_eigs(Eigenvalues, subn, "eigenvalue");
}

void calculate_subeigs2(double *mat, int n, int matsize, double *Eigenvalues, double *Det,
                        int *subIndex, int subn) {
  if (subn % 4 != 0 || n > 128 || subn < 1) {
    fprintf(stderr, " 'int n' must be divisible by 4, and less than or equal to 128! \n");
    return;
  }
  std::lock_guard<std::mutex> lock(g_legacy_mu);
  if (!quiet()) {
    printf("welcome to calculate subeigs2...\n");
    fflush(stdout);
  }
  if (legacy_solve(mat, matsize, subIndex, n, subn, Eigenvalues, n, Det, 1, nullptr, 0) != 0)
    return;
  print_eigs(Eigenvalues, n, "eigenvalue");
}

void calculate_subeigs3(double *mat, int subIndexLen, int matsize, double *Eigenvalues,
                        double *Det, int *subIndex, int subn) {
  if (subn % 4 != 0 || subn > 128 || subn < 1) {
    fprintf(stderr, " 'int subn' must be divisible by 4, and less than or equal to 128! \n");
    return;
  }
  if (subIndexLen % 32 != 0) {
    fprintf(stderr, " 'subIndexLen' must be divisible by 32! \n");
    return;
  }
  std::lock_guard<std::mutex> lock(g_legacy_mu);
  if (!quiet()) {
    printf("welcome to calculate subeigs2...\n");
    fflush(stdout);
  }
  if (legacy_solve(mat, matsize, subIndex, subIndexLen, subn, Eigenvalues, subIndexLen, Det,
                   subIndexLen / subn, nullptr, 0) != 0)
    return;
  print_eigs(Eigenvalues, subIndexLen, "eigenvalue");
}

int pl_multi_subsvd(const double *g, int M, const int *subIndex, int L, int minor, double *eigs,
                    double *det, int *err) {
  if (!g || !subIndex || !eigs || !det || !err || M < 1 || minor < 1 || L < minor ||
      L % minor != 0)
    return -1;
  std::lock_guard<std::mutex> lock(g_legacy_mu);
  return legacy_solve(g, M, subIndex, L, minor, eigs, L, det, L / minor, err, L / minor);
}

}  // extern "C"
