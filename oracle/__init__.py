"""CPU oracle for the batched-submatrix Lanczos eigenvalue path.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` as the checker; the product
package ``paralanczos_amd`` never imports it.

Thin ctypes binding over ``oracle/liboracle.so`` (built from ``oracle.c`` by
``oracle/Makefile``). Every function takes and returns numpy arrays.  The C
source documents which reference lines each function restates and what pins it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

ST_BREAKDOWN = 1
ST_NONFINITE = 2
ST_NOCONV = 4
ST_BADINDEX = 8
ST_BADK = 16
TAU = 1.0 / 1073741824.0


def build(force: bool = False) -> str:
    """Compile ``liboracle.so`` with gcc (seconds)."""
    src = os.path.join(HERE, "oracle.c")
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(src):
        subprocess.run(["make", "-s"] + (["-B"] if force else []) + ["-C", HERE],
                       check=True, capture_output=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


P = ctypes.c_void_p
I32, I64, U32, U64, DBL = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_double


def _declare(L):
    sig = {
        "or_philox4x32_10": (None, [P, P, P]),
        "or_stream_word": (U32, [U64, U32, U64, U32]),
        "or_floyd": (None, [U64, U64, I32, I32, P]),
        "or_gen_subsets_uniform": (None, [U64, I64, I64, I32, I32, P]),
        "or_gen_ks": (None, [U64, I64, I64, P, I32, P]),
        "or_gen_subsets_csr": (None, [U64, I64, I64, P, I32, P]),
        "or_gen_phi_gauss": (None, [U64, I32, I32, P]),
        "or_gen_phi_bernoulli": (None, [U64, I32, I32, P]),
        "or_dct_rows": (None, [U64, I32, I32, P]),
        "or_gen_phi_dct": (None, [U64, I32, I32, P]),
        "or_gather": (U32, [P, I64, I64, P, I32, P]),
        "or_lanczos": (U32, [P, I32, P, P]),
        "or_lanczos_noreorth": (None, [P, I32, P, P]),
        "or_sturm_count": (ctypes.c_int, [P, P, I32, DBL]),
        "or_sturm_eig": (DBL, [P, P, I32, I32, P]),
        "or_tqlrat": (ctypes.c_int, [I32, P, P]),
        "or_solve_batch": (None, [P, I64, I64, P, P, I32, I64, P, P, P, P, I32]),
        "or_solve_batch_alleig": (None, [P, I64, I64, P, I32, I64, P, P, P, I32]),
        "or_solve_batch_faithful": (None, [P, I64, P, I32, I64, P, P, P, I32]),
        "or_faithful_extremes": (None, [P, I64, I64, P, P, I32, I64, P, P, P, I32]),
        "or_upper": (DBL, [P, I32, I32]),
        "or_upper_batch": (None, [P, I64, I64, P, P, I32, I64, I32, P, P, I32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def philox4x32_10(ctr, key):
    c = _c(ctr, np.uint32)
    k = _c(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().or_philox4x32_10(_p(c), _p(k), _p(out))
    return out


def stream_word(seed, tag, sid, w):
    return int(lib().or_stream_word(seed, tag, sid, w))


def floyd(seed, sid, n, k):
    out = np.zeros(k, np.int32)
    lib().or_floyd(seed, sid, n, k, _p(out))
    return out


def gen_subsets_uniform(seed, nsub, k, n, s0=0):
    idx = np.zeros(nsub * k, np.int32)
    lib().or_gen_subsets_uniform(seed, s0, nsub, k, n, _p(idx))
    return idx


def gen_ks(seed, nsub, klist, s0=0):
    kl = _c(klist, np.int32)
    ks = np.zeros(nsub, np.int32)
    lib().or_gen_ks(seed, s0, nsub, _p(kl), len(kl), _p(ks))
    return ks


def gen_subsets_csr(seed, offsets, n, s0=0):
    off = _c(offsets, np.int64)
    idx = np.zeros(int(off[-1]), np.int32)
    lib().or_gen_subsets_csr(seed, s0, len(off) - 1, _p(off), n, _p(idx))
    return idx


def gen_phi(kind, seed, m, n):
    """Column-major m x n sensing matrix, returned as an (n, m) array (row j = column j)."""
    out = np.zeros((n, m), np.float32)
    fn = {"gauss": lib().or_gen_phi_gauss, "bernoulli": lib().or_gen_phi_bernoulli,
          "dct": lib().or_gen_phi_dct}[kind]
    fn(seed, m, n, _p(out))
    return out


def dct_rows(seed, m, n):
    out = np.zeros(m, np.int32)
    lib().or_dct_rows(seed, m, n, _p(out))
    return out


def gather(G, idx, ld=None):
    G = _c(G, np.float32)
    n = G.shape[0]
    ld = G.shape[1] if ld is None else ld
    ix = _c(idx, np.int32)
    k = len(ix)
    A = np.zeros((k, k), np.float64)
    st = lib().or_gather(_p(G), n, ld, _p(ix), k, _p(A))
    return A, int(st)


def lanczos(A):
    A = _c(A, np.float64)
    k = A.shape[0]
    al = np.zeros(k)
    be = np.zeros(k)
    st = lib().or_lanczos(_p(A), k, _p(al), _p(be))
    return al, be, int(st)


def lanczos_noreorth(A):
    A = _c(A, np.float64)
    k = A.shape[0]
    al = np.zeros(k)
    be = np.zeros(k)
    lib().or_lanczos_noreorth(_p(A), k, _p(al), _p(be))
    return al, be


def sturm_count(alpha, beta, sigma):
    a = _c(alpha, np.float64)
    b = _c(beta, np.float64)
    return int(lib().or_sturm_count(_p(a), _p(b), len(a), sigma))


def sturm_eig(alpha, beta, t):
    a = _c(alpha, np.float64)
    b = _c(beta, np.float64)
    nit = np.zeros(1, np.int32)
    return float(lib().or_sturm_eig(_p(a), _p(b), len(a), t, _p(nit))), int(nit[0])


def tqlrat(d, e2):
    d = np.array(d, np.float64)
    e = np.array(e2, np.float64)
    ierr = lib().or_tqlrat(len(d), _p(d), _p(e))
    return d, int(ierr)


def solve_batch(G, idx, k=None, offsets=None, nthreads=1, ld=None):
    """lambda_min, lambda_max, delta (FP64) and status per subset."""
    G = _c(G, np.float32)
    n = G.shape[0]
    ld = G.shape[1] if ld is None else ld
    ix = _c(idx, np.int32)
    if offsets is not None:
        off = _c(offsets, np.int64)
        nsub = len(off) - 1
        offp = _p(off)
        kk = 0
    else:
        nsub = len(ix) // k
        offp = None
        kk = k
    lmin = np.zeros(nsub)
    lmax = np.zeros(nsub)
    delta = np.zeros(nsub)
    st = np.zeros(nsub, np.uint32)
    lib().or_solve_batch(_p(G), n, ld, _p(ix), offp, kk, nsub, _p(lmin), _p(lmax),
                         _p(delta), _p(st), nthreads)
    return lmin, lmax, delta, st


def solve_batch_alleig(G, idx, k, nthreads=1, ld=None):
    G = _c(G, np.float32)
    n = G.shape[0]
    ld = G.shape[1] if ld is None else ld
    ix = _c(idx, np.int32)
    nsub = len(ix) // k
    eig = np.zeros(nsub * k)
    det = np.zeros(nsub)
    st = np.zeros(nsub, np.uint32)
    lib().or_solve_batch_alleig(_p(G), n, ld, _p(ix), k, nsub, _p(eig), _p(det), _p(st),
                                nthreads)
    return eig, det, st


def solve_batch_faithful(G, idx, k, nthreads=1):
    G = _c(G, np.float64)
    n = G.shape[0]
    ix = _c(idx, np.int32)
    nsub = len(ix) // k
    eig = np.zeros(nsub * k)
    det = np.zeros(nsub)
    ierr = np.zeros(nsub, np.int32)
    lib().or_solve_batch_faithful(_p(G), n, _p(ix), k, nsub, _p(eig), _p(det), _p(ierr),
                                  nthreads)
    return eig, det, ierr


def upper(A, iters=10):
    """Power-iteration estimate of lambda_max of a dense matrix (FP64)."""
    A = _c(A, np.float64)
    return float(lib().or_upper(_p(A), A.shape[0], int(iters)))


def upper_batch(G, idx, k=None, offsets=None, iters=10, nthreads=1, ld=None):
    """Per-subset power-iteration estimates (FP64) and status."""
    G = _c(G, np.float32)
    n = G.shape[0]
    ld = G.shape[1] if ld is None else ld
    ix = _c(idx, np.int32)
    if offsets is not None:
        off = _c(offsets, np.int64)
        nsub = len(off) - 1
        offp, kk = _p(off), 0
    else:
        nsub = len(ix) // k
        offp, kk = None, k
    up = np.zeros(nsub)
    st = np.zeros(nsub, np.uint32)
    lib().or_upper_batch(_p(G), n, ld, _p(ix), offp, kk, nsub, int(iters), _p(up), _p(st),
                         nthreads)
    return up, st


def faithful_extremes(G, idx, k=None, offsets=None, nthreads=1, ld=None):
    """The reference's path as it stands (FP64, no reorthogonalisation,
    TQLRAT over all eigenvalues) on FP32 G: lambda_min, lambda_max, ierr."""
    G = _c(G, np.float32)
    n = G.shape[0]
    ld = G.shape[1] if ld is None else ld
    ix = _c(idx, np.int32)
    if offsets is not None:
        off = _c(offsets, np.int64)
        nsub = len(off) - 1
        offp, kk = _p(off), 0
    else:
        nsub = len(ix) // k
        offp, kk = None, k
    lo = np.zeros(nsub)
    hi = np.zeros(nsub)
    ierr = np.zeros(nsub, np.int32)
    lib().or_faithful_extremes(_p(G), n, ld, _p(ix), offp, kk, nsub, _p(lo), _p(hi), _p(ierr),
                               nthreads)
    return lo, hi, ierr
