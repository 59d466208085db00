"""PL_FORM_PLAIN_WAVE: one 64-lane wavefront per subset at every k (beyond
k = 64 it streams G_S and the basis from a global scratch), against the
oracle (relative eigenvalue error <= 1e-5), the other forms, and its
argument checks."""
import ctypes

import numpy as np
import pytest
import torch

import oracle
from paralanczos_amd import _lib, solver

pytestmark = pytest.mark.gpu

RTOL = 1e-5


def rel(x, ref):
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    return np.abs(x - ref) / np.maximum(np.abs(ref), 1e-30)


def _np(ts):
    return [t.cpu().numpy() for t in ts]


@pytest.mark.parametrize("k", [8, 64, 65, 100, 128, 129, 200, 256])
def test_plain_wave_vs_oracle(gpu, k):
    n = 1024
    G = solver.gram(solver.gen_phi("gauss", 300 + k, max(4 * k, 64), n))
    idx = solver.gen_subsets(31 + k, 40, k, n)
    olo, ohi, _, ost = oracle.solve_batch(G.cpu().numpy(), idx.cpu().numpy(), k=k, nthreads=8)
    assert np.all(ost == 0)
    lo, hi, de, st = _np(solver.solve(G, idx, k=k, form="plain_wave"))
    assert np.all(st == 0)
    assert rel(lo, olo).max() <= RTOL and rel(hi, ohi).max() <= RTOL
    np.testing.assert_allclose(de, np.maximum(hi - 1.0, 1.0 - lo), rtol=1e-6, atol=1e-7)
    if k <= 64:  # the same 64-lane kernel as the plain form
        p = _np(solver.solve(G, idx, k=k, form="plain"))
        assert np.array_equal(lo, p[0]) and np.array_equal(hi, p[1])


def test_plain_wave_mixed_batch(gpu):
    """A mixed-size batch (k in 8..256, one launch per k) agrees with the
    tuned form to the oracle tolerance; more subsets than resident
    wavefronts reuse the scratch slots."""
    n = 2048
    G = solver.gram(solver.gen_phi("bernoulli", 12, 1024, n))
    ks = solver.gen_ks(13, 3000, [8, 16, 64, 72, 128, 130, 256])
    off = solver.offsets_from_ks(ks)
    idx = solver.gen_subsets_csr(14, off, n)
    kmax = int(ks.max().item())
    w = _np(solver.solve(G, idx, k=kmax, offsets=off, form="plain_wave"))
    t = _np(solver.solve(G, idx, k=kmax, offsets=off))
    assert not w[3].any() and not t[3].any()
    assert rel(w[0], t[0]).max() <= RTOL and rel(w[1], t[1]).max() <= RTOL
    big = solver.gen_subsets(15, 2500, 100, n)  # > 2048 resident slots
    a = _np(solver.solve(G, big, k=100, form="plain_wave"))
    b = _np(solver.solve(G, big, k=100, form="plain"))
    assert not a[3].any()
    assert rel(a[0], b[0]).max() <= RTOL and rel(a[1], b[1]).max() <= RTOL


def test_plain_wave_flags(gpu):
    """Bad indices and sizes are flagged per subset; precision flags are
    refused (the form is FP32 only)."""
    n = 500
    G = solver.gram(solver.gen_phi("gauss", 2, 128, n))
    idx = solver.gen_subsets(3, 4, 100, n)
    idx[100 + 7] = n
    st = _np(solver.solve(G, idx, k=100, form="plain_wave"))[3]
    assert st[0] == 0 and st[2] == 0 and st[3] == 0 and st[1] == 8
    L = _lib.load()
    P = lambda x: ctypes.c_void_p(x.data_ptr())
    o = [torch.empty(4, device="cuda") for _ in range(3)] + [torch.empty(4, dtype=torch.int32, device="cuda")]
    rc = L.pl_solve(P(G), n, n, P(idx), None, 100, 4, *[P(x) for x in o], 2 | 0x200,
                    ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc != 0


def test_plain_wave_rank_deficient(gpu):
    """m = 96 < k: G_S has rank 96, the Krylov space of the start vector is
    exhausted early and the recurrence restarts (BREAKDOWN) or continues on
    the reorthogonalised remainder; the extremes still agree with LAPACK,
    lambda_min = 0 to 1e-5 of lambda_max."""
    n = 600
    G = solver.gram(solver.gen_phi("gauss", 5, 96, n))
    Gh = G.cpu().numpy()
    for k in (100, 200):
        idx = solver.gen_subsets(60 + k, 8, k, n)
        lo, hi, de, st = _np(solver.solve(G, idx, k=k, form="plain_wave"))
        assert np.all((st & ~1) == 0)
        ih = idx.cpu().numpy()
        for s in range(8):
            A, _ = oracle.gather(Gh, ih[s * k:(s + 1) * k])
            w = np.linalg.eigvalsh(A)
            assert abs(lo[s] - w[0]) <= 1e-5 * w[-1], (k, s)
            assert abs(hi[s] - w[-1]) <= 1e-5 * w[-1], (k, s)
