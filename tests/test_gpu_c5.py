"""Parity on the serving workload itself: Phi is 4096 seeded rows of the
32768-point orthonormal DCT-II scaled by sqrt(32768/4096), G = Phi^T Phi is
32768 x 32768 FP32 (4 GiB, built on the device by pl_gram), and a wave of
requests of log-uniform sizes with k in {16, 32, 64, 128, 256} is served by
the continuous-batching engine.

* Served results equal a direct tuned solve of each request bit for bit.
* For at least 100 subsets of every k: lambda_min and lambda_max within
  1e-5 relative of the FP64 oracle (the reference recurrence of
  ParaLanczos.cu:1285-1374 with full reorthogonalisation and Sturm
  bisection) and of FP64 LAPACK dsyevd on the gathered FP32 G_S; status 0.
* The plain form (one launch per k, FP32 throughout) on the same subsets
  meets the same bar.
* Subsets with a duplicated index (a singular G_S whose Krylov space from
  the ones vector misses the antisymmetric null vector) keep
  lambda_min = 0 to 1e-5 lambda_max and lambda_max to 1e-5.
"""
import numpy as np
import pytest
import torch

import oracle
from paralanczos_amd import solver

pytestmark = pytest.mark.gpu

N, M = 32768, 4096
KS = [16, 32, 64, 128, 256]
RTOL = 1e-5
PER_K = 100


def rel(x, ref):
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    return np.abs(x - ref) / np.maximum(np.abs(ref), 1e-30)


def _np(ts):
    return [t.cpu().numpy() for t in ts]


@pytest.fixture(scope="module")
def c5(gpu):
    phi = solver.gen_phi("dct", 0xC5, M, N)
    G = solver.gram(phi)
    del phi
    rng = np.random.default_rng(0xC5 + 3)
    reqs = []
    for r in range(20):
        nsub = max(1, int(np.exp(rng.uniform(0.0, np.log(1024.0)))))
        ks = solver.gen_ks(0xC5 + 1, nsub, KS, s0=r * 4096)
        off = solver.offsets_from_ks(ks)
        idx = solver.gen_subsets_csr(0xC5 + 1, off, N, s0=r * 4096)
        reqs.append((idx, off, nsub))
    torch.cuda.synchronize()
    Gh = G.cpu().numpy()
    yield G, Gh, reqs
    del G, Gh
    torch.cuda.empty_cache()


def _sample(reqs, outs):
    """The first PER_K subsets of each k across the wave, in request order:
    their index sets, sizes and served outputs."""
    sets, ks, got = [], [], []
    count = {k: 0 for k in KS}
    for (idx, off, nsub), out in zip(reqs, outs):
        o = off.cpu().numpy()
        ix = idx.cpu().numpy()
        for s in range(nsub):
            k = int(o[s + 1] - o[s])
            if count[k] >= PER_K:
                continue
            count[k] += 1
            sets.append(ix[o[s]:o[s + 1]])
            ks.append(k)
            got.append([a[s] for a in out])
    assert all(c == PER_K for c in count.values()), count
    return sets, np.array(ks), np.array(got, dtype=np.float64)


def test_c5_served_wave_vs_oracle_and_lapack(c5):
    G, Gh, reqs = c5
    srv = solver.Server(G, max_subsets=16384, max_indices=16384 * 256, depth=3)
    try:
        tickets = [srv.submit(idx, offsets=off, kmax=256) for idx, off, _ in reqs]
        served = [_np(srv.wait(t)[0]) for t in tickets]
    finally:
        srv.close()
    for (idx, off, _), out in zip(reqs, served):
        direct = _np(solver.solve(G, idx, k=256, offsets=off))
        for a, b in zip(out, direct):
            assert np.array_equal(a, b)
    sets, ks, got = _sample(reqs, served)
    offs = np.concatenate([[0], np.cumsum(ks)]).astype(np.int64)
    flat = np.concatenate(sets).astype(np.int32)
    olo, ohi, odel, ost = oracle.solve_batch(Gh, flat, offsets=offs, nthreads=16)
    assert np.all(ost == 0), np.unique(ost)
    assert np.all(got[:, 3] == 0)
    for k in KS:
        m = ks == k
        assert rel(got[m, 0], olo[m]).max() <= RTOL, k
        assert rel(got[m, 1], ohi[m]).max() <= RTOL, k
        assert np.abs(got[m, 2] - odel[m]).max() <= RTOL * np.maximum(1.0, ohi[m]).max(), k
    # FP64 LAPACK on the gathered FP32 G_S: every tenth sample subset
    for j in range(0, len(sets), 10):
        A, st = oracle.gather(Gh, sets[j])
        assert st == 0
        w = np.linalg.eigvalsh(A)
        assert rel(got[j, 0], w[0]) <= RTOL and rel(got[j, 1], w[-1]) <= RTOL, (j, ks[j])


def test_c5_plain_form_vs_oracle(c5):
    G, Gh, reqs = c5
    outs = [_np(solver.solve(G, idx, k=256, offsets=off, form="plain")) for idx, off, _ in reqs]
    sets, ks, got = _sample(reqs, outs)
    offs = np.concatenate([[0], np.cumsum(ks)]).astype(np.int64)
    flat = np.concatenate(sets).astype(np.int32)
    olo, ohi, _, ost = oracle.solve_batch(Gh, flat, offsets=offs, nthreads=16)
    assert np.all(ost == 0) and np.all(got[:, 3] == 0)
    for k in KS:
        m = ks == k
        assert rel(got[m, 0], olo[m]).max() <= RTOL, k
        assert rel(got[m, 1], ohi[m]).max() <= RTOL, k


@pytest.mark.parametrize("form", ["tuned", "plain"])
def test_c5_duplicate_index_subsets(c5, form):
    G, Gh, _ = c5
    rows = []
    for i, k in enumerate(KS):
        for r in range(4):
            s = oracle.floyd(0xC5 + 7, 10 * i + r, N, k)
            s[1 + r] = s[k - 1 - r]  # one duplicated index
            rows.append(s)
    ks = np.array([len(s) for s in rows])
    offs = np.concatenate([[0], np.cumsum(ks)]).astype(np.int64)
    flat = np.concatenate(rows).astype(np.int32)
    lo, hi, de, st = _np(solver.solve(G, torch.from_numpy(flat).cuda(), k=256,
                                      offsets=torch.from_numpy(offs).cuda(), form=form))
    for j, s in enumerate(rows):
        A, gst = oracle.gather(Gh, s)
        assert gst == 0
        w = np.linalg.eigvalsh(A)
        assert (st[j] & ~1) == 0, (j, st[j])
        assert abs(lo[j] - w[0]) <= RTOL * w[-1], (j, lo[j], w[0])
        assert rel(hi[j], w[-1]) <= RTOL, (j, hi[j], w[-1])
