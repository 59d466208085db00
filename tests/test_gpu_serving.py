"""The serving engine (continuous batching against a device-resident G):
every request's results equal a direct solve of the same request, alone or
in a mix, bit for bit; admission splits the queue into steps within the
arena and work budgets; the values match the oracle (relative error <= 1e-5)."""
import numpy as np
import pytest
import torch

import oracle
from paralanczos_amd import ParaLanczosError, solver

pytestmark = pytest.mark.gpu

RTOL = 1e-5
KS = [16, 32, 64, 128, 256]


def rel(x, ref):
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    return np.abs(x - ref) / np.maximum(np.abs(ref), 1e-30)


@pytest.fixture(scope="module")
def env(gpu):
    n = 4096
    G = solver.gram(solver.gen_phi("gauss", 0xC5, 1024, n))
    rng = np.random.default_rng(0xC5)
    reqs = []
    for r in range(30):
        nsub = int(np.exp(rng.uniform(0.0, np.log(300.0))))
        ks = solver.gen_ks(1000 + r, nsub, KS)
        off = solver.offsets_from_ks(ks)
        idx = solver.gen_subsets_csr(2000 + r, off, n)
        reqs.append((idx, off, int(ks.max().item()), nsub))
    torch.cuda.synchronize()
    return G, reqs


def _np(ts):
    return [t.cpu().numpy() for t in ts]


def test_alone_vs_mix(env):
    G, reqs = env
    direct = [_np(solver.solve(G, idx, k=256, offsets=off)) for idx, off, _, _ in reqs]
    srv = solver.Server(G, max_subsets=2048, max_indices=2048 * 256, depth=3)
    try:
        idx0, off0, km0, _ = reqs[0]
        out, lat, _ = srv.wait(srv.submit(idx0, offsets=off0, kmax=km0))
        alone = _np(out)
        for a, b in zip(alone, direct[0]):
            assert np.array_equal(a, b)
        # the mix: all 30, request 0 in the middle of them
        order = list(range(1, 15)) + [0] + list(range(15, 30))
        tickets = {r: srv.submit(reqs[r][0], offsets=reqs[r][1], kmax=reqs[r][2]) for r in order}
        lats = []
        for r in order:
            out, lat, q = srv.wait(tickets[r])
            assert lat >= q >= 0
            lats.append(lat)
            got = _np(out)
            for a, b in zip(got, direct[r]):
                assert np.array_equal(a, b), r
            if r == 0:
                for a, b in zip(got, alone):
                    assert np.array_equal(a, b)
        st = srv.stats()
        assert st["requests"] == 31 and st["subsets"] == sum(x[3] for x in reqs) + reqs[0][3]
        assert 0 < st["p50_s"] <= st["p99_s"] <= st["max_s"]
    finally:
        srv.close()
    # the values: the oracle on three requests
    Gh = G.cpu().numpy()
    for r in (0, 7, 21):
        idx, off, _, _ = reqs[r]
        olo, ohi, _, ost = oracle.solve_batch(Gh, idx.cpu().numpy(), offsets=off.cpu().numpy(),
                                              nthreads=8)
        assert np.all(ost == 0) and np.all(direct[r][3] == 0)
        assert rel(direct[r][0], olo).max() <= RTOL and rel(direct[r][1], ohi).max() <= RTOL


def test_admission_splits_steps(env):
    G, reqs = env
    # an arena of 64 subsets, 4 requests and a work budget of about two k = 256
    # subsets per step: the queue is cut into many steps, results unchanged
    budget = 2 * 154 * 1024 * 256
    srv = solver.Server(G, max_subsets=300, max_indices=300 * 256, max_requests=4, depth=2,
                        work_budget=budget)
    try:
        with pytest.raises(ParaLanczosError):
            big = solver.gen_subsets(5, 301, 16, G.shape[0])
            srv.submit(big, k=16)
        tickets = [srv.submit(idx, offsets=off, kmax=km) for idx, off, km, _ in reqs]
        for (idx, off, km, nsub), t in zip(reqs, tickets):
            out, lat, q = srv.wait(t)
            ref = _np(solver.solve(G, idx, k=256, offsets=off))
            for a, b in zip(_np(out), ref):
                assert np.array_equal(a, b)
        st = srv.stats(reset=True)
        assert st["requests"] == len(reqs) and st["steps"] >= len(reqs) // 4
        assert srv.stats()["requests"] == 0
    finally:
        srv.close()


def test_uniform_requests_and_precision_form(env):
    G, _ = env
    n = G.shape[0]
    srv = solver.Server(G, max_subsets=4096, max_indices=4096 * 256, depth=2, reorth="f32")
    try:
        jobs = []
        for i, k in enumerate(KS):
            idx = solver.gen_subsets(70 + i, 100 + 37 * i, k, n)
            jobs.append((idx, k, srv.submit(idx, k=k)))
        for idx, k, t in jobs:
            out, _, _ = srv.wait(t)
            ref = _np(solver.solve(G, idx, k=k, reorth="f32"))
            for a, b in zip(_np(out), ref):
                assert np.array_equal(a, b), k
    finally:
        srv.close()


def test_malformed_request_stays_in_its_own_results(env):
    """Requests whose offsets do not run from 0 to nidx (a slice of a larger
    CSR batch: off[0] > 0; a truncated one: off[nsub] < nidx) are flagged
    BADK with NaN outputs as a whole, and the requests packed beside them in
    the same step keep their results bit for bit."""
    G, reqs = env
    good = [reqs[r] for r in (2, 5, 9)]
    direct = [_np(solver.solve(G, idx, k=256, offsets=off)) for idx, off, _, _ in good]
    shifted = reqs[3][1] + 2                      # off[0] = 2, off[nsub] = nidx + 2
    trunc = reqs[4][1].clone()
    trunc[-1] -= 1                                # off[nsub] = nidx - 1
    bad = [(reqs[3][0], shifted, reqs[3][2]), (reqs[4][0], trunc, reqs[4][2])]
    srv = solver.Server(G, max_subsets=4096, max_indices=4096 * 256, depth=1)
    try:
        tickets = [srv.submit(good[0][0], offsets=good[0][1], kmax=good[0][2]),
                   srv.submit(bad[0][0], offsets=bad[0][1], kmax=bad[0][2]),
                   srv.submit(good[1][0], offsets=good[1][1], kmax=good[1][2]),
                   srv.submit(bad[1][0], offsets=bad[1][1], kmax=bad[1][2]),
                   srv.submit(good[2][0], offsets=good[2][1], kmax=good[2][2])]
        outs = [_np(srv.wait(t)[0]) for t in tickets]
    finally:
        srv.close()
    for got, ref in zip((outs[0], outs[2], outs[4]), direct):
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)
    for got in (outs[1], outs[3]):
        assert np.all(got[3] == solver.SUB_BADK)
        assert np.all(np.isnan(got[0])) and np.all(np.isnan(got[1])) and np.all(np.isnan(got[2]))
