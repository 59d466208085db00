"""Mixed-size batching on the GPU: device-side binning, persistent per-bin
launches, graph-captured plans and the autotuner.

Within a kernel family a subset's arithmetic does not depend on the lane
count or LDS padding (zero padding), so the tuned binned form, the plain
one-launch-per-k form and the graph replays agree bit for bit; the oracle
pins the values themselves (relative error <= 1e-5).
"""
import numpy as np
import pytest
import torch

import oracle
from paralanczos_amd import solver

pytestmark = pytest.mark.gpu

RTOL = 1e-5


def rel(x, ref):
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    return np.abs(x - ref) / np.maximum(np.abs(ref), 1e-30)


@pytest.fixture(scope="module")
def mixed(gpu):
    n = 2048
    G = solver.gram(solver.gen_phi("gauss", 0xC3, 512, n))
    ks = solver.gen_ks(0xB1, 3000, list(range(8, 129, 8)) + [1, 3, 5, 13, 31, 33, 65, 127])
    off = solver.offsets_from_ks(ks)
    idx = solver.gen_subsets_csr(0xB2, off, n)
    torch.cuda.synchronize()
    return G, ks, off, idx


def _np(ts):
    return [t.cpu().numpy() for t in ts]


def test_tuned_plain_variant_bitexact(mixed):
    """With FP32 Gram-Schmidt coefficients at every k (the tuned form switches
    to FP64 ones at k = 128 by default) the tuned and plain forms agree bit for
    bit up to k = 64, where both use the lane-group family. Beyond, the tuned
    form uses the 256-lane half-row kernel and the plain form the 128-lane
    kernel, each with its own summation order: they agree to 1e-5, and one
    launch of either kernel over every size agrees bit for bit with the form
    that uses it."""
    G, ks, off, idx = mixed
    tuned = _np(solver.solve(G, idx, k=128, offsets=off, form="tuned", reorth="f32"))
    plain = _np(solver.solve(G, idx, k=128, offsets=off, form="plain"))
    s128 = _np(solver.solve(G, idx, k=128, offsets=off, lanes=128, pad=4))
    s256 = _np(solver.solve(G, idx, k=128, offsets=off, lanes=256, pad=4))
    kk0 = ks.cpu().numpy()
    lo, hi = kk0 <= 64, kk0 > 64
    for a, b, c, d in zip(tuned, plain, s128, s256):
        assert np.array_equal(a[lo], b[lo], equal_nan=True)
        assert np.array_equal(a[hi], d[hi], equal_nan=True)
        assert np.array_equal(b[hi], c[hi], equal_nan=True)
    for x in (plain, s128, s256):
        assert rel(x[0], tuned[0]).max() <= RTOL and rel(x[1], tuned[1]).max() <= RTOL
    assert np.all(tuned[3] == 0)
    # the default tuned form differs only in the k = 128 subsets, within 1e-5
    dflt = _np(solver.solve(G, idx, k=128, offsets=off))
    kk = ks.cpu().numpy()
    assert np.array_equal(dflt[0][kk < 128], tuned[0][kk < 128])
    assert rel(dflt[0], tuned[0]).max() <= RTOL and rel(dflt[1], tuned[1]).max() <= RTOL
    # the values: the oracle on a slice of the batch
    o = off.cpu().numpy()
    sl = slice(0, 400)
    olo, ohi, _, ost = oracle.solve_batch(G.cpu().numpy(), idx.cpu().numpy()[:o[400]],
                                          offsets=o[:401], nthreads=8)
    assert np.all(ost == 0)
    assert rel(tuned[0][sl], olo).max() <= RTOL and rel(tuned[1][sl], ohi).max() <= RTOL


def test_sizes_beyond_kmax_flagged(mixed):
    """k larger than the declared maximum is BADK (its T would not fit the
    stride), sizes up to it are solved, in every form."""
    G, ks, off, idx = mixed
    kmax = 40
    kk = ks.cpu().numpy()
    for form in ("tuned", "plain"):
        lo, hi, de, st = _np(solver.solve(G, idx, k=kmax, offsets=off, form=form))
        big = kk > kmax
        assert np.all(st[big] == 16) and np.all(np.isnan(lo[big]))
        assert np.all(st[~big] == 0) and np.all(np.isfinite(lo[~big]))
    ref = _np(solver.solve(G, idx, k=128, offsets=off))
    assert np.array_equal(lo[~big], ref[0][~big])


def test_plan_graph_replay(mixed):
    """A tuned plan records a graph on its first execution and replays it;
    every replay, with new index sets written into the bound tensor and eager
    device work in between, equals an eager solve of the same inputs."""
    G, ks, off, idx = mixed
    n = G.shape[0]
    idx2 = idx.clone()
    plan = solver.Plan(G, idx2, k=128, offsets=off, form="tuned")
    try:
        for rep in range(4):
            idx2.copy_(solver.gen_subsets_csr(0xB2 + rep, off, n))
            junk = torch.randn(1 << 20, device="cuda").sum()  # eager work between replays
            got = _np(plan.execute())
            torch.cuda.synchronize()
            assert plan.is_graph
            ref = _np(solver.solve(G, idx2, k=128, offsets=off))
            for a, b in zip(got, ref):
                assert np.array_equal(a, b, equal_nan=True), rep
            assert np.isfinite(junk.item())
    finally:
        plan.close()


def test_plan_uniform_and_plain(mixed):
    G, ks, off, idx = mixed
    n = G.shape[0]
    u = solver.gen_subsets(0xB7, 5000, 24, n)
    plan = solver.Plan(G, u, k=24, form="tuned")
    ref = _np(solver.solve(G, u, k=24))
    for _ in range(3):
        got = _np(plan.execute())
        for a, b in zip(got, ref):
            assert np.array_equal(a, b, equal_nan=True)
    assert plan.is_graph
    plan.close()
    # the plain form of a mixed batch waits for bin counts: it runs eagerly
    pp = solver.Plan(G, idx, k=128, offsets=off, form="plain")
    got = _np(pp.execute())
    assert not pp.is_graph
    ref = _np(solver.solve(G, idx, k=128, offsets=off, form="plain"))
    for a, b in zip(got, ref):
        assert np.array_equal(a, b, equal_nan=True)
    pp.close()


def test_autotune(mixed):
    G, ks, off, idx = mixed
    before = _np(solver.solve(G, idx, k=128, offsets=off))
    try:
        res = solver.autotune(G, [8, 24, 40, 72], sample=4096)
        for k, (lanes, pad, ms) in res.items():
            assert lanes >= k and pad in (0, 1, 4) and ms > 0
        # unlisted sizes take the winner of the next listed size above them
        # when it is of their own kernel family (lane groups up to 64 lanes,
        # then the 256-lane half-row kernel), else their default
        fam = lambda lanes: 0 if lanes < 128 else lanes
        for k in range(1, 73):
            nxt = min(x for x in (8, 24, 40, 72) if x >= k)
            want = solver.tuning_get(nxt)
            dflt = (256 if k > 64 else max(8, 1 << (k - 1).bit_length()), 4)
            if fam(want[0]) != fam(dflt[0]):
                want = dflt
            assert solver.tuning_get(k) == want, k
        after = _np(solver.solve(G, idx, k=128, offsets=off))
        for a, b in zip(before, after):
            assert np.array_equal(a, b, equal_nan=True)
    finally:
        solver.tuning_reset()
    assert solver.tuning_get(24) == (32, 4)


def test_autotune_edges(mixed):
    """Tuned bin edges: at each listed size every variant that takes it
    competes across kernel families; unlisted sizes take the next listed
    size's winner. With the table fixed, the binned solve still matches the
    oracle, and each subset's result is the same alone and in the mix."""
    G, ks, off, idx = mixed
    klist = [8, 24, 40, 64, 72, 128]
    try:
        win, table = solver.autotune_edges(G, klist, sample=4096)
        for k in klist:
            lanes, pad, ms = win[k]
            assert ms > 0 and (lanes, pad) in table[k] and table[k][(lanes, pad)] == ms
            assert min(table[k].values()) == ms
            # every variant that takes k was timed, none that cannot
            for (l, p) in table[k]:
                assert (l if l < 256 else l // 2) >= k
            cands = {(l, p) for l in (8, 16, 32, 64) if l >= k for p in (0, 1, 4)}
            cands |= {(128, p) for p in (1, 4) if k <= 128} | {(256, 4)} | {(512, 4)}
            assert set(table[k]) == {c for c in cands if (c[0] if c[0] < 256 else c[0] // 2) >= k}
        for k in range(1, 129):
            nxt = min(x for x in klist if x >= k)
            assert solver.tuning_get(k) == solver.tuning_get(nxt), k
        out = _np(solver.solve(G, idx, k=128, offsets=off))
        assert not out[3].any()
        o = off.cpu().numpy()
        ih = idx.cpu().numpy()
        olo, ohi, _, _ = oracle.solve_batch(G.cpu().numpy(), ih, offsets=o, nthreads=8)
        assert rel(out[0], olo).max() <= RTOL and rel(out[1], ohi).max() <= RTOL
        # alone: each of 40 subsets solved by itself with the same table
        for s in range(0, 3000, 75):
            k = int(o[s + 1] - o[s])
            one = _np(solver.solve(G, idx[o[s]:o[s + 1]].contiguous(), k=k))
            assert one[0][0] == out[0][s] and one[1][0] == out[1][s], s
    finally:
        solver.tuning_reset()
    assert solver.tuning_get(24) == (32, 4) and solver.tuning_get(100) == (256, 4)


def test_padded_leading_dimension(gpu):
    """G stored with a leading dimension above n (rows padded with NaN that
    no kernel may read) gives the same outputs bit for bit as the contiguous
    G, through the C ABI, for sizes on every kernel family (lane groups, the
    256- and 512-lane kernels) in the tuned and plain forms."""
    import ctypes

    from paralanczos_amd import _lib

    n, pad = 1000, 13
    G = solver.gram(solver.gen_phi("gauss", 0xD1, 320, n))
    Gp = torch.full((n, n + pad), float("nan"), device="cuda")
    Gp[:, :n] = G
    ks = solver.gen_ks(0xD2, 400, [8, 16, 32, 64, 100, 128, 200, 256])
    off = solver.offsets_from_ks(ks)
    idx = solver.gen_subsets_csr(0xD3, off, n)
    nsub = off.numel() - 1
    L = _lib.load()
    P = lambda x: ctypes.c_void_p(x.data_ptr())
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for form in ("tuned", "plain"):
        ref = _np(solver.solve(G, idx, k=256, offsets=off, form=form))
        out = [torch.empty(nsub, device="cuda") for _ in range(3)] + \
              [torch.empty(nsub, dtype=torch.int32, device="cuda")]
        rc = L.pl_solve(P(Gp), n, n + pad, P(idx), P(off), 256, nsub, *[P(x) for x in out],
                        solver.FORMS[form], stream)
        assert rc == 0
        got = _np(out)
        assert not got[3].any()
        for a, b in zip(got, ref):
            assert np.array_equal(a, b), form
