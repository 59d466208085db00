"""CPU tests pinning the oracle (the checker of the GPU path).

Each oracle function is checked against something that does not restate its
own formula: published known-answer vectors, committed LAPACK fixtures from
the reference's example data, closed forms, library routines, invariants and
brute force on tiny inputs.
"""
import hashlib
import json
import math
import os

import numpy as np
import pytest
import scipy.fft
import scipy.linalg

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "golden.json")))


# ---- Philox4x32-10: Random123 known-answer vectors (Salmon et al., SC'11) ----
@pytest.mark.parametrize("ctr,key,expect", [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
])
def test_philox_kat(ctr, key, expect):
    assert list(oracle.philox4x32_10(ctr, key)) == expect


def test_stream_word_layout():
    seed, tag, sid = 0x1234567890ABCDEF, 3, 0x0000000500000007
    for w in range(12):
        blk = oracle.philox4x32_10([w >> 2, sid & 0xFFFFFFFF, sid >> 32, tag],
                                   [seed & 0xFFFFFFFF, seed >> 32])
        assert oracle.stream_word(seed, tag, sid, w) == blk[w & 3]


def test_floyd_invariants():
    for n, k in [(10, 10), (256, 8), (4096, 32), (32768, 256), (5, 1)]:
        for sid in range(20):
            s = oracle.floyd(42, sid, n, k)
            assert len(s) == k and len(set(s.tolist())) == k
            assert s.min() >= 0 and s.max() < n


def test_floyd_uniform_over_subsets():
    # n=6, k=3: all 20 subsets equally likely (chi-square, 19 dof, p ~ 1e-6 bound)
    idx = oracle.gen_subsets_uniform(7, 40000, 3, 6).reshape(-1, 3)
    keys = [tuple(sorted(r)) for r in idx.tolist()]
    counts = np.array([keys.count(c) for c in sorted(set(keys))])
    assert len(counts) == 20
    exp = 40000 / 20
    chi2 = ((counts - exp) ** 2 / exp).sum()
    assert chi2 < 55.0


def test_ks_uniform_and_digests():
    klist = list(range(8, 129, 8))
    ks = oracle.gen_ks(0xC5, 32000, klist)
    assert set(ks.tolist()) <= set(klist)
    counts = np.array([(ks == k).sum() for k in klist])
    exp = 32000 / 16
    assert ((counts - exp) ** 2 / exp).sum() < 45.0
    d = GOLD["index_digests"]["ks_0xc5"]
    ks64 = oracle.gen_ks(0xC5, 64, klist)
    assert hashlib.sha256(ks64.astype("<i4").tobytes()).hexdigest() == d["sha256"]


@pytest.mark.parametrize("key", ["0xc2_k8_n256", "0xc3_k32_n4096"])
def test_index_stream_digests(key):
    seed_s, k_s, n_s = key.split("_")
    seed, k, n = int(seed_s, 16), int(k_s[1:]), int(n_s[1:])
    idx = oracle.gen_subsets_uniform(seed, 64, k, n)
    d = GOLD["index_digests"][key]
    assert hashlib.sha256(idx.astype("<i4").tobytes()).hexdigest() == d["sha256"]
    assert idx[:k].tolist() == d["first"]


def test_csr_matches_uniform():
    off = np.arange(0, 33 * 40, 33, dtype=np.int64)
    a = oracle.gen_subsets_csr(9, off, 1000)
    b = oracle.gen_subsets_uniform(9, 39, 33, 1000)
    assert np.array_equal(a, b)


def test_phi_gauss_moments():
    m, n = 256, 2048
    P = oracle.gen_phi("gauss", 0xC1, m, n)  # row j = column j of Phi
    np.testing.assert_allclose(np.linalg.norm(P.astype(np.float64), axis=1), 1.0, atol=1e-6)
    z = P.reshape(-1) * math.sqrt(m)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1.0) < 0.01
    kurt = ((z - z.mean()) ** 4).mean() / z.var() ** 2
    assert abs(kurt - 3.0) < 0.05


def test_phi_bernoulli():
    m, n = 2048, 64
    P = oracle.gen_phi("bernoulli", 0xC4, m, n)
    v = np.float32(1 / math.sqrt(m))
    assert set(np.unique(P).tolist()) == {float(-v), float(v)}
    assert abs((P > 0).mean() - 0.5) < 0.01


def test_phi_dct_is_orthonormal_dct():
    n = 64
    P = oracle.gen_phi("dct", 5, n, n)  # all rows: Phi = permuted orthonormal DCT-II
    rows = oracle.dct_rows(5, n, n)
    C = scipy.fft.dct(np.eye(n), type=2, norm="ortho", axis=0)  # C[r, j]
    np.testing.assert_allclose(P.T, C[rows, :], atol=2e-7)
    np.testing.assert_allclose(P.astype(np.float64) @ P.T.astype(np.float64), np.eye(n), atol=2e-6)
    Pm = oracle.gen_phi("dct", 6, 16, n)  # partial: unit columns on average after sqrt(n/m)
    assert abs(np.mean(np.linalg.norm(Pm, axis=1) ** 2) - 1.0) < 1e-6


def test_gather_indexing():
    rng = np.random.default_rng(1)
    G = rng.normal(size=(50, 50)).astype(np.float32)  # deliberately non-symmetric
    idx = np.array([7, 3, 41, 3, 0], np.int32)
    A, st = oracle.gather(G, idx)
    assert st == 0
    for r in range(5):
        for c in range(5):
            i, j = idx[r], idx[c]
            assert A[r, c] == G[min(i, j), max(i, j)]
    _, st = oracle.gather(G, np.array([1, 50], np.int32))
    assert st == oracle.ST_BADINDEX


def all_eigs(al, be):
    return np.array([oracle.sturm_eig(al, be, t)[0] for t in range(len(al))])


@pytest.mark.parametrize("name", ["full_0_7", "s1357", "s0123", "s4567"])
def test_lanczos_reference_examples(name):
    g = GOLD["subsets"][name]
    G = np.array(GOLD["example_matrix"])
    A = G[np.ix_(g["index"], g["index"])]
    al, be, st = oracle.lanczos(A)
    assert st == 0
    w = all_eigs(al, be)
    np.testing.assert_allclose(w, g["eig"], rtol=1e-12)
    assert abs(np.prod(w) - g["det"]) / g["det"] < 1e-10


def test_lanczos_first_step_closed_form():
    d = np.array([1.0, 2.0, 4.0, 8.0, 3.0])
    al, be, _ = oracle.lanczos(np.diag(d))
    assert abs(al[0] - d.mean()) < 1e-14
    assert abs(be[1] - d.std()) < 1e-14
    np.testing.assert_allclose(np.sort(all_eigs(al, be)), np.sort(d), rtol=1e-12)


def test_lanczos_2x2_closed_form():
    a, b, d = 3.0, 0.7, -1.5
    al, be, _ = oracle.lanczos(np.array([[a, b], [b, d]]))
    assert abs(al[0] - ((a + d) / 2 + b)) < 1e-14
    assert abs(al[1] - ((a + d) / 2 - b)) < 1e-14
    assert abs(be[1] - abs(a - d) / 2) < 1e-14


def test_lanczos_random_spd_vs_dsyevd_and_breakdown():
    rng = np.random.default_rng(3)
    for k in (3, 17, 64, 128, 256):
        X = rng.normal(size=(2 * k, k))
        A = X.T @ X / (2 * k)
        al, be, st = oracle.lanczos(A)
        assert st == 0
        np.testing.assert_allclose(all_eigs(al, be), np.linalg.eigvalsh(A), rtol=1e-9, atol=1e-12)
    # identity: ones/sqrt(k) is an eigenvector -> restarts, all eigenvalues 1
    al, be, st = oracle.lanczos(np.eye(16))
    assert st & oracle.ST_BREAKDOWN
    np.testing.assert_allclose(all_eigs(al, be), 1.0, atol=1e-14)
    # duplicated column: singular, lambda_min = 0 recovered via the restart
    X = rng.normal(size=(40, 12))
    X[:, 5] = X[:, 2]
    A = X.T @ X
    al, be, st = oracle.lanczos(A)
    w = np.linalg.eigvalsh(A)
    np.testing.assert_allclose(all_eigs(al, be), w, atol=1e-10 * w[-1])


def test_sturm_count_brute_force():
    rng = np.random.default_rng(5)
    for k in (1, 2, 5, 9):
        a = rng.normal(size=k)
        b = np.concatenate([[0.0], rng.normal(size=k - 1)])
        T = np.diag(a) + np.diag(b[1:], 1) + np.diag(b[1:], -1)
        w = np.linalg.eigvalsh(T)
        for sigma in rng.normal(size=20) * 2:
            assert oracle.sturm_count(a, b, sigma) == int((w < sigma).sum())


@pytest.mark.parametrize("k", [1, 2, 8, 33, 256])
def test_toeplitz_closed_form(k):
    a0, b0 = 0.9, 0.35
    al = np.full(k, a0)
    be = np.concatenate([[0.0], np.full(k - 1, b0)])
    exact = np.sort(a0 + 2 * b0 * np.cos(np.arange(1, k + 1) * np.pi / (k + 1)))
    np.testing.assert_allclose(all_eigs(al, be), exact, rtol=1e-12, atol=1e-14)
    d, ierr = oracle.tqlrat(al, be ** 2)
    assert ierr == 0
    np.testing.assert_allclose(d, exact, rtol=1e-12, atol=1e-14)


def test_tridiagonal_fixtures_sturm_and_tqlrat():
    for t in GOLD["tridiagonal"]:
        al, be, w = np.array(t["alpha"]), np.array(t["beta"]), np.array(t["eig"])
        np.testing.assert_allclose(all_eigs(al, be), w, rtol=1e-11, atol=1e-13)
        d, ierr = oracle.tqlrat(al, be ** 2)
        assert ierr == 0
        np.testing.assert_allclose(d, w, rtol=1e-11, atol=1e-13)


def test_faithful_variant_on_examples():
    G = np.array(GOLD["example_matrix"])
    for name in ("full_0_7", "s1357"):
        g = GOLD["subsets"][name]
        k = len(g["index"])
        eig, det, ierr = oracle.solve_batch_faithful(G, np.array(g["index"] * 2, np.int32), k)
        assert np.all(ierr == 0)
        np.testing.assert_allclose(eig.reshape(2, k), [g["eig"]] * 2, rtol=1e-8)
        np.testing.assert_allclose(det, g["det"], rtol=1e-8)


def test_batch_drivers():
    rng = np.random.default_rng(11)
    X = rng.normal(size=(64, 200)).astype(np.float32)
    X /= np.linalg.norm(X, axis=0)
    G = (X.T @ X).astype(np.float32)
    G = ((G + G.T) * np.float32(0.5)).astype(np.float32)
    idx = oracle.gen_subsets_uniform(4, 50, 12, 200)
    lo, hi, de, st = oracle.solve_batch(G, idx, k=12, nthreads=4)
    eig, det, st2 = oracle.solve_batch_alleig(G, idx, 12, nthreads=2)
    assert np.all(st == 0) and np.all(st2 == 0)
    for s in range(50):
        A = G[np.ix_(idx[s * 12:(s + 1) * 12], idx[s * 12:(s + 1) * 12])].astype(np.float64)
        w = np.linalg.eigvalsh(A)
        assert abs(lo[s] - w[0]) < 1e-12 and abs(hi[s] - w[-1]) < 1e-12
        assert abs(de[s] - max(w[-1] - 1, 1 - w[0])) < 1e-12
        np.testing.assert_allclose(eig[s * 12:(s + 1) * 12], w, atol=1e-12)
    ks = np.array([3, 12, 1, 7], np.int64)
    off = np.concatenate([[0], np.cumsum(ks)])
    idx2 = oracle.gen_subsets_csr(8, off, 200)
    lo2, hi2, _, st3 = oracle.solve_batch(G, idx2, offsets=off)
    for s in range(4):
        S = idx2[off[s]:off[s + 1]]
        w = np.linalg.eigvalsh(G[np.ix_(S, S)].astype(np.float64))
        assert abs(lo2[s] - w[0]) < 1e-12 and abs(hi2[s] - w[-1]) < 1e-12


def test_planted_mistakes_are_caught():
    """The pins above must reject plausible oracle mistakes: a dropped
    three-term term, a wrong Sturm index, a transposed gather."""
    G = np.array(GOLD["example_matrix"])
    A = G[np.ix_([1, 3, 5, 7], [1, 3, 5, 7])]
    al, be, _ = oracle.lanczos(A)
    w = np.array(GOLD["subsets"]["s1357"]["eig"])
    bad_beta = be.copy()
    bad_beta[2] = 0.0  # dropped coupling term
    assert np.max(np.abs(all_eigs(al, bad_beta) - w)) > 1e-3
    wrong_index = np.array([oracle.sturm_eig(al, be, min(t + 1, 3))[0] for t in range(4)])
    assert np.max(np.abs(wrong_index - w)) > 1e-3
    rng = np.random.default_rng(2)
    Gn = rng.normal(size=(20, 20)).astype(np.float32)
    A1, _ = oracle.gather(Gn, np.array([2, 9, 4], np.int32))
    assert not np.array_equal(A1, Gn[np.ix_([2, 9, 4], [2, 9, 4])].T) or np.allclose(Gn, Gn.T)


# ---- power-iteration estimate (UPPER_KER, ParaLanczos.cu:1484-1626) ----
@pytest.mark.parametrize("iters", [1, 2, 3, 10, 12])
def test_upper_diagonal_closed_form(iters):
    """v_0 = 1/sqrt(k): after L steps v is D^L 1 normalised, so the last
    ||D v|| is ||D^L 1|| / ||D^(L-1) 1||."""
    d = np.array([3.0, 1.0, 0.5, 2.0, 2.9, 0.1])
    expect = math.sqrt(np.sum(d ** (2 * iters)) / np.sum(d ** (2 * (iters - 1))))
    assert abs(oracle.upper(np.diag(d), iters) - expect) <= 1e-13 * expect


def test_upper_rank_one_and_identity():
    """A = u u^T: the first step gives ||u|| |sum u| / sqrt(k), every later
    one ||u||^2 exactly; c I gives c."""
    u = np.array([0.5, -1.0, 2.0, 0.25, 1.5])
    A = np.outer(u, u)
    k = len(u)
    assert abs(oracle.upper(A, 1) - np.linalg.norm(u) * abs(u.sum()) / math.sqrt(k)) < 1e-13
    for L in (2, 5, 10):
        assert abs(oracle.upper(A, L) - u @ u) < 1e-12
    assert abs(oracle.upper(2.5 * np.eye(7), 10) - 2.5) < 1e-14


def test_upper_bounded_by_and_converges_to_lambda_max():
    rng = np.random.default_rng(5)
    for k in (8, 40, 128):
        X = rng.normal(size=(2 * k, k))
        A = X.T @ X / (2 * k)
        w = np.linalg.eigvalsh(A)
        assert oracle.upper(A, 10) <= w[-1] * (1 + 1e-14)
        assert abs(oracle.upper(A, 20000) - w[-1]) <= 1e-9 * w[-1]


def test_upper_batch_matches_gathered_subsets():
    rng = np.random.default_rng(8)
    X = rng.normal(size=(48, 120)).astype(np.float32)
    G = (X.T @ X).astype(np.float32)
    G = ((G + G.T) * np.float32(0.5)).astype(np.float32)
    ks = np.array([5, 17, 1, 33], np.int64)
    off = np.concatenate([[0], np.cumsum(ks)])
    idx = oracle.gen_subsets_csr(3, off, 120)
    up, st = oracle.upper_batch(G, idx, offsets=off, iters=10, nthreads=2)
    assert np.all(st == 0)
    for s in range(4):
        A, _ = oracle.gather(G, idx[off[s]:off[s + 1]])
        assert up[s] == oracle.upper(A, 10)
    bad = idx.copy()
    bad[0] = 500
    up2, st2 = oracle.upper_batch(G, bad, offsets=off)
    assert st2[0] == oracle.ST_BADINDEX and np.isnan(up2[0]) and st2[1] == 0


def test_upper_planted_mistakes_are_caught():
    """The closed forms reject an iteration count off by one and a returned
    squared norm (planted here in a numpy restatement, not in the oracle)."""
    d = np.array([3.0, 1.0, 0.5, 2.0])

    def planted(A, iters, extra=0, square=False):
        v = np.ones(len(A)) / math.sqrt(len(A))
        b = 0.0
        for _ in range(iters + extra):
            w = A @ v
            b = w @ w if square else math.sqrt(w @ w)
            v = w / math.sqrt(w @ w)
        return b

    expect = math.sqrt(np.sum(d ** 20) / np.sum(d ** 18))
    assert abs(planted(np.diag(d), 10) - expect) < 1e-12
    assert abs(planted(np.diag(d), 10, extra=1) - expect) > 1e-6
    assert abs(planted(np.diag(d), 10, extra=-1) - expect) > 1e-6
    assert abs(planted(np.diag(d), 10, square=True) - expect) > 1e-3


def test_faithful_extremes_driver():
    """The FP32-G, CSR form of the reference-faithful path equals the FP64-G
    driver (pinned on the example matrix above) on the same data, and on
    well-conditioned Gram subsets its extremes agree with dsyevd."""
    rng = np.random.default_rng(21)
    X = rng.normal(size=(96, 150)).astype(np.float32)
    X /= np.linalg.norm(X, axis=0)
    G = (X.T @ X).astype(np.float32)
    G = ((G + G.T) * np.float32(0.5)).astype(np.float32)
    idx = oracle.gen_subsets_uniform(6, 30, 10, 150)
    lo, hi, ierr = oracle.faithful_extremes(G, idx, k=10, nthreads=2)
    eig, _, ierr2 = oracle.solve_batch_faithful(G.astype(np.float64), idx, 10)
    assert np.all(ierr == 0) and np.all(ierr2 == 0)
    e = eig.reshape(30, 10)
    assert np.array_equal(lo, e[:, 0]) and np.array_equal(hi, e[:, -1])
    ks = np.array([4, 16, 1, 9], np.int64)
    off = np.concatenate([[0], np.cumsum(ks)])
    idx2 = oracle.gen_subsets_csr(9, off, 150)
    lo2, hi2, ierr3 = oracle.faithful_extremes(G, idx2, offsets=off)
    assert np.all(ierr3 == 0)
    for s in range(4):
        S = idx2[off[s]:off[s + 1]]
        w = np.linalg.eigvalsh(G[np.ix_(S, S)].astype(np.float64))
        assert abs(lo2[s] - w[0]) < 1e-8 and abs(hi2[s] - w[-1]) < 1e-8
    bad = idx2.copy()
    bad[0] = 1000
    lo3, _, ierr4 = oracle.faithful_extremes(G, bad, offsets=off)
    assert ierr4[0] == -1 and np.isnan(lo3[0]) and ierr4[1] == 0
