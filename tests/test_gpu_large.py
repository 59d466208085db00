"""Subsets of 128 < k <= 256 (the 512-lane kernel with its basis split between
LDS and registers) and the FP64 Gram-Schmidt coefficient forms, against the
oracle and FP64 LAPACK (relative eigenvalue error <= 1e-5)."""
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


def _np(ts):
    return [t.cpu().numpy() for t in ts]


def test_gather_dump_halfrow_bit_exact(gpu):
    """The half-row kernels' gathers (256 lanes: k <= 128, 512 lanes: k <= 256)."""
    n = 1024
    G = solver.gram(solver.gen_phi("gauss", 3, 256, n))
    Gh = G.cpu().numpy()
    for lanes, k in ((256, 65), (256, 100), (256, 128), (512, 129), (512, 200), (512, 256)):
        idx = solver.gen_subsets(40 + k, 9, k, n)
        out = solver.gather(G, idx, k, "dump", lanes, 4).cpu().numpy().reshape(-1, k, k)
        ih = idx.cpu().numpy()
        for s in range(9):
            A, st = oracle.gather(Gh, ih[s * k:(s + 1) * k])
            assert st == 0 and np.array_equal(out[s], A.astype(np.float32)), (lanes, k, s)
        cs = solver.gather(G, idx, k, "checksum", lanes, 4).cpu().numpy()
        np.testing.assert_allclose(cs, out.reshape(9, -1).astype(np.float64).sum(1), rtol=1e-4)


@pytest.mark.parametrize("k", [100, 129, 160, 200, 255, 256])
def test_k256_vs_oracle(gpu, k):
    n = 1024
    G = solver.gram(solver.gen_phi("gauss", 900 + k, 4 * k, n))
    idx = solver.gen_subsets(77 + k, 24, k, n)
    Gh = G.cpu().numpy()
    ih = idx.cpu().numpy()
    olo, ohi, _, ost = oracle.solve_batch(Gh, ih, k=k, nthreads=8)
    assert np.all(ost == 0)
    runs = [solver.solve(G, idx, k=k, lanes=512, pad=4)]
    if k > 128:
        runs += [solver.solve(G, idx, k=k, reorth=r) for r in ("f32", "f64", None)]
    for r in runs:
        lo, hi, de, st = _np(r)
        assert np.all(st == 0)
        assert rel(lo, olo).max() <= RTOL and rel(hi, ohi).max() <= RTOL
        np.testing.assert_allclose(de, np.maximum(hi - 1.0, 1.0 - lo), rtol=1e-6, atol=1e-7)


@pytest.mark.parametrize("k", [148, 149, 150, 151, 152, 165, 166, 167, 182, 230, 245, 246, 247, 254])
def test_512_lane_register_blocks_vs_oracle(gpu, k):
    """128 < k <= 256 on the 512-lane kernel: basis columns 0..147 in LDS,
    column 148 + 2 m + h in register slot m / 8 of lane m mod 8 of each
    8-row block of part h. A column takes part in the Gram-Schmidt passes
    of steps >= its index and the last step runs none, so k = 150 and 151
    are the first sizes that use a register column of part 0 and of part 1,
    k = 166 / 167 the first that use slot 1, k = 246 / 247 slot 6, and
    k = 148, 149 have none; every Gram-Schmidt coefficient form (forced FP32,
    forced FP64, the default) against the oracle."""
    assert solver.tuning_get(k) == (512, 4)
    n = 1024
    G = solver.gram(solver.gen_phi("gauss", 500 + k, 4 * k, n))
    idx = solver.gen_subsets(61 + k, 24, k, n)
    olo, ohi, _, ost = oracle.solve_batch(G.cpu().numpy(), idx.cpu().numpy(), k=k, nthreads=8)
    assert np.all(ost == 0)
    for r in ("f32", "f64", None):
        lo, hi, de, st = _np(solver.solve(G, idx, k=k, reorth=r))
        assert np.all(st == 0), r
        assert rel(lo, olo).max() <= RTOL and rel(hi, ohi).max() <= RTOL, r
        np.testing.assert_allclose(de, np.maximum(hi - 1.0, 1.0 - lo), rtol=1e-6, atol=1e-7)


@pytest.mark.parametrize("k", [66, 92, 93, 94, 111, 127, 128])
def test_256_lane_register_columns_vs_oracle(gpu, k):
    """64 < k <= 128 on the 256-lane kernel: basis columns 0..91 in LDS, the
    rest in registers (two per lane and row). k = 93, 94 form the first
    register column of each row part; every Gram-Schmidt coefficient form
    (forced FP32, forced FP64, the default) against the oracle."""
    assert solver.tuning_get(k) == (256, 4)
    n = 768
    G = solver.gram(solver.gen_phi("gauss", 300 + k, 4 * k, n))
    idx = solver.gen_subsets(31 + k, 40, k, n)
    olo, ohi, _, ost = oracle.solve_batch(G.cpu().numpy(), idx.cpu().numpy(), k=k, nthreads=8)
    assert np.all(ost == 0)
    for r in ("f32", "f64", None):
        lo, hi, de, st = _np(solver.solve(G, idx, k=k, reorth=r))
        assert np.all(st == 0), r
        assert rel(lo, olo).max() <= RTOL and rel(hi, ohi).max() <= RTOL, r
        np.testing.assert_allclose(de, np.maximum(hi - 1.0, 1.0 - lo), rtol=1e-6, atol=1e-7)


def test_k256_rank_deficient_alleig(gpu):
    """m = 160 < k: G_S has rank 160 and the Krylov space of the start vector
    is exhausted early. Whether beta then falls under the restart threshold
    (flagged BREAKDOWN) or stays at FP32 rounding level (the reorthogonalised
    remainder continues the recurrence), every eigenvalue agrees with LAPACK,
    the zero ones to 1e-5 of lambda_max."""
    n = 600
    G = solver.gram(solver.gen_phi("gauss", 5, 160, n))
    Gh = G.cpu().numpy()
    for k in (200, 256):
        idx = solver.gen_subsets(6 + k, 8, k, n)
        eig, det, err = _np(solver.solve_alleig(G, idx, k))
        ih = idx.cpu().numpy()
        assert np.all((err & ~1) == 0)
        for s in range(8):
            A, _ = oracle.gather(Gh, ih[s * k:(s + 1) * k])
            w = np.linalg.eigvalsh(A)
            e = eig[s * k:(s + 1) * k]
            assert np.all(np.diff(e) >= 0)
            assert np.max(np.abs(e - w)) <= 1e-5 * w[-1], (k, s)


def test_alleig_k256_full_rank(gpu):
    n = 2048
    G = solver.gram(solver.gen_phi("gauss", 8, 1024, n))
    Gh = G.cpu().numpy()
    k = 224
    idx = solver.gen_subsets(9, 6, k, n)
    eig, det, err = _np(solver.solve_alleig(G, idx, k))
    ih = idx.cpu().numpy()
    assert np.all(err == 0)
    for s in range(6):
        A, _ = oracle.gather(Gh, ih[s * k:(s + 1) * k])
        w = np.linalg.eigvalsh(A)
        assert np.max(np.abs(eig[s * k:(s + 1) * k] - w)) <= 1e-5 * w[-1]
        assert rel(det[s], np.prod(w)) <= 1e-3


def test_mixed_up_to_256(gpu):
    """k from {16, 32, 64, 128, 256} in one CSR batch: the tuned form (FP64
    coefficients from 128) within 1e-5 of the oracle; with FP32 coefficients
    forced it is bit-identical to the plain one-launch-per-k form wherever
    both use the same kernel (k <= 64: the lane groups; k = 256: the 512-lane
    kernel) and within 1e-5 at k = 128 (256- vs 128-lane kernel)."""
    n = 4096
    G = solver.gram(solver.gen_phi("gauss", 0xC5, 1024, n))
    ks = solver.gen_ks(0xC5 + 1, 600, [16, 32, 64, 128, 256])
    off = solver.offsets_from_ks(ks)
    idx = solver.gen_subsets_csr(0xC5 + 2, off, n)
    tuned = _np(solver.solve(G, idx, k=256, offsets=off))
    plain = _np(solver.solve(G, idx, k=256, offsets=off, form="plain"))
    t32 = _np(solver.solve(G, idx, k=256, offsets=off, reorth="f32"))
    same = ks.cpu().numpy() != 128
    for a, b in zip(t32, plain):
        assert np.array_equal(a[same], b[same])
    assert rel(t32[0], plain[0]).max() <= RTOL and rel(t32[1], plain[1]).max() <= RTOL
    o = off.cpu().numpy()
    olo, ohi, _, ost = oracle.solve_batch(G.cpu().numpy(), idx.cpu().numpy(), offsets=o,
                                          nthreads=8)
    assert np.all(ost == 0)
    for r in (tuned, plain):
        assert np.all(r[3] == 0)
        assert rel(r[0], olo).max() <= RTOL and rel(r[1], ohi).max() <= RTOL


def test_c4_bernoulli_sample(gpu):
    """The precision workload's shape: Bernoulli +-1/sqrt(2048) Phi of
    2048 x 16384, G = Phi^T Phi (1 GiB), k = 128 and 256; both coefficient
    precisions within 1e-5 of dsyevd on the gathered submatrices."""
    m, n = 2048, 16384
    G = solver.gram(solver.gen_phi("bernoulli", 0xC4, m, n))
    for k in (128, 256):
        idx = solver.gen_subsets(0xC4 + 2, 48, k, n)
        A = solver.gather(G, idx, k, "dump", 512, 4).cpu().numpy().reshape(-1, k, k)
        w = np.linalg.eigvalsh(A.astype(np.float64))
        for r in ("f32", "f64"):
            lo, hi, de, st = _np(solver.solve(G, idx, k=k, reorth=r))
            assert np.all(st == 0), (k, r)
            assert rel(lo, w[:, 0]).max() <= RTOL and rel(hi, w[:, -1]).max() <= RTOL, (k, r)
    del G
    torch.cuda.empty_cache()


def test_rip_stats_per_size(gpu):
    """Per-size RIP statistics of a solved mixed batch equal numpy's over the
    same outputs; a subset with an out-of-range index is skipped (row 0)."""
    n = 2048
    G = solver.gram(solver.gen_phi("gauss", 0xA1, 512, n))
    ks = solver.gen_ks(0xA2, 700, [16, 32, 64, 128, 256])
    off = solver.offsets_from_ks(ks)
    idx = solver.gen_subsets_csr(0xA3, off, n)
    idx[int(off[5].item())] = n + 3  # subset 5 gets BADINDEX
    lo, hi, de, st = solver.solve(G, idx, k=256, offsets=off)
    out = solver.rip_stats(lo, hi, de, st, offsets=off, kmax=256).cpu().numpy()
    lo, hi, de, st, kk = [t.cpu().numpy() for t in (lo, hi, de, st, ks)]
    assert st[5] & 8 and out[0, 0] == 1
    ok = (st & (2 | 8 | 16)) == 0
    for k in (16, 32, 64, 128, 256):
        m = ok & (kk == k)
        r = out[k]
        assert r[0] == m.sum()
        np.testing.assert_allclose(r[[1, 2, 3, 4]], [lo[m].astype(np.float64).mean(),
                                                    lo[m].astype(np.float64).std(),
                                                    hi[m].astype(np.float64).mean(),
                                                    hi[m].astype(np.float64).std()], rtol=1e-9)
        assert r[5] == de[m].max() and r[6] == lo[m].min() and r[7] == hi[m].max()
    assert np.isnan(out[17, 1]) and out[17, 0] == 0
    # a second pass over the same outputs: counts and extremes identical, sums
    # to the last bits of FP64 atomics
    u = solver.rip_stats(*[torch.from_numpy(a).cuda() for a in (lo, hi, de, st)], offsets=off,
                         kmax=256).cpu().numpy()
    assert np.array_equal(u[:, [0, 5, 6, 7]], out[:, [0, 5, 6, 7]], equal_nan=True)
    np.testing.assert_allclose(u[1:, :], out[1:, :], rtol=1e-12, equal_nan=True)


def test_multisection_agrees_with_bisection(gpu):
    """Launches of up to 16384 subsets find the extremes by 9-way
    multisection, larger ones by bisection. Both stop at the same width
    (2^-27 relative) around the eigenvalue of the same T, so a subset's
    extremes agree to that width and the FP32 rounding of the outputs."""
    n = 4096
    G = solver.gram(solver.gen_phi("gauss", 17, 512, n))
    for k in (8, 40, 129, 256):
        idx = solver.gen_subsets(70 + k, 16385, k, n)
        big = _np(solver.solve(G, idx, k=k))  # 16385 subsets: bisection
        small = _np(solver.solve(G, idx[:300 * k].contiguous(), k=k))  # multisection
        assert not small[3].any() and not big[3][:300].any()
        for a, b in ((small[0], big[0][:300]), (small[1], big[1][:300])):
            assert rel(a, b).max() <= 2.0 ** -22, k


@pytest.mark.parametrize("reorth", ["f32", "f64"])
@pytest.mark.parametrize("k", [100, 200, 250])
def test_breakdown_both_coefficient_forms(gpu, k, reorth):
    """The restart path with each Gram-Schmidt coefficient form, including
    steps whose basis has register columns (k > 148 on the 512-lane kernel):
    the identity (the start vector is an eigenvector, so every step restarts)
    and a subset with a duplicated index (singular G_S, lambda_min = 0),
    against FP64 LAPACK."""
    n = 512
    G = torch.eye(n, device="cuda", dtype=torch.float32).contiguous()
    idx = torch.arange(k, dtype=torch.int32, device="cuda")
    lo, hi, de, st = _np(solver.solve(G, idx, k=k, reorth=reorth))
    assert st[0] & 1 and (st[0] & ~1) == 0
    assert abs(lo[0] - 1) < 1e-6 and abs(hi[0] - 1) < 1e-6 and abs(de[0]) < 1e-6
    G = solver.gram(solver.gen_phi("gauss", 4, 2 * k, n))
    Gh = G.cpu().numpy()
    rows = []
    for r in range(3):
        s = oracle.floyd(11, r, n, k)
        s[5 + r] = s[k - 3 - r]
        rows.append(s)
    flat = np.concatenate(rows).astype(np.int32)
    lo, hi, de, st = _np(solver.solve(G, torch.from_numpy(flat).cuda(), k=k, reorth=reorth))
    for j, s in enumerate(rows):
        A, _ = oracle.gather(Gh, s)
        w = np.linalg.eigvalsh(A)
        assert (st[j] & ~1) == 0, (j, st[j])
        assert abs(lo[j] - w[0]) <= RTOL * w[-1], (j, lo[j], w[0])
        assert rel(hi[j], w[-1]) <= RTOL, (j, hi[j], w[-1])


@pytest.mark.parametrize("form", ["tuned", "plain_wave"])
@pytest.mark.parametrize("k", [100, 200])
def test_edge_cases_large(gpu, k, form):
    """The edge cases of the small-k kernels at k > 64 (the 256- and 512-lane
    kernels, and the one-wavefront plain form): identity (a restart at every
    step), duplicated indices (singular G_S, lambda_min = 0), an out-of-range
    index and a non-finite entry (flagged, neighbours unaffected)."""
    n = 512
    G = torch.eye(n, device="cuda", dtype=torch.float32).contiguous()
    idx = torch.arange(k, dtype=torch.int32, device="cuda")
    lo, hi, de, st = _np(solver.solve(G, idx, k=k, form=form))
    assert st[0] & 1 and (st[0] & ~1) == 0
    assert abs(lo[0] - 1) < 1e-6 and abs(hi[0] - 1) < 1e-6 and abs(de[0]) < 1e-6
    G = solver.gram(solver.gen_phi("gauss", 3, 2 * k, n))
    base = oracle.floyd(5, 0, n, k)
    dup = base.copy()
    dup[3] = dup[k - 7]
    lo, hi, de, st = _np(solver.solve(G, torch.from_numpy(dup).cuda(), k=k, form=form))
    A, _ = oracle.gather(G.cpu().numpy(), dup)
    w = np.linalg.eigvalsh(A)
    assert (st[0] & ~1) == 0
    assert abs(lo[0] - w[0]) <= 1e-5 * w[-1] and rel(hi[0], w[-1]) <= RTOL
    two = np.concatenate([base, base]).astype(np.int32)
    two[k + 2] = n + 5
    lo, hi, de, st = _np(solver.solve(G, torch.from_numpy(two).cuda(), k=k, form=form))
    assert st[1] & 8 and np.isnan(lo[1]) and st[0] == 0 and np.isfinite(lo[0])
    Gn = G.clone()
    Gn[base[0], base[1]] = float("nan")
    Gn[base[1], base[0]] = float("nan")
    lo, hi, de, st = _np(solver.solve(Gn, torch.from_numpy(base).cuda(), k=k, form=form))
    assert st[0] & 2 and np.isnan(lo[0])


def test_mixed_batch_on_a_side_stream_is_stream_ordered(gpu):
    """A tuned mixed batch with k > 128 forks its smaller bins to a helper
    stream and joins them back before the bisection: work queued after the
    call on the caller's stream sees the complete outputs, repeatedly, on a
    non-default stream as on the default one. The fork is checked to have
    happened (the library's fork counter), and the outputs equal those of the
    same batch with every bin on the caller's stream."""
    n = 2048
    G = solver.gram(solver.gen_phi("gauss", 0xD1, 512, n))
    ks = solver.gen_ks(0xD2, 900, [16, 64, 128, 256])
    off = solver.offsets_from_ks(ks)
    idx = solver.gen_subsets_csr(0xD3, off, n)
    f0 = solver.overlap_forks()
    seq = _np(solver.solve(G, idx, k=256, offsets=off, overlap=False))
    assert solver.overlap_forks() == f0
    ref = _np(solver.solve(G, idx, k=256, offsets=off))
    assert solver.overlap_forks() == f0 + 1
    for a, b in zip(ref, seq):
        assert np.array_equal(a, b, equal_nan=True)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    copies = []
    with torch.cuda.stream(side):
        out = solver.solve(G, idx, k=256, offsets=off)
        for _ in range(3):
            for t in out:
                t.fill_(0)
            solver.solve(G, idx, k=256, offsets=off, out=out)
            copies.append([t.clone() for t in out])
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    assert solver.overlap_forks() == f0 + 5
    for c in copies:
        for a, b in zip(_np(c), seq):
            assert np.array_equal(a, b, equal_nan=True)
