"""The configurations at their full sizes, checked by properties that hold at
any size (the oracle and LAPACK check seeded samples of them elsewhere and,
here, a spread sample of each full batch):

* every status word is clean (no error bits; the serving G may restart);
* Rayleigh bounds from the subset's own diagonal: lambda_min <= min_i G_ii
  and lambda_max >= max_i G_ii (e_i is a unit vector), lambda_max <=
  trace(G_S) and lambda_min >= 0 (G_S is a Gram matrix, so positive
  semidefinite), each to the 1e-5 relative bar;
* delta = max(lambda_max - 1, 1 - lambda_min) to FP32 rounding;
* the extremes of a principal submatrix do not depend on the order of its
  indices (the ones start vector is permutation invariant, so the reversed
  order gives the same T up to rounding): within 1e-5 relative;
* two runs of the same batch agree bit for bit;
* FP64 LAPACK on the G_S of 256 subsets spread over the batch: 1e-5.

Sizes (each workload at its full size): k = 32 on G 4096^2 with 2^20 subsets;
k uniform in {8, 16, .., 128} on G 8192^2 with 2^20 subsets; k = 128 and
k = 256 on the Bernoulli G 16384^2 with 131,072 subsets each, both
Gram-Schmidt coefficient forms; one serving step (32,768 subsets, k in
{16, 32, 64, 128, 256}) on the DCT G 32768^2.
"""
import numpy as np
import pytest
import torch

from paralanczos_amd import solver

pytestmark = pytest.mark.gpu

RTOL = 1e-5
NSAMPLE = 256


def rel(x, ref):
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    return np.abs(x - ref) / np.maximum(np.abs(ref), 1e-30)


def _np(ts):
    return [t.cpu().numpy() for t in ts]


def _diag_stats(G, idx, offsets):
    """Per subset: min, max and sum of G's diagonal over its indices."""
    d = torch.diagonal(G)[idx.long()].double().cpu().numpy()
    starts = offsets[:-1].cpu().numpy()
    return (np.minimum.reduceat(d, starts), np.maximum.reduceat(d, starts),
            np.add.reduceat(d, starts))


def _reversed(idx, offsets, nsub, uniform):
    """The first nsub subsets with each one's indices in reverse order."""
    if uniform:
        k = int(offsets[1].item())
        return idx[:nsub * k].view(nsub, k).flip(1).contiguous().view(-1), None
    ih = idx.cpu().numpy()
    off = offsets.cpu().numpy()
    out = np.concatenate([ih[off[s]:off[s + 1]][::-1] for s in range(nsub)]).astype(np.int32)
    return torch.from_numpy(out).cuda(), offsets[:nsub + 1].contiguous()


def _lapack(G, idx, offsets, picks):
    off = offsets.cpu().numpy()
    ih = idx.cpu().numpy()
    lo = np.empty(len(picks))
    hi = np.empty(len(picks))
    for n, s in enumerate(picks):
        sel = torch.from_numpy(ih[off[s]:off[s + 1]].astype(np.int64)).cuda()
        A = G[sel][:, sel].double().cpu().numpy()  # G is exactly symmetric
        w = np.linalg.eigvalsh(A)
        lo[n], hi[n] = w[0], w[-1]
    return lo, hi


def _check(G, idx, offsets, kmax, restart_ok=False, uniform=False, **kw):
    """uniform: every subset has kmax indices and the batch goes through the
    uniform-size entry (offsets then only serve the checks)."""
    nsub = offsets.numel() - 1
    soff = None if uniform else offsets
    out = _np(solver.solve(G, idx, k=kmax, offsets=soff, **kw))
    lo, hi, de, st = out
    if restart_ok:
        assert np.all((st & ~1) == 0), np.unique(st, return_counts=True)
    else:
        assert np.all(st == 0), np.unique(st, return_counts=True)
    dmin, dmax, dsum = _diag_stats(G, idx, offsets)
    lo64, hi64 = lo.astype(np.float64), hi.astype(np.float64)
    assert np.all(lo64 <= hi64)
    assert np.all(lo64 <= dmin + RTOL * dmax), np.max(lo64 - dmin)
    assert np.all(hi64 >= dmax * (1 - RTOL)), np.min(hi64 - dmax)
    assert np.all(hi64 <= dsum * (1 + RTOL))
    assert np.all(lo64 >= -RTOL * hi64), lo64.min()
    dref = np.maximum(hi64 - 1.0, 1.0 - lo64)
    assert np.abs(de - dref).max() <= 2.0 ** -22 * np.maximum(1.0, np.abs(dref)).max()
    again = _np(solver.solve(G, idx, k=kmax, offsets=soff, **kw))
    for a, b in zip(out, again):
        assert np.array_equal(a, b)
    nrev = min(nsub, 65536)
    ridx, roff = _reversed(idx, offsets, nrev, uniform)
    rlo, rhi, _, rst = _np(solver.solve(G, ridx, k=kmax, offsets=roff, **kw))
    assert np.all((rst & ~1) == 0)
    assert np.all(np.abs(rlo - lo[:nrev]) <= RTOL * hi[:nrev])  # lambda_min to lambda_max's scale
    assert rel(rhi, hi[:nrev]).max() <= RTOL
    picks = np.linspace(0, nsub - 1, NSAMPLE).astype(np.int64)
    llo, lhi = _lapack(G, idx, offsets, picks)
    assert np.all(np.abs(lo[picks] - llo) <= RTOL * lhi)
    assert rel(hi[picks], lhi).max() <= RTOL
    return out


def _uniform_offsets(nsub, k):
    return torch.arange(0, (nsub + 1) * k, k, dtype=torch.int64, device="cuda")


def test_full_k32_g4096(gpu):
    G = solver.gram(solver.gen_phi("gauss", 0xC2, 512, 4096))
    nsub, k = 1 << 20, 32
    idx = solver.gen_subsets(0xC2 + 1, nsub, k, 4096)
    _check(G, idx, _uniform_offsets(nsub, k), k, uniform=True)


def test_full_mixed_k_up_to_128_g8192(gpu):
    G = solver.gram(solver.gen_phi("gauss", 0xC3, 1024, 8192))
    nsub = 1 << 20
    ks = solver.gen_ks(0xC3 + 1, nsub, list(range(8, 129, 8)))
    off = solver.offsets_from_ks(ks)
    idx = solver.gen_subsets_csr(0xC3 + 1, off, 8192)
    _check(G, idx, off, 128)


@pytest.mark.parametrize("reorth", ["f32", "f64"])
def test_full_bernoulli_k128_k256_g16384(gpu, reorth):
    G = solver.gram(solver.gen_phi("bernoulli", 0xC4, 2048, 16384))
    nsub = 131072
    for k in (128, 256):
        idx = solver.gen_subsets(0xC4 + 1, nsub, k, 16384, s0=0 if k == 128 else nsub)
        _check(G, idx, _uniform_offsets(nsub, k), k, uniform=True, reorth=reorth)
    del G
    torch.cuda.empty_cache()


def test_full_serving_step_dct_g32768(gpu):
    G = solver.gram(solver.gen_phi("dct", 0xC5, 4096, 32768))
    nsub = 32768
    ks = solver.gen_ks(0xC5 + 1, nsub, [16, 32, 64, 128, 256])
    off = solver.offsets_from_ks(ks)
    idx = solver.gen_subsets_csr(0xC5 + 1, off, 32768)
    _check(G, idx, off, 256, restart_ok=True)
    del G
    torch.cuda.empty_cache()
