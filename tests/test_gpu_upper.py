"""pl_upper, the counterpart of UPPER_KER (ParaLanczos.cu:1484-1626): the
power-iteration estimate of lambda_max per subset, against the FP64 oracle
(or_upper on the gathered FP32 G_S, pinned by closed forms in
tests/test_oracle.py) to 1e-5 relative; the diagonal closed form; the full
matrix as the identity index set (the reference's own case, n <= 128);
bounded by dsyevd's lambda_max; bad inputs flagged."""
import math

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


def test_upper_mixed_sizes_vs_oracle(gpu):
    n = 4096
    G = solver.gram(solver.gen_phi("gauss", 0xB1, 1024, n))
    ks = torch.tensor([1, 2, 7, 16, 33, 64, 65, 100, 128, 129, 200, 256] * 25, dtype=torch.int32,
                      device="cuda")
    off = solver.offsets_from_ks(ks)
    idx = solver.gen_subsets_csr(0xB2, off, n)
    Gh = G.cpu().numpy()
    for iters in (1, 10, 30):
        up, st = solver.upper(G, idx, offsets=off, iters=iters)
        up, st = up.cpu().numpy(), st.cpu().numpy()
        ref, rst = oracle.upper_batch(Gh, idx.cpu().numpy(), offsets=off.cpu().numpy(),
                                      iters=iters, nthreads=8)
        assert np.all(st == 0) and np.all(rst == 0)
        assert rel(up, ref).max() <= RTOL, iters
    # bounded by the largest eigenvalue of each gathered G_S
    o = off.cpu().numpy()
    ix = idx.cpu().numpy()
    for s in range(0, len(o) - 1, 7):
        A, _ = oracle.gather(Gh, ix[o[s]:o[s + 1]])
        assert up[s] <= np.linalg.eigvalsh(A)[-1] * (1 + RTOL)
    # uniform k, each kernel capacity
    for k in (40, 128, 256):
        idu = solver.gen_subsets(0xB3 + k, 300, k, n)
        up, st = solver.upper(G, idu, k=k)
        ref, _ = oracle.upper_batch(Gh, idu.cpu().numpy(), k=k, iters=10, nthreads=8)
        assert np.all(st.cpu().numpy() == 0)
        assert rel(up.cpu().numpy(), ref).max() <= RTOL, k


def test_upper_diagonal_and_full_matrix(gpu):
    d = np.linspace(0.25, 3.0, 128)
    rng = np.random.default_rng(4)
    rng.shuffle(d)
    G = torch.diag(torch.tensor(d, dtype=torch.float32)).cuda().contiguous()
    full = torch.arange(128, dtype=torch.int32, device="cuda")  # UPPER_KER's full matrix
    df = np.float64(np.float32(d))
    for iters in (1, 10):
        up, st = solver.upper(G, full, k=128, iters=iters)
        expect = math.sqrt(np.sum(df ** (2 * iters)) / np.sum(df ** (2 * (iters - 1))))
        assert st.item() == 0 and rel(up.item(), expect) <= RTOL
    # a dense full matrix of the reference's largest size
    X = rng.normal(size=(300, 128))
    A = (X.T @ X / 300).astype(np.float32)
    up, st = solver.upper(torch.from_numpy(A).cuda(), full, k=128)
    assert st.item() == 0 and rel(up.item(), oracle.upper(np.float64(A), 10)) <= RTOL
    # many steps converge to lambda_max
    up, _ = solver.upper(torch.from_numpy(A).cuda(), full, k=128, iters=3000)
    assert rel(up.item(), np.linalg.eigvalsh(np.float64(A))[-1]) <= 1e-5


def test_upper_flags_bad_inputs(gpu):
    n = 256
    G = solver.gram(solver.gen_phi("gauss", 0xB4, 128, n))
    k = 20
    idx = solver.gen_subsets(0xB5, 3, k, n)
    idx[k + 4] = n + 3                       # subset 1: out of range
    Gn = G.clone()
    i0, i1 = idx[2 * k].item(), idx[2 * k + 1].item()
    Gn[min(i0, i1), max(i0, i1)] = float("inf")  # subset 2: non-finite
    up, st = solver.upper(Gn, idx, k=k)
    up, st = up.cpu().numpy(), st.cpu().numpy()
    assert st[0] == 0 and np.isfinite(up[0])
    assert st[1] == solver.SUB_BADINDEX and np.isnan(up[1])
    assert st[2] == solver.SUB_NONFINITE and np.isnan(up[2])
    ks = torch.tensor([5, 0, 40], dtype=torch.int32, device="cuda")
    off = solver.offsets_from_ks(ks)
    idx2 = solver.gen_subsets_csr(0xB6, off, n)
    up, st = solver.upper(G, idx2, k=32, offsets=off)  # 40 > k: BADK, 0: BADK
    st = st.cpu().numpy()
    assert st[0] == 0 and st[1] == solver.SUB_BADK and st[2] == solver.SUB_BADK
