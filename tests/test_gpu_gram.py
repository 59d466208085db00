"""pl_gram (G = Phi^T Phi) and pl_subgram (G_S = Phi_S^T Phi_S per index set)
on the matrix cores, against FP64 products of the same FP32 frame.

The bound is the classical one for a sum of m products in FP32, whatever the
order: |G_ij - (Phi^T Phi)_ij| <= gamma_{m+1} (|Phi|^T |Phi|)_ij with
gamma_t = t u / (1 - t u), u = 2^-24 (one rounding per product and per
addition). Symmetry is exact: each off-diagonal entry is computed once and
stored twice.
"""
import ctypes

import numpy as np
import pytest
import torch

from paralanczos_amd import _lib, solver

pytestmark = pytest.mark.gpu

U = 2.0 ** -24


def _check(G, Phi64):
    m = Phi64.shape[0]
    ref = Phi64.T @ Phi64
    A = np.abs(Phi64)
    bound = (m + 1) * U / (1 - (m + 1) * U) * (A.T @ A) + 1e-30
    err = np.abs(G.astype(np.float64) - ref)
    assert np.all(err <= bound), float((err / bound).max())
    assert np.array_equal(G, G.T)


@pytest.mark.parametrize("m,n", [(1, 1), (3, 5), (16, 129), (64, 256), (100, 300), (513, 1000),
                                 (1024, 2048)])
@pytest.mark.parametrize("layout", ["row", "col"])
def test_gram_vs_fp64(gpu, m, n, layout):
    g = torch.Generator(device="cuda").manual_seed(1000 * m + n)
    Phi = torch.randn((m, n), device="cuda", generator=g, dtype=torch.float32)
    src = Phi if layout == "row" else Phi.T.contiguous()
    G = solver.gram(src, layout=layout).cpu().numpy()
    _check(G, Phi.cpu().numpy().astype(np.float64))


def test_gram_padded_leading_dimensions(gpu):
    """ld > row length for Phi in both layouts and ldg > n for G; the padding
    of G is left untouched."""
    L = _lib.load()
    m, n = 77, 203
    g = torch.Generator(device="cuda").manual_seed(5)
    Phi = torch.randn((m, n), device="cuda", generator=g, dtype=torch.float32)
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for layout, buf, ld in ((0, torch.zeros((m, n + 9), device="cuda"), n + 9),
                            (1, torch.zeros((n, m + 3), device="cuda"), m + 3)):
        if layout == 0:
            buf[:, :n] = Phi
        else:
            buf[:, :m] = Phi.T
        Gb = torch.full((n, n + 5), 7.0, device="cuda")
        assert L.pl_gram(P(buf), m, n, ld, layout, P(Gb), n + 5, st) == 0
        torch.cuda.synchronize()
        Gh = Gb.cpu().numpy()
        assert np.all(Gh[:, n:] == 7.0)
        _check(np.ascontiguousarray(Gh[:, :n]), Phi.cpu().numpy().astype(np.float64))


def test_gram_of_generated_frames(gpu):
    """gen_phi's Gaussian frame has unit columns, so diag(G) = 1; for every
    kind the Gram matrix of the generated frame meets the FP32 bound."""
    for kind in ("gauss", "bernoulli", "dct"):
        phiT = solver.gen_phi(kind, 0x5A, 256, 1024)
        G = solver.gram(phiT).cpu().numpy()
        _check(G, phiT.cpu().numpy().astype(np.float64).T)
        if kind == "gauss":
            assert np.abs(np.diag(G) - 1).max() < 1e-5


def test_gram_argument_errors(gpu):
    L = _lib.load()
    t = torch.zeros(16, device="cuda")
    P = ctypes.c_void_p(t.data_ptr())
    assert L.pl_gram(None, 4, 4, 4, 0, P, 4, None) == 1
    assert L.pl_gram(P, 0, 4, 4, 0, P, 4, None) == 1
    assert L.pl_gram(P, 4, 4, 3, 0, P, 4, None) == 1   # row layout: ld < n
    assert L.pl_gram(P, 4, 2, 3, 1, P, 2, None) == 1   # column layout: ld < m
    assert L.pl_gram(P, 4, 4, 4, 2, P, 4, None) == 1   # unknown layout
    assert L.pl_gram(P, 4, 4, 4, 0, P, 3, None) == 1   # ldg < n


@pytest.mark.parametrize("k", [1, 5, 32, 64, 65, 128, 200, 256])
@pytest.mark.parametrize("layout", ["row", "col"])
def test_subgram_vs_fp64_and_gram(gpu, k, layout):
    """Every sub-Gram meets the FP32 bound against FP64 and equals, bit for
    bit, the G_S that pl_gather takes from pl_gram's G (same accumulation
    order over the frame's rows)."""
    m, n, nsub = 300, 700, 12
    g = torch.Generator(device="cuda").manual_seed(77 + k)
    Phi = torch.randn((m, n), device="cuda", generator=g, dtype=torch.float32)
    src = Phi if layout == "row" else Phi.T.contiguous()
    idx = solver.gen_subsets(900 + k, nsub, k, n)
    out, st = solver.subgram(src, idx, k, layout=layout)
    out = out.cpu().numpy()
    assert np.all(st.cpu().numpy() == 0)
    P64 = Phi.cpu().numpy().astype(np.float64)
    ih = idx.cpu().numpy().reshape(nsub, k)
    for s in range(nsub):
        _check(out[s], P64[:, ih[s]])
    G = solver.gram(src, layout=layout)
    via_g = solver.gather(G, idx, k, "dump", 0, 4).cpu().numpy().reshape(nsub, k, k)
    assert np.array_equal(out, via_g)


def test_subgram_bad_index_flagged(gpu):
    m, n, k = 64, 100, 16
    g = torch.Generator(device="cuda").manual_seed(3)
    phiT = torch.randn((n, m), device="cuda", generator=g, dtype=torch.float32)
    idx = solver.gen_subsets(4, 3, k, n)
    idx[k + 5] = n  # subset 1 reaches past the frame
    out, st = solver.subgram(phiT, idx, k)
    st = st.cpu().numpy()
    assert st[0] == 0 and st[2] == 0 and st[1] == 8
