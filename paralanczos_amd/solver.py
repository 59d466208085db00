"""Device-tensor front end of the batched-submatrix Lanczos path.

Thin wrappers over the C ABI of ``include/paralanczos.h``: torch supplies the
device memory and the current HIP stream, the library does all the work.
Every function raises ``ParaLanczosError`` when the native call fails; there
is no host fallback.
"""
from __future__ import annotations

import ctypes

import torch

from ._lib import ParaLanczosError, check, load

FORMS = {"tuned": 0, "plain": 1}
PHI_KINDS = {"gauss": 0, "bernoulli": 1, "dct": 2}

SUB_BREAKDOWN = 1
SUB_NONFINITE = 2
SUB_NOCONV = 4
SUB_BADINDEX = 8
SUB_BADK = 16


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dev(t: torch.Tensor, dtype, name):
    if not t.is_cuda:
        raise ParaLanczosError(f"{name} must be a device tensor")
    if t.dtype != dtype:
        raise ParaLanczosError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ParaLanczosError(f"{name} must be contiguous")
    return t


def max_k() -> int:
    return int(load().pl_max_k())


def gen_subsets(seed: int, nsub: int, k: int, n: int, s0: int = 0, device="cuda"):
    """nsub seeded index sets of size k drawn from [0, n), concatenated (int32)."""
    idx = torch.empty(nsub * k, dtype=torch.int32, device=device)
    check(load().pl_gen_subsets(seed, s0, nsub, k, n, _p(idx), _stream()), "pl_gen_subsets")
    return idx


def gen_ks(seed: int, nsub: int, klist, s0: int = 0, device="cuda"):
    """Per-subset sizes drawn uniformly from klist (at most 32 entries)."""
    kl = (ctypes.c_int32 * len(klist))(*[int(x) for x in klist])
    ks = torch.empty(nsub, dtype=torch.int32, device=device)
    check(load().pl_gen_ks(seed, s0, nsub, ctypes.cast(kl, ctypes.c_void_p), len(klist), _p(ks),
                           _stream()), "pl_gen_ks")
    return ks


def offsets_from_ks(ks: torch.Tensor) -> torch.Tensor:
    off = torch.zeros(ks.numel() + 1, dtype=torch.int64, device=ks.device)
    torch.cumsum(ks.to(torch.int64), 0, out=off[1:])
    return off


def gen_subsets_csr(seed: int, offsets: torch.Tensor, n: int, s0: int = 0):
    _dev(offsets, torch.int64, "offsets")
    total = int(offsets[-1].item())
    idx = torch.empty(total, dtype=torch.int32, device=offsets.device)
    check(load().pl_gen_subsets_csr(seed, s0, offsets.numel() - 1, _p(offsets), n, _p(idx),
                                    _stream()), "pl_gen_subsets_csr")
    return idx


def gen_phi(kind: str, seed: int, m: int, n: int, device="cuda"):
    """Sensing matrix Phi (m x n) returned transposed: an (n, m) tensor whose
    row j is column j of Phi (the library's column-major layout)."""
    phi = torch.empty((n, m), dtype=torch.float32, device=device)
    check(load().pl_gen_phi(PHI_KINDS[kind], seed, m, n, _p(phi), _stream()), "pl_gen_phi")
    return phi


def gram(phiT: torch.Tensor) -> torch.Tensor:
    """G = Phi^T Phi in FP32 (a plain library GEMM; no TF32/XF32), made exactly
    symmetric by averaging with its transpose."""
    prev = torch.backends.cuda.matmul.allow_tf32
    torch.backends.cuda.matmul.allow_tf32 = False
    try:
        G = phiT @ phiT.T
    finally:
        torch.backends.cuda.matmul.allow_tf32 = prev
    G.add_(G.T.contiguous()).mul_(0.5)
    return G


def solve(G, idx, k=None, offsets=None, form="tuned", lanes=None, pad=None, out=None):
    """lambda_min, lambda_max, delta (FP32) and status (int32 view of the u32
    status word) per subset. Uniform size k, or CSR offsets (k = largest size)."""
    L = load()
    _dev(G, torch.float32, "G")
    _dev(idx, torch.int32, "idx")
    n, ld = G.shape[0], G.stride(0)
    if offsets is not None:
        _dev(offsets, torch.int64, "offsets")
        nsub = offsets.numel() - 1
        if k is None:
            k = int((offsets[1:] - offsets[:-1]).max().item()) if nsub > 0 else 1
    else:
        if k is None:
            raise ParaLanczosError("k is required without offsets")
        nsub = idx.numel() // k
    if out is None:
        lmin = torch.empty(nsub, dtype=torch.float32, device=G.device)
        lmax = torch.empty_like(lmin)
        delta = torch.empty_like(lmin)
        status = torch.empty(nsub, dtype=torch.int32, device=G.device)
    else:
        lmin, lmax, delta, status = out
    args = (_p(G), n, ld, _p(idx), _p(offsets), int(k), nsub, _p(lmin), _p(lmax), _p(delta),
            _p(status))
    if lanes is not None or pad is not None:
        lanes = lanes if lanes is not None else 64
        pad = pad if pad is not None else 4
        check(L.pl_solve_variant(*args, int(lanes), int(pad), _stream()), "pl_solve_variant")
    else:
        check(L.pl_solve(*args, FORMS[form], _stream()), "pl_solve")
    return lmin, lmax, delta, status


def solve_alleig(G, idx, k):
    """All k eigenvalues ascending (FP64), det and status per subset."""
    L = load()
    _dev(G, torch.float32, "G")
    _dev(idx, torch.int32, "idx")
    nsub = idx.numel() // k
    eig = torch.empty(nsub * k, dtype=torch.float64, device=G.device)
    det = torch.empty(nsub, dtype=torch.float64, device=G.device)
    err = torch.empty(nsub, dtype=torch.int32, device=G.device)
    check(L.pl_solve_alleig(_p(G), G.shape[0], G.stride(0), _p(idx), int(k), nsub, _p(eig),
                            _p(det), _p(err), _stream()), "pl_solve_alleig")
    return eig, det, err


def gather(G, idx, k, mode="dump", lanes=0, pad=4):
    """Gather stage alone: the k x k submatrices (mode 'dump') or their entry
    sums (mode 'checksum')."""
    L = load()
    _dev(G, torch.float32, "G")
    _dev(idx, torch.int32, "idx")
    nsub = idx.numel() // k
    size = nsub * k * k if mode == "dump" else nsub
    out = torch.empty(size, dtype=torch.float32, device=G.device)
    check(L.pl_gather(_p(G), G.shape[0], G.stride(0), _p(idx), int(k), nsub, _p(out),
                      0 if mode == "dump" else 1, int(lanes), int(pad), _stream()), "pl_gather")
    return out
