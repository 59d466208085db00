"""pl_solve_noalloc (no device allocation: its workspace is the library's
static per-GPU scratch) against pl_solve, through the C ABI: bit-identical
outputs on two calls in a row, the second with a smaller input, in the tuned
and plain forms; oversized calls and the one-wavefront form are rejected
before any device work."""
import ctypes

import numpy as np
import pytest
import torch

import paralanczos_amd
from paralanczos_amd import solver

pytestmark = pytest.mark.gpu

KS = [16, 32, 64, 128, 256]
MAX_BYTES = 32 << 20


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _outs(nsub):
    f = [torch.full((nsub,), -7.0, dtype=torch.float32, device="cuda") for _ in range(3)]
    return f + [torch.full((nsub,), 12345, dtype=torch.int32, device="cuda")]


def _noalloc(L, G, idx, off, k, nsub, form):
    out = _outs(nsub)
    st = L.pl_solve_noalloc(_p(G), G.shape[0], G.stride(0), _p(idx), _p(off), int(k), int(nsub),
                            *[_p(t) for t in out], int(form),
                            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    return st, out


def _np(ts):
    return [t.cpu().numpy() for t in ts]


def test_noalloc_matches_pl_solve_two_calls(gpu):
    L = paralanczos_amd.load()
    assert L.pl_init(0) == 0
    n = 4096
    G = solver.gram(solver.gen_phi("gauss", 0xA1, 1024, n))
    # call 1: a mixed batch of 3000 subsets up to k = 256
    ks = solver.gen_ks(0xA2, 3000, KS)
    off = solver.offsets_from_ks(ks)
    idx = solver.gen_subsets_csr(0xA3, off, n)
    # call 2, smaller: 500 subsets of k = 100
    idx2 = solver.gen_subsets(0xA4, 500, 100, n)
    for form, name in ((0, "tuned"), (1, "plain")):
        assert L.pl_solve_workspace_size(3000, 256) <= MAX_BYTES
        st, got = _noalloc(L, G, idx, off, 256, 3000, form)
        assert st == 0, L.pl_last_error()
        ref = _np(solver.solve(G, idx, k=256, offsets=off, form=name))
        for a, b in zip(_np(got), ref):
            assert np.array_equal(a, b, equal_nan=True), name
        assert np.all(ref[3] == 0)
        st, got2 = _noalloc(L, G, idx2, None, 100, 500, form)
        assert st == 0, L.pl_last_error()
        ref2 = _np(solver.solve(G, idx2, k=100, form=name))
        for a, b in zip(_np(got2), ref2):
            assert np.array_equal(a, b, equal_nan=True), name
        assert np.all(ref2[3] == 0)


def test_noalloc_rejects_before_device_work(gpu):
    L = paralanczos_amd.load()
    n = 512
    G = torch.zeros((n, n), dtype=torch.float32, device="cuda")
    nsub = MAX_BYTES // (8 * 256) + 1  # T alone exceeds the scratch at k = 256
    assert L.pl_solve_workspace_size(nsub, 256) > MAX_BYTES
    idx = torch.zeros(1, dtype=torch.int32, device="cuda")
    out = _outs(1)
    fake = ctypes.c_void_p(out[0].data_ptr())
    st = L.pl_solve_noalloc(_p(G), n, n, _p(idx), None, 256, nsub, fake, fake, fake, fake, 0,
                            None)
    assert st == 1 and b"PL_NOALLOC_SCRATCH_BYTES" in L.pl_last_error()
    st = L.pl_solve_noalloc(_p(G), n, n, _p(idx), None, 1, 1, *[_p(t) for t in out], 2, None)
    assert st == 4 and b"one-wavefront" in L.pl_last_error()
    torch.cuda.synchronize()
    assert out[0].item() == -7.0 and out[3].item() == 12345
