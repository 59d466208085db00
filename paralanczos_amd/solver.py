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

FORMS = {"tuned": 0, "plain": 1, "plain_wave": 2}
REORTH = {None: 0, "auto": 0, "f32": 0x100, "f64": 0x200}
NO_OVERLAP = 0x400
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


LAYOUTS = {"row": 0, "col": 1}


def gram(phi: torch.Tensor, layout: str = "col") -> torch.Tensor:
    """G = Phi^T Phi (FP32, exactly symmetric) on the matrix cores (pl_gram).
    layout 'col': phi is the (n, m) tensor gen_phi returns (row j = column j
    of Phi); 'row': phi is Phi itself, (m, n), as the reference's frame."""
    L = load()
    _dev(phi, torch.float32, "phi")
    if phi.dim() != 2 or phi.stride(1) != 1:
        raise ParaLanczosError("phi must be a 2-D tensor with unit column stride")
    if layout == "col":
        n, m = phi.shape
    elif layout == "row":
        m, n = phi.shape
    else:
        raise ParaLanczosError("layout must be 'row' or 'col'")
    G = torch.empty((n, n), dtype=torch.float32, device=phi.device)
    check(L.pl_gram(_p(phi), m, n, phi.stride(0), LAYOUTS[layout], _p(G), n, _stream()), "pl_gram")
    return G


def subgram(phi: torch.Tensor, idx: torch.Tensor, k: int, layout: str = "col"):
    """G_S = Phi_S^T Phi_S of every index set S (uniform k) straight from the
    frame (pl_subgram): returns (nsub, k, k) FP32 and the int32 status view."""
    L = load()
    _dev(phi, torch.float32, "phi")
    _dev(idx, torch.int32, "idx")
    if phi.dim() != 2 or phi.stride(1) != 1:
        raise ParaLanczosError("phi must be a 2-D tensor with unit column stride")
    if layout == "col":
        n, m = phi.shape
    elif layout == "row":
        m, n = phi.shape
    else:
        raise ParaLanczosError("layout must be 'row' or 'col'")
    nsub = idx.numel() // k
    out = torch.empty((nsub, k, k), dtype=torch.float32, device=phi.device)
    status = torch.empty(nsub, dtype=torch.int32, device=phi.device)
    check(L.pl_subgram(_p(phi), m, n, phi.stride(0), LAYOUTS[layout], _p(idx), int(k), nsub,
                       _p(out), _p(status), _stream()), "pl_subgram")
    return out, status


def solve(G, idx, k=None, offsets=None, form="tuned", lanes=None, pad=None, out=None,
          reorth=None, overlap=True):
    """lambda_min, lambda_max, delta (FP32) and status (int32 view of the u32
    status word) per subset. Uniform size k, or CSR offsets (k = largest size).
    reorth: None/'auto' (FP64 Gram-Schmidt coefficients from k = 128 on in the
    tuned form), 'f32' or 'f64' (forced for k > 64). overlap=False keeps every
    bin of a tuned mixed batch on the current stream (no helper stream)."""
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
        flags = FORMS[form] | REORTH[reorth] | (0 if overlap else NO_OVERLAP)
        check(L.pl_solve(*args, flags, _stream()), "pl_solve")
    return lmin, lmax, delta, status


def overlap_forks() -> int:
    """Mixed tuned batches split over the caller's stream and a helper stream
    since the library was loaded."""
    return int(load().pl_overlap_forks())


def upper(G, idx, k=None, offsets=None, iters=10):
    """Power-iteration estimate of lambda_max per subset (FP32; the
    reference's UPPER_KER runs iters = 10) and the status (int32)."""
    L = load()
    _dev(G, torch.float32, "G")
    _dev(idx, torch.int32, "idx")
    if offsets is not None:
        _dev(offsets, torch.int64, "offsets")
        nsub = offsets.numel() - 1
        if k is None:
            k = int((offsets[1:] - offsets[:-1]).max().item()) if nsub > 0 else 1
    else:
        if k is None:
            raise ParaLanczosError("k is required without offsets")
        nsub = idx.numel() // k
    up = torch.empty(nsub, dtype=torch.float32, device=G.device)
    status = torch.empty(nsub, dtype=torch.int32, device=G.device)
    check(L.pl_upper(_p(G), G.shape[0], G.stride(0), _p(idx), _p(offsets), int(k), nsub,
                     int(iters), _p(up), _p(status), _stream()), "pl_upper")
    return up, status


def workspace_size(nsub, k):
    return int(load().pl_solve_workspace_size(int(nsub), int(k)))


def solve_stage(G, idx, stage, ws, k=None, offsets=None, form="tuned", reorth=None, out=None):
    """One stage of a solve on the caller's workspace tensor (uint8, at least
    workspace_size(nsub, k) bytes): 1 = gather + Lanczos (T kept in ws),
    2 = bisection of T. Returns the output tensors (written by stage 2)."""
    L = load()
    _dev(G, torch.float32, "G")
    _dev(idx, torch.int32, "idx")
    if offsets is not None:
        _dev(offsets, torch.int64, "offsets")
        nsub = offsets.numel() - 1
    else:
        nsub = idx.numel() // k
    if out is None:
        out = (torch.empty(nsub, dtype=torch.float32, device=G.device),
               torch.empty(nsub, dtype=torch.float32, device=G.device),
               torch.empty(nsub, dtype=torch.float32, device=G.device),
               torch.empty(nsub, dtype=torch.int32, device=G.device))
    check(L.pl_solve_stage(_p(G), G.shape[0], G.stride(0), _p(idx), _p(offsets), int(k), nsub,
                           _p(out[0]), _p(out[1]), _p(out[2]), _p(out[3]),
                           FORMS[form] | REORTH[reorth], int(stage), _p(ws), ws.numel(),
                           _stream()), "pl_solve_stage")
    return out


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


RIP_FIELDS = ("count", "lmin_mean", "lmin_std", "lmax_mean", "lmax_std", "delta_max",
              "lmin_min", "lmax_max")


def rip_stats(lmin, lmax, delta, status, k=None, offsets=None, kmax=None):
    """Per-size statistics of a solved batch: a (kmax+1, 8) FP64 device tensor,
    row k = (count, mean/std lambda_min, mean/std lambda_max, max delta,
    min lambda_min, max lambda_max); row 0 counts the skipped subsets."""
    L = load()
    nsub = lmin.numel()
    if kmax is None:
        kmax = k if offsets is None else int((offsets[1:] - offsets[:-1]).max().item())
    out = torch.empty((kmax + 1, 8), dtype=torch.float64, device=lmin.device)
    check(L.pl_rip_stats(_p(lmin), _p(lmax), _p(delta), _p(status), _p(offsets),
                         int(k or 0), nsub, int(kmax), _p(out), _stream()), "pl_rip_stats")
    return out


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


class Plan:
    """A solve bound to fixed device tensors (``pl_plan_*``): the first
    ``execute`` of a tuned plan records a HIP graph, later ones replay it. The
    tensors' contents may change between executions, their storage may not."""

    def __init__(self, G, idx, k=None, offsets=None, form="tuned", reorth=None):
        L = load()
        _dev(G, torch.float32, "G")
        _dev(idx, torch.int32, "idx")
        if offsets is not None:
            _dev(offsets, torch.int64, "offsets")
            nsub = offsets.numel() - 1
            if k is None:
                raise ParaLanczosError("k (the largest subset size) is required")
        else:
            if k is None:
                raise ParaLanczosError("k is required without offsets")
            nsub = idx.numel() // k
        self.nsub = nsub
        self.lmin = torch.empty(nsub, dtype=torch.float32, device=G.device)
        self.lmax = torch.empty_like(self.lmin)
        self.delta = torch.empty_like(self.lmin)
        self.status = torch.empty(nsub, dtype=torch.int32, device=G.device)
        # keep the bound tensors alive as long as the plan
        self._keep = (G, idx, offsets)
        h = ctypes.c_void_p()
        check(L.pl_plan_create(ctypes.byref(h), _p(G), G.shape[0], G.stride(0), _p(idx),
                               _p(offsets), int(k), nsub, _p(self.lmin), _p(self.lmax),
                               _p(self.delta), _p(self.status), FORMS[form] | REORTH[reorth]),
              "pl_plan_create")
        self._h = h

    def execute(self):
        check(load().pl_plan_execute(self._h, _stream()), "pl_plan_execute")
        return self.lmin, self.lmax, self.delta, self.status

    @property
    def is_graph(self) -> bool:
        return bool(load().pl_plan_is_graph(self._h))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            check(load().pl_plan_destroy(self._h), "pl_plan_destroy")
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(G, klist, sample=65536, seed=0x7A, stream=True):
    """Time every (lanes, pad) variant per size in klist on seeded subsets of
    G and store the winners in the library's tuning table. Returns
    {k: (lanes, pad, ms)}."""
    L = load()
    _dev(G, torch.float32, "G")
    ks = [int(x) for x in klist]
    kl = (ctypes.c_int32 * len(ks))(*ks)
    best = (ctypes.c_double * len(ks))()
    check(L.pl_autotune(_p(G), G.shape[0], G.stride(0), ctypes.cast(kl, ctypes.c_void_p),
                        len(ks), int(sample), int(seed), ctypes.cast(best, ctypes.c_void_p),
                        _stream()), "pl_autotune")
    out = {}
    for i, k in enumerate(ks):
        lanes, pad = tuning_get(k)
        out[k] = (lanes, pad, float(best[i]))
    return out


TUNE_LANES = (8, 16, 32, 64, 128, 256, 512)
TUNE_PADS = (0, 1, 4)


def autotune_edges(G, klist, sample=65536, seed=0x7A):
    """pl_autotune_edges: every variant that takes a listed size competes
    (across kernel families), unlisted sizes take the next listed size's
    winner. Returns ({k: (lanes, pad, ms)}, {k: {(lanes, pad): ms}})."""
    L = load()
    _dev(G, torch.float32, "G")
    ks = [int(x) for x in klist]
    kl = (ctypes.c_int32 * len(ks))(*ks)
    best = (ctypes.c_double * len(ks))()
    nv = len(TUNE_LANES) * len(TUNE_PADS)
    sweep = (ctypes.c_double * (len(ks) * nv))()
    check(L.pl_autotune_edges(_p(G), G.shape[0], G.stride(0), ctypes.cast(kl, ctypes.c_void_p),
                              len(ks), int(sample), int(seed), ctypes.cast(best, ctypes.c_void_p),
                              ctypes.cast(sweep, ctypes.c_void_p), _stream()),
          "pl_autotune_edges")
    win, table = {}, {}
    for i, k in enumerate(ks):
        lanes, pad = tuning_get(k)
        win[k] = (lanes, pad, float(best[i]))
        table[k] = {(TUNE_LANES[v // 3], TUNE_PADS[v % 3]): float(sweep[i * nv + v])
                    for v in range(nv) if sweep[i * nv + v] >= 0}
    return win, table


def tuning_get(k):
    lanes, pad = ctypes.c_int(), ctypes.c_int()
    check(load().pl_tuning_get(int(k), ctypes.byref(lanes), ctypes.byref(pad)), "pl_tuning_get")
    return lanes.value, pad.value


def tuning_set(k, lanes, pad):
    check(load().pl_tuning_set(int(k), int(lanes), int(pad)), "pl_tuning_set")


def tuning_reset():
    check(load().pl_tuning_reset(), "pl_tuning_reset")


class Server:
    """Continuous-batching serving engine (``pl_server_*``) over a
    device-resident G. ``submit`` takes device tensors (index sets, and CSR
    offsets or a uniform k) and returns a ticket; ``wait`` returns the outputs
    and the request's latency in seconds."""

    def __init__(self, G, max_subsets=65536, max_indices=1 << 23, max_requests=1024,
                 max_k=256, depth=3, form="tuned", reorth=None, work_budget=0.0):
        from ._lib import ServerConfig
        L = load()
        _dev(G, torch.float32, "G")
        cfg = ServerConfig(int(max_subsets), int(max_indices), int(max_requests), int(max_k),
                           int(depth), FORMS[form] | REORTH[reorth], float(work_budget))
        h = ctypes.c_void_p()
        check(L.pl_server_create(_p(G), G.shape[0], G.stride(0), ctypes.byref(cfg),
                                 ctypes.byref(h)), "pl_server_create")
        self._h = h
        self._G = G
        self._pending = {}

    def submit(self, idx, k=None, offsets=None, kmax=None, out=None):
        _dev(idx, torch.int32, "idx")
        if offsets is not None:
            _dev(offsets, torch.int64, "offsets")
            nsub = offsets.numel() - 1
            if kmax is None:
                raise ParaLanczosError("kmax is required with offsets")
            kk = 0
        else:
            if k is None:
                raise ParaLanczosError("k is required without offsets")
            nsub = idx.numel() // k
            kk, kmax = int(k), int(k)
        dev = idx.device
        if out is None:
            out = (torch.empty(nsub, dtype=torch.float32, device=dev),
                   torch.empty(nsub, dtype=torch.float32, device=dev),
                   torch.empty(nsub, dtype=torch.float32, device=dev),
                   torch.empty(nsub, dtype=torch.int32, device=dev))
        t = ctypes.c_int64()
        check(load().pl_server_submit(self._h, _p(idx), _p(offsets), kk, nsub, idx.numel(),
                                      int(kmax), _p(out[0]), _p(out[1]), _p(out[2]),
                                      _p(out[3]), _stream(), ctypes.byref(t)),
              "pl_server_submit")
        self._pending[t.value] = (idx, offsets, out)
        return t.value

    def wait(self, ticket):
        lat, q = ctypes.c_double(), ctypes.c_double()
        check(load().pl_server_wait(self._h, int(ticket), ctypes.byref(lat), ctypes.byref(q)),
              "pl_server_wait")
        _, _, out = self._pending.pop(ticket)
        return out, lat.value, q.value

    def poll(self, ticket):
        """True once the request is complete (non-blocking)."""
        done = ctypes.c_int()
        check(load().pl_server_poll(self._h, int(ticket), ctypes.byref(done)), "pl_server_poll")
        return bool(done.value)

    def stats(self, reset=False):
        from ._lib import ServerStats
        st = ServerStats()
        check(load().pl_server_query(self._h, ctypes.byref(st), int(bool(reset))),
              "pl_server_query")
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            check(load().pl_server_destroy(self._h), "pl_server_destroy")
            self._h = None
            self._pending.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
