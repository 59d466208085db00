"""ctypes binding of libparalanczos.so (the C ABI of include/paralanczos.h and
include/ParaLanczos.h). Loading fails loudly when the library is missing:
there is no fallback path."""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libparalanczos.so")

P = ctypes.c_void_p
I32, I64, U64, DBL = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_double
INT = ctypes.c_int
ST = ctypes.c_int  # pl_status

# name -> (restype, argtypes); every symbol the two public headers declare
SIGNATURES = {
    # paralanczos.h
    "pl_version": (ctypes.c_char_p, []),
    "pl_status_string": (ctypes.c_char_p, [ST]),
    "pl_last_error": (ctypes.c_char_p, []),
    "pl_init": (ST, [INT]),
    "pl_max_k": (INT, []),
    "pl_overlap_forks": (I64, []),
    "pl_solve": (ST, [P, I64, I64, P, P, INT, I64, P, P, P, P, INT, P]),
    "pl_solve_workspace_size": (ctypes.c_size_t, [I64, INT]),
    "pl_solve_noalloc": (ST, [P, I64, I64, P, P, INT, I64, P, P, P, P, INT, P]),
    "pl_solve_ws": (ST, [P, I64, I64, P, P, INT, I64, P, P, P, P, INT, P, ctypes.c_size_t, P]),
    "pl_solve_stage": (ST, [P, I64, I64, P, P, INT, I64, P, P, P, P, INT, INT, P,
                            ctypes.c_size_t, P]),
    "pl_solve_variant": (ST, [P, I64, I64, P, P, INT, I64, P, P, P, P, INT, INT, P]),
    "pl_solve_alleig": (ST, [P, I64, I64, P, INT, I64, P, P, P, P]),
    "pl_upper": (ST, [P, I64, I64, P, P, INT, I64, INT, P, P, P]),
    "pl_plan_create": (ST, [P, P, I64, I64, P, P, INT, I64, P, P, P, P, INT]),
    "pl_plan_execute": (ST, [P, P]),
    "pl_plan_is_graph": (INT, [P]),
    "pl_plan_destroy": (ST, [P]),
    "pl_autotune": (ST, [P, I64, I64, P, INT, I64, U64, P, P]),
    "pl_autotune_edges": (ST, [P, I64, I64, P, INT, I64, U64, P, P, P]),
    "pl_tuning_get": (ST, [INT, P, P]),
    "pl_tuning_set": (ST, [INT, INT, INT]),
    "pl_tuning_reset": (ST, []),
    "pl_server_create": (ST, [P, I64, I64, P, P]),
    "pl_server_submit": (ST, [P, P, P, INT, I64, I64, INT, P, P, P, P, P, P]),
    "pl_server_wait": (ST, [P, I64, P, P]),
    "pl_server_poll": (ST, [P, I64, P]),
    "pl_server_query": (ST, [P, P, INT]),
    "pl_server_destroy": (ST, [P]),
    "pl_rip_stats": (ST, [P, P, P, P, P, INT, I64, INT, P, P]),
    "pl_gather": (ST, [P, I64, I64, P, INT, I64, P, INT, INT, INT, P]),
    "pl_gen_subsets": (ST, [U64, I64, I64, INT, INT, P, P]),
    "pl_gen_subsets_csr": (ST, [U64, I64, I64, P, INT, P, P]),
    "pl_gen_ks": (ST, [U64, I64, I64, P, INT, P, P]),
    "pl_gen_phi": (ST, [INT, U64, INT, INT, P, P]),
    "pl_gram": (ST, [P, I64, I64, I64, INT, P, I64, P]),
    "pl_subgram": (ST, [P, I64, I64, I64, INT, P, INT, I64, P, P, P]),
    # ParaLanczos.h
    "print_tridiagonal": (None, [P, P, INT]),
    "calculate_eigs": (None, [P, INT, P, P]),
    "calculate_subeigs": (None, [P, INT, P, P, P, INT]),
    "calculate_subeigs2": (None, [P, INT, INT, P, P, P, INT]),
    "calculate_subeigs3": (None, [P, INT, INT, P, P, P, INT]),
    "pl_multi_subsvd": (INT, [P, INT, P, INT, INT, P, P, P]),
}



class ServerConfig(ctypes.Structure):
    _fields_ = [("max_subsets", I64), ("max_indices", I64), ("max_requests", I32),
                ("max_k", I32), ("depth", I32), ("form", I32), ("work_budget", DBL)]


class ServerStats(ctypes.Structure):
    _fields_ = [("requests", I64), ("subsets", I64), ("steps", I64), ("span_s", DBL),
                ("p50_s", DBL), ("p99_s", DBL), ("max_s", DBL)]


_lib = None


class ParaLanczosError(RuntimeError):
    pass


def load():
    """Load libparalanczos.so (after torch, so both share one HIP runtime)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ParaLanczosError(
            f"{LIB_PATH} is missing: build it with `python -m paralanczos_amd.build` "
            "(hipcc, gfx950). There is no CPU fallback.")
    try:  # pragma: no cover - torch is optional for the C-ABI symbol checks
        import torch  # noqa: F401
    except Exception:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int, what: str):
    if status != 0:
        L = load()
        raise ParaLanczosError(
            f"{what}: {L.pl_status_string(status).decode()} ({L.pl_last_error().decode()})")
