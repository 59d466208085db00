"""Build recipe for libparalanczos.so (hipcc, gfx950), in-tree.

    python -m paralanczos_amd.build            # build (always recompiles)
    python -m paralanczos_amd.build --asm DIR  # also dump gfx950 assembly into DIR

The library lands next to this file so that it travels with the source tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libparalanczos.so")
SOURCES = ["pl_kernels.hip", "pl_capi.hip", "pl_server.hip", "pl_gram.hip"]
ARCH = "gfx950"


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found; a ROCm toolchain is required to build paralanczos_amd")


def flags() -> list[str]:
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-Wall",
        "-Wno-unused-function",
        # every FMA in the kernels is explicit; implicit contraction would break
        # the TwoSum error terms and make partner lanes compute different bits
        "-ffp-contract=off",
        # no packed-FP32 instructions anywhere in the device code
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", os.path.join(HERE, "..", "include"),
    ]


def build(verbose: bool = False, asm: str | None = None) -> str:
    """Compile libparalanczos.so in place; with asm, also write the device
    assembly of pl_kernels.hip into that directory."""
    t0 = time.time()
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + flags() + ["-shared", "-o", tmp] + [os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libparalanczos.so")
    os.replace(tmp, LIB)
    if asm:
        out = asm
        os.makedirs(out, exist_ok=True)
        acmd = [hipcc()] + flags() + ["--cuda-device-only", "-S", "-o",
                                      os.path.join(out, "pl_kernels.s"),
                                      os.path.join(CSRC, "pl_kernels.hip")]
        subprocess.run(acmd, check=True)
    if verbose:
        print(f"built {LIB} in {time.time() - t0:.1f}s", flush=True)
    return LIB


if __name__ == "__main__":
    # python -m paralanczos_amd.build [--asm DIR]
    a = sys.argv[sys.argv.index("--asm") + 1] if "--asm" in sys.argv else None
    build(verbose=True, asm=a)
