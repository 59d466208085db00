"""CPU tests of the C ABI boundary (no GPU work is launched).

* libparalanczos.so exports every function the public headers declare.
* Argument errors are rejected before any device call, with the reference's
  messages for the drop-in entry points (ParaLanczos.cu:1814-1824, :1734,
  :1695) and the Matlab front-ends' checks (Matlab/gpu_*.m).
* print_tridiagonal reproduces the reference's layout (ParaLanczos.cu:1642-1682).
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import paralanczos_amd
from paralanczos_amd import _lib, legacy

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("ParaLanczos.h", "paralanczos.h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    names.discard("defined")
    return names


def test_headers_declare_expected_surface():
    names = declared_functions()
    for ref in ("calculate_subeigs3", "calculate_subeigs", "calculate_subeigs2", "calculate_eigs",
                "print_tridiagonal"):
        assert ref in names
    assert {"pl_solve", "pl_solve_alleig", "pl_gen_subsets", "pl_gen_phi"} <= names


def test_library_exports_every_declared_symbol():
    L = paralanczos_amd.load()
    names = declared_functions()
    nm = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                        text=True, check=True).stdout
    exported = {line.split()[-1] for line in nm.splitlines() if line.strip()}
    missing = sorted(n for n in names if n not in exported)
    assert not missing, missing
    for n in names:
        getattr(L, n)
        assert n in _lib.SIGNATURES, f"{n} has no ctypes signature"


def test_headers_compile_as_plain_c(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "ParaLanczos.h"\n#include "paralanczos.h"\n'
                   "int main(void){return (int)(sizeof(&calculate_subeigs3) + sizeof(&pl_solve)) > 0 ? 0 : 1;}\n")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    "-c", str(src), "-o", str(tmp_path / "t.o")], check=True)


def test_status_strings_and_version():
    L = paralanczos_amd.load()
    assert L.pl_version().startswith(b"paralanczos-amd")
    assert L.pl_status_string(0) == b"ok"
    assert L.pl_status_string(1) == b"invalid argument"
    assert L.pl_max_k() >= 64


def test_device_api_rejects_bad_arguments_before_any_device_call():
    L = paralanczos_amd.load()
    fake = ctypes.c_void_p(0x1000)
    # null G
    assert L.pl_solve(None, 16, 16, fake, None, 8, 4, fake, fake, fake, fake, 0, None) == 1
    assert b"non-null" in L.pl_last_error()
    # ld < n
    assert L.pl_solve(fake, 16, 8, fake, None, 8, 4, fake, fake, fake, fake, 0, None) == 1
    # k above the supported maximum
    assert L.pl_solve(fake, 1024, 1024, fake, None, 100000, 4, fake, fake, fake, fake, 0,
                      None) == 4
    # unknown form
    assert L.pl_solve(fake, 16, 16, fake, None, 8, 4, fake, fake, fake, fake, 9, None) == 1
    # variant with fewer lanes than k
    assert L.pl_solve_variant(fake, 64, 64, fake, None, 32, 4, fake, fake, fake, fake, 16, 4,
                              None) == 1
    # k > n in the generator
    assert L.pl_gen_subsets(1, 0, 10, 9, 8, fake, None) == 1
    assert L.pl_gen_ks(1, 0, 10, fake, 0, fake, None) == 1
    assert L.pl_gen_phi(7, 1, 4, 4, fake, None) == 1


def test_legacy_argument_messages(capfd):
    g = np.eye(8)
    eig, det = legacy.calculate_subeigs3(g, np.zeros(32, np.int32), 6)
    err = capfd.readouterr().err
    assert "'int subn' must be divisible by 4, and less than or equal to 128!" in err
    assert np.all(eig == 0)
    legacy.calculate_subeigs3(g, np.zeros(20, np.int32), 4)
    assert "'subIndexLen' must be divisible by 32!" in capfd.readouterr().err
    legacy.calculate_eigs(np.eye(6))
    assert "'int n' must be divisible by 4" in capfd.readouterr().err
    legacy.calculate_subeigs(np.eye(8), np.zeros(3, np.int32))
    assert "'int n' must be divisible by 4" in capfd.readouterr().err


def test_matlab_front_end_checks(capsys):
    g = np.eye(16)
    assert legacy.gpu_subsvd(np.ones((4, 5)), [0, 1]) == (-1, -1, -1)
    assert "don't agree" in capsys.readouterr().out
    assert legacy.gpu_subsvd(g, [0, 1, 2]) == (-1, -1, -1)
    assert "not dyadic" in capsys.readouterr().out
    assert legacy.gpu_multi_subsvd(g, list(range(12)), 8) == (-1, -1, -1)
    assert "minor must divide" in capsys.readouterr().out
    assert legacy.gpu_multi_subsvd(g, list(range(12)), 3) == (-1, -1, -1)
    assert "minor is not dyadic" in capsys.readouterr().out
    assert legacy.gpu_multi_subsvd(np.eye(300), list(range(256)), 256) == (-1, -1, -1)
    assert "> 128 not supported" in capsys.readouterr().out
    assert legacy.gpu_multi_subsvd2(g, list(range(16)), 4) == (-1, -1, -1)
    assert "divisible by 32" in capsys.readouterr().out
    assert legacy.gpu_svd(np.eye(12)) == (-1, -1, -1)
    assert "dimensions are not dyadic" in capsys.readouterr().out
    assert legacy.gpu_svd(np.eye(256)) == (-1, -1, -1)
    assert "larger than 128" in capsys.readouterr().out


def test_print_tridiagonal_layout(capfd):
    legacy.print_tridiagonal([1.0, 2.0, 3.0], [0.0, 0.5, 0.25])
    out = capfd.readouterr().out
    assert out == (" [  1 ,  0.5 ,  0 ; \n"
                   " 0.5 ,  2 ,  0.25 ; \n"
                   " 0 ,  0.25 ,  3 ] \n")


def test_matlab_wrappers_match_python_restatement():
    """The .m front-ends carry the same argument messages as the Python
    restatement of the reference's Matlab checks (no Matlab here to run them)."""
    import inspect
    import re

    from paralanczos_amd import legacy

    mdir = os.path.join(ROOT, "paralanczos_amd", "matlab")
    for name in ("gpu_subsvd", "gpu_multi_subsvd", "gpu_multi_subsvd2", "gpu_svd"):
        text = open(os.path.join(mdir, name + ".m")).read()
        assert f"function [eigs, det1, err1] = {name}(" in text
        m_msgs = [s.replace("''", "'") for s in re.findall(r"disp\('((?:[^']|'')*)'\)", text)]
        py_msgs = re.findall(r'print\("([^"]*)"\)', inspect.getsource(getattr(legacy, name)))
        assert m_msgs == py_msgs, name
        assert "pl_multi_subsvd_call" in text
    call = open(os.path.join(mdir, "pl_multi_subsvd_call.m")).read()
    assert "calllib('libparalanczos', 'pl_multi_subsvd'" in call


def test_python_form_constants_match_the_header():
    """solver.FORMS / REORTH / layout values are the header's #defines."""
    from paralanczos_amd import solver
    text = open(os.path.join(ROOT, "include", "paralanczos.h")).read()
    defs = {m.group(1): int(m.group(2), 0)
            for m in re.finditer(r"#define\s+(PL_\w+)\s+(0x[0-9a-fA-F]+|\d+)", text)}
    assert solver.FORMS == {"tuned": defs["PL_FORM_TUNED"], "plain": defs["PL_FORM_PLAIN"],
                            "plain_wave": defs["PL_FORM_PLAIN_WAVE"]}
    assert solver.REORTH["f32"] == defs["PL_FORM_REORTH_F32"]
    assert solver.REORTH["f64"] == defs["PL_FORM_REORTH_F64"]
    assert solver.LAYOUTS["row"] == defs["PL_LAYOUT_ROW"] and solver.LAYOUTS["col"] == defs["PL_LAYOUT_COL"]
