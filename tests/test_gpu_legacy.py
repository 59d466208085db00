"""The drop-in host entry points (ParaLanczos.h) and the Matlab front-end
semantics on the GPU, against the reference examples' known answers (the
8 x 8 example matrix of example1.cu:19-26; FP64 LAPACK values of the
FP32-rounded matrix, relative error <= 1e-5) and LAPACK on larger inputs."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
from paralanczos_amd import legacy

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "golden.json")))
A = np.array(GOLD["example_matrix"])
S = GOLD["subsets"]


def close(got, want, rtol=1e-5):
    got, want = np.asarray(got, np.float64), np.asarray(want, np.float64)
    return np.all(np.abs(got - want) <= rtol * np.abs(want))


def test_c_example_program(gpu, tmp_path):
    """legacy_examples.c: the reference's example calls from plain C,
    compiled against include/ParaLanczos.h and linked to the library."""
    lib = os.path.join(ROOT, "paralanczos_amd")
    exe = str(tmp_path / "legacy_examples")
    subprocess.run(["gcc", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"),
                    os.path.join(lib, "examples", "legacy_examples.c"), "-L", lib,
                    "-lparalanczos", f"-Wl,-rpath,{lib}", "-lm", "-o", exe], check=True)
    env = dict(os.environ, PARALANCZOS_QUIET="1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=120)
    sys.stdout.write(r.stdout)
    assert r.returncode == 0 and "all example checks passed" in r.stdout


def test_calculate_entry_points(gpu, monkeypatch):
    monkeypatch.setenv("PARALANCZOS_QUIET", "1")
    eig, det = legacy.calculate_subeigs(A, S["s1357"]["index"])
    assert close(eig, S["s1357"]["eig_fp32_input"]) and close(det, S["s1357"]["det"], 1e-4)
    eig, det = legacy.calculate_eigs(A)
    assert close(eig, S["full_0_7"]["eig_fp32_input"]) and close(det, S["full_0_7"]["det"], 1e-4)
    idx = np.arange(32) % 8
    eig, det = legacy.calculate_subeigs3(A, idx, 4)
    for s in range(8):
        ref = S["s0123"] if s % 2 == 0 else S["s4567"]
        assert close(eig[4 * s:4 * s + 4], ref["eig_fp32_input"])
        assert close(det[s], ref["det"], 1e-4)
    eig, det = legacy.calculate_subeigs2(A, idx[:8], 4)
    assert close(eig[:4], S["s0123"]["eig_fp32_input"]) and close(eig[4:], S["s4567"]["eig_fp32_input"])
    assert close(det, S["s0123"]["det"], 1e-4)


def test_matlab_front_ends(gpu):
    eigs, det, err = legacy.gpu_subsvd(A, S["s1357"]["index"])
    assert close(eigs, S["s1357"]["eig_fp32_input"]) and np.all(err == 0)
    idx = np.arange(32) % 8
    for fn in (legacy.gpu_multi_subsvd, legacy.gpu_multi_subsvd2):
        eigs, det, err = fn(A, idx, 4)
        assert eigs.shape == (32,) and det.shape == (8,) and err.shape == (8,)
        assert close(eigs[:4], S["s0123"]["eig_fp32_input"]) and np.all(err == 0)
        assert close(det[1], S["s4567"]["det"], 1e-4)
    eigs, det, err = legacy.gpu_svd(A)
    assert close(eigs, S["full_0_7"]["eig_fp32_input"]) and close(det[0], S["full_0_7"]["det"], 1e-4)


def test_subeigs3_larger_vs_lapack(gpu, monkeypatch):
    """64 subsets of 32 from a 300 x 300 Gram matrix: every eigenvalue of
    every subset against dsyevd of the FP32-rounded submatrix."""
    monkeypatch.setenv("PARALANCZOS_QUIET", "1")
    rng = np.random.default_rng(3)
    phi = rng.standard_normal((128, 300))
    phi /= np.linalg.norm(phi, axis=0)
    g = phi.T @ phi
    idx = np.concatenate([rng.choice(300, 32, replace=False) for _ in range(64)]).astype(np.int32)
    eig, det = legacy.calculate_subeigs3(g, idx, 32)
    g32 = g.astype(np.float32)
    for s in range(64):
        A_, st = oracle.gather(g32, idx[32 * s:32 * s + 32])
        w = np.linalg.eigvalsh(A_)
        assert np.max(np.abs(eig[32 * s:32 * s + 32] - w)) <= 1e-5 * w[-1]
        assert abs(det[s] - np.prod(w)) <= 1e-3 * abs(np.prod(w))
