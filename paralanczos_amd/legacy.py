"""Host-array front ends with the reference's names and semantics.

* ``calculate_*`` / ``print_tridiagonal`` call the identically named C entry
  points of ``include/ParaLanczos.h`` (reference ParaLanczos.h:39-47) through
  ctypes with numpy buffers sized exactly as the reference's callers size
  them.
* ``gpu_subsvd`` / ``gpu_multi_subsvd`` / ``gpu_multi_subsvd2`` / ``gpu_svd``
  restate the Matlab front-ends (Matlab/gpu_subsvd.m:14-59,
  gpu_multi_subsvd.m:13-68, gpu_multi_subsvd2.m:13-73, gpu_svd.m:13-57): same
  argument checks and messages, the same ``-1, -1, -1`` return on a failed
  check, then one call of ``pl_multi_subsvd`` on host arrays (the reference
  launched its kernels from PTX through the Parallel Computing Toolbox).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from ._lib import load


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _mat(mat):
    m = np.ascontiguousarray(mat, dtype=np.float64)
    if m.ndim != 2 or m.shape[0] != m.shape[1]:
        raise ValueError("mat must be a square 2-D array")
    return m


def print_tridiagonal(alpha, beta, n=None):
    a = np.ascontiguousarray(alpha, dtype=np.float64)
    b = np.ascontiguousarray(beta, dtype=np.float64)
    n = len(a) if n is None else n
    load().print_tridiagonal(_p(a), _p(b), n)


def calculate_eigs(mat):
    """ParaLanczos.cu:1685-1722: all eigenvalues of the n x n matrix."""
    m = _mat(mat)
    n = m.shape[0]
    eig = np.zeros(n)
    det = np.zeros(1)
    load().calculate_eigs(_p(m), n, _p(eig), _p(det))
    return eig, det[0]


def calculate_subeigs(mat, subIndex):
    """ParaLanczos.cu:1725-1765 (example1.cu:47): one subset."""
    m = _mat(mat)
    idx = np.ascontiguousarray(subIndex, dtype=np.int32)
    subn = len(idx)
    eig = np.zeros(max(m.shape[0], subn))  # example1.cu allocates n doubles
    det = np.zeros(1)
    load().calculate_subeigs(_p(m), m.shape[0], _p(eig), _p(det), _p(idx), subn)
    return eig[:subn], det[0]


def calculate_subeigs2(mat, subIndex, subn):
    """ParaLanczos.cu:1767-1808: len(subIndex) indices, one determinant back."""
    m = _mat(mat)
    idx = np.ascontiguousarray(subIndex, dtype=np.int32)
    n = len(idx)
    eig = np.zeros(n)
    det = np.zeros(1)
    load().calculate_subeigs2(_p(m), n, m.shape[0], _p(eig), _p(det), _p(idx), subn)
    return eig, det[0]


def calculate_subeigs3(mat, subIndex, subn):
    """ParaLanczos.cu:1811-1863 (example2.cu:54, example3.cu:54)."""
    m = _mat(mat)
    idx = np.ascontiguousarray(subIndex, dtype=np.int32)
    L = len(idx)
    eig = np.zeros(L)
    det = np.zeros(max(L // max(subn, 1), 1))
    load().calculate_subeigs3(_p(m), L, m.shape[0], _p(eig), _p(det), _p(idx), subn)
    return eig, det


def _dyadic(x) -> bool:
    return x > 0 and math.floor(math.log2(x)) - math.log2(x) == 0


def _run(g, idx, minor):
    g = np.ascontiguousarray(g, dtype=np.float64)
    ix = np.ascontiguousarray(idx, dtype=np.int32)
    L = len(ix)
    eigs = np.zeros(L)
    det = np.zeros(L // minor)
    err = np.zeros(L // minor, dtype=np.int32)
    rc = load().pl_multi_subsvd(_p(g), g.shape[0], _p(ix), L, minor, _p(eigs), _p(det), _p(err))
    if rc != 0:
        return -1, -1, -1
    return eigs, det, err


def gpu_subsvd(g, subIndex):
    """Matlab/gpu_subsvd.m:14-59: one subset of dyadic length <= 128."""
    g = np.asarray(g)
    M, N = g.shape
    L = len(subIndex)
    if M != N:
        print("dimensions of input for matrix don't agree! ")
        return -1, -1, -1
    if not _dyadic(L):
        print("length of index is not dyadic!")
        return -1, -1, -1
    if L > 128:
        print("subindex of length > 128 not supported! ")
        return -1, -1, -1
    return _run(g, subIndex, L)


def gpu_multi_subsvd(g, subIndex, minor):
    """Matlab/gpu_multi_subsvd.m:13-68: L/minor subsets of dyadic size minor."""
    g = np.asarray(g)
    M, N = g.shape
    L = len(subIndex)
    if M != N:
        print("dimensions of input for matrix don't agree! ")
        return -1, -1, -1
    if not _dyadic(minor):
        print("length of minor is not dyadic!")
        return -1, -1, -1
    if L % minor != 0:
        print("minor must divide into the length of the index! ")
        return -1, -1, -1
    if minor > 128:
        print("minor of length > 128 not supported! ")
        return -1, -1, -1
    return _run(g, subIndex, minor)


def gpu_multi_subsvd2(g, subIndex, minor):
    """Matlab/gpu_multi_subsvd2.m:13-73: as gpu_multi_subsvd, and the index
    length must be divisible by 32."""
    g = np.asarray(g)
    M, N = g.shape
    L = len(subIndex)
    if M != N:
        print("dimensions of input for matrix don't agree! ")
        return -1, -1, -1
    if L % 32 != 0:
        print("length of subIndex must be divisible by 32! ")
        return -1, -1, -1
    if not _dyadic(minor):
        print("length of minor is not dyadic!")
        return -1, -1, -1
    if L % minor != 0:
        print("minor must divide into the length of the index! ")
        return -1, -1, -1
    if minor > 128:
        print("minor of length > 128 not supported! ")
        return -1, -1, -1
    return _run(g, subIndex, minor)


def gpu_svd(g):
    """Matlab/gpu_svd.m:13-57: all eigenvalues of a dyadic matrix <= 128."""
    g = np.asarray(g)
    M, N = g.shape
    if M != N:
        print("dimensions of input don't agree! ")
        return -1, -1, -1
    if not _dyadic(M):
        print("dimensions are not dyadic!")
        return -1, -1, -1
    if M > 128:
        print("Matrices larger than 128 x 128 are not supported!")
        return -1, -1, -1
    return _run(g, np.arange(M, dtype=np.int32), M)
