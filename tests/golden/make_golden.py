"""Writes tests/golden/golden.json: known answers computed in this repository
with numpy / scipy LAPACK (dsyevd, dstebz) -- nothing from the reference is
executed. Inputs are the reference's own example data:
  * the 8 x 8 SPD matrix of example1.cu:19-26 (= Matlab/for_example_3.m:3-10),
  * the index sets of example1.cu:39-42 / example2.cu:42-48 ({1,3,5,7}),
    example3.cu:43-49 ({0..7}) and how2use.m:33-style upper-left {0,1,2,3};
plus seeded random tridiagonal matrices for the Sturm/QL solvers and the
SHA-256 digests of the first 64 index sets of each workload seed.

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np
import scipy.linalg

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

EXAMPLE_G = [
    [11.6150, 12.6022, 2.7449, 0.6018, 5.7419, -0.8989, 6.2368, 1.4880],
    [12.6022, 32.5907, 6.8222, 3.8457, 3.8868, -1.3879, 7.2553, 1.3472],
    [2.7449, 6.8222, 11.3095, 1.9900, -3.8608, 0.9067, 3.4324, 3.3344],
    [0.6018, 3.8457, 1.9900, 4.7412, -0.0663, -1.2597, -0.2133, -1.6123],
    [5.7419, 3.8868, -3.8608, -0.0663, 18.0182, -1.7439, -1.5991, 1.2881],
    [-0.8989, -1.3879, 0.9067, -1.2597, -1.7439, 1.4386, 1.4156, 1.8073],
    [6.2368, 7.2553, 3.4324, -0.2133, -1.5991, 1.4156, 8.2406, 2.8208],
    [1.4880, 1.3472, 3.3344, -1.6123, 1.2881, 1.8073, 2.8208, 4.7334],
]


def main():
    import oracle

    G = np.array(EXAMPLE_G)
    out = {"example_matrix": EXAMPLE_G, "subsets": {}}
    for name, S in (("full_0_7", list(range(8))), ("s1357", [1, 3, 5, 7]), ("s0123", [0, 1, 2, 3]),
                    ("s4567", [4, 5, 6, 7])):
        A = G[np.ix_(S, S)]
        w = np.linalg.eigvalsh(A)
        A32 = A.astype(np.float32).astype(np.float64)
        w32 = np.linalg.eigvalsh(A32)
        out["subsets"][name] = {"index": S, "eig": w.tolist(), "det": float(np.linalg.det(A)),
                                "eig_fp32_input": w32.tolist()}
    rng = np.random.default_rng(20261203)
    tri = []
    for k in (1, 2, 3, 7, 16, 33, 64, 128, 256):
        a = rng.normal(1.0, 0.5, k)
        b = np.abs(rng.normal(0.3, 0.2, max(k - 1, 0)))
        w = scipy.linalg.eigvalsh_tridiagonal(a, b, lapack_driver="stebz") if k > 1 else a.copy()
        tri.append({"alpha": a.tolist(), "beta": [0.0] + b.tolist(), "eig": list(map(float, w))})
    out["tridiagonal"] = tri
    digests = {}
    for seed, k, n in ((0xC2, 8, 256), (0xC3, 32, 4096)):
        idx = oracle.gen_subsets_uniform(seed, 64, k, n)
        digests[f"{seed:#x}_k{k}_n{n}"] = {
            "sha256": hashlib.sha256(idx.astype("<i4").tobytes()).hexdigest(),
            "first": idx[:k].tolist()}
    klist = list(range(8, 129, 8))
    ks = oracle.gen_ks(0xC5, 64, klist)
    digests["ks_0xc5"] = {"sha256": hashlib.sha256(ks.astype("<i4").tobytes()).hexdigest(),
                          "first": ks[:16].tolist()}
    out["index_digests"] = digests
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", os.path.join(HERE, "golden.json"))


if __name__ == "__main__":
    main()
