#!/usr/bin/env python3
"""Benchmark of the batched principal-submatrix Lanczos eigenvalue path on one MI355X.

Workload (the plain batched GPU configuration): Phi is 512 x 4096 with
N(0, 1/512) FP32 entries and unit-norm columns (seed 0xC2); G = Phi^T Phi is
4096 x 4096 FP32 (64 MiB); 2^20 seeded subsets of k = 32 column indices
(seed 0xC3); 32-step Lanczos with full reorthogonalisation; one launch per step.
A step is one solve of all 2^20 subsets, inputs already resident in HBM.

Prints one JSON line: eig-solves/s, per-step p50/p99 latency, the plain
(one-wavefront-per-subset) form beside it, the relative error of a seeded
4096-subset sample of the timed outputs against FP64 LAPACK dsyevd, the HBM
roofline of the solve kernel, and the CPU oracle timed on the host cores.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "eig-solves/s + p99 latency, mixed k<=256, rel err<=1e-5 vs FP64, vs plain"
N, M, K, NSUB = 4096, 512, 32, 1 << 20
SEED_PHI, SEED_IDX, SEED_SAMPLE = 0xC2, 0xC3, 0xC4
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def bytes_per_solve(k):  # FP32 G_S square + int32 indices + 3 FP32 outputs
    return 4 * k * k + 4 * k + 12


def flops_per_solve(k):  # k matvecs + one full reorthogonalisation pass
    return 4 * k ** 3


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 accuracy sample")
    ap.add_argument("--cpu-sample", type=int, default=1 << 18)
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="first run one device copy of this many MiB (profiler counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); this path is "
              "single-GPU by design, see DESIGN.md for how it shards.", file=sys.stderr)
        return 2

    import torch

    from paralanczos_amd import solver

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.copy_mib:
        src = torch.ones(args.copy_mib << 18, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    phi = solver.gen_phi("gauss", SEED_PHI, M, N)
    G = solver.gram(phi)
    del phi
    idx = solver.gen_subsets(SEED_IDX, NSUB, K, N)
    out = solver.solve(G, idx, k=K, form="tuned")
    torch.cuda.synchronize()

    def run(form, o):
        solver.solve(G, idx, k=K, form=form, out=o)

    # warmup, then K timed steps bracketed by synchronize; per-step events on
    # the stream the kernel is launched on (torch's current stream)
    for _ in range(args.warmup):
        run("tuned", out)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    t0 = time.perf_counter()
    ev[0].record(stream)
    for i in range(args.steps):
        run("tuned", out)
        ev[i + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    step_ms = np.array([ev[i].elapsed_time(ev[i + 1]) for i in range(args.steps)])
    kernel_ms = float(ev[0].elapsed_time(ev[-1])) / args.steps
    ms_per_step = wall * 1e3 / args.steps
    value = NSUB / (ms_per_step / 1e3)

    # plain form: one 64-lane wavefront per subset, same launch count
    plain_out = solver.solve(G, idx, k=K, form="plain")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.plain_steps):
        run("plain", plain_out)
    torch.cuda.synchronize()
    plain_ms = (time.perf_counter() - t1) * 1e3 / args.plain_steps
    plain_value = NSUB / (plain_ms / 1e3)

    # gather stage alone (stage (a) + a checksum of each G_S, no Lanczos)
    solver.gather(G, idx, K, "checksum")
    torch.cuda.synchronize()
    g0 = torch.cuda.Event(enable_timing=True)
    g1 = torch.cuda.Event(enable_timing=True)
    g0.record(stream)
    for _ in range(3):
        solver.gather(G, idx, K, "checksum")
    g1.record(stream)
    torch.cuda.synchronize()
    gather_ms = g0.elapsed_time(g1) / 3

    # accuracy of the timed outputs on a seeded sample vs FP64 dsyevd
    lmin, lmax, delta, status = [t.cpu().numpy() for t in out]
    Gh = G.cpu().numpy()
    ih = idx.cpu().numpy().reshape(NSUB, K)
    rng = np.random.default_rng(SEED_SAMPLE)
    sample = np.sort(rng.choice(NSUB, 16 if args.no_check else 4096, replace=False))
    errs = []
    for s in sample:
        S = ih[s]
        w = np.linalg.eigvalsh(Gh[np.ix_(S, S)].astype(np.float64))
        errs.append(max(abs(lmin[s] - w[0]) / abs(w[0]), abs(lmax[s] - w[-1]) / abs(w[-1])))
    errs = np.array(errs)
    nonzero_status = int(np.count_nonzero(status))

    B = NSUB * bytes_per_solve(K)
    F = NSUB * flops_per_solve(K)
    achieved = B / (kernel_ms / 1e3) / 1e9
    traffic = None
    prof = os.path.join(ROOT, "profiles", "solve_k32_n4096_traffic.json")
    if os.path.exists(prof):
        with open(prof) as f:
            traffic = json.load(f).get("traffic_bytes_per_launch")

    cpu = None
    if not args.no_cpu:
        import oracle

        cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        ns = min(args.cpu_sample, NSUB)
        flat = np.ascontiguousarray(ih[:ns].reshape(-1))
        oracle.solve_batch(Gh, flat[: 64 * K], k=K, nthreads=cores)  # warm the pool
        t2 = time.perf_counter()
        oracle.solve_batch(Gh, flat, k=K, nthreads=cores)
        cpu_s = time.perf_counter() - t2
        cpu = {"value": ns / cpu_s, "unit": "eig-solves/s", "cores": cores, "kind": "port",
               "sample": f"first {ns} of the 2^20 k=32 subsets, FP64 oracle (full reorth + "
                         f"Sturm bisection), OpenMP over subsets, {cpu_s:.2f} s"}

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "eig-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "plain batched GPU: Phi 512x4096 N(0,1/512) unit columns, "
                               "G = Phi^T Phi 4096x4096 FP32, 2^20 subsets k=32, 32-step "
                               "Lanczos with full reorthogonalisation, one launch",
                   "n": N, "m": M, "k": K, "nsub": NSUB, "form": "tuned",
                   "parallelism": "single-gpu"},
        "p50_ms": float(np.percentile(step_ms, 50)),
        "p99_ms": float(np.percentile(step_ms, 99)),
        "plain": {"value": plain_value, "ms_per_step": plain_ms},
        "vs_plain": value / plain_value,
        "accuracy": {"sample": len(sample), "max_rel_err": float(errs.max()),
                     "median_rel_err": float(np.median(errs)), "vs": "FP64 dsyevd",
                     "nonzero_status": nonzero_status},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "roofline_fp32": {"achieved": F / (kernel_ms / 1e3) / 1e12, "peak": FP32_PEAK_TFLOPS,
                          "unit": "TFLOP/s",
                          "frac": F / (kernel_ms / 1e3) / 1e12 / FP32_PEAK_TFLOPS},
        "kernel_ms": kernel_ms,
        "gather": {"ms": gather_ms,
                   "algorithmic_GBs": NSUB * (4 * K * K + 4 * K + 4) / (gather_ms / 1e3) / 1e9},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
