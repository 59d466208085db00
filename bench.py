#!/usr/bin/env python3
"""Benchmark of the batched principal-submatrix Lanczos eigenvalue path on one MI355X.

Workload (the serving configuration, mixed k <= 256): Phi is 4096 seeded rows
(seed 0xC5) of the orthonormal 32768-point DCT-II scaled by sqrt(32768/4096);
G = Phi^T Phi is 32768 x 32768 FP32 (4 GiB), resident in HBM. A step serves
one wave of R requests (default 64) through the continuous-batching engine:
request r holds floor(exp(U(0, ln 4096))) subsets (numpy seed 0xC5+3), each
of size k drawn uniformly from {16, 32, 64, 128, 256} and index sets drawn
without replacement (Philox streams 0xC5+1), all generated on the device
before timing. The engine is driven closed-loop by one client per request:
each client resubmits its request as soon as the previous submission
completes, so the queue never drains; a step is one round of all R clients
(one wave's worth of work), and the value is subsets solved per second over
K steps.

Reported beside it, in the same JSON line:
  * latency: p50/p99 of the closed-loop requests, and of Poisson arrivals at
    fractions of the measured saturation rate (0.3 .. 1.0);
  * served_waves: the same work as waves that drain between them (a wave is
    submitted after the previous one completed);
  * plain: the same subsets solved as one CSR batch by the plain form, one
    launch per distinct k, FP32 throughout, one 64-lane wavefront per subset
    (PL_FORM_PLAIN_WAVE: beyond k = 64 the wavefront streams G_S and its
    basis from a global scratch); vs_plain is the served rate over it.
    plain_multiwave / vs_plain_multiwave: the same against the plain form
    that runs k > 64 on the multi-wavefront kernels (PL_FORM_PLAIN). And
    the tuned batch without the serving engine;
  * served_f32: the same waves served with FP32 Gram-Schmidt coefficients at
    every k (the plain forms' arithmetic), against PL_FORM_PLAIN;
  * accuracy: relative error of lambda_min / lambda_max of a 4096-subset
    sample of the served outputs against FP64 LAPACK dsyevd;
  * roofline of the dominant kernel (the k = 256 gather + Lanczos kernel,
    FP64 Gram-Schmidt coefficients), timed with HIP events on its stream,
    against the FP32 vector peak, HBM, and (roofline_lds) the LDS bytes the
    profiler counts for the same launch against the LDS peak;
  * the gather stage alone over the same subsets: algorithmic GB/s and the
    HBM fetch rate the profiler's counters give for it (profiles/);
  * gram: G = Phi^T Phi of the serving frame by pl_gram (matrix cores) against
    one FP32 library GEMM of the full product (set-up, outside the timing);
  * cpu_baseline: the FP64 oracle on the GPU host's cores over a bounded
    sample of the same subsets, and (c1) over the whole k = 8 configuration
    on one thread and on all cores;
  * c1, c2: the uniform-k configurations (k = 8 on a 256^2 G, k = 32 on a
    4096^2 G over 2^20 subsets), tuned vs plain and accuracy vs dsyevd;
  * batching: the mixed k in {8..128} configuration (tuned binned persistent
    form eager / from a graph / after autotuning, vs the plain form);
  * precision: k = 128 and 256 with FP32 vs FP64 Gram-Schmidt coefficients,
    throughput and error vs FP64 dsyevd.

    python bench.py [--gpus 1] [--steps 5] [--warmup 2] [--requests 64] [--no-cpu]
"""
from __future__ import annotations

import argparse
import gc
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "eig-solves/s + p99 latency, mixed k<=256, rel err<=1e-5 vs FP64, vs plain"
N, M = 32768, 4096
KLIST = [16, 32, 64, 128, 256]
SEED_PHI, SEED_SUB, SEED_NSUB, SEED_SAMPLE = 0xC5, 0xC5 + 1, 0xC5 + 3, 0xC5 + 2
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
LDS_PEAK_TBS = 256 * 128 * 2.4e9 / 1e12  # 256 CUs x 128 bytes per clock x 2.4 GHz = 78.6
DOMINANT_K = 256


def bytes_per_solve(k):  # FP32 G_S square + int32 indices + 3 FP32 outputs
    return 4 * k * k + 4 * k + 12


def flops_per_solve(k):  # k matvecs + one full reorthogonalisation pass
    return 4 * k ** 3


def make_requests(solver, torch, R):
    rng = np.random.default_rng(SEED_NSUB)
    reqs = []
    for r in range(R):
        nsub = max(1, int(math.exp(rng.uniform(0.0, math.log(4096.0)))))
        ks = solver.gen_ks(SEED_SUB, nsub, KLIST, s0=r * 4096)
        off = solver.offsets_from_ks(ks)
        idx = solver.gen_subsets_csr(SEED_SUB, off, N, s0=r * 4096)
        reqs.append({"idx": idx, "off": off, "ks": ks, "nsub": nsub})
    torch.cuda.synchronize()
    return reqs


def serve_wave(srv, reqs):
    tickets = [srv.submit(q["idx"], offsets=q["off"], kmax=max(KLIST)) for q in reqs]
    return [srv.wait(t) for t in tickets]


def serve_closed_loop(srv, reqs, rounds):
    """Closed loop with one client per request of the step: each client
    resubmits its request as soon as the previous submission completes, so
    the engine's queue never drains between waves; `rounds` submissions per
    client. Returns the outputs of each client's last round."""
    kmax = max(KLIST)
    left = [rounds] * len(reqs)
    live = {}
    for c, q in enumerate(reqs):
        live[srv.submit(q["idx"], offsets=q["off"], kmax=kmax)] = c
        left[c] -= 1
    last = [None] * len(reqs)
    while live:
        for t in list(live):
            if not srv.poll(t):
                continue
            c = live.pop(t)
            out, _, _ = srv.wait(t)
            if left[c] > 0:
                left[c] -= 1
                q = reqs[c]
                live[srv.submit(q["idx"], offsets=q["off"], kmax=kmax)] = c
            else:
                last[c] = (out,)
    return last


def poisson_run(srv, reqs, rate, seed, passes=5):
    """Submit `passes` rounds of the requests at Poisson arrival times of the
    given rate (requests/s); returns (subsets/s over the run, p50, p99 latency
    in s)."""
    rng = np.random.default_rng(seed)
    arrivals = reqs * passes
    gaps = rng.exponential(1.0 / rate, size=len(arrivals))
    srv.stats(reset=True)
    t0 = time.perf_counter()
    due = t0
    tickets = []
    for q, g in zip(arrivals, gaps):
        due += g
        while time.perf_counter() < due:
            pass
        tickets.append(srv.submit(q["idx"], offsets=q["off"], kmax=max(KLIST)))
    for t in tickets:
        srv.wait(t)
    st = srv.stats()
    return st["subsets"] / st["span_s"], st["p50_s"], st["p99_s"]


def batching_leg(solver, torch, autotune=True):
    """Mixed sizes k in {8, 16, ..., 128}: Phi 1024 x 8192 N(0,1/1024) with
    unit columns (seed 0xC3), 2^20 subsets (sizes and index sets from streams
    0xC3+1). Tuned binned persistent form (eager, and replayed from a graph),
    the plain one-launch-per-k forms (one wavefront per subset; and with the
    multi-wavefront kernels beyond k = 64), and the tuned form after
    autotuning."""
    n, nsub = 8192, 1 << 20
    G = solver.gram(solver.gen_phi("gauss", 0xC3, 1024, n))
    ks = solver.gen_ks(0xC3 + 1, nsub, list(range(8, 129, 8)))
    off = solver.offsets_from_ks(ks)
    idx = solver.gen_subsets_csr(0xC3 + 1, off, n)
    torch.cuda.synchronize()

    def ms(fn, reps=2):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / reps

    out = solver.solve(G, idx, k=128, offsets=off)
    res = {"workload": "Phi 1024x8192 N(0,1/1024) unit columns, G 8192^2 FP32, 2^20 subsets, "
                       "k uniform in {8,16,...,128}", "subsets": nsub}
    res["plain_ms"] = ms(lambda: solver.solve(G, idx, k=128, offsets=off, form="plain_wave",
                                              out=out), reps=1)
    res["plain_multiwave_ms"] = ms(lambda: solver.solve(G, idx, k=128, offsets=off, form="plain",
                                                        out=out))
    res["tuned_ms"] = ms(lambda: solver.solve(G, idx, k=128, offsets=off, out=out))
    plan = solver.Plan(G, idx, k=128, offsets=off)
    res["tuned_graph_ms"] = ms(plan.execute)
    res["graph"] = plan.is_graph
    plan.close()
    if autotune:
        t0 = time.perf_counter()
        tune = solver.autotune(G, list(range(8, 129, 8)), sample=32768)
        res["autotune_s"] = time.perf_counter() - t0
        res["autotuned"] = {k: list(v[:2]) for k, v in tune.items()}
        res["autotuned_ms"] = ms(lambda: solver.solve(G, idx, k=128, offsets=off, out=out))
        solver.tuning_reset()
        # bin edges tuned too: every variant that takes a size competes at it
        t0 = time.perf_counter()
        win, table = solver.autotune_edges(G, list(range(8, 129, 8)), sample=32768)
        res["autotune_edges_s"] = time.perf_counter() - t0
        fam = lambda lanes: "lane groups" if lanes < 128 else f"{lanes} lanes"
        res["edges_sweep_ms"] = {k: {f"{l}/{p}": round(t, 4) for (l, p), t in sorted(v.items())}
                                 for k, v in table.items()}
        res["edges_winners"] = {k: list(v[:2]) for k, v in win.items()}
        ks_sorted = sorted(win)
        res["edges"] = [{"from_k": k, "family": fam(win[k][0])} for i, k in enumerate(ks_sorted)
                        if i == 0 or fam(win[k][0]) != fam(win[ks_sorted[i - 1]][0])]
        res["autotuned_edges_ms"] = ms(lambda: solver.solve(G, idx, k=128, offsets=off, out=out))
        solver.tuning_reset()
    best = min(res["tuned_ms"], res["tuned_graph_ms"], res.get("autotuned_ms", 1e30),
               res.get("autotuned_edges_ms", 1e30))
    res["solves_s_tuned_best"] = nsub / (best / 1e3)
    res["solves_s_plain"] = nsub / (res["plain_ms"] / 1e3)
    res["vs_plain"] = res["plain_ms"] / best
    res["vs_plain_multiwave"] = res["plain_multiwave_ms"] / best
    del G
    torch.cuda.empty_cache()
    return res


def small_configs_leg(solver, torch, sample=4096):
    """The two uniform-k configurations. C1: Phi 64 x 256 N(0,1/64) with unit
    columns (seed 0xC1), 1024 subsets of k = 8 (stream 0xC1+1). C2: Phi
    512 x 4096 N(0,1/512) with unit columns (seed 0xC2), 2^20 subsets of
    k = 32 (stream 0xC2+1). Tuned and plain forms (at k <= 64 both plain forms
    are the same one-wavefront kernel) timed with HIP events on
    torch's stream, accuracy of lambda_min / lambda_max against FP64 dsyevd on
    every C1 subset and on a 4096-subset C2 sample (stream 0xC2+2 positions)."""
    res = {}
    for name, seed, m, n, k, nsub, reps in (("c1", 0xC1, 64, 256, 8, 1024, 50),
                                            ("c2", 0xC2, 512, 4096, 32, 1 << 20, 5)):
        G = solver.gram(solver.gen_phi("gauss", seed, m, n))
        idx = solver.gen_subsets(seed + 1, nsub, k, n)
        out = solver.solve(G, idx, k=k)
        torch.cuda.synchronize()

        def ev_ms(fn):
            fn()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                fn()
            e1.record()
            torch.cuda.synchronize()
            return e0.elapsed_time(e1) / reps

        r = {"workload": f"Phi {m}x{n} N(0,1/{m}) unit columns, G {n}^2 FP32, {nsub} subsets of k={k}",
             "subsets": nsub, "k": k}
        r["tuned_ms"] = ev_ms(lambda: solver.solve(G, idx, k=k, out=out))
        r["plain_ms"] = ev_ms(lambda: solver.solve(G, idx, k=k, form="plain", out=out))
        r["solves_s_tuned"] = nsub / (r["tuned_ms"] / 1e3)
        r["solves_s_plain"] = nsub / (r["plain_ms"] / 1e3)
        r["vs_plain"] = r["plain_ms"] / r["tuned_ms"]
        out = solver.solve(G, idx, k=k)
        lmin, lmax, st = out[0].cpu().numpy(), out[1].cpu().numpy(), out[3].cpu().numpy()
        rng = np.random.default_rng(seed + 2)
        sel = np.sort(rng.choice(nsub, min(nsub, sample), replace=False))
        A = solver.gather(G, idx.reshape(nsub, k)[torch.from_numpy(sel).to(idx.device)].reshape(-1),
                          k, "dump", 0, 4).cpu().numpy().reshape(-1, k, k)
        w = eigvalsh64(A)
        e = np.maximum(np.abs(lmin[sel] - w[:, 0]) / np.abs(w[:, 0]),
                       np.abs(lmax[sel] - w[:, -1]) / np.abs(w[:, -1]))
        r["accuracy"] = {"sample": int(len(sel)), "max_rel_err": float(e.max()),
                         "median_rel_err": float(np.median(e)), "vs": "FP64 dsyevd",
                         "nonzero_status": int(np.count_nonzero(st))}
        res[name] = r
        del G, idx, out
        torch.cuda.empty_cache()
    return res


def eigvalsh64(A):
    """FP64 LAPACK eigenvalues (dsyevd through numpy) of a stack of symmetric
    matrices, ascending per matrix: chunks over a thread pool with one BLAS
    thread each (numpy's LAPACK loops release the GIL; the per-matrix results
    do not depend on the split)."""
    from concurrent.futures import ThreadPoolExecutor
    from threadpoolctl import threadpool_limits
    A = np.asarray(A, dtype=np.float64)
    nw = max(1, min(16, os.cpu_count() or 1))
    chunks = [c for c in np.array_split(np.arange(len(A)), nw * 4) if len(c)]
    with threadpool_limits(1), ThreadPoolExecutor(nw) as ex:
        return np.concatenate(list(ex.map(lambda c: np.linalg.eigvalsh(A[c]), chunks)))


def precision_leg(solver, torch, dev, sample=4096):
    """k = 128 and 256 on Phi = Bernoulli +-1/sqrt(2048), 2048 x 16384 (seed
    0xC4; G 16384^2 FP32, 1 GiB): FP32 vs FP64 Gram-Schmidt coefficients,
    throughput over 262,144 subsets of each k and the max and median relative
    error of lambda_min/lambda_max against FP64 dsyevd of the gathered
    submatrices of the first 4096 (seeded) subsets."""
    n = 16384
    G = solver.gram(solver.gen_phi("bernoulli", 0xC4, 2048, n))
    res = {"workload": "Phi Bernoulli +-1/sqrt(2048) 2048x16384, G 16384^2 FP32 (1 GiB)"}
    for k, nsub in ((128, 1 << 18), (256, 1 << 18)):
        idx = solver.gen_subsets(0xC4 + 1, nsub, k, n)
        A = solver.gather(G, idx[:sample * k], k, "dump", 512 if k > 128 else 128, 4)
        w = eigvalsh64(A.cpu().numpy().reshape(-1, k, k))
        del A
        for form in ("f32", "f64"):
            out = solver.solve(G, idx, k=k, reorth=form)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(2):
                solver.solve(G, idx, k=k, reorth=form, out=out)
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t0) / 2
            lo, hi = out[0][:sample].cpu().numpy(), out[1][:sample].cpu().numpy()
            e = np.maximum(np.abs(lo - w[:, 0]) / np.abs(w[:, 0]),
                           np.abs(hi - w[:, -1]) / np.abs(w[:, -1]))
            res[f"k{k}_{form}"] = {"solves_s": nsub / dt, "ms": dt * 1e3, "subsets": nsub,
                                   "max_rel_err": float(e.max()),
                                   "median_rel_err": float(np.median(e)), "sample": sample,
                                   "nonzero_status": int(torch.count_nonzero(out[3]).item())}
    del G
    torch.cuda.empty_cache()
    return res


def gram_leg(solver, torch, phi, G):
    """G = Phi^T Phi of the serving frame (m = 4096, n = 32768, column layout):
    pl_gram (matrix cores, upper-triangle tiles mirrored) against one FP32
    library GEMM of the full product, HIP events on torch's stream."""
    n, m = phi.shape
    tiles = -(-n // 128)
    flops = tiles * (tiles + 1) // 2 * 2.0 * 128 * 128 * m  # computed tiles
    ev = lambda: torch.cuda.Event(enable_timing=True)

    def timed(f, reps):
        f()
        torch.cuda.synchronize()
        e0, e1 = ev(), ev()
        e0.record()
        for _ in range(reps):
            out = f()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps, out

    ms, _ = timed(lambda: solver.gram(phi), 3)
    prev = torch.backends.cuda.matmul.allow_tf32
    torch.backends.cuda.matmul.allow_tf32 = False
    try:
        lib_ms, Gl = timed(lambda: phi @ phi.T, 2)
    finally:
        torch.backends.cuda.matmul.allow_tf32 = prev
    diff = (Gl - G).abs().max().item()
    del Gl
    torch.cuda.empty_cache()
    return {"workload": f"G = Phi^T Phi, Phi {m} x {n} FP32 (column layout), G {n} x {n}",
            "ms": ms, "tflops": flops / (ms / 1e3) / 1e12, "peak_tflops": FP32_PEAK_TFLOPS,
            "frac": flops / (ms / 1e3) / 1e12 / FP32_PEAK_TFLOPS, "flops_per_launch": flops,
            "library_gemm_ms": lib_ms, "vs_library": lib_ms / ms, "max_abs_diff_vs_library": diff}


def host_buffers_leg(solver, torch, G, idx_all, off_all, tot, reps=3):
    """Rates with the inputs and outputs in host memory (PCIe included).
    main: the step's CSR batch with indices, offsets and the four outputs in
    pinned host memory; each call uploads the indices and offsets, solves the
    tuned form on the resident G and downloads the outputs, on torch's stream.
    legacy_c1 / legacy_c2: pl_multi_subsvd (the reference's host entry point)
    on the uniform configurations, FP64 G and int32 indices in host memory,
    all eigenvalues back; each call converts G to FP32, allocates, uploads,
    solves and downloads."""
    from paralanczos_amd import legacy

    res = {}
    idx_h = torch.empty(idx_all.shape, dtype=idx_all.dtype, pin_memory=True)
    off_h = torch.empty(off_all.shape, dtype=off_all.dtype, pin_memory=True)
    idx_h.copy_(idx_all)
    off_h.copy_(off_all)
    idx_d, off_d = torch.empty_like(idx_all), torch.empty_like(off_all)
    o = solver.solve(G, idx_all, k=256, offsets=off_all)
    oh = [torch.empty(t.shape, dtype=t.dtype, pin_memory=True) for t in o]

    def call():
        idx_d.copy_(idx_h, non_blocking=True)
        off_d.copy_(off_h, non_blocking=True)
        solver.solve(G, idx_d, k=256, offsets=off_d, out=o)
        for a, b in zip(oh, o):
            a.copy_(b, non_blocking=True)

    call()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        call()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        solver.solve(G, idx_all, k=256, offsets=off_all, out=o)
    torch.cuda.synchronize()
    ms_dev = (time.perf_counter() - t0) * 1e3 / reps
    moved = idx_h.numel() * 4 + off_h.numel() * 8 + sum(t.numel() * 4 for t in oh)
    res["main"] = {"value": tot / (ms / 1e3), "ms": ms, "device_resident_ms": ms_dev,
                   "bytes_moved": moved, "subsets": tot,
                   "workload": "the step's CSR batch, pinned host indices/offsets/outputs"}
    del idx_h, off_h, idx_d, off_d, o, oh
    for name, seed, m, n, k, nsub in (("legacy_c1", 0xC1, 64, 256, 8, 1024),
                                      ("legacy_c2", 0xC2, 512, 4096, 32, 1 << 20)):
        g = solver.gram(solver.gen_phi("gauss", seed, m, n)).double().cpu().numpy()
        ix = solver.gen_subsets(seed + 1, nsub, k, n).cpu().numpy()
        eigs, _, err = legacy.gpu_multi_subsvd(g, ix, k)  # warm
        n_rep = 20 if nsub < 4096 else 3
        t0 = time.perf_counter()
        for _ in range(n_rep):
            legacy.gpu_multi_subsvd(g, ix, k)
        ms = (time.perf_counter() - t0) * 1e3 / n_rep
        res[name] = {"value": nsub / (ms / 1e3), "ms": ms, "subsets": nsub, "k": k,
                     "err_nonzero": int(np.count_nonzero(err)),
                     "workload": f"pl_multi_subsvd, FP64 G {n}^2 and {nsub} subsets of k={k} "
                                 f"in host memory, all eigenvalues returned"}
    torch.cuda.empty_cache()
    return res


def kernel_times(solver, torch, dev):
    """Stage 1 of the k = 128 and k = 256 kernels on the serving G: 6200
    subsets each (one served wave's worth), FP64 and FP32 Gram-Schmidt
    coefficients, 3 timed launches after a warm one (HIP events on the
    launching stream); microseconds per subset and CU; the eigenvalues of the
    first 48 subsets of each against the FP64 oracle."""
    import oracle

    G = solver.gram(solver.gen_phi("dct", SEED_PHI, M, N))
    cus = torch.cuda.get_device_properties(dev).multi_processor_count
    stream = torch.cuda.current_stream()
    out = {"workload": "serving G (DCT-II 4096 x 32768), 6200 seeded subsets per k", "cus": cus}
    Gh = None
    for k in (128, 256):
        nsub = 6200
        idx = solver.gen_subsets(SEED_SAMPLE + k, nsub, k, N)
        ws = torch.empty(solver.workspace_size(nsub, k), dtype=torch.uint8, device=dev)
        for reorth in ("f64", "f32"):
            o = solver.solve_stage(G, idx, 1, ws, k=k, reorth=reorth)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(3):
                solver.solve_stage(G, idx, 1, ws, k=k, reorth=reorth, out=o)
            e1.record(stream)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / 3
            lo, hi, _, st = solver.solve(G, idx[:48 * k].contiguous(), k=k, reorth=reorth)
            if Gh is None:
                Gh = G.cpu().numpy()
            olo, ohi, _, ost = oracle.solve_batch(Gh, idx[:48 * k].cpu().numpy(), k=k,
                                                  nthreads=16)
            err = max(float(np.max(np.abs(lo.cpu().numpy() - olo) / olo)),
                      float(np.max(np.abs(hi.cpu().numpy() - ohi) / ohi)))
            out[f"k{k}_{reorth}"] = {
                "ms": ms, "us_per_subset_cu": ms * 1e3 * cus / nsub,
                "tflops": nsub * flops_per_solve(k) / (ms / 1e3) / 1e12,
                "frac": nsub * flops_per_solve(k) / (ms / 1e3) / 1e12 / FP32_PEAK_TFLOPS,
                "max_rel_err_48": err, "nonzero_status": int(np.count_nonzero(st.cpu().numpy()))}
            print(json.dumps({f"k{k}_{reorth}": out[f"k{k}_{reorth}"]}), flush=True)
    print(json.dumps(out), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--requests", type=int, default=64, help="requests per step")
    ap.add_argument("--rates", default="0.3,0.5,0.7,0.85,1.0",
                    help="Poisson arrival rates as fractions of the saturation rate")
    ap.add_argument("--passes", type=int, default=5,
                    help="rounds of the step's requests per Poisson rate")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 accuracy sample")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="bound on the CPU baseline's work")
    ap.add_argument("--profile-dominant", action="store_true",
                    help="only run the dominant kernel's launches (for rocprof counter runs)")
    ap.add_argument("--no-legs", action="store_true",
                    help="skip the batching and precision configurations")
    ap.add_argument("--profile-k128", action="store_true",
                    help="only run stage 1 over the step's k = 128 subsets (the 256-lane kernel, "
                         "FP64 coefficients) 1 + 3 times, for rocprof")
    ap.add_argument("--profile-c2", action="store_true",
                    help="only run stage 1 and the gather alone of the k = 32 configuration "
                         "(2^20 subsets of a 4096^2 G), for rocprof")
    ap.add_argument("--profile-gather", action="store_true",
                    help="only run the gather stage over the dominant subsets (rocprof counters)")
    ap.add_argument("--profile-gram", action="store_true",
                    help="run pl_gram on the serving frame 1 + 3 times and exit (for rocprofv3)")
    ap.add_argument("--kernel-times", action="store_true",
                    help="time stage 1 (gather + Lanczos) of the k = 128 and k = 256 kernels in "
                         "both coefficient precisions on the serving G, check a sample against "
                         "the oracle, print one JSON line and exit")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="only copy this many MiB device to device and exit (calibrates the "
                         "profiler's byte counters against a transfer of known size)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): the path runs "
              "on a single GPU (subsets are independent: a sharded run would replicate G "
              "and deal subsets out to the GPUs).", file=sys.stderr)
        return 2

    import torch

    from paralanczos_amd import solver

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.copy_mib:
        src = torch.ones(args.copy_mib << 18, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": src.numel() * 4}), flush=True)
        return 0
    if args.profile_c2:
        # the k = 32 configuration: Phi 512 x 4096 (seed 0xC2), 2^20 subsets;
        # stage 1 (gather + Lanczos, 32-lane groups) and the gather alone,
        # 1 + 3 launches each
        G = solver.gram(solver.gen_phi("gauss", 0xC2, 512, 4096))
        idx = solver.gen_subsets(0xC2 + 1, 1 << 20, 32, 4096)
        ws = torch.empty(solver.workspace_size(1 << 20, 32), dtype=torch.uint8, device=dev)
        o = solver.solve_stage(G, idx, 1, ws, k=32)
        for _ in range(3):
            solver.solve_stage(G, idx, 1, ws, k=32, out=o)
        for _ in range(4):
            solver.gather(G, idx, 32, "checksum", 32, 4)
        torch.cuda.synchronize()
        print(json.dumps({"profile_c2": True, "subsets": 1 << 20, "k": 32}), flush=True)
        return 0
    if args.kernel_times:
        return kernel_times(solver, torch, dev)
    t_setup = time.perf_counter()
    phi = solver.gen_phi("dct", SEED_PHI, M, N)
    if args.profile_gram:
        G = solver.gram(phi)
        torch.cuda.synchronize()
        for _ in range(3):
            solver.gram(phi)
        torch.cuda.synchronize()
        print(json.dumps({"profile_gram": True, "m": M, "n": N}), flush=True)
        return 0
    G = solver.gram(phi)
    profiling = args.profile_dominant or args.profile_gather or args.profile_k128
    gram = None if profiling else gram_leg(solver, torch, phi, G)
    torch.cuda.empty_cache()
    reqs = make_requests(solver, torch, args.requests)
    tot = sum(q["nsub"] for q in reqs)
    stream = torch.cuda.current_stream()

    if args.profile_k128:
        # the second kernel of the mix: stage 1 over as many k = 128 subsets
        # as one step holds
        n128 = int(sum(int((q["ks"] == 128).sum().item()) for q in reqs))
        idx128 = solver.gen_subsets(SEED_SAMPLE + 1, n128, 128, N)
        ws = torch.empty(solver.workspace_size(n128, 128), dtype=torch.uint8, device=dev)
        o128 = solver.solve_stage(G, idx128, 1, ws, k=128)
        for _ in range(3):
            solver.solve_stage(G, idx128, 1, ws, k=128, out=o128)
        torch.cuda.synchronize()
        print(json.dumps({"profile_k128": True, "subsets": n128}), flush=True)
        return 0

    # the dominant kernel alone: stage 1 (gather + Lanczos) over as many
    # k = 256 subsets as one step holds, on torch's current stream
    n256 = int(sum(int((q["ks"] == DOMINANT_K).sum().item()) for q in reqs))
    idx256 = solver.gen_subsets(SEED_SAMPLE, n256, DOMINANT_K, N)
    if gram is not None:
        # the same subsets' G_S straight from the frame (pl_subgram), against
        # gathering them from the resident G
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        sg = solver.subgram(phi, idx256, DOMINANT_K)
        del sg  # its blocks are reused from torch's cache by the timed call
        torch.cuda.synchronize()
        e0.record()
        sg = solver.subgram(phi, idx256, DOMINANT_K)
        e1.record()
        torch.cuda.synchronize()
        gram["subgram_k256_ms"] = e0.elapsed_time(e1)
        gram["subgram_subsets"] = n256
        del sg
    del phi
    torch.cuda.empty_cache()
    ws = torch.empty(solver.workspace_size(n256, DOMINANT_K), dtype=torch.uint8, device=dev)
    o256 = solver.solve_stage(G, idx256, 1, ws, k=DOMINANT_K)
    torch.cuda.synchronize()
    if args.profile_dominant:
        for _ in range(3):
            solver.solve_stage(G, idx256, 1, ws, k=DOMINANT_K, out=o256)
        torch.cuda.synchronize()
        print(json.dumps({"profile_dominant": True, "subsets": n256}), flush=True)
        return 0
    # stage (a) alone over the same subsets: the gather-and-checksum launch
    gsum = solver.gather(G, idx256, DOMINANT_K, "checksum", 512, 4)
    torch.cuda.synchronize()
    if args.profile_gather:
        for _ in range(3):
            solver.gather(G, idx256, DOMINANT_K, "checksum", 512, 4)
        torch.cuda.synchronize()
        print(json.dumps({"profile_gather": True, "subsets": n256}), flush=True)
        return 0
    g0, g1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    g0.record(stream)
    for _ in range(3):
        solver.gather(G, idx256, DOMINANT_K, "checksum", 512, 4)
    g1.record(stream)
    torch.cuda.synchronize()
    gather_ms = g0.elapsed_time(g1) / 3
    del gsum
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(3):
        solver.solve_stage(G, idx256, 1, ws, k=DOMINANT_K, out=o256)
    e1.record(stream)
    torch.cuda.synchronize()
    dom_ms = e0.elapsed_time(e1) / 3
    del ws
    # the second kernel of the mix: stage 1 over as many k = 128 subsets as
    # the step holds (the 256-lane kernel, FP64 coefficients)
    n128 = int(sum(int((q["ks"] == 128).sum().item()) for q in reqs))
    idx128 = solver.gen_subsets(SEED_SAMPLE + 1, n128, 128, N)
    ws = torch.empty(solver.workspace_size(n128, 128), dtype=torch.uint8, device=dev)
    o128 = solver.solve_stage(G, idx128, 1, ws, k=128)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(3):
        solver.solve_stage(G, idx128, 1, ws, k=128, out=o128)
    e1.record(stream)
    torch.cuda.synchronize()
    k128_ms = e0.elapsed_time(e1) / 3
    del ws, idx128, o128

    srv = solver.Server(G, max_subsets=16384, max_indices=16384 * 256, max_requests=1024,
                        max_k=256, depth=3)
    for _ in range(args.warmup):
        serve_wave(srv, reqs)
    torch.cuda.synchronize()
    # the same work as waves that drain between them (each wave submitted
    # after the previous one completed): the engine's rate without a queue
    t0 = time.perf_counter()
    for _ in range(args.steps):
        serve_wave(srv, reqs)
    torch.cuda.synchronize()
    waves_wall = time.perf_counter() - t0
    srv.stats(reset=True)
    gc.collect()
    gc.disable()  # no collector pauses in the host loop that submits and waits
    # timed: `steps` rounds of the closed loop (each of the step's 64 clients
    # resubmits as soon as its request completes)
    t0 = time.perf_counter()
    res = serve_closed_loop(srv, reqs, args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gc.enable()
    closed = srv.stats()
    ms_per_step = wall * 1e3 / args.steps
    value = tot * args.steps / wall
    sat_rate = len(reqs) * args.steps / wall  # requests/s at saturation

    latency = []
    for i, frac in enumerate(float(x) for x in args.rates.split(",") if x):
        thr, p50, p99 = poisson_run(srv, reqs, frac * sat_rate, 100 + i, args.passes)
        latency.append({"rate_frac": frac, "rate_req_s": frac * sat_rate,
                        "requests": len(reqs) * args.passes,
                        "solves_s": thr, "p50_ms": p50 * 1e3, "p99_ms": p99 * 1e3})
    srv.close()

    # the same waves served with FP32 Gram-Schmidt coefficients at every k: the
    # arithmetic of the multi-wavefront plain form, to compare the engine and
    # kernels with it like for like
    srv = solver.Server(G, max_subsets=16384, max_indices=16384 * 256, max_requests=1024,
                        max_k=256, depth=3, reorth="f32")
    serve_wave(srv, reqs)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    serve_closed_loop(srv, reqs, args.steps)
    torch.cuda.synchronize()
    f32_wall = time.perf_counter() - t1
    srv.close()
    served_f32 = {"value": tot * args.steps / f32_wall, "ms_per_step": f32_wall * 1e3 / args.steps}

    # the same subsets as one CSR batch: plain form, and tuned without the engine
    offs = torch.cat([q["off"][1:] - q["off"][:-1] for q in reqs])
    off_all = solver.offsets_from_ks(offs)
    idx_all = torch.cat([q["idx"] for q in reqs])
    # the plain forms through plans: their workspace (and the one-wavefront
    # form's scratch) allocated once, outside the timed calls
    batch = {}
    for form in ("plain_wave", "plain", "tuned", "tuned_plan"):
        if form == "tuned":
            o = solver.solve(G, idx_all, k=256, offsets=off_all, form=form)
            run = lambda: solver.solve(G, idx_all, k=256, offsets=off_all, form="tuned", out=o)
        else:
            plan = solver.Plan(G, idx_all, k=256, offsets=off_all,
                               form="tuned" if form == "tuned_plan" else form)
            o = plan.execute()
            run = plan.execute
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(2):
            run()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t1) * 1e3 / 2
        batch[form] = {"value": tot / (ms / 1e3), "ms": ms}
        if form == "plain_wave":  # the plain form's outputs, for its own accuracy
            plain_out = [t.cpu().numpy() for t in (o[0], o[1], o[3])]
        if form != "tuned":
            plan.close()

    # accuracy of the served outputs (last timed step) on a 4096-subset sample
    lmin = torch.cat([r[0][0] for r in res]).cpu().numpy()
    lmax = torch.cat([r[0][1] for r in res]).cpu().numpy()
    status = torch.cat([r[0][3] for r in res]).cpu().numpy()
    ks_all = offs.cpu().numpy()
    rng = np.random.default_rng(SEED_SAMPLE)
    sample = np.sort(rng.choice(tot, min(tot, 256 if args.no_check else 4096), replace=False))
    off_h = off_all.cpu().numpy()
    idx_h = idx_all.cpu().numpy()
    errs, errs_plain = [], []
    for k in KLIST:
        sel = sample[ks_all[sample] == k]
        if len(sel) == 0:
            continue
        sub = np.concatenate([idx_h[off_h[s]:off_h[s + 1]] for s in sel]).astype(np.int32)
        A = solver.gather(G, torch.from_numpy(sub).to(dev), k, "dump", 512 if k > 128 else
                          (128 if k > 64 else 64), 4).cpu().numpy().reshape(-1, k, k)
        w = eigvalsh64(A)
        e = np.maximum(np.abs(lmin[sel] - w[:, 0]) / np.abs(w[:, 0]),
                       np.abs(lmax[sel] - w[:, -1]) / np.abs(w[:, -1]))
        errs.append(e)
        errs_plain.append(np.maximum(np.abs(plain_out[0][sel] - w[:, 0]) / np.abs(w[:, 0]),
                                     np.abs(plain_out[1][sel] - w[:, -1]) / np.abs(w[:, -1])))
    errs = np.concatenate(errs)
    errs_plain = np.concatenate(errs_plain)

    # roofline of the dominant kernel (algorithmic work per launch / duration)
    B = n256 * bytes_per_solve(DOMINANT_K)
    F = n256 * flops_per_solve(DOMINANT_K)
    tflops = F / (dom_ms / 1e3) / 1e12
    gbs = B / (dom_ms / 1e3) / 1e9
    traffic, lds, traffic_split = None, None, None
    prof = os.path.join(ROOT, "profiles", "solve_k256_traffic.json")
    if os.path.exists(prof):
        with open(prof) as f:
            tj = json.load(f)
        traffic = tj.get("traffic_bytes_per_launch")
        traffic_split = {"hbm_read": tj.get("hbm_read_bytes_per_launch"),
                         "infinity_cache_read": tj.get("infinity_cache_read_bytes_per_launch"),
                         "dram_destined_read": tj.get("dram_destined_read_bytes_per_launch"),
                         "write": tj.get("write_bytes_per_launch")}
        # LDS bytes the lanes request per launch (rocprof SQ_INSTS_LDS_*_BANDWIDTH
        # of the same launch), over this run's launch time
        if tj.get("lds_load_bytes_per_launch"):
            lb = tj["lds_load_bytes_per_launch"] + tj.get("lds_store_bytes_per_launch", 0.0)
            tbs = lb / (dom_ms / 1e3) / 1e12
            lds = {"achieved": tbs, "peak": LDS_PEAK_TBS, "unit": "TB/s", "frac": tbs / LDS_PEAK_TBS,
                   "bytes_per_launch": lb,
                   "busy": tj["lds_idx_active_cycles_per_launch"]
                   / (tj["grbm_gui_active_per_launch"] / 8 * 256)}
    # gather stage: algorithmic bytes (the k x k FP32 G_S + indices + one
    # output), and the HBM bytes the profiler counts for the same launch
    gB = n256 * (4 * DOMINANT_K * DOMINANT_K + 4 * DOMINANT_K + 4)
    gather = {"ms": gather_ms, "subsets": n256,
              "algorithmic_GBs": gB / (gather_ms / 1e3) / 1e9, "peak_GBs": HBM_PEAK_GBS}
    gprof = os.path.join(ROOT, "profiles", "gather_k256_traffic.json")
    if os.path.exists(gprof):
        with open(gprof) as f:
            gj = json.load(f)
        # DRAM-destined line fills (TCC_EA0_RDREQ_DRAM x 128 B), of which the
        # Infinity Cache serves the modelled share; hbm_frac uses the rest
        db, hb = gj.get("dram_destined_bytes_per_launch"), gj.get("hbm_read_bytes_per_launch")
        if db and hb:
            gather["dram_destined_GBs"] = db / (gather_ms / 1e3) / 1e9
            gather["infinity_cache_frac_model"] = gj.get("infinity_cache_hit_fraction_model")
            gather["hbm_GBs"] = hb / (gather_ms / 1e3) / 1e9
            gather["hbm_frac"] = gather["hbm_GBs"] / HBM_PEAK_GBS
            gather["copy_1gib_read_plus_write_GBs"] = gj.get("copy_1gib_read_plus_write_tbs", 0) * 1e3

    legs = {}
    if not args.no_legs:
        legs.update(small_configs_leg(solver, torch))
        legs["batching"] = batching_leg(solver, torch)
        legs["precision"] = precision_leg(solver, torch, dev)
        legs["host_buffers"] = host_buffers_leg(solver, torch, G, idx_all, off_all, tot)

    cpu = None
    if not args.no_cpu:
        import oracle

        cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        # a bounded sample: leading subsets of the step, in chunks, until the
        # time bound; the oracle reads G on the host (4 GiB copy)
        Gh = G.cpu().numpy()
        done, spent, ch = 0, 0.0, 64
        oracle.solve_batch(Gh, idx_h[:off_h[8]], offsets=off_h[:9], nthreads=cores)  # warm
        while spent < args.cpu_seconds and done < tot:
            hi = min(tot, done + ch)
            t2 = time.perf_counter()
            oracle.solve_batch(Gh, idx_h[off_h[done]:off_h[hi]], offsets=off_h[done:hi + 1]
                               - off_h[done], nthreads=cores)
            spent += time.perf_counter() - t2
            done = hi
            ch = min(ch * 2, 4096)
        cpu = {"value": done / spent, "unit": "eig-solves/s", "cores": cores, "kind": "port",
               "sample": f"first {done} subsets of the step's request wave (same k mix), FP64 "
                         f"oracle (full reorth + Sturm bisection), OpenMP over subsets, "
                         f"{spent:.1f} s"}
        # the reference's own algorithm as it stands: FP64, no
        # reorthogonalisation, TQLRAT over all eigenvalues, on the same leading
        # subsets; its extremes against the FP64 oracle's on that sample
        fd, fspent, ch = 0, 0.0, 64
        f_lo, f_hi, f_err = [], [], []
        while fspent < args.cpu_seconds / 2 and fd < tot:
            hi = min(tot, fd + ch)
            t2 = time.perf_counter()
            lo_, hi_, ie_ = oracle.faithful_extremes(
                Gh, idx_h[off_h[fd]:off_h[hi]], offsets=off_h[fd:hi + 1] - off_h[fd],
                nthreads=cores)
            fspent += time.perf_counter() - t2
            f_lo.append(lo_)
            f_hi.append(hi_)
            f_err.append(ie_)
            fd = hi
            ch = min(ch * 2, 4096)
        f_lo, f_hi, f_err = np.concatenate(f_lo), np.concatenate(f_hi), np.concatenate(f_err)
        nchk = min(fd, 2048)
        o_lo, o_hi, _, _ = oracle.solve_batch(Gh, idx_h[:off_h[nchk]], offsets=off_h[:nchk + 1],
                                              nthreads=cores)
        ferr = np.maximum(np.abs(f_lo[:nchk] - o_lo) / np.abs(o_lo),
                          np.abs(f_hi[:nchk] - o_hi) / np.abs(o_hi))
        cpu["reference_faithful"] = {
            "value": fd / fspent, "unit": "eig-solves/s", "cores": cores,
            "sample": f"first {fd} subsets of the step's request wave, FP64 Lanczos without "
                      f"reorthogonalisation + TQLRAT (all eigenvalues), OpenMP over subsets, "
                      f"{fspent:.1f} s",
            "ierr_nonzero": int(np.count_nonzero(f_err)),
            "max_rel_err_vs_oracle": float(np.nanmax(ferr)),
            "median_rel_err_vs_oracle": float(np.nanmedian(ferr)), "checked": int(nchk)}
        del Gh
        # the small uniform configuration runs whole on the host: Phi 64 x 256
        # (seed 0xC1), 1024 subsets of k = 8 (stream 0xC1+1), one thread and
        # all cores
        G1 = solver.gram(solver.gen_phi("gauss", 0xC1, 64, 256)).cpu().numpy()
        i1 = solver.gen_subsets(0xC1 + 1, 1024, 8, 256).cpu().numpy()
        c1 = {}
        for th in (1, cores):
            oracle.solve_batch(G1, i1, k=8, nthreads=th)  # warm
            reps, t2 = 0, time.perf_counter()
            while time.perf_counter() - t2 < 0.5:
                oracle.solve_batch(G1, i1, k=8, nthreads=th)
                reps += 1
            c1[f"solves_s_{th}_threads"] = 1024 * reps / (time.perf_counter() - t2)
            oracle.faithful_extremes(G1, i1, k=8, nthreads=th)  # warm
            reps, t2 = 0, time.perf_counter()
            while time.perf_counter() - t2 < 0.5:
                oracle.faithful_extremes(G1, i1, k=8, nthreads=th)
                reps += 1
            c1[f"faithful_solves_s_{th}_threads"] = 1024 * reps / (time.perf_counter() - t2)
        cpu["c1"] = dict(c1, workload="Phi 64x256, 1024 subsets of k=8, FP64 oracle (full "
                                      "reorth + Sturm) and the reference-faithful path (no "
                                      "reorth, TQLRAT)")

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
        "config": {"workload": "serving: partial DCT-II Phi 4096x32768 (seed 0xC5), G = Phi^T Phi "
                               "32768x32768 FP32 (4 GiB) resident; per step a wave of requests "
                               "of log-uniform 1..4096 subsets, k uniform in {16,32,64,128,256}, "
                               "served closed-loop (64 clients, each resubmitting on completion) by "
                               "the continuous-batching engine (tuned "
                               "form, FP64 Gram-Schmidt coefficients from k=128)",
                   "n": N, "m": M, "requests_per_step": len(reqs), "subsets_per_step": tot,
                   "k": KLIST, "form": "tuned+served", "parallelism": "single-gpu"},
        "p50_ms": closed["p50_s"] * 1e3,
        "p99_ms": closed["p99_s"] * 1e3,
        "served_waves": {"value": tot * args.steps / waves_wall,
                         "ms_per_step": waves_wall * 1e3 / args.steps,
                         "note": "the same requests as waves that drain between them"},
        "latency_poisson": latency,
        "plain": dict(batch["plain_wave"], accuracy={
            "sample": int(len(errs_plain)), "max_rel_err": float(errs_plain.max()),
            "median_rel_err": float(np.median(errs_plain)), "vs": "FP64 dsyevd",
            "nonzero_status": int(np.count_nonzero(plain_out[2]))}),
        "plain_multiwave": batch["plain"],
        "tuned_batch": batch["tuned"],
        "tuned_plan": batch["tuned_plan"],
        "vs_plain": value / batch["plain_wave"]["value"],
        "vs_plain_multiwave": value / batch["plain"]["value"],
        "served_f32": dict(served_f32, vs_plain_multiwave=served_f32["value"] / batch["plain"]["value"]),
        "accuracy": {"sample": int(len(errs)), "max_rel_err": float(errs.max()),
                     "median_rel_err": float(np.median(errs)), "vs": "FP64 dsyevd",
                     "nonzero_status": int(np.count_nonzero(status))},
        "roofline": {"bound": "valu", "achieved": tflops, "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": tflops / FP32_PEAK_TFLOPS, "traffic": traffic,
                     "traffic_split": traffic_split,
                     "kernel": "solve_rowpart_kernel<256, 2, TRIDIAG, double>",
                     "launch_ms": dom_ms, "subsets_per_launch": n256},
        "roofline_k128": {"bound": "valu", "kernel": "solve_rowpart_kernel<128, 2, TRIDIAG, double>",
                          "launch_ms": k128_ms, "subsets_per_launch": n128,
                          "achieved": n128 * flops_per_solve(128) / (k128_ms / 1e3) / 1e12,
                          "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                          "frac": n128 * flops_per_solve(128) / (k128_ms / 1e3) / 1e12
                          / FP32_PEAK_TFLOPS},
        "roofline_hbm": {"achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS},
        "roofline_lds": lds,
        "gather": gather,
        "gram": gram,
        **legs,
        "cpu_baseline": cpu,
        "setup_s": time.perf_counter() - t_setup - wall,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
