#!/usr/bin/env python3
"""Throughput of the batched FP64 node-ODE evolver on one MI355X.

Workload: 1e5 independent isolated-halo node ODE systems (30 FP64 state
variables each, structure-of-arrays in HBM), halo mass log-uniform in
[1e9, 1e13] Msun, hot-gas fraction uniform in [0.05, 0.17], metallicity
log-uniform in [1e-4, 0.02], interval uniform in [0.1, 2] Gyr, Cash-Karp
with relative tolerance 1e-3 and absolute tolerance 1e-6 x scale, cooling
table 200 (T) x 50 (Z).  One step = one evolve call over the whole batch,
inputs already resident in HBM.

Metric: node-Gyr evolved per second = sum over nodes of (t_end - t_start)
divided by the wall time of a step.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "node-Gyr evolved/sec: FP64 batched RK vs MLP+fallback at ≤1% M* error"
FP64_PEAK_TFLOPS = 78.6


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(batch, physics, table, tol_abs, tol_rel, flags, target_seconds=15.0):
    """The CPU restatement (oracle) on a bounded sample of the same batch, all host cores."""
    import oracle
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit():
        cores = min(cores, int(omp))
    cores = max(1, cores)
    # calibrate on 200 nodes with one thread
    m0 = min(200, batch.n)
    sl = slice(0, m0)
    t = time.perf_counter()
    oracle.evolve(batch.y[:, sl], batch.node_par[:, sl], batch.t_start[sl], batch.t_end[sl], batch.h_step[sl],
                  physics, table, tol_abs, tol_rel, flags, root_tol_rel=1e-6, nthreads=1)
    per_node = (time.perf_counter() - t) / m0
    m = int(min(batch.n, max(m0, target_seconds * cores / max(per_node, 1e-9))))
    sl = slice(0, m)
    t = time.perf_counter()
    _, _, _, st, _ = oracle.evolve(batch.y[:, sl], batch.node_par[:, sl], batch.t_start[sl], batch.t_end[sl],
                                   batch.h_step[sl], physics, table, tol_abs, tol_rel, flags, root_tol_rel=1e-6,
                                   nthreads=cores)
    wall = time.perf_counter() - t
    gyr = float(np.sum(np.abs(batch.t_end[sl] - batch.t_start[sl])))
    return {"value": gyr / wall, "unit": "node-Gyr/s", "cores": cores, "kind": "port",
            "sample": f"first {m} of the {batch.n} nodes ({gyr:.1f} node-Gyr), oracle/oracle.c with OpenMP "
                      f"schedule(dynamic) over nodes, cooling-radius root tolerance 1e-6 as the reference, "
                      f"{wall:.2f} s wall"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nodes", type=int, default=100000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every box in this setup has one MI355X, and the batched evolver shards "
              f"by whole node batches or forests (see DESIGN.md). --gpus {args.gpus} is not supported.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from galacticus_amd import params as gp, workloads as wl
    from galacticus_amd.cooling import synthetic_cie_table
    from galacticus_amd.evolver import NodeEvolverBatched
    from galacticus_amd import flops as gflops

    if not torch.cuda.is_available():
        log("no GPU visible")
        sys.exit(1)
    tol_abs, tol_rel = 1.0e-6, 1.0e-3
    physics = gp.physics_params()
    table = synthetic_cie_table()
    batch = wl.isolated_halos(args.nodes, seed=args.seed, params=physics)
    n = batch.n
    node_gyr = batch.node_gyr
    ev = NodeEvolverBatched(n, physics, table, odeToleranceAbsolute=tol_abs, odeToleranceRelative=tol_rel,
                            reuseODEStepSize=True)
    dev = torch.device("cuda:0")
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    y0, npar, t0, t1, h0 = d(batch.y), d(batch.node_par), d(batch.t_start), d(batch.t_end), d(batch.h_step)
    total = args.warmup + args.steps
    ys = [y0.clone() for _ in range(total)]
    hs = [h0.clone() for _ in range(total)]
    scale = torch.empty_like(y0)
    status = torch.empty(n, dtype=torch.int32, device=dev)
    counters = torch.zeros(3 * n, dtype=torch.int64, device=dev)
    torch.cuda.synchronize()

    for i in range(args.warmup):
        ev.evolve_device(ys[i], npar, t0, t1, hs[i], scale=scale, status=status, counters=counters)
    torch.cuda.synchronize()
    log(f"warmup done ({args.warmup} steps)")

    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_wall = time.perf_counter()
    for k in range(args.steps):
        i = args.warmup + k
        starts[k].record()
        ev.evolve_device(ys[i], npar, t0, t1, hs[i], scale=scale, status=status, counters=counters)
        ends[k].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall
    kernel_ms = float(np.mean([s.elapsed_time(e) for s, e in zip(starts, ends)]))
    ms_per_step = wall * 1e3 / args.steps
    st = status.cpu().numpy()
    if not (st == 0).all():
        log(f"non-success statuses: {np.unique(st, return_counts=True)}")
        sys.exit(1)
    cnt = counters.view(3, n).cpu().numpy()
    attempts = int(cnt[0].sum())
    flop_step = gflops.attempt_flops()
    achieved_tflops = attempts * flop_step / (kernel_ms * 1e-3) / 1e12
    log(f"step {ms_per_step:.3f} ms (kernel {kernel_ms:.3f} ms), attempts {attempts}, max/node {cnt[0].max()}")

    result = {
        "metric": METRIC,
        "value": node_gyr / (ms_per_step * 1e-3),
        "unit": "node-Gyr/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "batched plain RK: isolated-halo node ODE systems (FP64 Cash-Karp, rel 1e-3, abs 1e-6 x scale)",
            "n_nodes": n, "n_vars": 30, "node_gyr": node_gyr, "seed": args.seed,
            "halo_mass": "log-uniform 1e9-1e13 Msun", "f_hot": "uniform 0.05-0.17", "metallicity": "log-uniform 1e-4-0.02",
            "dt_gyr": "uniform 0.1-2", "cooling_table": "200x50 FP64", "parallelism": "single GPU, one node per lane",
        },
        "roofline": {
            "bound": "mfma",
            "compute": "fp64 (VALU; the FP64 vector and matrix peaks are both 78.6 TFLOP/s on MI355X)",
            "achieved": achieved_tflops,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / FP64_PEAK_TFLOPS,
            "traffic": None,
            "flop_per_attempt": flop_step,
            "attempts_per_launch": attempts,
            "kernel_ms": kernel_ms,
        },
    }
    if not args.no_cpu:
        cb = cpu_baseline(batch, physics, table, tol_abs, tol_rel, gp.FLAG_REUSE_STEP_SIZE, args.cpu_seconds)
        result["cpu_baseline"] = cb
        log(f"cpu baseline {cb['value']:.4g} node-Gyr/s on {cb['cores']} cores; GPU/CPU = {result['value']/cb['value']:.1f}")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
