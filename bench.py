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

# Independent batches overlap on separate streams only as far as the HIP
# runtime gives those streams separate hardware queues; with its default of 4
# queues per process, four streams measured 18.6 ms per step, with 8 or 16
# queues 10.0 ms.  The later legs create streams of their own (surrogate,
# concurrent forests): with 8 queues the four forest streams shared queues
# (64 trees in 0.29 s), with 16 they did not (0.17 s).  Read when the HIP
# runtime starts, so it is set before torch is.
# Up to 32 evolve streams, each on a queue of its own: 32 queues (the most
# this pool allows).  With 22, the 20 timed batches ran at 14.5 ms per step
# instead of 11.1; with 20, two of them shared a queue (19.6 ms).  A process
# keeps a queue for every stream it ever used, and beyond about 24 the
# hardware scheduler time-slices them, which makes persistent tree launches
# 1.5-1.8x slower: the tree leg runs before the others create their streams.
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 32:
    os.environ["GPU_MAX_HW_QUEUES"] = "32"

METRIC = "node-Gyr evolved/sec: FP64 batched RK vs MLP+fallback at ≤1% M* error"
FP64_PEAK_TFLOPS = 78.6

# HBM bytes per evolve launch on the 1e5-node isolated-halo batch, by launch
# shape, from rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE runs of this
# script (profiles/evolve_kernel_hbm_traffic_isolated_halos_1e5.md), each
# counter scaled by what it reads for one elementwise kernel that reads and
# writes 1 GiB (FETCH_SIZE x 2.0000, WRITE_SIZE x 1.0000).
HBM_TRAFFIC_1E5 = {
    (64, 1, 2): (241314.75 * 2.0 + 223554.75 * 1.0) * 1024,
    (64, 2, 2): (1922487.5 * 2.0 + 3820465.0 * 1.0) * 1024,
}


_T_START = time.perf_counter()


def log(msg):
    print(f"[{time.perf_counter() - _T_START:7.1f} s] {msg}", file=sys.stderr, flush=True)


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo", encoding="utf-8") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(batch, physics, table, tol_abs, tol_rel, flags, target_seconds=15.0):
    """The CPU restatement (oracle) on a bounded sample of the same batch: one
    warm-up run and the median of five timed runs on all host cores, and the
    same on one core, built for this host (-O3 -march=native)."""
    import tempfile
    import oracle
    build = "prebuilt oracle/lib/liboracle.so (-O3 -march=x86-64-v3)"
    try:
        oracle.use_library(oracle.build_native(os.path.join(tempfile.gettempdir(), "galacticus_amd_oracle")))
        build = "oracle/oracle.c built on this host with gcc -O3 -march=native -fopenmp -fno-math-errno"
    except Exception as e:      # a host without gcc keeps the portable build
        log(f"cpu baseline: native build failed ({e}); using the prebuilt oracle")
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit():
        cores = min(cores, int(omp))
    cores = max(1, cores)

    def run(m, threads):
        sl = slice(0, m)
        t = time.perf_counter()
        _, _, _, st, _ = oracle.evolve(batch.y[:, sl], batch.node_par[:, sl], batch.t_start[sl], batch.t_end[sl],
                                       batch.h_step[sl], physics, table, tol_abs, tol_rel, flags, root_tol_rel=1e-6,
                                       nthreads=threads)
        return time.perf_counter() - t, float(np.sum(np.abs(batch.t_end[sl] - batch.t_start[sl])))

    # calibrate on 200 nodes with one thread
    m0 = min(200, batch.n)
    per_node = run(m0, 1)[0] / m0
    out = {}
    for label, threads, share in (("all", cores, 0.65), ("one", 1, 0.35)):
        per_run = target_seconds * share / 6.0
        m = int(min(batch.n, max(m0, per_run * threads / max(per_node, 1e-9))))
        walls = [run(m, threads) for _ in range(6)]
        gyr = walls[0][1]
        med = float(np.median([w for w, _ in walls[1:]]))
        out[label] = (gyr / med, m, gyr, med)
    v_all, m_all, gyr_all, med_all = out["all"]
    v_one, m_one, gyr_one, med_one = out["one"]
    return {"value": v_all, "unit": "node-Gyr/s", "cores": cores, "kind": "port",
            "one_core_value": v_one, "nproc": os.cpu_count(), "cpu_model": _cpu_model(),
            "sample": f"first {m_all} of the {batch.n} nodes ({gyr_all:.1f} node-Gyr) on {cores} threads, median "
                      f"{med_all:.2f} s of 5 runs after one warm-up; first {m_one} nodes ({gyr_one:.1f} node-Gyr) on "
                      f"one core, median {med_one:.2f} s; {build}; OpenMP schedule(dynamic) over nodes; "
                      f"cooling-radius root tolerance 1e-6 as the reference"}


MASS_HEADS = ("DISK_MASS_STELLAR", "SPHEROID_MASS_STELLAR", "DISK_MASS_GAS", "SPHEROID_MASS_GAS", "HOT_MASS", "BH_MASS")


_STREAMS = []


def stream_pool(n):
    """n streams from one pool for the whole run.  Every stream holds a
    hardware queue for the rest of the process (destroying it does not give
    the queue back), and with more than about 24 queues in a process the
    hardware scheduler time-slices them: a persistent tree launch then takes
    1.5-1.8x as long.  The legs therefore share the integrator leg's streams
    instead of creating their own."""
    import torch
    while len(_STREAMS) < n:
        _STREAMS.append(torch.cuda.Stream())
    return _STREAMS[:n]


def _timed(fn, reps=2, all_times=None):
    import torch
    best = None
    out = None
    for _ in range(reps):
        torch.cuda.synchronize()
        t = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        if all_times is not None:
            all_times.append(dt)
        best = dt if best is None else min(best, dt)
    return best, out


def parse_tree_splits(text):
    """'4,8,4x2' -> ((4, 4), (8, 8), (4, 2)): F forests on S streams (S = F
    when omitted); splits into fewer than two forests are dropped."""
    out = []
    for item in str(text).split(","):
        item = item.strip()
        if not item:
            continue
        f, _, st = item.partition("x")
        n_forests = int(f)
        n_streams = int(st) if st else n_forests
        if n_streams < 1:
            raise ValueError(f"tree split {item!r}: at least one stream")
        if n_forests > 1:
            out.append((n_forests, n_streams))
    return tuple(out)


def surrogate_leg(physics, table, tol_abs, tol_rel, args):
    """Surrogate training on integrator trajectories of merger trees drawn as
    the 64-tree workload (roots log-uniform 1e10-1e12 Msun, resolution 1e9
    Msun, z=5 -> 0; forests of 1024 such trees, seeds 100, 101, ...), then
    surrogate+fallback against the plain integrator on held-out tree
    intervals (seeds 200, ...), the 64-tree forest evolved with the surrogate
    in front of the integrator, and chained isolated-halo histories.  Of the
    networks trained from `train_seeds` seeds, the one that rejects fewest
    validation intervals (seeds 300, ...) is kept."""
    import torch
    from galacticus_amd import workloads as wl, surrogate as sg
    from galacticus_amd.evolver import NodeEvolverBatched
    ev = NodeEvolverBatched(args.eval_pairs, physics, table, odeToleranceAbsolute=tol_abs,
                            odeToleranceRelative=tol_rel)
    t = time.perf_counter()
    pairs = sg.generate_tree_pairs(physics, table, args.train_pairs, seed=100, tol_abs=tol_abs, tol_rel=tol_rel)
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - t
    log(f"surrogate: {pairs.n} training pairs from tree trajectories in {gen_s:.2f} s")
    # Gate: predicted error 0.05, baryon budget 1.5%, metal budget 10%,
    # negatives down to -1 s.  In the measured sweep it keeps the 64-tree
    # forest's z=0 stellar masses within about 0.5% (median) and 4% (max) of
    # the integrator-only forest; a 2% baryon budget runs faster but lets
    # single roots drift by 10%.
    gate = sg.GateConfig(error_max=0.05, mass_tolerance=1.5e-2, metal_tolerance=1e-1, negative_tolerance=1.0)
    # Model selection: networks trained from different initialisations and
    # shuffles reject different fractions of intervals (8-12% in repeated
    # trainings); each is scored on validation intervals of their own forests
    # (seeds 300, ...) and the one that rejects fewest is kept.  The
    # evaluation intervals below (seeds 200, ...) take no part in the choice.
    val = sg.generate_tree_pairs(physics, table, args.val_pairs, seed=300, tol_abs=tol_abs, tol_rel=tol_rel)
    h_val = -torch.ones(val.n, dtype=torch.float64, device=val.t0.device)
    candidates = []
    for s_train in range(max(1, args.train_seeds)):
        mlp_s, norm_s, rep_s = sg.train_surrogate(ev, pairs, epochs=args.train_epochs, batch=8192, seed=s_train)
        y_v, h_v = val.y0.clone(), h_val.clone()
        _, _, n_rej_v = sg.SurrogateEvolver(ev, mlp_s, norm_s, gate).evolve_device(y_v, val.node_par, val.t0,
                                                                                   val.t1, h_v)
        candidates.append((n_rej_v / val.n, s_train, mlp_s, norm_s, rep_s))
        log(f"surrogate: network seed {s_train} trained {rep_s.steps} steps in {rep_s.seconds:.2f} s "
            f"({rep_s.samples_per_s:.3g} samples/s, {rep_s.tflops:.2f} TFLOP/s), loss {rep_s.loss_first:.4g} -> "
            f"{rep_s.loss_last:.4g}, rejects {n_rej_v / val.n:.3f} of {val.n} validation intervals")
        del y_v, h_v
    rejected_by_seed = {str(c[1]): c[0] for c in candidates}
    _, chosen_seed, mlp, norm, rep = min(candidates, key=lambda c: c[0])
    for c in candidates:
        if c[1] != chosen_seed:
            c[2].close()
    del pairs, val, h_val, candidates
    test = sg.generate_tree_pairs(physics, table, args.eval_pairs, seed=200, tol_abs=tol_abs, tol_rel=tol_rel)
    n = test.n
    gyr = float((test.t1 - test.t0).sum())
    h0 = -torch.ones(n, dtype=torch.float64, device=test.t0.device)
    holder = {}

    launch = ev.autotune(test.y0, test.node_par, test.t0, test.t1, h0)

    def run_rk():
        y, h = test.y0.clone(), h0.clone()
        ev.evolve_device(y, test.node_par, test.t0, test.t1, h)
        holder["rk"] = y
    t_rk, _ = _timed(run_rk)
    se = sg.SurrogateEvolver(ev, mlp, norm, gate)

    def run_sur():
        y, h = test.y0.clone(), h0.clone()
        st, acc, n_rej = se.evolve_device(y, test.node_par, test.t0, test.t1, h)
        holder["sur"] = (y, acc, n_rej, st)
    t_sur, _ = _timed(run_sur)
    y_rk = holder["rk"]
    y_s, acc, n_rej, st = holder["sur"]
    V = wl.V
    accb = acc.bool()
    per_q = {}
    for name in MASS_HEADS:
        ref, got = y_rk[V[name]], y_s[V[name]]
        sig = ref.abs() > 1e-6 * test.y0[V["HOT_MASS"]].abs().clamp_min(1.0)
        m = accb & sig
        rel = ((got - ref).abs() / ref.abs().clamp_min(1e-300))[m]
        per_q[name] = {"median": float(rel.median()) if rel.numel() else 0.0,
                       "p90": float(torch.quantile(rel[:2 ** 24].float(), 0.9)) if rel.numel() else 0.0}
    ms_ref = y_rk[V["DISK_MASS_STELLAR"]] + y_rk[V["SPHEROID_MASS_STELLAR"]]
    ms_sur = y_s[V["DISK_MASS_STELLAR"]] + y_s[V["SPHEROID_MASS_STELLAR"]]
    ms_err = ((ms_sur - ms_ref).abs() / ms_ref.clamp_min(1e-300))
    # the same surrogate on chained node histories from z=5 to z=0, and in
    # the 64-tree forest
    chain = chained_histories(ev, se, physics, args)
    forest = surrogate_in_forest(se, physics, table, tol_abs, tol_rel)
    rk_rate, sur_rate = gyr / t_rk, gyr / t_sur
    piped = pipelined_surrogate(ev, mlp, norm, gate, test, h0, launch, physics, table, tol_abs, tol_rel, args)
    log(f"surrogate: RK {rk_rate:.4g} node-Gyr/s, surrogate+fallback {sur_rate:.4g} ({sur_rate / rk_rate:.2f}x), "
        f"rejected {n_rej / n:.3f}, median M* error {float(ms_err.median()):.2e}")
    log(f"surrogate pipelined: RK {piped['rk_s_per_batch'] * 1e3:.2f} ms/batch "
        f"({piped['rk_streams']} streams, launch {piped['rk_launch']}), surrogate+fallback "
        f"{piped['surrogate_s_per_batch'] * 1e3:.2f} ms/batch ({piped['surrogate_streams']} streams, "
        f"launch {piped['surrogate_launch']}) "
        f"({piped['rk_s_per_batch'] / piped['surrogate_s_per_batch']:.2f}x)")
    args.trained_surrogate = se
    return {
        "workload": "node intervals of merger trees drawn as the 64-tree workload (forests of 1024 trees, roots "
                    "1e10-1e12 Msun, z=5 -> 0; training seeds 100+, evaluation seeds 200+), FP64 integrator vs "
                    "bf16 MLP + gate + FP64 fallback",
        "training": {"selection": {"networks": len(rejected_by_seed), "validation_pairs": args.val_pairs,
                                   "validation_rejected_by_seed": rejected_by_seed, "chosen_seed": chosen_seed},
                     "pairs": int(rep.steps * rep.batch / max(rep.epochs, 1)), "epochs": rep.epochs, "batch": rep.batch,
                     "steps": rep.steps, "seconds": rep.seconds, "samples_per_s": rep.samples_per_s,
                     "data_generation_s": gen_s, "loss_first": rep.loss_first, "loss_last": rep.loss_last,
                     "roofline": {"bound": "mfma", "achieved": rep.tflops, "peak": 2500.0, "unit": "TFLOP/s",
                                  "frac": rep.tflops / 2500.0, "flop_per_sample": sg.FLOPS_TRAIN_PER_ROW}},
        "inference": {"intervals": n, "node_gyr": gyr, "launch": list(launch), "rk_node_gyr_per_s": rk_rate,
                      "surrogate_node_gyr_per_s": sur_rate, "speedup": sur_rate / rk_rate,
                      "rejected_fraction": n_rej / n, "statuses_ok": bool((st == 0).all()),
                      "gate": dataclasses_asdict(gate),
                      "median_stellar_mass_error_all": float(ms_err.median()),
                      "accepted_relative_error": per_q},
        "chained_to_z0": chain,
        "forest_64_trees": forest,
        "pipelined": dict(piped, rk_node_gyr_per_s=gyr / piped["rk_s_per_batch"],
                          surrogate_node_gyr_per_s=gyr / piped["surrogate_s_per_batch"],
                          speedup=piped["rk_s_per_batch"] / piped["surrogate_s_per_batch"]),
    }


def dataclasses_asdict(x):
    import dataclasses
    return dataclasses.asdict(x)


PIPELINE_SHAPES = ((64, 1, 1), (64, 2, 1), (64, 4, 1), (64, 1, 2), (64, 2, 2))


def pipelined_surrogate(ev, mlp, norm, gate, test, h0, launch, physics, table, tol_abs, tol_rel, args):
    """Throughput of both paths as bench's main leg measures the integrator:
    independent copies of the evaluation batch alternate over streams, each
    stream with its own integrator handle (and network handle), so one
    batch's tail of stiff nodes overlaps the next batch's start.  Each path
    runs with the number of streams and the integrator launch shape that are
    fastest for it in this overlapped regime (the fallback batch is a sparse
    subset of stiff intervals, whose best configuration differs from the full
    batch's): every combination gets one timed pass, the fastest is timed
    afresh (median of three passes)."""
    import torch
    from galacticus_amd import surrogate as sg
    from galacticus_amd.evolver import NodeEvolverBatched
    counts = sorted({max(1, c) for c in args.pipeline_streams})
    n_streams, batches = max(counts), args.pipeline_batches
    n = test.n
    evs = [ev] + [NodeEvolverBatched(n, physics, table, odeToleranceAbsolute=tol_abs, odeToleranceRelative=tol_rel)
                  for _ in range(n_streams - 1)]
    params = mlp.get_params()
    mlps = [mlp] + [sg.SurrogateMLP(8192, params=params) for _ in range(n_streams - 1)]
    ses = [sg.SurrogateEvolver(e, m, norm, gate) for e, m in zip(evs, mlps)]
    streams = stream_pool(n_streams)
    status = [torch.zeros(n, dtype=torch.int32, device=test.t0.device) for _ in range(n_streams)]
    accepted = [torch.ones(n, dtype=torch.int32, device=test.t0.device) for _ in range(n_streams)]
    out = {}
    for name in ("rk", "surrogate"):
        ys = [test.y0.clone() for _ in range(batches)]
        hs = [h0.clone() for _ in range(batches)]

        def launch_one(k, used):
            j = k % used
            if name == "rk":
                evs[j].evolve_device(ys[k], test.node_par, test.t0, test.t1, hs[k], status=status[j], stream=streams[j])
            else:
                ses[j].evolve_device(ys[k], test.node_par, test.t0, test.t1, hs[k], status=status[j],
                                     accepted=accepted[j], stream=streams[j], count_rejected=False)

        def timed_pass(used):
            for k in range(batches):
                ys[k].copy_(test.y0)
                hs[k].copy_(h0)
            torch.cuda.synchronize()
            t = time.perf_counter()
            for k in range(batches):
                launch_one(k, used)
            torch.cuda.synchronize()
            return (time.perf_counter() - t) / batches

        per_config = {}
        for used in counts:
            for shape in PIPELINE_SHAPES:
                for e in evs:
                    e.set_launch(*shape)
                for j in range(used):      # warm-up (allocations of first use)
                    launch_one(j % batches, used)
                torch.cuda.synchronize()
                per_config[(used, shape)] = timed_pass(used)
        used, best = min(per_config, key=per_config.get)
        for e in evs:
            e.set_launch(*best)
        out[f"{name}_s_per_batch"] = float(np.median([timed_pass(used) for _ in range(3)]))
        out[f"{name}_launch"] = list(best)
        out[f"{name}_streams"] = used
        out[f"{name}_s_per_batch_by_config"] = {f"{u} streams, {a},{b},{c}": v
                                                  for (u, (a, b, c)), v in per_config.items()}
        if name == "surrogate":
            out["rejected_fraction_last_batch"] = int((accepted[0] == 0).sum()) / n
        out[f"{name}_statuses_ok"] = all(bool((st == 0).all()) for st in status)
        del ys, hs
    ev.set_launch(*launch)
    out.update({"stream_counts": counts, "batches": batches, "intervals_per_batch": n})
    for m in mlps[1:]:
        m.close()
    return out


def surrogate_in_forest(se, physics, table, tol_abs, tol_rel):
    """The 64-tree workload evolved by the tree executor with and without the
    surrogate in front of the integrator: z=0 stellar masses of the roots."""
    import torch
    from galacticus_amd import trees as tr, workloads as wl
    from galacticus_amd.evolver import NodeEvolverBatched
    from galacticus_amd.surrogate import SurrogateEvolver
    cosmo = tr.cosmology(physics)
    t0 = float(wl.scale_factor_to_cosmic_time(1.0))
    t5 = float(wl.scale_factor_to_cosmic_time(1.0 / 6.0))
    rng = np.random.Generator(np.random.PCG64(0))
    forest = tr.Forest.build_pch(64, 10.0 ** rng.uniform(10.0, 12.0, 64), t0, cosmo, tr.pch_params(1e9, t5), seed=0)
    wp = tr.walk_params(t0)
    sched = tr.Schedule(forest, cosmo, wp)
    y0 = forest.initial_state(cosmo, wp)
    ev = NodeEvolverBatched(forest.n_nodes, physics, table, odeToleranceAbsolute=tol_abs, odeToleranceRelative=tol_rel)
    se_f = SurrogateEvolver(ev, se.mlp, se.norm, se.gate)
    out = {}
    for name, sur in (("rk", None), ("surrogate", se_f)):
        y = torch.from_numpy(y0).cuda()
        torch.cuda.synchronize()
        t = time.perf_counter()
        tr.evolve_forest(ev, sched, y, surrogate=sur)
        torch.cuda.synchronize()
        out[name] = (time.perf_counter() - t, y.cpu().numpy())
    V = wl.V
    r = forest.roots()
    ms = {k: v[1][V["DISK_MASS_STELLAR"], r] + v[1][V["SPHEROID_MASS_STELLAR"], r] for k, v in out.items()}
    rel = np.abs(ms["surrogate"] - ms["rk"]) / np.maximum(ms["rk"], 1e-300)
    n_ev = int((sched.ops["kind"] == 0).sum())
    res = {"roots": int(len(r)), "evolve_intervals": n_ev, "accepted_intervals": se_f.tree_accepted,
           "median_z0_stellar_mass_error": float(np.median(rel)), "p90_z0_stellar_mass_error": float(np.quantile(rel, 0.9)),
           "max_z0_stellar_mass_error": float(rel.max()), "rk_seconds": out["rk"][0],
           "surrogate_seconds": out["surrogate"][0]}
    log(f"surrogate in the 64-tree forest: {se_f.tree_accepted} of {n_ev} intervals accepted, z=0 M* error median "
        f"{np.median(rel):.2e}, p90 {np.quantile(rel, 0.9):.2e}; {out['surrogate'][0]:.3f} s vs {out['rk'][0]:.3f} s")
    return res


def chained_histories(ev, se, physics, args, n=20000):
    """Node histories from z=5 to z=0, interval by interval, once with the
    integrator alone and once with the surrogate in front of it."""
    import torch
    from galacticus_amd import workloads as wl
    hist = wl.halo_histories(n, seed=2, params=physics, start=float(wl.scale_factor_to_cosmic_time(1.0 / 6.0)))
    b = hist.batch
    dev = torch.device("cuda:0")
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    npar = d(b.node_par)
    out = {}
    times = {}
    for mode in ("rk", "surrogate"):
        rng = np.random.Generator(np.random.PCG64(99))
        y, h = d(b.y), d(b.h_step)
        t = b.t_start.copy()
        n_rej = 0
        n_int = 0
        torch.cuda.synchronize()
        t_wall = time.perf_counter()
        while (t < hist.t_final - 1e-9).any():
            t1 = hist.intervals(t, rng)
            t0d, t1d = d(t), d(t1)
            if mode == "rk":
                ev.evolve_device(y, npar, t0d, t1d, h)
            else:
                _, _, r = se.evolve_device(y, npar, t0d, t1d, h)
                n_rej += r
            n_int += n
            t = t1
        torch.cuda.synchronize()
        times[mode] = time.perf_counter() - t_wall
        out[mode] = (y[wl.V["DISK_MASS_STELLAR"]] + y[wl.V["SPHEROID_MASS_STELLAR"]]).cpu().numpy()
        if mode == "surrogate":
            rej = n_rej / n_int
    err = np.abs(out["surrogate"] - out["rk"]) / np.maximum(out["rk"], 1e-300)
    return {"histories": n, "start": "z=5", "median_z0_stellar_mass_error": float(np.median(err)),
            "p90_z0_stellar_mass_error": float(np.quantile(err, 0.9)), "rejected_fraction": rej,
            "rk_seconds": times["rk"], "surrogate_seconds": times["surrogate"]}


def _forest(physics, n_trees, log_m, z_start):
    """A PCH forest of the tree workloads (seed 0, resolution 1e9 Msun, roots
    log-uniform in 10^log_m Msun at z=0, from z_start), its walk and initial
    state; returns (forest, schedule, y0, build seconds, schedule seconds,
    cosmology, walk parameters)."""
    from galacticus_amd import trees as tr, workloads as wl
    cosmo = tr.cosmology(physics)
    t0 = float(wl.scale_factor_to_cosmic_time(1.0))
    t_start = float(wl.scale_factor_to_cosmic_time(1.0 / (1.0 + z_start)))
    rng = np.random.Generator(np.random.PCG64(0))
    roots = 10.0 ** rng.uniform(log_m[0], log_m[1], n_trees)
    t = time.perf_counter()
    forest = tr.Forest.build_pch(n_trees, roots, t0, cosmo, tr.pch_params(1e9, t_start), seed=0)
    build_s = time.perf_counter() - t
    wp = tr.walk_params(t0)
    t = time.perf_counter()
    sched = tr.Schedule(forest, cosmo, wp)
    sched_s = time.perf_counter() - t
    return forest, sched, forest.initial_state(cosmo, wp), build_s, sched_s, cosmo, wp


def _forest_run(physics, table, tol_abs, tol_rel, n_trees, log_m, z_start, label, cpu_threads, cpu=True, split=(),
                waves=True):
    """Build a PCH forest, record its walk, evolve it on the GPU as one
    operation graph (and, with waves, wave by wave for comparison) and (cpu)
    with the CPU oracle's tree runner over cpu_threads threads; compare the
    z=0 stellar masses of the roots."""
    import torch
    from galacticus_amd import trees as tr
    from galacticus_amd.evolver import NodeEvolverBatched
    forest, sched, y0, build_s, sched_s, cosmo, wp = _forest(physics, n_trees, log_m, z_start)
    ev = NodeEvolverBatched(forest.n_nodes, physics, table, odeToleranceAbsolute=tol_abs, odeToleranceRelative=tol_rel)
    holder = {}

    def run():
        y = torch.from_numpy(y0).cuda()
        holder["st"] = tr.evolve_forest(ev, sched, y)
        holder["y"] = y
    # the first call also builds the schedule's operation graph on the
    # device (successor links on the host, one upload), which later calls on
    # the same schedule reuse
    times = []
    gpu_s, _ = _timed(run, reps=3, all_times=times)
    st = holder["st"].cpu().numpy()
    yg = holder["y"].cpu().numpy()
    gyr = sched.node_gyr()
    k = np.bincount(sched.ops["kind"], minlength=4)
    out = {"workload": label, "nodes": forest.n_nodes, "operations": sched.n_ops, "waves": sched.n_waves,
           "evolve_intervals": int(k[0]), "promotions": int(k[1]), "halo_mergers": int(k[2]),
           "galaxy_mergers": int(k[3]), "node_gyr": gyr, "executor": "operation graph, one persistent launch",
           "launch": list(ev.launch()), "gpu_seconds": gpu_s, "gpu_seconds_first_call": times[0],
           "node_gyr_per_s": gyr / gpu_s,
           "host_build_seconds": build_s, "host_schedule_seconds": sched_s, "statuses_ok": bool((st == 0).all())}
    log(f"trees ({label}): {forest.n_nodes} nodes, {sched.n_ops} operations in {sched.n_waves} waves, "
        f"GPU {gpu_s:.3f} s ({gyr / gpu_s:.4g} node-Gyr/s) as one operation graph (first call, which builds "
        f"the graph on the device, {times[0]:.3f} s)")
    # the same schedule wave by wave: every wave waits for its slowest interval
    hold_w = {}

    def run_waves():
        y = torch.from_numpy(y0).cuda()
        hold_w["st"] = tr.evolve_forest(ev, sched, y, executor="waves")
        hold_w["y"] = y
    if waves:
        waves_s, _ = _timed(run_waves, reps=1)
        same_w = bool(np.array_equal(hold_w["y"].cpu().numpy(), yg))
        out["wave_executor"] = {"gpu_seconds": waves_s, "node_gyr_per_s": gyr / waves_s,
                                "identical_to_graph": same_w}
        log(f"trees ({label}): wave by wave {waves_s:.3f} s ({gyr / waves_s:.4g} node-Gyr/s), "
            f"identical states: {same_w}")
    del holder, hold_w, ev
    # the same trees as F independent forests, evolved concurrently over S
    # streams (forest k on stream k mod S), each as its own operation graph:
    # each (F, S) in `split` is timed and the fastest is kept.  One graph
    # already lets every operation start as soon as its own nodes allow, so
    # the default runs none.
    for n_forests, n_streams in split:
        parts = forest.split(n_forests)
        scheds = [tr.Schedule(f, cosmo, wp) for f, _ in parts]
        evs = [NodeEvolverBatched(f.n_nodes, physics, table, odeToleranceAbsolute=tol_abs,
                                  odeToleranceRelative=tol_rel) for f, _ in parts]
        pool = stream_pool(min(n_streams, len(parts)))
        streams = [pool[k % len(pool)] for k in range(len(parts))]
        hold = {}

        def run_split():
            ys = [torch.from_numpy(np.ascontiguousarray(y0[:, n0:n0 + f.n_nodes])).cuda() for f, n0 in parts]
            hold["st"] = tr.evolve_forests(evs, scheds, ys, streams)
            hold["y"] = ys
        split_s, _ = _timed(run_split)
        ys = np.concatenate([y.cpu().numpy() for y in hold["y"]], axis=1)
        same = bool(np.array_equal(ys, yg))
        rec = {"forests": len(parts), "streams": len(pool), "waves_max": max(sc.n_waves for sc in scheds),
               "gpu_seconds": split_s, "node_gyr_per_s": gyr / split_s, "identical_to_single_forest": same}
        out.setdefault("split_by_forests", {})[f"{len(parts)}x{len(pool)}"] = rec
        log(f"trees ({label}): as {len(parts)} concurrent forests on {len(pool)} streams {split_s:.3f} s "
            f"({gyr / split_s:.4g} node-Gyr/s), identical states: {same}")
        if same and split_s < out["gpu_seconds"]:
            out["gpu_seconds"], out["node_gyr_per_s"] = split_s, gyr / split_s
            out["split"] = rec
        del hold, evs
    if cpu:
        import oracle
        from galacticus_amd import params as gp
        t = time.perf_counter()
        yc, _, stc, bad = oracle.tree_run(sched.ops, y0, -np.ones(forest.n_nodes), forest.tree_of_node(),
                                          forest.n_trees, physics, table, tol_abs, tol_rel, gp.FLAG_REUSE_STEP_SIZE,
                                          0.25, root_tol_rel=1.0e-6, nthreads=cpu_threads)
        cpu_s = time.perf_counter() - t
        from galacticus_amd.workloads import V
        r = forest.roots()
        mg = yg[V["DISK_MASS_STELLAR"], r] + yg[V["SPHEROID_MASS_STELLAR"], r]
        mc = yc[V["DISK_MASS_STELLAR"], r] + yc[V["SPHEROID_MASS_STELLAR"], r]
        rel = np.abs(mg - mc) / np.maximum(mc, 1e-300)
        bins = np.arange(6.0, 12.51, 0.25)
        hg = np.histogram(np.log10(np.maximum(mg, 1e-30)), bins)[0]
        hc = np.histogram(np.log10(np.maximum(mc, 1e-30)), bins)[0]
        smf_diff = int(np.abs(hg - hc).sum())
        out["cpu_baseline"] = {"value": gyr / cpu_s, "unit": "node-Gyr/s", "cores": cpu_threads, "kind": "port",
                               "seconds": cpu_s, "failed_evolves": int(bad),
                               "sample": "the whole forest, oracle/oracle.c orc_tree_run (trees in parallel, "
                                         "operations in schedule order), cooling-radius root tolerance 1e-6"}
        out["gpu_over_cpu"] = cpu_s / out["gpu_seconds"]
        out["z0_stellar_mass_vs_cpu"] = {"median_relative": float(np.median(rel)), "max_relative": float(rel.max()),
                                         "smf_bins_0.25dex_count_moved": smf_diff, "roots": int(len(r))}
        log(f"trees ({label}): CPU {cpu_s:.2f} s on {cpu_threads} threads, GPU/CPU {cpu_s / out['gpu_seconds']:.1f}x, "
            f"root M* median rel diff {np.median(rel):.2e}, SMF galaxies moved between bins {smf_diff}")
    return out


def surrogate_forest_route(se, physics, table, tol_abs, tol_rel, n_trees):
    """The larger forest of the tree leg with the trained surrogate in front
    of the integrator for the intervals of non-satellite halos (per-wave
    route), against the same forest's operation graph without it: time, and
    z=0 stellar masses of the roots."""
    import torch
    from galacticus_amd import trees as tr
    from galacticus_amd.evolver import NodeEvolverBatched
    from galacticus_amd.surrogate import SurrogateEvolver
    forest, sched, y0, _, _, _, _ = _forest(physics, n_trees, (11.0, 13.0), 8.0)
    ev = NodeEvolverBatched(forest.n_nodes, physics, table, odeToleranceAbsolute=tol_abs, odeToleranceRelative=tol_rel)
    se_t = SurrogateEvolver(ev, se.mlp, se.norm, se.gate)
    hold = {}

    def run(sur):
        y = torch.from_numpy(y0).cuda()
        hold["st"] = tr.evolve_forest(ev, sched, y, surrogate=sur)
        hold["y"] = y
    plain_s, _ = _timed(lambda: run(None))
    y_plain = hold["y"].cpu().numpy()
    sur_s, _ = _timed(lambda: run(se_t))
    y_sur = hold["y"].cpu().numpy()
    from galacticus_amd.workloads import V
    r = forest.roots()
    mp = y_plain[V["DISK_MASS_STELLAR"], r] + y_plain[V["SPHEROID_MASS_STELLAR"], r]
    ms = y_sur[V["DISK_MASS_STELLAR"], r] + y_sur[V["SPHEROID_MASS_STELLAR"], r]
    rel = np.abs(ms - mp) / np.maximum(mp, 1e-300)
    bins = np.arange(6.0, 12.51, 0.25)
    moved = int(np.abs(np.histogram(np.log10(np.maximum(ms, 1e-30)), bins)[0] -
                       np.histogram(np.log10(np.maximum(mp, 1e-30)), bins)[0]).sum())
    gyr = sched.node_gyr()
    n_ev = int((sched.ops["kind"] == 0).sum())
    n_ns = int(((sched.ops["kind"] == 0) & (sched.ops["node_par"][:, 4] == 0)).sum())
    out = {"gpu_seconds": sur_s, "node_gyr_per_s": gyr / sur_s, "plain_graph_gpu_seconds_here": plain_s,
           "accepted_intervals": se_t.tree_accepted, "evolve_intervals": n_ev, "non_satellite_intervals": n_ns,
           "accepted_fraction": se_t.tree_accepted / max(n_ev, 1),
           "statuses_ok": bool((hold["st"].cpu().numpy() == 0).all()),
           "z0_stellar_mass_vs_plain_gpu": {"median_relative": float(np.median(rel)),
                                            "p90_relative": float(np.quantile(rel, 0.9)),
                                            "max_relative": float(rel.max()), "smf_bins_0.25dex_count_moved": moved,
                                            "roots": int(len(r))},
           "note": "run after the integrator and surrogate legs (their streams hold hardware queues); the plain "
                   "operation graph is timed again beside it"}
    log(f"trees ({n_trees} trees): with the surrogate {sur_s:.3f} s against the plain graph's {plain_s:.3f} s here, "
        f"{se_t.tree_accepted} of {n_ev} intervals accepted; z=0 root M* vs plain median {np.median(rel):.2e}, "
        f"p90 {np.quantile(rel, 0.9):.2e}, SMF galaxies moved {moved}")
    return out


def large_forest_stream(physics, table, tol_abs, tol_rel, n_trees, per_forest, cpu_threads=0, walkers=3):
    """The large-forest workload streamed: n_trees PCH trees (roots drawn as
    the 4096-tree workload's, log-uniform 1e11-1e14 Msun, resolution 1e9
    Msun, z=8 -> 0) as consecutive forests of per_forest trees (forest k from
    builder seed k).  While the GPU evolves forest k as one operation graph,
    `walkers` host threads build and walk forests k+1 .. k+walkers (the
    library calls release the interpreter lock; each walk spreads its trees
    over the host's cores, and concurrent walks overlap one forest's largest
    trees and page faults with the others).  Reports the totals, the GPU and host time per forest
    and the z=0 stellar-mass function of the roots; with cpu_threads, also
    the CPU oracle's tree runner on every forest (trees in parallel) and its
    z=0 stellar masses against the GPU's."""
    from concurrent.futures import ThreadPoolExecutor
    import torch
    from galacticus_amd import params as gp, trees as tr, workloads as wl
    from galacticus_amd.evolver import NodeEvolverBatched
    cosmo = tr.cosmology(physics)
    t0 = float(wl.scale_factor_to_cosmic_time(1.0))
    t8 = float(wl.scale_factor_to_cosmic_time(1.0 / 9.0))
    wp = tr.walk_params(t0)
    roots = 10.0 ** np.random.Generator(np.random.PCG64(0)).uniform(11.0, 14.0, 4096)[:n_trees]
    n_forests = (n_trees + per_forest - 1) // per_forest

    def prepare(k):
        t = time.perf_counter()
        f = tr.Forest.build_pch(min(per_forest, n_trees - k * per_forest), roots[k * per_forest:(k + 1) * per_forest],
                                t0, cosmo, tr.pch_params(1e9, t8), seed=k)
        sc = tr.Schedule(f, cosmo, wp)
        return f, sc, f.initial_state(cosmo, wp), sc.node_gyr(), time.perf_counter() - t

    totals = {"nodes": 0, "operations": 0, "node_gyr": 0.0, "gpu_seconds": 0.0, "host_seconds": 0.0}
    if cpu_threads:
        totals["cpu_seconds"] = 0.0
    rel_all = []
    moved_cpu = 0
    per = []
    ms_roots = []
    ok = True
    t_wall = time.perf_counter()
    # the next forests' walks overlap this one's GPU run, unless the CPU
    # runner is timed too (it would share the host's cores with the walks)
    ahead = 0 if cpu_threads else max(1, int(walkers))
    pool = ThreadPoolExecutor(max_workers=ahead) if ahead else None
    futs = {}
    cur = prepare(0)
    for k in range(n_forests):
        if pool:
            for j in range(k + 1, min(n_forests, k + 1 + ahead)):
                if j not in futs:
                    futs[j] = pool.submit(prepare, j)
        f, sc, y0, gyr, host_s = cur
        ev = NodeEvolverBatched(f.n_nodes, physics, table, odeToleranceAbsolute=tol_abs, odeToleranceRelative=tol_rel)
        y = torch.from_numpy(y0).cuda()
        torch.cuda.synchronize()
        t = time.perf_counter()
        st = tr.evolve_forest(ev, sc, y)
        torch.cuda.synchronize()
        gpu_s = time.perf_counter() - t
        yh = y.cpu().numpy()
        ok = ok and bool((st.cpu().numpy() == 0).all())
        r = f.roots()
        from galacticus_amd.workloads import V
        mg = yh[V["DISK_MASS_STELLAR"], r] + yh[V["SPHEROID_MASS_STELLAR"], r]
        ms_roots.append(mg)
        rec = {"trees": int(f.n_trees), "nodes": int(f.n_nodes), "operations": int(sc.n_ops), "gpu_seconds": gpu_s,
               "host_seconds": host_s}
        if cpu_threads:
            import oracle
            t = time.perf_counter()
            yc, _, _, bad = oracle.tree_run(sc.ops, y0, -np.ones(f.n_nodes), f.tree_of_node(), f.n_trees, physics,
                                            table, tol_abs, tol_rel, gp.FLAG_REUSE_STEP_SIZE, 0.25,
                                            root_tol_rel=1.0e-6, nthreads=cpu_threads)
            cpu_s = time.perf_counter() - t
            mc = yc[V["DISK_MASS_STELLAR"], r] + yc[V["SPHEROID_MASS_STELLAR"], r]
            rel_all.append(np.abs(mg - mc) / np.maximum(mc, 1e-300))
            bins_c = np.arange(6.0, 13.01, 0.25)
            moved_cpu += int(np.abs(np.histogram(np.log10(np.maximum(mg, 1e-30)), bins_c)[0] -
                                    np.histogram(np.log10(np.maximum(mc, 1e-30)), bins_c)[0]).sum())
            rec["cpu_seconds"] = cpu_s
            rec["cpu_failed_evolves"] = int(bad)
            totals["cpu_seconds"] += cpu_s
            del yc
        per.append(rec)
        for key in ("nodes", "operations"):
            totals[key] += rec[key]
        totals["node_gyr"] += gyr
        totals["gpu_seconds"] += gpu_s
        totals["host_seconds"] += host_s
        log(f"large-forest stream: forest {k + 1}/{n_forests}: {f.n_nodes} nodes, {sc.n_ops} operations, GPU "
            f"{gpu_s:.2f} s, host build+walk {host_s:.2f} s"
            + (f", CPU {rec['cpu_seconds']:.2f} s (root M* median rel diff {np.median(rel_all[-1]):.2e})"
               if cpu_threads else ""))
        sc.close()
        del ev, y, yh, st, f, sc, y0, cur
        if k + 1 < n_forests:
            cur = futs.pop(k + 1).result() if pool else prepare(k + 1)
    if pool:
        pool.shutdown()
    wall = time.perf_counter() - t_wall
    ms = np.concatenate(ms_roots)
    bins = np.arange(6.0, 13.01, 0.25)
    smf = np.histogram(np.log10(np.maximum(ms, 1e-30)), bins)[0]
    out = dict(totals, workload=f"{n_trees} PCH trees of the large-forest workload's roots (log-uniform "
                                f"1e11-1e14 Msun, resolution 1e9 Msun, z=8 -> 0) as {n_forests} forests of "
                                f"{per_forest}, forest k from builder seed k", trees=n_trees, forests=n_forests,
               wall_seconds=wall, node_gyr_per_s_gpu=totals["node_gyr"] / totals["gpu_seconds"],
               node_gyr_per_s_wall=totals["node_gyr"] / wall, statuses_ok=ok,
               smf_log10_mstar_bins_0_25dex={"lower_edges": [float(b) for b in bins[:-1]],
                                             "counts": [int(c) for c in smf]},
               per_forest=per)
    if cpu_threads:
        rel = np.concatenate(rel_all)
        out["cpu_baseline"] = {"value": totals["node_gyr"] / totals["cpu_seconds"], "unit": "node-Gyr/s",
                               "cores": cpu_threads, "kind": "port", "seconds": totals["cpu_seconds"],
                               "sample": "every forest of the stream, oracle/oracle.c orc_tree_run (trees in "
                                         "parallel), cooling-radius root tolerance 1e-6"}
        out["gpu_over_cpu"] = totals["cpu_seconds"] / totals["gpu_seconds"]
        out["z0_stellar_mass_vs_cpu"] = {"median_relative": float(np.median(rel)), "max_relative": float(rel.max()),
                                         "smf_bins_0.25dex_count_moved": moved_cpu, "roots": int(len(rel))}
        log(f"large-forest stream: CPU {totals['cpu_seconds']:.1f} s on {cpu_threads} threads, GPU/CPU "
            f"{out['gpu_over_cpu']:.1f}x, root M* median rel diff {np.median(rel):.2e}, SMF moved {moved_cpu}")
    log(f"large-forest stream: {n_trees} trees, {totals['nodes']} nodes, {totals['operations']} operations; GPU "
        f"{totals['gpu_seconds']:.1f} s ({out['node_gyr_per_s_gpu']:.4g} node-Gyr/s), host build+walk "
        f"{totals['host_seconds']:.1f} s, wall {wall:.1f} s")
    return out


def trees_leg(physics, table, tol_abs, tol_rel, args):
    """Merger trees in the reference's walk order as operation graphs: the
    64-tree workload (roots 1e10-1e12 Msun, z=5 -> 0), a larger forest (roots
    1e11-1e13 Msun, z=8 -> 0) and a sample of the 4096-tree workload (roots
    1e11-1e14 Msun, z=8 -> 0), each against the CPU oracle's tree runner."""
    threads = min(16, os.cpu_count() or 1)
    small = _forest_run(physics, table, tol_abs, tol_rel, 64, (10.0, 12.0), 5.0,
                        "64 PCH trees, seed 0, roots log-uniform 1e10-1e12 Msun, resolution 1e9 Msun, z=5 -> 0",
                        threads, cpu=not args.no_cpu, split=args.tree_splits)
    out = dict(small)
    if args.large_trees > 0:
        out["large"] = _forest_run(physics, table, tol_abs, tol_rel, args.large_trees, (11.0, 13.0), 8.0,
                                   f"{args.large_trees} PCH trees, seed 0, roots log-uniform 1e11-1e13 Msun, "
                                   f"resolution 1e9 Msun, z=8 -> 0", threads, cpu=not args.no_cpu,
                                   split=args.tree_splits)
    if args.c5_stream_trees > 0:
        out["c5_stream"] = large_forest_stream(physics, table, tol_abs, tol_rel, args.c5_stream_trees,
                                               args.c5_forest_trees, 0 if args.no_cpu else threads,
                                               walkers=args.c5_walkers)
    if args.c5_trees > 0:
        # a sample of the 4096-tree large-forest workload: the seeds and
        # roots of its first trees (the builder seeds every tree by its
        # index; its rate table spans the call's largest root, so the
        # branchings differ in detail from a build of all 4096), the same
        # epochs.  The whole workload's ~4e9 operations do
        # not fit a host schedule here; the sample's wall time on the CPU is
        # set by its largest tree, as for the reference's OpenMP over trees.
        out["c5_sample"] = _forest_run(physics, table, tol_abs, tol_rel, args.c5_trees, (11.0, 14.0), 8.0,
                                       f"{args.c5_trees} of the 4096 PCH trees of the large-forest workload, seed 0, "
                                       f"roots log-uniform 1e11-1e14 Msun, resolution 1e9 Msun, z=8 -> 0", threads,
                                       cpu=not args.no_cpu, waves=False)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--nodes", type=int, default=100000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--streams", default="auto",
                    help="independent batches alternate over this many streams and evolver handles, or 'auto' "
                         "for one per timed step (up to 20)")
    ap.add_argument("--no-plan", action="store_true",
                    help="order each overlapped batch's claims inside its own evolve call instead of planning all "
                         "batches before the evolves launch")
    ap.add_argument("--launch", default="auto",
                    help="launch shape 'nodes_per_workgroup,workgroups_per_cu[,rhs_waves]', or 'auto' to time them")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-surrogate", action="store_true", help="skip the surrogate training/inference legs")
    ap.add_argument("--no-trees", action="store_true", help="skip the merger-tree leg")
    ap.add_argument("--tree-splits", default="",
                    help="comma-separated splits of a tree workload into concurrent forests, F (F forests on F "
                         "streams) or FxS (F forests on S streams); the fastest is reported")
    ap.add_argument("--large-trees", type=int, default=256,
                    help="trees in the larger forest of the tree leg (0 skips it)")
    ap.add_argument("--c5-stream-trees", type=int, default=0,
                    help="trees of the large-forest workload streamed as consecutive forests (0 skips it; 4096 is "
                         "the whole workload)")
    ap.add_argument("--c5-forest-trees", type=int, default=64, help="trees per forest of the streamed leg")
    ap.add_argument("--c5-walkers", type=int, default=3,
                    help="forests of the streamed leg built and walked on the host at once, ahead of the GPU")
    ap.add_argument("--c5-trees", type=int, default=16,
                    help="trees of the 4096-tree large-forest workload (roots 1e11-1e14 Msun, z=8) run as a sample "
                         "(0 skips it)")
    ap.add_argument("--train-pairs", type=int, default=16000000)
    ap.add_argument("--train-epochs", type=int, default=30)
    ap.add_argument("--train-seeds", type=int, default=3,
                    help="networks trained from different seeds; the one rejecting fewest validation intervals is kept")
    ap.add_argument("--val-pairs", type=int, default=262144)
    ap.add_argument("--eval-pairs", type=int, default=1000000)
    ap.add_argument("--pipeline-streams", default="4,8",
                    help="stream counts tried for each path of the pipelined surrogate comparison")
    ap.add_argument("--pipeline-batches", type=int, default=16,
                    help="evaluation batches per path in the pipelined surrogate comparison")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="run one elementwise copy of this many MiB first (reference kernel for HBM counters)")
    args = ap.parse_args()
    args.tree_splits = parse_tree_splits(args.tree_splits)
    args.pipeline_streams = tuple(int(k) for k in str(args.pipeline_streams).split(",") if k.strip())
    if args.gpus != 1:
        print(f"bench.py measures one GPU. Across GPUs the engine would shard by whole node batches or whole "
              f"forests with no exchange during evolution; that path is not built. --gpus {args.gpus} is not "
              f"supported.", file=sys.stderr)
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
    # The tree leg runs first, on the null stream, before the integrator leg
    # gives the process its 20 streams: persistent tree launches slow down
    # 1.5-1.8x once a process holds more than about 24 hardware queues.
    trees_out = None if args.no_trees else trees_leg(physics, table, tol_abs, tol_rel, args)
    batch = wl.isolated_halos(args.nodes, seed=args.seed, params=physics)
    n = batch.n
    node_gyr = batch.node_gyr
    # Every timed batch in flight at once (up to 20): each batch's tail -- the
    # few stiffest nodes, whose chains set the launch time -- then overlaps
    # the others instead of queueing behind a batch on the same stream.  More
    # streams than about 24 hardware queues time-slice: 40 batches on 32
    # streams took 21-25 ms per batch against 11-13 ms on 20.
    n_streams = min(max(1, args.steps), 20) if args.streams == "auto" else max(1, int(args.streams))
    evs = [NodeEvolverBatched(n, physics, table, odeToleranceAbsolute=tol_abs, odeToleranceRelative=tol_rel,
                              reuseODEStepSize=True) for _ in range(n_streams)]
    ev = evs[0]
    dev = torch.device("cuda:0")
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    y0, npar, t0, t1, h0 = d(batch.y), d(batch.node_par), d(batch.t_start), d(batch.t_end), d(batch.h_step)
    if args.copy_calibration_mib > 0:
        n_el = args.copy_calibration_mib * (1 << 20) // 8
        src = torch.ones(n_el, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        torch.mul(src, 1.0, out=dst)          # reads and writes n_el * 8 bytes
        torch.cuda.synchronize()
        log(f"copy calibration: {n_el * 8} bytes read and written by one elementwise kernel")
        del src, dst
    # One stream per evolver handle, created once: the runtime maps streams
    # to its hardware queues in creation order, so tuning and the timed region
    # share the same queues.
    streams = stream_pool(n_streams)
    tuning = {}
    plan_ahead = not args.no_plan

    def overlapped_round(items, before=None, after=None):
        """One batch per stream, all in flight together: items are
        (stream index, y, h_step, scale, status, counters).  With plan_ahead
        every batch's claim order is computed (gal_ode_batch_plan_d) before
        any persistent evolve launches, so that no batch's ordering kernels
        wait for CUs that other batches' evolves hold; the evolves then wait
        for all the plans.  before/after(k) record events around item k."""
        if plan_ahead:
            done = []
            for k, (j, y, h, sc, st, cnt) in enumerate(items):
                if before:
                    before(k)
                evs[j].plan_device(y, npar, t0, t1, scale=sc, stream=streams[j])
                e = torch.cuda.Event()
                e.record(streams[j])
                done.append(e)
            for k, (j, y, h, sc, st, cnt) in enumerate(items):
                for e in done:
                    streams[j].wait_event(e)
                evs[j].evolve_device(y, npar, t0, t1, h, scale=sc, status=st, counters=cnt, stream=streams[j],
                                     planned=True)
                if after:
                    after(k)
        else:
            for k, (j, y, h, sc, st, cnt) in enumerate(items):
                if before:
                    before(k)
                evs[j].evolve_device(y, npar, t0, t1, h, scale=sc, status=st, counters=cnt, stream=streams[j])
                if after:
                    after(k)

    shapes = ((64, 1, 1), (64, 2, 1), (64, 4, 1), (64, 1, 2), (64, 2, 2))
    if args.launch != "auto":
        v = tuple(int(x) for x in args.launch.split(","))
        shapes = (v if len(v) == 3 else v + (1,),)
    if args.launch == "auto":
        single = ev.autotune(y0, npar, t0, t1, h0)
        tuning["single_launch_choice"] = list(single)
    if len(shapes) > 1:
        # The library's autotuner times single launches; the timed region
        # overlaps independent batches on n_streams streams, where a shape
        # with more resident lanes can win although it loses alone.  Time
        # each shape on one round of n_streams concurrent batches and keep
        # the fastest per batch.
        best = None
        for shape in shapes:
            for e in evs:
                e.set_launch(*shape)
            yt = [y0.clone() for _ in range(n_streams)]
            ht = [h0.clone() for _ in range(n_streams)]
            torch.cuda.synchronize()
            t_a = time.perf_counter()
            overlapped_round([(j, yt[j], ht[j], None, None, None) for j in range(n_streams)])
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t_a) * 1e3 / n_streams
            tuning[f"{shape[0]},{shape[1]},{shape[2]}"] = ms
            if best is None or ms < best[0]:
                best = (ms, shape)
            del yt, ht
        launch = best[1]
    else:
        launch = shapes[0]
    for e in evs:
        e.set_launch(*launch)
    log(f"launch: {launch[0]} nodes per workgroup, {launch[1]} workgroups per CU, {launch[2]} waves per group, "
        f"{n_streams} streams (tuning {tuning})")
    total = args.warmup + args.steps
    ys = [y0.clone() for _ in range(total)]
    hs = [h0.clone() for _ in range(total)]
    # Consecutive steps are independent batches: they alternate over n_streams
    # evolver handles and streams, so one batch's tail (the few stiffest
    # nodes) overlaps the next batch's start.
    scales = [torch.empty_like(y0) for _ in range(n_streams)]
    statuses = [torch.zeros(n, dtype=torch.int32, device=dev) for _ in range(n_streams)]   # unused streams stay 0
    counters = [torch.zeros(3 * n, dtype=torch.int64, device=dev) for _ in range(n_streams)]
    torch.cuda.synchronize()

    def launch_step(i, k):
        j = k % n_streams
        evs[j].evolve_device(ys[i], npar, t0, t1, hs[i], scale=scales[j], status=statuses[j], counters=counters[j],
                             stream=streams[j])
        return j

    for i in range(args.warmup):
        launch_step(i, i)
    torch.cuda.synchronize()
    log(f"warmup done ({args.warmup} steps)")
    # One batch alone on one stream: the rate without any overlap of batches.
    single_ms = []
    for rep in range(3):
        ys1, hs1 = y0.clone(), h0.clone()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(streams[0])
        evs[0].evolve_device(ys1, npar, t0, t1, hs1, scale=scales[0], status=statuses[0], counters=counters[0],
                             stream=streams[0])
        e1.record(streams[0])
        torch.cuda.synchronize()
        single_ms.append(e0.elapsed_time(e1))
        del ys1, hs1
    single_launch_ms = float(np.median(single_ms))
    log(f"one batch alone: {single_launch_ms:.3f} ms ({node_gyr / (single_launch_ms * 1e-3):.4g} node-Gyr/s)")

    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_wall = time.perf_counter()
    # rounds of one batch per stream (one round when steps <= streams)
    for r0 in range(0, args.steps, n_streams):
        ks = list(range(r0, min(args.steps, r0 + n_streams)))
        overlapped_round([(k % n_streams, ys[args.warmup + k], hs[args.warmup + k], scales[k % n_streams],
                           statuses[k % n_streams], counters[k % n_streams]) for k in ks],
                         before=lambda m, ks=ks: starts[ks[m]].record(streams[ks[m] % n_streams]),
                         after=lambda m, ks=ks: ends[ks[m]].record(streams[ks[m] % n_streams]))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall
    kernel_ms = float(np.mean([s.elapsed_time(e) for s, e in zip(starts, ends)]))
    ms_per_step = wall * 1e3 / args.steps
    for st in statuses:
        st = st.cpu().numpy()
        if not (st == 0).all():
            log(f"non-success statuses: {np.unique(st, return_counts=True)}")
            sys.exit(1)
    cnt = counters[0].view(3, n).cpu().numpy()
    attempts = int(cnt[0].sum())
    rhs_evals = int(cnt[2].sum())
    # algorithmic FLOPs of a launch from its measured counters: F_rhs per RHS
    # evaluation (rejected attempts take five, accepted six) and F_comb per attempt
    flop_launch = gflops.evaluated_flops(rhs_evals, attempts)
    state_bytes = 8 * (gp.NVAR + npar.shape[0] + 3) + 8 * (2 * gp.NVAR + 1) + 4 + 3 * 8
    achieved_tflops = flop_launch / (kernel_ms * 1e-3) / 1e12
    aggregate_tflops = flop_launch * args.steps / wall / 1e12
    log(f"step {ms_per_step:.3f} ms (one launch {kernel_ms:.3f} ms, {n_streams} streams), attempts {attempts}, "
        f"max/node {cnt[0].max()}")

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
            "dt_gyr": "uniform 0.1-2", "cooling_table": "200x50 FP64", "parallelism": ("single GPU, one node per lane"
                                                                   + (", the RHS split over two waves"
                                                                      if launch[2] == 2 else "")),
            "launch": {"nodes_per_workgroup": launch[0], "workgroups_per_cu": launch[1], "rhs_waves": launch[2],
                       "streams": n_streams,
                       "claim_order": ("planned for every batch of a round before its evolves launch" if plan_ahead
                                       else "inside each evolve call"),
                       "tuning_ms_per_step": tuning},
        },
        "single_launch": {
            "ms": single_launch_ms,
            "value": node_gyr / (single_launch_ms * 1e-3),
            "unit": "node-Gyr/s",
            "note": "one batch alone on one stream (median of 3); value overlaps independent batches on "
                    f"{n_streams} streams",
        },
        "roofline": {
            "bound": "fp64-valu",
            "compute": "FP64 vector ALU (the MI355X FP64 vector peak, 78.6 TFLOP/s)",
            "achieved": achieved_tflops,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / FP64_PEAK_TFLOPS,
            "traffic": HBM_TRAFFIC_1E5.get(tuple(launch)) if n == 100000 and args.seed == 0 else None,
            "traffic_unit": "bytes per launch (rocprofv3 counters, calibrated)",
            "flop_per_rhs_evaluation": gflops.rhs_flops(),
            "flop_per_attempt_combination": gflops.combination_flops(),
            "flop_per_launch": flop_launch,
            "attempts_per_launch": attempts,
            "accepted_steps_per_launch": int(cnt[1].sum()),
            "rhs_evaluations_per_launch": int(cnt[2].sum()),
            "rhs_evaluations_per_s": int(cnt[2].sum()) * args.steps / wall,
            "accepted_steps_per_s": int(cnt[1].sum()) * args.steps / wall,
            "kernel_ms": kernel_ms,
            "aggregate_tflops": aggregate_tflops,
            # node state in and out of HBM: y, node parameters, t0, t1, h in;
            # y, h, scales, status, counters out
            "state_bytes_per_node": state_bytes,
            "state_gbs": state_bytes * n / (kernel_ms * 1e-3) / 1e9,
            "hbm_peak_gbs": 8000.0,
        },
    }
    del ys, hs
    if not args.no_surrogate:
        result["surrogate"] = surrogate_leg(physics, table, tol_abs, tol_rel, args)
    if trees_out is not None:
        if not args.no_surrogate and args.large_trees > 0:
            trees_out["large"]["surrogate"] = surrogate_forest_route(args.trained_surrogate, physics, table, tol_abs,
                                                                     tol_rel, args.large_trees)
        result["trees"] = trees_out
    if not args.no_cpu:
        cb = cpu_baseline(batch, physics, table, tol_abs, tol_rel, gp.FLAG_REUSE_STEP_SIZE, args.cpu_seconds)
        result["cpu_baseline"] = cb
        log(f"cpu baseline {cb['value']:.4g} node-Gyr/s on {cb['cores']} cores; GPU/CPU = {result['value']/cb['value']:.1f}")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
