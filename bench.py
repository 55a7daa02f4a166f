"""Benchmark of the M1 + chemistry radiation step on one MI355X.

Workloads (seeded synthetic states; see emma_amd/synth.py):
  c3 (default) 128^3 periodic level (level 7, 2,097,152 cells), 3 photon groups,
               64 point sources, lognormal n_H; mixed precision.  A step is one
               radiative sub-cycle of the level; the K timed sub-cycles are one
               replay of their captured graph (captured and run once on the
               same starting state, untimed).
  c2           64^3 periodic level, 1 group; FP64 plain.  A step is one sub-cycle.
  c5           256^3 coarse level (level 8) + 300000/75000 refined octs, 3 groups,
               4096 sources, c/100; timed like c4.
  c4           AMR levels 7-11 with 262144/80000/25000/6000/1200 octs, 3 groups,
               mixed precision.  A step is one coarse step of the RADSTEP level
               recursion (level l takes 2^(l-7) sub-steps, 31 launches), replayed as
               one HIP graph from the state after the first coarse step (run_amr).

Prints ONE JSON line.  value = leaf cell-updates/s of the timed region (state
resident in HBM; the library call is synchronous, so the wall clock brackets
complete device work).  roofline.achieved = algorithmic bytes of the timed
level-step launches / their device time, taken from the same graph replays run
once more right after the timed region between HIP events on the library's
stream (the events add host overhead, so the value's replays run without them;
the gaps between launches inside the graph count as kernel time);
mean_kernel_us = that device time / launches.  fp64_plain: the same
workload stepped by the FP64 kernel.  cpu_baseline: the reference's CPU path (oracle/_ref, compiled from the
reference's own sources) or, where that build is absent, the oracle (its CPU
restatement), on up to 16 host cores, one process each, on a bounded sample
drawn like the workload.

--stencil-path measures the drop-in path instead (advanceradGPU's contract: host
struct RGRID batches in, rfieldnew and fluxes back) on the 64^3, 1-group level in
batches of 8192 octs, --steps sweeps over the level after --warmup; its line
carries the PCIe-inclusive rate under its own metric name.

--calibrate-copy runs a plain device copy of known size (for scaling the
FETCH_SIZE / WRITE_SIZE counters of rocprofv3 --pmc runs).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def alg_bytes_per_cell(G: int, precision: str) -> int:
    """Reads N,F (4G), src (G), n_H, x, eint; writes N,F (4G), x, eint, T."""
    return (72 * G + 48) if precision == "fp64" else (36 * G + 48)


def make_workload(name: str, seed: int):
    from emma_amd import synth

    if name == "c3":
        return synth.config_c3(seed=seed, nsub=1)
    if name == "c2":
        return synth.config_c2(seed=seed)
    if name == "c4":
        return synth.config_c4(seed=seed, nsub=1)
    if name == "c5":
        return synth.config_c5(seed=seed, nsub=1)
    raise SystemExit(f"unknown workload {name}")


WORKLOAD_NAMES = {
    "c3": "128^3 periodic level (2,097,152 cells), 3 groups, 64 sources, lognormal n_H",
    "c2": "64^3 periodic level (262,144 cells), 1 group, lognormal n_H",
    "c4": "AMR levels 7-11 (262144/80000/25000/6000/1200 octs), 3 groups, lognormal sigma=1.5",
    "c5": "256^3 coarse level + 2 refined levels (300000/75000 octs), 3 groups, 4096 sources, c/100, lognormal sigma=2",
}


def _cpu_worker(args):
    """One core's share of the CPU baseline: RadSolver sub-cycles on a level-5
    sample (32^3 cells) drawn like the workload, restarted from the drawn state
    every 4 sub-cycles (without sources the sample drains to the EMIN floor,
    where the reference aborts, chem_utils.c:382).  The reference's own CPU
    functions when their build is present, else the oracle."""
    name, seconds, seed, use_ref = args
    from emma_amd import synth
    from oracle import Oracle, Reference

    if name in ("c3", "c4"):
        W = synth.uniform_workload(5, "Starburst99_1grpT_3grpS", seed=seed, nsrc=1)
    else:
        W = synth.uniform_workload(5, "SB99_equiv_mono_TH", seed=seed)
    s = Reference(W.params.ngrp) if use_ref else Oracle()
    dt = W.params.dt_rad(5)
    n, el = 0, 0.0
    while el < seconds:
        s.setup(W.params, W.tree, W.fields)
        t0 = time.perf_counter()
        for _ in range(4):
            s.clean_new(5)
            s.radsolver(5, dt)
        el += time.perf_counter() - t0
        n += 4
    return W.tree.level(5).ncell * n, el


def cpu_baseline(name: str, seconds: float = 10.0, cores: int = 0):
    """The reference CPU path (oracle/_ref, compiled from the reference's own
    sources) or, where that build is absent, the oracle, on `cores` host
    cores at once (one process each, spawned)."""
    import multiprocessing as mp

    from oracle import ref_available

    G = 1 if name == "c2" else 3
    use_ref = ref_available(G)
    if cores <= 0:
        cores = max(1, min(16, os.cpu_count() or 1))
    ctx = mp.get_context("spawn")
    try:
        with ctx.Pool(cores) as pool:
            res = pool.map(_cpu_worker, [(name, seconds, 3 + k, use_ref) for k in range(cores)])
    except Exception as exc:  # the reference build could not run here: time the oracle
        if not use_ref:
            raise
        print(f"cpu_baseline: reference build failed ({exc}); timing the oracle", file=sys.stderr)
        use_ref = False
        with ctx.Pool(cores) as pool:
            res = pool.map(_cpu_worker, [(name, seconds, 3 + k, use_ref) for k in range(cores)])
    rate = sum(c / e for c, e in res)
    nsub = sum(c for c, _ in res) // 32768
    return {
        "value": rate,
        "unit": "cell-updates/s",
        "cores": cores,
        "kind": "reference" if use_ref else "port",
        "sample": f"{'reference' if use_ref else 'oracle'} RadSolver (clean_new_rad + RadSolver) on a 32^3 periodic "
                  f"level (level 5, 32768 cells) drawn like the workload, one process per core, {nsub} sub-cycles "
                  f"in {seconds:.0f} s per core",
    }


def load_traffic(key: str):
    p = os.path.join(ROOT, "profiles", f"level_step_{key}_traffic.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f).get("bytes_per_launch")
    return None


def level_bytes(W, precision):
    """algorithmic bytes of one step of each level (leaf cells x bytes/cell)"""
    Bc = alg_bytes_per_cell(W.params.ngrp, precision)
    return {l: W.tree.leaf_cells(l) * Bc for l in range(W.tree.lmin, W.tree.lmax + 1)}


def run_uniform(W, precision, steps, warmup):
    from emma_amd import rt

    lvl = W.tree.lmin
    g = rt.RadGPU(W.params, rt.MIXED if precision == "mixed" else rt.FP64)
    g.setup(W.params, W.tree, W.fields)
    dt = W.params.dt_rad(lvl)
    for _ in range(warmup):
        g.radsolver(lvl, dt)
    # capture (and run) the graph of `steps` sub-cycles, then restore the
    # state it started from into the buffer the graph starts from: the timed
    # call replays that graph on the same state (an odd `steps` flips the
    # buffer parity, so the graph of the other parity is captured as well)
    start = g.get(lvl)
    g.coarse_step(1.0, steps)
    if steps % 2:
        g.coarse_step(1.0, steps)
    g.set(lvl, 0, start)
    t0 = time.perf_counter()
    g.coarse_step(1.0, steps)  # single level: `steps` sub-cycles, one launch each
    wall = time.perf_counter() - t0
    st = g.stats()
    # the same replay again, between HIP events, for the device time
    g.set(lvl, 0, start)
    g.enable_timing(rt.TIMING_GRAPH)
    g.coarse_step(1.0, steps)
    kernel_s = g.stats()["kernel_ms"] * 1e-3
    bytes_total = level_bytes(W, precision)[lvl] * st["level_steps"]
    return wall, st, kernel_s, bytes_total, st["level_steps"]


def run_amr(W, precision, steps, warmup):
    """Every timed coarse step is the second coarse step of the run: on this
    seeded input dense neutral cells drain to the EMIN floor and the reference
    path aborts in the third coarse step (chem_utils.c:382-385; the oracle
    reproduces it), so the state after the first coarse step (taken from the
    drawn state, untimed) is restored (host -> device, outside the timed region)
    before every timed coarse step.  A coarse step is one replay of the captured
    graph."""
    from emma_amd import rt

    g = rt.RadGPU(W.params, rt.MIXED if precision == "mixed" else rt.FP64)
    g.setup(W.params, W.tree, W.fields)
    g.coarse_step(1.0, 1)  # first coarse step, from the drawn state
    start = {l: g.get(l) for l in W.fields}

    def restore():
        for l, F in start.items():
            g.set(l, 0, F)

    # capture + instantiate the graphs of both buffer parities of the coarse level
    for _ in range(max(warmup, 2)):
        restore()
        g.coarse_step(1.0, 1)
    wall = 0.0
    tot = {"cell_updates": 0, "chem_substeps": 0, "level_steps": 0}
    for _ in range(steps):
        restore()
        t0 = time.perf_counter()
        g.coarse_step(1.0, 1)  # timed: one graph replay
        wall += time.perf_counter() - t0
        st = g.stats()
        for k in tot:
            tot[k] += st[k]
    # the same replays again, between HIP events, for the device time
    g.enable_timing(rt.TIMING_GRAPH)
    kernel_ms = 0.0
    for _ in range(steps):
        restore()
        g.coarse_step(1.0, 1)
        kernel_ms += g.stats()["kernel_ms"]
    lb = level_bytes(W, precision)
    per_coarse = sum(b << (l - W.tree.lmin) for l, b in lb.items())
    return wall, tot, kernel_ms * 1e-3, per_coarse * steps, tot["level_steps"]


def stencil_path(steps: int, warmup: int, stride: int = 8192) -> dict:
    """The drop-in path on the 64^3 level: every oct gathered on the host into
    struct RGRID batches of `stride` octs (untimed), then per sweep over the
    level one emma_rt_advance_stencil call per batch (host -> device copy,
    sweeps + update + chemistry, device -> host copy of the New part), timed
    by the wall clock.  The batch handed over lives in page-locked memory
    (emma_rt_create_pinned_stencil, as EMMA's create_pinned_stencil_rad)."""
    import numpy as np
    from emma_amd import rgrid, rt, synth

    W = synth.config_c2()
    lvl = W.tree.lmin
    L, F = W.tree.level(lvl), W.fields[lvl]
    batches = [rgrid.gather_complete(L, F, np.arange(o, min(L.noct, o + stride)), W.params.ngrp)
               for o in range(0, L.noct, stride)]
    g = rt.RadGPU(W.params, rt.FP64)
    g.create_radstencil(stride, 8)
    pinned = g.pinned_stencil(stride, 8)
    dt = W.params.dt_rad(lvl)

    def sweep(timed):
        t = 0.0
        for b in batches:
            w = pinned.array[: b.shape[0]]
            w[...] = b  # every sweep starts from the gathered state
            t0 = time.perf_counter()
            g.advance_stencil(w, w.shape[0], 0.5**lvl, dt)
            t += time.perf_counter() - t0
        return t

    for _ in range(warmup):
        sweep(False)
    wall = sum(sweep(True) for _ in range(steps))
    pinned.free()
    cells = L.ncell * steps
    # what crosses PCIe per oct: the 7 oct slots and the centre cells' rfieldnew
    # to the device, the centre cells' rfieldnew and rflux back
    dtp = rgrid.rgrid_dtype(W.params.ngrp)
    cellflux = dtp["New"].base
    rtype = cellflux["rfieldnew"].itemsize
    moved = (int(dtp["oct"].itemsize + 8 * rtype), int(8 * (rtype + cellflux["rflux"].itemsize)))
    return {
        "metric": "drop-in stencil path cell-updates/s (host RGRID batches, PCIe-inclusive)",
        "value": cells / wall,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": wall * 1e3 / steps,
        "higher_is_better": True,
        "dtype": "fp64",
        "data": "synthetic",
        "config": {"workload": WORKLOAD_NAMES["c2"] + ", struct RGRID batches of %d octs" % stride,
                   "bytes_per_oct_to_device": moved[0], "bytes_per_oct_to_host": moved[1]},
    }


def copy_calibration(nbytes: int = 1 << 30, reps: int = 10):
    import torch

    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda:0").fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "reps": reps,
                      "GBps": 2 * nbytes * reps / el / 1e9}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=32)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--workload", default="c3", choices=["c3", "c2", "c4", "c5"])
    ap.add_argument("--precision", default=None, choices=["mixed", "fp64"])
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--cpu-cores", type=int, default=0, help="host cores for the CPU baseline (0: up to 16)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64-plain", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--stencil-path", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library runs the radiation step on one GPU's shard "
              "(the multi-GPU halo exchange is not part of it)", file=sys.stderr)
        sys.exit(2)
    if args.stencil_path:
        print(json.dumps(stencil_path(args.steps, args.warmup)))
        return
    if args.calibrate_copy:
        copy_calibration()
        return

    precision = args.precision or ("fp64" if args.workload == "c2" else "mixed")
    W = make_workload(args.workload, args.seed)
    runner = run_amr if args.workload in ("c4", "c5") else run_uniform
    wall, st, kernel_s, bytes_per_unit, launches = runner(W, precision, args.steps, args.warmup)
    cell_updates = st["cell_updates"]
    value = cell_updates / wall
    # device time of the timed graph replays (HIP events on the library's
    # stream around each replay) over their algorithmic bytes and launches
    achieved = bytes_per_unit / kernel_s / 1e9
    mean_launch_us = kernel_s / max(launches, 1) * 1e6
    out = {
        "metric": "M1+chem cell-updates/s (mixed+batched vs FP64 plain, 1 GPU) and % HBM peak",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64" if precision == "mixed" else "fp64",
        "data": "synthetic",
        "config": {
            "workload": WORKLOAD_NAMES[args.workload],
            "cells": sum(W.tree.level(l).ncell for l in range(W.tree.lmin, W.tree.lmax + 1)),
            "groups": W.params.ngrp,
            "precision": precision,
            "step": ("second coarse step (31 level steps, HIP graph)" if args.workload == "c4"
                     else "second coarse step (7 level steps, HIP graph)" if args.workload == "c5"
                     else "radiative sub-cycle"),
            "chem_substeps_per_cell_update": st["chem_substeps"] / max(cell_updates, 1),
            "mean_kernel_us": mean_launch_us,
            "parallelism": "1 GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": load_traffic(f"{args.workload}_{precision}"),
            "alg_bytes_per_cell_update": alg_bytes_per_cell(W.params.ngrp, precision),
        },
    }
    if precision == "mixed" and not args.no_fp64_plain:
        nsteps = max(4, args.steps // 4)
        w2, st2, _, _, _ = runner(W, "fp64", nsteps, args.warmup)
        v2 = st2["cell_updates"] / w2
        out["fp64_plain"] = {"value": v2, "unit": "cell-updates/s", "steps": nsteps, "speedup_of_value": value / v2}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.workload, args.cpu_seconds, args.cpu_cores)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
