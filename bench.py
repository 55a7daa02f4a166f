"""Benchmark of the M1 + chemistry radiation step on one MI355X.

A step is one radiative sub-cycle of the workload's level (every leaf cell
advanced once: transport, update and chemistry).  Default workload: 128^3
periodic level (level 7, 2,097,152 cells), 3 photon groups, 64 point
sources, seeded lognormal density; mixed precision (FP32 photon density and
flux, FP64 chemistry).  --steps sub-cycles run in the timed region.

Prints ONE JSON line (see the keys below).  value = cell-updates/s of the
whole job with the state already resident in HBM.  roofline.achieved is the
algorithmic bytes of one launch of the level-step kernel divided by its mean
device time (HIP events around every launch on the library's stream).
cpu_baseline times the oracle (the CPU restatement of the reference path,
one core) on a bounded sample of the same workload.
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
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(name: str, seconds: float = 10.0):
    """Oracle (scalar C restatement, 1 core) on a level-5 sample (32^3 cells)
    drawn like the workload; sub-cycles repeat until `seconds` have passed."""
    from emma_amd import synth
    from oracle import Oracle

    if name == "c3":
        W = synth.uniform_workload(5, "Starburst99_1grpT_3grpS", seed=3, nsrc=1)
    else:
        W = synth.uniform_workload(5, "SB99_equiv_mono_TH", seed=2)
    o = Oracle()
    dt = W.params.dt_rad(5)
    n = 0
    el = 0.0
    # restart from the drawn state every 4 sub-cycles: without sources the
    # sample drains to the EMIN floor, where the reference aborts (chem_utils.c:382)
    while el < seconds:
        o.setup(W.params, W.tree, W.fields)
        t0 = time.perf_counter()
        for _ in range(4):
            o.radsolver(5, dt)
        el += time.perf_counter() - t0
        n += 4
    cells = W.tree.level(5).ncell
    return {
        "value": cells * n / el,
        "unit": "cell-updates/s",
        "cores": 1,
        "kind": "port",
        "sample": f"oracle RadSolver on a 32^3 periodic level (level 5, {cells} cells) drawn like the workload, "
                  f"{n} sub-cycles in {el:.1f} s",
    }


def load_traffic(workload: str):
    p = os.path.join(ROOT, "profiles", f"level_step_{workload}_traffic.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f).get("bytes_per_launch")
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=32)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--workload", default="c3", choices=["c3", "c2"])
    ap.add_argument("--precision", default=None, choices=["mixed", "fp64"])
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the radiation step is measured on one GPU's shard "
              "(multi-GPU halo exchange is not part of this build)", file=sys.stderr)
        sys.exit(2)

    from emma_amd import rt

    precision = args.precision or ("mixed" if args.workload == "c3" else "fp64")
    W = make_workload(args.workload, args.seed)
    lvl = W.tree.lmin
    G = W.params.ngrp
    ncell = W.tree.level(lvl).ncell
    g = rt.RadGPU(W.params, rt.MIXED if precision == "mixed" else rt.FP64)
    g.setup(W.params, W.tree, W.fields)
    dt = W.params.dt_rad(lvl)
    for _ in range(args.warmup):
        g.radsolver(lvl, dt)
    # timed region: K sub-cycles, HIP events around every level-step launch
    g.enable_timing(True)
    t0 = time.perf_counter()
    g.coarse_step(1.0, args.steps)
    t1 = time.perf_counter()
    st = g.stats()
    wall = t1 - t0
    cell_updates = st["cell_updates"]
    value = cell_updates / wall
    Bc = alg_bytes_per_cell(G, precision)
    launches = st["level_steps"]
    mean_ms = st["kernel_ms"] / max(launches, 1)
    achieved = Bc * ncell / (mean_ms * 1e-3) / 1e9
    out = {
        "metric": "M1+chem cell-updates/s (1 GPU) and % HBM peak",
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
            "workload": "128^3 periodic level, 3 groups, 64 sources, lognormal n_H (C3)" if args.workload == "c3"
            else "64^3 periodic level, 1 group (C2)",
            "cells": ncell,
            "groups": G,
            "precision": precision,
            "substeps_per_cell_update": st["chem_substeps"] / max(cell_updates, 1),
            "kernel_ms_per_step": mean_ms,
            "parallelism": "1 GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": load_traffic(f"{args.workload}_{precision}"),
            "alg_bytes_per_cell_update": Bc,
        },
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.workload, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
