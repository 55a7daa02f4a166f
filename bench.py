"""Benchmark of the M1 + chemistry radiation step on one MI355X.

Workloads (seeded synthetic states; see emma_amd/synth.py):
  c5 (default) reionisation-like box: 256^3 coarse level (level 8, 16.8 M cells)
               + 300000/75000 refined octs (levels 9-10), lognormal sigma=2,
               3 groups, 4096 sources with dN/dL ~ L^-2, c/100; mixed precision.
               A step is one coarse step of the RADSTEP level recursion (level l
               takes 2^(l-8) level steps, 7 launches), one replay of its HIP
               graph, from the state after the first coarse step (run_amr).
  c4           AMR levels 7-11 with 262144/80000/25000/6000/1200 octs, 3 groups,
               mixed precision; a step is one coarse step (31 launches), timed
               like c5.  The c5 line carries a c4 block as well.
  c3           128^3 periodic level (level 7, 2,097,152 cells), 3 groups, 64
               sources; a step is one radiative sub-cycle, the K timed
               sub-cycles one replay of their captured graph.
  c2           64^3 periodic level, 1 group; FP64 plain.

Prints ONE JSON line.  value = leaf cell-updates/s of the timed region (state
resident in HBM; the library call is synchronous, so the wall clock brackets
complete device work).  roofline = the dominant level-step kernel (the level
with the most device time per step): algorithmic bytes of one launch (leaf
cells x bytes per cell-update) over its mean launch duration, measured right
after the timed region with HIP events around every kernel on the library's
stream (coarse steps launched eagerly for this); config.device_ms_per_step is
the device time of one graph replay between a pair of HIP events, and
config.step_alg_frac the whole step's algorithmic bytes over it.  fp64_plain:
the same workload stepped by the FP64 kernel.  cpu_baseline: the reference's
CPU path (oracle/_ref, compiled from the reference's own sources) or, where that
build is absent, the oracle (its CPU restatement), one process per host core
of the job's CPU share, on a bounded sample drawn like the workload.

--stencil-path measures the drop-in path instead (advanceradGPU's contract: host
struct RGRID batches in, rfieldnew and fluxes back) on the 64^3, 1-group level in
batches of 8192 octs, --steps sweeps over the level after --warmup; its line
carries the PCIe-inclusive rate under its own metric name.

--calibrate-copy runs a plain device copy kernel of known size (for scaling the
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

    if name in ("c3", "c4", "c5"):
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


def host_cores() -> int:
    """Host cores this job may use: the thread count the machine allots to
    the job (OMP_NUM_THREADS, which the GPU machines set to the job's CPU
    share) or else the CPUs the process may run on (its affinity mask);
    os.cpu_count() counts the whole host, which other jobs share."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except AttributeError:
        return max(1, os.cpu_count() or 1)


def cpu_baseline(name: str, seconds: float = 10.0, cores: int = 0):
    """The reference CPU path (oracle/_ref, compiled from the reference's own
    sources) or, where that build is absent, the oracle, on `cores` host
    cores at once (one process each, spawned)."""
    import multiprocessing as mp

    from oracle import ref_available

    G = 1 if name == "c2" else 3
    use_ref = ref_available(G)
    if cores <= 0:
        cores = host_cores()
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
        "per_core": rate / cores,
        "host_cpus_visible": os.cpu_count(),
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


def _dominant(W, precision, level_us, launches_per_step):
    """The level-step kernel with the most device time per step: its level,
    mean launch duration and roofline (leaf cells x algorithmic bytes per
    cell-update over that duration)."""
    lb = level_bytes(W, precision)
    per_step = {l: level_us[l] * launches_per_step[l] for l in level_us}
    l = max(per_step, key=per_step.get)
    achieved = lb[l] / (level_us[l] * 1e-6) / 1e9
    return l, achieved


KERNEL_NAMES = {"listed": "emma::level_step_tiled<{G}, false>", "perdir": "emma::level_step_tiled<{G}, true>",
                "latency": "emma::level_step_tiled_latency<{G}>"}


def _kernel_name(W, level, precision, tuning):
    if precision != "mixed" or W.params.ngrp not in (1, 3):
        return f"emma::level_step_kernel<{'double' if precision == 'fp64' else 'float'}, {W.params.ngrp}, ...>"
    t = tuning[level]
    name = KERNEL_NAMES[t["variant"]].format(G=W.params.ngrp)
    return name + (" + emma::chem_straggler_kernel<%d>" % W.params.ngrp if t["chem_budget"] > 0 else "")


def _configure(g, W, precision, tuning):
    """apply the tuning table (mixed precision) and report each level's
    variant and chemistry budget"""
    from emma_amd import rt, tune

    if tuning is not None and precision == "mixed":
        tune.apply(g, tuning)
    out = {}
    for l in range(W.tree.lmin, W.tree.lmax + 1):
        if precision == "mixed" and W.params.ngrp in (1, 3):
            t = g.level_tuning(l)
            out[l] = {"variant": rt.VARIANT_NAMES[t["variant"]], "chem_budget": t["chem_budget"],
                      "chem_keep": t["chem_keep"]}
    return out


def _level_table(W, precision, level_us, launches_per_step, tuning):
    lb = level_bytes(W, precision)
    return {str(l): dict({"leaf_cells": W.tree.leaf_cells(l), "launches_per_step": launches_per_step[l],
                          "us_per_launch": level_us[l],
                          "frac": lb[l] / (level_us[l] * 1e-6) / 1e9 / HBM_PEAK_GBS if level_us[l] > 0 else None},
                         **tuning.get(l, {}))
            for l in sorted(level_us)}


def run_uniform(W, precision, steps, warmup, tuning=None):
    """A step is one radiative sub-cycle of the level.  The K timed sub-cycles
    are one replay of their captured graph (captured and run once on the same
    starting state, untimed)."""
    from emma_amd import rt

    lvl = W.tree.lmin
    g = rt.RadGPU(W.params, rt.MIXED if precision == "mixed" else rt.FP64)
    g.setup(W.params, W.tree, W.fields)
    used = _configure(g, W, precision, tuning)
    dt = W.params.dt_rad(lvl)
    for _ in range(warmup):
        g.radsolver(lvl, dt)
    # capture (and run) the graph of `steps` sub-cycles, then restore the
    # state it started from into the buffer the graph starts from: the timed
    # call replays that graph on the same state (an odd `steps` flips the
    # buffer parity, so the graph of the other parity is captured as well)
    start = g.get(lvl)
    g.coarse_step(1.0, steps)
    st = g.stats()  # counted on the device (cell updates, chemistry sub-steps)
    _check_counted(st, W.leaf_cell_updates_per_coarse_step() * steps)
    # timed without the kernels' diagnostic counters (same results bit for
    # bit; their graphs are captured here, untimed)
    g.enable_counters(False)
    g.set(lvl, 0, start)
    g.coarse_step(1.0, steps)
    if steps % 2:
        g.coarse_step(1.0, steps)
    g.set(lvl, 0, start)
    t0 = time.perf_counter()
    g.coarse_step(1.0, steps)  # single level: `steps` sub-cycles, one launch each
    wall = time.perf_counter() - t0
    # the same replay again, between one pair of HIP events
    g.set(lvl, 0, start)
    g.enable_timing(rt.TIMING_GRAPH)
    g.coarse_step(1.0, steps)
    graph_ms = g.stats()["kernel_ms"]
    # and launched eagerly with events around every kernel (kernel durations)
    g.set(lvl, 0, start)
    g.enable_timing(rt.TIMING_LAUNCH)
    g.coarse_step(1.0, steps)
    ms, n = g.level_times()[lvl]
    g.close()
    return {"wall": wall, "stats": st, "graph_ms": graph_ms, "level_us": {lvl: ms / max(n, 1) * 1e3},
            "launches_per_step": {lvl: 1}, "steps": steps, "alg_bytes_per_step": level_bytes(W, precision)[lvl],
            "tuning": used}


def _check_counted(stats, expected):
    """the device's count of leaf cell updates is the host's (every leaf of
    every level step advanced)"""
    if stats["status"] != 0 or stats["cell_updates"] != expected:
        raise RuntimeError(f"device counted {stats['cell_updates']} leaf cell updates (status {stats['status']}), "
                           f"expected {expected}")


def run_amr(W, precision, steps, warmup, tuning=None, tune_here=False, log=None, schedule=None):
    """A step is one coarse step of the RADSTEP level recursion (level l takes
    2^(l-lmin) level steps), one replay of its captured HIP graph: per-level
    launches, or one persistent launch fed by a device work queue (the
    schedule the tuning table holds for the tree, or the tuner picks; the
    other schedule is timed after the timed region for the report).  Every
    timed coarse step is the second coarse step of the run: on these seeded
    inputs dense neutral cells drain to the EMIN floor, where the reference
    path aborts within a few coarse steps (chem_utils.c:382-385; the oracle
    reproduces it), so the state after the first coarse step (taken from the
    drawn state, untimed) is restored from its device-side copy
    (emma_rt_restore_state, outside the timed region) before every timed
    coarse step."""
    from emma_amd import rt

    g = rt.RadGPU(W.params, rt.MIXED if precision == "mixed" else rt.FP64)
    g.setup(W.params, W.tree, W.fields)
    used = _configure(g, W, precision, tuning)
    g.coarse_step(1.0, 1)  # first coarse step, from the drawn state
    g.save_state()
    restore = g.restore_state
    tuned = sched_tuned = None
    if tune_here and precision == "mixed" and W.params.ngrp in (1, 3):
        from emma_amd import tune

        tuned = tune.tune(g, reps=3, log=log)  # leaves the context set to the fastest per level
        used = _configure(g, W, precision, None)
        sched_tuned = tune.tune_schedule(g, reps=5, log=log)  # and on the faster schedule
    if schedule is not None:
        g.set_schedule({v: k for k, v in rt.SCHED_NAMES.items()}[schedule])
    sched = g.schedule()["sched"]

    # one coarse step counted on the device (cell updates, chemistry
    # sub-steps), untimed; the timed ones run without the kernels' diagnostic
    # counters (same results bit for bit)
    restore()
    g.coarse_step(1.0, 1)
    counted = g.stats()
    _check_counted(counted, W.leaf_cell_updates_per_coarse_step())
    g.enable_counters(False)
    # capture + instantiate the graphs of both buffer parities of the coarse level
    for _ in range(max(warmup, 2)):
        restore()
        g.coarse_step(1.0, 1)
    wall = 0.0
    for _ in range(steps):
        restore()
        t0 = time.perf_counter()
        g.coarse_step(1.0, 1)  # timed: one graph replay
        wall += time.perf_counter() - t0
        if g.stats()["status"] != 0:
            raise RuntimeError(f"coarse step status {g.stats()['status']}")
    tot = {k: counted[k] * steps for k in ("cell_updates", "chem_substeps", "level_steps")}
    # the same replays again, each between one pair of HIP events
    g.enable_timing(rt.TIMING_GRAPH)
    graph_ms = 0.0
    for _ in range(steps):
        restore()
        g.coarse_step(1.0, 1)
        graph_ms += g.stats()["kernel_ms"]
    sched_info = g.schedule()
    schedules = {sched: graph_ms / steps}
    # the other schedule, timed the same way (not the value)
    other = rt.SCHED_PERSISTENT if sched == "graph" else rt.SCHED_GRAPH
    g.enable_timing(rt.TIMING_OFF)
    g.set_schedule(other)
    for _ in range(2):
        restore()
        g.coarse_step(1.0, 1)
    g.enable_timing(rt.TIMING_GRAPH)
    oms = 0.0
    for _ in range(steps):
        restore()
        g.coarse_step(1.0, 1)
        oms += g.stats()["kernel_ms"]
    schedules[rt.SCHED_NAMES[other]] = oms / steps
    if other == rt.SCHED_PERSISTENT:
        sched_info = g.schedule()
    # and launched eagerly with events around every kernel (kernel durations;
    # per-level launches)
    g.enable_timing(rt.TIMING_OFF)
    g.set_schedule(rt.SCHED_GRAPH)
    g.enable_timing(rt.TIMING_LAUNCH)
    acc = {}
    nrep = max(1, min(steps, 5))
    for _ in range(nrep):
        restore()
        g.coarse_step(1.0, 1)
        for l, (ms, n) in g.level_times().items():
            a = acc.setdefault(l, [0.0, 0])
            a[0] += ms
            a[1] += n
    g.close()
    level_us = {l: a[0] / max(a[1], 1) * 1e3 for l, a in acc.items() if a[1] > 0}
    launches = {l: 1 << (l - W.tree.lmin) for l in level_us}
    lb = level_bytes(W, precision)
    per_coarse = sum(b << (l - W.tree.lmin) for l, b in lb.items())
    sched_info.pop("sched", None)
    return {"wall": wall, "stats": tot, "graph_ms": graph_ms, "level_us": level_us, "launches_per_step": launches,
            "steps": steps, "alg_bytes_per_step": per_coarse, "tuning": used, "tuned": tuned,
            "schedule": sched, "schedules_ms": schedules, "persistent_plan": sched_info,
            "sched_tuned": sched_tuned}


def measure(name, W, precision, steps, warmup, tuning=None, tune_here=False, schedule=None):
    """One workload's timed region and device times, as the JSON fields of a
    bench line (value, roofline of the dominant kernel, per-level table)."""
    if name in ("c4", "c5"):
        r = run_amr(W, precision, steps, warmup, tuning, tune_here,
                    log=lambda m: print("bench: " + m, file=sys.stderr, flush=True), schedule=schedule)
    else:
        r = run_uniform(W, precision, steps, warmup, tuning)
    cu = r["stats"]["cell_updates"]
    lvl, achieved = _dominant(W, precision, r["level_us"], r["launches_per_step"])
    step_ms = r["graph_ms"] / steps
    return r, {
        "value": cu / r["wall"],
        "ms_per_step": r["wall"] * 1e3 / steps,
        "device_ms_per_step": step_ms,
        "step_alg_frac": r["alg_bytes_per_step"] / (step_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "chem_substeps_per_cell_update": r["stats"]["chem_substeps"] / max(cu, 1),
        "dominant_level": lvl,
        "kernel": _kernel_name(W, lvl, precision, r["tuning"]),
        "kernel_us": r["level_us"][lvl],
        "achieved": achieved,
        "levels": _level_table(W, precision, r["level_us"], r["launches_per_step"], r["tuning"]),
        "tuned": r.get("tuned"),
        "sched_tuned": r.get("sched_tuned"),
        "schedule": r.get("schedule"),
        "schedules_ms": r.get("schedules_ms"),
        "persistent_plan": r.get("persistent_plan"),
    }


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
    """A shader copy of known size: an elementwise kernel reading `nbytes` and
    writing `nbytes` (b = a + 1; a same-dtype copy_ may go to the DMA engines,
    which the shader counters do not see)."""
    import torch

    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda:0").fill_(1.0)
    b = torch.empty_like(a)
    torch.add(a, 1.0, out=b)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        torch.add(a, 1.0, out=b)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "reps": reps + 1,
                      "kernel": "torch elementwise add (b = a + 1), float32",
                      "GBps": 2 * nbytes * reps / el / 1e9}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=32)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--workload", default="c5", choices=["c3", "c2", "c4", "c5"])
    ap.add_argument("--precision", default=None, choices=["mixed", "fp64"])
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--cpu-cores", type=int, default=0,
                    help="host cores for the CPU baseline (0: the job's CPU share, see host_cores())")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-c4", action="store_true", help="c5 only: leave out the c4 block")
    ap.add_argument("--no-fp64-plain", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--stencil-path", action="store_true")
    ap.add_argument("--no-tuning", action="store_true",
                    help="the library's default per-level choices instead of the tuning table")
    ap.add_argument("--schedule", default=None, choices=["graph", "persistent"],
                    help="coarse-step schedule of the timed region (c4, c5; default: the tuning table's)")
    ap.add_argument("--tune", default=None, metavar="PATH",
                    help="run the autotuner on the workload first (outside the timed region), use its choices "
                         "and merge them into the table at PATH")
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

    from emma_amd import tune

    precision = args.precision or ("fp64" if args.workload == "c2" else "mixed")
    W = make_workload(args.workload, args.seed)
    table = None if (args.no_tuning or args.tune) else tune.load()
    tuning_src = ("autotuned in this run" if args.tune else
                  "library defaults" if not table else "table emma_amd/tuning.json")
    print(f"bench: workload {args.workload} ({precision}), {args.steps} steps after {args.warmup} warm-up, "
          f"per-level tuning: {tuning_src}", file=sys.stderr, flush=True)
    _, m = measure(args.workload, W, precision, args.steps, args.warmup, table, bool(args.tune), args.schedule)
    if args.tune and m["tuned"]:
        tune.save(m["tuned"], WORKLOAD_NAMES[args.workload], args.tune, m["sched_tuned"])
    sched_txt = {"graph": "one HIP graph replay of per-level launches",
                 "persistent": "one HIP graph replay of one persistent launch fed by a device work queue"}
    step = {"c4": "coarse step: 31 level steps (levels 7-11), ",
            "c5": "coarse step: 7 level steps (levels 8-10), "}.get(args.workload, "radiative sub-cycle")
    if args.workload in ("c4", "c5"):
        step += sched_txt[m["schedule"]]
    out = {
        "metric": "M1+chem cell-updates/s (mixed+batched vs FP64 plain, 1 GPU) and % HBM peak",
        "value": m["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64" if precision == "mixed" else "fp64",
        "data": "synthetic",
        "config": {
            "workload": WORKLOAD_NAMES[args.workload],
            "cells": sum(W.tree.level(l).ncell for l in range(W.tree.lmin, W.tree.lmax + 1)),
            "leaf_cell_updates_per_step": W.leaf_cell_updates_per_coarse_step(),
            "groups": W.params.ngrp,
            "precision": precision,
            "step": step,
            "device_ms_per_step": m["device_ms_per_step"],
            "step_alg_frac": m["step_alg_frac"],
            "chem_substeps_per_cell_update": m["chem_substeps_per_cell_update"],
            "levels": m["levels"],
            "tuning": tuning_src,
            "schedule": m["schedule"],
            "schedules_device_ms_per_step": m["schedules_ms"],
            "persistent_plan": m["persistent_plan"],
            "parallelism": "1 GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": m["achieved"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": m["achieved"] / HBM_PEAK_GBS,
            "traffic": load_traffic(f"{args.workload}_{precision}"),
            "kernel": m["kernel"],
            "level": m["dominant_level"],
            "kernel_us": m["kernel_us"],
            "alg_bytes_per_cell_update": alg_bytes_per_cell(W.params.ngrp, precision),
        },
    }
    if precision == "mixed" and not args.no_fp64_plain:
        nsteps = max(4, args.steps // 4)
        print(f"bench: fp64 plain, {nsteps} steps", file=sys.stderr, flush=True)
        _, m2 = measure(args.workload, W, "fp64", nsteps, min(args.warmup, 2))
        out["fp64_plain"] = {"value": m2["value"], "unit": "cell-updates/s", "steps": nsteps,
                             "device_ms_per_step": m2["device_ms_per_step"], "kernel": m2["kernel"],
                             "kernel_us": m2["kernel_us"], "frac": m2["achieved"] / HBM_PEAK_GBS,
                             "speedup_of_value": m["value"] / m2["value"]}
    if args.workload == "c5" and not args.no_c4:
        # the batched AMR tree of levels 7-11 beside it (more, smaller levels)
        del W
        W4 = make_workload("c4", 4)
        n4 = min(args.steps, 10)
        print(f"bench: c4 block, {n4} steps", file=sys.stderr, flush=True)
        _, m4 = measure("c4", W4, precision, n4, min(args.warmup, 3), table, bool(args.tune))
        if args.tune and m4["tuned"]:
            tune.save(m4["tuned"], WORKLOAD_NAMES["c4"], args.tune, m4["sched_tuned"])
        out["c4"] = {"workload": WORKLOAD_NAMES["c4"], "step": "coarse step: 31 level steps, " + sched_txt[m4["schedule"]],
                     "schedule": m4["schedule"], "schedules_device_ms_per_step": m4["schedules_ms"],
                     "persistent_plan": m4["persistent_plan"],
                     "value": m4["value"], "unit": "cell-updates/s", "steps": n4, "ms_per_step": m4["ms_per_step"],
                     "device_ms_per_step": m4["device_ms_per_step"], "step_alg_frac": m4["step_alg_frac"],
                     "kernel": m4["kernel"], "level": m4["dominant_level"], "kernel_us": m4["kernel_us"],
                     "frac": m4["achieved"] / HBM_PEAK_GBS, "levels": m4["levels"]}
    if not args.no_cpu_baseline:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(args.workload, args.cpu_seconds, args.cpu_cores)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
