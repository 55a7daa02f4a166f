"""Autotuner of the mixed-precision level step.

The reference picks its GPU launch shape from run parameters (threads per
block ``nthread``, streams ``nstream``, stencil batch ``hstride``:
``parameters.c:571-572, 653-656``, ``rad_utils_gpu.cu:1025-1043``).  Here the
per-level choices are

* the halo variant of the tiled kernel (``VARIANT_LISTED`` / ``_PERDIR`` /
  ``_LATENCY``, see ``emma_rt_set_level_variant``), and
* the chemistry sub-step budget after which a leaf goes to the straggler
  pass (``emma_rt_set_chem_budget``; 0 = no pass), and the factor beyond
  which a leaf's chain counts as long and stays in the first pass
  (``emma_rt_set_chem_keep``), and
* on a complete Morton-ordered cube, the tile order of the tiled kernel
  (``emma_rt_set_level_order``: Morton or raster),

which change the speed and not the results.  :func:`tune` times every
candidate on the workload's own state (event-timed level steps of repeated
coarse steps from one saved device state) and keeps, per level, the fastest.
For the whole tree, :func:`tune_schedule` picks the coarse-step schedule:
the captured graph of per-level launches or the persistent kernel fed by a
device work queue (``emma_rt_set_schedule``; the same results bit for bit).
The table (:func:`save` / :func:`load`) is keyed by the level's shape --
groups, cells, complete or sparse -- and by the tree's shape for the
schedule, and :func:`apply` sets a context from it.
"""
from __future__ import annotations

import json
import os

from . import rt

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")
BUDGETS = (0, 2, 3, 4, 6, 8)
KEEPS = (0.0, 8.0, 32.0)


ORDER_NAMES = {rt.ORDER_MORTON: "morton", rt.ORDER_RASTER: "raster"}


def level_key(ngrp: int, ncell: int, halo_per_dir: bool) -> str:
    return f"G{ngrp}-{'sparse' if halo_per_dir else 'complete'}-{ncell}"


def tree_key(ngrp: int, tree) -> str:
    return f"tree-G{ngrp}-" + "-".join(str(tree.level(l).ncell) for l in range(tree.lmin, tree.lmax + 1))


def _uncounted(f):
    """run f(g, ...) with the kernels' diagnostic counters off, as the bench
    times its coarse steps, and restore the context's setting"""
    def run(g, *args):
        was = getattr(g, "counters", True)
        g.enable_counters(False)
        try:
            return f(g, *args)
        finally:
            g.enable_counters(was)
    return run


@_uncounted
def _step_ms(g, reps):
    """mean device time (ms) of one coarse step (one event pair around each
    replay) over `reps` coarse steps from the saved state, after two untimed
    ones (the graphs of both buffer parities)"""
    for _ in range(2):
        g.restore_state()
        g.coarse_step(1.0, 1)
    g.enable_timing(rt.TIMING_GRAPH)
    ms = 0.0
    for _ in range(reps):
        g.restore_state()
        g.coarse_step(1.0, 1)
        ms += g.stats()["kernel_ms"]
    g.enable_timing(rt.TIMING_OFF)
    return ms / reps


def tune_schedule(g: "rt.RadGPU", reps: int = 5, log=None) -> dict:
    """The faster coarse-step schedule for the context's tree (its levels as
    they are set), from a saved state; leaves the context on it."""
    t = {}
    for sched in (rt.SCHED_GRAPH, rt.SCHED_PERSISTENT):
        g.set_schedule(sched)
        t[sched] = _step_ms(g, reps)
        if log:
            log(f"tune: schedule {rt.SCHED_NAMES[sched]}: {t[sched]:.3f} ms per coarse step")
    best = min(t, key=t.get)
    g.set_schedule(best)
    return {"key": tree_key(g.params.ngrp, g.tree), "sched": rt.SCHED_NAMES[best],
            "ms": {rt.SCHED_NAMES[k]: round(v, 4) for k, v in t.items()}}


@_uncounted
def _level_times(g, levels, reps):
    """mean device time (us) of each level's step over `reps` coarse steps
    from the saved state"""
    acc = {l: [0.0, 0] for l in levels}
    g.enable_timing(rt.TIMING_LAUNCH)
    for _ in range(reps):
        g.restore_state()
        g.coarse_step(1.0, 1)
        for l, (ms, n) in g.level_times().items():
            if l in acc:
                acc[l][0] += ms
                acc[l][1] += n
    g.enable_timing(rt.TIMING_OFF)
    return {l: a[0] / max(a[1], 1) * 1e3 for l, a in acc.items()}


def tune(g: "rt.RadGPU", reps: int = 3, budgets=BUDGETS, keeps=KEEPS, log=None) -> dict:
    """Per level of the context's tree, the (variant, budget, keep) with the
    smallest event-timed level step.  The context must hold a mixed-precision
    tree with 1 or 3 groups and a saved state (``g.save_state()``) to start
    every timed coarse step from.  Returns {level: {variant, chem_budget,
    chem_keep, us, default_us, candidates}}."""
    tree = g.tree
    levels = [l for l in range(tree.lmin, tree.lmax + 1) if tree.level(l).noct > 0]
    shape = {l: g.level_tuning(l) for l in levels}
    default = {l: (shape[l]["variant"], shape[l]["chem_budget"], shape[l]["chem_keep"]) for l in levels}
    t_default = _level_times(g, levels, reps)
    cands = {l: {} for l in levels}
    variants = [rt.VARIANT_LISTED, rt.VARIANT_PERDIR, rt.VARIANT_LATENCY]
    combos = [(v, b, k) for v in variants for b in budgets for k in (keeps if b > 0 else (0.0,))]
    for v, b, k in combos:
        for l in levels:  # a variant a level cannot take keeps its default
            ok = not (v == rt.VARIANT_LISTED and shape[l]["halo_per_dir"])
            g.set_level_variant(v if ok else default[l][0], level=l)
            g.set_chem_budget(b, level=l)
            g.set_chem_keep(k, level=l)
        t = _level_times(g, levels, reps)
        for l in levels:
            if v == rt.VARIANT_LISTED and shape[l]["halo_per_dir"]:
                continue
            cands[l][(v, b, k)] = t[l]
        if log:
            log(f"tune: variant {rt.VARIANT_NAMES[v]} budget {b} keep {k:g}: " +
                ", ".join(f"{l}: {t[l]:.1f} us" for l in levels))
    out = {}
    for l in levels:
        (v, b, k), us = min(cands[l].items(), key=lambda kv: kv[1])
        g.set_level_variant(v, level=l)
        g.set_chem_budget(b, level=l)
        g.set_chem_keep(k, level=l)
        L = tree.level(l)
        out[l] = {"key": level_key(g.params.ngrp, L.ncell, shape[l]["halo_per_dir"]),
                  "variant": rt.VARIANT_NAMES[v], "chem_budget": b, "chem_keep": k, "us": us,
                  "default_us": t_default[l],
                  "default": {"variant": rt.VARIANT_NAMES[default[l][0]], "chem_budget": default[l][1],
                              "chem_keep": default[l][2]},
                  "candidates": {f"{rt.VARIANT_NAMES[cv]}/{cb}/{ck:g}": round(ct, 2)
                                 for (cv, cb, ck), ct in cands[l].items()}}
    # the tile order of a complete Morton cube, with the level's choices
    # above: Morton order keeps more of the halo in L2 on some cubes, the
    # raster order on others
    for l in levels:
        if g.level_order(l)["cube_side_log2"] == 0:
            continue
        t_o = {}
        for o in (rt.ORDER_MORTON, rt.ORDER_RASTER):
            g.set_level_order(o, level=l)
            t_o[o] = _level_times(g, [l], reps)[l]
        best = min(t_o, key=t_o.get)
        g.set_level_order(best, level=l)
        out[l]["order"] = ORDER_NAMES[best]
        out[l]["order_us"] = {ORDER_NAMES[o]: round(t, 2) for o, t in t_o.items()}
        out[l]["us"] = t_o[best]
        if log:
            log(f"tune: level {l} tile order " + ", ".join(f"{ORDER_NAMES[o]} {t:.1f} us" for o, t in t_o.items()))
    return out


def load(path: str = TABLE) -> dict:
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def save(result: dict, workload: str, path: str = TABLE, schedule: dict = None, base: str = None) -> dict:
    """merge a tune() result (and a tune_schedule() result) into the table at
    `path` (entries keyed by level / tree shape; `workload` names where they
    were measured); a table not yet at `path` starts as a copy of `base`"""
    table = load(path) if os.path.exists(path) or base is None else load(base)
    if schedule is not None:
        table[schedule["key"]] = {"sched": schedule["sched"], "ms": schedule["ms"], "workload": workload}
    for l, r in result.items():
        table[r["key"]] = {"variant": r["variant"], "chem_budget": r["chem_budget"],
                           "chem_keep": r.get("chem_keep", 0.0), "us": round(r["us"], 2),
                           "default_us": round(r["default_us"], 2), "workload": workload, "level": int(l)}
        if "order" in r:
            table[r["key"]]["order"] = r["order"]
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
        f.write("\n")
    return table


def apply(g: "rt.RadGPU", table: dict = None) -> dict:
    """set every level whose shape the table holds, and the tree's schedule
    when the table holds the tree; returns {level: entry}"""
    if table is None:
        table = load()
    names = {v: k for k, v in rt.VARIANT_NAMES.items()}
    used = {}
    e = table.get(tree_key(g.params.ngrp, g.tree))
    if e is not None:  # the tree's schedule (g.schedule() reports it)
        g.set_schedule({v: k for k, v in rt.SCHED_NAMES.items()}[e["sched"]])
    if g.precision != rt.MIXED or g.params.ngrp not in (1, 3):
        return used
    for l in range(g.tree.lmin, g.tree.lmax + 1):
        L = g.tree.level(l)
        if L.noct == 0:
            continue
        sh = g.level_tuning(l)
        e = table.get(level_key(g.params.ngrp, L.ncell, sh["halo_per_dir"]))
        if e is None:
            continue
        g.set_level_variant(names[e["variant"]], level=l)
        g.set_chem_budget(int(e["chem_budget"]), level=l)
        g.set_chem_keep(float(e.get("chem_keep", 0.0)), level=l)
        if "order" in e and g.level_order(l)["cube_side_log2"] > 0:
            g.set_level_order({v: k for k, v in ORDER_NAMES.items()}[e["order"]], level=l)
        used[l] = e
    return used


def main(argv=None) -> int:
    """python -m emma_amd.tune [--schedule-only] [--table PATH] [--out PATH] c4 c5 ...:
    tune the named workloads (from the state after their first coarse step)
    and merge the choices into the table"""
    import argparse
    import sys

    from . import synth

    ap = argparse.ArgumentParser(prog="python -m emma_amd.tune")
    ap.add_argument("workloads", nargs="+", choices=["c4", "c5"])
    ap.add_argument("--schedule-only", action="store_true",
                    help="keep the table's per-level choices, tune the coarse-step schedule only")
    ap.add_argument("--table", default=TABLE, help="table the per-level choices are read from")
    ap.add_argument("--out", default=None, help="table to write (default: --table)")
    ap.add_argument("--reps", type=int, default=5)
    args = ap.parse_args(argv)
    out = args.out or args.table
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    names = {"c4": "AMR levels 7-11 (262144/80000/25000/6000/1200 octs), 3 groups, lognormal sigma=1.5",
             "c5": "256^3 coarse level + 2 refined levels (300000/75000 octs), 3 groups, 4096 sources, c/100, "
                   "lognormal sigma=2"}
    for w in args.workloads:
        W = synth.config_c4(seed=4, nsub=1) if w == "c4" else synth.config_c5(seed=3, nsub=1)  # the bench's seeds
        g = rt.RadGPU(W.params, rt.MIXED)
        g.setup(W.params, W.tree, W.fields)
        apply(g, load(args.table))
        g.set_schedule(rt.SCHED_GRAPH)
        g.coarse_step(1.0, 1)
        g.save_state()
        levels = {} if args.schedule_only else tune(g, reps=3, log=log)
        sched = tune_schedule(g, reps=args.reps, log=log)
        log(f"{w}: schedule {sched['sched']} {sched['ms']}")
        save(levels, names[w], out, sched, base=args.table)
        g.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
