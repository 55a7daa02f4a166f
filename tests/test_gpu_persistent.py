"""The persistent coarse step (EMMA_RT_SCHED_PERSISTENT) on the GPU.

A coarse step as one launch of persistent work-groups fed by a device work
queue runs the same level-step code on the same tiles as the graph of
per-level launches, every tile after the tiles it depends on; reflux
accumulators have one writer per level step and are summed in pair order.  So
every field of every level must equal the graph schedule's bit for bit: on
AMR trees in both precisions (tiled mixed kernel for 1 and 3 groups, per-cell
kernels for FP64 and 9 groups), in RADSTEP and COARSERAD mode, with domain
faces, with the hydro coupling and the test source (FillRad's block and the
hydro injection fused into the items), with star particles and with the
reionisation times; and, against the oracle, within the parity tolerances.
At full size (C4, C5) the two schedules agree bit for bit as well.

The persistent kernel holds the listed and per-direction halo variants of
the tiled kernel; a level the graph schedule would run with the latency
variant (same arithmetic, another load order, which the compiler schedules
into a different FP32 rounding) runs per-direction, so the graph runs here
take the per-direction variant on those levels too.
"""
import time

import numpy as np
import pytest

from emma_amd import rt, synth
from emma_amd import tree as T
from emma_amd.synth import Fields
from test_gpu_parity import FP64_TOL, MIXED_TOL, field_errors

pytestmark = pytest.mark.gpu

SMALL = (3, [60, 100, 80])


def small(table, seed, nsrc=3, faces=None):
    W = synth.amr_workload(SMALL[0], SMALL[1], table, seed=seed, sigma=1.5, nsrc=nsrc, grid=16)
    if faces == "trans_ref":
        tb = T.with_boundaries(W.tree, trans=range(6), reflect=(0, 2, 4))
        W = synth.Workload(W.name, W.params, tb, W.fields, W.nsub)
    return W


def no_latency_variant(g, tree):
    """levels on the latency variant take the per-direction one (what the
    persistent kernel runs there)"""
    if g.precision != rt.MIXED or g.params.ngrp not in (1, 3):
        return
    for l in range(tree.lmin, tree.lmax + 1):
        if g.level_tuning(l)["variant"] == rt.VARIANT_LATENCY:
            g.set_level_variant(rt.VARIANT_PERDIR, l)


def run(W, precision, sched, calls=2, ncoarse=1, mode=rt.RADSTEP, prepare=None):
    g = rt.RadGPU(W.params, precision)
    g.setup(W.params, W.tree, W.fields, mode=mode)
    no_latency_variant(g, W.tree)
    if prepare is not None:
        prepare(g)
    g.set_schedule(sched)
    for _ in range(calls):
        g.coarse_step(1.0, ncoarse)
    st = g.stats()
    out = {l: g.get(l) for l in range(W.tree.lmin, W.tree.lmax + 1)}
    info = g.schedule()
    g.close()
    return out, st, info


def assert_same(a, b):
    for l in a:
        for k in Fields.NAMES:
            x, y = getattr(a[l], k), getattr(b[l], k)
            assert np.array_equal(x, y), (l, k, np.flatnonzero(x != y)[:5])


def both(W, precision, **kw):
    ga, sa, _ = run(W, precision, rt.SCHED_GRAPH, **kw)
    pa, sp, info = run(W, precision, rt.SCHED_PERSISTENT, **kw)
    assert sa["status"] == 0 and sp["status"] == 0, (sa, sp)
    assert sa["cell_updates"] == sp["cell_updates"] and sa["chem_substeps"] == sp["chem_substeps"], (sa, sp)
    assert info["sched"] == "persistent" and info["items"] > 0 and info["launches"] >= 1, info
    assert_same(ga, pa)
    return pa, info


@pytest.mark.parametrize("table,G", [("SB99_equiv_mono_TH", 1), ("Starburst99_1grpT_3grpS", 3),
                                     ("Starburst99_3grpT_3grpS", 9)])
@pytest.mark.parametrize("precision", [rt.FP64, rt.MIXED])
@pytest.mark.parametrize("ncoarse", [1, 2])
def test_radstep_same_as_graph(gpu_lib, table, G, precision, ncoarse):
    W = small(table, 11 + G)
    _, info = both(W, precision, calls=2, ncoarse=ncoarse)
    print("radstep G=%d prec=%d ncoarse=%d" % (G, precision, ncoarse), info)


@pytest.mark.parametrize("table,G", [("SB99_equiv_mono_TH", 1), ("Starburst99_1grpT_3grpS", 3)])
@pytest.mark.parametrize("precision", [rt.FP64, rt.MIXED])
@pytest.mark.parametrize("faces", ["periodic", "trans_ref"])
def test_coarserad_same_as_graph(gpu_lib, table, G, precision, faces):
    W = small(table, 20 + G, faces=faces)
    _, info = both(W, precision, calls=3, mode=rt.COARSERAD)
    assert info["launches"] == 2, info  # finer levels, then the coarse level (transfers between)


@pytest.mark.parametrize("G", [1, 3, 9])
@pytest.mark.parametrize("precision", [rt.FP64, rt.MIXED])
def test_domain_faces_same_as_graph(gpu_lib, G, precision):
    table = {1: "SB99_equiv_mono_TH", 3: "Starburst99_1grpT_3grpS", 9: "Starburst99_3grpT_3grpS"}[G]
    both(small(table, 29, nsrc=4, faces="trans_ref"), precision, calls=2)


@pytest.mark.parametrize("table,G", [("SB99_equiv_mono_TH", 1), ("Starburst99_1grpT_3grpS", 3)])
@pytest.mark.parametrize("precision", [rt.FP64, rt.MIXED])
def test_coupled_same_as_graph(gpu_lib, table, G, precision):
    """FillRad's block (hydro -> radiation, Stromgren source, E2T) fused
    before every item and the hydro injection after it"""
    W = small(table, 13, nsrc=0)
    rng = np.random.default_rng(7)
    hyd = {l: synth.hydro_like(F, rng) for l, F in W.fields.items()}

    def prepare(g):
        g.set_tphy(1e-3)
        g.set_coupling(True)
        g.set_stromgren_source(synth.Star(x=0.5, y=0.5, z=0.5))
        for l, H in hyd.items():
            g.set_hydro(l, H)
            g.set_oct_pos(l, synth.oct_corners(W.tree.level(l)))

    runs = {}
    for sched in (rt.SCHED_GRAPH, rt.SCHED_PERSISTENT):
        g = rt.RadGPU(W.params, precision)
        g.setup(W.params, W.tree, W.fields)
        no_latency_variant(g, W.tree)
        prepare(g)
        g.set_schedule(sched)
        g.coarse_step(1.0, 1)
        g.coarse_step(1.0, 1)
        assert g.stats()["status"] == 0
        runs[sched] = ({l: g.get(l) for l in W.fields}, {l: g.get_hydro(l) for l in W.fields})
        g.close()
    assert_same(runs[rt.SCHED_GRAPH][0], runs[rt.SCHED_PERSISTENT][0])
    for l in W.fields:
        for a, b in zip(runs[rt.SCHED_GRAPH][1][l], runs[rt.SCHED_PERSISTENT][1][l]):
            assert np.array_equal(a, b), l


@pytest.mark.parametrize("precision", [rt.FP64, rt.MIXED])
@pytest.mark.parametrize("ncoarse", [1, 2])
def test_coarserad_coupled_same_as_graph(gpu_lib, precision, ncoarse):
    """COARSERAD with the hydro coupling and the test source: in the
    persistent schedule the coarse level's hydro injection runs inside its
    items, before distribE (it reads the coarse level's x, eint; distribE
    writes the finer levels' N, F); several coarse steps per call give two
    persistent launches each"""
    W = small("Starburst99_1grpT_3grpS", 31, nsrc=0, faces="trans_ref")
    rng = np.random.default_rng(9)
    hyd = {l: synth.hydro_like(F, rng) for l, F in W.fields.items()}
    runs = {}
    for sched in (rt.SCHED_GRAPH, rt.SCHED_PERSISTENT):
        g = rt.RadGPU(W.params, precision)
        g.setup(W.params, W.tree, W.fields, tphy=1e-3, mode=rt.COARSERAD)
        no_latency_variant(g, W.tree)
        g.set_coupling(True)
        g.set_stromgren_source(synth.Star(x=0.02, y=0.02, z=0.02))
        for l, H in hyd.items():
            g.set_hydro(l, H)
            g.set_oct_pos(l, synth.oct_corners(W.tree.level(l)))
        g.set_schedule(sched)
        for _ in range(2):
            g.coarse_step(1.0, ncoarse)
        assert g.stats()["status"] == 0
        if sched == rt.SCHED_PERSISTENT:
            assert g.schedule()["launches"] == 2 * ncoarse
        runs[sched] = ({l: g.get(l) for l in W.fields}, {l: g.get_hydro(l) for l in W.fields})
        g.close()
    assert_same(runs[rt.SCHED_GRAPH][0], runs[rt.SCHED_PERSISTENT][0])
    for l in W.fields:
        for a, b in zip(runs[rt.SCHED_GRAPH][1][l], runs[rt.SCHED_PERSISTENT][1][l]):
            assert np.array_equal(a, b), l


@pytest.mark.parametrize("precision", [rt.FP64, rt.MIXED])
def test_stars_and_xion_times_same_as_graph(gpu_lib, precision):
    from test_stars import star_setup, workload

    W, stars = workload(3, 33)
    res = {}
    for sched in (rt.SCHED_GRAPH, rt.SCHED_PERSISTENT):
        g = rt.RadGPU(W.params, precision)
        star_setup(g, W, stars, False)
        no_latency_variant(g, W.tree)
        g.enable_xion_times(True)
        g.set_schedule(sched)
        for _ in range(2):
            g.coarse_step(1.0, 1)
        assert g.stats()["status"] == 0
        res[sched] = ({l: g.get(l) for l in W.fields}, {l: g.xion_times(l) for l in W.fields})
        g.close()
    assert_same(res[rt.SCHED_GRAPH][0], res[rt.SCHED_PERSISTENT][0])
    for l in W.fields:
        for a, b in zip(res[rt.SCHED_GRAPH][1][l], res[rt.SCHED_PERSISTENT][1][l]):
            assert np.array_equal(a, b), l


@pytest.mark.parametrize("table,G", [("SB99_equiv_mono_TH", 1), ("Starburst99_1grpT_3grpS", 3)])
@pytest.mark.parametrize("precision", [rt.FP64, rt.MIXED])
def test_persistent_vs_oracle(gpu_lib, oracle, table, G, precision):
    W = small(table, 41)
    g = rt.RadGPU(W.params, precision)
    g.setup(W.params, W.tree, W.fields)
    g.set_schedule(rt.SCHED_PERSISTENT)
    oracle.setup(W.params, W.tree, W.fields)
    for _ in range(3):
        g.coarse_step(1.0, 1)
        oracle.coarse_step()
    tol = FP64_TOL if precision == rt.FP64 else MIXED_TOL
    for l in range(W.tree.lmin, W.tree.lmax + 1):
        err = field_errors(g.get(l), oracle.get(l), W.params.cloc())
        print("persistent vs oracle G=%d prec=%d level %d" % (G, precision, l), err)
        assert err["x"] < tol and err["temp"] < tol and err["e"] < tol and err["eint"] < tol, (l, err)
        assert max(err["fx"], err["fy"], err["fz"]) < tol, (l, err)


def test_persistent_repeatable_and_switchable(gpu_lib):
    """two runs agree bit for bit; switching schedules between calls on one
    context continues the same state"""
    W = small("Starburst99_1grpT_3grpS", 17)
    a, _, _ = run(W, rt.MIXED, rt.SCHED_PERSISTENT, calls=3)
    b, _, _ = run(W, rt.MIXED, rt.SCHED_PERSISTENT, calls=3)
    assert_same(a, b)
    g = rt.RadGPU(W.params, rt.MIXED)
    g.setup(W.params, W.tree, W.fields)
    no_latency_variant(g, W.tree)
    for sched in (rt.SCHED_PERSISTENT, rt.SCHED_GRAPH, rt.SCHED_PERSISTENT):
        g.set_schedule(sched)
        g.coarse_step(1.0, 1)
    c = {l: g.get(l) for l in W.fields}
    g.close()
    assert_same(a, c)


def test_persistent_status_emin(gpu_lib):
    """an error the kernels raise (photon density below its floor) comes back
    from the persistent launch as from the graph"""
    W = small("SB99_equiv_mono_TH", 3)
    for F in W.fields.values():
        F.e[:] = 1e-30
        F.fx[:] = F.fy[:] = F.fz[:] = 0.0
    codes = []
    for sched in (rt.SCHED_GRAPH, rt.SCHED_PERSISTENT):
        g = rt.RadGPU(W.params, rt.FP64)
        g.setup(W.params, W.tree, W.fields)
        g.set_schedule(sched)
        with pytest.raises(rt.EmmaRTError) as ei:
            g.coarse_step(1.0, 1)
        codes.append(ei.value.status)
        g.close()
    assert codes[0] == codes[1] == rt.ERR_EMIN, codes


@pytest.mark.parametrize("sched", [rt.SCHED_GRAPH, rt.SCHED_PERSISTENT], ids=["graph", "persistent"])
def test_replays_start_from_cleared_words(gpu_lib, sched):
    """a replayed coarse step clears the status word, the straggler list
    lengths and the counters at the head of its graph: replays of the same
    graphs alternate between a state that raises EMIN and one that does not,
    and each reports its own outcome and counts"""
    W = small("Starburst99_1grpT_3grpS", 7)
    bad = {l: F.copy() for l, F in W.fields.items()}
    for F in bad.values():
        F.e[:] = 1e-30
        F.fx[:] = F.fy[:] = F.fz[:] = 0.0
    g = rt.RadGPU(W.params, rt.MIXED)
    g.setup(W.params, W.tree, W.fields)
    g.set_schedule(sched)
    g.set_chem_budget(2)
    for rep in range(3):
        for fields, ok in ((bad, False), (W.fields, True)):
            for l, F in fields.items():
                g.set(l, 0, F)
            if ok:
                g.coarse_step(1.0, 1)
                st = g.stats()
                assert st["status"] == 0 and st["cell_updates"] == W.leaf_cell_updates_per_coarse_step(), (rep, st)
            else:
                with pytest.raises(rt.EmmaRTError) as ei:
                    g.coarse_step(1.0, 1)
                assert ei.value.status == rt.ERR_EMIN, (rep, ei.value.status)
    g.close()


@pytest.mark.parametrize("config", ["c4", "c5"])
def test_full_size_same_as_graph(gpu_lib, config):
    """the benchmark trees (C4: levels 7-11, 31 level steps; C5: 256^3 + two
    refined levels), one coarse step from the drawn state, library defaults"""
    W = synth.config_c4(seed=4, nsub=1) if config == "c4" else synth.config_c5(seed=5, nsub=1)
    res = {}
    for sched in (rt.SCHED_GRAPH, rt.SCHED_PERSISTENT):
        g = rt.RadGPU(W.params, rt.MIXED)
        g.setup(W.params, W.tree, W.fields)
        no_latency_variant(g, W.tree)
        g.set_schedule(sched)
        g.save_state()
        g.coarse_step(1.0, 1)
        assert g.stats()["status"] == 0
        res[sched] = {l: g.get(l) for l in range(W.tree.lmin, W.tree.lmax + 1)}
        ts = []
        for _ in range(3):
            g.restore_state()
            t0 = time.perf_counter()
            g.coarse_step(1.0, 1)
            ts.append(time.perf_counter() - t0)
        print(config, rt.SCHED_NAMES[sched], "coarse step ms", [round(t * 1e3, 3) for t in ts], g.schedule(),
              flush=True)
        g.close()
    assert_same(res[rt.SCHED_GRAPH], res[rt.SCHED_PERSISTENT])


@pytest.mark.parametrize("sched", [rt.SCHED_GRAPH, rt.SCHED_PERSISTENT], ids=["graph", "persistent"])
def test_counters_off_same_results(gpu_lib, sched):
    """with the kernels' diagnostic counters off the coarse steps give the
    same state bit for bit and the counter fields read -1; on again, the
    counts are back (the host's leaf count, and sub-steps)"""
    W = small("Starburst99_1grpT_3grpS", seed=5)

    def counters(on):
        def prepare(g):
            g.enable_counters(on)
        return prepare

    res = [run(W, rt.MIXED, sched, prepare=counters(on)) for on in (True, False, True)]
    for (out, st, _), on in zip(res, (True, False, True)):
        assert st["status"] == 0, st
        if on:
            assert st["cell_updates"] == W.leaf_cell_updates_per_coarse_step() and st["chem_substeps"] > 0, st
        else:
            assert st["cell_updates"] == -1 and st["chem_substeps"] == -1 and st["deferred_cells"] == -1, st
    assert_same(res[0][0], res[1][0])
    assert_same(res[0][0], res[2][0])
