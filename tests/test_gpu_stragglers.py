"""The chemistry straggler pass of the tiled mixed kernel (GPU).

With a budget b per level, a leaf still sub-stepping after b tries keeps its
transported N, F and is finished by a second kernel that restarts its
chemistry from the same inputs.  The pass changes the speed, not the result:
against the same run without the pass, every cell's x_HII, eint, T agree to
FP64 rounding and N, F to FP32 rounding, and the accepted sub-step count (the
trajectory of chemrad's sub-step controller, chem_utils.c:453-767) is the
same; against the oracle the mixed-precision bar holds as without the pass.
"""
import numpy as np
import pytest

from emma_amd import rt, synth
from test_gpu_parity import MIXED_TOL, field_errors

pytestmark = pytest.mark.gpu


def _run(W, budget, steps, graph=True, keep=0.0):
    g = rt.RadGPU(W.params, rt.MIXED)
    g.setup(W.params, W.tree, W.fields)
    if budget:
        g.set_chem_budget(budget)
        g.set_chem_keep(keep)
    sub = 0
    for _ in range(steps):
        if graph:
            g.coarse_step(1.0, 1)
        else:
            from oracle import coarse_step

            coarse_step(g, W.tree, W.params)
        sub += g.stats()["chem_substeps"]
    return g, sub


def _same(A, B, c):
    """A and B equal up to the rounding of the two code paths"""
    e = field_errors(A, B, c)
    assert e["x"] < 1e-12 and e["eint"] < 1e-12 and e["temp"] < 1e-12, e
    assert e["e"] < 1e-6 and max(e["fx"], e["fy"], e["fz"]) < 1e-6, e
    return e


@pytest.mark.parametrize("table,G", [("SB99_equiv_mono_TH", 1), ("Starburst99_1grpT_3grpS", 3)])
@pytest.mark.parametrize("budget", [1, 3])
def test_uniform_budget(gpu_lib, oracle, table, G, budget):
    W = synth.uniform_workload(5, table, seed=31, nsrc=8)
    g0, s0 = _run(W, 0, 3)
    g1, s1 = _run(W, budget, 3)
    assert g1.chem_budget(5) == budget
    c = W.params.cloc()
    e = _same(g1.get(5), g0.get(5), c)
    assert s1 == s0, (s0, s1)
    # and against the oracle, as the mixed path without the pass
    oracle.setup(W.params, W.tree, W.fields)
    for _ in range(3):
        oracle.coarse_step()
    eo = field_errors(g1.get(5), oracle.get(5), c)
    print("budget %d G=%d vs no pass" % (budget, G), e, "vs oracle", eo, "sub-steps", s1)
    assert eo["x"] < MIXED_TOL and eo["temp"] < MIXED_TOL and eo["e"] < MIXED_TOL, eo
    assert max(eo["fx"], eo["fy"], eo["fz"]) < MIXED_TOL, eo


@pytest.mark.parametrize("table,G", [("SB99_equiv_mono_TH", 1), ("Starburst99_1grpT_3grpS", 3)])
def test_amr_budget(gpu_lib, table, G):
    """AMR tree (partial tiles, coarse faces, refluxes folded in before the
    chemistry), graph replays and host-driven level steps"""
    W = synth.amr_workload(3, [60, 100, 80], table, seed=11, sigma=1.5, nsrc=2, grid=16)
    g0, s0 = _run(W, 0, 2)
    for graph in (True, False):
        g1, s1 = _run(W, 2, 2, graph=graph)
        if graph:  # (host-driven steps: the stats are those of the last RadSolver call)
            assert s1 == s0, (s0, s1)
        for l in range(W.tree.lmin, W.tree.lmax + 1):
            _same(g1.get(l), g0.get(l), W.params.cloc())


def test_budget_full_level(gpu_lib):
    """a 128^3 level with 64 sources: many stragglers, the list cleared
    between level steps (four sub-cycles in one graph)"""
    W = synth.config_c3(seed=3, nsub=1)
    g0, s0 = _run(W, 0, 4)
    g1, s1 = _run(W, 2, 4)
    assert s1 == s0, (s0, s1)
    _same(g1.get(7), g0.get(7), W.params.cloc())
    st = g1.stats()
    assert st["status"] == 0
    assert st["deferred_cells"] > 0 and g0.stats()["deferred_cells"] == 0, (st, g0.stats())
    # long chains kept in the first pass: fewer leaves deferred, same results
    g2, s2 = _run(W, 2, 4, keep=8.0)
    assert s2 == s0, (s0, s2)
    assert 0 < g2.stats()["deferred_cells"] < st["deferred_cells"], (g2.stats(), st)
    assert g2.level_tuning(7)["chem_keep"] == 8.0
    _same(g2.get(7), g0.get(7), W.params.cloc())


def test_budget_args(gpu_lib):
    W = synth.uniform_workload(3, "SB99_equiv_mono_TH", seed=1)
    g = rt.RadGPU(W.params, rt.FP64)
    g.setup(W.params, W.tree, W.fields)
    with pytest.raises(rt.EmmaRTError):
        g.set_chem_budget(2)  # the FP64 path has no straggler pass
    with pytest.raises(rt.EmmaRTError):
        g.set_chem_keep(8.0)
    g = rt.RadGPU(W.params, rt.MIXED)
    g.setup(W.params, W.tree, W.fields)
    with pytest.raises(rt.EmmaRTError):
        g.set_chem_budget(-1)
    with pytest.raises(rt.EmmaRTError):
        g.set_chem_budget(2, level=9)
    g.set_chem_budget(4, level=3)
    assert g.chem_budget(3) == 4
    with pytest.raises(rt.EmmaRTError):
        g.set_chem_keep(-1.0)


def test_budget_after_list_words_grow(gpu_lib):
    """A call with more level steps than there are straggler list words
    reallocates them; a coarse-step shape captured before (its graph held the
    old words) is captured again, and the pass still gives the run without it
    (one coarse step, then 40 in one call: 280 level steps against 256 words,
    then one again)"""
    W = synth.amr_workload(3, [60, 100, 80], "Starburst99_1grpT_3grpS", seed=11, sigma=1.5, nsrc=2, grid=16)
    runs = []
    for budget in (0, 2):
        g = rt.RadGPU(W.params, rt.MIXED)
        g.setup(W.params, W.tree, W.fields)
        if budget:
            g.set_chem_budget(budget)
        sub = []
        for n in (1, 40, 1):
            g.coarse_step(1.0, n)
            st = g.stats()
            assert st["status"] == 0, (budget, n, st)
            sub.append(st["chem_substeps"])
        runs.append((g, sub))
    (g0, s0), (g1, s1) = runs
    assert s1 == s0, (s0, s1)
    for l in range(W.tree.lmin, W.tree.lmax + 1):
        _same(g1.get(l), g0.get(l), W.params.cloc())
