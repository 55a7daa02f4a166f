"""Halo variants of the tiled mixed kernel and the autotuner (GPU).

The variants differ in when a tile loads its halo (the neighbours outside
the tile) and the chemistry inputs, and LISTED evaluates the update in
another order (the cell's own shares cancelled on faces that carry flux);
with contraction allowed in the FP32 transport the compiler fuses each
kernel's arithmetic its own way, so they agree to FP32 rounding, and every
variant meets the mixed-precision bar against the oracle.
"""
import numpy as np
import pytest

from emma_amd import rt, synth, tune
from emma_amd.synth import Fields
from test_gpu_parity import MIXED_TOL, field_errors

pytestmark = pytest.mark.gpu


def _run(W, variant, steps=2, budget=0):
    g = rt.RadGPU(W.params, rt.MIXED)
    g.setup(W.params, W.tree, W.fields)
    g.set_level_variant(variant)
    if budget:
        g.set_chem_budget(budget)
    for _ in range(steps):
        g.coarse_step(1.0, 1)
        assert g.stats()["status"] == 0
    return g


def _close(A, B, c):
    """A and B equal up to FP32 rounding of the transport"""
    err = field_errors(A, B, c)
    assert err["e"] < 1e-5 and max(err["fx"], err["fy"], err["fz"]) < 1e-5, err
    assert err["x"] < 1e-5 and err["temp"] < 1e-5 and err["eint"] < 1e-5, err
    assert err["src"] == 0 and err["nh"] == 0, err


@pytest.mark.parametrize("table,G", [("SB99_equiv_mono_TH", 1), ("Starburst99_1grpT_3grpS", 3)])
def test_variants_amr(gpu_lib, oracle, table, G):
    W = synth.amr_workload(3, [60, 100, 80], table, seed=11, sigma=1.5, nsrc=2, grid=16)
    gp = _run(W, rt.VARIANT_PERDIR)
    gl = _run(W, rt.VARIANT_LATENCY)
    oracle.setup(W.params, W.tree, W.fields)
    for _ in range(2):
        oracle.coarse_step()
    c = W.params.cloc()
    for l in range(W.tree.lmin, W.tree.lmax + 1):
        assert gl.level_tuning(l)["variant"] == rt.VARIANT_LATENCY
        _close(gl.get(l), gp.get(l), c)
        err = field_errors(gl.get(l), oracle.get(l), c)
        print("latency variant G=%d level %d" % (G, l), err)
        assert err["x"] < MIXED_TOL and err["temp"] < MIXED_TOL and err["e"] < MIXED_TOL, (l, err)
        assert max(err["fx"], err["fy"], err["fz"]) < MIXED_TOL, (l, err)


def test_variants_complete_level(gpu_lib, oracle):
    """a complete 32^3 level: all three variants against the oracle"""
    W = synth.uniform_workload(5, "Starburst99_1grpT_3grpS", seed=5, nsrc=8)
    oracle.setup(W.params, W.tree, W.fields)
    for _ in range(2):
        oracle.coarse_step()
    B = oracle.get(5)
    gs = _run(W, rt.VARIANT_LISTED)
    assert gs.level_tuning(5)["variant"] == rt.VARIANT_LISTED and not gs.level_tuning(5)["halo_per_dir"]
    gp = _run(W, rt.VARIANT_PERDIR)
    gl = _run(W, rt.VARIANT_LATENCY, budget=3)
    for name, g in (("listed", gs), ("perdir", gp), ("latency+budget", gl)):
        err = field_errors(g.get(5), B, W.params.cloc())
        print(name, err)
        assert err["x"] < MIXED_TOL and err["temp"] < MIXED_TOL and err["e"] < MIXED_TOL, (name, err)
        assert max(err["fx"], err["fy"], err["fz"]) < MIXED_TOL, (name, err)
    _close(gl.get(5), gp.get(5), W.params.cloc())
    # the straggler pass changes FP64 rounding only
    err = field_errors(gl.get(5), _run(W, rt.VARIANT_LATENCY).get(5), W.params.cloc())
    assert err["x"] < 1e-12 and err["eint"] < 1e-12 and err["e"] < 1e-6, err


def test_variant_args(gpu_lib):
    W = synth.amr_workload(3, [60, 100, 80], "SB99_equiv_mono_TH", seed=11, sigma=1.5, grid=16)
    g = rt.RadGPU(W.params, rt.MIXED)
    g.setup(W.params, W.tree, W.fields)
    sparse = [l for l in range(4, 7) if g.level_tuning(l)["halo_per_dir"]]
    assert sparse
    with pytest.raises(rt.EmmaRTError):
        g.set_level_variant(rt.VARIANT_LISTED, level=sparse[0])
    with pytest.raises(rt.EmmaRTError):
        g.set_level_variant(7)
    g.set_level_variant(rt.VARIANT_PERDIR, level=sparse[0])
    assert g.level_tuning(sparse[0])["variant"] == rt.VARIANT_PERDIR
    g.set_level_variant(rt.VARIANT_AUTO)
    assert g.level_tuning(3)["variant"] == rt.VARIANT_LISTED


def test_tuner(gpu_lib, tmp_path):
    """the tuner picks per level from its candidates, and the table it saves
    sets a fresh context's levels"""
    W = synth.amr_workload(3, [60, 100, 80], "Starburst99_1grpT_3grpS", seed=11, sigma=1.5, nsrc=2, grid=16)
    g = rt.RadGPU(W.params, rt.MIXED)
    g.setup(W.params, W.tree, W.fields)
    g.coarse_step(1.0, 1)
    g.save_state()
    res = tune.tune(g, reps=1, budgets=(0, 3))
    for l, r in res.items():
        assert abs(r["us"] - min(r["candidates"].values())) < 0.01, r
    path = str(tmp_path / "t.json")
    tune.save(res, "test tree", path)
    g2 = rt.RadGPU(W.params, rt.MIXED)
    g2.setup(W.params, W.tree, W.fields)
    used = tune.apply(g2, tune.load(path))
    assert set(used) == set(res)
    names = {v: k for k, v in rt.VARIANT_NAMES.items()}
    orders = {v: k for k, v in tune.ORDER_NAMES.items()}
    for l, r in res.items():
        t = g2.level_tuning(l)
        assert t["variant"] == names[r["variant"]] and t["chem_budget"] == r["chem_budget"]
        cube = g2.level_order(l)["cube_side_log2"] > 0
        assert ("order" in r) == cube, (l, r)
        if cube:  # the tile order of a complete cube is tuned too
            assert set(r["order_us"]) == {"morton", "raster"}, r
            assert abs(r["us"] - min(r["order_us"].values())) < 0.01, r
            assert g2.level_order(l)["order"] == orders[r["order"]], (l, r)


def test_tuner_tile_order(gpu_lib, tmp_path):
    """on a complete Morton-ordered cube the tuner also picks the tile order,
    and the saved table sets it"""
    W = synth.uniform_workload(5, "Starburst99_1grpT_3grpS", seed=19, nsrc=6)
    g = rt.RadGPU(W.params, rt.MIXED)
    g.setup(W.params, W.tree, W.fields)
    g.coarse_step(1.0, 1)
    g.save_state()
    res = tune.tune(g, reps=1, budgets=(0,))
    r = res[W.tree.lmin]
    assert r["order"] in ("morton", "raster") and set(r["order_us"]) == {"morton", "raster"}, r
    path = str(tmp_path / "t.json")
    tune.save(res, "cube", path)
    for order in ("morton", "raster"):  # either entry is applied as written
        table = tune.load(path)
        table[r["key"]]["order"] = order
        g2 = rt.RadGPU(W.params, rt.MIXED)
        g2.setup(W.params, W.tree, W.fields)
        tune.apply(g2, table)
        assert g2.level_order(W.tree.lmin)["order"] == {v: k for k, v in tune.ORDER_NAMES.items()}[order]
        g2.close()
    g.close()


@pytest.mark.parametrize("amr", [False, True])
def test_raster_tile_order(gpu_lib, amr):
    """A complete Morton-ordered cube takes the Morton tile order by default
    and can take the raster order; the order changes which block runs which
    tile, not the results"""
    if amr:
        W = synth.amr_workload(4, [300, 500], "Starburst99_1grpT_3grpS", seed=19, sigma=1.5, nsrc=3, grid=16)
    else:
        W = synth.uniform_workload(5, "Starburst99_1grpT_3grpS", seed=19, nsrc=6)
    res = {}
    for order in (rt.ORDER_RASTER, rt.ORDER_MORTON):
        g = rt.RadGPU(W.params, rt.MIXED)
        g.setup(W.params, W.tree, W.fields)
        info = g.level_order(W.tree.lmin)
        assert info["cube_side_log2"] == W.tree.lmin - 1 and info["order"] == rt.ORDER_MORTON, info
        if W.tree.lmax > W.tree.lmin:
            assert g.level_order(W.tree.lmax)["cube_side_log2"] == 0
            with pytest.raises(rt.EmmaRTError):
                g.set_level_order(rt.ORDER_RASTER, W.tree.lmax)
        g.set_level_order(order, W.tree.lmin)
        for _ in range(2):
            g.coarse_step(1.0, 1)
        assert g.stats()["status"] == 0
        res[order] = {l: g.get(l) for l in range(W.tree.lmin, W.tree.lmax + 1)}
        g.close()
    for l in res[rt.ORDER_RASTER]:
        for k in Fields.NAMES:
            assert np.array_equal(getattr(res[rt.ORDER_RASTER][l], k), getattr(res[rt.ORDER_MORTON][l], k)), (l, k)
