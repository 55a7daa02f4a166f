"""emma_rt_dump_grid (GPU): the library's grid output of its device state is
byte for byte the reference's own dumpalloct_serial (io.c:1233-1363) of a
struct OCT tree holding the same state, for every field id the library holds,
in FP64 and mixed precision, on an AMR tree with the hydro coupling and the
reionisation times switched on."""
import filecmp
import os

import numpy as np
import pytest

from emma_amd import rt, rtio, synth
from oracle import ref_dump_grid, ref_io_available

pytestmark = pytest.mark.gpu

IDS = [0, 1, 2, 3, 4, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 27, 28, 29, 30, 31, 32, 33, 38, 39, 40]


@pytest.mark.parametrize("precision", [rt.FP64, rt.MIXED], ids=["fp64", "mixed"])
def test_dump_matches_reference_writer(gpu_lib, tmp_path, precision):
    if not ref_io_available(3):
        pytest.fail("reference io build (oracle/_ref/libemma_ref_io_g3.so) missing")
    W = synth.amr_workload(3, [60, 100, 80], "Starburst99_1grpT_3grpS", seed=13, sigma=1.5, nsrc=2, grid=16)
    levels = range(W.tree.lmin, W.tree.lmax + 1)
    g = rt.RadGPU(W.params, precision)
    g.setup(W.params, W.tree, W.fields, tphy=3.0e6)
    g.enable_xion_times(True)
    opos, H = {}, {}
    rng = np.random.default_rng(5)
    for l in levels:
        opos[l] = synth.oct_corners(W.tree.level(l))
        H[l] = synth.hydro_like(W.fields[l], rng)
        g.set_oct_pos(l, opos[l])
        g.set_hydro(l, H[l])
    g.set_coupling(1)
    g.coarse_step(1.0, 1)
    mine, ref = str(tmp_path / "mine"), str(tmp_path / "ref")
    os.makedirs(mine)
    os.makedirs(ref)
    g.dump_grid(mine, 4, IDS, rank=1, tsim=2.5e6)
    state = {l: g.get(l) for l in levels}
    hyd = {}
    for l in levels:
        p, dX, E = g.get_hydro(l)
        hyd[l] = (H[l].d, H[l].u, H[l].v, H[l].w, p, dX)
    xion = {l: g.xion_times(l) for l in levels}
    ref_dump_grid(3, ref, 4, IDS, W.tree, state, opos, rank=1, tsim=2.5e6, hydro=hyd, xion=xion)
    for fid in IDS:
        a, b = rtio.path(mine, fid, 4, 1), rtio.path(ref, fid, 4, 1)
        assert filecmp.cmp(a, b, shallow=False), (fid, rtio.FIELD_NAMES[fid])
    for fid in (0, 4, 18, 31, 32, 39):  # the same values through emma_rt_grid_field
        assert np.array_equal(g.grid_field(fid, rank=1), rtio.read(rtio.path(mine, fid, 4, 1))["data"]), fid
    r = rtio.read(rtio.path(mine, 32, 4, 1))
    assert r["n"] == sum(int(((W.tree.level(l).child < 0) | (l == W.tree.lmax)).sum()) for l in levels)
    assert np.all((r["data"] >= 0) & (r["data"] <= 1))


def test_dump_refuses_fields_not_held(gpu_lib, tmp_path):
    W = synth.uniform_workload(3, "SB99_equiv_mono_TH", seed=1)
    g = rt.RadGPU(W.params, rt.FP64)
    g.setup(W.params, W.tree, W.fields)
    with pytest.raises(rt.EmmaRTError):  # no oct corners yet
        g.dump_grid(str(tmp_path), 0, [18])
    g.set_oct_pos(3, synth.oct_corners(W.tree.level(3)))
    for bad in ([5], [22], [13], [39], [41], [42], [-1]):  # gravity, group 1 of 1, no hydro, no xion, rho_first
        with pytest.raises(rt.EmmaRTError):
            g.dump_grid(str(tmp_path), 0, bad)
    g.dump_grid(str(tmp_path), 0, [0, 18, 32, 38])
    assert rtio.read(rtio.path(str(tmp_path), 18, 0))["n"] == W.tree.level(3).ncell
