"""Parity of the HIP resident-level path with the oracle (GPU).

FP64 mode: every field within relative 1e-5 of the oracle (the oracle is
bit-identical to the reference CPU path, see test_oracle_ref.py); flux
components are compared relative to c*N of their cell, since a flux
component can cross zero.
Mixed mode (FP32 photon density/flux/source, FP64 chemistry): x_HII and T
within 1e-4 relative, N within 1e-4 relative, F within 1e-4 of c*N.
"""
import numpy as np
import pytest

from conftest import max_rel
from emma_amd import synth, rt
from emma_amd.synth import Fields

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-5
MIXED_TOL = 1e-4


def field_errors(A: Fields, B: Fields, c: float):
    """relative errors of A against the reference B"""
    err = {}
    err["e"] = max_rel(A.e, B.e)
    scale = np.maximum(c * B.e, 1e-300)
    for k in ("fx", "fy", "fz"):
        err[k] = float(np.max(np.abs(getattr(A, k) - getattr(B, k)) / scale))
    err["x"] = max_rel(A.nhplus / A.nh, B.nhplus / B.nh)
    err["eint"] = max_rel(A.eint, B.eint)
    err["temp"] = max_rel(A.temp, B.temp)
    err["src"] = max_rel(A.src, B.src)
    err["nh"] = max_rel(A.nh, B.nh)
    return err


def run_pair(oracle, W, precision, steps, mode="radsolver"):
    """mode: "radsolver" (host-driven level steps), "graph" or "persistent"
    (emma_rt_coarse_step under that schedule)"""
    g = rt.RadGPU(W.params, precision)
    g.setup(W.params, W.tree, W.fields)
    if mode == "persistent":
        g.set_schedule(rt.SCHED_PERSISTENT)
    oracle.setup(W.params, W.tree, W.fields)
    for _ in range(steps):
        if mode == "radsolver":
            # the per-level recursion driven from the host, one RadSolver call per level step
            from oracle import coarse_step

            coarse_step(g, W.tree, W.params)
        else:
            g.coarse_step(1.0, 1)
        oracle.coarse_step()
    return g


@pytest.mark.parametrize("table,G", [("SB99_equiv_mono_TH", 1), ("Starburst99_1grpT_3grpS", 3)])
def test_uniform_fp64(gpu_lib, oracle, table, G):
    W = synth.uniform_workload(4, table, seed=21, nsrc=4)
    g = run_pair(oracle, W, rt.FP64, steps=3)
    c = W.params.cloc()
    err = field_errors(g.get(4), oracle.get(4), c)
    print("uniform fp64 G=%d" % G, err, g.stats())
    assert max(err.values()) < FP64_TOL, err


@pytest.mark.parametrize("table,G", [("SB99_equiv_mono_TH", 1), ("Starburst99_1grpT_3grpS", 3)])
@pytest.mark.parametrize("mode", ["radsolver", "graph", "persistent"])
def test_amr_fp64(gpu_lib, oracle, table, G, mode):
    W = synth.amr_workload(3, [60, 100, 80], table, seed=11, sigma=1.5, nsrc=2, grid=16)
    g = run_pair(oracle, W, rt.FP64, steps=2, mode=mode)
    c = W.params.cloc()
    for l in range(W.tree.lmin, W.tree.lmax + 1):
        err = field_errors(g.get(l), oracle.get(l), c)
        print("amr fp64 G=%d level %d" % (G, l), err)
        assert max(err.values()) < FP64_TOL, (l, err)


def test_mixed_bound(gpu_lib, oracle):
    W = synth.uniform_workload(5, "Starburst99_1grpT_3grpS", seed=5, nsrc=8)
    g = run_pair(oracle, W, rt.MIXED, steps=4)
    err = field_errors(g.get(5), oracle.get(5), W.params.cloc())
    print("mixed", err)
    assert err["x"] < MIXED_TOL and err["temp"] < MIXED_TOL and err["eint"] < MIXED_TOL, err
    assert err["e"] < MIXED_TOL, err
    assert max(err["fx"], err["fy"], err["fz"]) < MIXED_TOL, err


@pytest.mark.parametrize("table,G", [("SB99_equiv_mono_TH", 1), ("Starburst99_1grpT_3grpS", 3)])
@pytest.mark.parametrize("mode", ["graph", "persistent"])
def test_amr_mixed_bound(gpu_lib, oracle, table, G, mode):
    """mixed precision on a 4-level tree: partial LDS tiles, coarse faces,
    FP32 reflux accumulators"""
    W = synth.amr_workload(3, [60, 100, 80], table, seed=11, sigma=1.5, nsrc=2, grid=16)
    g = run_pair(oracle, W, rt.MIXED, steps=2, mode=mode)
    for l in range(W.tree.lmin, W.tree.lmax + 1):
        err = field_errors(g.get(l), oracle.get(l), W.params.cloc())
        print("amr mixed G=%d level %d" % (G, l), err)
        assert err["x"] < MIXED_TOL and err["temp"] < MIXED_TOL and err["e"] < MIXED_TOL, (l, err)
        assert max(err["fx"], err["fy"], err["fz"]) < MIXED_TOL, (l, err)


def test_photon_budget_transport_only(gpu_lib):
    """Transport conserves sum N dx^3 on a periodic level up to the chemistry
    absorption: with n_H -> tiny and x -> 1 the absorption vanishes and the FP64
    budget is conserved to round-off."""
    W = synth.uniform_workload(4, "SB99_equiv_mono_TH", seed=3)
    F = W.fields[4]
    F.nh[:] = 1e-12
    F.nhplus[:] = F.nh * (1 - 1e-9)
    F.eint[:] = W.params.eint_code(F.nh, 1.0, 1e4)
    g = rt.RadGPU(W.params, rt.FP64)
    g.setup(W.params, W.tree, W.fields)
    p0 = float(np.sum(F.e) * 0.5 ** (3 * 4))
    for _ in range(5):
        g.radsolver(4, W.params.dt_rad(4))
    p1 = g.stats()["photons"]
    assert abs(p1 / p0 - 1) < 1e-12, (p0, p1)


def test_status_emin(gpu_lib):
    W = synth.uniform_workload(3, "SB99_equiv_mono_TH", seed=1)
    W.fields[3].e[0, 5] = 1e-30  # below EMIN after transport? set all neighbours low too
    W.fields[3].e[:] = 1e-30
    W.fields[3].fx[:] = W.fields[3].fy[:] = W.fields[3].fz[:] = 0
    g = rt.RadGPU(W.params, rt.FP64)
    g.setup(W.params, W.tree, W.fields)
    with pytest.raises(rt.EmmaRTError) as ei:
        g.radsolver(3, W.params.dt_rad(3))
    assert ei.value.status == rt.ERR_EMIN


def test_status_noconv(gpu_lib):
    W = synth.uniform_workload(3, "SB99_equiv_mono_TH", seed=1)
    W.params.fudgecool = 1e-21  # the controller's floor test fires on the first sub-step
    g = rt.RadGPU(W.params, rt.FP64)
    g.setup(W.params, W.tree, W.fields)
    with pytest.raises(rt.EmmaRTError) as ei:
        g.radsolver(3, W.params.dt_rad(3))
    assert ei.value.status == rt.ERR_NOCONV


def test_status_negative_photons(gpu_lib):
    W = synth.uniform_workload(3, "SB99_equiv_mono_TH", seed=1)
    F = W.fields[3]
    F.e[0, 9] = -1.0
    g = rt.RadGPU(W.params, rt.FP64)
    g.setup(W.params, W.tree, W.fields)
    with pytest.raises(rt.EmmaRTError) as ei:
        g.radsolver(3, W.params.dt_rad(3))
    assert ei.value.status == rt.ERR_NEG_PHOTON


@pytest.mark.parametrize("table,G", [("SB99_equiv_mono_TH", 1), ("Starburst99_1grpT_3grpS", 3)])
def test_level_size_limit(gpu_lib, table, G):
    """A level of more than min(2^29, (2^31-1)/ngrp) cells is refused before
    any table is read or any memory allocated (32-bit cell indices and byte
    offsets); the context stays usable."""
    import ctypes
    W = synth.uniform_workload(3, table, seed=1)
    g = rt.RadGPU(W.params, rt.MIXED)
    ip = ctypes.POINTER(ctypes.c_int)
    dummy = np.zeros(8, dtype=np.int32)
    limit = min(2**29, (2**31 - 1) // G)
    for noct in (limit // 8 + 1, 2**28):
        n = np.array([noct], dtype=np.int32)
        st = g.lib.emma_rt_tree_create(g.h, 3, 3, n.ctypes.data_as(ip), dummy.ctypes.data_as(ip),
                                       dummy.ctypes.data_as(ip), rt.MIXED)
        assert st == rt.ERR_ARG, (noct, st)
    g.setup(W.params, W.tree, W.fields)
    g.radsolver(3, W.params.dt_rad(3))
    g.close()


@pytest.mark.parametrize("table,G", [("SB99_equiv_mono_TH", 1), ("Starburst99_1grpT_3grpS", 3)])
@pytest.mark.parametrize("mode", ["host", "graph", "persistent"])
@pytest.mark.parametrize("precision", [rt.FP64, rt.MIXED])
@pytest.mark.parametrize("faces", ["periodic", "trans_ref"])
def test_coarserad(gpu_lib, oracle, table, G, mode, precision, faces):
    """COARSERAD: transport on the coarse level only, chemistry on every leaf,
    putE2coarse before and distribE after the coarse step; also with the
    domain faces of EMMA's Stromgren setup (TRANS* on all six, REF* on the
    minus faces), the flag set init/stromgen/param.mk combines with
    COARSERAD."""
    from emma_amd import tree as T
    from oracle import COARSERAD, coarse_step_coarserad

    W = synth.amr_workload(3, [60, 100, 80], table, seed=20 + G, sigma=1.5, nsrc=3, grid=16)
    if faces == "trans_ref":
        tb = T.with_boundaries(W.tree, trans=range(6), reflect=(0, 2, 4))
        W = synth.Workload(W.name, W.params, tb, W.fields, W.nsub)
    g = rt.RadGPU(W.params, precision)
    g.setup(W.params, W.tree, W.fields, mode=rt.COARSERAD)
    if mode == "persistent":
        g.set_schedule(rt.SCHED_PERSISTENT)
    oracle.setup(W.params, W.tree, W.fields, mode=COARSERAD)
    for _ in range(3):
        if mode in ("graph", "persistent"):
            g.coarse_step(1.0, 1)
        else:
            coarse_step_coarserad(g, W.tree, W.params)
        oracle.coarse_step()
    tol = FP64_TOL if precision == rt.FP64 else MIXED_TOL
    for l in range(W.tree.lmin, W.tree.lmax + 1):
        err = field_errors(g.get(l), oracle.get(l), W.params.cloc())
        print("coarserad G=%d prec=%d level %d" % (G, precision, l), err)
        assert err["x"] < tol and err["temp"] < tol and err["e"] < tol and err["eint"] < tol, (l, err)
        assert max(err["fx"], err["fy"], err["fz"]) < tol, (l, err)


def coupled_pair(oracle, W, precision, star, tphy, mode, steps=2, seed=7):
    rng = np.random.default_rng(seed)
    hyd = {l: synth.hydro_like(F, rng) for l, F in W.fields.items()}
    g = rt.RadGPU(W.params, precision)
    g.setup(W.params, W.tree, W.fields, tphy)
    if mode == "persistent":
        g.set_schedule(rt.SCHED_PERSISTENT)
    oracle.setup(W.params, W.tree, W.fields, tphy)
    for s in (g, oracle):
        s.set_coupling(True)
        s.set_stromgren_source(star)
        for l, H in hyd.items():
            s.set_hydro(l, H)
            s.set_oct_pos(l, synth.oct_corners(W.tree.level(l)))
    from oracle import coarse_step

    for _ in range(steps):
        if mode == "radsolver":
            coarse_step(g, W.tree, W.params)  # FillRad + RadSolver per level step from the host
        else:
            g.coarse_step(1.0, 1)
        oracle.coarse_step()
    return g


@pytest.mark.parametrize("table,G", [("SB99_equiv_mono_TH", 1), ("Starburst99_1grpT_3grpS", 3)])
@pytest.mark.parametrize("mode", ["radsolver", "graph", "persistent"])
@pytest.mark.parametrize("precision", [rt.FP64, rt.MIXED])
def test_hydro_coupling_and_source(gpu_lib, oracle, table, G, mode, precision):
    """FillRad (hydro -> radiation, Stromgren source, E2T) before every level
    step and the injection into p, dX, E after it, on a 4-level tree."""
    W = synth.amr_workload(3, [60, 100, 80], table, seed=13, sigma=1.5, grid=16)
    g = coupled_pair(oracle, W, precision, synth.Star(x=0.5, y=0.5, z=0.5), 1e-3, mode)
    tol = FP64_TOL if precision == rt.FP64 else MIXED_TOL
    c = W.params.cloc()
    for l in range(W.tree.lmin, W.tree.lmax + 1):
        err = field_errors(g.get(l), oracle.get(l), c)
        herr = [max_rel(a, b) for a, b in zip(g.get_hydro(l), oracle.get_hydro(l))]
        print("coupled G=%d %s level %d" % (G, mode, l), err, herr)
        assert max(err.values()) < tol, (l, err)
        assert max(herr) < tol, (l, herr)



@pytest.mark.parametrize("years", [1e6, 1e7, 1e8])
def test_source_time_groups(gpu_lib, oracle, years):
    """Nine groups = three time groups (0 / 3.6 / 36 Myr) of three frequency
    groups: the device source fills only the three groups of the time group
    the run's time lies in (src_utils.c:366-392), as the oracle does (pinned
    to the reference's NGRP_TIME = 3 WRADHYD build, test_coupling.py)."""
    W = synth.amr_workload(3, [60, 100, 80], "Starburst99_3grpT_3grpS", seed=13, sigma=1.5, grid=16)
    tphy = years / (W.params.unit_t / 3.1536e13 * 1e6)
    g = coupled_pair(oracle, W, rt.FP64, synth.Star(x=0.5, y=0.5, z=0.5), tphy, "graph")
    it = [1e6, 1e7, 1e8].index(years)
    lit = 0
    for l in range(W.tree.lmin, W.tree.lmax + 1):
        A = g.get(l)
        err = field_errors(A, oracle.get(l), W.params.cloc())
        assert max(err.values()) < FP64_TOL, (l, err)
        on = set(np.nonzero(np.any(A.src > 0, axis=1))[0].tolist())
        assert on <= {3 * it, 3 * it + 1, 3 * it + 2}, (years, on)
        lit += int(np.count_nonzero(A.src[3 * it] > 0))
    assert lit > 0


@pytest.mark.parametrize("precision", [rt.FP64, rt.MIXED])
def test_save_restore_state(gpu_lib, precision):
    """emma_rt_save_state / emma_rt_restore_state: a coarse step repeated from
    the restored device copy gives the same state bit for bit, whatever ran
    in between (including a different buffer parity)."""
    W = synth.amr_workload(3, [60, 100, 80], "Starburst99_1grpT_3grpS", seed=13, sigma=1.5, nsrc=3, grid=16)
    g = rt.RadGPU(W.params, precision)
    g.setup(W.params, W.tree, W.fields)
    with pytest.raises(rt.EmmaRTError):
        g.restore_state()  # nothing saved yet
    g.coarse_step(1.0, 1)
    g.save_state()
    g.coarse_step(1.0, 1)
    A = {l: g.get(l) for l in W.fields}
    g.coarse_step(1.0, 3)
    g.restore_state()
    g.coarse_step(1.0, 1)
    for l in W.fields:
        B = g.get(l)
        for k in Fields.NAMES:
            assert np.array_equal(getattr(A[l], k), getattr(B, k)), (l, k)


@pytest.mark.parametrize("G", [3, 9])
@pytest.mark.parametrize("precision", [rt.FP64, rt.MIXED])
@pytest.mark.parametrize("decrease", [False, True], ids=["steady", "decrease"])
def test_star_sources(gpu_lib, oracle, G, precision, decrease):
    """Star particles as sources (stars_sources) on the device: FillRad's
    sources against the oracle (pinned to the reference's STARS builds,
    test_stars.py), then two coarse steps with FillRad before every level
    step (captured graph) against the oracle's."""
    from test_stars import star_setup, workload

    W, stars = workload(G, 33)
    g = rt.RadGPU(W.params, precision)
    for s in (g, oracle):
        star_setup(s, W, stars, decrease)
    srctol = 1e-12 if precision == rt.FP64 else 1e-7
    for l in range(W.tree.lmin, W.tree.lmax + 1):
        g.fillrad(l)
        oracle.fillrad(l)
        a, b = g.get(l).src, oracle.get(l).src
        assert np.count_nonzero(b) > 0 or W.tree.leaf_cells(l) == 0
        assert np.array_equal(a == 0, b == 0), l
        assert max_rel(a, b) < srctol, (l, max_rel(a, b))
    for _ in range(2):
        g.coarse_step(1.0, 1)
        oracle.coarse_step()
    tol = FP64_TOL if precision == rt.FP64 else MIXED_TOL
    for l in range(W.tree.lmin, W.tree.lmax + 1):
        err = field_errors(g.get(l), oracle.get(l), W.params.cloc())
        print("stars G=%d prec=%d level %d" % (G, precision, l), err)
        assert err["x"] < tol and err["temp"] < tol and err["e"] < tol and err["eint"] < tol, (l, err)
        assert max(err["fx"], err["fy"], err["fz"]) < tol, (l, err)


def outlier_errors(A: Fields, B: Fields, c: float):
    """per field: (max relative error, number of cells beyond MIXED_TOL)"""
    out = {}
    xa, xb = A.nhplus / A.nh, B.nhplus / B.nh
    pairs = {"e": (A.e, B.e, np.abs(B.e)), "x": (xa, xb, np.abs(xb)), "eint": (A.eint, B.eint, np.abs(B.eint)),
             "temp": (A.temp, B.temp, np.abs(B.temp))}
    for k in ("fx", "fy", "fz"):
        pairs[k] = (getattr(A, k), getattr(B, k), np.maximum(c * B.e, 1e-300))
    for k, (a, b, scale) in pairs.items():
        r = np.abs(a - b) / np.maximum(scale, 1e-300)
        if r.ndim == 2:
            r = r.max(0)
        out[k] = (float(r.max()), int(np.count_nonzero(r > MIXED_TOL)))
    return out


@pytest.mark.parametrize("config", ["c3", "c4", "c5"])
def test_full_size_mixed_vs_fp64(gpu_lib, config):
    """At the benchmark sizes (C3: 128^3, 3 groups, 64 sources; C4: AMR levels
    7-11; C5: 256^3 + two refined levels, 4096 sources, c/100), the
    mixed-precision tiled kernels against the FP64 kernels (pinned to the
    oracle at small sizes) after two sub-cycles (C3) or one coarse step of the
    level recursion (C4, C5).  The chemistry sub-step controller of the
    reference is discontinuous (test_conditioning.py: a 1e-6 perturbation of
    the state moves single cells of C3 by 0.5 % in the FP64 oracle itself), so
    the bound is on the cells: at most 1 in 20,000 (and at least 5 allowed per
    level: a cell that flips also moves its neighbours in later sub-steps of
    the level) beyond the mixed tolerance, none beyond 5 %."""
    W = {"c3": lambda: synth.config_c3(nsub=2), "c4": lambda: synth.config_c4(nsub=1),
         "c5": lambda: synth.config_c5(nsub=1)}[config]()
    runs = {}
    for prec in (rt.FP64, rt.MIXED):
        g = rt.RadGPU(W.params, prec)
        g.setup(W.params, W.tree, W.fields)
        g.coarse_step(1.0, W.nsub if config == "c3" else 1)
        runs[prec] = {l: g.get(l) for l in range(W.tree.lmin, W.tree.lmax + 1)}
        g.close()
    c = W.params.cloc()
    for l in range(W.tree.lmin, W.tree.lmax + 1):
        err = outlier_errors(runs[rt.MIXED][l], runs[rt.FP64][l], c)
        n = runs[rt.FP64][l].ncell
        print(config, "level", l, n, err)
        for k, (mx, nbad) in err.items():
            assert mx < 0.05, (l, k, mx)
            assert nbad <= max(5, n // 20000), (l, k, nbad)


def test_timing_modes(gpu_lib):
    """TIMING_GRAPH brackets a coarse step's graph replay with one event pair,
    TIMING_LAUNCH every level-step kernel of an eager coarse step; both give
    the same state, a positive device time and the same step counts, and the
    number of event-bracketed spans tells which path ran."""
    W = synth.uniform_workload(4, "SB99_equiv_mono_TH", seed=2)
    out = {}
    for mode in (rt.TIMING_LAUNCH, rt.TIMING_GRAPH):
        g = rt.RadGPU(W.params, rt.MIXED)
        g.setup(W.params, W.tree, W.fields)
        g.coarse_step(1.0, 4)  # capture (graph mode) or eager
        g.set(4, 0, W.fields[4])
        g.enable_timing(mode)
        g.coarse_step(1.0, 4)
        st = g.stats()
        assert st["kernel_ms"] > 0 and st["level_steps"] == 4, (mode, st)
        assert st["timed_spans"] == (4 if mode == rt.TIMING_LAUNCH else 1), (mode, st)
        out[mode] = g.get(4)
        with pytest.raises(rt.EmmaRTError):
            g.enable_timing(3)
        g.close()
    for k in Fields.NAMES:
        assert np.array_equal(getattr(out[rt.TIMING_LAUNCH], k), getattr(out[rt.TIMING_GRAPH], k)), k


def refined_block_workload(table, seed):
    """Level 3 complete, its 4^3 cells at the origin refined into a block of
    64 level-4 octs (two 32-oct tiles): every wave of level 4 has at most 64
    halo faces, so the level runs the complete-level kernel variant, and the
    waves on the block's surface have coarse faces (injection, refluxing)."""
    from emma_amd import tree as T
    from emma_amd.params import make_params
    rng = np.random.default_rng(seed)
    L3 = T.uniform_level(3)
    split = np.nonzero(np.all(L3.cell_pos() < 4, axis=1))[0]
    L4 = T.refine(L3, split)
    tr = T.build_tree([L3, L4])
    params = make_params(table, clight=0.1)
    fields = {}
    for l in (3, 4):
        L = tr.level(l)
        fields[l] = synth.random_fields(params, np.exp(rng.normal(0.0, 1.0, L.ncell)), rng)
    synth.place_sources(params, fields[4], tr.level(4), 3, np.array([1e49, 1e50, 3e50]), rng)
    return synth.Workload("refined-block", params, tr, fields, 1)


def halo_faces_per_wave(L):
    """leaving faces per wave of 64 cells whose neighbour lies outside the
    cell's 32-oct tile, and whether any of them is a coarse face"""
    nei = L.nei_oct.reshape(-1, 6)
    cell = np.arange(L.ncell)
    oct, ic, oct0 = cell >> 3, cell & 7, (cell // 256) * 32
    cnt = np.zeros(L.ncell, int)
    coarse = np.zeros(L.ncell, bool)
    for d in range(3):
        no = nei[oct, 2 * d + ((ic >> d) & 1)]
        cnt += ~((no >= oct0) & (no < oct0 + 32))
        coarse |= no < 0
    return cnt.reshape(-1, 64).sum(1), coarse.reshape(-1, 64).any(1)


@pytest.mark.parametrize("table,G", [("SB99_equiv_mono_TH", 1), ("Starburst99_1grpT_3grpS", 3)])
@pytest.mark.parametrize("precision", [rt.FP64, rt.MIXED])
def test_refined_block_complete_level_kernel(gpu_lib, oracle, table, G, precision):
    W = refined_block_workload(table, seed=31 + G)
    per_wave, has_coarse = halo_faces_per_wave(W.tree.level(4))
    assert per_wave.max() <= 64 and has_coarse.any()
    g = run_pair(oracle, W, precision, steps=3, mode="graph")
    c = W.params.cloc()
    for l in (3, 4):
        err = field_errors(g.get(l), oracle.get(l), c)
        print("refined block G=%d prec=%d level %d" % (G, precision, l), err)
        if precision == rt.FP64:
            assert max(err.values()) < FP64_TOL, (l, err)
        else:
            assert err["x"] < MIXED_TOL and err["temp"] < MIXED_TOL and err["e"] < MIXED_TOL, (l, err)
            assert max(err["fx"], err["fy"], err["fz"]) < MIXED_TOL, (l, err)


@pytest.mark.parametrize("precision", [rt.FP64, rt.MIXED])
@pytest.mark.parametrize("mode", ["radsolver", "graph", "persistent"])
def test_nine_groups(gpu_lib, oracle, precision, mode):
    """Nine photon groups (Starburst99_3grpT_3grpS: three time groups of three
    frequency groups) run the kernels' generic group count; the oracle agrees
    with the reference's NGRP=9 build bit for bit (test_oracle_ref.py)."""
    W = synth.amr_workload(3, [60, 100, 80], "Starburst99_3grpT_3grpS", seed=13, sigma=1.5, nsrc=3, grid=16)
    assert W.params.ngrp == 9
    g = run_pair(oracle, W, precision, steps=2, mode=mode)
    c = W.params.cloc()
    for l in range(W.tree.lmin, W.tree.lmax + 1):
        err = field_errors(g.get(l), oracle.get(l), c)
        print("nine groups prec=%d %s level %d" % (precision, mode, l), err)
        if precision == rt.FP64:
            assert max(err.values()) < FP64_TOL, (l, err)
        else:
            assert err["x"] < MIXED_TOL and err["temp"] < MIXED_TOL and err["e"] < MIXED_TOL, (l, err)
            assert max(err["fx"], err["fy"], err["fz"]) < MIXED_TOL, (l, err)


@pytest.mark.parametrize("table,G", [("SB99_equiv_mono_TH", 1), ("Starburst99_1grpT_3grpS", 3),
                                     ("Starburst99_3grpT_3grpS", 9)])
@pytest.mark.parametrize("precision", [rt.FP64, rt.MIXED])
@pytest.mark.parametrize("sched", [rt.SCHED_GRAPH, rt.SCHED_PERSISTENT], ids=["graph", "persistent"])
def test_amr_bit_reproducible(gpu_lib, table, G, precision, sched):
    """Refluxing gives every (coarse leaf, face) pair its own accumulators
    with one writer per launch, summed in pair order by the coarse leaf: two
    runs of a 4-level tree (per-level transport, two coarse steps: host-driven
    level steps, then graph replays) agree bit for bit on every level, in
    both precisions and for the tiled (1, 3 groups) and per-cell (9 groups)
    mixed kernels."""
    W = synth.amr_workload(3, [60, 100, 80], table, seed=17, sigma=1.5, nsrc=3, grid=16)
    runs = []
    for _ in range(2):
        g = rt.RadGPU(W.params, precision)
        g.setup(W.params, W.tree, W.fields)
        from oracle import coarse_step

        coarse_step(g, W.tree, W.params)
        g.set_schedule(sched)
        g.coarse_step(1.0, 1)
        g.coarse_step(1.0, 1)
        runs.append({l: g.get(l) for l in range(W.tree.lmin, W.tree.lmax + 1)})
        g.close()
    for l in runs[0]:
        for k in Fields.NAMES:
            assert np.array_equal(getattr(runs[0][l], k), getattr(runs[1][l], k)), (l, k)


@pytest.mark.parametrize("table,G,reflect", [("SB99_equiv_mono_TH", 1, ()), ("Starburst99_1grpT_3grpS", 3, (0, 2, 4)),
                                             ("Starburst99_3grpT_3grpS", 9, (0, 2, 4))])
@pytest.mark.parametrize("precision", [rt.FP64, rt.MIXED])
@pytest.mark.parametrize("amr", [False, True])
@pytest.mark.parametrize("mode", ["graph", "persistent"])
def test_domain_boundaries(gpu_lib, oracle, table, G, reflect, precision, amr, mode):
    """Non-periodic domain faces (EMMA_RT_NEI_TRANSMISSIVE / _REFLECTIVE, the
    reference's TRANS*/REF* gather; the oracle agrees with the reference's
    TRANS builds bit for bit, test_oracle_ref.py): a complete 16^3 level (the
    tiled kernel's listed halo) and a 4-level tree (its per-direction halo,
    refined octs on the domain faces, no refluxing through them); 9 groups run
    the per-cell kernel.  Two coarse steps, graph mode."""
    from emma_amd import tree as T

    if amr:
        W = synth.amr_workload(3, [60, 100, 80], table, seed=29, sigma=1.5, nsrc=4, grid=16)
    else:
        W = synth.uniform_workload(4, table, seed=29, nsrc=6)
    tb = T.with_boundaries(W.tree, trans=range(6), reflect=reflect)
    W = synth.Workload(W.name, W.params, tb, W.fields, W.nsub)
    g = run_pair(oracle, W, precision, steps=2, mode=mode)
    c = W.params.cloc()
    for l in range(W.tree.lmin, W.tree.lmax + 1):
        err = field_errors(g.get(l), oracle.get(l), c)
        print("boundaries G=%d prec=%d amr=%d %s level %d" % (G, precision, amr, mode, l), err)
        if precision == rt.FP64:
            assert max(err.values()) < FP64_TOL, (l, err)
        else:
            assert err["x"] < MIXED_TOL and err["temp"] < MIXED_TOL and err["e"] < MIXED_TOL, (l, err)
            assert max(err["fx"], err["fy"], err["fz"]) < MIXED_TOL, (l, err)


def test_mixed_photon_floor(gpu_lib, oracle):
    """Cells whose photon density the chemistry clamps to its floor
    (EMIN * factgrp / unit_N, chem_utils.c:382, 772) keep running: stored in
    FP32, the floor 1e-11 would round down below itself (float(1e-11) <
    1e-11) and stop the next step at the floor test; the mixed path stores
    the smallest float at or above the floor instead, after the chemistry and
    when the state is set from FP64.  A level at the floor
    with no sources, four sub-cycles, against the oracle (one group: with
    several, group 0's floor EMIN * factgrp[0] lies below the test's EMIN and
    the reference itself stops there)."""
    from emma_amd.params import e_min_code

    W = synth.uniform_workload(3, "SB99_equiv_mono_TH", seed=31)
    F = W.fields[3]
    F.e[:] = e_min_code(W.params)
    F.fx[:] = 0.0
    F.fy[:] = 0.0
    F.fz[:] = 0.0
    g = run_pair(oracle, W, rt.MIXED, steps=4, mode="graph")
    err = field_errors(g.get(3), oracle.get(3), W.params.cloc())
    assert err["x"] < MIXED_TOL and err["temp"] < MIXED_TOL and err["e"] < MIXED_TOL, err
