"""The checker: a CPU restatement of EMMA's radiation + chemistry hot path
(``emma_oracle.c``) and, where it was built, the reference's own CPU functions
(``_ref/libemma_ref_g{1,3}.so``, see ``Makefile``).

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s cpu_baseline leg, where it is the thing results are
compared against (or timed as the CPU baseline) -- never the product path.

Both expose the same level-stepping interface so that tests can drive them
side by side:

* ``clean_new(l)``                      -- ``clean_new_rad`` (rad_utils.c:2050)
* ``radsolver(l, dt, aexp, chemonly)``  -- ``RadSolver`` (rad_utils.c:1861)
* ``coarse_step(aexp)``                 -- the RADSTEP level recursion of
  ``Advance_level`` (advanceamr.c:243-1300): per level, clean, recurse into
  the finer level (two sub-steps), then RadSolver.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from emma_amd.synth import Fields

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "libemma_oracle.so")
REF_SO = {1: os.path.join(HERE, "_ref", "libemma_ref_g1.so"), 3: os.path.join(HERE, "_ref", "libemma_ref_g3.so"),
          9: os.path.join(HERE, "_ref", "libemma_ref_g9.so")}
REF_SO_CR = {1: os.path.join(HERE, "_ref", "libemma_ref_g1_cr.so"), 3: os.path.join(HERE, "_ref", "libemma_ref_g3_cr.so")}
REF_SO_HYD = {1: os.path.join(HERE, "_ref", "libemma_ref_g1_hyd.so"),
              3: os.path.join(HERE, "_ref", "libemma_ref_g3_hyd.so"),
              9: os.path.join(HERE, "_ref", "libemma_ref_g9_hyd.so")}
# non-periodic domain faces: TRANS* on all six faces (G=1); plus REF* on the
# minus faces (G=3)
REF_SO_TRANS = {1: os.path.join(HERE, "_ref", "libemma_ref_g1_tr.so"),
                3: os.path.join(HERE, "_ref", "libemma_ref_g3_trref.so")}
# GPUAXL builds: the reference's RadSolver takes its GPU branch, advanceradGPU
# and the device set-up coming from emma_amd/dropin/rad_utils_hip.c over the
# HIP library (RADSTEP, and COARSERAD)
REF_SO_GPU = {1: os.path.join(HERE, "_ref", "libemma_ref_g1_gpu.so"),
              3: os.path.join(HERE, "_ref", "libemma_ref_g3_gpu.so")}
REF_SO_GPU_CR = {1: os.path.join(HERE, "_ref", "libemma_ref_g1_gpucr.so"),
                 3: os.path.join(HERE, "_ref", "libemma_ref_g3_gpucr.so")}
# STARS builds: FillRad's sources from star particles (stars_sources); the
# *dec build adds DECREASE_EMMISIVITY_AFTER_TLIFE
REF_SO_STARS = {3: os.path.join(HERE, "_ref", "libemma_ref_g3_stars.so"),
                9: os.path.join(HERE, "_ref", "libemma_ref_g9_stars.so")}
REF_SO_STARS_DEC = {9: os.path.join(HERE, "_ref", "libemma_ref_g9_starsdec.so")}

# the serial grid output of the reference (io.c, dumpalloct_serial) driven by io_driver.c
REF_SO_IO = {1: os.path.join(HERE, "_ref", "libemma_ref_io_g1.so"), 3: os.path.join(HERE, "_ref", "libemma_ref_io_g3.so")}

RADSTEP = 0
COARSERAD = 1

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


class _OrcFields(ctypes.Structure):
    _fields_ = [("ncell", ctypes.c_int)] + [(k, _dp) for k in Fields.NAMES]


class _OrcHydro(ctypes.Structure):
    _fields_ = [(k, _dp) for k in ("d", "u", "v", "w", "p", "dX", "E")]


class _OrcStars(ctypes.Structure):
    _fields_ = [("on", ctypes.c_int), ("decrease", ctypes.c_int), ("srcint", ctypes.c_double),
                ("tlife", ctypes.c_double), ("nstar", ctypes.c_int), ("cell", _ip), ("mass", _dp), ("age", _dp),
                ("radiative_state", _ip), ("is_star", _ip)]


class _OrcStar(ctypes.Structure):
    _fields_ = [("on", ctypes.c_int)] + [(k, ctypes.c_double) for k in
                                        ("x", "y", "z", "mass", "lifetime", "srcint")]


def _chydro(H) -> _OrcHydro:
    s = _OrcHydro()
    for k in _OrcHydro._fields_:
        a = getattr(H, k[0])
        assert a.dtype == np.float64 and a.flags.c_contiguous
        setattr(s, k[0], _d(a))
    return s


def _cfields(F: Fields) -> _OrcFields:
    s = _OrcFields()
    s.ncell = F.ncell
    for k in Fields.NAMES:
        a = getattr(F, k)
        assert a.dtype == np.float64 and a.flags.c_contiguous
        setattr(s, k, _d(a))
    return s


def build(ref: bool = False) -> None:
    """Compile the oracle (and the reference driver when asked)."""
    import subprocess

    targets = ["oracle"] + (["ref"] if ref else [])
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


class Oracle:
    """ctypes front end of emma_oracle.c plus a host-side level state."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            build()
        self.lib = ctypes.CDLL(path)
        L = self.lib
        L.orc_eddington.restype = ctypes.c_double
        L.orc_eddington.argtypes = [ctypes.c_double] * 5 + [ctypes.c_int, ctypes.c_int]
        L.orc_chem_cells.argtypes = [ctypes.c_void_p, ctypes.POINTER(_OrcFields), ctypes.c_double, ctypes.c_double,
                                     ctypes.c_int, _ip, _ip]
        L.orc_level_advance.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, _ip, ctypes.POINTER(_OrcFields),
                                        ctypes.POINTER(_OrcFields), ctypes.POINTER(_OrcFields),
                                        ctypes.POINTER(_OrcFields), ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_int, _ip, _ip]
        L.orc_level_final.argtypes = [ctypes.c_int, ctypes.c_int, _ip, ctypes.POINTER(_OrcFields),
                                      ctypes.POINTER(_OrcFields), ctypes.POINTER(_OrcFields), ctypes.c_double,
                                      _dp, _dp, _dp, _dp]
        L.orc_level_fluxes.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _ip, ctypes.POINTER(_OrcFields),
                                       ctypes.POINTER(_OrcFields), ctypes.c_double, _dp]
        L.orc_eddington_n.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_double, _ip, _ip, _dp]
        L.orc_rates_n.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_e2t_n.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, ctypes.c_double, _dp]
        L.orc_dtrad.restype = ctypes.c_double
        L.orc_dtrad.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_void_p]
        L.orc_getcellnei.argtypes = [ctypes.c_int, _ip, _ip]
        L.orc_level_advance_mode.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip, _ip,
                                             ctypes.POINTER(_OrcFields), ctypes.POINTER(_OrcFields),
                                             ctypes.POINTER(_OrcFields), ctypes.POINTER(_OrcFields), ctypes.c_double,
                                             ctypes.c_double, ctypes.c_double, ctypes.c_int, _ip, _ip]
        L.orc_level_final_mode.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip,
                                           ctypes.POINTER(_OrcFields), ctypes.POINTER(_OrcFields),
                                           ctypes.POINTER(_OrcFields), ctypes.c_double, _dp, _dp, _dp, _dp]
        L.orc_putE2coarse.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.POINTER(_OrcFields)]
        L.orc_distribE.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_fillrad_cells.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_OrcFields), _ip,
                                        _dp, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                        ctypes.c_double]
        L.orc_rad_to_hydro.argtypes = [ctypes.POINTER(_OrcFields), ctypes.POINTER(_OrcHydro)]
        self.params = None
        self.tree = None
        self.mode = RADSTEP
        self.coupling = False
        self.hydro = {}
        self.opos = {}
        self.star = None
        self.stars = {}
        self.star_params = None

    # ---- scalar / vector helpers
    def eddington(self, fx, fy, fz, e, c, i, j):
        fx, fy, fz, e = [np.ascontiguousarray(a, dtype=np.float64) for a in (fx, fy, fz, e)]
        n = fx.size
        i = np.ascontiguousarray(np.broadcast_to(i, (n,)), dtype=np.int32)
        j = np.ascontiguousarray(np.broadcast_to(j, (n,)), dtype=np.int32)
        out = np.empty(n)
        self.lib.orc_eddington_n(n, _d(fx), _d(fy), _d(fz), _d(e), c, _i(i), _i(j), _d(out))
        return out

    def rates(self, temp, x, nH):
        temp, x, nH = [np.ascontiguousarray(np.broadcast_to(a, np.shape(temp)), dtype=np.float64) for a in (temp, x, nH)]
        n = temp.size
        outs = [np.empty(n) for _ in range(5)]
        self.lib.orc_rates_n(n, _d(temp), _d(x), _d(nH), *[_d(o) for o in outs])
        return dict(zip(("alpha_a", "alpha_b", "beta", "lambda", "tcool"), outs))

    def e2t(self, params, nhplus, eint, nh, aexp=1.0):
        cp = params.to_c()
        nhplus, eint, nh = [np.ascontiguousarray(a, dtype=np.float64) for a in (nhplus, eint, nh)]
        out = np.empty(nhplus.size)
        self.lib.orc_e2t_n(ctypes.byref(cp), nhplus.size, _d(nhplus), _d(eint), _d(nh), aexp, _d(out))
        return out

    def getcellnei(self, icell):
        vn = np.zeros(6, np.int32)
        vc = np.zeros(6, np.int32)
        self.lib.orc_getcellnei(icell, _i(vn), _i(vc))
        return vn.tolist(), vc.tolist()

    def dt_rad(self, params, level, aexp=1.0):
        cp = params.to_c()
        return self.lib.orc_dtrad(level, aexp, ctypes.byref(cp))

    def chem_cells(self, params, F: Fields, dtnew, aexp=1.0, chemonly=0):
        """chemrad on every cell of F (all treated as leaves); returns
        (new Fields, nsub per cell, status, bad cell)."""
        cp = params.to_c()
        out = F.copy()
        cf = _cfields(out)
        nsub = np.zeros(out.ncell, np.int32)
        bad = np.full(1, -1, np.int32)
        st = self.lib.orc_chem_cells(ctypes.byref(cp), ctypes.byref(cf), dtnew, aexp, chemonly, _i(nsub), _i(bad))
        return out, nsub, st, int(bad[0])

    def level_fluxes(self, params, level, F: Fields, coarse: Fields | None, aexp=1.0):
        L = self.tree.level(level) if self.tree is not None else None
        return self._fluxes(params.ngrp, L, F, coarse, params.cloc(aexp))

    def _fluxes(self, G, L, F, coarse, c):
        out = np.zeros(L.ncell * 30 * G)
        cf = _cfields(F)
        cc = _cfields(coarse) if coarse is not None else None
        nei = np.ascontiguousarray(L.nei_oct.reshape(-1), dtype=np.int32)
        ch = np.ascontiguousarray(L.child, dtype=np.int32)
        st = self.lib.orc_level_fluxes(G, L.noct, _i(nei), _i(ch), ctypes.byref(cf),
                                       ctypes.byref(cc) if cc is not None else None, c, _d(out))
        if st != 0:
            raise RuntimeError(f"orc_level_fluxes status {st}")
        return out.reshape(L.ncell, 6, G, 5)

    # ---- level state interface (mirrors the reference driver)
    def setup(self, params, tree, fields: dict, tphy: float = 0.0, mode: int = RADSTEP):
        self.params = params
        self.cparams = params.to_c()
        self.tree = tree
        self.mode = mode
        self.old = {l: f.copy() for l, f in fields.items()}
        self.new = {l: f.copy() for l, f in fields.items()}
        self.t_first = {l: np.full(f.ncell, -1.0) for l, f in fields.items()}
        self.t_last = {l: np.full(f.ncell, -1.0) for l, f in fields.items()}
        self.tphy = tphy
        self.nsub = {}
        self.coupling = False
        self.hydro = {}
        self.opos = {}
        self.star = None
        self.stars = {}
        self.star_params = None

    def get(self, level, which=0) -> Fields:
        return (self.new if which else self.old)[level].copy()

    def set(self, level, which, F: Fields):
        (self.new if which else self.old)[level] = F.copy()

    def clean_new(self, level):
        self.new[level] = self.old[level].copy()

    def advancerad(self, level, dtnew, aexp=1.0, chemonly=0):
        t = self.tree
        L = t.level(level)
        G = self.params.ngrp
        co = self.old.get(level - 1) if level > t.lmin else None
        cn = self.new.get(level - 1) if level > t.lmin else None
        a, b = _cfields(self.old[level]), _cfields(self.new[level])
        ca = _cfields(co) if co is not None else None
        cb = _cfields(cn) if cn is not None else None
        nsub = np.zeros(L.ncell, np.int32)
        bad = np.full(1, -1, np.int32)
        nei = np.ascontiguousarray(L.nei_oct.reshape(-1), dtype=np.int32)
        ch = np.ascontiguousarray(L.child, dtype=np.int32)
        dx = 0.5**level
        st = self.lib.orc_level_advance_mode(ctypes.byref(self.cparams), self.mode, int(level == t.lmin), L.noct,
                                             _i(nei), _i(ch), ctypes.byref(a), ctypes.byref(b),
                                             ctypes.byref(ca) if ca else None, ctypes.byref(cb) if cb else None, dx,
                                             dtnew, aexp, chemonly, _i(nsub), _i(bad))
        self.nsub[level] = nsub
        if st != 0:
            raise OracleError(st, int(bad[0]))
        return L.noct

    def final(self, level, chemonly=0):
        t = self.tree
        L = t.level(level)
        fo = self.old.get(level + 1)
        ch = np.ascontiguousarray(L.child, dtype=np.int32)
        a, b = _cfields(self.old[level]), _cfields(self.new[level])
        c = _cfields(fo) if fo is not None else None
        tf, tl = self.t_first[level], self.t_last[level]
        ftf = self.t_first.get(level + 1)
        ftl = self.t_last.get(level + 1)
        self.lib.orc_level_final_mode(self.params.ngrp, self.mode, int(level == t.lmin), chemonly, L.noct, _i(ch),
                                      ctypes.byref(a), ctypes.byref(b), ctypes.byref(c) if c else None, self.tphy,
                                      _d(tf), _d(tl), _d(ftf) if ftf is not None else None,
                                      _d(ftl) if ftl is not None else None)
        if self.mode == COARSERAD and level == t.lmin:
            self._levels_call(self.lib.orc_distribE)

    def _levels_call(self, fn, *extra):
        t = self.tree
        ls = range(t.lmin, t.lmax + 1)
        childs = [np.ascontiguousarray(t.level(l).child, dtype=np.int32) for l in ls]
        cf = [_cfields(self.old[l]) for l in ls]
        arr_c = (ctypes.POINTER(ctypes.c_int) * len(childs))(*[_i(c) for c in childs])
        arr_f = (ctypes.POINTER(_OrcFields) * len(cf))(*[ctypes.pointer(x) for x in cf])
        self._keep = (childs, cf, arr_c, arr_f)
        fn(self.params.ngrp, t.lmin, len(childs), ctypes.cast(arr_c, ctypes.c_void_p),
           ctypes.cast(arr_f, ctypes.c_void_p), *extra)

    # ---- hydro coupling (WRADHYD) and the WRADTEST source
    def set_hydro(self, level, H):
        self.hydro[level] = H.copy()

    def get_hydro(self, level):
        H = self.hydro[level]
        return H.p.copy(), H.dX.copy(), H.E.copy()

    def set_coupling(self, hydro=True):
        self.coupling = bool(hydro)

    def set_oct_pos(self, level, xyz):
        self.opos[level] = np.ascontiguousarray(xyz, dtype=np.float64)

    def set_stromgren_source(self, star):
        self.star = star
        if star is not None:
            self.star_params = None

    def set_stars(self, level, S):
        self.stars[level] = S

    def set_star_sources(self, srcint, tlife, decrease_after_tlife=False, enable=True):
        self.star_params = (float(srcint), float(tlife), bool(decrease_after_tlife)) if enable else None
        if enable:
            self.star = None

    def fillrad(self, level, aexp=1.0):
        """FillRad: the COARSERAD block (split coarse cells gather their leaves'
        E, F), then the per-cell hydro copy / test source when enabled"""
        if self.mode == COARSERAD and level == self.tree.lmin:
            nc = _cfields(self.new[level])
            self._levels_call(self.lib.orc_putE2coarse, ctypes.byref(nc))
        if not (self.coupling or self.star is not None or self.star_params is not None):
            return
        L = self.tree.level(level)
        if L.noct == 0:
            return
        ch = np.ascontiguousarray(L.child, dtype=np.int32)
        H = _chydro(self.hydro[level]) if self.coupling else None
        st = None
        if self.star is not None:
            st = _OrcStar(1, *[float(getattr(self.star, k)) for k in
                               ("x", "y", "z", "mass", "lifetime", "srcint")])
        sp = None
        if self.star_params is not None and level in self.stars:
            a = self.stars[level].arrays()
            self._star_arrays = a
            sp = _OrcStars(1, int(self.star_params[2]), self.star_params[0], self.star_params[1], self.stars[level].n,
                           _i(a[0]), _d(a[1]), _d(a[2]), _i(a[3]), _i(a[4]))
        elif self.star_params is not None:  # no particles on this level: sources cleared
            sp = _OrcStars(1, int(self.star_params[2]), self.star_params[0], self.star_params[1], 0,
                           None, None, None, None, None)
        opos = self.opos.get(level, np.zeros(3 * L.noct))
        for F in (self.old[level], self.new[level]):  # rfield and rfieldnew
            cf = _cfields(F)
            self.lib.orc_fillrad_cells(ctypes.byref(self.cparams), level, self.tree.lmin, ctypes.byref(cf), _i(ch),
                                       _d(opos), ctypes.byref(H) if H is not None else None,
                                       ctypes.byref(st) if st is not None else None,
                                       ctypes.byref(sp) if sp is not None else None, self.tphy, aexp)

    def radsolver(self, level, dtnew, aexp=1.0, chemonly=0):
        n = self.advancerad(level, dtnew, aexp, chemonly)
        if n > 0:
            self.final(level, chemonly)
            if self.coupling:
                cf = _cfields(self.old[level])
                ch = _chydro(self.hydro[level])
                self.lib.orc_rad_to_hydro(ctypes.byref(cf), ctypes.byref(ch))

    def coarse_step(self, aexp=1.0):
        if self.mode == COARSERAD:
            coarse_step_coarserad(self, self.tree, self.params, aexp)
        else:
            coarse_step(self, self.tree, self.params, aexp)


class OracleError(RuntimeError):
    def __init__(self, status, cell):
        super().__init__(f"oracle status {status} at cell {cell}")
        self.status = status
        self.cell = cell


def coarse_step(solver, tree, params, aexp=1.0):
    """One coarse step of the RADSTEP level recursion (advanceamr.c:243-1300):
    each level cleans rfieldnew, advances the finer level by two half steps,
    then runs FillRad and RadSolver with its own dt_rad."""

    def adv(l):
        dt = params.dt_rad(l, aexp)
        for _ in range(1 if l == tree.lmin else 2):
            solver.clean_new(l)
            if l < tree.lmax and tree.level(l + 1).noct > 0:
                adv(l + 1)
            solver.fillrad(l, aexp)  # advanceamr.c:1035
            solver.radsolver(l, dt, aexp, 0)

    adv(tree.lmin)


def coarse_step_coarserad(solver, tree, params, aexp=1.0):
    """One coarse step of the COARSERAD level recursion (Advance_level_RAD,
    advanceamr.c:1309-1537): every level takes one step of the coarse level's
    dt_rad; per level: clean, recurse into the finer level, FillRad (the
    coarse level gathers its leaves' E, F), RadSolver with chemonly=0 (the
    transport runs on the coarse level only)."""
    dt = params.dt_rad(tree.lmin, aexp)

    def adv(l):
        solver.clean_new(l)
        if l < tree.lmax and tree.level(l + 1).noct > 0:
            adv(l + 1)
        solver.fillrad(l, aexp)
        solver.radsolver(l, dt, aexp, 0)

    adv(tree.lmin)


def _ref_table(coarserad: bool, hydro: bool, boundary: bool, gpu: bool = False, stars: int = 0) -> dict:
    if stars:
        return REF_SO_STARS_DEC if stars == 2 else REF_SO_STARS
    if gpu:
        return REF_SO_GPU_CR if coarserad else REF_SO_GPU
    return REF_SO_TRANS if boundary else REF_SO_HYD if hydro else REF_SO_CR if coarserad else REF_SO


def ref_available(ngrp: int, coarserad: bool = False, hydro: bool = False, boundary: bool = False,
                  gpu: bool = False, stars: int = 0) -> bool:
    return os.path.exists(_ref_table(coarserad, hydro, boundary, gpu, stars).get(ngrp, ""))


class Reference:
    """The reference's own CPU functions, driven through ref_driver.c.
    coarserad: the COARSERAD build; hydro: the WRADHYD build (FillRad with the
    hydro copy and the Stromgren source, RadSolver's injection into the hydro
    fields); boundary: the builds with non-periodic domain faces (REF_SO_TRANS),
    which decide from the octs' positions (set_oct_pos); gpu: the GPUAXL
    builds, whose RadSolver reaches the GPU through advanceradGPU of the
    drop-in file (set_params then also runs emma.c's device set-up)."""

    def __init__(self, ngrp: int, coarserad: bool = False, hydro: bool = False, boundary: bool = False,
                 gpu: bool = False, stars: int = 0):
        path = _ref_table(coarserad, hydro, boundary, gpu, stars)[ngrp]
        self.gpu = gpu
        self.stars_build = stars
        self.mode = COARSERAD if coarserad else RADSTEP
        self.hydro_build = hydro
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.lib = ctypes.CDLL(path)
        L = self.lib
        assert L.refdrv_ngrp() == ngrp
        self.ngrp = ngrp
        L.refdrv_set_params.argtypes = [_dp, ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp, _dp, _dp, _dp,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp]
        L.refdrv_build.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _ip, _ip, _ip]
        L.refdrv_set_level.argtypes = [ctypes.c_int, ctypes.c_int] + [_dp] * 9
        L.refdrv_get_level.argtypes = [ctypes.c_int, ctypes.c_int] + [_dp] * 9
        L.refdrv_get_txion.argtypes = [ctypes.c_int, _dp, _dp]
        L.refdrv_clean_new_rad.argtypes = [ctypes.c_int, ctypes.c_double]
        L.refdrv_radsolver.restype = ctypes.c_double
        L.refdrv_radsolver.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.refdrv_advancerad.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.refdrv_eddington.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_double, _ip, _ip, _dp]
        L.refdrv_getcellnei.argtypes = [ctypes.c_int, _ip, _ip]
        L.refdrv_e2t.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_double, _dp]
        L.refdrv_layout.argtypes = [ctypes.POINTER(ctypes.c_long)]
        L.refdrv_stencil_sweeps.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                            ctypes.c_double]
        L.refdrv_stencil_update.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                            ctypes.c_double]
        L.refdrv_stencil_chem.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_int]
        L.refdrv_set_tphy.argtypes = [ctypes.c_double]
        L.refdrv_rates_n.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.refdrv_fillrad_coarse.argtypes = [ctypes.c_int, ctypes.c_double]
        L.refdrv_set_oct_pos.argtypes = [ctypes.c_int, _dp]
        assert L.refdrv_coarserad() == (1 if coarserad else 0)
        assert L.refdrv_wradhyd() == (1 if hydro else 0)
        assert L.refdrv_gpuaxl() == (1 if gpu else 0)
        assert L.refdrv_stars() == (1 if stars else 0)
        if gpu:
            L.refdrv_gpu_init.argtypes = [ctypes.c_int]
        if stars:
            L.refdrv_set_star_params.argtypes = [ctypes.c_double, ctypes.c_double]
            L.refdrv_set_stars.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _dp, _dp, _ip, _ip]
            L.refdrv_fillrad.argtypes = [ctypes.c_int, ctypes.c_double]
        if hydro:
            L.refdrv_set_hydro.argtypes = [ctypes.c_int] + [_dp] * 6
            L.refdrv_get_hydro.argtypes = [ctypes.c_int] + [_dp] * 3
            L.refdrv_set_stromgren.argtypes = [ctypes.c_double] * 6
            L.refdrv_fillrad.argtypes = [ctypes.c_int, ctypes.c_double]
        self.stride = 4096

    def layout(self):
        out = np.zeros(22, np.int64)
        self.lib.refdrv_layout(out.ctypes.data_as(ctypes.POINTER(ctypes.c_long)))
        return out

    def set_params(self, p, tphy=0.0):
        G = self.ngrp
        assert p.ngrp == G
        u = p.units7()
        arr = [np.ascontiguousarray(np.asarray(getattr(p, k), dtype=np.float64)) for k in ("hnu", "alphae", "alphai", "factgrp")]
        tb = np.ascontiguousarray(np.asarray(p.time_bound, dtype=np.float64))
        self._keep = (u, arr, tb)
        self.lib.refdrv_set_params(_d(u), p.clight, p.fudgecool, int(p.ncvgcool), *[_d(a) for a in arr], p.H0, p.om,
                                   p.ov, tphy, _d(tb))
        if self.gpu:  # create_param_GPU snapshots the parameters (rad_utils_gpu.cu:57-106)
            assert self.lib.refdrv_gpu_init(self.stride) == 0, "no GPU for the GPUAXL build"

    def setup(self, params, tree, fields: dict, tphy: float = 0.0):
        self.params = params
        self.tree = tree
        noct, parent, neicell, child = tree.flat()
        self._tree_arrays = (noct, parent, neicell, child)
        st = self.lib.refdrv_build(tree.lmin, tree.lmax, _i(noct), _i(parent), _i(neicell), _i(child))
        assert st == 0
        self.set_params(params, tphy)
        for l, F in fields.items():
            self.set(l, 0, F)
            self.set(l, 1, F)

    def set(self, level, which, F: Fields):
        F = F.copy()
        self.lib.refdrv_set_level(level, which, *[_d(getattr(F, k)) for k in Fields.NAMES])

    def get(self, level, which=0) -> Fields:
        n = self.tree.level(level).ncell
        F = Fields.zeros(self.ngrp, n)
        self.lib.refdrv_get_level(level, which, *[_d(getattr(F, k)) for k in Fields.NAMES])
        return F

    def txion(self, level):
        n = self.tree.level(level).ncell
        a, b = np.empty(n), np.empty(n)
        self.lib.refdrv_get_txion(level, _d(a), _d(b))
        return a, b

    def clean_new(self, level, aexp=1.0):
        self.lib.refdrv_clean_new_rad(level, aexp)

    def radsolver(self, level, dtnew, aexp=1.0, chemonly=0):
        self.lib.refdrv_radsolver(level, dtnew, aexp, chemonly, self.stride)

    def advancerad(self, level, dtnew, aexp=1.0, chemonly=0):
        return self.lib.refdrv_advancerad(level, dtnew, aexp, chemonly, self.stride)

    def coarse_step(self, aexp=1.0):
        if self.mode == COARSERAD:
            coarse_step_coarserad(self, self.tree, self.params, aexp)
        else:
            coarse_step(self, self.tree, self.params, aexp)

    def set_stars(self, level, S):
        a = S.arrays()
        self._stars_keep = a
        assert self.lib.refdrv_set_stars(level, S.n, _i(a[0]), _d(a[1]), _d(a[2]), _i(a[3]), _i(a[4])) == 0

    def set_star_sources(self, srcint, tlife, decrease_after_tlife=False, enable=True):
        assert enable and bool(decrease_after_tlife) == (self.stars_build == 2), \
            "star sources and DECREASE_EMMISIVITY_AFTER_TLIFE are compile-time flags of the reference"
        self.lib.refdrv_set_star_params(float(srcint), float(tlife))

    def fillrad(self, level, aexp=1.0):
        if self.hydro_build or self.stars_build:
            self.lib.refdrv_fillrad(level, aexp)
        elif self.mode == COARSERAD:
            self.lib.refdrv_fillrad_coarse(level, aexp)

    # ---- WRADHYD build only
    def set_hydro(self, level, H):
        a = [np.ascontiguousarray(getattr(H, k), dtype=np.float64) for k in ("d", "u", "v", "w", "p", "dX")]
        self.lib.refdrv_set_hydro(level, *[_d(x) for x in a])

    def get_hydro(self, level):
        n = self.tree.level(level).ncell
        p, dX, E = np.empty(n), np.empty(n), np.empty(n)
        self.lib.refdrv_get_hydro(level, _d(p), _d(dX), _d(E))
        return p, dX, E

    def set_coupling(self, hydro=True):
        assert hydro == self.hydro_build, "the coupling is a compile-time flag of the reference (WRADHYD)"

    def set_oct_pos(self, level, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        self.lib.refdrv_set_oct_pos(level, _d(xyz))

    def set_stromgren_source(self, star):
        self.lib.refdrv_set_stromgren(star.x, star.y, star.z, star.mass, star.lifetime, star.srcint)

    def eddington(self, fx, fy, fz, e, c, i, j):
        fx, fy, fz, e = [np.ascontiguousarray(a, dtype=np.float64) for a in (fx, fy, fz, e)]
        n = fx.size
        i = np.ascontiguousarray(np.broadcast_to(i, (n,)), dtype=np.int32)
        j = np.ascontiguousarray(np.broadcast_to(j, (n,)), dtype=np.int32)
        out = np.empty(n)
        self.lib.refdrv_eddington(n, _d(fx), _d(fy), _d(fz), _d(e), c, _i(i), _i(j), _d(out))
        return out

    def getcellnei(self, icell):
        vn = np.zeros(6, np.int32)
        vc = np.zeros(6, np.int32)
        self.lib.refdrv_getcellnei(icell, _i(vn), _i(vc))
        return vn.tolist(), vc.tolist()

    def e2t(self, nhplus, eint, nh, aexp=1.0):
        nhplus, eint, nh = [np.ascontiguousarray(a, dtype=np.float64) for a in (nhplus, eint, nh)]
        out = np.empty(nhplus.size)
        self.lib.refdrv_e2t(nhplus.size, _d(nhplus), _d(eint), _d(nh), aexp, _d(out))
        return out

    def rates(self, temp, x, nH):
        temp, x, nH = [np.ascontiguousarray(np.broadcast_to(a, np.shape(temp)), dtype=np.float64) for a in (temp, x, nH)]
        n = temp.size
        outs = [np.empty(n) for _ in range(5)]
        self.lib.refdrv_rates_n(n, _d(temp), _d(x), _d(nH), *[_d(o) for o in outs])
        return dict(zip(("alpha_a", "alpha_b", "beta", "lambda", "tcool"), outs))

    # struct RGRID batches (numpy arrays of emma_amd.rgrid.rgrid_dtype), in place
    def stencil_sweeps(self, st, dx, dt, c):
        self.lib.refdrv_stencil_sweeps(st.ctypes.data_as(ctypes.c_void_p), st.shape[0], dx, dt, c)

    def stencil_update(self, st, dx, dt, c):
        self.lib.refdrv_stencil_update(st.ctypes.data_as(ctypes.c_void_p), st.shape[0], dx, dt, c)

    def stencil_chem(self, st, dx, dt, aexp=1.0, chemonly=0):
        self.lib.refdrv_stencil_chem(st.ctypes.data_as(ctypes.c_void_p), st.shape[0], dx, dt, aexp, chemonly)

    def stencil_advance(self, st, dx, dt, aexp=1.0, chemonly=0):
        """advancerad's per-batch work (rad_utils.c:1814-1830) on a gathered batch"""
        c = self.params.cloc(aexp)
        self.stencil_sweeps(st, dx, dt, c)
        self.stencil_update(st, dx, dt, c)
        self.stencil_chem(st, dx, dt, aexp, chemonly)


def ref_io_available(ngrp: int) -> bool:
    return os.path.exists(REF_SO_IO.get(ngrp, ""))


def ref_dump_grid(ngrp: int, folder: str, ndump: int, field_ids, tree, fields: dict, opos: dict, rank: int = 0,
                  tsim: float = 0.0, hydro: dict = None, xion: dict = None) -> None:
    """The reference's own dumpalloct_serial (io.c:1233-1363) on a struct OCT
    tree built from `tree`, with the levels' Fields, oct corners {level:
    (noct, 3)}, optional hydro {level: (d, u, v, w, p, dX)} and reionisation
    times {level: (t_first, t_last)}: checker for emma_rt_dump_grid."""
    lib = ctypes.CDLL(REF_SO_IO[ngrp])
    dp = ctypes.POINTER(ctypes.c_double)
    lib.ref_dump_grid.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_int, _ip, _ip, _dp, ctypes.POINTER(dp), ctypes.c_int, _ip]
    levels = list(range(tree.lmin, tree.lmax + 1))
    noct = np.array([tree.level(l).noct for l in levels], dtype=np.int32)
    child = np.ascontiguousarray(np.concatenate([tree.level(l).child for l in levels]), dtype=np.int32)
    pos = np.ascontiguousarray(np.concatenate([np.asarray(opos[l], dtype=np.float64).reshape(-1, 3)
                                               for l in levels]))
    keep = []  # arrays referenced by the pointer table
    ptrs = (dp * (16 * len(levels)))()
    for k, l in enumerate(levels):
        F = fields[l]
        cols = [F.e, F.fx, F.fy, F.fz, F.src, F.nhplus, F.nh, F.temp]
        cols += list(xion[l]) if xion and l in xion else [None, None]
        cols += list(hydro[l]) if hydro and l in hydro else [None] * 6
        for j, a in enumerate(cols):
            if a is None:
                continue
            a = np.ascontiguousarray(a, dtype=np.float64)
            keep.append(a)
            ptrs[k * 16 + j] = a.ctypes.data_as(dp)
    ids = np.ascontiguousarray(np.asarray(field_ids, dtype=np.int32))
    if not folder.endswith("/"):
        folder += "/"
    st = lib.ref_dump_grid(folder.encode(), int(ndump), int(rank), float(tsim), tree.lmin, tree.lmax, ngrp,
                           _i(noct), _i(child), _d(pos), ptrs, int(ids.size), _i(ids))
    if st != 0:
        raise RuntimeError(f"ref_dump_grid: {st}")
