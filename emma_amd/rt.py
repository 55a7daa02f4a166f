"""Host mirror of EMMA's radiation-step interface over libemma_rt.so.

The reference's host calls ``RadSolver(level, param, firstoct, cpu, stencil,
stride, dtnew, aexp, chemonly)`` (``rad_utils.c:1861``) once per level and
sub-step, after ``clean_new_rad`` (``rad_utils.c:2050``), in the level
recursion of ``Advance_level`` (``advanceamr.c:243``).  :class:`RadGPU` keeps
those names and argument meanings over a device-resident tree:

* ``radsolver(level, dtnew, aexp, chemonly)`` -- one level step on the GPU;
* ``clean_new(level)`` -- nothing to do (the device state is double
  buffered), kept so host loops read like the reference's;
* ``coarse_step(aexp, ncoarse)`` -- the whole RADSTEP recursion of
  ``ncoarse`` coarse steps as one HIP graph.

Errors the reference turns into ``abort()`` come back as
:class:`EmmaRTError` with the library's status code.  There is no CPU
fallback: without the built library (``python -m emma_amd.build``) every
entry point raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .params import RunParams
from .synth import Fields
from . import tree as T

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libemma_rt.so")

OK = 0
ERR_ARG = -1
ERR_HIP = -2
ERR_NEG_PHOTON = -3
ERR_EMIN = -4
ERR_NOCONV = -5
ERR_NEIGHBOR = -6
ERR_LAYOUT = -7
ERR_NODEVICE = -8
ERR_STATE = -9
ERR_SCHED = -10

FP64 = 0
MIXED = 1

RADSTEP = 0
COARSERAD = 1

TIMING_OFF = 0
TIMING_LAUNCH = 1
TIMING_GRAPH = 2

VARIANT_AUTO = -1
VARIANT_LISTED = 0
VARIANT_PERDIR = 1
VARIANT_LATENCY = 2
VARIANT_NAMES = {VARIANT_LISTED: "listed", VARIANT_PERDIR: "perdir", VARIANT_LATENCY: "latency"}

ORDER_MORTON = 0
ORDER_RASTER = 1

SCHED_GRAPH = 0
SCHED_PERSISTENT = 1
SCHED_NAMES = {SCHED_GRAPH: "graph", SCHED_PERSISTENT: "persistent"}

EXPORTS = [
    "emma_rt_countdevices", "emma_rt_initlocaldevice", "emma_rt_checkdevice", "emma_rt_strerror",
    "emma_rt_create", "emma_rt_set_params", "emma_rt_destroy",
    "emma_rt_rgrid_layout", "emma_rt_check_layout", "emma_rt_create_radstencil", "emma_rt_destroy_radstencil",
    "emma_rt_advance_stencil", "emma_rt_create_pinned_stencil", "emma_rt_destroy_pinned_stencil",
    "emma_rt_tree_create", "emma_rt_set_level", "emma_rt_get_level", "emma_rt_enable_xion_times",
    "emma_rt_get_xion_times", "emma_rt_set_tphy", "emma_rt_radsolver", "emma_rt_coarse_step",
    "emma_rt_get_stats", "emma_rt_enable_timing", "emma_rt_enable_counters", "emma_rt_set_mode", "emma_rt_fillrad",
    "emma_rt_set_hydro", "emma_rt_get_hydro", "emma_rt_set_coupling", "emma_rt_set_oct_pos",
    "emma_rt_set_stromgren_source", "emma_rt_get_level_times", "emma_rt_alloc_pinned",
    "emma_rt_save_state", "emma_rt_restore_state", "emma_rt_set_stars", "emma_rt_set_star_sources",
    "emma_rt_set_chem_budget", "emma_rt_get_chem_budget", "emma_rt_set_level_variant", "emma_rt_get_level_tuning",
    "emma_rt_advance_stencil_noalloc", "emma_rt_set_chem_keep", "emma_rt_dump_grid", "emma_rt_grid_field",
    "emma_rt_set_schedule", "emma_rt_get_schedule", "emma_rt_plan_schedule", "emma_rt_set_level_order",
    "emma_rt_get_level_order",
]

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)


class Stats(ctypes.Structure):
    _fields_ = [
        ("cell_updates", ctypes.c_longlong),
        ("chem_substeps", ctypes.c_longlong),
        ("level_steps", ctypes.c_longlong),
        ("kernel_ms", ctypes.c_double),
        ("photons", ctypes.c_double),
        ("status", ctypes.c_int),
        ("bad_level", ctypes.c_int),
        ("bad_cell", ctypes.c_int),
        ("timed_spans", ctypes.c_int),
        ("deferred_cells", ctypes.c_longlong),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SchedInfo(ctypes.Structure):
    _fields_ = [
        ("items", ctypes.c_longlong),
        ("deps", ctypes.c_longlong),
        ("blocks", ctypes.c_int),
        ("launches", ctypes.c_int),
        ("est_span", ctypes.c_double),
        ("est_chain", ctypes.c_double),
        ("est_work", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class EmmaRTError(RuntimeError):
    def __init__(self, status: int, msg: str, stats=None):
        super().__init__(f"emma_rt status {status}: {msg}")
        self.status = status
        self.stats = stats


_lib = None


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libemma_rt.so; raise if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} missing: build it with `python -m emma_amd.build`")
    L = ctypes.CDLL(path)
    L.emma_rt_strerror.restype = ctypes.c_char_p
    L.emma_rt_countdevices.argtypes = [ctypes.c_int, _ip]
    L.emma_rt_initlocaldevice.argtypes = [ctypes.c_int, ctypes.c_int]
    L.emma_rt_checkdevice.argtypes = [ctypes.c_int, _ip]
    L.emma_rt_create.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
    L.emma_rt_set_params.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.emma_rt_destroy.argtypes = [ctypes.c_void_p]
    L.emma_rt_rgrid_layout.argtypes = [ctypes.c_int, ctypes.c_int, _lp]
    L.emma_rt_check_layout.argtypes = [ctypes.c_int, ctypes.c_int, _lp]
    L.emma_rt_create_radstencil.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    L.emma_rt_destroy_radstencil.argtypes = [ctypes.c_void_p]
    L.emma_rt_advance_stencil.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int]
    L.emma_rt_create_pinned_stencil.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                                ctypes.POINTER(ctypes.c_void_p)]
    L.emma_rt_destroy_pinned_stencil.argtypes = [ctypes.c_void_p]
    L.emma_rt_tree_create.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _ip, _ip, _ip, ctypes.c_int]
    L.emma_rt_set_level.argtypes = [ctypes.c_void_p, ctypes.c_int] + [_dp] * 9
    L.emma_rt_get_level.argtypes = [ctypes.c_void_p, ctypes.c_int] + [_dp] * 9
    L.emma_rt_enable_xion_times.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.emma_rt_get_xion_times.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
    L.emma_rt_set_tphy.argtypes = [ctypes.c_void_p, ctypes.c_double]
    L.emma_rt_radsolver.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int]
    L.emma_rt_coarse_step.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int]
    L.emma_rt_get_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(Stats)]
    L.emma_rt_enable_timing.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.emma_rt_enable_counters.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.emma_rt_set_mode.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.emma_rt_fillrad.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
    L.emma_rt_set_hydro.argtypes = [ctypes.c_void_p, ctypes.c_int] + [_dp] * 6
    L.emma_rt_get_hydro.argtypes = [ctypes.c_void_p, ctypes.c_int] + [_dp] * 3
    L.emma_rt_set_coupling.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.emma_rt_set_oct_pos.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
    L.emma_rt_set_stromgren_source.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_double] * 6
    L.emma_rt_get_level_times.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _ip]
    L.emma_rt_set_stars.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _ip, _dp, _dp, _ip, _ip]
    L.emma_rt_set_star_sources.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_int]
    L.emma_rt_save_state.argtypes = [ctypes.c_void_p]
    L.emma_rt_restore_state.argtypes = [ctypes.c_void_p]
    L.emma_rt_set_chem_budget.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    L.emma_rt_get_chem_budget.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip]
    L.emma_rt_set_level_variant.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    L.emma_rt_advance_stencil_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                                  ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                                  ctypes.c_int]
    L.emma_rt_get_level_tuning.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, _ip, _ip, _dp]
    L.emma_rt_set_chem_keep.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
    L.emma_rt_dump_grid.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                    ctypes.c_int, _ip]
    L.emma_rt_grid_field.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_float),
                                     ctypes.c_longlong, ctypes.POINTER(ctypes.c_longlong)]
    L.emma_rt_set_schedule.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.emma_rt_set_level_order.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    L.emma_rt_get_level_order.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, _ip]
    L.emma_rt_get_schedule.argtypes = [ctypes.c_void_p, _ip, ctypes.POINTER(SchedInfo)]
    _llp = ctypes.POINTER(ctypes.c_longlong)
    L.emma_rt_plan_schedule.argtypes = [ctypes.c_int, _ip, _ip, _ip, ctypes.c_int, _ip, _dp, _ip, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_longlong, ctypes.c_longlong, _llp, _llp, _ip, _ip, _ip,
                                        _ip, ctypes.POINTER(ctypes.c_ubyte), _ip, _ip, _ip, _dp]
    _lib = L
    return L


def plan_schedule(tree: T.Tree, steps, slots: int, tile_cost=None, chunk=None, nqueues: int = 1) -> dict:
    """The persistent coarse step's planner (host code, no GPU): for the
    level indices `steps` (0 = coarsest, sequential order) of `tree`, the
    queues' ticket ranges and the items by ticket (level step, first tile,
    tiles, flags, simulated start rank) with the tickets each waits for."""
    L = load()
    lv = [tree.level(l) for l in range(tree.lmin, tree.lmax + 1)]
    noct = np.array([x.noct for x in lv], dtype=np.int32)
    nei = np.ascontiguousarray(np.concatenate([x.nei_oct.reshape(-1) for x in lv]).astype(np.int32))
    child = np.ascontiguousarray(np.concatenate([x.child.reshape(-1) for x in lv]).astype(np.int32))
    st = np.ascontiguousarray(np.asarray(steps, dtype=np.int32))
    cost = None if tile_cost is None else np.ascontiguousarray(np.asarray(tile_cost, dtype=np.float64))
    cs = None if chunk is None else np.ascontiguousarray(np.asarray(chunk, dtype=np.int32))
    ni, nd = ctypes.c_longlong(0), ctypes.c_longlong(0)
    ip = lambda a: a.ctypes.data_as(_ip)
    args0 = (len(lv), ip(noct), ip(nei), ip(child), len(st), ip(st), None if cost is None else _d(cost),
             None if cs is None else ip(cs), int(slots), int(nqueues))
    r = L.emma_rt_plan_schedule(*args0, 0, 0, ctypes.byref(ni), ctypes.byref(nd), None, None, None, None, None,
                                None, None, None, None)
    if r != OK:
        raise EmmaRTError(r, strerror(r))
    n, m = ni.value, nd.value
    item_step = np.zeros(n, np.int32)
    item_tile = np.zeros(n, np.int32)
    item_ntile = np.zeros(n, np.int32)
    item_flags = np.zeros(n, np.uint8)
    item_rank = np.zeros(n, np.int32)
    qoff = np.zeros(nqueues + 1, np.int32)
    dep_off = np.zeros(n + 1, np.int32)
    deps = np.zeros(max(m, 1), np.int32)
    est = np.zeros(3)
    r = L.emma_rt_plan_schedule(*args0, n, max(m, 1), ctypes.byref(ni), ctypes.byref(nd), ip(qoff), ip(item_step),
                                ip(item_tile), ip(item_ntile), item_flags.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)),
                                ip(item_rank), ip(dep_off), ip(deps), _d(est))
    if r != OK:
        raise EmmaRTError(r, strerror(r))
    return {"qoff": qoff, "item_step": item_step, "item_tile": item_tile, "item_ntile": item_ntile,
            "item_flags": item_flags, "item_prio": item_flags & 1, "item_rank": item_rank, "dep_off": dep_off,
            "deps": deps[:m], "est_span": est[0], "est_chain": est[1], "est_work": est[2]}


def strerror(status: int) -> str:
    return load().emma_rt_strerror(status).decode()


def rgrid_layout(ngrp: int, real_bytes: int = 8) -> np.ndarray:
    out = np.zeros(22, dtype=np.int64)
    st = load().emma_rt_rgrid_layout(ngrp, real_bytes, out.ctypes.data_as(_lp))
    if st != OK:
        raise EmmaRTError(st, strerror(st))
    return out


def device_count() -> int:
    n = ctypes.c_int(0)
    load().emma_rt_countdevices(0, ctypes.byref(n))
    return n.value


def _d(a):
    return a.ctypes.data_as(_dp)


class PinnedStencil:
    """struct RGRID records in page-locked host memory, viewed as a numpy
    array (``array``); ``free`` releases the memory (destroy_pinned_stencil_rad)."""

    def __init__(self, lib, ptr, stride: int, ngrp: int, real_bytes: int):
        from . import rgrid
        self.lib, self.ptr = lib, ptr
        dt = rgrid.rgrid_dtype(ngrp, real_bytes)
        buf = (ctypes.c_uint8 * (stride * dt.itemsize)).from_address(ptr.value)
        self.array = np.frombuffer(buf, dtype=dt, count=stride)

    def free(self):
        if self.ptr:
            self.array = None
            self.lib.emma_rt_destroy_pinned_stencil(self.ptr)
            self.ptr = None


class RadGPU:
    """A device context holding one tree of levels (the resident path)."""

    def __init__(self, params: RunParams, precision: int = FP64):
        self.lib = load()
        self.params = params
        self.precision = precision
        self._cp = params.to_c()
        h = ctypes.c_void_p()
        self._check(self.lib.emma_rt_create(ctypes.byref(self._cp), ctypes.byref(h)))
        self.h = h
        self.tree = None
        self.counters = True  # the library's default (emma_rt_enable_counters)

    def _check(self, st, stats=None):
        if st != OK:
            raise EmmaRTError(st, strerror(st), stats)
        return st

    def close(self):
        if getattr(self, "h", None):
            self.lib.emma_rt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- set-up
    def set_params(self, params: RunParams):
        self.params = params
        self._cp = params.to_c()
        self._check(self.lib.emma_rt_set_params(self.h, ctypes.byref(self._cp)))

    def load_tree(self, tree: T.Tree):
        noct = np.array(tree.noct(), dtype=np.int32)
        nei = np.ascontiguousarray(
            np.concatenate([tree.level(l).nei_oct.reshape(-1) for l in range(tree.lmin, tree.lmax + 1)]), dtype=np.int32)
        child = np.ascontiguousarray(
            np.concatenate([tree.level(l).child for l in range(tree.lmin, tree.lmax + 1)]), dtype=np.int32)
        self._check(self.lib.emma_rt_tree_create(self.h, tree.lmin, tree.lmax, noct.ctypes.data_as(_ip),
                                                 nei.ctypes.data_as(_ip), child.ctypes.data_as(_ip),
                                                 self.precision))
        self.tree = tree

    def setup(self, params: RunParams, tree: T.Tree, fields: dict, tphy: float = 0.0, mode: int = RADSTEP):
        if params is not self.params:
            self.set_params(params)
        self.set_mode(mode)
        self.load_tree(tree)
        for l, F in fields.items():
            self.set(l, 0, F)
        self.set_tphy(tphy)

    def set(self, level: int, which: int, F: Fields):
        if which != 0:
            raise ValueError("the device holds one state per level (rfield); rfieldnew is internal")
        F = F.copy()
        self._check(self.lib.emma_rt_set_level(self.h, level, *[_d(getattr(F, k)) for k in Fields.NAMES]))

    def get(self, level: int, which: int = 0) -> Fields:
        n = self.tree.level(level).ncell
        F = Fields.zeros(self.params.ngrp, n)
        self._check(self.lib.emma_rt_get_level(self.h, level, *[_d(getattr(F, k)) for k in Fields.NAMES]))
        return F

    def save_state(self):
        """device-side restart copy of every level's state"""
        self._check(self.lib.emma_rt_save_state(self.h))

    def restore_state(self):
        self._check(self.lib.emma_rt_restore_state(self.h))

    def enable_xion_times(self, on: bool = True):
        self._check(self.lib.emma_rt_enable_xion_times(self.h, int(on)))

    def xion_times(self, level: int):
        n = self.tree.level(level).ncell
        a, b = np.empty(n), np.empty(n)
        self._check(self.lib.emma_rt_get_xion_times(self.h, level, _d(a), _d(b)))
        return a, b

    def set_tphy(self, tphy: float):
        self._check(self.lib.emma_rt_set_tphy(self.h, float(tphy)))

    def enable_timing(self, mode: int = TIMING_LAUNCH):
        """TIMING_OFF, TIMING_LAUNCH (events around each level-step kernel,
        coarse steps launched eagerly) or TIMING_GRAPH (coarse steps replay
        their graph between one pair of events); True/False map to
        TIMING_LAUNCH/TIMING_OFF"""
        self._check(self.lib.emma_rt_enable_timing(self.h, int(mode)))

    def enable_counters(self, on: bool = True):
        """the kernels' device counters (stats' cell_updates, chem_substeps,
        deferred_cells; on by default).  Off, the level steps skip their
        per-tile reduction and those fields read -1; same results bit for bit"""
        self._check(self.lib.emma_rt_enable_counters(self.h, int(bool(on))))
        self.counters = bool(on)

    def set_chem_budget(self, budget: int, level: int = -1):
        """chemistry tries before a leaf goes to the straggler pass (0: off),
        for one level or every level (-1); mixed precision, 1 or 3 groups"""
        self._check(self.lib.emma_rt_set_chem_budget(self.h, int(level), int(budget)))

    def chem_budget(self, level: int) -> int:
        b = ctypes.c_int(0)
        self._check(self.lib.emma_rt_get_chem_budget(self.h, int(level), ctypes.byref(b)))
        return b.value

    def set_level_variant(self, variant: int, level: int = -1):
        """halo variant of the tiled kernel (VARIANT_*) for one level or all (-1)"""
        self._check(self.lib.emma_rt_set_level_variant(self.h, int(level), int(variant)))

    def level_tuning(self, level: int) -> dict:
        v, b, h, k = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0), ctypes.c_double(0.0)
        self._check(self.lib.emma_rt_get_level_tuning(self.h, int(level), ctypes.byref(v), ctypes.byref(b),
                                                      ctypes.byref(h), ctypes.byref(k)))
        return {"variant": v.value, "chem_budget": b.value, "halo_per_dir": bool(h.value), "chem_keep": k.value}

    def set_chem_keep(self, keep: float, level: int = -1):
        """with a chemistry budget: leaves with more than `keep` sub-steps to
        go at their pace so far stay in the first pass (0: defer all)"""
        self._check(self.lib.emma_rt_set_chem_keep(self.h, int(level), float(keep)))

    def dump_grid(self, folder: str, ndump: int, field_ids, rank: int = 0, tsim: float = 0.0):
        """dumpalloct_serial's per-field files for `field_ids` under `folder`
        (see emma_amd.rtio for the format and the reader)"""
        if not folder.endswith("/"):
            folder += "/"
        ids = np.ascontiguousarray(np.asarray(field_ids, dtype=np.int32))
        self._check(self.lib.emma_rt_dump_grid(self.h, folder.encode(), int(ndump), int(rank), float(tsim),
                                               int(ids.size), ids.ctypes.data_as(_ip)))

    def grid_field(self, field_id: int, rank: int = 0) -> np.ndarray:
        """one grid-output field's values over the dumped cells (float32)"""
        n = ctypes.c_longlong(0)
        self._check(self.lib.emma_rt_grid_field(self.h, int(field_id), int(rank), None, 0, ctypes.byref(n)))
        out = np.empty(n.value, dtype=np.float32)
        self._check(self.lib.emma_rt_grid_field(self.h, int(field_id), int(rank),
                                                out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), out.size,
                                                ctypes.byref(n)))
        return out

    def set_mode(self, mode: int):
        """RADSTEP (per-level transport) or COARSERAD (coarse-level transport)"""
        self._check(self.lib.emma_rt_set_mode(self.h, int(mode)))
        self.mode = mode

    # ---- coupling to the hydro state and the test source
    def set_hydro(self, level: int, H):
        """struct Wtype fields of a level (emma_amd.synth.Hydro)"""
        a = [np.ascontiguousarray(getattr(H, k), dtype=np.float64) for k in ("d", "u", "v", "w", "p", "dX")]
        self._check(self.lib.emma_rt_set_hydro(self.h, level, *[_d(x) for x in a]))

    def get_hydro(self, level: int):
        """(p, dX, E) of a level after the last radsolver / coarse step"""
        n = self.tree.level(level).ncell
        p, dX, E = np.empty(n), np.empty(n), np.empty(n)
        self._check(self.lib.emma_rt_get_hydro(self.h, level, _d(p), _d(dX), _d(E)))
        return p, dX, E

    def set_coupling(self, hydro: bool = True):
        self._check(self.lib.emma_rt_set_coupling(self.h, int(hydro)))

    def set_oct_pos(self, level: int, xyz: np.ndarray):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        self._check(self.lib.emma_rt_set_oct_pos(self.h, level, _d(xyz)))

    def set_stromgren_source(self, star):
        """star: emma_amd.synth.Star, or None to switch the source off"""
        if star is None:
            self._check(self.lib.emma_rt_set_stromgren_source(self.h, 0, *([0.0] * 6)))
            return
        self._check(self.lib.emma_rt_set_stromgren_source(self.h, 1, star.x, star.y, star.z, star.mass,
                                                          star.lifetime, star.srcint))

    def set_stars(self, level: int, S):
        """star particles of a level's leaf cells (emma_amd.synth.StarParticles)"""
        a = S.arrays()
        self._check(self.lib.emma_rt_set_stars(self.h, level, S.n, a[0].ctypes.data_as(_ip), _d(a[1]), _d(a[2]),
                                               a[3].ctypes.data_as(_ip), a[4].ctypes.data_as(_ip)))

    def set_star_sources(self, srcint: float, tlife: float, decrease_after_tlife: bool = False,
                         enable: bool = True):
        """FillRad's star-particle sources (stars_sources)"""
        self._check(self.lib.emma_rt_set_star_sources(self.h, int(enable), float(srcint), float(tlife),
                                                      int(decrease_after_tlife)))

    # ---- the reference's calls
    def fillrad(self, level: int, aexp: float = 1.0):
        """FillRad: the COARSERAD gather of the leaves' E, F into the coarse
        level, then the hydro copy / test source when enabled"""
        self._check(self.lib.emma_rt_fillrad(self.h, level, aexp))

    def clean_new(self, level: int):
        """clean_new_rad: no device work (double-buffered state)."""
        return None

    def radsolver(self, level: int, dtnew: float, aexp: float = 1.0, chemonly: int = 0):
        st = self.lib.emma_rt_radsolver(self.h, level, dtnew, aexp, chemonly)
        self._check(st, self.stats() if st < 0 and st != ERR_HIP else None)

    def coarse_step(self, aexp: float = 1.0, ncoarse: int = 1):
        st = self.lib.emma_rt_coarse_step(self.h, aexp, ncoarse)
        self._check(st, self.stats() if st < 0 and st != ERR_HIP else None)

    def set_level_order(self, order: int, level: int = -1):
        """tile order of the tiled kernel on complete Morton cubes (ORDER_*)"""
        self._check(self.lib.emma_rt_set_level_order(self.h, int(level), int(order)))

    def level_order(self, level: int) -> dict:
        o, m = ctypes.c_int(0), ctypes.c_int(0)
        self._check(self.lib.emma_rt_get_level_order(self.h, int(level), ctypes.byref(o), ctypes.byref(m)))
        return {"order": o.value, "cube_side_log2": m.value}

    def set_schedule(self, sched: int):
        """coarse-step schedule: SCHED_GRAPH (one launch per level step, in a
        captured graph) or SCHED_PERSISTENT (one persistent launch fed by a
        device work queue)"""
        self._check(self.lib.emma_rt_set_schedule(self.h, int(sched)))

    def schedule(self) -> dict:
        s, info = ctypes.c_int(0), SchedInfo()
        self._check(self.lib.emma_rt_get_schedule(self.h, ctypes.byref(s), ctypes.byref(info)))
        d = info.as_dict()
        d["sched"] = SCHED_NAMES.get(s.value, s.value)
        return d

    def stats(self) -> dict:
        s = Stats()
        self.lib.emma_rt_get_stats(self.h, ctypes.byref(s))
        return s.as_dict()

    def level_times(self) -> dict:
        """{level: (device ms, launches)} of the last call's event-timed level steps"""
        nl = self.tree.lmax - self.tree.lmin + 1
        ms, n = np.zeros(nl), np.zeros(nl, dtype=np.int32)
        self._check(self.lib.emma_rt_get_level_times(self.h, nl, _d(ms), n.ctypes.data_as(_ip)))
        return {self.tree.lmin + k: (float(ms[k]), int(n[k])) for k in range(nl)}

    # ---- stencil path
    def create_radstencil(self, stride: int, real_bytes: int = 8):
        self._check(self.lib.emma_rt_create_radstencil(self.h, stride, real_bytes))

    def pinned_stencil(self, stride: int, real_bytes: int = 8) -> "PinnedStencil":
        """page-locked host buffer of `stride` RGRIDs (create_pinned_stencil_rad)"""
        p = ctypes.c_void_p()
        self._check(self.lib.emma_rt_create_pinned_stencil(self.h, stride, real_bytes, ctypes.byref(p)))
        return PinnedStencil(self.lib, p, stride, self.params.ngrp, real_bytes)

    def advance_stencil(self, stencil: np.ndarray, nread: int, dxcur: float, dtnew: float, aexp: float = 1.0,
                        chemonly: int = 0, level: int = None):
        """Run the sweeps, update and chemistry on a host buffer holding
        ``nread`` struct RGRID records of ``level`` (modified in place; the
        level defaults to the one of dxcur = 0.5^level)."""
        assert stencil.flags.c_contiguous
        if level is None:
            level = int(round(-np.log2(dxcur)))
        st = self.lib.emma_rt_advance_stencil(self.h, stencil.ctypes.data_as(ctypes.c_void_p), nread, level, dxcur,
                                              dtnew, aexp, chemonly)
        self._check(st, self.stats() if st < 0 and st != ERR_HIP else None)
