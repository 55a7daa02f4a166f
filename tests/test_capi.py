"""The C-ABI library without a GPU: it loads, exports every entry point
include/emma_rt.h declares, computes the reference's struct RGRID layout, and
refuses work cleanly when no device is visible."""
import ctypes
import os
import re

import numpy as np
import pytest

import golden_io as gio
from emma_amd import rgrid, rt
from emma_amd.params import ATOMIC_TABLES, ATOMIC_TIME_BOUNDS, parse_atomic, parse_atomic_time_bounds, make_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "emma_rt.h")).read()
    return sorted(set(re.findall(r"\b(emma_rt_\w+)\s*\(", txt)))


def test_exports_match_header():
    lib = rt.load()
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
    assert sorted(rt.EXPORTS) == syms


def test_layout_matches_reference():
    z = gio.load("layout.npz")
    for G in (1, 3):
        assert np.array_equal(rt.rgrid_layout(G, 8), z[f"layout_g{G}"])
        assert np.array_equal(rgrid.layout(G, 8), z[f"layout_g{G}"])
    for G in range(1, 10):
        for rb in (4, 8):
            assert np.array_equal(rt.rgrid_layout(G, rb), rgrid.layout(G, rb)), (G, rb)
            assert rgrid.empty(1, G, rb).itemsize == rgrid.layout(G, rb)[19]


def test_layout_check_refuses_mismatch():
    lib = rt.load()
    lay = rt.rgrid_layout(3, 8)
    p = lay.ctypes.data_as(ctypes.POINTER(ctypes.c_long))
    assert lib.emma_rt_check_layout(3, 8, p) == rt.OK
    bad = lay.copy()
    bad[19] += 8  # e.g. a host compiled with SUPERNOVAE
    assert lib.emma_rt_check_layout(3, 8, bad.ctypes.data_as(ctypes.POINTER(ctypes.c_long))) == rt.ERR_LAYOUT
    assert lib.emma_rt_check_layout(1, 8, p) == rt.ERR_LAYOUT
    assert lib.emma_rt_rgrid_layout(0, 8, p) == rt.ERR_ARG
    assert lib.emma_rt_rgrid_layout(3, 2, p) == rt.ERR_ARG


def test_strerror_names_every_status():
    for st in (rt.OK, rt.ERR_ARG, rt.ERR_NEG_PHOTON, rt.ERR_EMIN, rt.ERR_NOCONV, rt.ERR_NEIGHBOR, rt.ERR_LAYOUT,
               rt.ERR_NODEVICE, rt.ERR_STATE):
        s = rt.strerror(st)
        assert s and s != "unknown status", st


def test_no_device_is_an_error_not_a_fallback():
    if rt.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(rt.EmmaRTError) as ei:
        rt.RadGPU(make_params())
    assert ei.value.status == rt.ERR_NODEVICE


def test_null_arguments_rejected():
    lib = rt.load()
    assert lib.emma_rt_create(None, None) == rt.ERR_ARG
    assert lib.emma_rt_radsolver(None, 5, 1.0, 1.0, 0) == rt.ERR_STATE
    assert lib.emma_rt_coarse_step(None, 1.0, 1) == rt.ERR_STATE
    assert lib.emma_rt_advance_stencil(None, None, 1, 7, 1.0, 1.0, 1.0, 0) == rt.ERR_ARG


@pytest.mark.parametrize("name", sorted(ATOMIC_TABLES))
def test_parse_atomic_files(name):
    """readAtomic's parse (parameters.c:431-500) of the reference's data files"""
    txt = open(os.path.join(gio.GOLDEN, "atomic", name + ".dat")).read()
    assert parse_atomic(txt) == ATOMIC_TABLES[name]
    assert parse_atomic_time_bounds(txt) == ATOMIC_TIME_BOUNDS.get(name, (0.0,))


def test_time_groups_of_the_nine_group_table():
    """Starburst99_3grpT_3grpS: three time groups of three frequency groups,
    bounds 0 / 3.6 / 36 Myr converted to years and closed by infinity
    (parameters.c:455-468), handed to the library with the parameters"""
    p = make_params("Starburst99_3grpT_3grpS")
    assert p.ngrp == 9 and p.ngrp_time == 3
    assert tuple(p.time_bound) == (0.0, 3.6e6, 36e6, float("inf"))
    c = p.to_c()
    assert c.ngrp_time == 3 and c.time_bound[1] == 3.6e6 and c.time_bound[3] == float("inf")
    assert make_params("Starburst99_1grpT_3grpS").to_c().ngrp_time == 1


def test_readatomic_conversions():
    p = make_params("Starburst99_1grpT_3grpS", clight=0.1)
    assert p.ngrp == 3
    np.testing.assert_allclose(p.hnu[0], 17.9087139468 * 1.602176565e-19, rtol=1e-15)
    np.testing.assert_allclose(p.alphai[1], 0.1 * 6.06571839078e-23 * 299792458.0, rtol=1e-15)
    c = p.to_c()
    assert c.ngrp == 3 and c.factgrp[2] == p.factgrp[2]
