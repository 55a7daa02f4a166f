"""emma_rt_advance_stencil_noalloc (GPU): the drop-in stencil step staged in
the library's static per-GPU device area instead of the buffer
emma_rt_create_radstencil allocates.  Called through the C-ABI, it returns
the same values as emma_rt_advance_stencil on the same struct RGRID batch, on
two calls in a row (the second with fewer records, so the area holds the
first call's contents), and refuses batches larger than the area."""
import ctypes

import numpy as np
import pytest

from emma_amd import rgrid, rt, synth

pytestmark = pytest.mark.gpu

AREA = 32 << 20  # EMMA_RT_NOALLOC_STENCIL_BYTES


def _batch(W, lvl, n, G, o0=0):
    L, F = W.tree.level(lvl), W.fields[lvl]
    return rgrid.gather_complete(L, F, np.arange(o0, o0 + n), G)


def _fields(x, prefix=""):
    """every leaf field of a structured array (padding bytes are not data:
    numpy copies structured records field by field)"""
    if x.dtype.names is None:
        yield prefix, x
        return
    for k in x.dtype.names:
        yield from _fields(x[k], prefix + "." + k)


def _same_records(a, b):
    for (k, x), (_, y) in zip(_fields(a), _fields(b)):
        assert np.array_equal(x, y), k


def _noalloc(g, st, n, lvl, dt, rbytes=8):
    return g.lib.emma_rt_advance_stencil_noalloc(g.h, st.ctypes.data_as(ctypes.c_void_p), n, lvl, 0.5**lvl, dt,
                                                 1.0, 0, rbytes)


@pytest.mark.parametrize("G,table", [(1, "SB99_equiv_mono_TH"), (3, "Starburst99_1grpT_3grpS")])
def test_noalloc_matches_allocating_entry(gpu_lib, G, table):
    lvl = 5
    W = synth.uniform_workload(lvl, table, seed=60 + G, nsrc=4)
    dt = W.params.dt_rad(lvl)
    rsize = int(rgrid.layout(G, 8)[19])
    nmax = AREA // rsize
    n1 = min(nmax, W.tree.level(lvl).noct)
    n2 = n1 // 3
    g = rt.RadGPU(W.params, rt.FP64)
    g.create_radstencil(n1, 8)
    h = rt.RadGPU(W.params, rt.FP64)  # no stencil buffer: the no-alloc entry needs none
    for n, o0 in ((n1, 0), (n2, 7)):
        a = _batch(W, lvl, n, G, o0)
        b = a.copy()
        g.advance_stencil(a, n, 0.5**lvl, dt)
        assert _noalloc(h, b, n, lvl, dt) == rt.OK
        assert h.stats()["status"] == 0 and h.stats()["cell_updates"] == g.stats()["cell_updates"]
        _same_records(a, b)
    # a batch beyond the staging area is refused before anything runs
    big = rgrid.empty(nmax + 1, G, 8)
    assert _noalloc(h, big, nmax + 1, lvl, dt) == rt.ERR_ARG
    assert _noalloc(h, big, 0, lvl, dt) == rt.OK
    assert _noalloc(h, big, 1, lvl, dt, rbytes=2) == rt.ERR_ARG
