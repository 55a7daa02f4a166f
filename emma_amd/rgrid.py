"""numpy view of EMMA's gathered radiation stencil, ``struct RGRID``
(``prototypes.h:1165-1168``): 7 ``OCTLIGHT_R`` (6 face neighbours and the
centre oct in slot 6, ``rad_utils.c:1522-1528``) and the ``OCTFLUX_R`` ``New``
holding each centre cell's ``rfieldnew`` and its 6*NGRP*5 face fluxes.

The dtype reproduces the C layout for a given NGRP and sizeof(REAL) (without
SUPERNOVAE/HELIUM members) so that a buffer built here can be handed to the
library's stencil entry point exactly as EMMA's host would hand its
``struct RGRID *stencil``.
"""
from __future__ import annotations

import numpy as np


def rtype_dtype(G: int, rb: int = 8) -> np.dtype:
    f = np.float64 if rb == 8 else np.float32
    return np.dtype([("e", f, (G,)), ("fx", f, (G,)), ("fy", f, (G,)), ("fz", f, (G,)), ("src", f, (G,)),
                     ("nhplus", f), ("eint", f), ("nh", f), ("temp", f)])


def rgrid_dtype(G: int, rb: int = 8) -> np.dtype:
    f = np.float64 if rb == 8 else np.float32
    R = rtype_dtype(G, rb)
    celllight = np.dtype({"names": ["rfield", "split"], "formats": [R, np.int8], "offsets": [0, R.itemsize],
                          "itemsize": R.itemsize + rb})
    cellflux = np.dtype([("deltaR", R), ("rfield", R), ("rfieldnew", R), ("rflux", f, (30 * G,))])
    return np.dtype([("oct", celllight, (7, 8)), ("New", cellflux, (8,))])


def layout(G: int, rb: int = 8) -> np.ndarray:
    """The 22-number layout table (sizes/offsets) of the C-ABI's layout check."""
    R = rtype_dtype(G, rb)
    dt = rgrid_dtype(G, rb)
    cl = dt["oct"].base
    cf = dt["New"].base
    v = [R.itemsize] + [R.fields[k][1] for k in ("e", "fx", "fy", "fz", "src", "nhplus", "eint", "nh", "temp")]
    v += [cl.itemsize, cl.fields["split"][1]]
    v += [cf.itemsize] + [cf.fields[k][1] for k in ("deltaR", "rfield", "rfieldnew", "rflux")]
    v += [8 * cl.itemsize, 8 * cf.itemsize, dt.itemsize, dt.fields["oct"][1], dt.fields["New"][1]]
    return np.array(v, dtype=np.int64)


def empty(nread: int, G: int, rb: int = 8) -> np.ndarray:
    return np.zeros(nread, dtype=rgrid_dtype(G, rb))


def flux_view(st: np.ndarray, G: int) -> np.ndarray:
    """rflux as [nread, 8 cells, 6 faces, G, 5 variables]."""
    return st["New"]["rflux"].reshape(st.shape[0], 8, 6, G, 5)


def gather_complete(L, F, octs: np.ndarray, G: int, rb: int = 8) -> np.ndarray:
    """The host gather (gatherstencilrad, rad_utils.c:1482-1534) for octs of
    a complete periodic level: slot 6 holds the oct's own cells, slots 0..5
    the cells of the same-level neighbour oct across each face
    (recursive_neighbor_gather_oct_rad, rad_utils.c:1413-1477, with every
    neighbour present), ``split`` marks split cells, and ``New`` starts as
    the centre cells' state (rfieldnew = rfield after clean_new_rad) with zero
    fluxes.  L is the level (``nei_oct``, ``child``), F its fields."""
    octs = np.asarray(octs, dtype=np.int64)
    n = octs.shape[0]
    nei = L.nei_oct.reshape(-1, 6)
    if (nei[octs] < 0).any():
        raise ValueError("gather_complete: an oct has a coarse neighbour")
    slot_oct = np.empty((n, 7), dtype=np.int64)
    slot_oct[:, :6] = nei[octs]
    slot_oct[:, 6] = octs
    cells = slot_oct[:, :, None] * 8 + np.arange(8)[None, None, :]  # [n, 7, 8]
    st = empty(n, G, rb)
    R = st["oct"]["rfield"]
    for k in ("e", "fx", "fy", "fz", "src"):
        R[k] = np.moveaxis(getattr(F, k)[:, cells], 0, -1)
    for k in ("nhplus", "eint", "nh", "temp"):
        R[k] = getattr(F, k)[cells]
    st["oct"]["split"] = (L.child[cells] >= 0).astype(np.int8)
    centre = R[:, 6, :]
    st["New"]["rfield"] = centre
    st["New"]["rfieldnew"] = centre
    return st
