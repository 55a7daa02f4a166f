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
