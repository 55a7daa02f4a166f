"""EMMA's serial grid output format (the files ``emma_rt_dump_grid`` writes).

``dumpalloct_serial`` (``io.c:1233-1363``) writes one file per output field,
``<folder>grid_<name>/<name>.<ndump:05d>.p<rank:05d>``: an int ``n``, four
bytes of time (the first ``sizeof(float)`` bytes of the writer's REAL
``tsim``, io.c:1282), six floats ``xmin, xmax, ymin, ymax, zmin, zmax``, then
``n`` floats, one per dumped cell (leaf cells, and every cell of level lmax),
levels from the coarsest up, octs in level order, cells 0..7.  Field ids and
names are those of ``assign_grid_field`` (``io.c:32-200``) and
``readOutputParam_grid`` (``parameters.c:132-177``).
"""
from __future__ import annotations

import os

import numpy as np

FIELD_NAMES = [
    "x", "y", "z", "l", "cpu", "gdata_d", "density", "gdata.p", "res", "f0", "f1", "f2", "marked",
    "field_d", "field_u", "field_v", "field_w", "field_p", "rfield_e0", "rfield_fx0", "rfield_fy0",
    "rfield_fz0", "rfield_e1", "rfield_fx1", "rfield_fy1", "rfield_fz1", "rfield_snfb", "rfield_e2",
    "rfield_fx2", "rfield_fy2", "rfield_fz2", "rfield_src", "xion", "field_dX", "field_dXHE", "field_dXXHE",
    "field_xHE", "field_xxHE", "rfield_temp", "t_first_xion", "t_last_xion", "rho_first_xion",
]
FIELD_ID = {n: i for i, n in enumerate(FIELD_NAMES)}
HEADER = np.dtype([("n", "<i4"), ("tsim_bytes", "V4"), ("bounds", "<f4", (6,))])


def path(folder: str, field: int, ndump: int, rank: int = 0) -> str:
    name = FIELD_NAMES[field]
    return os.path.join(folder, f"grid_{name}", f"{name}.{ndump:05d}.p{rank:05d}")


def read(p: str) -> dict:
    """{n, tsim_low (the header's four time bytes as the low half of a
    double), bounds, data (float32[n])}"""
    raw = np.fromfile(p, dtype=np.uint8)
    h = np.frombuffer(raw[:HEADER.itemsize].tobytes(), dtype=HEADER)[0]
    n = int(h["n"])
    data = np.frombuffer(raw[HEADER.itemsize:].tobytes(), dtype="<f4")
    if data.size != n:
        raise ValueError(f"{p}: header says {n} cells, file holds {data.size}")
    return {"n": n, "tsim_bytes": bytes(h["tsim_bytes"]), "bounds": h["bounds"].copy(), "data": data}
