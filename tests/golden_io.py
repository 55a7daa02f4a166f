"""Loaders that turn the golden fixtures back into trees and fields."""
import os

import numpy as np

from emma_amd import tree as T
from emma_amd.params import make_params
from emma_amd.synth import Fields

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TABLES = {1: "SB99_equiv_mono_TH", 3: "Starburst99_1grpT_3grpS"}


def load(name):
    return np.load(os.path.join(GOLDEN, name))


def fields(z, prefix) -> Fields:
    return Fields(*[np.ascontiguousarray(z[prefix + k]) for k in Fields.NAMES])


def params(G):
    return make_params(TABLES[G])


def level_tree(z) -> T.Tree:
    level = int(z["level"])
    nei = z["nei_oct"].astype(np.int32)
    L = T.Level(level=level, pos=np.zeros((nei.shape[0], 3), np.int32), parent=np.arange(nei.shape[0], dtype=np.int32),
                neicell=nei.copy(), child=z["child"].astype(np.int32))
    L.nei_oct = nei
    return T.Tree(lmin=level, lmax=level, levels={level: L})


def amr_tree(z) -> T.Tree:
    lmin, lmax = int(z["lmin"]), int(z["lmax"])
    levels = {}
    for l in range(lmin, lmax + 1):
        nei = z[f"nei_oct{l}"].astype(np.int32)
        L = T.Level(level=l, pos=np.zeros((nei.shape[0], 3), np.int32), parent=z[f"parent{l}"].astype(np.int32),
                    neicell=z[f"neicell{l}"].astype(np.int32), child=z[f"child{l}"].astype(np.int32))
        L.nei_oct = nei
        levels[l] = L
    return T.Tree(lmin=lmin, lmax=lmax, levels=levels)


def stencil_as_level(st, G):
    """A gathered RGRID batch as a level the oracle can advance: stencil i
    becomes octs 7i..7i+6 (its 6 neighbour slots and the centre in 7i+6);
    split bytes become child >= 0.  Returns (tree, fields, centre cell index)."""
    nread = st.shape[0]
    noct = 7 * nread
    nei = np.zeros((noct, 6), np.int32)
    for i in range(nread):
        for s in range(7):
            nei[7 * i + s] = 7 * i + s  # neighbour slots point at themselves (never compared)
        nei[7 * i + 6] = 7 * i + np.arange(6)
    child = np.where(st["oct"]["split"].reshape(-1) != 0, 0, -1).astype(np.int32)
    L = T.Level(level=6, pos=np.zeros((noct, 3), np.int32), parent=np.arange(noct, dtype=np.int32), neicell=nei,
                child=child)
    L.nei_oct = nei
    tr = T.Tree(lmin=6, lmax=6, levels={6: L})
    R = st["oct"]["rfield"].reshape(noct * 8)
    n = noct * 8
    F = Fields.zeros(G, n)
    for k in Fields.NAMES:
        a = R[k]
        setattr(F, k, np.ascontiguousarray(a.reshape(n, G).T) if a.ndim == 2 else np.ascontiguousarray(a))
    centre = (7 * np.arange(nread)[:, None] + 6) * 8 + np.arange(8)[None, :]
    return tr, F, centre.reshape(-1)
