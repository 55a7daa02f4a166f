"""Flat oct-tree description of the levels the radiation path advances.

EMMA keeps its mesh as pointer-linked ``struct OCT`` records
(``prototypes.h:1028-1050``): each oct has 8 cells, a parent cell at level-1,
six face-neighbour cells at level-1 (``nei[6]``) and per-cell ``child``
pointers.  The device path needs index tables instead of pointers, so a level
here is:

* ``pos``     int32 [noct, 3]   oct coordinates on the level's oct grid
* ``parent``  int32 [noct]      parent cell index at level-1
* ``neicell`` int32 [noct, 6]   face-neighbour cell index at level-1
* ``child``   int32 [noct * 8]  child oct index at level+1, -1 for a leaf
* ``nei_oct`` int32 [noct, 6]   derived: the same-level oct on that face, or
  ``-1 - neicell`` when that level-1 cell is a leaf (the coarse cell the
  reference injects into the stencil, ``rad_utils.c:1447-1476``)

Cells are numbered ``oct * 8 + icell`` with ``icell = ix + 2 iy + 4 iz``.  The
coarsest level ``lmin`` is complete and periodic; its parents form a
structural layer at ``lmin - 1`` whose cell ``s`` is the parent of oct ``s``.
"""
from __future__ import annotations

import dataclasses

import numpy as np

# nei_oct entries of domain-boundary faces (WRADTEST TRANS*/REF* flags,
# rad_utils.c:1065-1409): the ghost across the face is the cell's own copy,
# with the normal flux negated when reflective; no refluxing through it
NEI_TRANSMISSIVE = -(2**31)
NEI_REFLECTIVE = -(2**31) + 1

# face f = 2d + s: axis d, s=0 minus side, s=1 plus side
FACE_DIR = np.array([[-1, 0, 0], [1, 0, 0], [0, -1, 0], [0, 1, 0], [0, 0, -1], [0, 0, 1]], dtype=np.int64)
CELL_OFF = np.array([[i & 1, (i >> 1) & 1, i >> 2] for i in range(8)], dtype=np.int64)


def getcellnei(icell: int):
    """Face neighbours of a cell inside its oct (``oct.c:116-185``): returns
    (vnei, vcell) where vnei[f] is 6 for a neighbour in the same oct, else the
    face index, and vcell[f] is the neighbour's cell index."""
    vnei, vcell = [], []
    for f in range(6):
        d, s = f >> 1, f & 1
        b = (icell >> d) & 1
        vcell.append(icell ^ (1 << d))
        vnei.append(6 if b != s else f)
    return vnei, vcell


def _part1by2(v: np.ndarray) -> np.ndarray:
    v = v.astype(np.uint64) & np.uint64(0x1FFFFF)
    v = (v | (v << np.uint64(32))) & np.uint64(0x1F00000000FFFF)
    v = (v | (v << np.uint64(16))) & np.uint64(0x1F0000FF0000FF)
    v = (v | (v << np.uint64(8))) & np.uint64(0x100F00F00F00F00F)
    v = (v | (v << np.uint64(4))) & np.uint64(0x10C30C30C30C30C3)
    v = (v | (v << np.uint64(2))) & np.uint64(0x1249249249249249)
    return v


def morton(pos: np.ndarray) -> np.ndarray:
    """Morton (Z-order) key of integer coordinates [n, 3]."""
    return _part1by2(pos[:, 0]) | (_part1by2(pos[:, 1]) << np.uint64(1)) | (_part1by2(pos[:, 2]) << np.uint64(2))


@dataclasses.dataclass
class Level:
    level: int
    pos: np.ndarray
    parent: np.ndarray
    neicell: np.ndarray
    child: np.ndarray
    nei_oct: np.ndarray = None

    @property
    def noct(self) -> int:
        return int(self.pos.shape[0])

    @property
    def ncell(self) -> int:
        return 8 * self.noct

    def cell_pos(self) -> np.ndarray:
        """Integer cell coordinates [ncell, 3] on the level's cell grid."""
        return (2 * self.pos[:, None, :] + CELL_OFF[None, :, :]).reshape(-1, 3)

    def leaf_mask(self) -> np.ndarray:
        return self.child < 0


class _Lookup:
    """Lookup of oct coordinates on one level (periodic grid of n^3 octs):
    a dense index grid up to 2^24 octs, sorted keys beyond."""

    DENSE_MAX = 1 << 24

    def __init__(self, pos: np.ndarray, n: int):
        self.n = n
        key = (pos[:, 0].astype(np.int64) * n + pos[:, 1]) * n + pos[:, 2]
        if n**3 <= self.DENSE_MAX:
            self.grid = np.full(n**3, -1, dtype=np.int64)
            self.grid[key] = np.arange(key.size)
        else:
            self.grid = None
            self.order = np.argsort(key, kind="stable")
            self.keys = key[self.order]

    def find(self, q: np.ndarray) -> np.ndarray:
        n = self.n
        q = np.mod(q, n)
        key = (q[:, 0].astype(np.int64) * n + q[:, 1]) * n + q[:, 2]
        if self.grid is not None:
            return self.grid[key]
        i = np.searchsorted(self.keys, key)
        i = np.minimum(i, len(self.keys) - 1)
        ok = self.keys[i] == key
        return np.where(ok, self.order[i], -1)


@dataclasses.dataclass
class Tree:
    lmin: int
    lmax: int
    levels: dict

    def level(self, l: int) -> Level:
        return self.levels[l]

    def noct(self):
        return [self.levels[l].noct for l in range(self.lmin, self.lmax + 1)]

    def leaf_cells(self, l: int) -> int:
        return int(np.count_nonzero(self.levels[l].child < 0))

    def flat(self):
        """Concatenated per-level arrays (parent, neicell, child) in level order."""
        ls = [self.levels[l] for l in range(self.lmin, self.lmax + 1)]
        return (
            np.array([L.noct for L in ls], dtype=np.int32),
            np.ascontiguousarray(np.concatenate([L.parent for L in ls]).astype(np.int32)),
            np.ascontiguousarray(np.concatenate([L.neicell.reshape(-1) for L in ls]).astype(np.int32)),
            np.ascontiguousarray(np.concatenate([L.child for L in ls]).astype(np.int32)),
        )

    def validate(self):
        """Structural checks: parent/child symmetry, neighbour existence."""
        for l in range(self.lmin, self.lmax + 1):
            L = self.levels[l]
            assert L.nei_oct.shape == (L.noct, 6)
            if l == self.lmin:
                assert ((L.nei_oct >= 0) | (L.nei_oct <= NEI_REFLECTIVE)).all(), "the coarsest level must be complete"
            if l < self.lmax:
                F = self.levels[l + 1]
                split = np.nonzero(L.child >= 0)[0]
                assert len(split) == F.noct
                assert (F.parent[L.child[split]] == split).all()
            else:
                assert (L.child < 0).all()


def _finish_level(L: Level, coarse_child: np.ndarray | None):
    """nei_oct from neicell and the level-1 child table."""
    if coarse_child is None:
        L.nei_oct = L.neicell.astype(np.int32).copy()
    else:
        ch = coarse_child[L.neicell]
        L.nei_oct = np.where(ch >= 0, ch, -1 - L.neicell).astype(np.int32)


def uniform_level(level: int) -> Level:
    """Complete periodic level: 2**(level-1) octs a side, octs in Morton order."""
    n = 1 << (level - 1)
    g = np.arange(n, dtype=np.int64)
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pos = pos[np.argsort(morton(pos), kind="stable")]
    look = _Lookup(pos, n)
    nei = np.stack([look.find(pos + FACE_DIR[f]) for f in range(6)], 1)
    assert (nei >= 0).all()
    noct = pos.shape[0]
    L = Level(
        level=level,
        pos=pos.astype(np.int32),
        parent=np.arange(noct, dtype=np.int32),
        neicell=nei.astype(np.int32),
        child=np.full(noct * 8, -1, dtype=np.int32),
    )
    return L


def refine(coarse: Level, split_cells: np.ndarray) -> Level:
    """Build level l+1 from the cells of level l listed in ``split_cells``;
    sets ``coarse.child``.  Every face neighbour of a split cell must exist at
    level l (EMMA's 2:1 rule, ``amr.c:746``)."""
    l = coarse.level
    ncs = 1 << l  # cells a side at level l
    cpos = coarse.cell_pos()[split_cells]
    pos = cpos  # child oct coordinates at level l+1 equal the parent cell's
    order = np.argsort(morton(pos), kind="stable")
    pos = pos[order]
    parent = np.asarray(split_cells)[order]
    look = _Lookup(coarse.pos.astype(np.int64), ncs // 2)
    neicell = np.empty((pos.shape[0], 6), dtype=np.int64)
    for f in range(6):
        q = np.mod(pos + FACE_DIR[f], ncs)
        o = look.find(q >> 1)
        if (o < 0).any():
            raise ValueError("refine: a split cell lacks a face neighbour (2:1 rule)")
        neicell[:, f] = o * 8 + (q[:, 0] & 1) + 2 * (q[:, 1] & 1) + 4 * (q[:, 2] & 1)
    child = coarse.child.copy()
    child[parent] = np.arange(pos.shape[0], dtype=np.int32)
    coarse.child = child
    return Level(
        level=l + 1,
        pos=pos.astype(np.int32),
        parent=parent.astype(np.int32),
        neicell=neicell.astype(np.int32),
        child=np.full(pos.shape[0] * 8, -1, dtype=np.int32),
    )


def refinable(L: Level) -> np.ndarray:
    """Cells whose six face neighbours exist on their level."""
    ncs = 1 << L.level
    look = _Lookup(L.pos.astype(np.int64), ncs // 2)
    cp = L.cell_pos()
    ok = np.ones(cp.shape[0], dtype=bool)
    for f in range(6):
        q = np.mod(cp + FACE_DIR[f], ncs)
        ok &= look.find(q >> 1) >= 0
    return ok


def build_tree(levels: list) -> Tree:
    t = Tree(lmin=levels[0].level, lmax=levels[-1].level, levels={L.level: L for L in levels})
    for i, L in enumerate(levels):
        _finish_level(L, None if i == 0 else levels[i - 1].child)
    return t


def uniform_tree(level: int) -> Tree:
    return build_tree([uniform_level(level)])


def amr_tree(lmin: int, counts: list, score_fn, rng: np.random.Generator) -> Tree:
    """Clustered AMR tree: level lmin complete; level lmin+k+1 refines the
    ``counts[k]`` refinable leaves of level lmin+k with the highest
    ``score_fn(level, cell_centres)``."""
    levels = [uniform_level(lmin)]
    for k, want in enumerate(counts):
        L = levels[-1]
        ok = refinable(L)
        centres = (L.cell_pos() + 0.5) / float(1 << L.level)
        score = score_fn(L.level, centres) * (1.0 + 1e-6 * rng.random(L.ncell))
        score = np.where(ok, score, -np.inf)
        want = min(int(want), int(ok.sum()))
        if want <= 0:
            break
        pick = np.argpartition(-score, want - 1)[:want]
        levels.append(refine(L, np.sort(pick)))
    return build_tree(levels)


def with_boundaries(tree: Tree, trans=(0, 1, 2, 3, 4, 5), reflect=()) -> Tree:
    """The tree with non-periodic domain faces, as EMMA's WRADTEST build flags
    TRANSXM..TRANSZP and REFXM/REFYM/REFZM set them up (rad_utils.c:1065-1409):
    an oct whose lower corner lies on the domain face f in `trans`
    (x == 0 for f = 0, x + 2 dx == 1 for f = 1, ...) sees its own cells
    across that face, with the flux component normal to f negated for f in
    `reflect` (the reference defines reflection on the minus faces 0, 2, 4
    only), and refluxes nothing through it (rad_utils.c:1700-1734).  The
    faces' nei_oct entries become NEI_TRANSMISSIVE / NEI_REFLECTIVE;
    ``neicell`` keeps the periodic neighbours (the reference decides from the
    octs' positions)."""
    assert set(reflect) <= set(trans) and set(trans) <= set(range(6))
    levels = {}
    for l, L in tree.levels.items():
        nei = L.nei_oct.copy()
        last = (1 << (l - 1)) - 1  # the last oct coordinate on level l's grid
        for f in trans:
            d, s = f >> 1, f & 1
            on = L.pos[:, d] == (last if s else 0)
            nei[on, f] = NEI_REFLECTIVE if f in reflect else NEI_TRANSMISSIVE
        levels[l] = dataclasses.replace(L, nei_oct=nei)
    return Tree(tree.lmin, tree.lmax, levels)
