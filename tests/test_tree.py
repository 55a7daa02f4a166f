"""The flat oct-tree tables: neighbour symmetry, Morton order, parent/child
consistency and EMMA's 2:1 rule on generated AMR trees."""
import numpy as np
import pytest

from emma_amd import synth
from emma_amd import tree as T

OPP = [1, 0, 3, 2, 5, 4]


@pytest.mark.parametrize("level", [2, 3, 5])
def test_uniform_level_neighbours(level):
    L = T.uniform_tree(level).level(level)
    n = 1 << (level - 1)
    assert L.noct == n**3
    nei = L.nei_oct
    for f in range(6):
        assert np.array_equal(nei[nei[:, f], OPP[f]], np.arange(L.noct))
        d = (L.pos[nei[:, f]] - L.pos) % n
        exp = T.FACE_DIR[f] % n
        assert (d == exp).all()
    k = T.morton(L.pos.astype(np.int64))
    assert (np.diff(k.astype(np.int64)) > 0).all()


def test_cell_neighbour_rule():
    """getcellnei: crossing face 2d+s flips bit d of the cell index"""
    for c in range(8):
        vn, vc = T.getcellnei(c)
        for f in range(6):
            d, s = f >> 1, f & 1
            assert vc[f] == c ^ (1 << d)
            inside = ((c >> d) & 1) != s
            assert vn[f] == (6 if inside else f)


def test_amr_tree_consistency():
    W = synth.amr_workload(4, [300, 200, 100], "SB99_equiv_mono_TH", seed=3, sigma=1.5, grid=16)
    tr = W.tree
    tr.validate()
    assert tr.noct() == [512, 300, 200, 100]
    for l in range(tr.lmin + 1, tr.lmax + 1):
        L, C = tr.level(l), tr.level(l - 1)
        # every face neighbour of a fine oct exists at level-1 (2:1 rule)
        assert (L.neicell >= 0).all() and (L.neicell < C.ncell).all()
        # geometry: neighbour cell = parent cell shifted by one along the face
        ncs = 1 << (l - 1)
        cp = C.cell_pos()
        for f in range(6):
            exp = (cp[L.parent] + T.FACE_DIR[f]) % ncs
            assert (cp[L.neicell[:, f]] == exp).all()
        # nei_oct encodes split neighbours by oct index, leaves by -1 - cell
        split = C.child[L.neicell] >= 0
        assert np.array_equal(L.nei_oct[split], C.child[L.neicell][split])
        assert np.array_equal(L.nei_oct[~split], -1 - L.neicell[~split])
        assert (~split).any()  # coarse-fine faces exist


def test_c4_level_counts():
    """the batched-AMR workload: 262144/80000/25000/6000/1200 octs"""
    import emma_amd.synth as S

    rng = np.random.default_rng(4)
    g = S.gaussian_field(64, rng)
    tr = T.amr_tree(7, [80000, 25000, 6000, 1200], S.lognormal_sampler(g, 1.5), rng)
    assert tr.noct() == [262144, 80000, 25000, 6000, 1200]
    tr.validate()


def test_domain_boundary_codes():
    """with_boundaries marks exactly the faces on the domain faces asked for
    (by the reference's position test, rad_utils.c:1106-1213), reflective
    where asked, on every level, and leaves the periodic neighbour table
    (neicell) and the other entries alone; the codes are the C-ABI's."""
    import os
    import re

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    hdr = open(os.path.join(root, "include", "emma_rt.h")).read()
    defs = dict(re.findall(r"#define (EMMA_RT_NEI_\w+) \((-?\d+)(?: - 1)?\)", hdr))
    assert T.NEI_TRANSMISSIVE == -2147483647 - 1 and "EMMA_RT_NEI_TRANSMISSIVE" in defs
    assert T.NEI_REFLECTIVE == int(defs["EMMA_RT_NEI_REFLECTIVE"])
    W = synth.amr_workload(3, [60, 100], "SB99_equiv_mono_TH", seed=3, sigma=1.5, grid=16)
    tb = T.with_boundaries(W.tree, trans=(0, 1, 3, 4), reflect=(0, 4))
    for l in range(W.tree.lmin, W.tree.lmax + 1):
        L, B = W.tree.level(l), tb.level(l)
        last = (1 << (l - 1)) - 1
        for f in range(6):
            d, s = f >> 1, f & 1
            on = L.pos[:, d] == (last if s else 0)
            code = B.nei_oct[:, f]
            if f in (0, 1, 3, 4):
                want = T.NEI_REFLECTIVE if f in (0, 4) else T.NEI_TRANSMISSIVE
                assert (code[on] == want).all(), (l, f)
                assert (code[~on] == L.nei_oct[~on, f]).all(), (l, f)
            else:
                assert (code == L.nei_oct[:, f]).all(), (l, f)
        assert np.array_equal(B.neicell, L.neicell)
    corners = synth.oct_corners(W.tree.level(W.tree.lmin))
    assert np.count_nonzero(tb.level(W.tree.lmin).nei_oct[:, 0] == T.NEI_REFLECTIVE) == np.count_nonzero(corners[:, 0] == 0.0)
    tb.validate()
