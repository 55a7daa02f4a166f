"""Star particles as radiation sources (stars_sources, src_utils.c:209-335,
called by FillRad through putsource, 446-520): the oracle against the
reference's own STARS builds, bit for bit, and against the closed form.

Particles of several kinds sit in random leaf cells of every level, several
to a cell: radiating stars (isStar 6..8), particles that do not radiate
(isStar < 6) and AGN (isStar 100, whose branch is empty without AGN); ages
before the run's time (radiating), after it (not yet born: t < 0) and older
than the lifetime (radiating on; with DECREASE_EMMISIVITY_AFTER_TLIFE fading
as t^-4).  Nine groups = three time groups: each particle feeds the three
groups of its radiative_state."""
import numpy as np
import pytest

from emma_amd import synth
from emma_amd.synth import Fields
from oracle import Oracle, Reference, ref_available

TABLES = {3: "Starburst99_1grpT_3grpS", 9: "Starburst99_3grpT_3grpS"}
SRCINT = 5e48 / (2e2 * 1.989e30)
TPHY, TLIFE = 1.0, 0.4


def star_setup(solver, W, stars, decrease):
    solver.setup(W.params, W.tree, W.fields, TPHY)
    solver.set_star_sources(SRCINT, TLIFE, decrease)
    for l, S in stars.items():
        solver.set_stars(l, S)


def workload(G, seed):
    W = synth.amr_workload(3, [40, 80, 50], TABLES[G], seed=seed, sigma=1.5, grid=16)
    rng = np.random.default_rng(seed + 100)
    stars = {l: synth.random_stars(W.tree.level(l), 150, rng, W.params.ngrp_time, TPHY, TLIFE)
             for l in range(W.tree.lmin, W.tree.lmax + 1)}
    return W, stars


@pytest.mark.parametrize("G,build", [(3, 1), (9, 1), (9, 2)], ids=["g3", "g9", "g9-decrease"])
def test_star_sources_vs_reference(G, build):
    if not ref_available(G, stars=build):
        pytest.skip("reference STARS build (oracle/_ref) not present")
    W, stars = workload(G, 31)
    o, r = Oracle(), Reference(G, stars=build)
    for s in (o, r):
        star_setup(s, W, stars, build == 2)
    lit = 0
    for l in range(W.tree.lmin, W.tree.lmax + 1):
        o.fillrad(l)
        r.fillrad(l)
        for which in (0, 1):
            A, B = o.get(l, which), r.get(l, which)
            for k in Fields.NAMES:
                assert np.array_equal(getattr(A, k), getattr(B, k)), (l, which, k)
        lit += int(np.count_nonzero(o.get(l).src))
    assert lit > 0


def closed_form(W, S, level, decrease):
    """stars_sources summed per cell by numpy (the reference's formula)"""
    p = W.params
    L = W.tree.level(level)
    ns = p.ngrp // p.ngrp_time
    src = np.zeros((p.ngrp, L.ncell))
    for k in range(S.n):
        if S.is_star[k] < 6 or S.is_star[k] == 100 or L.child[S.cell[k]] >= 0:
            continue
        age = TPHY - S.age[k]
        t = age / TLIFE
        v = (S.mass[k] * p.unit_d) * SRCINT / (0.5**level) ** 3 * p.unit_t / p.unit_N * 1.0
        if decrease and age >= TLIFE:
            v *= (1.0 if t < 1 else t**-4)
        v *= 1.0 if t > 0 else 0.0
        g0 = S.radiative_state[k] * ns
        src[g0:g0 + ns, S.cell[k]] += v
    return src


@pytest.mark.parametrize("decrease", [False, True])
def test_star_sources_closed_form(decrease):
    W, stars = workload(9, 32)
    o = Oracle()
    star_setup(o, W, stars, decrease)
    for l in range(W.tree.lmin, W.tree.lmax + 1):
        o.fillrad(l)
        ref = closed_form(W, stars[l], l, decrease)
        got = o.get(l).src
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=0)
        # only groups of the particles' time groups carry light
        assert np.all(got[:, W.tree.level(l).child >= 0] == 0)
