"""How well-conditioned the reference path is: the chemistry sub-step
controller (chemrad, chem_utils.c:477-716: a sub-step is rejected when a
photon density turns negative, x leaves [0, 1], eint turns negative or
|eint'/eint - 1| > FRAC_VAR) is discontinuous in its inputs.  A cell sitting
on a decision boundary takes one sub-step or two, and its x_HII and N move
at the per-cent level, under any change of rounding.  This bounds what the
mixed-precision path can promise cell by cell (its GPU tests bound the
fraction of such cells instead) and is shown here on the FP64 oracle alone."""
import numpy as np
import pytest

from emma_amd import synth
from oracle import Oracle


@pytest.mark.timeout(300)
def test_substep_controller_discontinuity_c3():
    """C3 (128^3, 3 groups, 64 sources): after one sub-cycle, perturb the state
    by 1e-6 relative and take the second sub-cycle.  Most cells move by about
    1e-6, the sub-step count changes in tens of thousands of cells, and cells
    on a sub-step boundary move by hundreds of times the perturbation (up to
    0.5 % in x_HII for one cell, depending on the draw of the perturbation)."""
    W = synth.config_c3(nsub=2)
    o = Oracle()
    o.setup(W.params, W.tree, W.fields)
    dt = W.params.dt_rad(7)
    o.radsolver(7, dt)
    S0 = o.get(7)
    o.radsolver(7, dt)
    F = o.get(7)
    n_ref = o.nsub[7].copy()
    rng = np.random.default_rng(1)
    P = S0.copy()
    for k in ("e", "fx", "fy", "fz", "eint", "nhplus"):
        a = getattr(P, k)
        a *= 1 + 1e-6 * rng.standard_normal(a.shape)
    o.setup(W.params, W.tree, {7: P})
    o.radsolver(7, dt)
    G = o.get(7)
    x0, x1 = F.nhplus / F.nh, G.nhplus / G.nh
    dx = np.abs(x1 - x0) / x0
    print("max |dx|/x %.3g, cells beyond 1e-4: %d, sub-step count changed in %d cells"
          % (dx.max(), np.count_nonzero(dx > 1e-4), np.count_nonzero(o.nsub[7] != n_ref)))
    assert np.median(dx) < 1e-5
    assert np.count_nonzero(o.nsub[7] != n_ref) > 1000
    assert dx.max() > 1e-4  # over 100x the perturbation
    assert np.count_nonzero(dx > 1e-4) <= 20
