"""Worker of the full-size parity test (not a test module): one oracle run
of a benchmark workload, optionally from a perturbed state, in its own
process, its x_HII, T and N written as .npy files."""
import time

import numpy as np

PERTURBED = ("e", "fx", "fy", "fz", "src", "eint", "nhplus")


def run(args):
    """args = (workload, seed, steps, eps, perturbation seed, output prefix)"""
    name, seed, steps, eps, pseed, out = args
    from emma_amd import synth
    from oracle import Oracle

    t0 = time.time()
    W = {"c3": synth.config_c3, "c4": synth.config_c4, "c5": synth.config_c5}[name](seed=seed, nsub=1)
    if eps > 0:
        rng = np.random.default_rng(pseed)
        for F in W.fields.values():
            for k in PERTURBED:
                a = getattr(F, k)
                a *= 1 + eps * rng.uniform(-1.0, 1.0, a.shape)
    o = Oracle()
    o.setup(W.params, W.tree, W.fields)
    for _ in range(steps):
        o.coarse_step()
    for l in range(W.tree.lmin, W.tree.lmax + 1):
        F = o.get(l)
        np.save(f"{out}_{l}_x.npy", F.nhplus / F.nh)
        np.save(f"{out}_{l}_temp.npy", F.temp)
        np.save(f"{out}_{l}_e.npy", F.e)
    return time.time() - t0
