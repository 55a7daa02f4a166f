"""Write the golden vectors of the radiation path from the reference's own CPU
functions (rad_utils.c, chem_utils.c, oct.c compiled by oracle/Makefile into
oracle/_ref/).  Run in the build container, where /root/reference exists:

    make -C oracle ref && python tests/golden/make_golden.py

Every fixture holds inputs and the reference's outputs (npz, float64 unless
stated):

layout.npz        struct RGRID/Rtype/CELL* sizes and offsets (G=1,3; REAL=double)
getcellnei.npz    oct.c getcellnei table
eddington.npz     Eddington() on random and edge-case states, all (i, j)
rates.npz         cucompute_alpha_a/_b, cucompute_beta, cuCompCooling on a T,x,n_H grid
e2t.npz           E2T() on random states
chem_g{1,3}.npz   chemrad() on single cells over a state grid incl. the rejection regime
stencil_g{1,3}.npz rad_sweepX/Y/Z + updatefieldrad + chemrad on a gathered RGRID batch
level_g{1,3}.npz  RadSolver on a complete periodic level, after 1 and 5 sub-cycles
amr_g{1,3}.npz    two coarse steps of the RADSTEP recursion on a 4-level tree
                  (coarse injection, 1/8 refluxing, restriction, reionisation times)
c2_sample.npz     one RadSolver step of the 64^3 level (seeded workload): input and
                  output checksums plus sampled output cells
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from emma_amd import synth, rgrid  # noqa: E402
from emma_amd.params import make_params  # noqa: E402
from emma_amd.synth import Fields  # noqa: E402
from oracle import Reference  # noqa: E402

TABLES = {1: "SB99_equiv_mono_TH", 3: "Starburst99_1grpT_3grpS"}


def save(name, **kw):
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **kw)
    print(f"{name}: {os.path.getsize(path) / 1024:.0f} KiB")


def fields_dict(prefix, F: Fields):
    return {f"{prefix}{k}": getattr(F, k) for k in Fields.NAMES}


def g_layout():
    out = {}
    for G in (1, 3):
        out[f"layout_g{G}"] = Reference(G).layout()
    save("layout.npz", **out)


def g_getcellnei():
    r = Reference(1)
    vn = np.zeros((8, 6), np.int32)
    vc = np.zeros((8, 6), np.int32)
    for c in range(8):
        a, b = r.getcellnei(c)
        vn[c], vc[c] = a, b
    save("getcellnei.npz", vnei=vn, vcell=vc)


def g_eddington():
    rng = np.random.default_rng(101)
    n = 4000
    c = 0.1
    e = 10.0 ** rng.uniform(-8, 2, n)
    f = rng.uniform(0, 1, n)
    f[:50] = 0.0
    f[50:100] = 1.0
    u = rng.normal(size=(3, n))
    u /= np.linalg.norm(u, axis=0)
    u[:, 100:130] = np.array([1.0, 0, 0])[:, None]
    F = u * f * c * e
    e[130:150] = 0.0
    e[150:160] = -1e-3
    i = rng.integers(0, 3, n).astype(np.int32)
    j = rng.integers(0, 3, n).astype(np.int32)
    out = Reference(1).eddington(F[0], F[1], F[2], e, c, i, j)
    save("eddington.npz", fx=F[0], fy=F[1], fz=F[2], e=e, c=c, i=i, j=j, out=out)


def g_rates():
    T = 10.0 ** np.linspace(0, 9, 400)
    xs = np.array([0.0, 1e-3, 0.5, 1 - 1e-6, 1.0])
    nhs = np.array([1e-3, 1.0, 1e3, 1e6])
    TT, XX, NN = [a.reshape(-1) for a in np.meshgrid(T, xs, nhs, indexing="ij")]
    out = Reference(1).rates(TT, XX, NN)
    save("rates.npz", temp=TT, x=XX, nH=NN, **out)


def g_e2t():
    rng = np.random.default_rng(7)
    p = make_params()
    r = Reference(1)
    r.set_params(p)
    n = 2000
    nh = np.exp(rng.normal(0, 2, n))
    x = rng.uniform(0, 1, n)
    eint = p.eint_code(nh, x, 10.0 ** rng.uniform(1, 7, n))
    aexp = np.where(np.arange(n) < n // 2, 1.0, 0.25)
    temp = np.concatenate([r.e2t(x[: n // 2] * nh[: n // 2], eint[: n // 2], nh[: n // 2], 1.0),
                           r.e2t(x[n // 2:] * nh[n // 2:], eint[n // 2:], nh[n // 2:], 0.25)])
    save("e2t.npz", nhplus=x * nh, eint=eint, nh=nh, aexp=aexp, temp=temp)


def chem_states(G, p, rng):
    """single-cell states over a grid of n_H, x, T, N, src and dt"""
    nhs = [0.1, 1.0, 30.0]
    xs = [1e-6, 1e-3, 0.3, 0.99, 1 - 1e-9]
    Ts = [1e2, 3e3, 1e4, 1e5, 1e6]
    Ns = [1e-9, 1e-4, 1e-1, 3.0]
    srcs = [0.0, 1e3]
    grid = [(a, b, c, d, s) for a in nhs for b in xs for c in Ts for d in Ns for s in srcs]
    n = len(grid) + (-len(grid)) % 8
    grid += grid[: n - len(grid)]
    nh = np.array([g[0] for g in grid])
    x = np.array([g[1] for g in grid])
    T = np.array([g[2] for g in grid])
    N = np.array([g[3] for g in grid])
    S = np.array([g[4] for g in grid])
    F = Fields.zeros(G, n)
    F.nh[:] = nh
    F.nhplus[:] = x * nh
    F.eint[:] = p.eint_code(nh, x, T)
    u = rng.normal(size=(G, 3, n))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    for g in range(G):
        F.e[g] = N * (1.0 + 0.3 * g)
        f = u[g] * rng.uniform(0, 1, n) * p.cloc() * F.e[g]
        F.fx[g], F.fy[g], F.fz[g] = f
        F.src[g] = S
    F.temp[:] = T
    return F


def fields_to_stencil(F: Fields, G):
    """cells of F as the centre cells of ncell/8 gathered octs (New.rfieldnew)"""
    n = F.ncell
    st = rgrid.empty(n // 8, G)
    R = st["New"]["rfieldnew"]  # [nread, 8] view (a reshape would copy)
    for k in Fields.NAMES:
        a = getattr(F, k)
        R[k] = a.T.reshape(n // 8, 8, G) if a.ndim == 2 else a.reshape(n // 8, 8)
    return st


def stencil_to_fields(st, G):
    R = st["New"]["rfieldnew"]
    n = R.shape[0] * 8
    F = Fields.zeros(G, n)
    for k in Fields.NAMES:
        a = R[k]
        setattr(F, k, np.ascontiguousarray(a.reshape(n, G).T) if a.ndim == 3 else np.ascontiguousarray(a.reshape(n)))
    return F


def g_chem(G):
    rng = np.random.default_rng(300 + G)
    p = make_params(TABLES[G])
    r = Reference(G)
    r.set_params(p)
    r.params = p
    F = chem_states(G, p, rng)
    out = {}
    for k, dtf in enumerate((1.0, 100.0)):
        dt = p.dt_rad(5) * dtf
        st = fields_to_stencil(F, G)
        r.stencil_chem(st, 0.5**5, dt, 1.0, 0)
        out.update(fields_dict(f"out{k}_", stencil_to_fields(st, G)))
        out[f"dt{k}"] = dt
    save(f"chem_g{G}.npz", **fields_dict("in_", F), **out)


def g_stencil(G):
    rng = np.random.default_rng(500 + G)
    p = make_params(TABLES[G])
    r = Reference(G)
    r.set_params(p)
    r.params = p
    nread = 48 if G == 1 else 24
    st = rgrid.empty(nread, G)
    c = p.cloc()
    for slot in range(7):
        for cell in range(8):
            nh = np.exp(rng.normal(0, 1, nread))
            W = synth.random_fields(p, nh, rng)
            R = st["oct"]["rfield"][:, slot, cell]
            for k in Fields.NAMES:
                a = getattr(W, k)
                R[k] = a.T if a.ndim == 2 else a
            st["oct"]["split"][:, slot, cell] = (rng.random(nread) < (0.1 if slot == 6 else 0.2)).astype(np.int8)
    new = st["New"]["rfieldnew"]
    centre = st["oct"]["rfield"][:, 6, :]
    for k in ("e", "fx", "fy", "fz", "src", "nhplus", "eint", "nh", "temp"):
        new[k] = centre[k]
    new["e"] *= 1.0 + 0.01 * rng.random(new["e"].shape)  # rfieldnew holds refluxes on top of rfield
    inp = st.copy()
    dx = 0.5**6
    dt = p.dt_rad(6)
    r.stencil_sweeps(st, dx, dt, c)
    r.stencil_update(st, dx, dt, c)
    r.stencil_chem(st, dx, dt, 1.0, 0)
    save(f"stencil_g{G}.npz", inp=inp.view(np.uint8), out=st.view(np.uint8), dx=dx, dt=dt, c=c, nread=nread)


def g_level(G):
    level = 4 if G == 1 else 3
    W = synth.uniform_workload(level, TABLES[G], seed=600 + G, nsrc=3)
    r = Reference(G)
    r.setup(W.params, W.tree, W.fields)
    dt = W.params.dt_rad(level)
    out = {}
    for n in range(1, 6):
        r.clean_new(level)
        r.radsolver(level, dt)
        if n in (1, 5):
            out.update(fields_dict(f"step{n}_", r.get(level)))
    L = W.tree.level(level)
    save(f"level_g{G}.npz", level=level, nei_oct=L.nei_oct, child=L.child, dt=dt, **fields_dict("in_", W.fields[level]),
         **out)


def g_amr(G):
    W = synth.amr_workload(3, [60, 100, 80], TABLES[G], seed=700 + G, sigma=1.5, nsrc=2, grid=16)
    tr = W.tree
    r = Reference(G)
    r.setup(W.params, tr, W.fields)
    out = {}
    for step, tphy in enumerate((1.5, 2.5)):
        r.lib.refdrv_set_tphy(tphy)
        r.coarse_step()
    for l in range(tr.lmin, tr.lmax + 1):
        out.update(fields_dict(f"out{l}_", r.get(l)))
        tf, tl = r.txion(l)
        out[f"tfirst{l}"] = tf
        out[f"tlast{l}"] = tl
        L = tr.level(l)
        out[f"nei_oct{l}"] = L.nei_oct
        out[f"child{l}"] = L.child
        out[f"parent{l}"] = L.parent
        out[f"neicell{l}"] = L.neicell
        out.update(fields_dict(f"in{l}_", W.fields[l]))
    save(f"amr_g{G}.npz", lmin=tr.lmin, lmax=tr.lmax, tphy=np.array([1.5, 2.5]), **out)


def checksums(F: Fields):
    return np.array([[np.sum(getattr(F, k)), np.sum(getattr(F, k) ** 2)] for k in Fields.NAMES])


def g_c2():
    W = synth.config_c2()
    L = W.tree.level(6)
    r = Reference(1)
    r.setup(W.params, W.tree, W.fields)
    r.clean_new(6)
    r.radsolver(6, W.params.dt_rad(6))
    F = r.get(6)
    rng = np.random.default_rng(9)
    idx = np.sort(rng.choice(L.ncell, 4000, replace=False))
    save("c2_sample.npz", seed=2, in_checksums=checksums(W.fields[6]), out_checksums=checksums(F), idx=idx,
         **fields_dict("out_", F.take(idx)), **fields_dict("in_", W.fields[6].take(idx)))


if __name__ == "__main__":
    g_layout()
    g_getcellnei()
    g_eddington()
    g_rates()
    g_e2t()
    for G in (1, 3):
        g_chem(G)
        g_stencil(G)
        g_level(G)
        g_amr(G)
    g_c2()
