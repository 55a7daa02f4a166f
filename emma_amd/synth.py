"""Seeded synthetic radiation/chemistry states for the benchmark workloads.

Every workload is synthetic (no simulation snapshots travel): photon density
N, flux F, ionised fraction x, temperature T and hydrogen density n_H are drawn
per cell from the distributions the workloads name, on a static (a = 1) box
in the units of :func:`emma_amd.params.static_units`.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .params import RunParams, make_params, e_min_code
from . import tree as T


@dataclasses.dataclass
class Fields:
    """One level's radiation state, struct Rtype as SoA (float64).
    Group fields are [G, ncell]; the others are [ncell]."""

    e: np.ndarray
    fx: np.ndarray
    fy: np.ndarray
    fz: np.ndarray
    src: np.ndarray
    nhplus: np.ndarray
    eint: np.ndarray
    nh: np.ndarray
    temp: np.ndarray

    NAMES = ("e", "fx", "fy", "fz", "src", "nhplus", "eint", "nh", "temp")

    @property
    def ngrp(self) -> int:
        return self.e.shape[0]

    @property
    def ncell(self) -> int:
        return self.e.shape[1]

    def copy(self) -> "Fields":
        return Fields(*[np.array(getattr(self, k), dtype=np.float64, copy=True, order="C") for k in self.NAMES])

    @classmethod
    def zeros(cls, ngrp: int, ncell: int) -> "Fields":
        g = lambda: np.zeros((ngrp, ncell))
        s = lambda: np.zeros(ncell)
        return cls(g(), g(), g(), g(), g(), s(), s(), s(), s())

    def xion(self) -> np.ndarray:
        return self.nhplus / self.nh

    def take(self, idx) -> "Fields":
        return Fields(*[getattr(self, k)[..., idx] for k in self.NAMES])


@dataclasses.dataclass
class Hydro:
    """One level's hydro state as the radiation step reads and writes it
    (struct Wtype, float64 SoA [ncell]): density d (= n_H in code units with
    YHE = 0), velocity u, v, w, pressure p, ionised density dX, total energy E."""

    d: np.ndarray
    u: np.ndarray
    v: np.ndarray
    w: np.ndarray
    p: np.ndarray
    dX: np.ndarray
    E: np.ndarray

    NAMES = ("d", "u", "v", "w", "p", "dX", "E")

    def copy(self) -> "Hydro":
        return Hydro(*[np.array(getattr(self, k), dtype=np.float64, copy=True) for k in self.NAMES])


GAMMA = 5.0 / 3.0


def hydro_like(F: "Fields", rng: np.random.Generator, vmax: float = 0.1) -> Hydro:
    """A hydro state the radiation state F is consistent with: d = n_H,
    p = (GAMMA-1) eint, dX = x n_H, random velocities |u_i| < vmax."""
    n = F.ncell
    u, v, w = (rng.uniform(-vmax, vmax, n) for _ in range(3))
    p = (GAMMA - 1.0) * F.eint
    return Hydro(d=F.nh.copy(), u=u, v=v, w=w, p=p, dX=F.nhplus.copy(), E=np.zeros(n))


@dataclasses.dataclass
class Star:
    """The WRADTEST point source (stromgren_source, src_utils.c:339-412):
    position in box units, mass [solar masses], lifetime [yr] and srcint
    [photons/s/kg]; its time groups are the atomic table's (RunParams)."""

    x: float = 0.5
    y: float = 0.5
    z: float = 0.5
    mass: float = 2e2
    lifetime: float = 1e10
    srcint: float = 5e48 / (2e2 * 1.989e30)


@dataclasses.dataclass
class StarParticles:
    """Star particles of one level's leaf cells (struct PART fields the
    radiation sources read, prototypes.h:722-759), sorted by cell and, within
    a cell, in the order of the cell's particle list."""

    cell: np.ndarray             # int32 level cell index
    mass: np.ndarray             # code units
    age: np.ndarray              # formation time, the unit of tphy
    radiative_state: np.ndarray  # int32 time group
    is_star: np.ndarray          # int32 isStar (>= 6 radiates, 100 = AGN)

    @property
    def n(self) -> int:
        return int(self.cell.shape[0])

    def arrays(self):
        return (np.ascontiguousarray(self.cell, dtype=np.int32), np.ascontiguousarray(self.mass, dtype=np.float64),
                np.ascontiguousarray(self.age, dtype=np.float64),
                np.ascontiguousarray(self.radiative_state, dtype=np.int32),
                np.ascontiguousarray(self.is_star, dtype=np.int32))


def random_stars(L: T.Level, n: int, rng: np.random.Generator, ngrp_time: int = 1, tphy: float = 1.0,
                 tlife: float = 1.0) -> StarParticles:
    """n particles in random leaf cells of level L (several per cell happen),
    sorted by cell with a stable order inside a cell: radiating stars
    (isStar 6..8), non-radiating particles (isStar 0..5) and AGN (100);
    ages before and after tphy and beyond tphy - tlife."""
    leaves = np.nonzero(L.child < 0)[0]
    cell = np.sort(rng.choice(leaves, size=n, replace=True)).astype(np.int32)
    kind = rng.random(n)
    is_star = np.where(kind < 0.7, rng.integers(6, 9, n), np.where(kind < 0.9, rng.integers(0, 6, n), 100))
    age = tphy - rng.uniform(-0.5, 3.0, n) * tlife
    return StarParticles(cell=cell, mass=10.0 ** rng.uniform(-9, -6, n), age=age,
                         radiative_state=rng.integers(0, ngrp_time, n).astype(np.int32),
                         is_star=is_star.astype(np.int32))


def oct_corners(L: T.Level) -> np.ndarray:
    """Lower corners of a level's octs in box units (struct OCT x, y, z)."""
    return np.ascontiguousarray(L.pos.astype(np.float64) * 0.5 ** (L.level - 1))


def random_fields(params: RunParams, nh: np.ndarray, rng: np.random.Generator, aexp: float = 1.0) -> Fields:
    """x ~ U[1e-6, 1], T log-U[1e2, 1e5] K, N log-U[1e-10, 1] (code units),
    F = |F| u with u uniform on the sphere and |F| ~ U[0, c N]; no sources."""
    G = params.ngrp
    n = nh.shape[0]
    c = params.cloc(aexp)
    x = rng.uniform(1e-6, 1.0, n)
    temp = 10.0 ** rng.uniform(2.0, 5.0, n)
    e = 10.0 ** rng.uniform(-10.0, 0.0, (G, n))
    e = np.maximum(e, 2.0 * e_min_code(params, aexp))
    u = rng.normal(size=(G, 3, n))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    fmag = rng.uniform(0.0, 1.0, (G, n)) * c * e
    f = u * fmag[:, None, :]
    eint = params.eint_code(nh, x, temp, aexp)
    F = Fields(
        e=e,
        fx=np.ascontiguousarray(f[:, 0]),
        fy=np.ascontiguousarray(f[:, 1]),
        fz=np.ascontiguousarray(f[:, 2]),
        src=np.zeros((G, n)),
        nhplus=x * nh,
        eint=eint,
        nh=nh.astype(np.float64).copy(),
        temp=temp,
    )
    e2t(F, params, aexp)
    return F


def e2t(F: Fields, params: RunParams, aexp: float = 1.0):
    """Temperature from eint, n_H and x (``chem_utils.c:21-32``), in place."""
    from .params import KBOLTZ

    pstar = params.unit_n * params.unit_d * params.unit_v**2
    nH = F.nh / aexp**3 * params.unit_N
    eint = F.eint / aexp**5 * pstar
    F.temp = eint / (1.5 * nH * KBOLTZ * (1.0 + F.nhplus / F.nh))


def gaussian_field(n: int, rng: np.random.Generator, smooth_cells: float = 2.0) -> np.ndarray:
    """Periodic unit-variance Gaussian random field on an n^3 grid
    (white noise smoothed by a Gaussian of ``smooth_cells``)."""
    w = rng.normal(size=(n, n, n))
    k = np.fft.fftfreq(n) * n
    kx, ky, kz = np.meshgrid(k, k, k[: n // 2 + 1], indexing="ij")
    filt = np.exp(-0.5 * (kx**2 + ky**2 + kz**2) * (2 * np.pi * smooth_cells / n) ** 2)
    g = np.fft.irfftn(np.fft.rfftn(w) * filt, s=(n, n, n), axes=(0, 1, 2))
    g -= g.mean()
    g /= g.std()
    return g


def lognormal_sampler(g: np.ndarray, sigma: float):
    """Density field rho = exp(sigma g - sigma^2/2) sampled nearest-grid at
    fractional positions in [0, 1)^3."""
    n = g.shape[0]

    def f(level, centres):
        idx = np.minimum((centres * n).astype(np.int64), n - 1)
        return np.exp(sigma * g[idx[:, 0], idx[:, 1], idx[:, 2]] - 0.5 * sigma**2)

    return f


@dataclasses.dataclass
class Workload:
    name: str
    params: RunParams
    tree: T.Tree
    fields: dict  # level -> Fields
    nsub: int  # radiative sub-cycles (coarse steps) per run
    description: str = ""

    def leaf_cell_updates_per_coarse_step(self) -> int:
        """Leaf cells x level steps per coarse step (2^(l-lmin) steps at level l)."""
        return sum(self.tree.leaf_cells(l) << (l - self.tree.lmin) for l in range(self.tree.lmin, self.tree.lmax + 1))


def place_sources(params: RunParams, F: Fields, L: T.Level, nsrc: int, lum: np.ndarray, rng) -> np.ndarray:
    """Put ``nsrc`` point sources (photons/s) in distinct leaf cells."""
    leaves = np.nonzero(L.child < 0)[0]
    cells = rng.choice(leaves, size=nsrc, replace=False)
    dx = 0.5**L.level
    F.src[:, cells] = params.src_code(lum, dx)[None, :]
    return cells


def uniform_workload(level: int, table: str, seed: int, nsrc: int = 0, lum_range=(1e48, 1e51),
                     sigma: float = 1.0, nsub: int = 1, clight: float = 0.1, name: str = "") -> Workload:
    """One complete periodic level with iid lognormal n_H (config C2/C3 shape)."""
    rng = np.random.default_rng(seed)
    params = make_params(table, clight=clight)
    tr = T.uniform_tree(level)
    L = tr.level(level)
    nh = np.exp(rng.normal(0.0, sigma, L.ncell))
    F = random_fields(params, nh, rng)
    if nsrc:
        lum = 10.0 ** rng.uniform(np.log10(lum_range[0]), np.log10(lum_range[1]), nsrc)
        place_sources(params, F, L, nsrc, lum, rng)
    return Workload(name or f"uniform{level}", params, tr, {level: F}, nsub)


def amr_workload(lmin: int, counts: list, table: str, seed: int, sigma: float = 1.5, nsrc: int = 0,
                 lum_range=(1e48, 1e51), nsub: int = 1, clight: float = 0.1, grid: int = 64,
                 name: str = "") -> Workload:
    """Clustered AMR tree on a lognormal density (config C4 shape)."""
    rng = np.random.default_rng(seed)
    params = make_params(table, clight=clight)
    g = gaussian_field(grid, rng)
    rho = lognormal_sampler(g, sigma)
    tr = T.amr_tree(lmin, counts, rho, rng)
    fields = {}
    for l in range(tr.lmin, tr.lmax + 1):
        L = tr.level(l)
        centres = (L.cell_pos() + 0.5) / float(1 << l)
        nh = rho(l, centres)
        fields[l] = random_fields(params, nh, rng)
    if nsrc:
        lum = 10.0 ** rng.uniform(np.log10(lum_range[0]), np.log10(lum_range[1]), nsrc)
        place_sources(params, fields[tr.lmax], tr.level(tr.lmax), min(nsrc, tr.leaf_cells(tr.lmax)), lum[: min(nsrc, tr.leaf_cells(tr.lmax))], rng)
    return Workload(name or f"amr{lmin}-{tr.lmax}", params, tr, fields, nsub)


# The benchmark configurations (sizes as named by the workloads)
def config_c2(seed: int = 2) -> Workload:
    """64^3 periodic level (level 6, 262144 cells), 1 group, no sources."""
    return uniform_workload(6, "SB99_equiv_mono_TH", seed, nsub=1, name="C2-64^3-G1-fp64")


def config_c3(seed: int = 3, nsub: int = 32) -> Workload:
    """128^3 periodic level (level 7, 2097152 cells), 3 groups, 64 sources,
    32 radiative sub-cycles."""
    return uniform_workload(7, "Starburst99_1grpT_3grpS", seed, nsrc=64, nsub=nsub, name="C3-128^3-G3-mixed")


def powerlaw_luminosities(rng, n: int, lo: float, hi: float) -> np.ndarray:
    """n luminosities with dN/dL ~ L^-2 on [lo, hi] (inverse transform)."""
    u = rng.random(n)
    return 1.0 / (1.0 / lo - u * (1.0 / lo - 1.0 / hi))


def config_c5(seed: int = 5, nsub: int = 100) -> Workload:
    """Reionisation-like box: 256^3 coarse cells (level 8, 16.8 M cells) and two
    refined levels with 300000 / 75000 octs (3 M cells) on a lognormal sigma=2
    density, 4096 point sources with dN/dL ~ L^-2 on [1e49, 1e53] s^-1 in the
    densest leaves of the coarse level, 3 groups, c~ = c/100."""
    rng = np.random.default_rng(seed)
    params = make_params("Starburst99_1grpT_3grpS", clight=0.01)
    g = gaussian_field(128, rng)
    rho = lognormal_sampler(g, 2.0)
    tr = T.amr_tree(8, [300000, 75000], rho, rng)
    fields = {}
    for l in range(tr.lmin, tr.lmax + 1):
        L = tr.level(l)
        centres = (L.cell_pos() + 0.5) / float(1 << l)
        fields[l] = random_fields(params, rho(l, centres), rng)
    L = tr.level(tr.lmin)
    leaves = np.nonzero(L.child < 0)[0]
    dens = fields[tr.lmin].nh[leaves]
    cells = leaves[np.argsort(-dens, kind="stable")[:4096]]
    lum = powerlaw_luminosities(rng, cells.size, 1e49, 1e53)
    fields[tr.lmin].src[:, cells] = params.src_code(lum, 0.5**tr.lmin)[None, :]
    return Workload("C5-256^3+2-G3", params, tr, fields, nsub)


def config_c4(seed: int = 4, nsub: int = 40) -> Workload:
    """Levels 7-11 with 262144/80000/25000/6000/1200 octs, lognormal sigma=1.5."""
    return amr_workload(7, [80000, 25000, 6000, 1200], "Starburst99_1grpT_3grpS", seed, sigma=1.5, nsub=nsub,
                        name="C4-amr7-11-G3")
