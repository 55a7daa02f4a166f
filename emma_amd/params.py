"""Run parameters for the radiative-transfer path.

Mirrors the parts of EMMA's ``struct RUNPARAMS`` the M1 + chemistry path
reads (``prototypes.h:299-413``): ``struct UNITS`` (``prototypes.h:171-179``),
``clight``/``fudgecool``/``ncvgcool`` (``prototypes.h:366-371``), the atomic
group table filled by ``readAtomic`` (``parameters.c:418-509``) and the
``COSMOPARAM`` entries used by ``chemrad`` (H0, om, ov).
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
from typing import Sequence

import numpy as np

# physical constants of param.h:25-37
LIGHT_SPEED_IN_M_PER_S = 299792458.0
KBOLTZ = 1.3806e-23
PARSEC = 3.085677e16
MYR = 3.1536e13
PROTON_MASS = 1.67262158e-27
ELECTRONVOLT = 1.602176565e-19
# param.h:5,12,20-21
CFL = 0.85
FRAC_VAR = 0.25
PMIN = 1e-10
EMIN = 1e-8
GAMMA = 5.0 / 3.0

MAX_GRP = 9

# Group tables of the reference's atomic data files (src/atomic_data/*.dat):
# columns hnu [eV], alphae [m^2], alphai [m^2], factgrp.
ATOMIC_TABLES = {
    # SB99_equiv_mono_TH.dat
    "SB99_equiv_mono_TH": [(20.64, 2.83e-22, 2.381e-22, 1.0)],
    # Starburst99_1grpT_1grpS.dat
    "Starburst99_1grpT_1grpS": [(23.4209817443, 2.35188329313e-22, 1.82867991612e-22, 1.0)],
    # Starburst99_1grpT_3grpS.dat
    "Starburst99_1grpT_3grpS": [
        (17.9087139468, 3.42023557809e-22, 3.19218060719e-22, 0.769923932738),
        (32.5864317739, 6.62279605918e-23, 6.06571839078e-23, 0.230075891068),
        (55.6621660507, 1.16142701989e-23, 1.15856478863e-23, 1.76194762129e-07),
    ],
    # Starburst99_3grpT_3grpS.dat: three time groups of three frequency groups
    "Starburst99_3grpT_3grpS": [
        (18.0960837205, 3.33247782214e-22, 3.1029498972e-22, 0.711623692435),
        (32.7734105767, 6.54066737569e-23, 5.97338219337e-23, 0.288376214572),
        (58.0179019951, 1.03517562247e-23, 1.02588318683e-23, 9.29939343807e-08),
        (17.5908242307, 3.56894560696e-22, 3.34329166724e-22, 0.923901199349),
        (31.0925705306, 7.27896047138e-23, 6.80343540065e-23, 0.0760984046781),
        (54.6277040944, 1.21686713315e-23, 1.21682856377e-23, 3.95972432576e-07),
        (15.1064199164, 4.9465484405e-22, 4.85759793145e-22, 0.999990281992),
        (25.8970145725, 1.08028137438e-22, 1.07501523182e-22, 9.71800798434e-06),
        (77.9786617129, 5.72753801219e-24, 4.92806933216e-24, 2.87618856869e-51),
    ],
}


# TIME_BOUND(MYR) lines of the tables with more than one time group (the
# others have one group from 0 Myr)
ATOMIC_TIME_BOUNDS = {
    "Starburst99_3grpT_3grpS": (0.0, 3.6, 36.0),
}


def parse_atomic_time_bounds(text: str) -> tuple:
    """The TIME_BOUND(MYR) entries of an EMMA atomic data file, in Myr
    (``parameters.c:455-468`` reads them and closes them with infinity)."""
    tok = text.split()
    k = tok.index("NGRP_TIME")
    ntime = int(tok[k + 1])
    return tuple(float(t) for t in tok[k + 3: k + 3 + ntime])


def parse_atomic(text: str) -> list[tuple[float, float, float, float]]:
    """Parse an EMMA atomic data file the way ``readAtomic`` does
    (``parameters.c:431-500``): a ``NGRP_SPACE n`` header with its bounds,
    a ``NGRP_TIME m`` header with its bounds, four column names, then n*m rows
    of ``hnu alphae alphai factgrp``."""
    tok = text.split()
    k = 0
    if tok[k] != "NGRP_SPACE":
        raise ValueError("atomic file: expected NGRP_SPACE")
    nspace = int(tok[k + 1])
    k += 2 + 1 + nspace + 1  # SPACE_BOUND(eV), bounds, inf
    if tok[k] != "NGRP_TIME":
        raise ValueError("atomic file: expected NGRP_TIME")
    ntime = int(tok[k + 1])
    k += 2 + 1 + ntime + 1
    k += 4  # column header
    rows = []
    for _ in range(nspace * ntime):
        rows.append(tuple(float(t) for t in tok[k : k + 4]))
        k += 4
    return rows


class _CParams(ctypes.Structure):
    """C layout shared with ``emma_rt_params`` (include/emma_rt.h) and
    ``orc_params`` (oracle/emma_oracle.h)."""

    _fields_ = [
        ("ngrp", ctypes.c_int),
        ("ncvgcool", ctypes.c_int),
        ("fudgecool", ctypes.c_double),
        ("clight", ctypes.c_double),
        ("unit_l", ctypes.c_double),
        ("unit_v", ctypes.c_double),
        ("unit_t", ctypes.c_double),
        ("unit_n", ctypes.c_double),
        ("unit_mass", ctypes.c_double),
        ("unit_d", ctypes.c_double),
        ("unit_N", ctypes.c_double),
        ("hnu", ctypes.c_double * MAX_GRP),
        ("alphae", ctypes.c_double * MAX_GRP),
        ("alphai", ctypes.c_double * MAX_GRP),
        ("factgrp", ctypes.c_double * MAX_GRP),
        ("H0", ctypes.c_double),
        ("om", ctypes.c_double),
        ("ov", ctypes.c_double),
        ("lcoarse", ctypes.c_int),
        ("ngrp_time", ctypes.c_int),
        ("time_bound", ctypes.c_double * (MAX_GRP + 1)),
    ]


@dataclasses.dataclass
class RunParams:
    """Radiation/chemistry run parameters (code units as in EMMA)."""

    unit_l: float
    unit_v: float
    unit_t: float
    unit_n: float
    unit_mass: float
    unit_d: float
    unit_N: float
    clight: float = 0.1
    fudgecool: float = 1.0
    ncvgcool: int = 0
    hnu: Sequence[float] = ()
    alphae: Sequence[float] = ()
    alphai: Sequence[float] = ()
    factgrp: Sequence[float] = ()
    H0: float = 0.0
    om: float = 1.0
    ov: float = 0.0
    lcoarse: int = 0
    # time groups of the atomic table: ngrp = NGRP_SPACE x ngrp_time, bounds
    # in years (readAtomic, parameters.c:455-468), the last one infinite
    time_bound: Sequence[float] = (0.0, math.inf)

    @property
    def ngrp(self) -> int:
        return len(self.hnu)

    @property
    def ngrp_time(self) -> int:
        return len(self.time_bound) - 1

    @classmethod
    def from_table(cls, rows, clightorg: float, time_bound_myr=(0.0,), **units) -> "RunParams":
        """Apply ``readAtomic``'s conversions (``parameters.c:455-497``):
        hnu eV -> J; cross-sections m^2 -> m^3/s at the reduced light speed;
        time bounds Myr -> yr, closed by infinity."""
        hnu = [r[0] * ELECTRONVOLT for r in rows]
        ae = [clightorg * (r[1] * LIGHT_SPEED_IN_M_PER_S) for r in rows]
        ai = [clightorg * (r[2] * LIGHT_SPEED_IN_M_PER_S) for r in rows]
        fg = [r[3] for r in rows]
        if len(rows) > MAX_GRP:
            raise ValueError(f"at most {MAX_GRP} photon groups")
        tb = tuple(t * 1e6 for t in time_bound_myr) + (math.inf,)
        if len(rows) % (len(tb) - 1):
            raise ValueError("the group count is not a multiple of the time groups")
        return cls(clight=clightorg, hnu=hnu, alphae=ae, alphai=ai, factgrp=fg, time_bound=tb, **units)

    def to_c(self) -> _CParams:
        c = _CParams()
        c.ngrp = self.ngrp
        c.ncvgcool = int(self.ncvgcool)
        c.fudgecool = self.fudgecool
        c.clight = self.clight
        for k in ("unit_l", "unit_v", "unit_t", "unit_n", "unit_mass", "unit_d", "unit_N", "H0", "om", "ov"):
            setattr(c, k, float(getattr(self, k)))
        for k in ("hnu", "alphae", "alphai", "factgrp"):
            arr = getattr(c, k)
            for i, v in enumerate(getattr(self, k)):
                arr[i] = v
        c.lcoarse = int(self.lcoarse)
        c.ngrp_time = self.ngrp_time
        for i, v in enumerate(self.time_bound):
            c.time_bound[i] = v
        return c

    def units7(self) -> np.ndarray:
        return np.array(
            [self.unit_l, self.unit_v, self.unit_t, self.unit_n, self.unit_mass, self.unit_d, self.unit_N]
        )

    def cloc(self, aexp: float = 1.0) -> float:
        """Reduced light speed in code units (``rad_utils.c:1784``)."""
        return aexp * self.clight * LIGHT_SPEED_IN_M_PER_S / self.unit_v

    def dt_rad(self, level: int, aexp: float = 1.0) -> float:
        """Radiative CFL time step of a level (``advanceamr.c:226``)."""
        return CFL / (aexp * self.clight * LIGHT_SPEED_IN_M_PER_S / self.unit_v) / 3.0 / float(1 << level)

    def eint_code(self, nh_code, x, temp, aexp: float = 1.0):
        """Internal energy in code units for a temperature, inverting ``E2T``
        (``chem_utils.c:21-32``)."""
        pstar = self.unit_n * self.unit_d * self.unit_v**2
        nh_phys = np.asarray(nh_code) / aexp**3 * self.unit_N
        return 1.5 * nh_phys * KBOLTZ * (1.0 + np.asarray(x)) * np.asarray(temp) / pstar * aexp**5

    def src_code(self, luminosity, dx: float, aexp: float = 1.0):
        """Emissivity of one cell holding a source of ``luminosity`` photons/s,
        as ``stromgren_source`` sets ``rfield.src`` (``src_utils.c:381``)."""
        return np.asarray(luminosity) / (dx * self.unit_l) ** 3 * self.unit_t / self.unit_N * aexp**2


def static_units(n_h: float = 1e3, box_kpc: float = 15.0) -> dict:
    """Units of a static (a=1) box: unit_v = c, unit_t = unit_l/c, unit_N = n_H
    [m^-3] so that code n_H is the overdensity (``emma.c:1180-1196,1281-1284``)."""
    unit_l = box_kpc * 1e3 * PARSEC
    unit_v = LIGHT_SPEED_IN_M_PER_S
    unit_d = n_h * PROTON_MASS
    return dict(
        unit_l=unit_l,
        unit_v=unit_v,
        unit_t=unit_l / unit_v,
        unit_n=1.0,
        unit_mass=unit_d * unit_l**3,
        unit_d=unit_d,
        unit_N=n_h,
    )


def make_params(table: str = "SB99_equiv_mono_TH", clight: float = 0.1, n_h: float = 1e3, **kw) -> RunParams:
    return RunParams.from_table(ATOMIC_TABLES[table], clight, ATOMIC_TIME_BOUNDS.get(table, (0.0,)),
                                **static_units(n_h), **kw)


def e_min_code(params: RunParams, aexp: float = 1.0) -> float:
    """Smallest photon density chemrad accepts (EMIN, ``chem_utils.c:382``)."""
    return EMIN / params.unit_N * aexp**3


def t_rec_myr(n_h: float, alpha_b: float = 2.59e-19) -> float:
    return 1.0 / (n_h * alpha_b) / MYR


def stromgren_radius_m(ndot: float, n_h: float, alpha_b: float = 2.59e-19) -> float:
    return (3.0 * ndot / (4.0 * math.pi * n_h**2 * alpha_b)) ** (1.0 / 3.0)
