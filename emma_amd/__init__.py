"""emma_amd -- MI355X-native M1 radiative transfer + hydrogen chemistry
(the EMMA ``RadSolver`` hot path) as hand-written HIP kernels for gfx950
behind a C-ABI library (``include/emma_rt.h``)."""

from .params import RunParams, make_params  # noqa: F401
from .synth import Fields, Workload  # noqa: F401
from . import tree  # noqa: F401

__all__ = ["RunParams", "make_params", "Fields", "Workload", "tree"]
