"""coldplate-solver MI355X backend: the SIMPLE hot path (U/V momentum Jacobi
with SOU deferred correction and Brinkman / 2.5-D sinks, pressure-correction
solve, velocity / pressure correction) and the 3-D conjugate-heat-transfer
solve, as hand-written HIP kernels for gfx950 behind the C ABI declared in
include/coldplate_hip.h (libcoldplate_hip.so, built in-tree).

The shared library is loaded on first use; every entry point raises if it is
missing or no GPU is present -- there is no CPU fallback.
"""
from . import _types, configs, exportfiles  # noqa: F401
from ._types import default_config, SimpleConfig, StepStats, Patch, field_shape  # noqa: F401
from . import _capi  # noqa: F401  (loads nothing until first use)
from .simple import SIMPLE, Batch, pressure_solve_csr, pressure_solve_csr_noalloc, momentum_jacobi_csr, pressure_drop  # noqa: F401,E501
from .thermal import ThermalSolver3D, thermal_objective, surrogate_adjoint  # noqa: F401
from . import coupled  # noqa: F401
from .coupled import CoupledObjective, coupled_f0, f0_smooth  # noqa: F401

__all__ = ["_types", "configs", "exportfiles", "default_config", "SimpleConfig", "StepStats", "Patch", "field_shape",
           "SIMPLE", "Batch", "pressure_solve_csr", "pressure_solve_csr_noalloc", "momentum_jacobi_csr", "pressure_drop", "ThermalSolver3D", "thermal_objective", "surrogate_adjoint",
           "coupled", "CoupledObjective", "coupled_f0", "f0_smooth"]
