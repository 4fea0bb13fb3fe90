"""Solver settings and data containers (``src/fem/solvers/data.py:23-218``).

Only the types used on the RVE / material-dynamics path are mirrored, with the
reference's validation rules.  ``LinearSolverSettings`` gains the PCG controls
that replace the sparse direct solve (``src/fem/solvers/linear.py:15-22``).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass(frozen=True)
class DirichletDofSet:
    dofs: np.ndarray
    values: np.ndarray


@dataclass(frozen=True)
class ArmijoSettings:
    """``src/fem/solvers/data.py:32-48``."""

    enabled: bool = False
    sufficient_decrease: float = 1.0e-4
    reduction_factor: float = 0.5
    max_backtracks: int = 20

    def __post_init__(self) -> None:
        if not 0.0 < self.sufficient_decrease < 1.0:
            raise ValueError("Armijo sufficient_decrease must lie in (0, 1).")
        if not 0.0 < self.reduction_factor < 1.0:
            raise ValueError("Armijo reduction_factor must lie in (0, 1).")
        if self.max_backtracks < 0:
            raise ValueError("Armijo max_backtracks must be nonnegative.")


@dataclass(frozen=True)
class AdaptiveTimeStepSettings:
    """``src/fem/solvers/data.py:51-74``."""

    enabled: bool = False
    minimum_factor: float = 1.0 / 16.0
    maximum_factor: float = 2.0
    cutback_factor: float = 0.5
    growth_factor: float = 2.0
    easy_iteration_limit: int = 4
    easy_step_count: int = 3


@dataclass(frozen=True)
class NonlinearNewtonSettings:
    """``src/fem/solvers/data.py:86-95``."""

    max_iterations: int
    residual_absolute_tolerance: float
    residual_relative_tolerance: float
    increment_absolute_tolerance: float
    increment_relative_tolerance: float
    armijo: ArmijoSettings = field(default_factory=ArmijoSettings)


@dataclass(frozen=True)
class LinearSolverSettings:
    """Block-Jacobi PCG controls of the RVE Newton step.

    ``relative_tolerance`` is on ||r_k|| / ||b||; the defaults drive the linear
    residual well below the Newton tolerances of the reference configs.
    """

    backend: str = "pcg"
    method: str = "spectral"  # "spectral" (FFT reference medium) or "block_jacobi"
    max_iterations: int = 20000
    relative_tolerance: float = 1.0e-12
    absolute_tolerance: float = 0.0
    check_interval: int = 8
    reference_bulk: float = 0.0  # <= 0: geometric mean over the phases
    reference_shear: float = 0.0
    # Inexact Newton: the k-th linear solve stops at
    # rtol_k = max(relative_tolerance, 0.5 tol_newton / ||r_k||, min(forcing_max, ||r_k|| / ||r_0||));
    # the Newton residual and increment tests are unchanged.
    adaptive_forcing: bool = False
    forcing_max: float = 0.1


@dataclass(frozen=True)
class TimeIntegrationSettings:
    """``src/fem/solvers/data.py:104-114``."""

    scheme: str
    time_step: float
    num_steps: int
    beta: float = 0.25
    gamma: float = 0.5
    cfl_safety_factor: float = 0.5
    adaptive: AdaptiveTimeStepSettings = field(default_factory=AdaptiveTimeStepSettings)


@dataclass(frozen=True)
class SolverProgressEvent:
    """``src/fem/solvers/progress.py:16-32``."""

    stage: str
    scheme: str
    accepted_step: int
    nominal_total_steps: int
    physical_time: float
    final_time: float
    newton_iterations: int
    armijo_backtracks: int
    elapsed_seconds: float
    residual_norm: float | None = None
    residual_tolerance: float | None = None
