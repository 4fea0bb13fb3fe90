"""YAML configurations of the reference, read into this package's settings types.

``load_yaml_config`` / ``write_yaml_config`` are ``src/fem/io/config.py:15-30``
(PyYAML ``safe_load`` / ``safe_dump``).  The builders turn the config blocks
that the reference's benchmark runners read into the dataclasses of
:mod:`concrete_impact_ai_amd.solvers` and :mod:`concrete_impact_ai_amd.materials`,
with the same keys and defaults:

* ``material_update_settings``  ``src/concrete_impact/benchmarks/rve.py:227-236``
                                (and the identical helpers of ``plasticity.py:401-410``,
                                ``dynamic_viscoplastic.py:173-182``)
* ``newton_settings``           ``src/concrete_impact/benchmarks/rve.py:204-222``,
                                ``dynamic_viscoplastic.py:207-224`` (armijo block
                                optional, its defaults as there)
* ``time_integration_settings`` ``src/concrete_impact/benchmarks/dynamic_viscoplastic.py:185-204``
"""

from __future__ import annotations

from pathlib import Path
from typing import Any

import yaml

from concrete_impact_ai_amd.materials import MaterialUpdateSettings, build_material
from concrete_impact_ai_amd.solvers import (AdaptiveTimeStepSettings, ArmijoSettings,
                                            NonlinearNewtonSettings, TimeIntegrationSettings)

__all__ = ["load_yaml_config", "write_yaml_config", "material_update_settings",
           "newton_settings", "time_integration_settings", "build_material"]


def load_yaml_config(path: str | Path) -> dict[str, Any]:
    """Load a YAML configuration file (``src/fem/io/config.py:15-20``)."""
    with Path(path).open("r", encoding="utf-8") as stream:
        return yaml.safe_load(stream)


def write_yaml_config(data: dict[str, Any], path: str | Path) -> Path:
    """Write a YAML configuration file (``src/fem/io/config.py:23-30``)."""
    out = Path(path)
    out.parent.mkdir(parents=True, exist_ok=True)
    with out.open("w", encoding="utf-8") as stream:
        yaml.safe_dump(data, stream, sort_keys=False)
    return out


def material_update_settings(block: dict[str, Any]) -> MaterialUpdateSettings:
    return MaterialUpdateSettings(
        max_iterations=int(block["max_iterations"]),
        yield_relative_tolerance=float(block["yield_relative_tolerance"]),
        residual_absolute_tolerance=float(block["residual_absolute_tolerance"]),
        residual_relative_tolerance=float(block["residual_relative_tolerance"]),
    )


def newton_settings(block: dict[str, Any]) -> NonlinearNewtonSettings:
    armijo = block.get("armijo", {})
    return NonlinearNewtonSettings(
        max_iterations=int(block["max_iterations"]),
        residual_absolute_tolerance=float(block["residual_absolute_tolerance"]),
        residual_relative_tolerance=float(block["residual_relative_tolerance"]),
        increment_absolute_tolerance=float(block["increment_absolute_tolerance"]),
        increment_relative_tolerance=float(block["increment_relative_tolerance"]),
        armijo=ArmijoSettings(
            enabled=bool(armijo.get("enabled", False)),
            sufficient_decrease=float(armijo.get("sufficient_decrease", 1.0e-4)),
            reduction_factor=float(armijo.get("reduction_factor", 0.5)),
            max_backtracks=int(armijo.get("max_backtracks", 20)),
        ),
    )


def time_integration_settings(block: dict[str, Any]) -> TimeIntegrationSettings:
    adaptive = block.get("adaptive_time_step", {})
    return TimeIntegrationSettings(
        scheme=str(block["scheme"]),
        time_step=float(block["time_step"]),
        num_steps=int(block["num_steps"]),
        beta=float(block["beta"]),
        gamma=float(block["gamma"]),
        cfl_safety_factor=float(block["cfl_safety_factor"]),
        adaptive=AdaptiveTimeStepSettings(
            enabled=bool(adaptive.get("enabled", False)),
            minimum_factor=float(adaptive.get("minimum_factor", 1.0 / 16.0)),
            maximum_factor=float(adaptive.get("maximum_factor", 2.0)),
            cutback_factor=float(adaptive.get("cutback_factor", 0.5)),
            growth_factor=float(adaptive.get("growth_factor", 2.0)),
            easy_iteration_limit=int(adaptive.get("easy_iteration_limit", 4)),
            easy_step_count=int(adaptive.get("easy_step_count", 3)),
        ),
    )
