"""ctypes binding of libcia_hip.so (declared in include/cia.h).

The library is the only compute path: when it is missing or no HIP device is
visible, :func:`lib` raises instead of falling back to anything else.
"""

from __future__ import annotations

import ctypes as C
import os
import threading
from pathlib import Path

_LIB_PATH = Path(__file__).resolve().parent / "libcia_hip.so"
_lock = threading.Lock()
_lib = None

# ---- status / enums (mirrors include/cia.h) --------------------------------
CIA_OK = 0
CIA_ERR_INVALID_ARGUMENT = 1
CIA_ERR_HIP = 2
CIA_ERR_POINT_FAILURE = 3
CIA_ERR_RVE_FAILURE = 4
CIA_ERR_OUT_OF_MEMORY = 5
CIA_ERR_NOT_CONVERGED = 6

CIA_MAT_LINEAR_ELASTIC = 0
CIA_MAT_J2_VISCOPLASTIC = 1

CIA_OUT_TANGENT = 1
CIA_OUT_ENERGY = 2
CIA_OUT_DIAGNOSTICS = 4

CIA_PREC_FP64 = 0
CIA_PREC_FP32_REFINED = 1

CIA_RNO_DIRECT_RATE = 0
CIA_RNO_MOBILITY_GRADIENT = 1
CIA_RNO_CS_SEMI_IMPLICIT = 0
CIA_RNO_CURRENT_STATE = 1

# Point failure reason strings, index = CIA_PT_* code (plasticity.py:543-854,918,1441-1465).
POINT_REASONS = [
    "ok",
    "non_finite_j2_invariant",
    "negative_j2_invariant",
    "invalid_yield_activation_tolerance",
    "non_finite_closed_form",
    "closed_form_solution_out_of_bounds",
    "invalid_iteration_limit",
    "non_finite_problem_scale",
    "invalid_corrected_overstress",
    "non_finite_iteration_value",
    "non_positive_derivative",
    "converged_solution_out_of_bounds",
    "non_finite_newton_candidate",
    "newton_candidate_out_of_bounds",
    "maximum_iterations_exceeded",
    "non_positive_tangent_overstress",
]

RVE_REASONS = [
    "ok",
    "maximum_iterations_exceeded",
    "non_finite_increment",
    "armijo_exhausted",
    "non_finite_effective_tangent_sensitivity",
    "material_point_failure",
]

# Diagnostic row names, index = CIA_D_* (plasticity.py:578-595,641-678,704-708,1321-1327).
DIAG_NAMES = [
    "yield_value",
    "plastic_work",
    "local_iterations",
    "active_surface",
    "plastic_multiplier",
    "viscoplastic_multiplier",
    "equivalent_stress",
    "trial_yield_value",
    "yield_scale",
    "yield_activation_threshold",
    "yield_activation_margin",
    "viscoplastic_active",
    "time_step",
    "equivalent_plastic_strain",
    "viscoplastic_hv",
    "plastic_power",
    "hardening_storage_rate",
    "local_residual",
    "root_lower_bound",
    "root_upper_bound",
    "incremental_work_density",
    "equivalent_stress_increment",
]
N_DIAG = len(DIAG_NAMES)

FIELD_MICRO_STRAIN = 0
FIELD_MICRO_STRAIN_RATE = 1
FIELD_MICRO_STRESS = 2
FIELD_FREE_ENERGY = 3
FIELD_DISSIPATION = 4
FIELD_DIAGNOSTICS = 5
FIELD_DISPLACEMENT = 6
FIELD_INTERNAL_FORCE = 7


class Material(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("pad", C.c_int32),
        ("young_modulus", C.c_double),
        ("poisson_ratio", C.c_double),
        ("yield_stress", C.c_double),
        ("hardening_modulus", C.c_double),
        ("time_scale", C.c_double),
        ("reference_stress", C.c_double),
        ("rate_exponent", C.c_double),
        ("density", C.c_double),
    ]


class UpdateSettings(C.Structure):
    _fields_ = [
        ("max_iterations", C.c_int32),
        ("pad", C.c_int32),
        ("yield_relative_tolerance", C.c_double),
        ("residual_absolute_tolerance", C.c_double),
        ("residual_relative_tolerance", C.c_double),
    ]


class PointFailure(C.Structure):
    _fields_ = [
        ("point_id", C.c_int64),
        ("reason", C.c_int32),
        ("iteration", C.c_int32),
        ("plastic_multiplier", C.c_double),
        ("corrected_overstress", C.c_double),
        ("residual", C.c_double),
        ("derivative", C.c_double),
        ("newton_candidate", C.c_double),
        ("trial_yield", C.c_double),
        ("physical_upper_bound", C.c_double),
        ("residual_tolerance", C.c_double),
        ("trial_equivalent_stress", C.c_double),
        ("yield_radius", C.c_double),
        ("equivalent_plastic_strain", C.c_double),
        ("radicand", C.c_double),
        ("trial_stress", C.c_double * 6),
        ("strain", C.c_double * 6),
    ]


class NewtonSettings(C.Structure):
    _fields_ = [
        ("max_iterations", C.c_int32),
        ("armijo_enabled", C.c_int32),
        ("residual_absolute_tolerance", C.c_double),
        ("residual_relative_tolerance", C.c_double),
        ("increment_absolute_tolerance", C.c_double),
        ("increment_relative_tolerance", C.c_double),
        ("armijo_sufficient_decrease", C.c_double),
        ("armijo_reduction_factor", C.c_double),
        ("armijo_max_backtracks", C.c_int32),
        ("pad", C.c_int32),
    ]


class LinearSettings(C.Structure):
    _fields_ = [
        ("max_iterations", C.c_int32),
        ("check_interval", C.c_int32),
        ("relative_tolerance", C.c_double),
        ("absolute_tolerance", C.c_double),
        ("preconditioner", C.c_int32),
        ("adaptive_forcing", C.c_int32),
        ("reference_bulk", C.c_double),
        ("reference_shear", C.c_double),
        ("forcing_max", C.c_double),
    ]


CIA_PCG_SPECTRAL = 0
CIA_PCG_BLOCK_JACOBI = 1


class RVEDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int32),
        ("ny", C.c_int32),
        ("nz", C.c_int32),
        ("n_phases", C.c_int32),
        ("lengths", C.c_double * 3),
        ("phases", C.POINTER(Material)),
        ("element_phase", C.POINTER(C.c_int32)),
        ("newton", NewtonSettings),
        ("linear", LinearSettings),
        ("hill_mandel_power_scale", C.c_double),
        ("precision", C.c_int32),
        ("pad", C.c_int32),
    ]


class RVEResult(C.Structure):
    _fields_ = [
        ("reason", C.c_int32),
        ("newton_iterations", C.c_int32),
        ("armijo_backtracks", C.c_int32),
        ("linear_iterations", C.c_int32),
        ("gp_updates", C.c_int64),
        ("macro_stress", C.c_double * 6),
        ("effective_tangent", C.c_double * 36),
        ("free_energy_density", C.c_double),
        ("dissipation_density", C.c_double),
        ("micro_residual_norm", C.c_double),
        ("micro_residual_tolerance", C.c_double),
        ("volume", C.c_double),
        ("virtual_work_stress_error", C.c_double),
        ("periodic_fluctuation_error", C.c_double),
        ("reaction_antiperiodicity_error", C.c_double),
        ("periodic_class_reaction_error", C.c_double),
        ("macro_work_increment", C.c_double),
        ("micro_work_increment", C.c_double),
        ("hill_mandel_residual", C.c_double),
        ("hill_mandel_error", C.c_double),
        ("positive_work_density", C.c_double),
        ("negative_work_density", C.c_double),
        ("unloading_volume_fraction", C.c_double),
        ("viscoplastic_active_volume_fraction", C.c_double),
        ("yield_activation_margin_available", C.c_int32),
        ("pad", C.c_int32),
        ("minimum_yield_activation_margin", C.c_double),
        ("maximum_yield_activation_margin", C.c_double),
        ("failure_residual_norm", C.c_double),
        ("failure_residual_tolerance", C.c_double),
        ("failure_increment_norm", C.c_double),
        ("failure_increment_tolerance", C.c_double),
        ("failure_iteration", C.c_int32),
        ("pad2", C.c_int32),
        ("point_failure", PointFailure),
        ("phase_volume", C.c_double * 8),
        ("phase_average_stress", (C.c_double * 6) * 8),
        ("phase_free_energy_density", C.c_double * 8),
        ("phase_dissipation_density", C.c_double * 8),
    ]


_vp = C.c_void_p
_SIGNATURES = {
    "cia_abi_version": (C.c_int32, []),
    "cia_device_name": (C.c_int32, [C.c_char_p, C.c_int32]),
    "cia_last_error": (C.c_int32, [C.c_char_p, C.c_int32]),
    "cia_material_update_f64": (
        C.c_int32,
        [_vp, C.c_int64, C.POINTER(Material), C.c_int32, _vp, _vp, _vp, _vp, C.c_double,
         C.POINTER(UpdateSettings), C.c_uint32, _vp, _vp, _vp, _vp, _vp, _vp,
         C.POINTER(PointFailure)],
    ),
    "cia_material_update_f32": (
        C.c_int32,
        [_vp, C.c_int64, C.POINTER(Material), C.c_int32, _vp, _vp, _vp, _vp, C.c_double,
         C.POINTER(UpdateSettings), C.c_uint32, _vp, _vp, _vp, _vp, _vp, _vp,
         C.POINTER(PointFailure)],
    ),
    "cia_rve_create": (C.c_int32, [C.POINTER(RVEDesc), C.POINTER(_vp)]),
    "cia_rve_destroy": (C.c_int32, [_vp]),
    "cia_rve_sizes": (C.c_int32, [_vp, C.POINTER(C.c_int64)]),
    "cia_rve_solve": (
        C.c_int32,
        [_vp, _vp, _vp, C.POINTER(C.c_double), _vp, C.POINTER(C.c_double), C.c_double,
         C.POINTER(UpdateSettings), C.c_int32, _vp, _vp, C.POINTER(RVEResult)],
    ),
    "cia_rve_copy_field": (C.c_int32, [_vp, _vp, C.c_int32, _vp]),
    "cia_rve_batch_create": (C.c_int32, [C.POINTER(RVEDesc), C.c_int32, C.c_int32, C.POINTER(_vp)]),
    "cia_rve_batch_destroy": (C.c_int32, [_vp]),
    "cia_rve_batch_layout": (C.c_int32, [_vp, C.POINTER(C.c_int64)]),
    "cia_rve_batch_solve": (
        C.c_int32,
        [_vp, _vp, _vp, _vp, _vp, C.POINTER(C.c_double), C.POINTER(C.c_double),
         C.POINTER(C.c_double), C.POINTER(UpdateSettings), _vp, _vp, C.POINTER(RVEResult)],
    ),
    "cia_rve_batch_effective_tangent": (C.c_int32, [_vp, _vp, _vp, _vp]),
    "cia_rve_batch_tuning": (C.c_int32, [_vp, C.POINTER(C.c_int32), C.POINTER(C.c_double),
                                         C.c_int32]),
    "cia_rve_batch_set_tuning": (C.c_int32, [_vp, C.POINTER(C.c_int32), C.c_int32]),
    "cia_rve_batch_stats": (C.c_int32, [_vp, C.POINTER(C.c_int64), C.c_int32]),
    "cia_profile_enable": (C.c_int32, [C.c_int32]),
    "cia_profile_read": (C.c_int32, [C.POINTER(C.c_double), C.POINTER(C.c_int64), C.c_int32]),
}

class BoxDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int32),
        ("ny", C.c_int32),
        ("nz", C.c_int32),
        ("n_phases", C.c_int32),
        ("lengths", C.c_double * 3),
        ("phases", C.POINTER(Material)),
        ("element_phase", C.POINTER(C.c_int32)),
        ("precision", C.c_int32),
        ("pad", C.c_int32),
    ]


class SurrogateGate(C.Structure):
    _fields_ = [(name, C.c_double) for name in (
        "eps_scale", "rate_ref", "rate_scale", "inelastic_step", "d_eps_max", "rate_lo",
        "rate_hi", "latent_max", "gate_margin", "rate_margin", "commit_strain",
        "commit_time")] + [("stiffness", C.c_double * 36)]


N_HISTORY = 10
HISTORY_COLUMNS = ["kinetic", "free_energy", "dissipation_rate", "ext_du", "positive_work",
                   "negative_work", "unloading_volume", "active_volume", "margin_min",
                   "margin_max"]

_SIGNATURES.update({
    "cia_box_create": (C.c_int32, [C.POINTER(BoxDesc), C.POINTER(_vp)]),
    "cia_box_destroy": (C.c_int32, [_vp]),
    "cia_box_explicit_steps": (
        C.c_int32,
        [_vp, _vp, C.c_int32, C.c_double, C.c_int32, C.POINTER(C.c_double),
         C.POINTER(UpdateSettings), _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
         C.POINTER(PointFailure)],
    ),
    "cia_box_generic_strains": (
        C.c_int32, [_vp, _vp, C.c_int32, C.c_double, _vp, _vp, _vp, _vp, _vp, _vp]),
    "cia_box_generic_correct": (
        C.c_int32,
        [_vp, _vp, C.c_int32, C.c_double, C.c_double, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
         _vp, _vp, _vp, _vp, _vp, _vp]),
    "cia_box_strains_between": (C.c_int32, [_vp, _vp, C.c_double, _vp, _vp, _vp, _vp]),
    "cia_box_newmark_residual": (
        C.c_int32,
        [_vp, _vp, C.c_double, C.c_double, C.c_double, C.c_double, _vp, C.c_double, _vp, _vp,
         _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.POINTER(C.c_double)]),
    "cia_box_mass_apply": (C.c_int32, [_vp, _vp, C.c_double, _vp, _vp]),
    "cia_box_newmark_solve": (
        C.c_int32,
        [_vp, _vp, _vp, C.c_double, C.c_double, _vp, _vp, _vp, C.c_int32, C.c_int32, C.c_double,
         C.c_double, C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
    "cia_mlp_linear_forward": (
        C.c_int32, [_vp, C.c_int64, C.c_int64, C.c_int64, _vp, _vp, _vp, C.c_int32, _vp, _vp, _vp]),
    "cia_mlp_linear_backward_data": (
        C.c_int32, [_vp, C.c_int64, C.c_int64, C.c_int64, _vp, _vp, _vp, _vp, _vp]),
    "cia_mlp_linear_backward_weight": (
        C.c_int32, [_vp, C.c_int64, C.c_int64, C.c_int64, _vp, _vp, _vp, _vp, _vp, C.c_int64,
                    C.c_int32]),
    "cia_mlp_backward_weight_workspace": (C.c_int64, [C.c_int64, C.c_int64]),
    "cia_mlp_linear_backward_weight_noalloc": (
        C.c_int32, [_vp, C.c_int64, C.c_int64, C.c_int64, _vp, _vp, _vp, _vp, C.c_int32]),
    "cia_mlp_adam_step": (
        C.c_int32, [_vp, C.c_int64, _vp, _vp, _vp, _vp, _vp, C.c_float, C.c_float, C.c_float,
                    C.c_float, C.c_int32, C.c_float]),
    "cia_f32_to_bf16": (C.c_int32, [_vp, C.c_int64, _vp, _vp]),
    "cia_mlp_forward_fused": (C.c_int32, [_vp, C.c_int64, _vp, _vp, _vp, _vp]),
    "cia_surrogate_bptt_workspace": (C.c_int64, [C.c_int64, C.c_int64]),
    "cia_surrogate_bptt": (
        C.c_int32, [_vp, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_int64, _vp, _vp, _vp,
                    _vp, _vp, C.POINTER(C.c_float), C.c_float, _vp, _vp, _vp, _vp, _vp, _vp,
                    C.c_float, _vp, _vp, _vp, C.c_int64]),
    "cia_surrogate_update": (
        C.c_int32, [_vp, C.c_int64, _vp, C.c_double, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                    _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "cia_rno_packed_size": (C.c_int32, [C.c_int32, C.c_int32, C.POINTER(C.c_int64)]),
    "cia_rno_anchors": (C.c_int32, [_vp, _vp, C.c_int32, C.c_int32, _vp]),
    "cia_rno_update": (
        C.c_int32, [_vp, _vp, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                    C.POINTER(C.c_double), _vp, C.c_int64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                    _vp, _vp, _vp, _vp]),
})

BATCH_KERNELS = ["assemble", "gather", "pcg_elem", "pcg_node", "pcg_update", "symbol_solve",
                 "dft_x", "dft_y", "dft_z"]

PROFILE_SLOTS = ["assemble", "pcg_elem", "pcg_node", "pcg_update", "precond", "gather"]

# Entry points whose prototypes are declared by optional translation units.
_OPTIONAL: dict[str, tuple] = {}


def register_signatures(sigs: dict[str, tuple]) -> None:
    """Add prototypes of further entry points (used by the other modules)."""
    _OPTIONAL.update(sigs)
    if _lib is not None:
        _apply(_lib, sigs)


def _apply(handle, sigs):
    for name, (res, args) in sigs.items():
        fn = getattr(handle, name)
        fn.restype = res
        fn.argtypes = args


class LibraryUnavailable(RuntimeError):
    """libcia_hip.so is not built or cannot be loaded on this host."""


def library_path() -> Path:
    return _LIB_PATH


def lib():
    """Load libcia_hip.so once; raise LibraryUnavailable when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not _LIB_PATH.exists():
                raise LibraryUnavailable(
                    f"{_LIB_PATH} is missing: build it with `python -m concrete_impact_ai_amd.build`"
                )
            handle = C.CDLL(str(_LIB_PATH), mode=os.RTLD_NOW | os.RTLD_GLOBAL)
            _apply(handle, _SIGNATURES)
            _apply(handle, _OPTIONAL)
            _lib = handle
    return _lib


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    lib().cia_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


class CIAError(RuntimeError):
    def __init__(self, status: int, where: str):
        super().__init__(f"{where} failed with status {status}: {last_error()}")
        self.status = status


def check(status: int, where: str) -> None:
    if status not in (CIA_OK, CIA_ERR_POINT_FAILURE, CIA_ERR_RVE_FAILURE):
        raise CIAError(status, where)


def exported_symbols() -> list[str]:
    return sorted(set(_SIGNATURES) | set(_OPTIONAL))
