"""Material-point protocol and GPU-backed constitutive models.

Mirrors the reference interface so that callers can switch without changes:

* data types            ``src/fem/materials/data.py:16-108``
* failure exception     ``src/fem/materials/errors.py:15-47``
* J2 Perzyna material   ``src/fem/materials/plasticity.py:79-117,223-256``
* linear elastic        ``src/fem/materials/linear_elastic.py:25-62``
* phased dispatch       ``src/fem/materials/phased.py:27-199``
* builder registry      ``src/fem/materials/registry.py:24-43`` (the two kinds on this path)

``update`` accepts and returns NumPy arrays like the reference; the arithmetic
runs in ``libcia_hip.so`` (``cia_material_update_f64``).  Device-resident
callers use :func:`update_device` with torch tensors in SoA layout.
"""

from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import Any, Protocol

import numpy as np
import torch

from concrete_impact_ai_amd import _lib

VOIGT_DIM = 6


# ---------------------------------------------------------------------------
# data types (src/fem/materials/data.py)


@dataclass(frozen=True)
class MaterialUpdateSettings:
    max_iterations: int
    yield_relative_tolerance: float
    residual_absolute_tolerance: float
    residual_relative_tolerance: float


@dataclass(frozen=True)
class MaterialState:
    variables: dict


@dataclass(frozen=True)
class MaterialResponseRequirements:
    tangent: bool = False
    free_energy: bool = False
    dissipation: bool = False


DEFAULT_RESPONSE_REQUIREMENTS = MaterialResponseRequirements()


@dataclass(frozen=True)
class MaterialPointRequest:
    strains: np.ndarray
    strain_rates: np.ndarray | None
    time_step: float
    kinematics: str
    update_settings: MaterialUpdateSettings


@dataclass(frozen=True)
class MaterialPointResponse:
    stresses: np.ndarray
    state: MaterialState
    tangents: np.ndarray | None = None
    free_energy: np.ndarray | None = None
    dissipation: np.ndarray | None = None
    diagnostics: dict = field(default_factory=dict)


class MaterialModel(Protocol):
    @property
    def name(self) -> str: ...

    @property
    def density(self) -> float: ...

    @property
    def maximum_wave_speed(self) -> float: ...

    def initialize_state(self, n_points: int) -> MaterialState: ...

    def update(self, request, state, requirements=DEFAULT_RESPONSE_REQUIREMENTS): ...


def compute_isotropic_pressure_wave_speed(young_modulus, poisson_ratio, density) -> float:
    """``src/fem/materials/data.py:100-108``."""
    bulk = young_modulus / (3.0 * (1.0 - 2.0 * poisson_ratio))
    shear = young_modulus / (2.0 * (1.0 + poisson_ratio))
    return float(np.sqrt((bulk + 4.0 * shear / 3.0) / density))


# ---------------------------------------------------------------------------
# errors (src/fem/materials/errors.py)


def _to_serializable(value: Any) -> Any:
    if isinstance(value, dict):
        return {str(k): _to_serializable(v) for k, v in value.items()}
    if isinstance(value, (list, tuple)):
        return [_to_serializable(v) for v in value]
    if hasattr(value, "item"):
        return _to_serializable(value.item())
    if isinstance(value, float) and not math.isfinite(value):
        return str(value)
    return value


class MaterialPointConvergenceError(RuntimeError):
    """Structured material-point failure (``src/fem/materials/errors.py:15-34``)."""

    def __init__(self, message: str, reason: str, diagnostics: dict) -> None:
        super().__init__(message)
        self.reason = reason
        self.diagnostics = _to_serializable(diagnostics)

    def with_context(self, **context: Any) -> "MaterialPointConvergenceError":
        return MaterialPointConvergenceError(str(self), self.reason, {**self.diagnostics, **context})


# ---------------------------------------------------------------------------
# models


@dataclass(frozen=True)
class J2ViscoplasticMaterial:
    """Rate-dependent J2 (Perzyna) material (``plasticity.py:79-117``)."""

    name: str
    density: float
    young_modulus: float
    poisson_ratio: float
    yield_stress: float
    hardening_modulus: float
    time_scale: float
    reference_stress: float
    rate_exponent: float

    @property
    def maximum_wave_speed(self) -> float:
        return compute_isotropic_pressure_wave_speed(
            self.young_modulus, self.poisson_ratio, self.density
        )

    def initialize_state(self, n_points: int) -> MaterialState:
        return MaterialState(
            variables={
                "plastic_strain": np.zeros((n_points, VOIGT_DIM)),
                "equivalent_plastic_strain": np.zeros(n_points),
                "viscoplastic_multiplier": np.zeros(n_points),
            }
        )

    def device_kernel(self, point_phase_ids=None) -> "PointMaterialKernel":
        return PointMaterialKernel(self, point_phase_ids)

    def update(self, request, state, requirements=DEFAULT_RESPONSE_REQUIREMENTS):
        return _update_host([("", self)], None, request, state, requirements, phased=False)


@dataclass(frozen=True)
class LinearElasticMaterial:
    """Linear elastic material (``linear_elastic.py:25-52``)."""

    name: str
    density: float
    young_modulus: float
    poisson_ratio: float

    @property
    def maximum_wave_speed(self) -> float:
        return compute_isotropic_pressure_wave_speed(
            self.young_modulus, self.poisson_ratio, self.density
        )

    def initialize_state(self, n_points: int) -> MaterialState:
        return MaterialState(variables={})

    def device_kernel(self, point_phase_ids=None) -> "PointMaterialKernel":
        return PointMaterialKernel(self, point_phase_ids)

    def update(self, request, state, requirements=DEFAULT_RESPONSE_REQUIREMENTS):
        return _update_host([("", self)], None, request, state, requirements, phased=False)


@dataclass(frozen=True)
class RVEPhase:
    name: str
    material: Any


@dataclass(frozen=True)
class PhasedMaterialModel:
    """Fixed-topology phase dispatch (``phased.py:35-143``)."""

    name: str
    phases: tuple
    point_phase_ids: np.ndarray
    effective_density: float

    @property
    def density(self) -> float:
        return self.effective_density

    @property
    def maximum_wave_speed(self) -> float:
        return max(p.material.maximum_wave_speed for p in self.phases)

    def initialize_state(self, n_points: int) -> MaterialState:
        if n_points != self.point_phase_ids.size:
            raise ValueError("Phased material point count does not match its phase map.")
        variables = {}
        for pid, phase in enumerate(self.phases):
            count = int(np.count_nonzero(self.point_phase_ids == pid))
            for key, values in phase.material.initialize_state(count).variables.items():
                variables[encode_phase_state_key(phase.name, key)] = values
        return MaterialState(variables)

    def device_kernel(self, point_phase_ids=None) -> "PointMaterialKernel":
        return PointMaterialKernel(self, point_phase_ids)

    def update(self, request, state, requirements=DEFAULT_RESPONSE_REQUIREMENTS):
        if request.strains.shape[0] != self.point_phase_ids.size:
            raise ValueError("Phased material request does not match its fixed phase map.")
        return _update_host(
            [(p.name, p.material) for p in self.phases],
            np.asarray(self.point_phase_ids),
            request,
            state,
            requirements,
            phased=True,
        )


def encode_phase_state_key(phase_name: str, variable_name: str) -> str:
    """``phased.py:158-160``."""
    return f"phase_{phase_name}__{variable_name}"


def decode_phase_state_key(key: str) -> tuple[str, str]:
    """``phased.py:163-168``."""
    prefix, variable = key.split("__", maxsplit=1)
    if not prefix.startswith("phase_"):
        raise ValueError(f"Invalid phased material state key: {key}.")
    return prefix.removeprefix("phase_"), variable


def build_j2_viscoplastic_material(spec: dict) -> J2ViscoplasticMaterial:
    """``plasticity.py:223-256`` (with its parameter validation)."""
    m = J2ViscoplasticMaterial(
        name=str(spec["name"]),
        density=float(spec["density"]),
        young_modulus=float(spec["young_modulus"]),
        poisson_ratio=float(spec["poisson_ratio"]),
        yield_stress=float(spec["yield_stress"]),
        hardening_modulus=float(spec["hardening_modulus"]),
        time_scale=float(spec["time_scale"]),
        reference_stress=float(spec["reference_stress"]),
        rate_exponent=float(spec["rate_exponent"]),
    )
    if m.young_modulus <= 0.0:
        raise ValueError("J2 viscoplasticity requires young_modulus > 0.")
    if not -1.0 < m.poisson_ratio < 0.5:
        raise ValueError("J2 viscoplasticity requires -1 < poisson_ratio < 0.5.")
    if m.yield_stress <= 0.0:
        raise ValueError("J2 viscoplasticity requires yield_stress > 0.")
    if m.hardening_modulus < 0.0:
        raise ValueError("J2 viscoplasticity requires hardening_modulus >= 0.")
    if m.time_scale <= 0.0:
        raise ValueError("J2 viscoplasticity requires time_scale > 0.")
    if m.reference_stress <= 0.0:
        raise ValueError("J2 viscoplasticity requires reference_stress > 0.")
    if m.rate_exponent < 1.0:
        raise ValueError("J2 viscoplasticity requires rate_exponent >= 1.")
    return m


def build_linear_elastic_material(spec: dict) -> LinearElasticMaterial:
    return LinearElasticMaterial(
        name=str(spec["name"]),
        density=float(spec["density"]),
        young_modulus=float(spec["young_modulus"]),
        poisson_ratio=float(spec["poisson_ratio"]),
    )


MATERIAL_BUILDERS = {
    "linear_elastic": build_linear_elastic_material,
    "j2_viscoplastic": build_j2_viscoplastic_material,
}


def build_material(spec: dict):
    """``src/fem/materials/registry.py:33-43`` for the kinds on the FE² path."""
    kind = str(spec["type"])
    if kind not in MATERIAL_BUILDERS:
        raise ValueError(f"Material type {kind!r} is not part of this FE² hot path.")
    return MATERIAL_BUILDERS[kind](spec)


# ---------------------------------------------------------------------------
# marshalling


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise _lib.LibraryUnavailable("no HIP device is visible; the GPU path cannot run")
    return torch.device("cuda", torch.cuda.current_device())


def stream_handle() -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def cia_material(m) -> _lib.Material:
    if isinstance(m, J2ViscoplasticMaterial):
        return _lib.Material(
            _lib.CIA_MAT_J2_VISCOPLASTIC, 0, m.young_modulus, m.poisson_ratio, m.yield_stress,
            m.hardening_modulus, m.time_scale, m.reference_stress, m.rate_exponent, m.density,
        )
    if isinstance(m, LinearElasticMaterial):
        return _lib.Material(
            _lib.CIA_MAT_LINEAR_ELASTIC, 0, m.young_modulus, m.poisson_ratio, 0.0, 0.0, 1.0, 1.0,
            1.0, m.density,
        )
    raise TypeError(f"{type(m).__name__} has no GPU kernel on this path")


def cia_settings(s: MaterialUpdateSettings) -> _lib.UpdateSettings:
    return _lib.UpdateSettings(
        int(s.max_iterations), 0, float(s.yield_relative_tolerance),
        float(s.residual_absolute_tolerance), float(s.residual_relative_tolerance),
    )


def _ptr(t: torch.Tensor | None) -> C.c_void_p:
    return C.c_void_p(0 if t is None else t.data_ptr())


def update_device(
    phases: list,
    phase_id: torch.Tensor | None,
    strain: torch.Tensor,
    strain_rate: torch.Tensor | None,
    state_in: torch.Tensor | None,
    time_step: float,
    settings: MaterialUpdateSettings,
    tangent: bool = False,
    energy: bool = False,
    diagnostics: bool = False,
):
    """Device-resident update on SoA tensors ([6][n] strain, [8][n] state).

    Returns ``(stress, state_out, tangent|None, free_energy|None, dissipation|None,
    diagnostics|None, failure)`` where ``failure`` is None or a PointFailure.
    """
    dtype = strain.dtype
    n = strain.shape[1]
    dev = strain.device
    # The C ABI reads raw row-major SoA buffers: every tensor must be contiguous.
    if strain.shape[0] != 6 or not strain.is_contiguous():
        raise ValueError("strain must be a contiguous [6][n] tensor")
    if state_in is not None and (state_in.shape != (8, n) or state_in.dtype != dtype
                                 or not state_in.is_contiguous()):
        raise ValueError("state must be a contiguous [8][n] tensor of the strain dtype")
    if strain_rate is not None and (strain_rate.shape != (6, n) or strain_rate.dtype != dtype
                                    or not strain_rate.is_contiguous()):
        raise ValueError("strain_rate must be a contiguous [6][n] tensor of the strain dtype")
    if phase_id is not None:
        if phase_id.dtype != torch.int32 or phase_id.shape != (n,):
            raise ValueError("phase_id must be an int32 [n] tensor")
        if n and (int(phase_id.min()) < 0 or int(phase_id.max()) >= len(phases)):
            raise ValueError("phase_id out of range of the phase table")
    mats = (_lib.Material * len(phases))(*[cia_material(m) for m in phases])
    flags = (
        (_lib.CIA_OUT_TANGENT if tangent else 0)
        | (_lib.CIA_OUT_ENERGY if energy else 0)
        | (_lib.CIA_OUT_DIAGNOSTICS if diagnostics else 0)
    )
    stress = torch.empty((6, n), dtype=dtype, device=dev)
    state_out = None if state_in is None else torch.empty_like(state_in)
    tan = torch.empty((36, n), dtype=dtype, device=dev) if tangent else None
    fe = torch.empty(n, dtype=dtype, device=dev) if energy else None
    di = torch.empty(n, dtype=dtype, device=dev) if energy else None
    dg = torch.empty((_lib.N_DIAG, n), dtype=dtype, device=dev) if diagnostics else None
    fail = _lib.PointFailure()
    settings_c = cia_settings(settings)
    fn = _lib.lib().cia_material_update_f64 if dtype == torch.float64 else _lib.lib().cia_material_update_f32
    status = fn(
        stream_handle(), n, mats, len(phases), _ptr(phase_id), _ptr(strain), _ptr(strain_rate),
        _ptr(state_in), float(time_step), C.byref(settings_c), flags, _ptr(stress),
        _ptr(state_out), _ptr(tan), _ptr(fe), _ptr(di), _ptr(dg), C.byref(fail),
    )
    _lib.check(status, "cia_material_update")
    return stress, state_out, tan, fe, di, dg, (fail if status == _lib.CIA_ERR_POINT_FAILURE else None)


def point_failure_error(fail: _lib.PointFailure, material, request, point_id: int, phase_material=None):
    """Build the reference's exception (reason strings and diagnostic keys of
    plasticity.py:549-563,601-612,857-904,1441-1465 with the context of :375-383)."""
    mat = phase_material if phase_material is not None else material
    reason = _lib.POINT_REASONS[fail.reason]
    strain = [float(v) for v in fail.strain]
    if reason in ("non_finite_j2_invariant", "negative_j2_invariant"):
        msg = ("J2 equivalent-stress invariant is non-finite." if reason == "non_finite_j2_invariant"
               else "J2 equivalent-stress invariant is negative.")
        ts = np.asarray(fail.trial_stress, dtype=np.float64)
        dev = ts - np.sum(ts[:3]) / 3.0 * np.asarray([1, 1, 1, 0, 0, 0.0])
        diag = {"algorithm": "j2_equivalent_stress", "radicand": fail.radicand,
                "deviatoric_stress": dev.tolist()}
    elif reason == "invalid_yield_activation_tolerance":
        msg = "J2 viscoplasticity received an invalid yield activation tolerance."
        diag = {"algorithm": "j2_viscoplastic_yield_activation",
                "yield_relative_tolerance": request.update_settings.yield_relative_tolerance,
                "trial_yield_value": fail.trial_yield}
    elif reason == "non_positive_tangent_overstress":
        msg = "J2 viscoplastic tangent requires positive corrected overstress."
        diag = {}
    else:
        msg = f"J2 viscoplastic local Newton failed: {reason}."
        g = mat.young_modulus / (2.0 * (1.0 + mat.poisson_ratio))
        diag = {
            "algorithm": "j2_viscoplastic_local_newton",
            "iteration": int(fail.iteration),
            "max_iterations": request.update_settings.max_iterations,
            "time_step": request.time_step,
            "rate_exponent": mat.rate_exponent,
            "time_scale": mat.time_scale,
            "reference_stress": mat.reference_stress,
            "yield_stress": mat.yield_stress,
            "hardening_modulus": mat.hardening_modulus,
            "shear_modulus": g,
            "hardening_slope": 3.0 * g + mat.hardening_modulus,
            "trial_yield": fail.trial_yield,
            "physical_lower_bound": 0.0,
            "physical_upper_bound": fail.physical_upper_bound,
            "plastic_multiplier": fail.plastic_multiplier,
            "corrected_overstress": fail.corrected_overstress,
            "residual": fail.residual,
            "residual_tolerance": fail.residual_tolerance,
            "derivative": fail.derivative,
            "newton_candidate": fail.newton_candidate,
        }
        diag.update(
            trial_equivalent_stress=fail.trial_equivalent_stress,
            yield_radius=fail.yield_radius,
            equivalent_plastic_strain=fail.equivalent_plastic_strain,
            trial_stress=[float(v) for v in fail.trial_stress],
        )
    err = MaterialPointConvergenceError(msg, reason, diag)
    return err.with_context(
        material_name=mat.name,
        material_type=type(mat).__name__,
        point_id=point_id,
        time_step=request.time_step,
        kinematics=request.kinematics,
        strain=strain,
    )


def _pack_state(phases, point_phase_ids, state: MaterialState, n: int, phased: bool) -> np.ndarray:
    packed = np.zeros((8, n))
    for pid, (pname, mat) in enumerate(phases):
        if not isinstance(mat, J2ViscoplasticMaterial):
            continue
        ids = np.arange(n) if point_phase_ids is None else np.flatnonzero(point_phase_ids == pid)
        pre = f"phase_{pname}__" if phased else ""
        v = state.variables
        packed[0:6, ids] = np.asarray(v[pre + "plastic_strain"]).T
        packed[6, ids] = v[pre + "equivalent_plastic_strain"]
        packed[7, ids] = v[pre + "viscoplastic_multiplier"]
    return packed


def unpack_state(phases, point_phase_ids, packed: np.ndarray, phased: bool) -> dict:
    n = packed.shape[1]
    out = {}
    for pid, (pname, mat) in enumerate(phases):
        if not isinstance(mat, J2ViscoplasticMaterial):
            continue
        ids = np.arange(n) if point_phase_ids is None else np.flatnonzero(point_phase_ids == pid)
        pre = f"phase_{pname}__" if phased else ""
        out[pre + "plastic_strain"] = np.ascontiguousarray(packed[0:6, ids].T)
        out[pre + "equivalent_plastic_strain"] = packed[6, ids].copy()
        out[pre + "viscoplastic_multiplier"] = packed[7, ids].copy()
    return out


def diagnostics_dict(phases, point_phase_ids, diag: np.ndarray, has_rates: bool,
                     phased: bool) -> dict:
    """Select the reference's diagnostic keys from the CIA_N_DIAG rows.

    Single J2 material: lazily created keys (``plasticity.py:1684-1694``);
    phased: sorted union of phase keys, NaN where undefined (``phased.py:118-134``).
    """
    n = diag.shape[1]
    names = set()
    j2_points = np.zeros(n, dtype=bool)
    for pid, (_, mat) in enumerate(phases):
        mask = np.ones(n, dtype=bool) if point_phase_ids is None else point_phase_ids == pid
        if not np.any(mask):
            continue
        if isinstance(mat, J2ViscoplasticMaterial):
            j2_points |= mask
            names.update(_lib.DIAG_NAMES[:20])
            if has_rates:
                names.update(_lib.DIAG_NAMES[20:])
        elif phased:
            names.update(["incremental_work_density", "equivalent_stress_increment",
                          "viscoplastic_active"])
    active_row = diag[_lib.DIAG_NAMES.index("viscoplastic_active")]
    if not np.any(j2_points & (active_row > 0.5)):
        names.discard("viscoplastic_hv")
    ordered = sorted(names) if phased else [k for k in _lib.DIAG_NAMES if k in names]
    return {k: diag[_lib.DIAG_NAMES.index(k)].copy() for k in ordered}


def _update_host(phases, point_phase_ids, request, state, requirements, phased: bool):
    if request.kinematics != "three_dimensional":
        raise ValueError("Plastic material updates currently require three_dimensional kinematics.")
    strains = np.asarray(request.strains, dtype=np.float64)
    if strains.ndim != 2 or strains.shape[1] != VOIGT_DIM:
        raise ValueError("Plastic materials require strain arrays with shape (n_points, 6).")
    n = strains.shape[0]
    mats = [m for _, m in phases]
    has_j2 = any(isinstance(m, J2ViscoplasticMaterial) for m in mats)
    if phased and request.strain_rates is None and any(
        isinstance(m, LinearElasticMaterial) for m in mats
    ):
        raise ValueError("Elastic RVE phase diagnostics require strain rates and tangents.")
    if n == 0:
        empty = np.zeros((0, VOIGT_DIM))
        return MaterialPointResponse(
            stresses=empty, state=state,
            tangents=np.zeros((0, 6, 6)) if requirements.tangent else None,
            free_energy=np.zeros(0) if requirements.free_energy else None,
            dissipation=np.zeros(0) if requirements.dissipation else None,
        )
    dev = device()
    strain_t = torch.as_tensor(np.ascontiguousarray(strains.T), device=dev)
    rate_t = (None if request.strain_rates is None
              else torch.as_tensor(np.ascontiguousarray(np.asarray(request.strain_rates).T), device=dev))
    state_t = (torch.as_tensor(_pack_state(phases, point_phase_ids, state, n, phased), device=dev)
               if has_j2 else None)
    pid_t = (None if point_phase_ids is None
             else torch.as_tensor(point_phase_ids.astype(np.int32), device=dev))
    need_tan = requirements.tangent or (phased and any(
        isinstance(m, LinearElasticMaterial) for m in mats))
    stress, st_out, tan, fe, di, dg, fail = update_device(
        mats, pid_t, strain_t, rate_t, state_t, request.time_step, request.update_settings,
        tangent=need_tan, energy=True, diagnostics=True,
    )
    if fail is not None:
        p = int(fail.point_id)
        phase_mat = mats[0] if point_phase_ids is None else mats[int(point_phase_ids[p])]
        local = p if point_phase_ids is None else int(
            np.count_nonzero(point_phase_ids[:p] == point_phase_ids[p]))
        raise point_failure_error(fail, phase_mat, request, local, phase_mat)
    stresses = stress.cpu().numpy().T.copy()
    diag = dg.cpu().numpy()
    variables = dict(state.variables)
    if has_j2:
        variables.update(unpack_state(phases, point_phase_ids, st_out.cpu().numpy(), phased))
    single_elastic = (not phased) and not has_j2
    return MaterialPointResponse(
        stresses=stresses,
        state=MaterialState(variables),
        tangents=(tan.cpu().numpy().T.reshape(n, 6, 6).copy() if requirements.tangent else None),
        free_energy=fe.cpu().numpy().copy() if requirements.free_energy else None,
        dissipation=di.cpu().numpy().copy() if requirements.dissipation else None,
        diagnostics={} if single_elastic else diagnostics_dict(
            phases, point_phase_ids, diag, request.strain_rates is not None, phased),
    )


# ---------------------------------------------------------------------------
# Device-resident material protocol used by the generic macro drivers
# (explicit dynamics with any MaterialModel, Newmark).  A material exposes
# ``device_kernel(point_phase_ids)`` returning an object with
#   initial_state(n) -> opaque committed state
#   update(strain [6][n], rate [6][n] | None, state, dt, settings, tangent) -> DeviceResponse
#   material_state(state) -> MaterialState (host view with the reference keys)


@dataclass
class DeviceResponse:
    """Point responses on the device (FP64 SoA): stress [6][n], tangent [36][n]
    (row-major 6x6 per point) or None, and the [n] point arrays that the energy
    histories integrate (free energy, dissipation, incremental work density,
    viscoplastic-active value, yield activation margin), each possibly None."""

    stress: torch.Tensor
    state: Any
    tangent: torch.Tensor | None = None
    free_energy: torch.Tensor | None = None
    dissipation: torch.Tensor | None = None
    work: torch.Tensor | None = None
    active: torch.Tensor | None = None
    margin: torch.Tensor | None = None


class PointMaterialKernel:
    """Device kernel of the point materials (J2 Perzyna, linear elastic, phased)."""

    def __init__(self, material, point_phase_ids=None):
        if isinstance(material, PhasedMaterialModel):
            self.phases = [(p.name, p.material) for p in material.phases]
            self.phased = True
            ppid = np.asarray(material.point_phase_ids if point_phase_ids is None else point_phase_ids)
        elif isinstance(material, (J2ViscoplasticMaterial, LinearElasticMaterial)):
            self.phases = [("", material)]
            self.phased = False
            ppid = None
        else:
            raise TypeError(f"{type(material).__name__} is not a point material")
        self.mats = [m for _, m in self.phases]
        self.has_j2 = any(isinstance(m, J2ViscoplasticMaterial) for m in self.mats)
        self.ppid = ppid
        self._pid_t = None if ppid is None else torch.as_tensor(ppid.astype(np.int32), device=device())

    def initial_state(self, n: int):
        return torch.zeros((8, n), dtype=torch.float64, device=device()) if self.has_j2 else None

    def update(self, strain, rate, state, dt, settings, tangent=False) -> DeviceResponse:
        stress, st_out, tan, fe, di, dg, fail = update_device(
            self.mats, self._pid_t, strain, rate, state, dt, settings,
            tangent=tangent, energy=True, diagnostics=True)
        if fail is not None:
            p = int(fail.point_id)
            ph = self.mats[0] if self.ppid is None else self.mats[int(self.ppid[p])]
            req = type("Req", (), {"time_step": dt, "kinematics": "three_dimensional",
                                   "update_settings": settings})
            raise point_failure_error(fail, ph, req, p, ph)
        d = {n: i for i, n in enumerate(_lib.DIAG_NAMES)}
        has_diag = self.has_j2 or self.phased
        return DeviceResponse(
            stress=stress, state=st_out if self.has_j2 else None, tangent=tan,
            free_energy=fe, dissipation=di,
            work=dg[d["incremental_work_density"]].contiguous() if has_diag else None,
            active=dg[d["viscoplastic_active"]].contiguous() if has_diag else None,
            margin=dg[d["yield_activation_margin"]].contiguous() if has_diag else None)

    def material_state(self, state) -> MaterialState:
        if state is None:
            return MaterialState({})
        return MaterialState(unpack_state(self.phases, self.ppid, state.cpu().numpy(), self.phased))

    def equivalent_plastic_strain(self, state):
        return None if state is None else state[6]
