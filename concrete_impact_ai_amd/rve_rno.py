"""The reference's energy-based RVE recurrent neural operator on the device.

Artifact-compatible inference of ``EnergyDissipationRVERNO``
(``src/concrete_impact/nn/models/rve_rno.py:162-418``): the same parameter names
(a state dict from ``best_model.pt`` loads as it is), the same update
(``update_batch``, ``:249-313``), anchors (``prepare_energy_anchors``,
``:224-247``) and outputs, computed by one fused FP32 matrix-core kernel per call
(``cia_rno_update``; the tangent, which the reference takes by six autograd
passes, by forward-mode differentiation in the same kernel).  Around it:

* ``load_rve_rno_state_dict``   ``src/concrete_impact/nn/deployment.py:30-65``
  (the strict architecture contract against ``artifact_metadata.json``);
* ``RVERNOMaterialAdapter``     ``deployment.py:68-165`` (the material-point
  protocol; also a device kernel for this package's dynamics drivers).

The kernel supports the reference's choices of evolution (``direct_rate``,
``mobility_gradient``) and integrator (``current_state_stress_forward_euler``,
``cs_semi_implicit_forward_euler``) at width 128 with 4 hidden layers (the
artifact's architecture) and latent dimensions 1..8; other widths are rejected.
"""

from __future__ import annotations

import ctypes as C
import json
from dataclasses import dataclass
from pathlib import Path

import numpy as np
import torch

from concrete_impact_ai_amd import _lib
from concrete_impact_ai_amd.materials import (DEFAULT_RESPONSE_REQUIREMENTS, DeviceResponse,
                                              MaterialPointResponse, MaterialState, device,
                                              stream_handle)

HIDDEN = 128
LAYERS = 4
KIN = 16

INTEGRATORS = {"cs_semi_implicit_forward_euler": _lib.CIA_RNO_CS_SEMI_IMPLICIT,
               "current_state_stress_forward_euler": _lib.CIA_RNO_CURRENT_STATE}
EVOLUTIONS = {"direct_rate": _lib.CIA_RNO_DIRECT_RATE,
              "mobility_gradient": _lib.CIA_RNO_MOBILITY_GRADIENT}


class SurrogateCapabilityError(RuntimeError):
    """A requested response is not available from the surrogate (``fem/surrogates``)."""


@dataclass(frozen=True)
class RVERNOModelConfig:
    """``RVERNOModelConfig`` (``src/concrete_impact/nn/config.py:91-113``)."""

    family: str = "energy_rve_rno"
    time_representation: str = "continuous"
    integrator: str = "current_state_stress_forward_euler"
    evolution: str = "direct_rate"
    reference_time_scale: float = 0.02
    latent_dimension: int = 6
    hidden_size: int = HIDDEN
    hidden_layers: int = LAYERS
    activation: str = "silu"
    strain_input_scale: tuple = (1.0,) * 6
    dtype: str = "float32"
    device: str = "cuda"

    def __post_init__(self):
        if self.family != "energy_rve_rno" or self.time_representation != "continuous":
            raise ValueError("RVE-RNO family must be energy_rve_rno with continuous time")
        if self.integrator not in INTEGRATORS:
            raise ValueError(f"unknown RVE-RNO integrator {self.integrator!r}")
        if self.evolution not in EVOLUTIONS:
            raise ValueError(f"unknown RVE-RNO evolution {self.evolution!r}")
        if self.activation != "silu":
            raise ValueError("RVE-RNO activation must be silu")
        if not self.reference_time_scale > 0.0:
            raise ValueError("RVE-RNO reference time scale must be positive")
        if len(self.strain_input_scale) != 6 or min(self.strain_input_scale) <= 0.0:
            raise ValueError("RVE-RNO strain input scale must contain six positive values.")
        if self.dtype != "float32":
            raise ValueError("the RVE-RNO kernel computes in float32 (the artifacts' dtype)")
        if self.device != "cuda":
            raise ValueError("this RVE-RNO runs on the HIP device only (device='cuda')")
        if self.hidden_size != HIDDEN or self.hidden_layers != LAYERS:
            raise ValueError(f"the RVE-RNO kernel is built for {LAYERS} hidden layers of width "
                             f"{HIDDEN}; received {self.hidden_layers} x {self.hidden_size}")
        if not 1 <= self.latent_dimension <= 8:
            raise ValueError("the RVE-RNO kernel supports latent dimensions 1..8")


@dataclass(frozen=True)
class EnergyAnchors:
    """Zero-state energy values and strain gradients (``rve_rno.py:23-30``)."""

    equilibrium_value: torch.Tensor
    equilibrium_strain_gradient: torch.Tensor
    nonequilibrium_value: torch.Tensor
    nonequilibrium_strain_gradient: torch.Tensor
    packed: torch.Tensor  # the 14 values in kernel order


@dataclass(frozen=True)
class RVERNOUpdateBatch:
    """One batched update (``rve_rno.py:33-45``)."""

    stress: torch.Tensor
    latent_state: torch.Tensor
    state_rate: torch.Tensor
    dimensionless_state_rate: torch.Tensor
    dissipation: torch.Tensor
    free_energy: torch.Tensor
    thermodynamic_violation: torch.Tensor
    nonequilibrium_energy_increment: torch.Tensor
    tangent: torch.Tensor | None = None


RVERNOUpdate = RVERNOUpdateBatch  # one point: the same fields without the batch axis


def _network_arrays(state: dict, prefix: str) -> tuple[list, np.ndarray, np.ndarray | None]:
    idx = sorted({int(k[len(prefix) + 1:].split(".")[0]) for k in state
                  if k.startswith(prefix + ".")})
    if len(idx) != LAYERS + 1:
        raise ValueError(f"{prefix}: expected {LAYERS} hidden layers and an output layer")
    hidden = [(np.asarray(state[f"{prefix}.{i}.weight"], np.float32),
               np.asarray(state[f"{prefix}.{i}.bias"], np.float32)) for i in idx[:-1]]
    last = idx[-1]
    bias = state.get(f"{prefix}.{last}.bias")
    return (hidden, np.asarray(state[f"{prefix}.{last}.weight"], np.float32),
            None if bias is None else np.asarray(bias, np.float32))


def _pack_network(hidden, w_out, b_out, n_in: int, n_out: int) -> np.ndarray:
    """One network in the kernel's layout (``include/cia.h``, ``cia_rno_update``)."""
    (w1, b1), rest = hidden[0], hidden[1:]
    if w1.shape != (HIDDEN, n_in) or any(w.shape != (HIDDEN, HIDDEN) for w, _ in rest):
        raise ValueError(f"network shapes do not match {n_in} -> {HIDDEN}^{LAYERS}")
    if w_out.shape != (n_out, HIDDEN):
        raise ValueError(f"output layer has shape {w_out.shape}, expected {(n_out, HIDDEN)}")
    w1p = np.zeros((HIDDEN, KIN), np.float32)
    w1p[:, :n_in] = w1
    rows = 16 * ((n_out + 15) // 16)
    wo = np.zeros((rows, HIDDEN), np.float32)
    wo[:n_out] = w_out
    bo = np.zeros(rows, np.float32)
    if b_out is not None:
        bo[:n_out] = b_out
    parts = [w1p.ravel(), b1]
    for w, b in rest:
        parts += [w.ravel(), b]
    for w, _ in rest:
        parts.append(np.ascontiguousarray(w.T).ravel())
    parts += [np.ascontiguousarray(w1p.T).ravel(), wo.ravel(), bo]
    return np.concatenate(parts).astype(np.float32)


def pack_state_dict(config: RVERNOModelConfig, state: dict) -> np.ndarray:
    """The reference's state dict (``EnergyDissipationRVERNO``) as the kernel's buffer."""
    m = config.latent_dimension
    direct = config.evolution == "direct_rate"
    evo_prefix = "state_evolution.network" if direct else "state_evolution.network.network"
    n_evo = m if direct else m * (m + 1) // 2
    eq = _network_arrays(state, "equilibrium_energy.network.network")
    ne = _network_arrays(state, "nonequilibrium_energy.network.network")
    ev = _network_arrays(state, evo_prefix)
    if eq[2] is not None or ne[2] is not None:
        raise ValueError("the energy networks have no output bias")
    packed = np.concatenate([_pack_network(*eq, 6, 1), _pack_network(*ne, 6 + m, 1),
                             _pack_network(*ev, 6 + m, n_evo)])
    size = C.c_int64()
    _lib.check(_lib.lib().cia_rno_packed_size(m, EVOLUTIONS[config.evolution], C.byref(size)),
               "cia_rno_packed_size")
    if packed.size != size.value:
        raise RuntimeError(f"packed RVE-RNO weights hold {packed.size} floats, the kernel "
                           f"expects {size.value}")
    return packed


def _p(t):
    return C.c_void_p(0 if t is None else t.data_ptr())


class EnergyDissipationRVERNO:
    """The reference model's inference on the device, FP32 (the artifact's dtype)."""

    def __init__(self, config: RVERNOModelConfig, state_dict: dict):
        self.config = config
        self._packed = torch.as_tensor(pack_state_dict(config, state_dict), device=device())
        self._scale = (C.c_double * 6)(*[float(v) for v in config.strain_input_scale])
        self._anchors = None

    @property
    def dtype(self) -> torch.dtype:
        return torch.float32

    @property
    def device(self) -> torch.device:
        return self._packed.device

    def initial_state_batch(self, batch_size: int) -> torch.Tensor:
        return torch.zeros((batch_size, self.config.latent_dimension), dtype=torch.float32,
                           device=self.device)

    def initial_state(self) -> torch.Tensor:
        return torch.zeros(self.config.latent_dimension, dtype=torch.float32, device=self.device)

    def prepare_energy_anchors(self) -> EnergyAnchors:
        """Zero-state anchors, evaluated once by the update kernel's own networks."""
        if self._anchors is None:
            a = torch.empty(14, dtype=torch.float32, device=self.device)
            _lib.check(_lib.lib().cia_rno_anchors(stream_handle(), _p(self._packed),
                                                  self.config.latent_dimension,
                                                  EVOLUTIONS[self.config.evolution], _p(a)),
                       "cia_rno_anchors")
            self._anchors = EnergyAnchors(a[0], a[1:7], a[7], a[8:14], a)
        return self._anchors

    def update_batch(self, strain: torch.Tensor, previous_latent: torch.Tensor,
                     time_step: torch.Tensor, anchors: EnergyAnchors | None = None,
                     compute_tangent: bool = False) -> RVERNOUpdateBatch:
        """Update all points at one time index (``rve_rno.py:249-313``)."""
        b = strain.shape[0]
        m = self.config.latent_dimension
        if strain.shape != (b, 6):
            raise ValueError(f"RVE-RNO strain must have shape [B, 6], received {tuple(strain.shape)}.")
        if previous_latent.shape != (b, m):
            raise ValueError("RVE-RNO latent state must have shape [B, latent_dimension].")
        if time_step.shape != (b,):
            raise ValueError("RVE-RNO time step must have shape [B].")
        anchors = self.prepare_energy_anchors() if anchors is None else anchors
        dev = self.device
        f32 = dict(dtype=torch.float32, device=dev)
        eps = strain.to(**f32).contiguous()
        z = previous_latent.to(**f32).contiguous()
        dt = time_step.to(**f32).contiguous()
        out = RVERNOUpdateBatch(
            stress=torch.empty((b, 6), **f32), latent_state=torch.empty((b, m), **f32),
            state_rate=torch.empty((b, m), **f32),
            dimensionless_state_rate=torch.empty((b, m), **f32),
            dissipation=torch.empty(b, **f32), free_energy=torch.empty(b, **f32),
            thermodynamic_violation=torch.empty(b, **f32),
            nonequilibrium_energy_increment=torch.empty(b, **f32),
            tangent=torch.empty((b, 6, 6), **f32) if compute_tangent else None)
        _lib.check(_lib.lib().cia_rno_update(
            stream_handle(), _p(self._packed), m, EVOLUTIONS[self.config.evolution],
            INTEGRATORS[self.config.integrator], float(self.config.reference_time_scale),
            self._scale, _p(anchors.packed), b, _p(eps), _p(z), _p(dt), _p(out.stress),
            _p(out.latent_state), _p(out.state_rate), _p(out.dimensionless_state_rate),
            _p(out.dissipation), _p(out.free_energy), _p(out.thermodynamic_violation),
            _p(out.nonequilibrium_energy_increment), _p(out.tangent)), "cia_rno_update")
        return out

    def update(self, strain: torch.Tensor, previous_latent: torch.Tensor, time_step: torch.Tensor,
               compute_tangent: bool = False, context=None) -> RVERNOUpdate:
        """The batched update with a batch of one (``rve_rno.py:420-452``)."""
        del context
        if tuple(strain.shape) != (6,):
            raise ValueError(f"RVE-RNO strain must have shape [6], received {tuple(strain.shape)}.")
        if tuple(previous_latent.shape) != (self.config.latent_dimension,):
            raise ValueError("RVE-RNO latent state has an incompatible shape.")
        r = self.update_batch(strain.reshape(1, 6), previous_latent.reshape(1, -1),
                              torch.as_tensor(time_step).reshape(1), None, compute_tangent)
        return RVERNOUpdate(*[None if v is None else v[0] for v in
                              (r.stress, r.latent_state, r.state_rate, r.dimensionless_state_rate,
                               r.dissipation, r.free_energy, r.thermodynamic_violation,
                               r.nonequilibrium_energy_increment, r.tangent)])


_CONTRACT = ("model_family", "integrator", "evolution", "reference_time_scale", "latent_dimension",
             "hidden_size", "hidden_layers", "activation", "strain_input_scale", "dtype",
             "deployment_integrator")


def _expected_contract(config: RVERNOModelConfig) -> dict:
    return {"model_family": config.family, "integrator": config.integrator,
            "evolution": config.evolution, "reference_time_scale": config.reference_time_scale,
            "latent_dimension": config.latent_dimension, "hidden_size": config.hidden_size,
            "hidden_layers": config.hidden_layers, "activation": config.activation,
            "strain_input_scale": tuple(config.strain_input_scale), "dtype": config.dtype,
            "deployment_integrator": "explicit_only"}


def _received_contract(meta: dict) -> dict:
    got = {k: meta.get(k) for k in _CONTRACT}
    got["strain_input_scale"] = tuple(meta.get("strain_input_scale", (1.0,) * 6))
    return got


def _read_state(model_path: Path) -> dict:
    if model_path.suffix == ".npz":
        with np.load(model_path, allow_pickle=False) as z:
            return {k: z[k] for k in z.files if k != "metadata"}
    checkpoint = torch.load(model_path, map_location="cpu", weights_only=True)
    return {k: v.detach().cpu().numpy() for k, v in checkpoint["model_state_dict"].items()}


def load_rve_rno_state_dict(model_config: RVERNOModelConfig, model_path, metadata_path
                            ) -> EnergyDissipationRVERNO:
    """Load one RVE-RNO after strict architecture-contract matching
    (``deployment.py:30-65``).  ``model_path``: the reference's ``best_model.pt``
    (read with ``torch.load(weights_only=True)``) or an ``.npz`` of its state dict."""
    meta = json.loads(Path(metadata_path).read_text(encoding="utf-8"))
    expected, received = _expected_contract(model_config), _received_contract(meta)
    if received != expected:
        raise ValueError("RVE-RNO artifact metadata does not match the requested architecture: "
                         f"expected={expected}; received={received}.")
    return EnergyDissipationRVERNO(model_config, _read_state(Path(model_path)))


def load_rve_rno_artifact(npz_path) -> EnergyDissipationRVERNO:
    """An artifact stored as one ``.npz`` (state-dict arrays and ``metadata``, the
    artifact's JSON): the config is the metadata's, checked by the same contract."""
    with np.load(npz_path, allow_pickle=False) as z:
        meta = json.loads(str(z["metadata"]))
        state = {k: z[k] for k in z.files if k != "metadata"}
    config = RVERNOModelConfig(
        integrator=meta["integrator"], evolution=meta["evolution"],
        reference_time_scale=float(meta["reference_time_scale"]),
        latent_dimension=int(meta["latent_dimension"]), hidden_size=int(meta["hidden_size"]),
        hidden_layers=int(meta["hidden_layers"]), activation=meta["activation"],
        strain_input_scale=tuple(meta.get("strain_input_scale", (1.0,) * 6)), dtype=meta["dtype"])
    if _received_contract(meta) != _expected_contract(config):
        raise ValueError("RVE-RNO artifact metadata is not a deployable contract")
    return EnergyDissipationRVERNO(config, state)


class RVERNOMaterialAdapter:
    """One trained RVE-RNO behind the material-point protocol (``deployment.py:68-165``)."""

    def __init__(self, model: EnergyDissipationRVERNO, name: str, density: float,
                 maximum_wave_speed: float):
        if density <= 0.0 or maximum_wave_speed <= 0.0:
            raise ValueError("RVE-RNO density and wave-speed bound must be positive.")
        self.model = model
        self.name = name
        self.density = float(density)
        self.maximum_wave_speed = float(maximum_wave_speed)

    def initialize_state(self, n_points: int) -> MaterialState:
        return MaterialState({"latent_state": np.zeros((n_points, self.model.config.latent_dimension),
                                                       dtype=np.float64)})

    def update(self, request, state: MaterialState,
               requirements=DEFAULT_RESPONSE_REQUIREMENTS) -> MaterialPointResponse:
        """Update independent points from committed states (``deployment.py:97-165``)."""
        if request.kinematics != "three_dimensional":
            raise ValueError("Fixed-cell RVE-RNO currently requires three-dimensional strain.")
        latent = np.asarray(state.variables["latent_state"])
        strains = np.asarray(request.strains)
        if strains.shape[0] != latent.shape[0]:
            raise ValueError("RVE-RNO strain and latent-state point counts differ.")
        dev = self.model.device
        b = strains.shape[0]
        r = self.model.update_batch(
            torch.as_tensor(strains, dtype=torch.float32, device=dev),
            torch.as_tensor(latent, dtype=torch.float32, device=dev),
            torch.full((b,), float(request.time_step), dtype=torch.float32, device=dev),
            None, compute_tangent=bool(requirements.tangent))
        if requirements.tangent and r.tangent is None:
            raise SurrogateCapabilityError("RVE-RNO tangent was requested but not returned.")

        def host(t):
            return t.double().cpu().numpy()
        return MaterialPointResponse(
            stresses=host(r.stress), state=MaterialState({"latent_state": host(r.latent_state)}),
            tangents=host(r.tangent) if requirements.tangent else None,
            dissipation=host(r.dissipation) if requirements.dissipation else None,
            free_energy=host(r.free_energy) if requirements.free_energy else None,
            diagnostics={"thermodynamic_violation": host(r.thermodynamic_violation),
                         "nonequilibrium_energy_increment": host(r.nonequilibrium_energy_increment)})

    def device_kernel(self, point_phase_ids=None):
        return _RNOKernel(self)


class _RNOKernel:
    """Device protocol of the dynamics drivers: strain [6][n] FP64 in, stress [6][n]
    FP64 out; the committed state is the latent [n][m] FP32 on the device."""

    def __init__(self, adapter: RVERNOMaterialAdapter):
        self.adapter = adapter
        self.model = adapter.model

    def initial_state(self, n: int):
        return self.model.initial_state_batch(n)

    def update(self, strain, rate, state, dt, settings, tangent=False) -> DeviceResponse:
        n = strain.shape[1]
        r = self.model.update_batch(strain.t(), state,
                                    torch.full((n,), float(dt), dtype=torch.float32,
                                               device=self.model.device),
                                    None, compute_tangent=tangent)
        return DeviceResponse(
            stress=r.stress.t().double().contiguous(), state=r.latent_state,
            tangent=r.tangent.reshape(n, 36).t().double().contiguous() if tangent else None,
            free_energy=r.free_energy.double(), dissipation=r.dissipation.double())

    def material_state(self, state) -> MaterialState:
        return MaterialState({"latent_state": state.double().cpu().numpy()})
