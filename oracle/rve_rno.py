"""Oracle: the reference's energy-based RVE recurrent neural operator in float64
(test infrastructure only, see ``oracle/__init__.py``).

Restates ``EnergyDissipationRVERNO`` (``src/concrete_impact/nn/models/rve_rno.py:
162-418``) for one time index of a batch of material points.  Networks
(``rve_rno.py:63-129``): ``hidden_layers`` dense layers of width ``hidden_size``
with SiLU, then a linear output (no bias for the two scalar energies; a bias for
the state evolution).  With e = strain / strain_input_scale,

    psi_eq(e), psi_neq([e, z])                         scalar energies
    anchors: values and strain gradients at e = 0, z = 0 (``:224-247``)
    beta  = d psi_neq / dz at (e, z_prev)
    direct_rate:       r = evo([e, z_prev])                        D = -beta . r / tau
    mobility_gradient: L = tril(evo([e, z_prev])), p = L^T (-beta),
                       r = L p,                                     D = p . p / tau
    z_cand = z_prev + (dt / tau) r,  state_rate = r / tau
    sigma = (d psi_eq/de - a_eq) / s + (d psi_neq/de at S - a_neq) / s
    with S = (e, z_cand) for ``current_state_stress_forward_euler`` and
    S = (e, z_prev) for ``cs_semi_implicit_forward_euler`` (``:389-406``).

The input gradient is the explicit SiLU chain of ``ScalarMLP.value_and_input_gradient``
(``:110-129``).  The tangent d sigma / d strain, which the reference takes by six
autograd passes (``:290-302``), is here the forward-mode derivative of that chain
(silu'' = s(1-s)(2 + a(1-2s)), s the logistic function) through the state update.

Pinning (tests/test_oracle_rve_rno.py): values and input gradients and Hessian
products of every network against torch's own ``nn.Linear``/``nn.SiLU`` modules
loaded strictly with the artifact's state-dict names, and torch autograd; zero
anchoring; stress equal to the autograd strain gradient of the anchored energy at
the candidate state (the reference test ``test_rve_rno_training.py:112-171``);
the tangent against central differences of the stress; D = -beta . state_rate for
both evolutions and D >= 0 for the mobility form; the latent increment, state rate,
violation and energy increment identities of ``test_rve_rno_training.py:51-110``;
and the artifact's stress and tangent along the reference-held FE² sample path.
The reference module itself cannot be imported here (gmsh is absent), so no
reference-computed vectors exist.
"""

from __future__ import annotations

import json
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

ENERGY_NETWORKS = ("equilibrium_energy", "nonequilibrium_energy")


def logistic(a: np.ndarray) -> np.ndarray:
    return 0.5 * (1.0 + np.tanh(0.5 * a))


def silu(a: np.ndarray) -> np.ndarray:
    return a * logistic(a)


def silu_d1(a: np.ndarray) -> np.ndarray:
    s = logistic(a)
    return s * (1.0 + a * (1.0 - s))


def silu_d2(a: np.ndarray) -> np.ndarray:
    s = logistic(a)
    return s * (1.0 - s) * (2.0 + a * (1.0 - 2.0 * s))


@dataclass
class Network:
    """Dense SiLU stack: ``hidden`` = [(W [out, in], b)], then (W_out, b_out or None)."""

    hidden: list
    out_weight: np.ndarray
    out_bias: np.ndarray | None

    @classmethod
    def from_state(cls, state: dict, prefix: str) -> "Network":
        """``prefix`` is the ``nn.Sequential`` path, e.g. ``equilibrium_energy.network.network``."""
        idx = sorted({int(k[len(prefix) + 1:].split(".")[0]) for k in state
                      if k.startswith(prefix + ".")})
        layers = [(np.asarray(state[f"{prefix}.{i}.weight"], np.float64),
                   np.asarray(state[f"{prefix}.{i}.bias"], np.float64)) for i in idx[:-1]]
        last = idx[-1]
        bias = state.get(f"{prefix}.{last}.bias")
        return cls(layers, np.asarray(state[f"{prefix}.{last}.weight"], np.float64),
                   None if bias is None else np.asarray(bias, np.float64))

    def forward(self, x: np.ndarray):
        """Output [B, out] and the preactivations of every hidden layer."""
        pre = []
        h = x
        for w, b in self.hidden:
            a = h @ w.T + b
            pre.append(a)
            h = silu(a)
        out = h @ self.out_weight.T
        if self.out_bias is not None:
            out = out + self.out_bias
        return out, pre

    def forward_jvp(self, x: np.ndarray, v: np.ndarray):
        """Output and its directional derivatives: x [B, n], v [B, K, n] -> [B, K, out]."""
        h, dh = x, v
        for w, b in self.hidden:
            a = h @ w.T + b
            da = dh @ w.T
            dh = silu_d1(a)[:, None, :] * da
            h = silu(a)
        out = h @ self.out_weight.T
        if self.out_bias is not None:
            out = out + self.out_bias
        return out, dh @ self.out_weight.T

    def value_and_gradient(self, x: np.ndarray):
        """Scalar energy and its input gradient (``rve_rno.py:110-129``)."""
        out, pre = self.forward(x)
        g = np.broadcast_to(self.out_weight[0], (x.shape[0], self.out_weight.shape[1]))
        for (w, _), a in zip(reversed(self.hidden), reversed(pre)):
            g = (g * silu_d1(a)) @ w
        return out[:, 0], g

    def gradient_hvp(self, x: np.ndarray, v: np.ndarray):
        """Value, input gradient and Hessian products H v: v [B, K, n] -> [B, K, n]."""
        h, dh = x, v
        pre, dpre = [], []
        for w, b in self.hidden:
            a = h @ w.T + b
            da = dh @ w.T
            pre.append(a)
            dpre.append(da)
            dh = silu_d1(a)[:, None, :] * da
            h = silu(a)
        value = (h @ self.out_weight.T)[:, 0]
        g = np.broadcast_to(self.out_weight[0], h.shape)
        dg = np.zeros(dh.shape)
        for (w, _), a, da in zip(reversed(self.hidden), reversed(pre), reversed(dpre)):
            d1 = silu_d1(a)
            dt = dg * d1[:, None, :] + (g * silu_d2(a))[:, None, :] * da
            g = (g * d1) @ w
            dg = dt @ w
        return value, g, dg


@dataclass
class RNOConfig:
    """``RVERNOModelConfig`` (``src/concrete_impact/nn/config.py:91-113``), the fields used."""

    integrator: str = "current_state_stress_forward_euler"
    evolution: str = "direct_rate"
    reference_time_scale: float = 0.02
    latent_dimension: int = 6
    strain_input_scale: tuple = field(default=(1.0,) * 6)


class RVERNO:
    """The float64 model over a state dict with the reference's parameter names."""

    def __init__(self, config: RNOConfig, state: dict):
        if config.integrator not in ("current_state_stress_forward_euler",
                                     "cs_semi_implicit_forward_euler"):
            raise ValueError(f"unknown integrator {config.integrator}")
        if config.evolution not in ("direct_rate", "mobility_gradient"):
            raise ValueError(f"unknown evolution {config.evolution}")
        self.config = config
        self.scale = np.asarray(config.strain_input_scale, np.float64)
        self.eq = Network.from_state(state, "equilibrium_energy.network.network")
        self.neq = Network.from_state(state, "nonequilibrium_energy.network.network")
        evo_prefix = ("state_evolution.network" if config.evolution == "direct_rate"
                      else "state_evolution.network.network")
        self.evo = Network.from_state(state, evo_prefix)
        m = config.latent_dimension
        self.tril = np.tril_indices(m)

    @classmethod
    def from_artifact(cls, path: str | Path) -> "RVERNO":
        """The artifact fixture: state-dict arrays and ``metadata`` (the artifact JSON)."""
        with np.load(path, allow_pickle=False) as z:
            state = {k: z[k] for k in z.files if k != "metadata"}
            meta = json.loads(str(z["metadata"]))
        config = RNOConfig(integrator=meta["integrator"], evolution=meta["evolution"],
                           reference_time_scale=float(meta["reference_time_scale"]),
                           latent_dimension=int(meta["latent_dimension"]),
                           strain_input_scale=tuple(meta.get("strain_input_scale", (1.0,) * 6)))
        return cls(config, state)

    def anchors(self):
        """Zero-state values and strain gradients (``rve_rno.py:224-247``)."""
        m = self.config.latent_dimension
        v_eq, g_eq = self.eq.value_and_gradient(np.zeros((1, 6)))
        v_ne, g_ne = self.neq.value_and_gradient(np.zeros((1, 6 + m)))
        return v_eq[0], g_eq[0], v_ne[0], g_ne[0, :6]

    def _rate(self, x_prev, beta, v=None):
        """Dimensionless rate r (and its directional derivatives for x directions v, with
        dbeta the matching derivatives of beta) for either evolution."""
        m = self.config.latent_dimension
        if self.config.evolution == "direct_rate":
            if v is None:
                return self.evo.forward(x_prev)[0], None
            r, dr = self.evo.forward_jvp(x_prev, v[0])
            return r, dr
        if v is None:
            entries = self.evo.forward(x_prev)[0]
            dentries = None
        else:
            entries, dentries = self.evo.forward_jvp(x_prev, v[0])
        b = x_prev.shape[0]
        low = np.zeros((b, m, m))
        low[:, self.tril[0], self.tril[1]] = entries
        f = -beta
        p = np.einsum("bji,bj->bi", low, f)
        r = np.einsum("bij,bj->bi", low, p)
        if v is None:
            return r, p
        dbeta = v[1]
        k = dentries.shape[1]
        dlow = np.zeros((b, k, m, m))
        dlow[:, :, self.tril[0], self.tril[1]] = dentries
        df = -dbeta
        dp = np.einsum("bkji,bj->bki", dlow, f) + np.einsum("bji,bkj->bki", low, df)
        dr = np.einsum("bkij,bj->bki", dlow, p) + np.einsum("bij,bkj->bki", low, dp)
        return r, dr

    def update_batch(self, strain, latent, time_step, tangent: bool = False) -> dict:
        """One explicit update of every point (``rve_rno.py:249-313,334-418``)."""
        cfg = self.config
        strain = np.asarray(strain, np.float64)
        latent = np.asarray(latent, np.float64)
        dt = np.asarray(time_step, np.float64)
        b, m = strain.shape[0], cfg.latent_dimension
        if strain.shape != (b, 6) or latent.shape != (b, m) or dt.shape != (b,):
            raise ValueError("strain [B, 6], latent [B, m] and time step [B] are required")
        tau = cfg.reference_time_scale
        a_eq_v, a_eq_g, a_ne_v, a_ne_g = self.anchors()
        e = strain / self.scale
        x_prev = np.concatenate([e, latent], axis=1)
        eye = np.broadcast_to(np.eye(6), (b, 6, 6))
        eye_x = np.concatenate([eye, np.zeros((b, 6, m))], axis=2)
        if tangent:
            v_eq, g_eq, h_eq = self.eq.gradient_hvp(e, eye)
            v_p, g_p, h_p = self.neq.gradient_hvp(x_prev, eye_x)
        else:
            v_eq, g_eq = self.eq.value_and_gradient(e)
            v_p, g_p = self.neq.value_and_gradient(x_prev)
        equilibrium = v_eq - a_eq_v - (e * a_eq_g).sum(1)
        sigma_eq = (g_eq - a_eq_g) / self.scale
        beta = g_p[:, 6:]
        if tangent:
            r, dr = self._rate(x_prev, beta, (eye_x, h_p[:, :, 6:]))
            if cfg.evolution == "mobility_gradient":
                p = np.einsum("bji,bj->bi", self._lower(x_prev), -beta)
        else:
            r, p = self._rate(x_prev, beta)
        state_rate = r / tau
        if cfg.evolution == "direct_rate":
            dissipation = -(beta * state_rate).sum(1)
        else:
            dissipation = (p * p).sum(1) / tau
        cand = latent + (dt / tau)[:, None] * r
        x_cand = np.concatenate([e, cand], axis=1)
        previous = v_p - a_ne_v - (e * a_ne_g).sum(1)
        if tangent and cfg.integrator == "current_state_stress_forward_euler":
            v_dir = np.concatenate([eye, (dt / tau)[:, None, None] * dr], axis=2)
            v_c, g_c, h_c = self.neq.gradient_hvp(x_cand, v_dir)
        else:
            v_c, g_c = self.neq.value_and_gradient(x_cand)
        candidate = v_c - a_ne_v - (e * a_ne_g).sum(1)
        if cfg.integrator == "current_state_stress_forward_euler":
            response, sigma_ne = candidate, (g_c[:, :6] - a_ne_g) / self.scale
        else:
            response, sigma_ne = previous, (g_p[:, :6] - a_ne_g) / self.scale
        out = {
            "stress": sigma_eq + sigma_ne,
            "latent_state": cand,
            "state_rate": state_rate,
            "dimensionless_state_rate": r,
            "dissipation": dissipation,
            "free_energy": equilibrium + response,
            "thermodynamic_violation": np.maximum(-dissipation, 0.0),
            "nonequilibrium_energy_increment": candidate - previous,
        }
        if tangent:
            # d sigma_i / d e_k, k the direction index: rows of the HVPs are directions
            d_eq = h_eq[:, :, :6]
            d_ne = (h_c[:, :, :6] if cfg.integrator == "current_state_stress_forward_euler"
                    else h_p[:, :, :6])
            ds_de = np.transpose(d_eq + d_ne, (0, 2, 1)) / self.scale[None, :, None]
            out["tangent"] = ds_de / self.scale[None, None, :]
        return out

    def _lower(self, x_prev):
        m = self.config.latent_dimension
        entries = self.evo.forward(x_prev)[0]
        low = np.zeros((x_prev.shape[0], m, m))
        low[:, self.tril[0], self.tril[1]] = entries
        return low

    def rollout(self, strains, time_steps, latent0=None) -> dict:
        """Sequential updates of one batch over a path: strains [T, B, 6], time steps [T, B]."""
        strains = np.asarray(strains, np.float64)
        t_n, b = strains.shape[:2]
        z = (np.zeros((b, self.config.latent_dimension)) if latent0 is None
             else np.asarray(latent0, np.float64))
        rows: dict = {}
        for n in range(t_n):
            u = self.update_batch(strains[n], z, np.asarray(time_steps, np.float64)[n])
            for k, v in u.items():
                rows.setdefault(k, []).append(v)
            z = u["latent_state"]
        return {k: np.stack(v) for k, v in rows.items()}
