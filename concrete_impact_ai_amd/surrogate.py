"""Recurrent MLP material-point surrogate with an error-gated FE² fallback.

Role in the reference: the recurrent material-point network behind the material
protocol (``EnergyDissipationRVERNO``, ``src/concrete_impact/nn/models/rve_rno.py:162-476``;
``RVERNOMaterialAdapter``, ``nn/deployment.py:68-165``), trained by back-propagation
through time on RVE responses (``nn/training/rve_rno.py:113-415``).  This network is
the build's own (13 -> 256 -> 256 -> 256 -> 256 -> 12, GELU):

    x_t = [d_eps_t / eps_scale (6), log10(rate_t / rate_ref) / rate_scale (1), z_t (6)]
    o_t = MLP(x_t)
    sigma_t = sigma_{t-1} + c stress_scale o_t[0:6],   z_{t+1} = z_t + o_t[6:12]

with c = STRESS_STEP and sigma_{-1} = z_0 = 0: the network predicts stress increments, so
the six internal state variables carry only the inelastic state (the elastic part of
an increment needs none); predicting the stress itself from the latent needs the
latent to integrate the total strain and measured 1.6x the held-out error.

Every matrix product runs in the HIP kernels of ``libcia_hip.so`` on BF16 matrix
cores with FP32 accumulation: layer-wise GEMMs with fused bias/GELU epilogues for
training (forward and backward through time, GELU' fused into the backward-data
GEMM, split-K weight gradients, Adam), and one fused all-layer kernel for inference.

The gate: a point whose inputs leave the envelope seen in training (strain
increment, rate feature) or whose latent state leaves the training range is handed
to the FE² adapter (``RVEHomogenizedMaterial``) for the rest of the run.  Its RVE
starts from the virgin state and is driven to the point's current strain in
``fallback_substeps`` proportional increments (the history before the trip is not
replayed).
"""

from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np
import torch

from concrete_impact_ai_amd import _lib
from concrete_impact_ai_amd.materials import (
    DEFAULT_RESPONSE_REQUIREMENTS,
    DeviceResponse,
    MaterialPointResponse,
    MaterialState,
    device,
    stream_handle,
)

SIZES = (13, 256, 256, 256, 256, 12)
N_LAYERS = len(SIZES) - 1
STRESS_STEP = 0.1  # scale of the network's stress-increment outputs


def _p(t):
    return C.c_void_p(None) if t is None else C.c_void_p(t.data_ptr())


def _check(status, where):
    _lib.check(status, where)


def _bf16(x: torch.Tensor) -> torch.Tensor:
    """fp32 -> bf16 by the library kernel (round to nearest even)."""
    x = x.contiguous()
    y = torch.empty(x.shape, dtype=torch.bfloat16, device=x.device)
    _check(_lib.lib().cia_f32_to_bf16(stream_handle(), x.numel(), _p(x), _p(y)), "cia_f32_to_bf16")
    return y


class SurrogateParams:
    """Weights and biases as two contiguous fp32 blobs (W1 [256][13] | W2..W4 [256][256] |
    W5 [12][256]; b1..b4 [256] | b5 [12]) with a bf16 copy of the weights, Adam moments
    and gradient blobs of the same layout."""

    def __init__(self, layers=None, seed: int = 0):
        if layers is None:
            rng = np.random.default_rng(seed)
            layers = []
            for li in range(N_LAYERS):
                fi, fo = SIZES[li], SIZES[li + 1]
                w = rng.standard_normal((fo, fi)) * math.sqrt(2.0 / fi)
                if li == N_LAYERS - 1:
                    w *= 0.1
                layers.append((w, np.zeros(fo)))
        dev = device()
        self.w_shapes = [(SIZES[i + 1], SIZES[i]) for i in range(N_LAYERS)]
        self.w_offsets = np.concatenate([[0], np.cumsum([a * b for a, b in self.w_shapes])])
        self.b_offsets = np.concatenate([[0], np.cumsum(SIZES[1:])])
        self.w = torch.as_tensor(np.concatenate([np.asarray(w, np.float64).ravel() for w, _ in layers]),
                                 dtype=torch.float32, device=dev).contiguous()
        self.b = torch.as_tensor(np.concatenate([np.asarray(b, np.float64).ravel() for _, b in layers]),
                                 dtype=torch.float32, device=dev).contiguous()
        self.w_bf16 = _bf16(self.w)
        self.m_w = torch.zeros_like(self.w)
        self.v_w = torch.zeros_like(self.w)
        self.m_b = torch.zeros_like(self.b)
        self.v_b = torch.zeros_like(self.b)
        self.step = 0

    def W(self, li, bf16=True):
        src = self.w_bf16 if bf16 else self.w
        o0, o1 = int(self.w_offsets[li]), int(self.w_offsets[li + 1])
        return src[o0:o1].view(self.w_shapes[li])

    def B(self, li):
        return self.b[int(self.b_offsets[li]):int(self.b_offsets[li + 1])]

    def layers_numpy(self):
        w = self.w.double().cpu().numpy()
        b = self.b.double().cpu().numpy()
        return [(w[int(self.w_offsets[i]):int(self.w_offsets[i + 1])].reshape(self.w_shapes[i]),
                 b[int(self.b_offsets[i]):int(self.b_offsets[i + 1])]) for i in range(N_LAYERS)]

    def refresh_bf16(self):
        self.w_bf16 = _bf16(self.w)

    def adam(self, gw, gb, lr, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0):
        self.step += 1
        lib = _lib.lib()
        st = stream_handle()
        _check(lib.cia_mlp_adam_step(st, self.w.numel(), _p(self.w), _p(gw.contiguous()), _p(self.m_w),
                                     _p(self.v_w), _p(self.w_bf16), lr, beta1, beta2, eps, self.step,
                                     weight_decay), "cia_mlp_adam_step")
        _check(lib.cia_mlp_adam_step(st, self.b.numel(), _p(self.b), _p(gb.contiguous()), _p(self.m_b),
                                     _p(self.v_b), _p(None), lr, beta1, beta2, eps, self.step, 0.0),
               "cia_mlp_adam_step")

    def state_dict(self):
        return {"w": self.w.cpu().numpy().copy(), "b": self.b.cpu().numpy().copy()}

    @classmethod
    def from_state_dict(cls, d):
        p = cls(seed=0)
        p.w.copy_(torch.as_tensor(d["w"], device=p.w.device))
        p.b.copy_(torch.as_tensor(d["b"], device=p.b.device))
        p.refresh_bf16()
        return p


class _Workspace:
    def __init__(self):
        self.buf = None

    def get(self, n):
        if self.buf is None or self.buf.numel() < n:
            self.buf = torch.empty(n, dtype=torch.float32, device=device())
        return self.buf


_WS = _Workspace()


class MLPFunction(torch.autograd.Function):
    """One network evaluation y [M][12] = MLP(x [M][13]) on the HIP kernels; backward
    returns d x and the gradients of the weight / bias blobs."""

    @staticmethod
    def forward(ctx, x, w_blob, b_blob, params: SurrogateParams):
        lib = _lib.lib()
        st = stream_handle()
        M = x.shape[0]
        dev = x.device
        acts = [_bf16(x)]
        zs = []
        for li in range(N_LAYERS - 1):
            N, K = params.w_shapes[li]
            z = torch.empty((M, N), dtype=torch.float32, device=dev)
            a = torch.empty((M, N), dtype=torch.bfloat16, device=dev)
            _check(lib.cia_mlp_linear_forward(st, M, N, K, _p(acts[-1]), _p(params.W(li)),
                                              _p(params.B(li)), 1, _p(z), _p(None), _p(a)),
                   "cia_mlp_linear_forward")
            zs.append(z)
            acts.append(a)
        N, K = params.w_shapes[-1]
        y = torch.empty((M, N), dtype=torch.float32, device=dev)
        _check(lib.cia_mlp_linear_forward(st, M, N, K, _p(acts[-1]), _p(params.W(N_LAYERS - 1)),
                                          _p(params.B(N_LAYERS - 1)), 0, _p(None), _p(y), _p(None)),
               "cia_mlp_linear_forward")
        ctx.params = params
        ctx.acts = acts
        ctx.zs = zs
        return y

    @staticmethod
    def backward(ctx, gy):
        params = ctx.params
        lib = _lib.lib()
        st = stream_handle()
        M = gy.shape[0]
        dev = gy.device
        gw = torch.zeros_like(params.w)
        gb = torch.zeros_like(params.b)
        ws_n = lib.cia_mlp_backward_weight_workspace(256, 256)
        ws = _WS.get(ws_n)
        g = _bf16(gy)
        for li in range(N_LAYERS - 1, -1, -1):
            N, K = params.w_shapes[li]
            o0, o1 = int(params.w_offsets[li]), int(params.w_offsets[li + 1])
            b0, b1 = int(params.b_offsets[li]), int(params.b_offsets[li + 1])
            _check(lib.cia_mlp_linear_backward_weight(st, M, N, K, _p(g), _p(ctx.acts[li]),
                                                      _p(gw[o0:o1]), _p(gb[b0:b1]), _p(ws), ws_n, 0),
                   "cia_mlp_linear_backward_weight")
            if li > 0:
                gnext = torch.empty((M, K), dtype=torch.bfloat16, device=dev)
                _check(lib.cia_mlp_linear_backward_data(st, M, N, K, _p(g), _p(params.W(li)),
                                                        _p(ctx.zs[li - 1]), _p(None), _p(gnext)),
                       "cia_mlp_linear_backward_data")
                g = gnext
            else:
                gx = torch.empty((M, K), dtype=torch.float32, device=dev)
                _check(lib.cia_mlp_linear_backward_data(st, M, N, K, _p(g), _p(params.W(0)),
                                                        _p(None), _p(gx), _p(None)),
                       "cia_mlp_linear_backward_data")
        return gx, gw, gb, None


def mlp(x, params: SurrogateParams, w_leaf=None, b_leaf=None):
    return MLPFunction.apply(x.contiguous(), params.w if w_leaf is None else w_leaf,
                             params.b if b_leaf is None else b_leaf, params)


def forward_fused(x: torch.Tensor, params: SurrogateParams) -> torch.Tensor:
    """Inference: all five layers in one kernel; x [M][13] fp32 -> [M][12] fp32."""
    x = x.contiguous().float()
    y = torch.empty((x.shape[0], 12), dtype=torch.float32, device=x.device)
    _check(_lib.lib().cia_mlp_forward_fused(stream_handle(), x.shape[0], _p(x), _p(params.w_bf16),
                                            _p(params.b), _p(y)), "cia_mlp_forward_fused")
    return y


@dataclass
class SurrogateScales:
    """Feature / target scaling and the training envelope used by the gate."""

    eps_scale: float
    rate_ref: float
    rate_scale: float
    stress_scale: float
    d_eps_max: float = 1.0        # max |d_eps| / eps_scale seen in training
    rate_lo: float = -1.0         # rate feature range seen in training
    rate_hi: float = 1.0
    latent_max: float = 1.0       # max |z| seen in training rollouts

    @classmethod
    def fit(cls, d_eps, rate, stress):
        eps_scale = float(np.max(np.abs(d_eps))) or 1.0
        lr = np.log10(np.maximum(rate, 1e-30))
        ref = float(np.median(lr))
        rs = float(np.max(np.abs(lr - ref))) or 1.0
        return cls(eps_scale, 10.0 ** ref, rs, float(np.max(np.abs(stress))) or 1.0, 1.0,
                   float(np.min((lr - ref) / rs)), float(np.max((lr - ref) / rs)), 1.0)

    def rate_feature(self, rate):
        if isinstance(rate, torch.Tensor):
            return torch.log10(torch.clamp(rate, min=1e-30) / self.rate_ref) / self.rate_scale
        return np.log10(np.maximum(rate, 1e-30) / self.rate_ref) / self.rate_scale


def rollout(params: SurrogateParams, scales: SurrogateScales, d_eps: torch.Tensor,
            rate: torch.Tensor, fused: bool = False, w_leaf=None, b_leaf=None):
    """d_eps [B][T][6] (physical), rate [B][T] -> stress [B][T][6] (physical), latents."""
    B, T, _ = d_eps.shape
    z = torch.zeros((B, 6), dtype=torch.float32, device=d_eps.device)
    s = torch.zeros((B, 6), dtype=torch.float32, device=d_eps.device)
    feats_e = (d_eps / scales.eps_scale).float()
    feats_r = scales.rate_feature(rate).float()
    out = []
    lat = [z]
    for t in range(T):
        x = torch.cat([feats_e[:, t], feats_r[:, t, None], z], dim=1)
        o = forward_fused(x, params) if fused else mlp(x, params, w_leaf, b_leaf)
        s = s + STRESS_STEP * o[:, :6]
        out.append(s)
        z = z + o[:, 6:]
        lat.append(z)
    return torch.stack(out, dim=1) * scales.stress_scale, torch.stack(lat, dim=1)


def train(params: SurrogateParams, scales: SurrogateScales, d_eps, rate, stress, iterations: int,
          batch_paths: int, window: int, lr: float = 1e-3, seed: int = 0, lr_final: float = 1e-5,
          log_every: int = 0, progress=None):
    """BPTT on random windows of the training paths (``train_rve_rno``,
    ``nn/training/rve_rno.py:200-415``).  The loss is the path-normalised stress MSE
    of ``_batch_loss`` (``:540-638``): each path's squared errors are divided by its
    peak squared stress, so low-amplitude paths weigh as much as high ones.  Windows
    start at step 0 so the latent starts from zero as in deployment.  Cosine
    learning-rate decay from ``lr`` to ``lr_final``."""
    dev = device()
    de = torch.as_tensor(d_eps, dtype=torch.float32, device=dev)
    rt = torch.as_tensor(rate, dtype=torch.float32, device=dev)
    sg = torch.as_tensor(stress, dtype=torch.float32, device=dev) / scales.stress_scale
    P, T, _ = de.shape
    path_w = 1.0 / torch.clamp(sg.abs().amax(dim=(1, 2)), min=1e-3) ** 2   # [P]
    rng = np.random.default_rng(seed)
    w_leaf = params.w.detach().requires_grad_(True)
    b_leaf = params.b.detach().requires_grad_(True)
    history = []
    for it in range(iterations):
        idx = torch.as_tensor(rng.choice(P, size=min(batch_paths, P), replace=P < batch_paths),
                              device=dev)
        L = min(window, T)
        pred, _ = rollout(params, SurrogateScales(scales.eps_scale, scales.rate_ref, scales.rate_scale, 1.0),
                          de[idx, :L], rt[idx, :L], False, w_leaf, b_leaf)
        loss = torch.mean(((pred - sg[idx, :L]) ** 2).mean(dim=(1, 2)) * path_w[idx])
        w_leaf.grad = None
        b_leaf.grad = None
        loss.backward()
        cur = lr_final + 0.5 * (lr - lr_final) * (1.0 + math.cos(math.pi * it / max(iterations - 1, 1)))
        params.adam(w_leaf.grad, b_leaf.grad, cur)
        if log_every and (it % log_every == 0 or it == iterations - 1):
            history.append((it, float(loss.detach())))
            if progress is not None:
                print(f"surrogate training: iteration {it} loss {history[-1][1]:.3e}", file=progress,
                      flush=True)
    return history


def envelope(params, scales, d_eps, rate):
    """Set the gate's envelope from the training inputs and the latents of their rollouts."""
    dev = device()
    de = torch.as_tensor(d_eps, dtype=torch.float32, device=dev)
    rt = torch.as_tensor(rate, dtype=torch.float32, device=dev)
    _, lat = rollout(params, scales, de, rt, fused=True)
    rf = scales.rate_feature(np.asarray(rate))
    scales.d_eps_max = float(np.max(np.abs(d_eps)) / scales.eps_scale)
    scales.rate_lo = float(np.min(rf))
    scales.rate_hi = float(np.max(rf))
    scales.latent_max = float(lat.abs().max())
    return scales


# ---------------------------------------------------------------------------
# MaterialModel with the gated FE² fallback


class SurrogateMaterial:
    """Material-point protocol around the trained surrogate.  State per point: the
    committed strain (6), the latent (6) and a fallback flag; points on the FE²
    fallback carry their RVE states in the adapter's batch."""

    def __init__(self, name, params: SurrogateParams, scales: SurrogateScales, density: float,
                 maximum_wave_speed: float, fallback=None, fallback_substeps: int = 4,
                 gate_margin: float = 1.25):
        self.name = name
        self.params = params
        self.scales = scales
        self.density = float(density)
        self.maximum_wave_speed = float(maximum_wave_speed)
        self.fallback = fallback
        self.fallback_substeps = int(fallback_substeps)
        self.gate_margin = float(gate_margin)
        self.stats = {"surrogate_points": 0, "fallback_points": 0, "gate_trips": 0}

    def initialize_state(self, n_points: int) -> MaterialState:
        return MaterialState({"surrogate_strain": np.zeros((n_points, 6)),
                              "surrogate_stress": np.zeros((n_points, 6)),
                              "surrogate_latent": np.zeros((n_points, 6)),
                              "surrogate_fallback": np.zeros(n_points)})

    def device_kernel(self, point_phase_ids=None):
        return _SurrogateKernel(self)

    def update(self, request, state, requirements=DEFAULT_RESPONSE_REQUIREMENTS):
        if requirements.tangent:
            raise ValueError("The surrogate material provides no consistent tangent.")
        dev = device()
        k = _SurrogateKernel(self)
        strain = torch.as_tensor(np.ascontiguousarray(np.asarray(request.strains).T), device=dev)
        rate = (None if request.strain_rates is None else
                torch.as_tensor(np.ascontiguousarray(np.asarray(request.strain_rates).T), device=dev))
        st = k.from_host(state)
        r = k.update(strain, rate, st, request.time_step, request.update_settings, False)
        return MaterialPointResponse(stresses=r.stress.cpu().numpy().T.copy(),
                                     state=k.material_state(r.state),
                                     free_energy=None, dissipation=None, diagnostics={})


class _SurrogateState:
    def __init__(self, strain, stress, latent, fallback_mask, rve_state=None, rve_index=None):
        self.strain = strain            # [n][6] fp64 committed strain
        self.stress = stress            # [n][6] fp64 committed stress
        self.latent = latent            # [n][6] fp32
        self.fallback = fallback_mask   # [n] bool
        self.rve_state = rve_state      # adapter batch state of the fallback points
        self.rve_index = rve_index      # np.ndarray of point ids in rve_state order


class _SurrogateKernel:
    def __init__(self, material: SurrogateMaterial):
        self.m = material

    def initial_state(self, n):
        dev = device()
        return _SurrogateState(torch.zeros((n, 6), dtype=torch.float64, device=dev),
                               torch.zeros((n, 6), dtype=torch.float64, device=dev),
                               torch.zeros((n, 6), dtype=torch.float32, device=dev),
                               torch.zeros(n, dtype=torch.bool, device=dev))

    def from_host(self, state):
        dev = device()
        v = state.variables
        return _SurrogateState(torch.as_tensor(v["surrogate_strain"], device=dev),
                               torch.as_tensor(v["surrogate_stress"], dtype=torch.float64, device=dev),
                               torch.as_tensor(v["surrogate_latent"], dtype=torch.float32, device=dev),
                               torch.as_tensor(np.asarray(v["surrogate_fallback"]) > 0.5, device=dev))

    def material_state(self, st):
        return MaterialState({"surrogate_strain": st.strain.cpu().numpy(),
                              "surrogate_stress": st.stress.cpu().numpy(),
                              "surrogate_latent": st.latent.double().cpu().numpy(),
                              "surrogate_fallback": st.fallback.double().cpu().numpy()})

    def update(self, strain, rate, st: _SurrogateState, dt, settings, tangent=False) -> DeviceResponse:
        if tangent:
            raise ValueError("The surrogate material provides no consistent tangent.")
        m = self.m
        sc = m.scales
        eps = strain.t().contiguous().double()             # [n][6]
        d_eps = eps - st.strain
        if rate is not None:
            rmag = torch.linalg.norm(rate.t().double(), dim=1)
        else:
            rmag = torch.linalg.norm(d_eps, dim=1) / dt
        feat_r = sc.rate_feature(rmag).float()
        x = torch.cat([(d_eps / sc.eps_scale).float(), feat_r[:, None], st.latent], dim=1)
        o = forward_fused(x, m.params)
        sig = st.stress + STRESS_STEP * sc.stress_scale * o[:, :6].double()
        lat = st.latent + o[:, 6:]
        trip = ((d_eps.abs().max(dim=1).values / sc.eps_scale > m.gate_margin * sc.d_eps_max)
                | (feat_r < sc.rate_lo - 0.5) | (feat_r > sc.rate_hi + 0.5)
                | (lat.abs().max(dim=1).values > m.gate_margin * sc.latent_max)
                | ~torch.isfinite(lat).all(dim=1))
        fb = st.fallback | trip
        rve_state, rve_index = st.rve_state, st.rve_index
        n_fb = int(fb.sum())
        new_trips = int((trip & ~st.fallback).sum())
        m.stats["gate_trips"] += new_trips
        m.stats["fallback_points"] += n_fb
        m.stats["surrogate_points"] += int(fb.numel() - n_fb)
        if n_fb:
            if m.fallback is None:
                raise RuntimeError(f"{n_fb} surrogate points left the training envelope and no "
                                   "FE² fallback is configured.")
            sig, rve_state, rve_index = self._fallback(eps, st, fb, sig, dt, settings)
        return DeviceResponse(stress=sig.t().contiguous(),
                              state=_SurrogateState(eps.clone(), sig, lat, fb, rve_state, rve_index))

    def _fallback(self, eps, st, fb, sig, dt, settings):
        """FE² for the points on the fallback: earlier trips continue from their
        committed RVE states; new trips start a virgin RVE driven to the current
        strain in proportional sub-increments."""
        from concrete_impact_ai_amd.fe2 import RVEBatchState
        from concrete_impact_ai_amd.materials import MaterialPointRequest

        ad = self.m.fallback
        ids = torch.nonzero(fb).flatten().cpu().numpy()
        old = st.rve_index if st.rve_index is not None else np.zeros(0, dtype=np.int64)
        new = np.setdiff1d(ids, old)
        eps_h = eps.cpu().numpy()
        sig = sig.clone()
        states, order = [], []
        if old.size:
            r = ad.update(MaterialPointRequest(eps_h[old], None, dt, "three_dimensional", settings),
                          st.rve_state)
            sig[torch.as_tensor(old, device=sig.device)] = torch.as_tensor(r.stresses, device=sig.device)
            states.append(r.state)
            order.append(old)
        if new.size:
            ns = ad.initialize_state(new.size)
            S = max(1, self.m.fallback_substeps)
            for k in range(1, S + 1):
                r = ad.update(MaterialPointRequest(eps_h[new] * (k / S), None, dt / S,
                                                   "three_dimensional", settings), ns)
                ns = r.state
            sig[torch.as_tensor(new, device=sig.device)] = torch.as_tensor(r.stresses, device=sig.device)
            states.append(ns)
            order.append(new)
        if len(states) == 1:
            batch = states[0]
        else:
            batch = RVEBatchState(torch.cat([s_.packed for s_ in states]).contiguous(),
                                  torch.cat([s_.w for s_ in states]).contiguous(),
                                  np.concatenate([s_.macro for s_ in states]), ad.operator_cache)
        return sig, batch, np.concatenate(order).astype(np.int64)
