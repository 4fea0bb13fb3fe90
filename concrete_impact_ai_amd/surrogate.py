"""Recurrent MLP material-point surrogate with an error-gated FE² fallback.

Role in the reference: the recurrent material-point network behind the material
protocol (``EnergyDissipationRVERNO``, ``src/concrete_impact/nn/models/rve_rno.py:162-476``;
``RVERNOMaterialAdapter``, ``nn/deployment.py:68-165``), trained by back-propagation
through time on RVE responses (``nn/training/rve_rno.py:113-415``).  This network is
the build's own (13 -> 256 -> 256 -> 256 -> 256 -> 12, GELU):

    x_t = [d_eps_t / eps_scale (6), log10(rate_t / rate_ref) / rate_scale (1), tanh(z_t) (6)]
    o_t = MLP(x_t)
    ein_t = ein_{t-1} + c eps_scale o_t[0:6],   z_{t+1} = z_t + o_t[6:12]
    sigma_t = C (eps_t - ein_t)

with c = INELASTIC_STEP, ein_{-1} = z_0 = 0 (the network sees the latent through tanh, a
bounded input, so that an untrained recurrence cannot feed back on itself without
bound over long paths) and C the homogenised elastic stiffness of
the RVE (its effective tangent at a small strain): the network predicts inelastic
strain increments, so an elastic increment, loading or unloading, needs no network
output, and the six internal state variables carry only the inelastic state.  On
macro impact strain paths this structure measured 0.7x the stress error of
predicting stress increments directly (same data and training budget); on held-out
training paths stress increments in turn measured 0.6x the error of predicting the
stress from the latent.

Every matrix product runs in the HIP kernels of ``libcia_hip.so`` on BF16 matrix
cores with FP32 accumulation: layer-wise GEMMs with fused bias/GELU epilogues for
training (forward and backward through time, GELU' fused into the backward-data
GEMM, split-K weight gradients, Adam), and one fused all-layer kernel for inference.

As a material (``SurrogateMaterial``) every point keeps an anchor, the state of its
last committed network step: each call evaluates the network once, in one fused
kernel, from the anchor with the whole increment since it, and commits only when
that increment reaches the size of a training step, so the network sees training-
sized increments whatever the macro time step.  The gate: a point whose inputs leave
the envelope seen in training (strain increment, rate above the trained range) or
whose latent state leaves the training range, and a monitor point whose stress
differs from its own FE² solution by more than the error bound (the error gate
judged by FE²), is handed to the FE² adapter (``RVEHomogenizedMaterial``) for the
rest of the run; see ``SurrogateMaterial`` for how its RVE state is formed.
"""

from __future__ import annotations

import ctypes as C
import math
import time
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
INELASTIC_STEP = 0.3  # scale of the network's inelastic-strain-increment outputs


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
    d_eps_commit: float = 0.0     # median |d_eps|_inf of a training step (physical strain)
    stiffness: tuple = ()         # homogenised elastic stiffness, 36 values row-major (Pa)

    @classmethod
    def fit(cls, d_eps, rate, stress, stiffness, rate_floor: float = 1.0):
        """Scales from training data and the RVE's homogenised elastic stiffness
        (6 x 6, engineering Voigt); rates below ``rate_floor`` (1/s, the
        quasi-static end: increments at the ends of smooth path segments and holds)
        count as ``rate_floor``, which becomes the lower clamp of the rate feature."""
        eps_scale = float(np.max(np.abs(d_eps))) or 1.0
        lr = np.log10(np.maximum(rate, rate_floor))
        ref = float(np.median(lr))
        rs = float(np.max(np.abs(lr - ref))) or 1.0
        return cls(eps_scale, 10.0 ** ref, rs, float(np.max(np.abs(stress))) or 1.0, 1.0,
                   float(np.min((lr - ref) / rs)), float(np.max((lr - ref) / rs)), 1.0,
                   float(np.median(np.max(np.abs(d_eps), axis=-1))),
                   tuple(float(v) for v in np.asarray(stiffness, dtype=np.float64).ravel()))

    def stiffness_matrix(self) -> np.ndarray:
        if len(self.stiffness) != 36:
            raise ValueError("the surrogate scales carry no 6 x 6 homogenised stiffness")
        return np.asarray(self.stiffness, dtype=np.float64).reshape(6, 6)

    def rate_feature(self, rate):
        """log10(rate / rate_ref) / rate_scale, clamped below at rate_lo (as the
        material kernel clamps it)."""
        if isinstance(rate, torch.Tensor):
            f = torch.log10(torch.clamp(rate, min=1e-30) / self.rate_ref) / self.rate_scale
            return torch.clamp(f, min=self.rate_lo)
        return np.maximum(np.log10(np.maximum(rate, 1e-30) / self.rate_ref) / self.rate_scale, self.rate_lo)


def rollout(params: SurrogateParams, scales: SurrogateScales, d_eps: torch.Tensor,
            rate: torch.Tensor, fused: bool = False, w_leaf=None, b_leaf=None,
            normalized: bool = False):
    """d_eps [B][T][6] (physical, from zero strain), rate [B][T] -> stress [B][T][6]
    (physical, or divided by stress_scale when ``normalized``), latents."""
    B, T, _ = d_eps.shape
    dev = d_eps.device
    z = torch.zeros((B, 6), dtype=torch.float32, device=dev)
    ein = torch.zeros((B, 6), dtype=torch.float32, device=dev)
    feats_e = (d_eps / scales.eps_scale).float()
    eps = torch.cumsum(feats_e, dim=1)                     # strain / eps_scale
    feats_r = scales.rate_feature(rate).float()
    Cs = torch.as_tensor(scales.stiffness_matrix() * (scales.eps_scale / scales.stress_scale),
                         dtype=torch.float32, device=dev)
    out = []
    lat = [z]
    for t in range(T):
        x = torch.cat([feats_e[:, t], feats_r[:, t, None], torch.tanh(z)], dim=1)
        o = forward_fused(x, params) if fused else mlp(x, params, w_leaf, b_leaf)
        ein = ein + INELASTIC_STEP * o[:, :6]
        out.append((eps[:, t] - ein) @ Cs.T)
        z = z + o[:, 6:]
        lat.append(z)
    sig = torch.stack(out, dim=1)
    return (sig if normalized else sig * scales.stress_scale), torch.stack(lat, dim=1)


class BpttData:
    """A training set on the device in the layout of ``cia_surrogate_bptt``:
    features [P][T][8] (d_eps / eps_scale, rate feature, 0), cumulative strain / eps_scale
    [P][T][6] and target stress / stress_scale [P][T][6], fp32."""

    def __init__(self, features, strain, target, stiffness_scaled):
        dev = device()
        self.feat = torch.as_tensor(np.asarray(features), dtype=torch.float32, device=dev).contiguous()
        self.strain = torch.as_tensor(np.asarray(strain), dtype=torch.float32, device=dev).contiguous()
        self.target = torch.as_tensor(np.asarray(target), dtype=torch.float32, device=dev).contiguous()
        self.P, self.T = int(self.feat.shape[0]), int(self.feat.shape[1])
        assert self.feat.shape == (self.P, self.T, 8)
        assert self.strain.shape == self.target.shape == (self.P, self.T, 6)
        cs = np.asarray(stiffness_scaled, dtype=np.float64).reshape(36)
        self.cs = (C.c_float * 36)(*[float(v) for v in cs])

    @classmethod
    def from_paths(cls, scales: SurrogateScales, d_eps, rate, stress):
        """Features, strains and targets of physical paths (d_eps [P][T][6], rate [P][T],
        stress [P][T][6]) in the scaled units of ``scales``."""
        d_eps = np.asarray(d_eps, dtype=np.float64)
        P, T, _ = d_eps.shape
        feat = np.zeros((P, T, 8), dtype=np.float32)
        feat[..., :6] = d_eps / scales.eps_scale
        feat[..., 6] = scales.rate_feature(np.asarray(rate, dtype=np.float64))
        strain = np.cumsum(d_eps / scales.eps_scale, axis=1)
        target = np.asarray(stress, dtype=np.float64) / scales.stress_scale
        return cls(feat, strain, target,
                   scales.stiffness_matrix() * (scales.eps_scale / scales.stress_scale))

    def path_weights(self, window: int) -> torch.Tensor:
        """1 / (peak |target| over the first ``window`` steps)^2 per path (floor 1e-3):
        the path normalisation of the loss over the steps it is trained on."""
        peak = self.target[:, :window].abs().amax(dim=(1, 2))
        return (1.0 / torch.clamp(peak, min=1e-3) ** 2).contiguous()


class BpttStep:
    """Loss and gradients of one batch through ``cia_surrogate_bptt`` (workspace,
    gradient blobs and the device loss held across calls)."""

    def __init__(self, params: SurrogateParams, paths: int, window: int):
        lib = _lib.lib()
        dev = device()
        self.paths, self.window = int(paths), int(window)
        self.ws_bytes = int(lib.cia_surrogate_bptt_workspace(self.paths, self.window))
        self.ws = torch.empty(self.ws_bytes, dtype=torch.uint8, device=dev)
        self.gw = torch.zeros_like(params.w)
        self.gb = torch.zeros_like(params.b)
        self.loss = torch.zeros(1, dtype=torch.float32, device=dev)
        self.grad_norm = torch.zeros(1, dtype=torch.float32, device=dev)
        self.max_grad_norm = 0.0   # > 0: clip the global gradient norm (see cia_surrogate_bptt)

    def __call__(self, params: SurrogateParams, data: BpttData, idx: torch.Tensor,
                 path_w: torch.Tensor, window_start: int = 0, state_in: torch.Tensor | None = None,
                 state_out: torch.Tensor | None = None):
        """Loss (device, 1 float) of the batch over steps window_start ..
        window_start + window - 1; gradients in ``gw`` / ``gb``.  ``state_in`` /
        ``state_out``: [paths][12] fp32 recurrence state (latent, inelastic strain) at
        the window start / end (None = zeros / not wanted)."""
        assert idx.dtype == torch.int32 and idx.numel() == self.paths and idx.is_contiguous()
        for st in (state_in, state_out):
            assert st is None or (st.dtype == torch.float32 and st.shape == (self.paths, 12)
                                  and st.is_contiguous())
        _check(_lib.lib().cia_surrogate_bptt(
            stream_handle(), data.P, data.T, self.paths, int(window_start), self.window, _p(idx),
            _p(data.feat), _p(data.strain), _p(data.target), _p(path_w), data.cs, INELASTIC_STEP,
            _p(params.w), _p(params.b), _p(self.gw), _p(self.gb), _p(self.loss),
            _p(self.grad_norm), float(self.max_grad_norm), _p(state_in), _p(state_out),
            _p(self.ws), self.ws_bytes), "cia_surrogate_bptt")
        return self.loss


def train(params: SurrogateParams, scales: SurrogateScales, d_eps, rate, stress, iterations: int,
          batch_paths: int, window: int, lr: float = 1e-3, seed: int = 0, lr_final: float = 1e-5,
          log_every: int = 0, progress=None, data: BpttData | None = None, carry: bool = False,
          first_window_iterations: int = 0, max_grad_norm: float = 1.0):
    """BPTT on the training paths (``train_rve_rno``, ``nn/training/rve_rno.py:200-415``).
    The loss is the path-normalised stress MSE of ``_batch_loss`` (``:540-638``): each
    path's squared errors are divided by its peak squared stress (over the trained
    window, or over the whole path when ``carry``), so low-amplitude paths weigh as much
    as high ones.  Cosine learning-rate decay from ``lr`` to ``lr_final``.  Every
    iteration is one ``cia_surrogate_bptt`` call (forward and backward through time,
    weight gradients) and one Adam step; batch indices are drawn on the host.

    ``carry=False``: every iteration trains the first ``window`` steps of a new batch
    from the zero state (as in deployment).  ``carry=True``: truncated BPTT over whole
    paths -- a batch is trained window by window (steps 0 .. window - 1, window .. 2
    window - 1, ...), the recurrence state carried from one window to the next and the
    gradient truncated at window starts; the first ``first_window_iterations``
    iterations train first windows only (an untrained recurrence drifts over later
    windows).

    Every step's global gradient norm is clipped to ``max_grad_norm`` (0 disables), and
    a step whose gradient is not finite is skipped (its gradient zeroed); the number of
    skipped steps is returned as the ``skipped_steps`` attribute of the history list --
    the reference raises instead (``_require_finite_gradients``, ``:1040-1067``)."""
    if data is None:
        data = BpttData.from_paths(scales, d_eps, rate, stress)
    P, T = data.P, data.T
    L = min(window, T)
    B = min(batch_paths, P)
    path_w = data.path_weights(T if carry else L)
    step = BpttStep(params, B, L)
    tail = T - (T // L) * L
    step_tail = BpttStep(params, B, tail) if carry and tail else None
    for st in (step, step_tail):
        if st is not None:
            st.max_grad_norm = max_grad_norm
    skipped = torch.zeros(1, dtype=torch.int32, device=device())
    rng = np.random.default_rng(seed)
    dev = device()
    states = [torch.zeros((B, 12), dtype=torch.float32, device=dev) for _ in range(2)]
    history = []
    idx = None
    s0 = 0
    for it in range(iterations):
        if not carry or it < first_window_iterations or s0 >= T:
            s0 = 0
        if s0 == 0:
            draw = rng.choice(P, size=B, replace=False)   # B <= P: distinct paths
            idx = torch.as_tensor(draw, dtype=torch.int32, device=dev)
        st = step if s0 + L <= T else step_tail
        if carry:
            loss = st(params, data, idx, path_w, s0, None if s0 == 0 else states[0], states[1])
            states.reverse()
            s0 += st.window
        else:
            loss = st(params, data, idx, path_w)
        cur = lr_final + 0.5 * (lr - lr_final) * (1.0 + math.cos(math.pi * it / max(iterations - 1, 1)))
        params.adam(st.gw, st.gb, cur)
        skipped += (~torch.isfinite(st.grad_norm)).int()
        if log_every and (it % log_every == 0 or it == iterations - 1):
            history.append((it, float(loss)))
            if progress is not None:
                print(f"surrogate training: iteration {it} loss {history[-1][1]:.3e}", file=progress,
                      flush=True)
    history = _History(history)
    history.skipped_steps = int(skipped)
    return history


class _History(list):
    """Training history [(iteration, loss)] with the count of skipped steps."""

    skipped_steps = 0


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
    scales.d_eps_commit = float(np.median(np.max(np.abs(d_eps), axis=-1)))
    return scales


# ---------------------------------------------------------------------------
# MaterialModel with the gated FE² fallback


class SurrogateMaterial:
    """Material-point protocol around the trained surrogate (the role of
    ``RVERNOMaterialAdapter``, ``nn/deployment.py:68-165``), one
    ``cia_surrogate_update`` kernel per call.

    Every point keeps an anchor (strain, inelastic strain, latent, age): each call
    evaluates the network once from the anchor with the whole strain increment since
    it and the mean rate over that time, and the anchor moves only when the increment
    reaches ``commit_fraction`` times the median training increment (or its age
    reaches ``commit_time`` > 0).  The network so always sees increments of the size
    it was trained on, whatever the macro time step; between commits the stress is
    C (strain - inelastic strain) with the anchor's inelastic strain advanced by one
    network step.  Rates below the training range are clamped to its lower end (the
    quasi-static limit of the viscoplastic response).

    Energies (the reference adapter's ``free_energy`` / ``dissipation``,
    ``nn/models/rve_rno.py:285-308``, in this network's variables): stored energy
    1/2 (strain - ein) . stress and inelastic power stress . d ein / dt, from the
    kernel.  Tangent: d stress / d strain = C (I - d ein / d strain) through the
    network's input Jacobian (layer-wise BF16 GEMMs, autograd), or the fallback
    RVE's effective tangent.

    Two gates send a point to the FE² adapter (``fallback``, a
    ``RVEHomogenizedMaterial``) for the rest of the run:

    * the input gate, in the kernel: an increment beyond ``gate_margin`` times the
      training envelope, a rate above it by more than ``rate_margin`` (in units of the
      rate feature), a latent beyond ``gate_margin`` times the range seen in
      training, or a non-finite result;
    * the error gate judged by FE²: ``monitor_points`` carry their own RVE states,
      advanced every ``check_interval`` calls to the current strain over the time
      since the last check (one FE² increment per check); a monitor whose surrogate
      stress differs from its FE² stress by more than ``error_bound`` times the
      stress scale (max norm) trips, and continues from that RVE state, i.e. from
      the FE² solution along its whole (checked) history.

    A point that trips the input gate without a monitor RVE replays its recorded
    path in a virgin RVE: zero strain at time 0, its anchor strain at the anchor's
    time (elapsed time minus the anchor's age), its current strain now, each segment
    in ``fallback_substeps`` increments over the segment's real duration, so the
    replayed rates are the rates the point saw (a point still on its first anchor is
    replayed along the straight path over the elapsed time).

    ``probe_points``: point ids whose strain, stress and time step are recorded at
    every call (``probe_history()``), for checking the run against the FE² oracle."""

    def __init__(self, name, params: SurrogateParams, scales: SurrogateScales, density: float,
                 maximum_wave_speed: float, fallback=None, fallback_substeps: int = 4,
                 gate_margin: float = 1.25, rate_margin: float = 0.5, commit_fraction: float = 0.5,
                 commit_time: float = 0.0, probe_points=None, monitor_points=None,
                 check_interval: int = 1, error_bound: float = 0.05):
        self.name = name
        self.params = params
        self.scales = scales
        self.density = float(density)
        self.maximum_wave_speed = float(maximum_wave_speed)
        self.fallback = fallback
        self.fallback_substeps = int(fallback_substeps)
        self.gate_margin = float(gate_margin)
        self.rate_margin = float(rate_margin)
        self.commit_fraction = float(commit_fraction)
        self.commit_time = float(commit_time)
        self.probe_points = None if probe_points is None else np.asarray(probe_points, dtype=np.int64)
        self.monitor_points = (None if monitor_points is None
                               else np.unique(np.asarray(monitor_points, dtype=np.int64)))
        if self.monitor_points is not None and fallback is None:
            raise ValueError("monitor points need an FE² fallback material to judge them")
        self.check_interval = max(1, int(check_interval))
        self.error_bound = float(error_bound)
        self._probe = {"strain": [], "stress": [], "dt": []}
        self.stats = {"surrogate_points": 0, "fallback_points": 0, "gate_trips": 0, "calls": 0,
                      "error_checks": 0, "error_trips": 0, "judged_error_max": 0.0,
                      "judged_error_last": []}
        self.call_times = []  # host clock after each call's device work (the trip count read)

    def gate(self) -> _lib.SurrogateGate:
        sc = self.scales
        g = _lib.SurrogateGate(sc.eps_scale, sc.rate_ref, sc.rate_scale, INELASTIC_STEP, sc.d_eps_max,
                               sc.rate_lo, sc.rate_hi, sc.latent_max, self.gate_margin, self.rate_margin,
                               self.commit_fraction * sc.d_eps_commit, self.commit_time)
        g.stiffness[:] = [float(v) for v in sc.stiffness_matrix().ravel()]
        return g

    def probe_history(self):
        """(strain (T, k, 6), stress (T, k, 6), dt (T,)) of the probe points, one row per call."""
        if not self._probe["dt"]:
            return np.zeros((0, 0, 6)), np.zeros((0, 0, 6)), np.zeros(0)
        return (torch.stack(self._probe["strain"]).transpose(1, 2).cpu().numpy(),
                torch.stack(self._probe["stress"]).transpose(1, 2).cpu().numpy(),
                np.asarray(self._probe["dt"]))

    def clear_probes(self):
        self._probe = {"strain": [], "stress": [], "dt": []}

    def initialize_state(self, n_points: int) -> MaterialState:
        return MaterialState({"surrogate_strain": np.zeros((n_points, 6)),
                              "surrogate_inelastic_strain": np.zeros((n_points, 6)),
                              "surrogate_current_inelastic_strain": np.zeros((n_points, 6)),
                              "surrogate_latent": np.zeros((n_points, 6)),
                              "surrogate_age": np.zeros(n_points),
                              "surrogate_time": np.zeros(n_points),
                              "surrogate_fallback": np.zeros(n_points)})

    def device_kernel(self, point_phase_ids=None):
        return _SurrogateKernel(self)

    def update(self, request, state, requirements=DEFAULT_RESPONSE_REQUIREMENTS):
        dev = device()
        k = _SurrogateKernel(self)
        strain = torch.as_tensor(np.ascontiguousarray(np.asarray(request.strains, dtype=np.float64).T),
                                 device=dev)
        st = k.from_host(state)
        r = k.update(strain, None, st, request.time_step, request.update_settings,
                     bool(requirements.tangent))
        tangents = None
        if requirements.tangent:
            tangents = r.tangent.t().reshape(-1, 6, 6).cpu().numpy().copy()
        return MaterialPointResponse(
            stresses=r.stress.cpu().numpy().T.copy(), state=k.material_state(r.state),
            tangents=tangents,
            free_energy=r.free_energy.cpu().numpy().copy() if requirements.free_energy else None,
            dissipation=r.dissipation.cpu().numpy().copy() if requirements.dissipation else None,
            diagnostics={})

    # -- tangent -----------------------------------------------------------------

    def network_tangent(self, strain: torch.Tensor, st: "_SurrogateState", dt: float) -> torch.Tensor:
        """d stress / d strain [36][n] (row-major 6 x 6 per point) of the network
        branch: stress = C (strain - ein'), ein' = anchor ein + c eps_scale o[0:6](x),
        x = [d / eps_scale, feature(|d|_2 / tau), latent], d = strain - anchor strain;
        the rate feature's derivative is d / (|d|_2^2 ln 10 rate_scale) where it is
        not clamped.  The six Jacobian rows come from six backward passes through the
        layer-wise kernels."""
        sc = self.scales
        n = strain.shape[1]
        d = (strain - st.strain).t()                         # [n][6]
        tau = st.age.double() + dt
        nrm = torch.linalg.norm(d, dim=1)
        rate = nrm / tau
        feat_raw = torch.log10(torch.clamp(rate, min=1e-30) / sc.rate_ref) / sc.rate_scale
        feat = torch.clamp(feat_raw, min=sc.rate_lo)
        x = torch.cat([d / sc.eps_scale, feat[:, None], torch.tanh(st.latent.t().double())],
                      dim=1).float()
        x = x.detach().requires_grad_(True)
        o = mlp(x, self.params)
        jac = torch.empty((n, 6, 13), dtype=torch.float64, device=strain.device)
        for i in range(6):
            (g,) = torch.autograd.grad(o[:, i].sum(), x, retain_graph=i < 5)
            jac[:, i, :] = g.double()
        dfeat = torch.where((feat_raw > sc.rate_lo) & (nrm > 0.0),
                            1.0 / (torch.clamp(nrm, min=1e-300) ** 2 * math.log(10.0) * sc.rate_scale),
                            torch.zeros_like(nrm))[:, None] * d           # [n][6]
        dein = INELASTIC_STEP * (jac[:, :, :6] + sc.eps_scale * jac[:, :, 6:7] * dfeat[:, None, :])
        Cm = torch.as_tensor(sc.stiffness_matrix(), device=strain.device)
        eye = torch.eye(6, dtype=torch.float64, device=strain.device)
        tan = torch.einsum("ij,njk->nik", Cm, eye[None] - dein)          # [n][6][6]
        return tan.reshape(n, 36).t().contiguous()


class _SurrogateState:
    """Anchors in the kernel's layout: strain / inelastic strain [6][n] fp64, latent
    [6][n] fp32, age [n] fp32, fallback [n] uint8; ``current`` [6][n] fp64 the last
    call's inelastic strain (for the dissipation); ``time`` the elapsed physical
    time (host); n_fallback the host's count of fallback points; rve_state /
    rve_index the FE² states of the fallback points; mon_state / mon_time /
    mon_calls the monitor RVEs, the time of their last check and the calls since."""

    def __init__(self, strain, inelastic, latent, age, fallback, n_fallback, rve_state=None,
                 rve_index=None, current=None, time=0.0, mon_state=None, mon_time=0.0, mon_calls=0):
        self.strain = strain
        self.inelastic = inelastic
        self.latent = latent
        self.age = age
        self.fallback = fallback
        self.n_fallback = int(n_fallback)
        self.rve_state = rve_state
        self.rve_index = rve_index
        self.current = current if current is not None else torch.zeros_like(inelastic)
        self.time = float(time)
        self.mon_state = mon_state
        self.mon_time = float(mon_time)
        self.mon_calls = int(mon_calls)


def _select(state, rows):
    """Rows of an FE² batch state (``fe2.RVEBatchState``)."""
    from concrete_impact_ai_amd.fe2 import RVEBatchState

    idx = torch.as_tensor(np.asarray(rows, dtype=np.int64), device=state.packed.device)
    return RVEBatchState(state.packed.index_select(0, idx).contiguous(),
                         state.w.index_select(0, idx).contiguous(),
                         np.asarray(state.macro)[np.asarray(rows, dtype=np.int64)], state._cache)


def _concat(states):
    from concrete_impact_ai_amd.fe2 import RVEBatchState

    states = [s for s in states if s is not None and len(s) > 0]
    if len(states) == 1:
        return states[0]
    return RVEBatchState(torch.cat([s.packed for s in states]).contiguous(),
                         torch.cat([s.w for s in states]).contiguous(),
                         np.concatenate([s.macro for s in states]), states[0]._cache)


class _SurrogateKernel:
    def __init__(self, material: SurrogateMaterial):
        self.m = material
        self.trips = torch.zeros(1, dtype=torch.int32, device=device())  # device counter
        self._trips_seen = 0

    def initial_state(self, n):
        dev = device()
        st = _SurrogateState(torch.zeros((6, n), dtype=torch.float64, device=dev),
                             torch.zeros((6, n), dtype=torch.float64, device=dev),
                             torch.zeros((6, n), dtype=torch.float32, device=dev),
                             torch.zeros(n, dtype=torch.float32, device=dev),
                             torch.zeros(n, dtype=torch.uint8, device=dev), 0)
        if self.m.monitor_points is not None:
            if self.m.monitor_points.size and self.m.monitor_points[-1] >= n:
                raise ValueError("monitor point ids beyond the number of points")
            st.mon_state = self.m.fallback.initialize_state(self.m.monitor_points.size)
        return st

    def from_host(self, state):
        dev = device()
        v = state.variables
        n = np.asarray(v["surrogate_strain"]).shape[0]
        fb = np.asarray(v["surrogate_fallback"]) > 0.5

        def rows(key, dtype):
            return torch.as_tensor(np.ascontiguousarray(np.asarray(v[key], dtype=np.float64).T),
                                   dtype=dtype, device=dev).contiguous()

        age = np.asarray(v["surrogate_age"]) if "surrogate_age" in v else np.zeros(n)
        cur = (rows("surrogate_current_inelastic_strain", torch.float64)
               if "surrogate_current_inelastic_strain" in v else None)
        t = float(np.asarray(v["surrogate_time"])[0]) if "surrogate_time" in v and n else 0.0
        st = _SurrogateState(rows("surrogate_strain", torch.float64),
                             rows("surrogate_inelastic_strain", torch.float64),
                             rows("surrogate_latent", torch.float32),
                             torch.as_tensor(age, dtype=torch.float32, device=dev),
                             torch.as_tensor(fb.astype(np.uint8), device=dev), int(fb.sum()),
                             v.get("_rve_state"), v.get("_rve_index"), cur, t,
                             v.get("_monitor_state"), float(v.get("_monitor_time", 0.0)),
                             int(v.get("_monitor_calls", 0)))
        if st.mon_state is None and self.m.monitor_points is not None:
            st.mon_state = self.m.fallback.initialize_state(self.m.monitor_points.size)
        return st

    def material_state(self, st):
        n = st.strain.shape[1]
        out = {"surrogate_strain": st.strain.t().cpu().numpy(),
               "surrogate_inelastic_strain": st.inelastic.t().cpu().numpy(),
               "surrogate_current_inelastic_strain": st.current.t().cpu().numpy(),
               "surrogate_latent": st.latent.t().double().cpu().numpy(),
               "surrogate_age": st.age.double().cpu().numpy(),
               "surrogate_time": np.full(n, st.time),
               "surrogate_fallback": st.fallback.double().cpu().numpy()}
        if st.rve_state is not None:
            out["_rve_state"] = st.rve_state
            out["_rve_index"] = st.rve_index
        if st.mon_state is not None:
            out["_monitor_state"] = st.mon_state
            out["_monitor_time"] = st.mon_time
            out["_monitor_calls"] = st.mon_calls
        return MaterialState(out)

    def update(self, strain, rate, st: _SurrogateState, dt, settings, tangent=False) -> DeviceResponse:
        """strain [6][n] fp64 on the device; ``rate`` is not used (the rate feature is
        the mean rate since the anchor)."""
        m = self.m
        n = strain.shape[1]
        dev = strain.device
        strain = strain.contiguous()
        if strain.dtype != torch.float64 or strain.shape[0] != 6:
            raise ValueError("surrogate strains must be [6][n] float64")
        stress = torch.empty((6, n), dtype=torch.float64, device=dev)
        psi = torch.empty(n, dtype=torch.float64, device=dev)
        diss = torch.empty(n, dtype=torch.float64, device=dev)
        nxt = _SurrogateState(torch.empty_like(st.strain), torch.empty_like(st.inelastic),
                              torch.empty_like(st.latent), torch.empty_like(st.age),
                              torch.empty_like(st.fallback), st.n_fallback, st.rve_state, st.rve_index,
                              torch.empty_like(st.current), st.time + float(dt), st.mon_state,
                              st.mon_time, st.mon_calls + 1)
        gate = m.gate()
        _check(_lib.lib().cia_surrogate_update(
            stream_handle(), n, _p(strain), float(dt), _p(m.params.w_bf16), _p(m.params.b), C.byref(gate),
            _p(st.strain), _p(st.inelastic), _p(st.latent), _p(st.age), _p(st.fallback), _p(stress),
            _p(nxt.strain), _p(nxt.inelastic), _p(nxt.latent), _p(nxt.age), _p(nxt.fallback),
            _p(self.trips), _p(st.current), _p(nxt.current), _p(psi), _p(diss)), "cia_surrogate_update")
        tan = m.network_tangent(strain, st, float(dt)) if tangent else None
        total = int(self.trips.item())
        m.call_times.append(time.perf_counter())
        new_trips = total - self._trips_seen
        self._trips_seen = total
        nxt.n_fallback = st.n_fallback + new_trips
        m.stats["calls"] += 1
        m.stats["gate_trips"] += new_trips
        mon_now = None
        if m.monitor_points is not None and len(m.monitor_points):
            mon_now = self._judge(strain, stress, st, nxt, dt, settings, new_trips > 0)
        if nxt.n_fallback:
            if m.fallback is None:
                raise RuntimeError(f"{nxt.n_fallback} surrogate points left the training envelope and "
                                   "no FE² fallback is configured.")
            self._fallback(strain, st, nxt, stress, psi, diss, tan, dt, settings, mon_now)
        m.stats["fallback_points"] += nxt.n_fallback
        m.stats["surrogate_points"] += n - nxt.n_fallback
        if m.probe_points is not None:
            ids = torch.as_tensor(m.probe_points, device=dev)
            m._probe["strain"].append(strain.index_select(1, ids))
            m._probe["stress"].append(stress.index_select(1, ids))
            m._probe["dt"].append(float(dt))
        return DeviceResponse(stress=stress, state=nxt, tangent=tan, free_energy=psi, dissipation=diss)

    # -- error gate judged by FE² ------------------------------------------------

    def _judge(self, strain, stress, st, nxt, dt, settings, envelope_trips):
        """Advance the monitor RVEs on check calls (and on calls where a monitor point
        has newly tripped the input gate), compare, trip.  Returns the monitor rows
        advanced to this call's strain (their FE² stresses) or None."""
        m = self.m
        mon = m.monitor_points
        fb_prev = st.fallback.index_select(0, torch.as_tensor(mon, device=strain.device)).cpu().numpy() > 0
        fb_now = nxt.fallback.index_select(0, torch.as_tensor(mon, device=strain.device)).cpu().numpy() > 0
        live = ~fb_prev  # monitors still on the surrogate before this call
        newly = live & fb_now
        is_check = nxt.mon_calls >= m.check_interval
        if not live.any() or (not is_check and not newly.any()):
            return None
        rows = np.flatnonzero(live) if is_check else np.flatnonzero(newly)
        ids = mon[rows]
        idx = torch.as_tensor(ids, device=strain.device)
        eps = strain.index_select(1, idx).t().cpu().numpy()
        span = nxt.time - st.mon_time  # one FE² increment since the monitors' last advance
        res, adv = m.fallback.solve_batch(eps, np.full(rows.size, span), settings,
                                          _select(st.mon_state, rows))
        fe2 = np.asarray([r.macro_stress[:] for r in res]).reshape(-1, 6)
        sur = stress.index_select(1, idx).t().cpu().numpy()
        err = np.max(np.abs(sur - fe2), axis=1) / m.scales.stress_scale
        # the advanced rows replace the monitors' states (points not advanced keep theirs)
        packed = st.mon_state.packed.clone()
        w = st.mon_state.w.clone()
        macro = np.asarray(st.mon_state.macro).copy()
        ridx = torch.as_tensor(rows, device=packed.device)
        packed.index_copy_(0, ridx, adv.packed)
        w.index_copy_(0, ridx, adv.w)
        macro[rows] = eps
        from concrete_impact_ai_amd.fe2 import RVEBatchState

        nxt.mon_state = RVEBatchState(packed, w, macro, st.mon_state._cache)
        if is_check:
            nxt.mon_time = nxt.time
            nxt.mon_calls = 0
            m.stats["error_checks"] += 1
            m.stats["judged_error_last"] = err.tolist()
            m.stats["judged_error_max"] = max(m.stats["judged_error_max"], float(np.max(err)))
            judged = err > m.error_bound
            trip_rows = rows[judged & ~newly[rows]]
            if trip_rows.size:
                nxt.fallback.index_fill_(0, torch.as_tensor(mon[trip_rows], device=strain.device), 1)
                nxt.n_fallback += int(trip_rows.size)
                m.stats["error_trips"] += int(trip_rows.size)
        else:
            nxt.mon_time = st.mon_time
        return {"rows": rows, "ids": ids, "stress": fe2, "results": res, "state": adv, "span": span}

    # -- FE² fallback ------------------------------------------------------------

    def _fallback(self, strain, st, nxt, stress, psi, diss, tan, dt, settings, mon_now):
        """FE² for the points on the fallback (their stresses, energies and tangents
        replace the network's): earlier trips continue from their committed RVE
        states; monitor points that trip continue from their monitor RVEs; other new
        trips replay their anchor path in a virgin RVE over its real durations."""
        m = self.m
        ad = m.fallback
        dev = strain.device
        need_tangent = tan is not None
        ids = torch.nonzero(nxt.fallback).flatten().cpu().numpy()
        old = nxt.rve_index if nxt.rve_index is not None else np.zeros(0, dtype=np.int64)
        new = np.setdiff1d(ids, old)
        mon_ids = m.monitor_points if m.monitor_points is not None else np.zeros(0, dtype=np.int64)
        new_mon = np.intersect1d(new, mon_ids)
        new_rep = np.setdiff1d(new, new_mon)
        states, order = [], []

        def put(points, res_list):
            pts = torch.as_tensor(points, device=dev)
            vals = np.asarray([r.macro_stress[:] for r in res_list]).reshape(-1, 6)
            stress[:, pts] = torch.as_tensor(np.ascontiguousarray(vals.T), device=dev)
            psi[pts] = torch.as_tensor(np.asarray([r.free_energy_density for r in res_list]), device=dev)
            diss[pts] = torch.as_tensor(np.asarray([r.dissipation_density for r in res_list]), device=dev)
            if need_tangent:
                t = np.asarray([r.effective_tangent[:] for r in res_list]).reshape(-1, 36)
                tan[:, pts] = torch.as_tensor(np.ascontiguousarray(t.T), device=dev)

        def eps_of(points):
            return strain.index_select(1, torch.as_tensor(points, device=dev)).t().cpu().numpy()

        if old.size:
            res, state = ad.solve_batch(eps_of(old), np.full(old.size, float(dt)), settings,
                                        nxt.rve_state, need_tangent)
            put(old, res)
            states.append(state)
            order.append(old)
        if new_mon.size:
            # the monitor RVEs were advanced to this call by _judge (error or input trips)
            pos = {int(p): k for k, p in enumerate(mon_now["ids"])}
            rows = [pos[int(p)] for p in new_mon]
            res = [mon_now["results"][k] for k in rows]
            if need_tangent:  # tangents are not formed in the monitor solve: solve again
                prev = _select(st.mon_state, [int(np.searchsorted(mon_ids, p)) for p in new_mon])
                res, _ = ad.solve_batch(eps_of(new_mon), np.full(new_mon.size, mon_now["span"]),
                                        settings, prev, True)
            put(new_mon, res)
            states.append(_select(mon_now["state"], rows))
            order.append(new_mon)
        if new_rep.size:
            e_now = eps_of(new_rep)
            pidx = torch.as_tensor(new_rep, device=dev)
            e_a = st.strain.index_select(1, pidx).t().cpu().numpy()
            age = st.age.index_select(0, pidx).double().cpu().numpy()
            t_now = nxt.time
            t_a = np.clip(st.time - age, 0.0, None)
            first = t_a <= 0.0  # still on the first anchor: the straight path, split in two
            t_a = np.where(first, 0.5 * t_now, t_a)
            e_a = np.where(first[:, None], 0.5 * e_now, e_a)
            S = max(1, m.fallback_substeps)
            ns = ad.initialize_state(new_rep.size)
            segments = ((np.zeros_like(e_a), e_a, np.zeros_like(t_a), t_a),
                        (e_a, e_now, t_a, np.full_like(t_a, t_now)))
            for seg, (e0, e1, t0, t1) in enumerate(segments):
                h = (t1 - t0) / S
                for k in range(1, S + 1):
                    final = seg == 1 and k == S
                    res, ns = ad.solve_batch(e0 + (e1 - e0) * (k / S), h, settings, ns,
                                             need_tangent and final)
            put(new_rep, res)
            states.append(ns)
            order.append(new_rep)
        nxt.rve_state = _concat(states)
        nxt.rve_index = np.concatenate(order).astype(np.int64)
