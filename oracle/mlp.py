"""Oracle: the recurrent material-point surrogate in float64 (test infrastructure only,
see ``oracle/__init__.py``).

The reference's surrogate (``EnergyDissipationRVERNO``,
``src/concrete_impact/nn/models/rve_rno.py:162-476``; rollout
``nn/training/rve_rno.py:113-198``) is a different network, so this oracle is the
build's own network restated in NumPy float64: per step

    x_t = [d_eps_t (6), rate_t (1), z_t (6)]          (13 features)
    h   = GELU(W4 GELU(W3 GELU(W2 GELU(W1 x + b1) + b2) + b3) + b4)
    o   = W5 h + b5                                   (12 outputs)
    sigma_t = sigma_{t-1} + c o[0:6],   z_{t+1} = z_t + o[6:12],   sigma_{-1} = z_0 = 0

(c = STRESS_STEP: the network predicts stress increments, so the latent only has
to carry the inelastic state, not the total strain)

with GELU(v) = v Phi(v) (erf form) and loss = mean over paths, steps and the six
components of (sigma_t - target_t)^2.  ``loss_and_grad`` is a hand-written
backpropagation through time.

Pinning (tests/test_oracle_mlp.py): gradients against central finite differences and
against torch-CPU float64 autograd of the same rollout; batched rollout equals the
per-path rollout (the pattern of ``test_rve_rno_training.py:285-315``).
"""

from __future__ import annotations

import numpy as np
from scipy.special import erf

SIZES = (13, 256, 256, 256, 256, 12)
STRESS_STEP = 0.1


def init_params(seed: int = 0, sizes=SIZES):
    """He-style initialisation: W_l ~ N(0, 2 / fan_in), b_l = 0; output layer scaled 0.1."""
    rng = np.random.default_rng(seed)
    params = []
    for li in range(len(sizes) - 1):
        fan_in, fan_out = sizes[li], sizes[li + 1]
        w = rng.standard_normal((fan_out, fan_in)) * np.sqrt(2.0 / fan_in)
        if li == len(sizes) - 2:
            w *= 0.1
        params.append((w, np.zeros(fan_out)))
    return params


def gelu(v):
    return 0.5 * v * (1.0 + erf(v / np.sqrt(2.0)))


def gelu_grad(v):
    return 0.5 * (1.0 + erf(v / np.sqrt(2.0))) + v * np.exp(-0.5 * v * v) / np.sqrt(2.0 * np.pi)


def mlp_forward(params, x):
    """One network evaluation; returns the output and the per-layer (input, pre-activation)."""
    cache = []
    a = x
    for li, (w, b) in enumerate(params):
        z = a @ w.T + b
        cache.append((a, z))
        a = gelu(z) if li < len(params) - 1 else z
    return a, cache


def rollout(params, d_eps, rate):
    """d_eps (B, T, 6), rate (B, T) -> stresses (B, T, 6), latents (B, T + 1, 6), caches."""
    B, T, _ = d_eps.shape
    z = np.zeros((B, 6))
    s = np.zeros((B, 6))
    sig = np.zeros((B, T, 6))
    lat = [z]
    caches = []
    for t in range(T):
        x = np.concatenate([d_eps[:, t], rate[:, t, None], z], axis=1)
        o, cache = mlp_forward(params, x)
        caches.append(cache)
        s = s + STRESS_STEP * o[:, :6]
        sig[:, t] = s
        z = z + o[:, 6:]
        lat.append(z)
    return sig, np.stack(lat, axis=1), caches


def loss_and_grad(params, d_eps, rate, target):
    """Mean squared stress error of the rollout and its gradient by BPTT."""
    B, T, _ = d_eps.shape
    sig, _, caches = rollout(params, d_eps, rate)
    n = B * T * 6
    diff = sig - target
    loss = float(np.sum(diff * diff) / n)
    grads = [(np.zeros_like(w), np.zeros_like(b)) for w, b in params]
    gz = np.zeros((B, 6))  # dL/dz_{t+1} flowing back
    gs = np.zeros((B, 6))  # dL/dsigma_t (sigma_{t+1} = sigma_t + c o[0:6] carries it back)
    for t in range(T - 1, -1, -1):
        go = np.zeros((B, 12))
        gs = gs + 2.0 * diff[:, t] / n
        go[:, :6] = STRESS_STEP * gs
        go[:, 6:] = gz                    # z_{t+1} = z_t + o[6:12]
        cache = caches[t]
        g = go
        for li in range(len(params) - 1, -1, -1):
            a, z = cache[li]
            if li < len(params) - 1:
                g = g * gelu_grad(z)
            gw, gb = grads[li]
            gw += g.T @ a
            gb += g.sum(axis=0)
            g = g @ params[li][0]
        gz = gz + g[:, 7:13]              # dL/dz_t = dL/dz_{t+1} + dL/dx_t[7:13]
    return loss, grads
