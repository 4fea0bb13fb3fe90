"""Oracle: the recurrent material-point surrogate in float64 (test infrastructure only,
see ``oracle/__init__.py``).

The reference's surrogate (``EnergyDissipationRVERNO``,
``src/concrete_impact/nn/models/rve_rno.py:162-476``; rollout
``nn/training/rve_rno.py:113-198``) is a different network, so this oracle is the
build's own network restated in NumPy float64: per step, in scaled units (strain
in units of the feature scale, stress in units of the stress scale),

    x_t = [d_eps_t (6), rate_t (1), tanh(z_t) (6)]    (13 features)
    h   = GELU(W4 GELU(W3 GELU(W2 GELU(W1 x + b1) + b2) + b3) + b4)
    o   = W5 h + b5                                   (12 outputs)
    eps_t = eps_{t-1} + d_eps_t,   ein_t = ein_{t-1} + c o[0:6]
    sigma_t = C (eps_t - ein_t),   z_{t+1} = z_t + o[6:12],   eps_{-1} = ein_{-1} = z_0 = 0

(c = INELASTIC_STEP; C = the homogenised elastic stiffness of the RVE: the network
predicts inelastic strain increments, so an elastic increment needs no network
output at all and the latent carries only the inelastic state; the network sees the
latent through tanh, a bounded input, so that an untrained recurrence cannot feed
back on itself without bound over long paths)

with GELU in its tanh form, 0.5 v (1 + tanh(sqrt(2/pi) (v + 0.044715 v^3))), and loss = mean over paths, steps and the six
components of (sigma_t - target_t)^2.  ``loss_and_grad`` is a hand-written
backpropagation through time.

Pinning (tests/test_oracle_mlp.py): gradients against central finite differences and
against torch-CPU float64 autograd of the same rollout; batched rollout equals the
per-path rollout (the pattern of ``test_rve_rno_training.py:285-315``); a zero output
layer gives the elastic response C eps exactly.  ``material_update`` (the anchored
material-point policy) is pinned to ``rollout``: with commits at every call it
reproduces the rollout, a path split into sub-steps reproduces the rollout at the
commit steps, low rates equal the clamped rate, and planted mistakes in the anchor
bookkeeping break those identities.
"""

from __future__ import annotations

import numpy as np

SIZES = (13, 256, 256, 256, 256, 12)
INELASTIC_STEP = 0.3


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


GELU_K = np.sqrt(2.0 / np.pi)
GELU_C = 0.044715


def gelu(v):
    """GELU, tanh form (Hendrycks & Gimpel): 0.5 v (1 + tanh(k (v + c v^3)))."""
    return 0.5 * v * (1.0 + np.tanh(GELU_K * (v + GELU_C * v ** 3)))


def gelu_grad(v):
    t = np.tanh(GELU_K * (v + GELU_C * v ** 3))
    return 0.5 * (1.0 + t) + 0.5 * v * (1.0 - t * t) * GELU_K * (1.0 + 3.0 * GELU_C * v * v)


def mlp_forward(params, x):
    """One network evaluation; returns the output and the per-layer (input, pre-activation)."""
    cache = []
    a = x
    for li, (w, b) in enumerate(params):
        z = a @ w.T + b
        cache.append((a, z))
        a = gelu(z) if li < len(params) - 1 else z
    return a, cache


def rollout(params, d_eps, rate, C, state=None):
    """d_eps (B, T, 6), rate (B, T), stiffness C (6, 6), all scaled -> stresses
    (B, T, 6), latents (B, T + 1, 6), caches.  ``state``: (z, ein, eps) at the start,
    each (B, 6) (default zeros: the start of a path)."""
    B, T, _ = d_eps.shape
    if state is None:
        z, ein, eps = np.zeros((B, 6)), np.zeros((B, 6)), np.zeros((B, 6))
    else:
        z, ein, eps = (np.array(v, dtype=np.float64) for v in state)
    sig = np.zeros((B, T, 6))
    lat = [z]
    caches = []
    for t in range(T):
        x = np.concatenate([d_eps[:, t], rate[:, t, None], np.tanh(z)], axis=1)
        o, cache = mlp_forward(params, x)
        caches.append(cache)
        eps = eps + d_eps[:, t]
        ein = ein + INELASTIC_STEP * o[:, :6]
        sig[:, t] = (eps - ein) @ C.T
        z = z + o[:, 6:]
        lat.append(z)
    return sig, np.stack(lat, axis=1), caches


def final_state(params, d_eps, rate, C, state=None):
    """(z, ein, eps) after rolling the recurrence over d_eps / rate from ``state``."""
    B, T, _ = d_eps.shape
    sig, lat, _ = rollout(params, d_eps, rate, C, state)
    eps = (np.zeros((B, 6)) if state is None else np.asarray(state[2])) + d_eps.sum(axis=1)
    ein = eps - sig[:, -1] @ np.linalg.inv(C).T        # sigma_T = C (eps_T - ein_T)
    return lat[:, -1], ein, eps


def loss_and_grad(params, d_eps, rate, target, C, path_weight=None, state=None):
    """Mean squared stress error of the rollout and its gradient by BPTT.  With
    ``path_weight`` (B,) each path's squared errors are multiplied by its weight
    (loss = sum_b w_b sum_{t,i} diff^2 / (B T 6), the path-normalised loss of
    ``nn/training/rve_rno.py:540-638`` when w_b = 1 / peak_b^2).  ``state`` as for
    ``rollout``: the gradient of a window that starts from a carried state, truncated at
    the window start (truncated BPTT)."""
    B, T, _ = d_eps.shape
    w_b = np.ones(B) if path_weight is None else np.asarray(path_weight, dtype=np.float64)
    sig, lat, caches = rollout(params, d_eps, rate, C, state)
    n = B * T * 6
    diff = sig - target
    loss = float(np.sum(w_b[:, None, None] * diff * diff) / n)
    grads = [(np.zeros_like(w), np.zeros_like(b)) for w, b in params]
    gz = np.zeros((B, 6))   # dL/dz_{t+1} flowing back
    ge = np.zeros((B, 6))   # dL/dein_t (ein_{t+1} = ein_t + c o[0:6] carries it back)
    for t in range(T - 1, -1, -1):
        go = np.zeros((B, 12))
        ge = ge - (2.0 * w_b[:, None] * diff[:, t] / n) @ C    # sigma_t = C (eps_t - ein_t)
        go[:, :6] = INELASTIC_STEP * ge
        go[:, 6:] = gz                          # z_{t+1} = z_t + o[6:12]
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
        # dL/dz_t = dL/dz_{t+1} + dL/dx_t[7:13] tanh'(z_t) (the network sees tanh(z_t))
        gz = gz + g[:, 7:13] * (1.0 - np.tanh(lat[:, t]) ** 2)
    return loss, grads


# ---------------------------------------------------------------------------
# Material-point update with anchors (the policy of the deployed surrogate)

GATE_FIELDS = ("eps_scale", "rate_ref", "rate_scale", "inelastic_step", "d_eps_max", "rate_lo",
               "rate_hi", "latent_max", "gate_margin", "rate_margin", "commit_strain", "commit_time")


def initial_anchor(n):
    return {"strain": np.zeros((n, 6)), "inelastic": np.zeros((n, 6)), "latent": np.zeros((n, 6)),
            "age": np.zeros(n), "fallback": np.zeros(n, dtype=bool), "current": np.zeros((n, 6))}


def material_energies(strain, stress, current_prev, current, dt):
    """Stored energy 1/2 (strain - ein') . stress and inelastic power
    stress . (ein' - ein_prev) / dt of one surrogate call (ein' = ``current``, the
    call's inelastic strain; ein_prev the previous call's).  The network has no
    hardening energy, so the stored energy is the elastic one of the homogenised
    stiffness (the free energy / dissipation pair the reference's adapter returns,
    ``nn/models/rve_rno.py:285-308``, in this network's state variables)."""
    psi = 0.5 * np.einsum("ni,ni->n", strain - current, stress)
    diss = np.einsum("ni,ni->n", stress, current - current_prev) / dt
    return psi, diss


def material_update(params, gate, strain, dt, anchor):
    """One call of the surrogate material at n points (float64, physical units).

    Every point keeps an anchor (strain, inelastic strain, latent, age = time since
    the anchor) and the network is evaluated once from it:

        d = strain - anchor_strain,  tau = age + dt,  rate = |d|_2 / tau
        x = [d / eps_scale, max(log10(rate / rate_ref) / rate_scale, rate_lo), latent]
        ein' = anchor_inelastic + inelastic_step eps_scale o[0:6],  z' = latent + o[6:12]
        stress = C (strain - ein')          (C = gate["stiffness"], 6 x 6)

    The anchor moves to (strain, ein', z', 0) when |d|_inf >= commit_strain or
    (commit_time > 0 and) tau >= commit_time, and otherwise stays with age tau.  A
    point trips when |d|_inf / eps_scale > gate_margin d_eps_max, the unclamped rate
    feature > rate_hi + rate_margin, |z'|_inf > gate_margin latent_max, or a result is
    not finite.  Returns (stress, next anchor, trip mask, number of new trips).

    The reference's deployed network is evaluated at every call from the committed
    latent (``RVERNOMaterialAdapter.update``, ``nn/deployment.py:97-165``); the anchors
    are this build's policy so that the network always sees increments of the size it
    was trained on, whatever the macro time step."""
    g = gate
    C = np.asarray(g["stiffness"], dtype=np.float64).reshape(6, 6)
    d = strain - anchor["strain"]
    tau = anchor["age"] + dt
    rate = np.linalg.norm(d, axis=1) / tau
    feat_raw = np.log10(np.maximum(rate, 1e-30) / g["rate_ref"]) / g["rate_scale"]
    feat = np.maximum(feat_raw, g["rate_lo"])
    x = np.concatenate([d / g["eps_scale"], feat[:, None], np.tanh(anchor["latent"])], axis=1)
    o, _ = mlp_forward(params, x)
    ein = anchor["inelastic"] + g["inelastic_step"] * g["eps_scale"] * o[:, :6]
    stress = (strain - ein) @ C.T
    z = anchor["latent"] + o[:, 6:]
    dinf = np.max(np.abs(d), axis=1)
    finite = np.all(np.isfinite(stress), axis=1) & np.all(np.isfinite(z), axis=1)
    trip = ((dinf / g["eps_scale"] > g["gate_margin"] * g["d_eps_max"])
            | (feat_raw > g["rate_hi"] + g["rate_margin"])
            | (np.max(np.abs(z), axis=1) > g["gate_margin"] * g["latent_max"]) | ~finite)
    commit = dinf >= g["commit_strain"]
    if g["commit_time"] > 0.0:
        commit |= tau >= g["commit_time"]
    c = commit[:, None]
    nxt = {"strain": np.where(c, strain, anchor["strain"]),
           "inelastic": np.where(c, ein, anchor["inelastic"]),
           "latent": np.where(c, z, anchor["latent"]),
           "age": np.where(commit, 0.0, tau),
           "fallback": anchor["fallback"] | trip,
           "current": ein}
    return stress, nxt, trip, int(np.sum(trip & ~anchor["fallback"]))
