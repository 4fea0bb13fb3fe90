"""Pins of the float64 surrogate oracle (CPU): BPTT gradients against central finite
differences and against torch-CPU float64 autograd of the same rollout, and batched
rollout == per-path rollout (pattern of test_rve_rno_training.py:285-315)."""

import numpy as np
import torch

from oracle import mlp as OM

SMALL = (13, 16, 16, 16, 16, 12)


def _data(seed=0, B=3, T=5):
    rng = np.random.default_rng(seed)
    d_eps = rng.standard_normal((B, T, 6)) * 0.5
    rate = rng.standard_normal((B, T)) * 0.3
    target = rng.standard_normal((B, T, 6)) * 0.2
    return d_eps, rate, target


def test_bptt_gradient_matches_central_differences():
    params = OM.init_params(1, SMALL)
    d_eps, rate, target = _data()
    _, grads = OM.loss_and_grad(params, d_eps, rate, target)
    rng = np.random.default_rng(5)
    h = 1e-6
    for li in (0, 2, 4):
        w, b = params[li]
        for _ in range(4):
            i, j = rng.integers(w.shape[0]), rng.integers(w.shape[1])
            old = w[i, j]
            w[i, j] = old + h
            lp, _ = OM.loss_and_grad(params, d_eps, rate, target)
            w[i, j] = old - h
            lm, _ = OM.loss_and_grad(params, d_eps, rate, target)
            w[i, j] = old
            fd = (lp - lm) / (2 * h)
            assert abs(fd - grads[li][0][i, j]) <= 1e-7 * max(1.0, abs(fd)), (li, i, j)
        k = rng.integers(b.shape[0])
        old = b[k]
        b[k] = old + h
        lp, _ = OM.loss_and_grad(params, d_eps, rate, target)
        b[k] = old - h
        lm, _ = OM.loss_and_grad(params, d_eps, rate, target)
        b[k] = old
        assert abs((lp - lm) / (2 * h) - grads[li][1][k]) <= 1e-7


def test_bptt_gradient_matches_torch_autograd():
    params = OM.init_params(2, SMALL)
    d_eps, rate, target = _data(3, B=4, T=6)
    loss, grads = OM.loss_and_grad(params, d_eps, rate, target)
    tp = [(torch.tensor(w, requires_grad=True), torch.tensor(b, requires_grad=True)) for w, b in params]
    de, rt, tg = (torch.tensor(a) for a in (d_eps, rate, target))
    z = torch.zeros(4, 6, dtype=torch.float64)
    s = torch.zeros(4, 6, dtype=torch.float64)
    tl = 0.0
    for t in range(de.shape[1]):
        a = torch.cat([de[:, t], rt[:, t, None], z], dim=1)
        for li, (w, b) in enumerate(tp):
            a = a @ w.T + b
            if li < len(tp) - 1:
                a = torch.nn.functional.gelu(a)
        s = s + 0.1 * a[:, :6]     # stress increments, step 0.1
        tl = tl + ((s - tg[:, t]) ** 2).sum()
        z = z + a[:, 6:]
    tl = tl / (4 * 6 * 6)
    tl.backward()
    assert abs(float(tl) - loss) < 1e-13
    for (gw, gb), (w, b) in zip(grads, tp):
        np.testing.assert_allclose(gw, w.grad.numpy(), rtol=1e-10, atol=1e-13)
        np.testing.assert_allclose(gb, b.grad.numpy(), rtol=1e-10, atol=1e-13)


def test_batched_rollout_equals_per_path_rollout():
    params = OM.init_params(4, SMALL)
    d_eps, rate, _ = _data(6, B=5, T=7)
    sig, lat, _ = OM.rollout(params, d_eps, rate)
    for p in range(5):
        s1, l1, _ = OM.rollout(params, d_eps[p:p + 1], rate[p:p + 1])
        np.testing.assert_allclose(sig[p], s1[0], rtol=0, atol=1e-15)
        np.testing.assert_allclose(lat[p], l1[0], rtol=0, atol=1e-15)


def test_planted_mistake_in_latent_recurrence_is_caught():
    """Dropping the latent term of the BPTT recursion changes the gradient."""
    params = OM.init_params(1, SMALL)
    d_eps, rate, target = _data()
    _, good = OM.loss_and_grad(params, d_eps, rate, target)
    orig = OM.loss_and_grad

    def broken(params, d_eps, rate, target):
        B, T, _ = d_eps.shape
        sig, _, caches = OM.rollout(params, d_eps, rate)
        n = B * T * 6
        diff = sig - target
        grads = [(np.zeros_like(w), np.zeros_like(b)) for w, b in params]
        for t in range(T - 1, -1, -1):
            go = np.zeros((B, 12))
            go[:, :6] = 2.0 * diff[:, t] / n
            g = go
            for li in range(len(params) - 1, -1, -1):
                a, z = caches[t][li]
                if li < len(params) - 1:
                    g = g * OM.gelu_grad(z)
                grads[li][0][...] += g.T @ a
                grads[li][1][...] += g.sum(axis=0)
                g = g @ params[li][0]
        return 0.0, grads

    _, bad = broken(params, d_eps, rate, target)
    assert orig is OM.loss_and_grad
    assert np.max(np.abs(bad[0][0] - good[0][0])) > 1e-6 * np.max(np.abs(good[0][0]))


def test_planted_mistake_in_stress_recurrence_is_caught():
    """Treating each step's stress as the network output alone (no accumulation of the
    increments) is a different model: its loss and gradient differ from the oracle's."""
    params = OM.init_params(5, SMALL)
    d_eps, rate, target = _data()
    loss, good = OM.loss_and_grad(params, d_eps, rate, target)
    sig, _, _ = OM.rollout(params, d_eps, rate)
    B, T, _ = d_eps.shape
    direct = np.zeros_like(sig)
    z = np.zeros((B, 6))
    for t in range(T):
        o, _ = OM.mlp_forward(params, np.concatenate([d_eps[:, t], rate[:, t, None], z], axis=1))
        direct[:, t] = OM.STRESS_STEP * o[:, :6]
        z = z + o[:, 6:]
    np.testing.assert_allclose(direct[:, 0], sig[:, 0], rtol=0, atol=1e-15)  # equal at t = 0
    assert np.max(np.abs(direct[:, 1:] - sig[:, 1:])) > 1e-6 * np.max(np.abs(sig))
    assert abs(float(np.mean((direct - target) ** 2)) - loss) > 1e-9 * loss
