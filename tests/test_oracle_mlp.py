"""Pins of the float64 surrogate oracle (CPU): BPTT gradients against central finite
differences and against torch-CPU float64 autograd of the same rollout, batched
rollout == per-path rollout (pattern of test_rve_rno_training.py:285-315), the elastic
limit, and the anchored material-point policy against the rollout."""

import numpy as np
import torch

from oracle import mlp as OM

SMALL = (13, 16, 16, 16, 16, 12)


def _stiffness(seed=0):
    """A symmetric positive definite 6 x 6 stiffness (scaled units)."""
    a = np.random.default_rng(100 + seed).standard_normal((6, 6))
    return a @ a.T / 6.0 + np.eye(6)


C0 = _stiffness()


def _data(seed=0, B=3, T=5):
    rng = np.random.default_rng(seed)
    d_eps = rng.standard_normal((B, T, 6)) * 0.5
    rate = rng.standard_normal((B, T)) * 0.3
    target = rng.standard_normal((B, T, 6)) * 0.2
    return d_eps, rate, target


def test_bptt_gradient_matches_central_differences():
    params = OM.init_params(1, SMALL)
    d_eps, rate, target = _data()
    _, grads = OM.loss_and_grad(params, d_eps, rate, target, C0)
    rng = np.random.default_rng(5)
    h = 1e-6
    for li in (0, 2, 4):
        w, b = params[li]
        for _ in range(4):
            i, j = rng.integers(w.shape[0]), rng.integers(w.shape[1])
            old = w[i, j]
            w[i, j] = old + h
            lp, _ = OM.loss_and_grad(params, d_eps, rate, target, C0)
            w[i, j] = old - h
            lm, _ = OM.loss_and_grad(params, d_eps, rate, target, C0)
            w[i, j] = old
            fd = (lp - lm) / (2 * h)
            assert abs(fd - grads[li][0][i, j]) <= 1e-7 * max(1.0, abs(fd)), (li, i, j)
        k = rng.integers(b.shape[0])
        old = b[k]
        b[k] = old + h
        lp, _ = OM.loss_and_grad(params, d_eps, rate, target, C0)
        b[k] = old - h
        lm, _ = OM.loss_and_grad(params, d_eps, rate, target, C0)
        b[k] = old
        assert abs((lp - lm) / (2 * h) - grads[li][1][k]) <= 1e-7


def test_path_weighted_loss_equals_repeated_paths_and_central_differences():
    """Integer path weights equal repeating the paths: weights (2, 1, 3) on 3 paths give
    twice the unweighted loss and gradients of the 6-path batch with each path repeated
    its weight's times (the normaliser B T 6 halves); the weighted gradient also matches
    central differences."""
    params = OM.init_params(7, SMALL)
    d_eps, rate, target = _data(8, B=3, T=4)
    w = np.array([2.0, 1.0, 3.0])
    loss_w, g_w = OM.loss_and_grad(params, d_eps, rate, target, C0, path_weight=w)
    rep = np.repeat(np.arange(3), w.astype(int))
    loss_r, g_r = OM.loss_and_grad(params, d_eps[rep], rate[rep], target[rep], C0)
    assert abs(loss_w - 2.0 * loss_r) <= 1e-13 * loss_w
    for (a, b), (c, d) in zip(g_w, g_r):
        np.testing.assert_allclose(a, 2.0 * c, rtol=1e-11, atol=1e-15)
        np.testing.assert_allclose(b, 2.0 * d, rtol=1e-11, atol=1e-15)
    h = 1e-6
    wt, _ = params[2]
    for i, j in ((0, 0), (5, 11), (15, 3)):
        old = wt[i, j]
        wt[i, j] = old + h
        lp, _ = OM.loss_and_grad(params, d_eps, rate, target, C0, path_weight=w)
        wt[i, j] = old - h
        lm, _ = OM.loss_and_grad(params, d_eps, rate, target, C0, path_weight=w)
        wt[i, j] = old
        assert abs((lp - lm) / (2 * h) - g_w[2][0][i, j]) <= 1e-7 * max(1.0, abs(g_w[2][0][i, j]))


def test_carried_window_continues_the_rollout_and_its_truncated_gradient():
    """Rolling [0, s0) and then [s0, T) from final_state reproduces the full rollout's
    stresses and latents; the window's loss gradient from a fixed carried state (the
    truncated-BPTT gradient) matches central differences with that state held fixed."""
    params = OM.init_params(9, SMALL)
    d_eps, rate, target = _data(10, B=3, T=7)
    s0 = 3
    full, lat_full, _ = OM.rollout(params, d_eps, rate, C0)
    st = OM.final_state(params, d_eps[:, :s0], rate[:, :s0], C0)
    np.testing.assert_allclose(st[0], lat_full[:, s0], rtol=1e-12, atol=1e-14)
    tail, lat_tail, _ = OM.rollout(params, d_eps[:, s0:], rate[:, s0:], C0, st)
    np.testing.assert_allclose(tail, full[:, s0:], rtol=1e-11, atol=1e-13)
    np.testing.assert_allclose(lat_tail, lat_full[:, s0:], rtol=1e-12, atol=1e-14)
    _, g = OM.loss_and_grad(params, d_eps[:, s0:], rate[:, s0:], target[:, s0:], C0, state=st)
    h = 1e-6
    w1, _ = params[1]
    for i, j in ((0, 0), (7, 3), (15, 15)):
        old = w1[i, j]
        w1[i, j] = old + h
        lp, _ = OM.loss_and_grad(params, d_eps[:, s0:], rate[:, s0:], target[:, s0:], C0, state=st)
        w1[i, j] = old - h
        lm, _ = OM.loss_and_grad(params, d_eps[:, s0:], rate[:, s0:], target[:, s0:], C0, state=st)
        w1[i, j] = old
        assert abs((lp - lm) / (2 * h) - g[1][0][i, j]) <= 1e-7 * max(1.0, abs(g[1][0][i, j]))


def test_bptt_gradient_matches_torch_autograd():
    params = OM.init_params(2, SMALL)
    d_eps, rate, target = _data(3, B=4, T=6)
    loss, grads = OM.loss_and_grad(params, d_eps, rate, target, C0)
    tp = [(torch.tensor(w, requires_grad=True), torch.tensor(b, requires_grad=True)) for w, b in params]
    de, rt, tg = (torch.tensor(a) for a in (d_eps, rate, target))
    z = torch.zeros(4, 6, dtype=torch.float64)
    ein = torch.zeros(4, 6, dtype=torch.float64)
    Ct = torch.tensor(C0)
    eps = torch.cumsum(de, dim=1)
    tl = 0.0
    for t in range(de.shape[1]):
        a = torch.cat([de[:, t], rt[:, t, None], torch.tanh(z)], dim=1)
        for li, (w, b) in enumerate(tp):
            a = a @ w.T + b
            if li < len(tp) - 1:
                a = torch.nn.functional.gelu(a, approximate="tanh")
        ein = ein + 0.3 * a[:, :6]           # inelastic strain increments, step 0.3
        s = (eps[:, t] - ein) @ Ct.T         # sigma = C (eps - ein)
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
    sig, lat, _ = OM.rollout(params, d_eps, rate, C0)
    for p in range(5):
        s1, l1, _ = OM.rollout(params, d_eps[p:p + 1], rate[p:p + 1], C0)
        # equal up to the summation order of the matrix products (BLAS blocking)
        np.testing.assert_allclose(sig[p], s1[0], rtol=0, atol=1e-13)
        np.testing.assert_allclose(lat[p], l1[0], rtol=0, atol=1e-14)


def test_planted_mistake_in_latent_recurrence_is_caught():
    """Dropping the latent term of the BPTT recursion changes the gradient."""
    params = OM.init_params(1, SMALL)
    d_eps, rate, target = _data()
    _, good = OM.loss_and_grad(params, d_eps, rate, target, C0)
    orig = OM.loss_and_grad

    def broken(params, d_eps, rate, target):
        B, T, _ = d_eps.shape
        sig, _, caches = OM.rollout(params, d_eps, rate, C0)
        n = B * T * 6
        diff = sig - target
        grads = [(np.zeros_like(w), np.zeros_like(b)) for w, b in params]
        ge = np.zeros((B, 6))
        for t in range(T - 1, -1, -1):
            go = np.zeros((B, 12))
            ge = ge - (2.0 * diff[:, t] / n) @ C0
            go[:, :6] = OM.INELASTIC_STEP * ge
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
    """Summing the network's outputs as stress increments (no elastic part), or
    dropping the accumulation of the inelastic strain, is a different model: the
    rollout, the loss and the gradient differ from the oracle's."""
    params = OM.init_params(5, SMALL)
    d_eps, rate, target = _data()
    loss, good = OM.loss_and_grad(params, d_eps, rate, target, C0)
    sig, _, _ = OM.rollout(params, d_eps, rate, C0)
    B, T, _ = d_eps.shape
    no_elastic = np.zeros_like(sig)
    no_accum = np.zeros_like(sig)
    z = np.zeros((B, 6))
    s = np.zeros((B, 6))
    eps = np.cumsum(d_eps, axis=1)
    for t in range(T):
        o, _ = OM.mlp_forward(params, np.concatenate([d_eps[:, t], rate[:, t, None], np.tanh(z)], axis=1))
        s = s + OM.INELASTIC_STEP * o[:, :6] @ C0.T
        no_elastic[:, t] = -s
        no_accum[:, t] = (eps[:, t] - OM.INELASTIC_STEP * o[:, :6]) @ C0.T
        z = z + o[:, 6:]
    np.testing.assert_allclose(no_accum[:, 0], sig[:, 0], rtol=0, atol=1e-14)  # equal at t = 0
    for bad in (no_elastic, no_accum):
        assert np.max(np.abs(bad[:, 1:] - sig[:, 1:])) > 1e-6 * np.max(np.abs(sig))
        assert abs(float(np.mean((bad - target) ** 2)) - loss) > 1e-9 * loss


def test_zero_output_layer_gives_the_elastic_response():
    """With the output layer zeroed the network adds no inelastic strain: the
    rollout is sigma_t = C eps_t exactly and the latent stays zero."""
    params = OM.init_params(6, SMALL)
    params[-1] = (np.zeros_like(params[-1][0]), np.zeros_like(params[-1][1]))
    d_eps, rate, _ = _data(7, B=4, T=6)
    sig, lat, _ = OM.rollout(params, d_eps, rate, C0)
    np.testing.assert_allclose(sig, np.cumsum(d_eps, axis=1) @ C0.T, rtol=1e-14, atol=1e-14)
    assert not lat.any()


# ---- anchored material-point policy ------------------------------------------------

C_PHYS = 3e10 * _stiffness(1)  # Pa


def _gate(**kw):
    g = {"eps_scale": 1e-4, "rate_ref": 100.0, "rate_scale": 1.0,
         "inelastic_step": OM.INELASTIC_STEP, "d_eps_max": 1.0, "rate_lo": -1.0, "rate_hi": 1.0,
         "latent_max": 1e9, "gate_margin": 1.25, "rate_margin": 0.5, "commit_strain": 0.0,
         "commit_time": 0.0, "stiffness": C_PHYS.ravel()}
    g.update(kw)
    return g


def _path(seed, P=4, T=6, lo=0.5):
    """Training-like increments (|d|_inf ~ eps_scale) at rates of 10..1000 /s."""
    rng = np.random.default_rng(seed)
    inc = rng.standard_normal((P, T, 6))
    inc *= (rng.uniform(lo, 1.0, (P, T)) / np.max(np.abs(inc), axis=2))[..., None] * 1e-4
    rate = 10.0 ** rng.uniform(1, 3, (P, T))
    dts = np.linalg.norm(inc, axis=2) / rate
    return inc, dts


def _drive(params, gate, inc, dts, substeps=1, mistake=None):
    """Drive the material along the path, each increment split into ``substeps``
    equal calls; returns the stress after every call."""
    P, T, _ = inc.shape
    a = OM.initial_anchor(P)
    eps = np.zeros((P, 6))
    out = []
    for t in range(T):
        for _ in range(substeps):
            eps = eps + inc[:, t] / substeps
            dt = dts[:, t] / substeps
            if mistake == "no_age":
                a = dict(a, age=np.zeros(P))
            s, a, _, _ = OM.material_update(params, gate, eps, dt, a)
            if mistake == "step_increment":
                a = dict(a, strain=eps.copy())
            out.append(s)
    return np.stack(out, axis=1)


def _rollout_physical(params, gate, inc, dts):
    rate = np.linalg.norm(inc, axis=2) / dts
    feat = np.maximum(np.log10(rate / gate["rate_ref"]) / gate["rate_scale"], gate["rate_lo"])
    stress_scale = 1e6  # any: the rollout's scaled stiffness carries it
    sig, _, _ = OM.rollout(params, inc / gate["eps_scale"], feat, C_PHYS * gate["eps_scale"] / stress_scale)
    return sig * stress_scale


def test_material_update_committing_every_call_is_the_rollout():
    params = OM.init_params(7, SMALL)
    g = _gate()
    inc, dts = _path(1)
    got = _drive(params, g, inc, dts)
    ref = _rollout_physical(params, g, inc, dts)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12 * np.max(np.abs(ref)))


def test_material_update_substeps_reproduce_the_rollout_at_commits():
    """Each increment split into 4 calls: the anchor commits exactly once per
    increment (commit_strain just below the smallest increment), and the stress at
    that call equals the rollout's; a wrong increment (per call instead of since the
    anchor) or a lost anchor age breaks the identity."""
    params = OM.init_params(8, SMALL)
    inc, dts = _path(2, lo=0.8)  # 3/4 of any increment stays below the commit strain
    g = _gate(commit_strain=0.999 * np.min(np.max(np.abs(inc), axis=2)))
    ref = _rollout_physical(params, g, inc, dts)
    got = _drive(params, g, inc, dts, substeps=4)[:, 3::4]
    scale = np.max(np.abs(ref))
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12 * scale)
    for mistake in ("step_increment", "no_age"):
        bad = _drive(params, g, inc, dts, substeps=4, mistake=mistake)[:, 3::4]
        assert np.max(np.abs(bad - ref)) > 1e-6 * scale, mistake


def test_material_update_between_commits_evaluates_from_the_anchor():
    """Calls below the commit strain leave the anchor unchanged (age advanced) and
    return the stress of one network step over the whole increment since it."""
    params = OM.init_params(9, SMALL)
    g = _gate(commit_strain=1.0)
    a = OM.initial_anchor(3)
    rng = np.random.default_rng(3)
    eps = rng.standard_normal((3, 6)) * 3e-5
    s1, a1, _, _ = OM.material_update(params, g, eps, 1e-6, a)
    s2, a2, _, _ = OM.material_update(params, g, 2 * eps, 1e-6, a1)
    for k in ("strain", "inelastic", "latent"):
        np.testing.assert_array_equal(a2[k], a[k])
    np.testing.assert_allclose(a2["age"], 2e-6, rtol=1e-15)
    one, _, _, _ = OM.material_update(params, g, 2 * eps, 2e-6, a)
    np.testing.assert_allclose(s2, one, rtol=1e-13, atol=1e-13 * np.max(np.abs(one)))
    assert np.max(np.abs(s2 - s1)) > 0.0


def test_material_update_rate_clamp_and_gate():
    params = OM.init_params(10, SMALL)
    g = _gate(latent_max=1e9)
    a = OM.initial_anchor(4)
    eps = np.zeros((4, 6))
    eps[:, 0] = [5e-5, 5e-5, 2e-4, 5e-5]
    dt = np.asarray([5e-5 / 1e-3, 5e-5 / 10.0, 2e-4 / 100.0, 5e-5 / 1e5])  # rates 1e-3, 10, 100, 1e5
    s, nxt, trip, n_new = OM.material_update(params, g, eps, dt, a)
    # rate 1e-3 /s is clamped to the lower end of the range (10 /s)
    np.testing.assert_allclose(s[0], s[1], rtol=1e-14)
    # |d| = 2 eps_scale > 1.25 d_eps_max trips; rate 1e5 /s (feature 3 > 1 + 0.5) trips
    assert list(trip) == [False, False, True, True] and n_new == 2
    # an already-tripped point is not counted again
    _, _, _, again = OM.material_update(params, g, eps, dt, dict(a, fallback=nxt["fallback"]))
    assert again == 0
    # latent range
    g2 = _gate(latent_max=1e-12)
    _, _, trip2, _ = OM.material_update(params, g2, eps[:2], dt[:2], OM.initial_anchor(2))
    assert trip2.all()


# ---- production path design (training paths) -------------------------------------

def test_production_paths_reach_amplitude_and_target_rate():
    """Closed forms of the reference's production design: the peak engineering-tensor
    norm is the amplitude, the peak strain rate approaches the target (smoothstep
    peak rate 1.5 per unit segment time), the knots are hit at segment ends, and a
    reversed path returns through zero."""
    from concrete_impact_ai_amd.microstructure import production_strain_path, tensor_norm

    for fam in ("monotonic", "load_unload_reload", "reverse", "hold"):
        t, e = production_strain_path(fam, "negative_triaxial", 2e-3, 100.0, 3000, seed=3, perturbation=0.1)
        assert abs(np.max(tensor_norm(e)) - 2e-3) < 1e-12
        rate = tensor_norm(np.diff(e, axis=0)) / np.diff(t)
        assert 99.0 < np.max(rate) <= 100.0 + 1e-9
        np.testing.assert_allclose(e[0], 0.0)
    t, e = production_strain_path("reverse", "deviatoric_uniaxial", 1e-3, 50.0, 3000)
    d = np.asarray([1.0, -0.5, -0.5, 0, 0, 0]) / tensor_norm([1.0, -0.5, -0.5, 0, 0, 0])
    np.testing.assert_allclose(e[1000], 1e-3 * d, atol=1e-15)      # knot 1 at a third of the time
    np.testing.assert_allclose(e[2000], -1e-3 * d, atol=1e-15)     # knot -1
    np.testing.assert_allclose(e[-1], 0.5e-3 * d, atol=1e-15)
    t, e = production_strain_path("nonproportional", "shear_dominant", 1e-3, 50.0, 200, seed=1,
                                  turning_angle=np.pi / 3)
    assert abs(np.max(tensor_norm(e)) - 1e-3) < 1e-12
