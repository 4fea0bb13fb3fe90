"""GPU tests of the surrogate's BF16 matrix-core kernels and training path.

* GEMM kernels on inputs that are exact in bf16, against float64 products: the
  FP32-accumulated result must agree to ~1e-6 relative, which pins the MFMA operand
  and accumulator lane layout, the tiling and the bounds handling (odd sizes).
* Fused bias / GELU / GELU' epilogues, split-K weight gradient with the bias column.
* The network forward against the float64 oracle within bf16 rounding (2e-2), the
  fused all-layer kernel against the layer-wise path, BPTT gradients against the
  oracle's hand-written BPTT (relative 5e-2, cosine > 0.998).
* Adam training lowers the rollout loss; the gate sends out-of-envelope points to
  the FE² fallback.
"""

import numpy as np
import pytest
import torch

from oracle import mlp as OM

pytestmark = pytest.mark.gpu


def _bf(a):
    """Round float64 values to the nearest bf16 (returned as float64)."""
    t = torch.tensor(np.asarray(a, dtype=np.float32)).to(torch.bfloat16)
    return t.to(torch.float64).numpy()


def _dev(a, dtype):
    return torch.tensor(np.asarray(a), dtype=dtype, device="cuda").contiguous()


def _p(t):
    import ctypes as C
    return C.c_void_p(None) if t is None else C.c_void_p(t.data_ptr())


def _gelu(v):
    return OM.gelu(v)


def _dgelu(v):
    return OM.gelu_grad(v)


@pytest.mark.parametrize("M,N,K", [(37, 29, 45), (130, 256, 13), (257, 12, 256), (1, 16, 32)])
def test_linear_forward_matches_float64(gpu, M, N, K):
    from concrete_impact_ai_amd import _lib

    rng = np.random.default_rng(M + N + K)
    x = _bf(rng.standard_normal((M, K)))
    w = _bf(rng.standard_normal((N, K)) * 0.3)
    b = rng.standard_normal(N).astype(np.float32).astype(np.float64)
    ref = x @ w.T + b
    lib = _lib.lib()
    st = torch.cuda.current_stream().cuda_stream
    X, W, Bv = _dev(x, torch.bfloat16), _dev(w, torch.bfloat16), _dev(b, torch.float32)
    Y = torch.empty((M, N), dtype=torch.float32, device="cuda")
    Z = torch.empty((M, N), dtype=torch.float32, device="cuda")
    Yb = torch.empty((M, N), dtype=torch.bfloat16, device="cuda")
    assert lib.cia_mlp_linear_forward(st, M, N, K, _p(X), _p(W), _p(Bv), 0, _p(None), _p(Y), _p(None)) == 0
    torch.cuda.synchronize()
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(Y.cpu().numpy() - ref)) <= 2e-6 * scale
    assert lib.cia_mlp_linear_forward(st, M, N, K, _p(X), _p(W), _p(Bv), 1, _p(Z), _p(Y), _p(Yb)) == 0
    torch.cuda.synchronize()
    assert np.max(np.abs(Z.cpu().numpy() - ref)) <= 2e-6 * scale
    g = _gelu(Z.double().cpu().numpy())
    assert np.max(np.abs(Y.cpu().numpy() - g)) <= 1e-5 * max(1.0, np.max(np.abs(g)))
    assert np.max(np.abs(Yb.double().cpu().numpy() - g)) <= 1e-2 * max(1.0, np.max(np.abs(g)))


@pytest.mark.parametrize("M,N,K", [(37, 29, 45), (300, 256, 256), (64, 12, 256)])
def test_linear_backward_data_matches_float64(gpu, M, N, K):
    from concrete_impact_ai_amd import _lib

    rng = np.random.default_rng(7 + M)
    dz = _bf(rng.standard_normal((M, N)))
    w = _bf(rng.standard_normal((N, K)) * 0.3)
    zprev = rng.standard_normal((M, K)).astype(np.float32).astype(np.float64)
    ref = dz @ w
    lib = _lib.lib()
    st = torch.cuda.current_stream().cuda_stream
    DZ, W, ZP = _dev(dz, torch.bfloat16), _dev(w, torch.bfloat16), _dev(zprev, torch.float32)
    DX = torch.empty((M, K), dtype=torch.float32, device="cuda")
    assert lib.cia_mlp_linear_backward_data(st, M, N, K, _p(DZ), _p(W), _p(None), _p(DX), _p(None)) == 0
    torch.cuda.synchronize()
    assert np.max(np.abs(DX.cpu().numpy() - ref)) <= 2e-6 * np.max(np.abs(ref))
    assert lib.cia_mlp_linear_backward_data(st, M, N, K, _p(DZ), _p(W), _p(ZP), _p(DX), _p(None)) == 0
    torch.cuda.synchronize()
    ref2 = ref * _dgelu(zprev)
    assert np.max(np.abs(DX.cpu().numpy() - ref2)) <= 1e-5 * np.max(np.abs(ref2))


@pytest.mark.parametrize("M,N,K", [(37, 29, 45), (5000, 256, 13), (4096, 256, 256)])
def test_linear_backward_weight_matches_float64(gpu, M, N, K):
    from concrete_impact_ai_amd import _lib

    rng = np.random.default_rng(11 + M)
    dz = _bf(rng.standard_normal((M, N)))
    x = _bf(rng.standard_normal((M, K)))
    lib = _lib.lib()
    st = torch.cuda.current_stream().cuda_stream
    DZ, X = _dev(dz, torch.bfloat16), _dev(x, torch.bfloat16)
    dW = torch.full((N, K), 0.5, dtype=torch.float32, device="cuda")
    db = torch.full((N,), 0.25, dtype=torch.float32, device="cuda")
    nws = lib.cia_mlp_backward_weight_workspace(N, K)
    ws = torch.empty(nws, dtype=torch.float32, device="cuda")
    assert lib.cia_mlp_linear_backward_weight(st, M, N, K, _p(DZ), _p(X), _p(dW), _p(db), _p(ws), nws, 0) == 0
    torch.cuda.synchronize()
    rw, rb = dz.T @ x, dz.sum(axis=0)
    assert np.max(np.abs(dW.cpu().numpy() - rw)) <= 2e-5 * np.max(np.abs(rw))
    assert np.max(np.abs(db.cpu().numpy() - rb)) <= 2e-5 * np.max(np.abs(rb))
    assert lib.cia_mlp_linear_backward_weight(st, M, N, K, _p(DZ), _p(X), _p(dW), _p(db), _p(ws), nws, 1) == 0
    torch.cuda.synchronize()
    assert np.max(np.abs(dW.cpu().numpy() - 2 * rw)) <= 4e-5 * np.max(np.abs(rw))
    assert np.max(np.abs(db.cpu().numpy() - 2 * rb)) <= 4e-5 * np.max(np.abs(rb))


def _inputs(rng, M):
    x = rng.standard_normal((M, 13)) * 0.5
    return x


def test_network_forward_fused_and_layerwise_match_oracle(gpu):
    from concrete_impact_ai_amd import surrogate as S

    rng = np.random.default_rng(0)
    layers = OM.init_params(3)
    params = S.SurrogateParams(layers)
    for M in (1, 127, 1000, 4097):
        x = _inputs(rng, M)
        ref, _ = OM.mlp_forward(layers, x)
        X = _dev(x, torch.float32)
        y_l = S.mlp(X, params).detach().double().cpu().numpy()
        y_f = S.forward_fused(X, params).double().cpu().numpy()
        scale = np.max(np.abs(ref))
        assert np.max(np.abs(y_l - ref)) <= 2e-2 * scale, M
        assert np.max(np.abs(y_f - ref)) <= 2e-2 * scale, M
        assert np.max(np.abs(y_f - y_l)) <= 2e-3 * scale, M


def test_bptt_gradient_matches_oracle(gpu):
    from concrete_impact_ai_amd import surrogate as S

    rng = np.random.default_rng(1)
    layers = OM.init_params(5)
    B, T = 64, 8
    d_eps = rng.standard_normal((B, T, 6)) * 0.5
    rate = rng.standard_normal((B, T)) * 0.3
    target = rng.standard_normal((B, T, 6)) * 0.2
    a = np.random.default_rng(11).standard_normal((6, 6))
    Cs = a @ a.T / 6.0 + np.eye(6)     # scaled stiffness, symmetric positive definite
    loss_ref, g_ref = OM.loss_and_grad(layers, d_eps, rate, target, Cs)
    params = S.SurrogateParams(layers)
    scales = S.SurrogateScales(1.0, 1.0, 1.0, 1.0)
    w_leaf = params.w.detach().requires_grad_(True)
    b_leaf = params.b.detach().requires_grad_(True)
    # unit scales: feed the rate feature directly (rate_ref 1, rate_scale 1 -> log10);
    # build the features by hand to match the oracle's raw inputs
    de = _dev(d_eps, torch.float32)
    rt = _dev(rate, torch.float32)
    z = torch.zeros((B, 6), dtype=torch.float32, device="cuda")
    ein = torch.zeros((B, 6), dtype=torch.float32, device="cuda")
    eps = torch.cumsum(de, dim=1)
    Ct = _dev(Cs, torch.float32)
    loss = 0.0
    tg = _dev(target, torch.float32)
    for t in range(T):
        x = torch.cat([de[:, t], rt[:, t, None], torch.tanh(z)], dim=1)
        o = S.mlp(x, params, w_leaf, b_leaf)
        ein = ein + S.INELASTIC_STEP * o[:, :6]
        sg = (eps[:, t] - ein) @ Ct.T
        loss = loss + ((sg - tg[:, t]) ** 2).sum()
        z = z + o[:, 6:]
    loss = loss / (B * T * 6)
    loss.backward()
    assert abs(float(loss) - loss_ref) <= 2e-2 * loss_ref
    gw = w_leaf.grad.double().cpu().numpy()
    gb = b_leaf.grad.double().cpu().numpy()
    rw = np.concatenate([g[0].ravel() for g in g_ref])
    rb = np.concatenate([g[1].ravel() for g in g_ref])
    for got, ref in ((gw, rw), (gb, rb)):
        cos = got @ ref / (np.linalg.norm(got) * np.linalg.norm(ref))
        assert cos > 0.998
        assert np.linalg.norm(got - ref) <= 5e-2 * np.linalg.norm(ref)
    del scales


def _fused_bptt(params, d_eps, rate, target, Cs, idx, window, path_w):
    """Loss and gradients of cia_surrogate_bptt with unit scales (the raw rate is the
    rate feature, as in the oracle)."""
    from concrete_impact_ai_amd import surrogate as S

    P, T, _ = d_eps.shape
    feat = np.zeros((P, T, 8))
    feat[..., :6] = d_eps
    feat[..., 6] = rate
    data = S.BpttData(feat, np.cumsum(d_eps, axis=1), target, Cs)
    step = S.BpttStep(params, len(idx), window)
    loss = step(params, data, torch.as_tensor(np.asarray(idx), dtype=torch.int32, device="cuda"),
                torch.as_tensor(path_w, dtype=torch.float32, device="cuda"))
    torch.cuda.synchronize()
    return float(loss), step.gw.double().cpu().numpy(), step.gb.double().cpu().numpy()


def _spd(seed):
    a = np.random.default_rng(seed).standard_normal((6, 6))
    return a @ a.T / 6.0 + np.eye(6)


def test_fused_bptt_matches_oracle_and_layerwise(gpu):
    """cia_surrogate_bptt (fused forward / backward through time, weight-gradient
    GEMMs) against the FP64 oracle's loss_and_grad at the bf16 tolerance of the
    layer-wise test, and against the layer-wise kernels driven by autograd (same bf16
    roundings) at 1e-2 of the gradient norm."""
    from concrete_impact_ai_amd import surrogate as S

    rng = np.random.default_rng(1)
    layers = OM.init_params(5)
    B, T = 64, 8
    d_eps = rng.standard_normal((B, T, 6)) * 0.5
    rate = rng.standard_normal((B, T)) * 0.3
    target = rng.standard_normal((B, T, 6)) * 0.2
    Cs = _spd(11)
    loss_ref, g_ref = OM.loss_and_grad(layers, d_eps, rate, target, Cs)
    params = S.SurrogateParams(layers)
    loss, gw, gb = _fused_bptt(params, d_eps, rate, target, Cs, np.arange(B), T, np.ones(B))
    rw = np.concatenate([g[0].ravel() for g in g_ref])
    rb = np.concatenate([g[1].ravel() for g in g_ref])
    assert abs(loss - loss_ref) <= 2e-2 * loss_ref
    for got, ref in ((gw, rw), (gb, rb)):
        cos = got @ ref / (np.linalg.norm(got) * np.linalg.norm(ref))
        assert cos > 0.998
        assert np.linalg.norm(got - ref) <= 5e-2 * np.linalg.norm(ref)
    # layer-wise kernels + autograd on the same batch
    w_leaf = params.w.detach().requires_grad_(True)
    b_leaf = params.b.detach().requires_grad_(True)
    de = _dev(d_eps, torch.float32)
    rt = _dev(rate, torch.float32)
    z = torch.zeros((B, 6), dtype=torch.float32, device="cuda")
    ein = torch.zeros((B, 6), dtype=torch.float32, device="cuda")
    eps = torch.cumsum(de, dim=1)
    Ct = _dev(Cs, torch.float32)
    tg = _dev(target, torch.float32)
    lw = 0.0
    for t in range(T):
        o = S.mlp(torch.cat([de[:, t], rt[:, t, None], torch.tanh(z)], dim=1), params, w_leaf, b_leaf)
        ein = ein + S.INELASTIC_STEP * o[:, :6]
        lw = lw + (((eps[:, t] - ein) @ Ct.T - tg[:, t]) ** 2).sum()
        z = z + o[:, 6:]
    lw = lw / (B * T * 6)
    lw.backward()
    assert abs(loss - float(lw.detach())) <= 2e-3 * loss
    lgw = w_leaf.grad.double().cpu().numpy()
    lgb = b_leaf.grad.double().cpu().numpy()
    assert np.linalg.norm(gw - lgw) <= 1e-2 * np.linalg.norm(lgw)
    assert np.linalg.norm(gb - lgb) <= 1e-2 * np.linalg.norm(lgb)


def test_fused_bptt_ragged_batch_window_and_path_weights(gpu):
    """A batch of 37 rows (not a multiple of the 16-row tile) drawn with repeats from 20
    paths of 9 steps, a window of 6 steps and unequal path weights: loss and gradients
    equal the oracle's path-weighted loss_and_grad on the gathered windows, with the
    workspace taken from NaN-filled memory (the tile's rows past the batch are read by
    the backward kernel)."""
    from concrete_impact_ai_amd import surrogate as S

    rng = np.random.default_rng(4)
    layers = OM.init_params(6)
    P, T, L = 20, 9, 6
    d_eps = rng.standard_normal((P, T, 6)) * 0.5
    rate = rng.standard_normal((P, T)) * 0.3
    target = rng.standard_normal((P, T, 6)) * 0.2
    pw = rng.uniform(0.5, 2.0, P)
    idx = rng.integers(0, P, 37)
    Cs = _spd(12)
    loss_ref, g_ref = OM.loss_and_grad(layers, d_eps[idx, :L], rate[idx, :L], target[idx, :L], Cs,
                                       path_weight=pw[idx])
    params = S.SurrogateParams(layers)
    # the workspace comes from blocks filled with NaN bit patterns: rows past the
    # batch must be written by the kernels before anything reads them
    from concrete_impact_ai_amd import _lib

    ws = int(_lib.lib().cia_surrogate_bptt_workspace(len(idx), L))
    for _ in range(4):
        torch.full((ws,), 255, dtype=torch.uint8, device="cuda")
    loss, gw, gb = _fused_bptt(params, d_eps, rate, target, Cs, idx, L, pw)
    rw = np.concatenate([g[0].ravel() for g in g_ref])
    rb = np.concatenate([g[1].ravel() for g in g_ref])
    assert abs(loss - loss_ref) <= 2e-2 * loss_ref
    for got, ref in ((gw, rw), (gb, rb)):
        cos = got @ ref / (np.linalg.norm(got) * np.linalg.norm(ref))
        assert cos > 0.998
        assert np.linalg.norm(got - ref) <= 5e-2 * np.linalg.norm(ref)


def test_fused_bptt_carried_windows_match_oracle(gpu):
    """Truncated BPTT over 12-step paths in windows of 5, 5 and 2 steps, the recurrence
    state carried by the kernel's own state_out: every window's loss and gradients equal
    the oracle's truncated-BPTT loss_and_grad from the oracle's carried state, and the
    carried latent and inelastic strain follow the oracle's."""
    from concrete_impact_ai_amd import surrogate as S

    rng = np.random.default_rng(8)
    layers = OM.init_params(9)
    P, T = 40, 12
    d_eps = rng.standard_normal((P, T, 6)) * 0.5
    rate = rng.standard_normal((P, T)) * 0.3
    target = rng.standard_normal((P, T, 6)) * 0.2
    pw = rng.uniform(0.5, 2.0, P)
    Cs = _spd(13)
    feat = np.zeros((P, T, 8))
    feat[..., :6] = d_eps
    feat[..., 6] = rate
    data = S.BpttData(feat, np.cumsum(d_eps, axis=1), target, Cs)
    params = S.SurrogateParams(layers)
    idx = torch.arange(P, dtype=torch.int32, device="cuda")
    pw_dev = torch.as_tensor(pw, dtype=torch.float32, device="cuda")
    states = [torch.zeros((P, 12), dtype=torch.float32, device="cuda") for _ in range(2)]
    st_ref = None
    for s0, L in ((0, 5), (5, 5), (10, 2)):
        sl = slice(s0, s0 + L)
        loss_ref, g_ref = OM.loss_and_grad(layers, d_eps[:, sl], rate[:, sl], target[:, sl], Cs,
                                           path_weight=pw, state=st_ref)
        st_ref = OM.final_state(layers, d_eps[:, sl], rate[:, sl], Cs, st_ref)
        step = S.BpttStep(params, P, L)
        loss = float(step(params, data, idx, pw_dev, s0, None if s0 == 0 else states[0], states[1]))
        states.reverse()
        got_state = states[0].double().cpu().numpy()
        for part, ref in ((got_state[:, :6], st_ref[0]), (got_state[:, 6:], st_ref[1])):
            assert np.max(np.abs(part - ref)) <= 2e-2 * np.max(np.abs(ref)), s0
        assert abs(loss - loss_ref) <= 2e-2 * loss_ref, s0
        rw = np.concatenate([g[0].ravel() for g in g_ref])
        rb = np.concatenate([g[1].ravel() for g in g_ref])
        gw, gb = step.gw.double().cpu().numpy(), step.gb.double().cpu().numpy()
        for got, ref in ((gw, rw), (gb, rb)):
            cos = got @ ref / (np.linalg.norm(got) * np.linalg.norm(ref))
            assert cos > 0.998, s0
            assert np.linalg.norm(got - ref) <= 5e-2 * np.linalg.norm(ref), s0


def test_fused_bptt_clips_the_gradient_norm_and_zeroes_non_finite_steps(gpu):
    """max_grad_norm scales the gradients to that global norm (direction unchanged) and
    reports the norm before clipping; a non-finite target makes the norm non-finite and
    the step's gradients zero."""
    from concrete_impact_ai_amd import surrogate as S

    rng = np.random.default_rng(21)
    P, T = 32, 6
    d_eps = rng.standard_normal((P, T, 6)) * 0.5
    rate = rng.standard_normal((P, T)) * 0.3
    target = rng.standard_normal((P, T, 6)) * 0.2
    feat = np.zeros((P, T, 8))
    feat[..., :6] = d_eps
    feat[..., 6] = rate
    params = S.SurrogateParams(OM.init_params(4))
    idx = torch.arange(P, dtype=torch.int32, device="cuda")
    pw = torch.ones(P, dtype=torch.float32, device="cuda")

    def run(tg, max_norm):
        data = S.BpttData(feat, np.cumsum(d_eps, axis=1), tg, _spd(3))
        step = S.BpttStep(params, P, T)
        step.max_grad_norm = max_norm
        step(params, data, idx, pw)
        torch.cuda.synchronize()
        return (step.gw.double().cpu().numpy(), step.gb.double().cpu().numpy(),
                float(step.grad_norm))

    gw, gb, norm = run(target, 0.0)
    full = np.sqrt(np.sum(gw ** 2) + np.sum(gb ** 2))
    assert abs(norm - full) <= 1e-4 * full
    cw, cb, cnorm = run(target, 0.25 * full)
    assert abs(cnorm - norm) <= 1e-6 * norm
    np.testing.assert_allclose(cw, 0.25 * gw, rtol=1e-5, atol=1e-12 * full)
    np.testing.assert_allclose(cb, 0.25 * gb, rtol=1e-5, atol=1e-12 * full)
    bad = target.copy()
    bad[5, 2, 1] = np.nan
    zw, zb, znorm = run(bad, 1.0)
    assert not np.isfinite(znorm)
    assert not zw.any() and not zb.any()


def test_training_reduces_rollout_loss(gpu):
    """Fit a viscoelastic-like recurrence sigma_t = 0.8 sigma_{t-1} + 1e10 d_eps_t
    (elastic stiffness 1e10 I: the inelastic strain takes up the relaxation)."""
    from concrete_impact_ai_amd import surrogate as S

    rng = np.random.default_rng(2)
    P, T = 256, 16
    d_eps = rng.standard_normal((P, T, 6)) * 1e-4
    rate = np.exp(rng.uniform(np.log(10.0), np.log(1000.0), (P, T)))
    sig = np.zeros((P, T, 6))
    s = np.zeros((P, 6))
    for t in range(T):
        s = 0.8 * s + 1e10 * d_eps[:, t]
        sig[:, t] = s
    scales = S.SurrogateScales.fit(d_eps, rate, sig, 1e10 * np.eye(6))
    params = S.SurrogateParams(seed=3)
    hist = S.train(params, scales, d_eps, rate, sig, iterations=300, batch_paths=128, window=T,
                   lr=2e-3, log_every=50)
    assert hist[-1][1] < 0.05 * hist[0][1]
    pred, _ = S.rollout(params, scales, _dev(d_eps, torch.float32), _dev(rate, torch.float32), fused=True)
    err = np.max(np.abs(pred.double().cpu().numpy() - sig)) / np.max(np.abs(sig))
    assert err < 0.2


def test_gate_sends_out_of_envelope_points_to_fe2(gpu):
    from concrete_impact_ai_amd import fe2
    from concrete_impact_ai_amd import materials as M
    from concrete_impact_ai_amd import rve as R
    from concrete_impact_ai_amd import solvers as Sv
    from concrete_impact_ai_amd import surrogate as S

    mat = M.J2ViscoplasticMaterial("j2", 2400.0, 25e9, 0.2, 20e6, 0.5e9, 1e-3, 20e6, 1.0)
    nodes, elements = R.build_structured_hex8_box((1.0, 1.0, 1.0), (1, 1, 1))
    model = R.StructuredHex8RVE(nodes, elements, (1, 1, 1), mat,
                                R.build_structured_hex8_periodic_constraint(nodes, (1, 1, 1)),
                                Sv.NonlinearNewtonSettings(20, 1e-3, 1e-10, 1e-14, 1e-10))
    ad = fe2.RVEHomogenizedMaterial(model, "fallback")
    scales = S.SurrogateScales(1e-4, 100.0, 1.0, 1e8, d_eps_max=1.0, rate_lo=-1.0, rate_hi=1.0,
                               latent_max=1e6, stiffness=tuple(3e10 * np.eye(6).ravel()))
    sm = S.SurrogateMaterial("sur", S.SurrogateParams(seed=0), scales, 2400.0, 4000.0, fallback=ad,
                             fallback_substeps=4)
    try:
        n = 6
        st = sm.initialize_state(n)
        eps = np.zeros((n, 6))
        eps[:, 0] = 5e-5
        eps[2, 0] = 3e-3          # increment 30x the training envelope -> FE²
        req = M.MaterialPointRequest(eps, None, 1e-5, "three_dimensional",
                                     M.MaterialUpdateSettings(50, 1e-10, 1e-12, 1e-10))
        r = sm.update(req, st)
        fb = r.state.variables["surrogate_fallback"]
        assert fb[2] == 1.0 and fb.sum() == 1.0
        # the point trips on its first anchor: the fallback stress equals a virgin RVE
        # driven along the straight path over the elapsed time, in two segments of
        # 4 increments (8 increments of dt / 8)
        ns = ad.initialize_state(1)
        for k in range(1, 9):
            rr = ad.update(M.MaterialPointRequest(eps[2:3] * k / 8, None, 1e-5 / 8, "three_dimensional",
                                                  M.MaterialUpdateSettings(50, 1e-10, 1e-12, 1e-10)), ns)
            ns = rr.state
        np.testing.assert_allclose(r.stresses[2], rr.stresses[0], rtol=1e-12)
        assert sm.stats["gate_trips"] == 1
    finally:
        ad.close()
