"""Pins of the float64 RVE-RNO oracle (``oracle/rve_rno.py``), which restates
``EnergyDissipationRVERNO`` (``src/concrete_impact/nn/models/rve_rno.py:162-418``).

* Networks: torch's own ``nn.Linear``/``nn.SiLU`` modules, arranged with the
  reference's module names and loaded strictly from the artifact's state dict,
  give the same values; torch autograd gives the same input gradients
  (``ScalarMLP.value_and_input_gradient``, ``rve_rno.py:110-129``; the reference
  test ``test_rve_rno_training.py:285-297``) and Hessian products.
* Update: zero anchoring, batch = scalar, the latent increment / state rate /
  violation identities (``test_rve_rno_training.py:51-110``); the current-state
  stress is the autograd strain gradient of the anchored energy at the candidate
  state (``:112-171``); the tangent equals central differences (``:155-171``,
  ``:173-224`` with a zero-preserving input scale); D = -beta . state_rate in
  both evolutions, D >= 0 for the mobility form.
* Artifact: its metadata contract (``nn/deployment.py:30-65``), and its stress and
  tangent along the reference-held FE² sample path; planted mistakes (transposed
  energy weights, a wrong SiLU derivative) leave that band.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch
from concrete_impact_ai_amd import hdf5 as H
from oracle import rve_rno as R
from rve_rno_cases import ARTIFACT, SAMPLE, Tree as _Tree, artifact_state as _artifact_state
from rve_rno_cases import artifact_tree as _artifact_tree

VARIANTS = [(e, i) for e in ("direct_rate", "mobility_gradient")
            for i in ("cs_semi_implicit_forward_euler", "current_state_stress_forward_euler")]


def _tiny(evolution, integrator, seed=2026):
    torch.manual_seed(seed)
    tree = _Tree(2, 16, 2, evolution).double()
    for p in tree.parameters():
        with torch.no_grad():
            p.normal_(0.0, 0.6)
    state = {k: v.detach().numpy().copy() for k, v in tree.state_dict().items()}
    cfg = R.RNOConfig(integrator=integrator, evolution=evolution, latent_dimension=2,
                      strain_input_scale=(0.01, 0.02, 0.03, 0.04, 0.05, 0.06))
    return R.RVERNO(cfg, state), tree


def _strains(rng, b, scale=0.01):
    return rng.standard_normal((b, 6)) * scale


def test_artifact_metadata_contract():
    _, meta = _artifact_state()
    assert meta["model_family"] == "energy_rve_rno"
    assert meta["integrator"] == "current_state_stress_forward_euler"
    assert meta["evolution"] == "direct_rate" and meta["activation"] == "silu"
    assert (meta["latent_dimension"], meta["hidden_size"], meta["hidden_layers"]) == (6, 128, 4)
    assert meta["voigt_order"] == ["xx", "yy", "zz", "yz", "xz", "xy"]
    assert meta["strain_convention"] == "engineering_shear"
    assert meta["deployment_integrator"] == "explicit_only" and meta["dtype"] == "float32"
    assert "strain_input_scale" not in meta  # the schema default, six ones
    m = R.RVERNO.from_artifact(ARTIFACT)
    np.testing.assert_array_equal(m.scale, np.ones(6))


@pytest.mark.parametrize("which", ["equilibrium_energy", "nonequilibrium_energy"])
def test_energy_networks_match_torch_modules_and_autograd(which):
    tree = _artifact_tree()
    m = R.RVERNO.from_artifact(ARTIFACT)
    net = m.eq if which == "equilibrium_energy" else m.neq
    n_in = 6 if which == "equilibrium_energy" else 12
    rng = np.random.default_rng(5)
    x = rng.standard_normal((7, n_in)) * 0.7
    v = rng.standard_normal((7, 3, n_in))
    value, grad, hvp = net.gradient_hvp(x, v)
    xt = torch.tensor(x, requires_grad=True)
    yt = getattr(tree, which)(xt)[:, 0]
    gt = torch.autograd.grad(yt.sum(), xt, create_graph=True)[0]
    np.testing.assert_allclose(value, yt.detach().numpy(), rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(grad, gt.detach().numpy(), rtol=1e-12, atol=1e-13)
    for k in range(3):
        hk = torch.autograd.grad((gt * torch.tensor(v[:, k])).sum(), xt, retain_graph=True)[0]
        np.testing.assert_allclose(hvp[:, k], hk.numpy(), rtol=1e-11, atol=1e-12)
    v2, g2 = net.value_and_gradient(x)
    np.testing.assert_array_equal(v2, value)
    np.testing.assert_array_equal(g2, grad)


def test_state_evolution_matches_torch_module_and_jvp():
    tree = _artifact_tree()
    m = R.RVERNO.from_artifact(ARTIFACT)
    rng = np.random.default_rng(6)
    x = rng.standard_normal((5, 12)) * 0.5
    v = rng.standard_normal((5, 2, 12))
    out, jvp = m.evo.forward_jvp(x, v)
    xt = torch.tensor(x)
    np.testing.assert_allclose(out, tree.state_evolution(xt).detach().numpy(), rtol=1e-13,
                               atol=1e-14)
    for k in range(2):
        _, jt = torch.func.jvp(tree.state_evolution, (xt,), (torch.tensor(v[:, k]),))
        np.testing.assert_allclose(jvp[:, k], jt.detach().numpy(), rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("evolution,integrator", VARIANTS)
def test_update_identities(evolution, integrator):
    m, tree = _tiny(evolution, integrator)
    rng = np.random.default_rng(7)
    s, z = _strains(rng, 6), rng.standard_normal((6, 2)) * 0.5
    dt = rng.uniform(1e-4, 5e-3, 6)
    u = m.update_batch(s, z, dt)
    tau = m.config.reference_time_scale
    np.testing.assert_allclose(u["latent_state"] - z,
                               (dt / tau)[:, None] * u["dimensionless_state_rate"], rtol=1e-13,
                               atol=1e-15)
    np.testing.assert_allclose(u["state_rate"], u["dimensionless_state_rate"] / tau, rtol=1e-15)
    np.testing.assert_array_equal(u["thermodynamic_violation"], np.maximum(-u["dissipation"], 0))
    # beta from torch autograd of the torch modules at the previous state
    xt = torch.tensor(np.concatenate([s / m.scale, z], 1), requires_grad=True)
    beta = torch.autograd.grad(tree.nonequilibrium_energy(xt).sum(), xt)[0][:, 6:].numpy()
    np.testing.assert_allclose(u["dissipation"], -(beta * u["state_rate"]).sum(1), rtol=1e-11,
                               atol=1e-12)
    if evolution == "mobility_gradient":
        assert np.all(u["dissipation"] >= 0.0)
    # energy increment = anchored energy at the candidate minus at the previous state
    _, _, a_v, a_g = m.anchors()
    e = s / m.scale

    def anchored(zz):
        x = torch.tensor(np.concatenate([e, zz], 1))
        return tree.nonequilibrium_energy(x)[:, 0].detach().numpy() - a_v - (e * a_g).sum(1)
    np.testing.assert_allclose(u["nonequilibrium_energy_increment"],
                               anchored(u["latent_state"]) - anchored(z), rtol=1e-10, atol=1e-13)
    # batch = scalar
    for i in range(6):
        ui = m.update_batch(s[i:i + 1], z[i:i + 1], dt[i:i + 1])
        for k in u:
            np.testing.assert_allclose(ui[k][0], u[k][i], rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("evolution", ["direct_rate", "mobility_gradient"])
def test_zero_anchoring(evolution):
    m, _ = _tiny(evolution, "cs_semi_implicit_forward_euler")
    u = m.update_batch(np.zeros((2, 6)), np.zeros((2, 2)), np.full(2, 1e-3))
    typical = np.abs(m.update_batch(np.full((1, 6), 0.01), np.zeros((1, 2)), np.full(1, 1e-3))
                     ["stress"]).max()
    assert np.abs(u["stress"]).max() < 1e-14 * typical
    a = R.RVERNO.from_artifact(ARTIFACT)
    u = a.update_batch(np.zeros((1, 6)), np.zeros((1, 6)), np.zeros(1))
    assert np.abs(u["stress"]).max() < 1e-14 and abs(u["free_energy"][0]) < 1e-14


@pytest.mark.parametrize("evolution", ["direct_rate", "mobility_gradient"])
def test_current_state_stress_is_gradient_of_anchored_energy_at_candidate(evolution):
    m, tree = _tiny(evolution, "current_state_stress_forward_euler")
    rng = np.random.default_rng(8)
    s, z, dt = _strains(rng, 3), rng.standard_normal((3, 2)) * 0.5, np.full(3, 2e-3)
    u = m.update_batch(s, z, dt)
    zero_step = m.update_batch(s, z, np.zeros(3))
    assert np.abs(u["stress"] - zero_step["stress"]).max() > 1e-6
    a_eq_v, a_eq_g, a_v, a_g = m.anchors()
    st = torch.tensor(s, requires_grad=True)
    e = st / torch.tensor(m.scale)
    psi = (tree.equilibrium_energy(e)[:, 0] - a_eq_v - (e * torch.tensor(a_eq_g)).sum(1)
           + tree.nonequilibrium_energy(torch.cat([e, torch.tensor(u["latent_state"])], 1))[:, 0]
           - a_v - (e * torch.tensor(a_g)).sum(1))
    grad = torch.autograd.grad(psi.sum(), st)[0].numpy()
    np.testing.assert_allclose(u["stress"], grad, rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(u["free_energy"], psi.detach().numpy(), rtol=1e-12, atol=1e-14)


def _central_difference(m, s, z, dt, h):
    cols = []
    for k in range(6):
        d = np.zeros_like(s)
        d[:, k] = h
        cols.append((m.update_batch(s + d, z, dt)["stress"]
                     - m.update_batch(s - d, z, dt)["stress"]) / (2 * h))
    return np.stack(cols, axis=2)


@pytest.mark.parametrize("evolution,integrator", VARIANTS)
def test_tangent_matches_central_differences(evolution, integrator):
    m, _ = _tiny(evolution, integrator)
    rng = np.random.default_rng(9)
    s, z, dt = _strains(rng, 4), rng.standard_normal((4, 2)) * 0.5, np.full(4, 4e-3)
    u = m.update_batch(s, z, dt, tangent=True)
    fd = _central_difference(m, s, z, dt, 1e-7)
    np.testing.assert_allclose(u["tangent"], fd, rtol=2e-6, atol=2e-6 * np.abs(fd).max())
    no_t = m.update_batch(s, z, dt)
    for k in no_t:
        np.testing.assert_allclose(u[k], no_t[k], rtol=1e-14, atol=1e-16)


def test_artifact_tangent_matches_central_differences():
    m = R.RVERNO.from_artifact(ARTIFACT)
    rng = np.random.default_rng(10)
    s, z, dt = _strains(rng, 3), rng.standard_normal((3, 6)) * 0.3, np.full(3, 5e-3)
    u = m.update_batch(s, z, dt, tangent=True)
    fd = _central_difference(m, s, z, dt, 1e-6)
    np.testing.assert_allclose(u["tangent"], fd, rtol=2e-6, atol=2e-6 * np.abs(fd).max())


def _sample_path():
    f = H.File(SAMPLE)
    p = f["paths"][f["paths"].keys()[0]]
    return (p["macro/strain"].read(), p["macro/stress"].read(), p["macro/time_step"].read(),
            p["macro/effective_tangent"].read())


def _sample_error(m, eps, sig, dt):
    r = m.rollout(eps[:, None, :], dt[:, None])
    return np.linalg.norm(r["stress"][:, 0] - sig) / np.linalg.norm(sig), r


def test_artifact_follows_reference_fe2_sample_path(monkeypatch):
    """The trained artifact on the reference-held elastic FE² path (64 increments of
    3.47e-5 s, strains up to 3.7e-4): relative stress error 0.279 and tangent error
    0.227 against the stored effective tangent, inside the artifact's own test-split
    maximum path stress error (0.4156)."""
    eps, sig, dt, ceff = _sample_path()
    m = R.RVERNO.from_artifact(ARTIFACT)
    err, r = _sample_error(m, eps, sig, dt)
    assert 0.2 < err < 0.35
    u = m.update_batch(eps[-8:], r["latent_state"][-9:-1, 0], dt[-8:], tangent=True)
    t_err = np.linalg.norm(u["tangent"] - ceff[-8:]) / np.linalg.norm(ceff[-8:])
    assert t_err < 0.3
    # planted: a transposed hidden layer of an energy network, a wrong SiLU derivative
    for net in ("eq", "neq"):
        bad = R.RVERNO.from_artifact(ARTIFACT)
        n = getattr(bad, net)
        n.hidden[1] = (n.hidden[1][0].T.copy(), n.hidden[1][1])
        assert _sample_error(bad, eps, sig, dt)[0] > 0.5
    monkeypatch.setattr(R, "silu_d1", lambda a: R.logistic(a))
    assert _sample_error(m, eps, sig, dt)[0] > 0.4
