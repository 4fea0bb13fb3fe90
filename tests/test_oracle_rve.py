"""Pins of the oracle's Hex8, periodic and RVE restatement (CPU only)."""

import numpy as np
import pytest

from oracle import hex8, j2vp as OJ, periodic, rve as OR

SET = OJ.UpdateSettings(50, 1e-10, 1e-12, 1e-10)
REF_PATH = np.asarray([
    [0.0020, -0.0005, -0.0005, 0.0000, 0.0000, 0.0100],
    [0.0040, -0.0010, -0.0010, 0.0020, -0.0030, 0.0200],
    [0.0030, -0.0008, -0.0008, 0.0010, -0.0020, 0.0120],
    [0.0045, -0.0012, -0.0012, -0.0010, 0.0020, -0.0100],
])
REF_MAT = OJ.J2VPParams(3339.9160679991464, 0.36994096308414537, 10.0, 80.0, 0.02, 10.0, 2.0)


def test_uniform_strain_reaction():
    """test_nonlinear_solid.py:29-77: reaction on x = 1 equals (lambda + 2G) eps at 1e-12."""
    nodes, el = hex8.structured_box((1.0, 1.0, 1.0), (1, 1, 1))
    b, jw = hex8.geometry_cache(nodes, el)
    u = np.zeros(3 * nodes.shape[0])
    u[0::3] = 1e-3 * nodes[:, 0]
    eps, _ = hex8.strain_fields(el, b, u, np.zeros_like(u), 1.0)
    c = OJ.elasticity_matrix(1000.0, 0.25)
    f = hex8.internal_force(el, b, jw, eps @ c.T, u.size)
    right = np.flatnonzero(np.isclose(nodes[:, 0], 1.0))
    g = 1000.0 / 2.5
    lam = 1000.0 * 0.25 / (1.25 * 0.5)
    assert np.sum(f[3 * right]) == pytest.approx((lam + 2 * g) * 1e-3, rel=1e-12, abs=1e-12)


def test_patch_test_volume_and_null_space():
    nodes, el = hex8.structured_box((1.0, 0.7, 1.3), (3, 2, 4))
    rng = np.random.default_rng(0)
    nodes = nodes + 0.02 * rng.normal(size=nodes.shape) * (nodes > 0) * (nodes < [1.0, 0.7, 1.3])
    b, jw = hex8.geometry_cache(nodes, el)
    assert np.sum(jw) == pytest.approx(1.0 * 0.7 * 1.3, rel=1e-12)  # interior jiggle keeps volume
    e0 = np.asarray([1e-3, -2e-4, 3e-4, 5e-4, -1e-4, 2e-4])
    u = (periodic.affine_matrix(nodes) @ e0)
    # affine field relative to min corner; the constant shift does not change strain
    eps, _ = hex8.strain_fields(el, b, u, u, 1.0)
    assert np.max(np.abs(eps - e0)) < 1e-15
    c = OJ.elasticity_matrix(30e9, 0.2)
    ke = hex8.element_stiffness(b[0], jw[0], c)
    assert np.allclose(ke, ke.T, rtol=0, atol=1e-6 * np.max(np.abs(ke)))
    ev = np.linalg.eigvalsh(0.5 * (ke + ke.T))
    assert np.sum(np.abs(ev) < 1e-9 * np.max(ev)) == 6


def test_lumped_mass_total():
    nodes, el = hex8.structured_box((2.0, 1.0, 1.0), (4, 2, 3))
    m = hex8.lumped_mass(nodes, el, 2400.0)
    assert np.sum(m) == pytest.approx(3 * 2400.0 * 2.0, rel=1e-12)


@pytest.mark.parametrize("div", [(1, 1, 1), (2, 2, 2), (4, 4, 4), (3, 2, 5)])
def test_periodic_constraint(div):
    """test_rve.py:36-63."""
    e = np.asarray([0.01, -0.003, 0.002, 0.004, -0.006, 0.008])
    nodes, _ = hex8.structured_box((1.0, 1.0, 1.0), div)
    con = periodic.structured_constraint(nodes, div)
    cc = int(np.prod(div))
    assert np.unique(con.class_ids).size == cc
    assert con.transformation.shape == (3 * nodes.shape[0], 3 * (cc - 1))
    if 1 < cc <= 27:
        assert np.linalg.matrix_rank(con.transformation.toarray()) == 3 * (cc - 1)
    u = (con.affine @ e).reshape(-1, 3)
    a = con.affine.reshape(nodes.shape[0], 3, 6)
    for node, c in enumerate(con.class_ids):
        rep = con.representatives[c]
        assert np.allclose(u[node] - u[rep], (a[node] - a[rep]) @ e, atol=1e-14)
    # affine consistency: u = x . eps with 1/2 on engineering shear
    x = nodes - nodes.min(axis=0)
    exp = np.stack([x[:, 0] * e[0] + 0.5 * x[:, 2] * e[4] + 0.5 * x[:, 1] * e[5],
                    x[:, 1] * e[1] + 0.5 * x[:, 2] * e[3] + 0.5 * x[:, 0] * e[5],
                    x[:, 2] * e[2] + 0.5 * x[:, 1] * e[3] + 0.5 * x[:, 0] * e[4]], axis=1)
    assert np.allclose(u, exp, atol=1e-14)


@pytest.mark.parametrize("div", [1, 2, 4])
def test_homogeneous_rve_equals_point(div):
    """test_rve.py:66-130 and single_phase_j2_viscoplastic.yaml tolerances."""
    m = OR.Model((1.0, 1.0, 1.0), (div,) * 3, [("m", REF_MAT)], OR.Newton(armijo=OR.Armijo(True)),
                 hill_mandel_power_scale=10.0)
    res = OR.solve_path(m, 1e-3 * np.arange(5), np.vstack([np.zeros(6), REF_PATH]), SET)
    st = (np.zeros((1, 6)), np.zeros(1), np.zeros(1))
    prev = np.zeros(6)
    for e, r in zip(REF_PATH, res):
        p = OJ.j2vp_update(REF_MAT, e[None], *st, 1e-3, SET, ((e - prev) / 1e-3)[None])
        st = (p["plastic_strain"], p["equivalent_plastic_strain"], p["viscoplastic_multiplier"])
        prev = e
        assert np.linalg.norm(r["macro_stress"] - p["stress"][0]) / 10.0 <= 1e-10
        assert np.allclose(r["effective_tangent"], p["tangent"][0], rtol=1e-10, atol=1e-9)
        assert np.max(np.abs(r["state"]["material_state"]["plastic_strain"] - p["plastic_strain"])) <= 1e-12
        d = r["diagnostics"]
        assert d["periodic_fluctuation_error"] <= 1e-14
        assert d["hill_mandel_error"] <= 1e-12
        assert d["virtual_work_stress_error"] <= 1e-10
        assert d["reaction_antiperiodicity_error"] <= 1e-10


def _two_phase(div, same=False, plastic=True, seed=3):
    rng = np.random.default_rng(seed)
    eph = (rng.random(int(np.prod(div))) < 0.4).astype(np.int64)
    mort = OJ.J2VPParams(25e9, 0.2, 20e6, 0.5e9, 1e-3, 20e6, 1.0) if plastic else OJ.ElasticParams(25e9, 0.2)
    agg = mort if same else OJ.ElasticParams(50e9, 0.2)
    return OR.Model((1.0, 1.0, 1.0), div, [("mortar", mort), ("aggregate", agg)],
                    OR.Newton(20, 1e-6, 1e-11, 1e-16, 1e-11, OR.Armijo(True)), element_phase_ids=eph)


def test_directional_tangent_fd():
    """src/concrete_impact/benchmarks/rve.py:239-278: effective tangent vs central FD (1e-7)."""
    m = _two_phase((3, 3, 3))
    st = m.initial_state()
    e1 = np.asarray([1.5e-3, -3e-4, -2e-4, 4e-4, -1e-4, 6e-4])
    st = OR.solve_increment(m, e1, 1e-5, SET, st)["state"]
    target = 1.6 * e1
    r = OR.solve_increment(m, target, 1e-5, SET, st)
    d = np.asarray([0.31, -0.22, 0.17, 0.41, -0.53, 0.59])
    d /= np.linalg.norm(d)
    h = 1e-7 * 3e-3 / 1e-2
    plus = OR.solve_increment(m, target + h * d, 1e-5, SET, st, need_tangent=False)
    minus = OR.solve_increment(m, target - h * d, 1e-5, SET, st, need_tangent=False)
    num = (plus["macro_stress"] - minus["macro_stress"]) / (2 * h)
    ana = r["effective_tangent"] @ d
    assert np.linalg.norm(num - ana) / np.linalg.norm(ana) < 1e-5
    assert r["diagnostics"]["viscoplastic_active_volume_fraction"] > 0.1


def test_equal_phases_equal_homogeneous():
    """rve_heterogeneous.py:79-99: identical phases give the homogeneous response."""
    m = _two_phase((3, 3, 3), same=True)
    e = np.asarray([2e-3, -5e-4, -5e-4, 1e-3, 0.0, 5e-4])
    r = OR.solve_increment(m, e, 1e-5, SET, m.initial_state())
    mat = m.phases[0][1]
    p = OJ.j2vp_update(mat, e[None], np.zeros((1, 6)), np.zeros(1), np.zeros(1), 1e-5, SET)
    assert np.allclose(r["macro_stress"], p["stress"][0], rtol=1e-10, atol=1e-6)


def test_elastic_two_phase_within_voigt_reuss():
    """rve_heterogeneous.py:133-145: C_eff between the Reuss and Voigt bounds."""
    m = _two_phase((4, 4, 4), plastic=False)
    r = OR.solve_increment(m, np.asarray([1e-4, 0, 0, 0, 0, 0]), 1e-5, SET, m.initial_state())
    f = np.mean(m.element_phase_ids == 1)
    c1, c2 = OJ.elasticity_matrix(25e9, 0.2), OJ.elasticity_matrix(50e9, 0.2)
    voigt = (1 - f) * c1 + f * c2
    reuss = np.linalg.inv((1 - f) * np.linalg.inv(c1) + f * np.linalg.inv(c2))
    c = r["effective_tangent"]
    sym = 0.5 * (c + c.T)
    assert np.min(np.linalg.eigvalsh(voigt - sym)) > -1e-6 * np.max(c)
    assert np.min(np.linalg.eigvalsh(sym - reuss)) > -1e-6 * np.max(c)


def test_armijo_reference_cases():
    """tests/unit/fem/test_nonlinear_controls.py:22-58."""
    s = OR.Armijo(True)
    assert OR.armijo_search(np.asarray([1.0]), lambda a: np.asarray([1.0 - a]), s) == (1.0, 0, 0.0)
    assert OR.armijo_search(np.asarray([1.0]), lambda a: np.asarray([1.0 - 4 * a]), s) == (0.25, 2, 0.0)
    with pytest.raises(OR.ArmijoExhausted) as info:
        OR.armijo_search(np.asarray([1.0]), lambda a: np.asarray([1.0]), OR.Armijo(True, max_backtracks=2))
    assert info.value.diagnostics["max_backtracks"] == 2
    assert info.value.diagnostics["last_candidate_residual_norm"] == 1.0


def test_literal_loop_model_equals_vectorised():
    """The reference's literal per-point / per-element loops (oracle/loops.py, timed by
    the CPU baseline) give the vectorised model's increments to round-off."""
    from oracle import loops as OLP

    mat = OJ.J2VPParams(30e9, 0.2, 30e6, 1e9, 1e-3, 30e6, 1.0)
    kw = dict(lengths=(1.0, 1.0, 1.0), divisions=(2, 2, 3), phases=[("m", mat)],
              newton=OR.Newton(20, 1e-10 * 30e6, 1e-10, 1e-12, 1e-10, OR.Armijo(True)))
    vec, loop = OR.Model(**kw), OLP.LoopModel(**kw)
    sv, sl = vec.initial_state(), loop.initial_state()
    dt = 2e-7
    for k in range(1, 4):
        eps = np.asarray([1.2e-3 * k, -2e-4 * k, 1e-4, 3e-4 * k, 0.0, -1e-4 * k])
        rv = OR.solve_increment(vec, eps, dt, OJ.UpdateSettings(), sv)
        rl = OR.solve_increment(loop, eps, dt, OJ.UpdateSettings(), sl)
        np.testing.assert_allclose(rl["macro_stress"], rv["macro_stress"], rtol=1e-12, atol=1e-3)
        np.testing.assert_allclose(rl["effective_tangent"], rv["effective_tangent"], rtol=1e-10,
                                   atol=1e-3)
        for key, v in rv["state"]["material_state"].items():
            np.testing.assert_allclose(rl["state"]["material_state"][key], v, rtol=1e-12, atol=1e-18)
        assert rl["diagnostics"]["newton_iterations"] == rv["diagnostics"]["newton_iterations"]
        sv, sl = rv["state"], rl["state"]
