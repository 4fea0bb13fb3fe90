"""Pins of the oracle's material-point restatement (CPU only).

Each check is independent of the oracle's own formula: closed forms,
finite differences, bisection, energy identities and the reference's own
test cases (tests/unit/fem/materials/test_plasticity.py).
"""

import numpy as np
import pytest

from oracle import j2vp as OJ

REF = dict(young_modulus=3339.9160679991464, poisson_ratio=0.36994096308414537, yield_stress=10.0,
           hardening_modulus=80.0)
SET = OJ.UpdateSettings(50, 1e-10, 1e-12, 1e-10)


def mat(tau=0.02, m=2.0, sref=10.0, **kw):
    p = dict(REF)
    p.update(kw)
    return OJ.J2VPParams(p["young_modulus"], p["poisson_ratio"], p["yield_stress"],
                         p["hardening_modulus"], tau, sref, m)


def one(m_, strain, state=None, dt=1e-3, settings=SET, rates=None):
    z = np.zeros((1, 6))
    ep, q, g = state if state is not None else (z, np.zeros(1), np.zeros(1))
    return OJ.j2vp_update(m_, np.asarray(strain, dtype=float).reshape(1, 6), ep, q, g, dt, settings,
                          rates)


def test_pure_shear_closed_form_reference_config():
    """plastic_j2_viscoplastic_material_point.yaml: 121 steps to 0.023, dt 2, tau 30, m 1."""
    m_ = mat(tau=30.0, m=1.0)
    gam = np.linspace(0.0, 0.023, 121)
    exact = OJ.pure_shear_recurrence(m_, gam, 2.0)
    st = (np.zeros((1, 6)), np.zeros(1), np.zeros(1))
    for k, g in enumerate(gam):
        r = one(m_, [0, 0, 0, 0, 0, g], st, dt=2.0)
        st = (r["plastic_strain"], r["equivalent_plastic_strain"], r["viscoplastic_multiplier"])
        assert r["stress"][0, 5] == pytest.approx(exact["shear_stress"][k], rel=1e-10, abs=1e-9)
        assert r["diagnostics"]["plastic_multiplier"][0] == pytest.approx(
            exact["plastic_multiplier"][k], rel=1e-10, abs=1e-9)
        assert r["equivalent_plastic_strain"][0] == pytest.approx(
            exact["equivalent_plastic_strain"][k], rel=1e-10, abs=1e-9)


def _uniaxial_recurrence(E, nu, sy, H, tau, sref, dt, strains):
    """Textbook radial return for uniaxial strain (Simo & Hughes): plastic strain
    p (1, -1/2, -1/2), q_trial = 3G |2e/3 - p|, sigma_xx = K e + 2G (2e/3 - p)."""
    G = E / (2 * (1 + nu))
    K = E / (3 * (1 - 2 * nu))
    p = qb = 0.0
    out = []
    for e in strains:
        s = 2 * G * (2 * e / 3 - p)
        qt = 1.5 * abs(s)
        f = qt - (sy + H * qb)
        if f > 1e-10 * max(sy, sref, qt):
            dg = f / (tau * sref / dt + 3 * G + H)
            p += np.sign(s) * dg
            qb += dg
        out.append((K * e + 2 * G * (2 * e / 3 - p), p, qb))
    return np.asarray(out)


def test_uniaxial_strain_c1_recurrence():
    """The C1 material (E 30 GPa, nu 0.2, sigma_y 30 MPa, H 1 GPa, eta 1e-3 s, m 1)
    ramped to 2e-3 at 100 1/s over 100 steps."""
    E, nu, sy, H, tau = 30e9, 0.2, 30e6, 1e9, 1e-3
    m_ = OJ.J2VPParams(E, nu, sy, H, tau, sy, 1.0)
    strains = 2e-3 * np.arange(1, 101) / 100
    dt = 2e-5 / 100 * 1  # 100 1/s: de = 2e-5 per step -> dt = 2e-7 s
    dt = 2e-7
    exact = _uniaxial_recurrence(E, nu, sy, H, tau, sy, dt, strains)
    st = (np.zeros((1, 6)), np.zeros(1), np.zeros(1))
    for k, e in enumerate(strains):
        r = one(m_, [e, 0, 0, 0, 0, 0], st, dt=dt)
        st = (r["plastic_strain"], r["equivalent_plastic_strain"], r["viscoplastic_multiplier"])
        assert r["stress"][0, 0] == pytest.approx(exact[k, 0], rel=1e-12)
        assert r["plastic_strain"][0, 0] == pytest.approx(exact[k, 1], rel=1e-10, abs=1e-18)
        assert r["plastic_strain"][0, 1] == pytest.approx(-0.5 * exact[k, 1], rel=1e-10, abs=1e-18)
        assert r["equivalent_plastic_strain"][0] == pytest.approx(exact[k, 2], rel=1e-10, abs=1e-18)
    assert exact[-1, 2] > 0.0  # the path yields


def _bisect_root(f, hs, dt, tau, sref, m):
    lo = np.longdouble(0.0)
    hi = np.longdouble(f) / np.longdouble(hs)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        r = mid - np.longdouble(dt) * ((np.longdouble(f) - np.longdouble(hs) * mid) / sref) ** m / tau
        if r > 0:
            hi = mid
        else:
            lo = mid
    return float(0.5 * (lo + hi))


@pytest.mark.parametrize("m", [2.0, 3.5])
def test_newton_multiplier_matches_bisection(m):
    m_ = mat(m=m)
    rng = np.random.default_rng(1)
    strains = rng.normal(scale=0.01, size=(200, 6))
    r = OJ.j2vp_update(m_, strains, np.zeros((200, 6)), np.zeros(200), np.zeros(200), 1e-3, SET)
    G = OJ.shear_modulus(m_.young_modulus, m_.poisson_ratio)
    hs = 3 * G + m_.hardening_modulus
    pl = r["plastic_mask"]
    assert pl.sum() > 50
    for i in np.flatnonzero(pl):
        f = r["diagnostics"]["trial_yield_value"][i]
        root = _bisect_root(f, hs, 1e-3, m_.time_scale, m_.reference_stress, m)
        tol = max(1e-12, 1e-10 * f / hs)
        assert abs(r["diagnostics"]["plastic_multiplier"][i] - root) <= 2 * tol
        assert abs(r["diagnostics"]["local_residual"][i]) <= tol


def _fd_tangent(m_, strain, dt):
    t = np.zeros((6, 6))
    for c in range(6):
        h = 1e-8 * max(1.0, abs(strain[c]))
        ep = strain.copy()
        em = strain.copy()
        ep[c] += h
        em[c] -= h
        t[:, c] = (one(m_, ep, dt=dt)["stress"][0] - one(m_, em, dt=dt)["stress"][0]) / (2 * h)
    return t


@pytest.mark.parametrize("m,tau,dt", [(1.0, 30.0, 2.0), (2.0, 0.02, 1e-3)])
def test_tangent_matches_finite_differences(m, tau, dt):
    """test_plasticity.py:179-256 (rtol 1e-5, atol 1e-6) plus symmetry."""
    m_ = mat(tau=tau, m=m)
    strain = np.asarray([0.004, -0.0015, -0.001, 0.002, -0.003, 0.02])
    r = one(m_, strain, dt=dt)
    assert r["diagnostics"]["active_surface"][0] == 4.0
    t = r["tangent"][0]
    assert np.allclose(t, _fd_tangent(m_, strain, dt), rtol=1e-5, atol=1e-6)
    assert np.allclose(t, t.T, rtol=0, atol=1e-9 * np.max(np.abs(t)))


def test_elastic_branch_tangent_is_elasticity():
    m_ = mat()
    r = one(m_, [1e-5, 0, 0, 0, 0, 1e-5])
    assert r["diagnostics"]["viscoplastic_active"][0] == 0.0
    assert np.array_equal(r["tangent"][0], OJ.elasticity_matrix(REF["young_modulus"], REF["poisson_ratio"]))


def test_energy_identities_quadratic_rate():
    """psi = 1/2 sigma.(eps - eps_p) + 1/2 H q^2 and D = power - storage >= 0 (1e-12)."""
    m_ = mat(m=2.0)
    strain = np.asarray([0.004, -0.0015, -0.001, 0.002, -0.003, 0.02])
    r = one(m_, strain)
    ee = strain - r["plastic_strain"][0]
    q = r["equivalent_plastic_strain"][0]
    psi = 0.5 * r["stress"][0] @ ee + 0.5 * m_.hardening_modulus * q**2
    assert r["free_energy"][0] == pytest.approx(psi, rel=1e-12)
    d = r["diagnostics"]
    assert r["dissipation"][0] >= 0.0
    assert r["dissipation"][0] == pytest.approx(d["plastic_power"][0] - d["hardening_storage_rate"][0],
                                                rel=1e-12, abs=1e-12)
    assert d["local_residual"][0] == pytest.approx(0.0, abs=1e-12)
    assert 0.0 < d["plastic_multiplier"][0] < d["root_upper_bound"][0]


def test_activation_uses_relative_stress_scale():
    """test_plasticity.py:89-138."""
    m_ = mat()
    G = OJ.shear_modulus(REF["young_modulus"], REF["poisson_ratio"])
    settings = OJ.UpdateSettings(50, 1e-3, 1e-12, 1e-10)
    out = []
    for over in (5e-3, 2e-2):
        s = np.zeros(6)
        s[5] = (10.0 + over) / (np.sqrt(3.0) * G)
        out.append(one(m_, s, settings=settings)["diagnostics"])
    el, pl = out
    assert 0.0 < el["trial_yield_value"][0] < el["yield_activation_threshold"][0]
    assert el["plastic_multiplier"][0] == 0.0 and el["yield_activation_margin"][0] < 0.0
    assert el["viscoplastic_active"][0] == 0.0
    assert pl["trial_yield_value"][0] > pl["yield_activation_threshold"][0]
    assert pl["plastic_multiplier"][0] > 0.0 and pl["yield_activation_margin"][0] > 0.0
    assert pl["viscoplastic_active"][0] == 1.0


def test_multiplier_grows_with_time_step():
    """test_plasticity.py:63-86."""
    m_ = mat(tau=30.0, m=1.0)
    s = [0, 0, 0, 0, 0, 0.02]
    small = one(m_, s, dt=0.5)
    large = one(m_, s, dt=5.0)
    assert large["diagnostics"]["viscoplastic_multiplier"][0] > small["diagnostics"]["viscoplastic_multiplier"][0]
    assert large["stress"][0, 5] < small["stress"][0, 5]


@pytest.mark.parametrize("case,reason", [
    ("maxit", "maximum_iterations_exceeded"),
    ("nonfinite", "non_finite_iteration_value"),
    ("nan", "non_finite_j2_invariant"),
])
def test_failure_reasons(case, reason):
    """test_plasticity.py:259-345."""
    m_ = mat(sref=1e-300) if case == "nonfinite" else mat()
    strain = np.asarray([0.004, -0.0015, -0.001, 0.002, -0.003, 0.02])
    settings = OJ.UpdateSettings(1, 1e-10, 1e-30, 1e-30) if case == "maxit" else SET
    if case == "nan":
        strain = np.zeros(6)
        strain[5] = np.nan
    with pytest.raises(OJ.OracleMaterialError) as info:
        one(m_, strain, settings=settings)
    assert info.value.reason == reason
    assert info.value.point_id == 0


def test_planted_sign_error_is_caught(monkeypatch):
    """A plausible mistake (tensor instead of engineering plastic strain) fails the
    pure-shear pin: the pins bite."""
    monkeypatch.setattr(OJ, "T2E", np.eye(6))
    m_ = mat(tau=30.0, m=1.0)
    gam = np.linspace(0.0, 0.023, 121)
    exact = OJ.pure_shear_recurrence(m_, gam, 2.0)
    st = (np.zeros((1, 6)), np.zeros(1), np.zeros(1))
    worst = 0.0
    for k, g in enumerate(gam):
        r = one(m_, [0, 0, 0, 0, 0, g], st, dt=2.0)
        st = (r["plastic_strain"], r["equivalent_plastic_strain"], r["viscoplastic_multiplier"])
        worst = max(worst, abs(r["stress"][0, 5] - exact["shear_stress"][k]) / abs(exact["shear_stress"][k] + 1e-30))
    assert worst > 1e-3
