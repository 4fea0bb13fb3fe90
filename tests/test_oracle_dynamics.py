"""Pins of the oracle's explicit dynamics (CPU only)."""

import numpy as np
import pytest

from oracle import dynamics as OD, j2vp as OJ


def _bar(phases, eph=None, div=(8, 2, 2), lengths=(1.0, 0.25, 0.25)):
    return OD.Box(lengths, div, phases, eph)


def test_elastic_energy_balance_and_reference_bound():
    """Elastic slab under a half-sine pulse: K + psi - W_ext stays ~0 (Verlet, O(dt^2))
    and within the reference's explicit acceptance bound 0.05 (test_plasticity_benchmarks.py:58-75)."""
    mat = OJ.ElasticParams(3339.9160679991464, 0.36994096308414537, density=1.0)
    box = _bar([("m", mat)], div=(6, 3, 3), lengths=(0.6, 0.3, 0.3))
    faces = OD.box_faces(box.divisions)
    fixed = (3 * faces["bottom"][:, None] + np.arange(3)).ravel()
    funit = OD.top_patch_unit_pressure(box.lengths, box.divisions, ((1, 5), (0, 3)))
    c = np.sqrt((3339.9 / (3 * (1 - 2 * 0.37)) + 4 / 3 * 3339.9 / (2 * 1.37)) / 1.0)
    dt = 0.25 * 0.1 / c
    out = OD.explicit(box, dt, 120, fixed, funit, OD.half_sine(5.0, 60 * dt), OJ.UpdateSettings())
    w = out["external_work"]
    assert w[-1] > 0.0
    assert np.max(np.abs(out["energy_residual"])) < 0.05 * np.max(np.abs(w))
    assert np.max(np.abs(out["energy_residual"])) < 2e-3 * np.max(np.abs(w))


def test_viscoplastic_dissipation_and_plastic_flow():
    mat = OJ.J2VPParams(3339.9160679991464, 0.36994096308414537, 10.0, 80.0, 0.02, 10.0, 2.0, 1.0)
    box = _bar([("m", mat)], div=(6, 3, 3), lengths=(0.6, 0.3, 0.3))
    faces = OD.box_faces(box.divisions)
    fixed = (3 * faces["bottom"][:, None] + np.arange(3)).ravel()
    funit = OD.top_patch_unit_pressure(box.lengths, box.divisions, ((1, 5), (0, 3)))
    dt = 5e-4
    out = OD.explicit(box, dt, 80, fixed, funit, OD.half_sine(60.0, 40 * dt), OJ.UpdateSettings())
    assert np.all(np.diff(out["dissipated"]) >= -1e-14)
    assert out["dissipated"][-1] > 0.0
    assert np.max(out["state"]["equivalent_plastic_strain"]) > 0.0
    assert np.max(np.abs(out["energy_residual"])) < 0.05 * np.max(np.abs(out["external_work"]))


def test_free_bar_momentum_is_conserved():
    """No supports, no load, initial velocity field: the total momentum sum m v is constant."""
    mat = OJ.ElasticParams(1000.0, 0.25, density=2.0)
    box = _bar([("m", mat)], div=(5, 2, 2), lengths=(1.0, 0.4, 0.4))
    n = box.n_dof
    from oracle import dynamics as D

    # one step manually with an initial velocity: reuse explicit() with a zero load by editing v0
    u = np.zeros(n)
    v = np.zeros(n)
    v[0::3] = np.linspace(-1, 1, n // 3) * 1e-3
    state = box.init_state()
    asm = box.assemble(u, u, 1e-4, OJ.UpdateSettings(), state)
    a = -asm["internal_force"] / box.mass
    p0 = np.sum((box.mass * v).reshape(-1, 3), axis=0)
    dt = 1e-4
    for _ in range(20):
        un = u + dt * v + 0.5 * dt * dt * a
        asm = box.assemble(un, u, dt, OJ.UpdateSettings(), state)
        an = -asm["internal_force"] / box.mass
        v = v + 0.5 * dt * (a + an)
        u, a = un, an
    p1 = np.sum((box.mass * v).reshape(-1, 3), axis=0)
    assert np.allclose(p1, p0, atol=1e-15)


def test_cfl_rejection():
    mat = OJ.ElasticParams(1000.0, 0.25, density=1.0)
    box = _bar([("m", mat)])
    c = np.sqrt((1000 / (3 * 0.5) + 4 / 3 * 400) / 1.0)
    with pytest.raises(ValueError):
        OD.explicit(box, 10.0 * 0.125 / c, 1, np.arange(3), np.zeros(box.n_dof), lambda t: 0.0,
                    OJ.UpdateSettings(), cfl=0.5, wave_speed=c)


def _slab(mat, density=None, eph=None, phases=None):
    phases = phases or [("m", mat)]
    return OD.Box((0.6, 0.3, 0.3), (6, 3, 3), phases, eph, density=density)


def _slab_load(box):
    faces = OD.box_faces(box.divisions)
    fixed = (3 * faces["bottom"][:, None] + np.arange(3)).ravel()
    funit = OD.top_patch_unit_pressure(box.lengths, box.divisions, ((1, 5), (0, 3)))
    return fixed, funit


def test_consistent_mass_row_sums_and_total():
    """Row sums of the consistent mass are the lumped mass (elasticity.py:73-77) and the
    total is 3 rho V (one rigid translation per component)."""
    from oracle import hex8

    box = _slab(OJ.ElasticParams(1.0, 0.25, density=2.5))
    M = hex8.consistent_mass(box.nodes, box.elements, 2.5)
    np.testing.assert_allclose(np.asarray(M.sum(axis=1)).ravel(), box.mass, rtol=1e-14)
    assert abs(M.sum() - 3 * 2.5 * 0.6 * 0.3 * 0.3) < 1e-13
    # symmetric positive definite: smallest eigenvalue of a 1-element mass is rho V/216
    one = OD.Box((1.0, 1.0, 1.0), (1, 1, 1), [("m", OJ.ElasticParams(1.0, 0.25, density=1.0))])
    Me = hex8.consistent_mass(one.nodes, one.elements, 1.0).toarray()
    assert np.allclose(Me, Me.T)
    assert abs(np.linalg.eigvalsh(Me).min() - 1.0 / 216.0) < 1e-15


def test_newmark_average_acceleration_conserves_elastic_energy():
    """Average-acceleration Newmark (beta 1/4, gamma 1/2) on a linear elastic body
    satisfies K + psi - W_ext = 0 to solver round-off with the trapezoidal work
    measure of material.py:497-499; the reference bound is 0.01 of the work
    (test_plasticity_benchmarks.py:58-75).  A wrong beta breaks it."""
    from oracle.rve import Armijo, Newton

    mat = OJ.ElasticParams(3339.9160679991464, 0.36994096308414537, density=1.0)
    box = _slab(mat)
    fixed, funit = _slab_load(box)
    newton = Newton(20, 1e-9, 1e-10, 1e-14, 1e-10, Armijo(True))
    out = OD.newmark(box, 2e-3, 30, fixed, funit, OD.half_sine(5.0, 0.02), OJ.UpdateSettings(), newton)
    w = np.max(np.abs(out["external_work"]))
    assert w > 0.0
    assert np.max(np.abs(out["energy_residual"])) < 1e-9 * w
    bad = OD.newmark(box, 2e-3, 30, fixed, funit, OD.half_sine(5.0, 0.02), OJ.UpdateSettings(),
                     newton, beta=0.3, gamma=0.6)
    assert np.max(np.abs(bad["energy_residual"])) > 1e-4 * w


def test_newmark_viscoplastic_bound_and_dissipation():
    from oracle.rve import Armijo, Newton

    mat = OJ.J2VPParams(3339.9160679991464, 0.36994096308414537, 10.0, 80.0, 0.02, 10.0, 2.0, 1.0)
    box = _slab(mat)
    fixed, funit = _slab_load(box)
    newton = Newton(20, 1e-9, 1e-10, 1e-14, 1e-10, Armijo(True))
    out = OD.newmark(box, 1e-3, 40, fixed, funit, OD.half_sine(60.0, 0.02), OJ.UpdateSettings(), newton)
    assert out["dissipated"][-1] > 0.0
    assert np.all(np.diff(out["dissipated"]) >= -1e-14)
    assert np.max(np.abs(out["energy_residual"])) < 0.01 * np.max(np.abs(out["external_work"]))


def test_time_step_rules_match_the_reference_cases():
    """``test_convergence_time_step_rules_respect_configured_bounds``
    (tests/unit/fem/test_nonlinear_controls.py:61-68)."""
    ad = OD.Adaptive(True)
    assert OD.cut_back(1.0, 1.0, ad) == 0.5
    assert OD.cut_back(0.0625, 1.0, ad) == 0.0625
    assert OD.grow(1.0, 1.0, ad) == 2.0
    assert OD.grow(2.0, 1.0, ad) == 2.0


def _adaptive_case(max_iterations, adaptive, steps=30):
    from oracle.rve import Armijo, Newton

    mat = OJ.J2VPParams(3339.9160679991464, 0.36994096308414537, 10.0, 80.0, 0.02, 10.0, 2.0, 1.0)
    box = _slab(mat)
    fixed, funit = _slab_load(box)
    newton = Newton(max_iterations, 1e-9, 1e-10, 1e-14, 1e-10, Armijo(True))
    return OD.newmark(box, 1e-3, steps, fixed, funit, OD.half_sine(60.0, 0.02), OJ.UpdateSettings(),
                      newton, adaptive=adaptive)


def test_newmark_adaptive_cut_back_and_growth():
    """material.py:377-430: with three Newton iterations allowed, steps that need
    more are rejected and retried at half the step; after three easy accepted
    steps the step doubles (at most twice the initial one); the last step is cut
    to reach the final time.  The controller's decisions are replayed from the
    recorded Newton counts and rejections, and the plastic run stays within the
    reference's implicit energy bound (0.01 of the work)."""
    ad = OD.Adaptive(True)
    out = _adaptive_case(3, ad)
    h = out["time_steps"][1:]
    assert out["times"][-1] == pytest.approx(0.03, rel=1e-14)
    assert np.any(out["rejected_counts"] > 0) and np.any(h > 1e-3) and np.any(h < 1e-3)
    assert set(r for rr in out["rejected_reasons"] for r in rr) == {"maximum_iterations_exceeded"}
    # replay the controller
    cur, easy, t = 1e-3, 0, 0.0
    for k in range(h.size):
        trial = min(cur, 0.03 - t)
        for _ in range(out["rejected_counts"][k + 1]):
            trial = OD.cut_back(trial, 1e-3, ad)
        trial = min(trial, 0.03 - t)
        assert h[k] == pytest.approx(trial, rel=1e-12), k
        t += h[k]
        is_easy = (out["rejected_counts"][k + 1] == 0 and out["armijo_backtracks"][k + 1] == 0
                   and out["newton_iterations"][k + 1] <= ad.easy_iteration_limit)
        easy = easy + 1 if is_easy else 0
        if easy >= ad.easy_step_count:
            cur, easy = OD.grow(h[k], 1e-3, ad), 0
        else:
            cur = h[k]
    assert np.max(np.abs(out["energy_residual"])) < 0.01 * np.max(np.abs(out["external_work"]))
    # without the controller the same Newton limit fails the run
    with pytest.raises(OD.NewmarkStepError) as err:
        _adaptive_case(3, OD.Adaptive(False))
    assert err.value.reason == "maximum_iterations_exceeded"


def test_newmark_adaptive_without_failures_equals_fixed_steps():
    fixed = _adaptive_case(20, OD.Adaptive(False), steps=12)
    still = _adaptive_case(20, OD.Adaptive(True, easy_step_count=10**6), steps=12)
    for key in ("u", "v", "a", "stress", "dissipated"):
        assert np.array_equal(fixed[key], still[key]), key
