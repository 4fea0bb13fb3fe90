"""Pin the oracle's heterogeneous plastic RVE by the laminate recurrence.

A periodic RVE layered along x has a piecewise-uniform exact solution that the
Hex8 space contains (``oracle/laminate.py``); the oracle's SuperLU Newton
(``src/fem/rve/solver.py:62-227`` restated) must reproduce the per-phase
recurrence: macro stress, every Gauss point's strain, stress and plastic state,
along a 20-step plastic load / unload path, for rate exponents 1 and 2 and for
two plastic phases.  A planted mistake (the two phases' committed states
exchanged after step 5) must break the agreement.
"""

import numpy as np
import pytest

import laminate_cases as LC
from oracle import j2vp as OJ
from oracle import laminate as OL
from oracle import rve as OR

SET = OJ.UpdateSettings(50, 1e-10, 1e-12, 1e-10)


def _oracle_model(name):
    div, lengths, phases, fractions, eph = LC.layer_setup(name)
    named = [(f"p{i}", p) for i, p in enumerate(phases)]
    model = OR.Model(lengths, div, named, OR.Newton(25, 1e-7, 1e-13, 1e-16, 1e-13, OR.Armijo(True)),
                     element_phase_ids=eph)
    return model, phases, fractions, eph


def _check_against(model, phases, eph, lam, steps, tol):
    """Largest relative deviation of the oracle RVE from the recurrence."""
    state = model.initial_state()
    times, strains = steps
    pid = LC.point_phase_rows(eph)
    worst = {"stress": 0.0, "strain": 0.0, "state": 0.0}
    for k in range(1, strains.shape[0]):
        r = OR.solve_increment(model, strains[k], times[k] - times[k - 1], SET, state, need_tangent=False)
        ref = lam[k - 1]
        smax = np.max(np.abs(ref["macro_stress"]))
        worst["stress"] = max(worst["stress"], np.max(np.abs(r["macro_stress"] - ref["macro_stress"])) / smax)
        emax = np.max(np.abs(ref["phase_strain"]))
        worst["strain"] = max(worst["strain"],
                              np.max(np.abs(r["micro_strains"] - ref["phase_strain"][pid])) / emax)
        for p, par in enumerate(phases):
            if not isinstance(par, OJ.J2VPParams):
                continue
            for var in ("plastic_strain", "equivalent_plastic_strain", "viscoplastic_multiplier"):
                got = r["state"]["material_state"][f"phase_p{p}__{var}"]
                want = ref["phase_state"][p][var]
                want = np.repeat(want, got.shape[0], axis=0)
                worst["state"] = max(worst["state"],
                                     np.max(np.abs(got - want)) / max(np.max(np.abs(want)), 1e-6))
        state = r["state"]
    return worst


@pytest.mark.parametrize("name", sorted(LC.CASES))
def test_oracle_rve_equals_laminate_recurrence(name):
    model, phases, fractions, eph = _oracle_model(name)
    times, strains = LC.load_unload_path()
    lam = OL.laminate_path(phases, fractions, times, strains, SET)
    # the path must exercise plasticity in every J2 phase and unloading
    for p, par in enumerate(phases):
        if isinstance(par, OJ.J2VPParams):
            assert lam[-1]["phase_state"][p]["equivalent_plastic_strain"][0] > 1e-5
    worst = _check_against(model, phases, eph, lam, (times, strains), 1e-10)
    assert worst["stress"] < 1e-10, worst
    assert worst["strain"] < 1e-10, worst
    assert worst["state"] < 1e-9, worst


def test_planted_state_swap_breaks_the_laminate_check():
    """Exchanging the two J2 phases' committed states after step 5, when both
    have yielded by different amounts (a phase or state mapping mistake along
    the history), must be detected."""
    name = "mortar_soft_m12"
    model, phases, fractions, eph = _oracle_model(name)
    times, strains = LC.load_unload_path()
    bad = OL.laminate_path(phases, fractions, times, strains, SET, swap_states_after=5)
    worst = _check_against(model, phases, eph, bad, (times, strains), 1e-10)
    assert worst["stress"] > 1e-4 and worst["state"] > 1e-2, worst


def test_laminate_recurrence_elastic_limit_is_the_reuss_voigt_mix():
    """Two elastic phases: the out-of-plane response is the series (Reuss) mix and
    the in-plane response the parallel (Voigt) mix of the layer stiffnesses,
    a closed form independent of the recurrence's Newton."""
    a, b = OJ.ElasticParams(30.0e9, 0.2), OJ.ElasticParams(60.0e9, 0.3)
    f = np.asarray([0.3, 0.7])
    ca, cb = OJ.elasticity_matrix(a.young_modulus, a.poisson_ratio), OJ.elasticity_matrix(
        b.young_modulus, b.poisson_ratio)
    eps = np.asarray([1e-3, -4e-4, 2e-4, 3e-4, -5e-4, 6e-4])
    lam = OL.laminate_path([a, b], f, [0.0, 1.0], np.stack([np.zeros(6), eps]))
    # closed form: unknown out-of-plane strains solve continuity of C[o,:] e
    o, i = OL.OUT_OF_PLANE, np.asarray([1, 2, 3])
    # x_a, x_b with f_a x_a + f_b x_b = eps_o and ca[o,o] x_a + ca[o,i] eps_i = cb[o,o] x_b + cb[o,i] eps_i
    lhs = np.block([[f[0] * np.eye(3), f[1] * np.eye(3)], [ca[np.ix_(o, o)], -cb[np.ix_(o, o)]]])
    rhs = np.concatenate([eps[o], (cb[np.ix_(o, i)] - ca[np.ix_(o, i)]) @ eps[i]])
    x = np.linalg.solve(lhs, rhs)
    ea, eb = eps.copy(), eps.copy()
    ea[o], eb[o] = x[:3], x[3:]
    np.testing.assert_allclose(lam[0]["phase_strain"], np.stack([ea, eb]), rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(lam[0]["macro_stress"], f[0] * ca @ ea + f[1] * cb @ eb, rtol=1e-12)
    # equal phases: uniform strain
    lam2 = OL.laminate_path([a, a], f, [0.0, 1.0], np.stack([np.zeros(6), eps]))
    np.testing.assert_allclose(lam2[0]["phase_strain"], np.stack([eps, eps]), rtol=1e-13, atol=1e-19)
