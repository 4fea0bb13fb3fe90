"""GPU parity of the batched RVE solver (cia_rve_batch_solve).

* a batch of uneven two-phase RVEs (different divisions, lengths, seeds, paths
  and time steps) against the FP64 oracle's SuperLU Newton, RVE by RVE (stress
  1e-9 relative, state 1e-9) and against the single-RVE device solver
  (iteration counts equal, stress 1e-10)
* FP32 element work with FP64 residuals: error of the homogenised stress against
  the FP64 oracle (2e-6, as for the single-RVE solver) and of the effective
  tangent (2e-7)
* Newton failure reasons per RVE equal those of the single-RVE solver
* batched effective tangents (cia_rve_batch_effective_tangent, the Schur
  complement of solver.py:438-461) against the oracle's SuperLU tangent (1e-8)
  and the single-RVE device tangent (1e-9), along loading and unloading paths;
  NaN for the RVEs whose Newton solve failed
* the autotuned grids change no bit of the result
"""

import numpy as np
import pytest

from oracle import j2vp as OJ
from oracle import rve as OR

pytestmark = pytest.mark.gpu

SIZES = [((4, 4, 4), (1.0, 1.0, 1.0)), ((5, 3, 4), (1.0, 0.8, 1.2)), ((3, 3, 3), (1.0, 1.0, 1.0)),
         ((6, 5, 4), (0.9, 1.0, 1.1)), ((8, 8, 8), (1.0, 1.0, 1.0))]


def _settings():
    from concrete_impact_ai_amd import materials as M

    return M.MaterialUpdateSettings(50, 1e-10, 1e-12, 1e-10)


def _models(sizes, precision="fp64", newton_rtol=1e-11, max_its=20, adaptive=True):
    from concrete_impact_ai_amd import materials as M
    from concrete_impact_ai_amd import rve as R
    from concrete_impact_ai_amd.solvers import (ArmijoSettings, LinearSolverSettings,
                                                NonlinearNewtonSettings)

    newton = NonlinearNewtonSettings(max_its, 1e-6, newton_rtol, 1e-16, newton_rtol,
                                     ArmijoSettings(True, 1e-4, 0.5, 20))
    lin = LinearSolverSettings(method="spectral", adaptive_forcing=adaptive)
    mortar = M.J2ViscoplasticMaterial("mortar", 2400.0, 25.0e9, 0.2, 20.0e6, 0.5e9, 1.0e-3, 20.0e6, 1.0)
    agg = M.LinearElasticMaterial("aggregate", 2400.0, 50.0e9, 0.2)
    models, omodels = [], []
    for i, (n, lengths) in enumerate(sizes):
        rng = np.random.default_rng(100 + i)
        eph = (rng.random(n[0] * n[1] * n[2]) < 0.4).astype(np.int64)
        nodes, elements = R.build_structured_hex8_box(lengths, n)
        models.append(R.StructuredHex8RVE(
            nodes, elements, tuple(n),
            M.PhasedMaterialModel("rve", (M.RVEPhase("mortar", mortar), M.RVEPhase("aggregate", agg)),
                                  np.repeat(eph, 8), 2400.0),
            R.build_structured_hex8_periodic_constraint(nodes, n), newton, element_phase_ids=eph,
            phase_names=("mortar", "aggregate"), linear_settings=lin, precision=precision))
        omodels.append(OR.Model(lengths, tuple(n),
                                [("mortar", OJ.J2VPParams(25.0e9, 0.2, 20.0e6, 0.5e9, 1.0e-3, 20.0e6, 1.0)),
                                 ("aggregate", OJ.ElasticParams(50.0e9, 0.2))],
                                OR.Newton(max_its, 1e-6, newton_rtol, 1e-16, newton_rtol, OR.Armijo(True)),
                                element_phase_ids=eph))
    return models, omodels


def _paths(n_rves, steps, seed=7):
    """Per RVE: a loading ramp in a random direction, then partial unloading;
    a different time step per RVE."""
    rng = np.random.default_rng(seed)
    strains = np.zeros((n_rves, steps + 1, 6))
    for i in range(n_rves):
        d = rng.normal(size=6)
        d /= np.max(np.abs(d))
        for k in range(1, steps + 1):
            strains[i, k] = 3e-3 * d * min(k, steps - 1) / (steps - 1)
        strains[i, -1] = strains[i, -2] - 0.6e-3 * d
    dts = 10.0 ** rng.uniform(-6, -4, size=n_rves)
    times = np.cumsum(np.concatenate([np.zeros((n_rves, 1)), np.repeat(dts[:, None], steps, 1)], 1), 1)
    return times, strains


def _rel(a, b):
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300)


def test_batch_matches_oracle_and_single_rve(gpu):
    from concrete_impact_ai_amd import batch as B
    from concrete_impact_ai_amd import rve as R

    models, omodels = _models(SIZES)
    steps = 5
    times, strains = _paths(len(models), steps)
    batch = B.RVEBatch(models)
    bstate = batch.initial_state()
    caches = [R.build_rve_operator_cache(m) for m in models]
    sstates = [R.initialize_rve_state(m, c) for m, c in zip(models, caches)]
    ostates = [om.initial_state() for om in omodels]
    osettings = OJ.UpdateSettings(50, 1e-10, 1e-12, 1e-10)
    plastic = 0.0
    for k in range(1, steps + 1):
        dts = times[:, k] - times[:, k - 1]
        resp = B.solve_rve_batch(batch, strains[:, k], dts, _settings(), bstate)
        assert all(r == "ok" for r in resp.reasons)
        for i, (m, c) in enumerate(zip(models, caches)):
            single = R.solve_rve_microequilibrium(
                m, R.RVERequest(strains[i, k], float(dts[i]), _settings(),
                                R.MaterialResponseRequirements(tangent=False)),
                sstates[i], c, retain_micro_fields=False)
            ref = OR.solve_increment(omodels[i], strains[i, k], float(dts[i]), osettings, ostates[i],
                                     need_tangent=False)
            assert _rel(resp.macro_stress[i], ref["macro_stress"]) < 1e-9, (k, i)
            assert _rel(resp.macro_stress[i], single.macro_stress) < 1e-10, (k, i)
            assert resp.diagnostics["newton_iterations"][i] == single.diagnostics["newton_iterations"]
            assert resp.free_energy_density[i] == pytest.approx(ref["free_energy_density"], rel=1e-9)
            assert resp.dissipation_density[i] == pytest.approx(
                ref["dissipation_density"], rel=1e-8, abs=1e-9 * abs(ref["free_energy_density"]) + 1e-30)
            od = ref["diagnostics"]
            for key in ("macro_work_increment", "micro_work_increment",
                        "viscoplastic_active_volume_fraction", "unloading_volume_fraction",
                        "positive_work_density", "negative_work_density"):
                assert resp.diagnostics[key][i] == pytest.approx(
                    od[key], rel=1e-8, abs=1e-12 * max(1.0, abs(od[key]))), (key, i)
            assert resp.diagnostics["hill_mandel_error"][i] <= max(1e-9, 10 * od["hill_mandel_error"])
            rs = resp.state.rve(i)
            for key, val in ref["state"]["material_state"].items():
                got = rs.material_state.variables[key]
                assert np.max(np.abs(got - val)) <= 1e-9 * max(np.max(np.abs(val)), 1e-3), (key, i)
            assert _rel(rs.fluctuation_dofs, ref["state"]["fluctuation_dofs"]) < 1e-7
            plastic = max(plastic, od["viscoplastic_active_volume_fraction"])
            sstates[i], ostates[i] = single.state, ref["state"]
        bstate = resp.state
    assert plastic > 0.3


def test_batch_effective_tangent_matches_oracle_and_single_rve(gpu):
    from concrete_impact_ai_amd import batch as B
    from concrete_impact_ai_amd import rve as R

    models, omodels = _models(SIZES)
    steps = 4
    times, strains = _paths(len(models), steps, seed=9)
    batch = B.RVEBatch(models)
    bstate = batch.initial_state()
    caches = [R.build_rve_operator_cache(m) for m in models]
    sstates = [R.initialize_rve_state(m, c) for m, c in zip(models, caches)]
    ostates = [om.initial_state() for om in omodels]
    osettings = OJ.UpdateSettings(50, 1e-10, 1e-12, 1e-10)
    plastic = 0.0
    for k in range(1, steps + 1):
        dts = times[:, k] - times[:, k - 1]
        plain = B.solve_rve_batch(batch, strains[:, k], dts, _settings(), bstate)
        resp = B.solve_rve_batch(batch, strains[:, k], dts, _settings(), bstate, need_tangent=True)
        assert resp.effective_tangent.shape == (len(models), 6, 6)
        assert np.array_equal(resp.macro_stress, plain.macro_stress)
        assert np.all(resp.diagnostics["linear_iterations"] > plain.diagnostics["linear_iterations"])
        for i, (m, c) in enumerate(zip(models, caches)):
            single = R.solve_rve_microequilibrium(
                m, R.RVERequest(strains[i, k], float(dts[i]), _settings(),
                                R.MaterialResponseRequirements(tangent=True)),
                sstates[i], c, retain_micro_fields=False)
            ref = OR.solve_increment(omodels[i], strains[i, k], float(dts[i]), osettings, ostates[i],
                                     need_tangent=True)
            got = resp.effective_tangent[i]
            assert _rel(got, ref["effective_tangent"]) < 1e-8, (k, i)
            assert _rel(got, single.effective_tangent) < 1e-9, (k, i)
            plastic = max(plastic, ref["diagnostics"]["viscoplastic_active_volume_fraction"])
            sstates[i], ostates[i] = single.state, ref["state"]
        bstate = resp.state
    assert plastic > 0.3
    # the tangent of the last solve again: the same bits
    res = (B._lib.RVEResult * len(models))()
    status, again = B.effective_tangents_raw(batch, res)
    assert status == B._lib.CIA_OK
    assert np.array_equal(again, resp.effective_tangent)


def test_batch_fp32_error_against_fp64_oracle(gpu):
    from concrete_impact_ai_amd import batch as B

    sizes = [((6, 6, 6), (1.0, 1.0, 1.0)), ((5, 7, 6), (1.0, 1.2, 1.0)), ((8, 8, 8), (1.0, 1.0, 1.0))]
    models, omodels = _models(sizes, precision="fp32", newton_rtol=1e-6)
    steps = 6
    times, strains = _paths(len(models), steps, seed=11)
    batch = B.RVEBatch(models)
    st = batch.initial_state()
    ost = [om.initial_state() for om in omodels]
    worst = worst_t = 0.0
    for k in range(1, steps + 1):
        dts = times[:, k] - times[:, k - 1]
        resp = B.solve_rve_batch(batch, strains[:, k], dts, _settings(), st, need_tangent=True)
        for i, om in enumerate(omodels):
            ref = OR.solve_increment(om, strains[i, k], float(dts[i]), OJ.UpdateSettings(), ost[i],
                                     need_tangent=True)
            worst = max(worst, _rel(resp.macro_stress[i], ref["macro_stress"]))
            worst_t = max(worst_t, _rel(resp.effective_tangent[i], ref["effective_tangent"]))
            ost[i] = ref["state"]
        st = resp.state
    print(f"fp32 batch: stress {worst:.2e}, effective tangent {worst_t:.2e}", flush=True)
    assert worst < 2e-6, worst
    assert worst_t < 2e-7, worst_t


def test_batch_failure_reasons_match_single_rve(gpu):
    """Two Newton iterations at most: the RVEs that need more fail with the
    reference's reason, exactly as the single-RVE solver reports them."""
    from concrete_impact_ai_amd import batch as B
    from concrete_impact_ai_amd import rve as R

    models, _ = _models(SIZES, max_its=2, adaptive=False)
    batch = B.RVEBatch(models)
    eps = np.tile(np.asarray([3e-3, -1e-3, -5e-4, 1e-3, -2e-3, 5e-4]), (len(models), 1))
    eps[2] *= 0.01  # elastic: converges at once
    resp = B.solve_rve_batch(batch, eps, np.full(len(models), 1e-5), _settings(),
                             batch.initial_state(), raise_on_failure=False)
    expected = []
    for i, m in enumerate(models):
        c = R.build_rve_operator_cache(m)
        try:
            R.solve_rve_microequilibrium(m, R.RVERequest(eps[i], 1e-5, _settings(),
                                                         R.MaterialResponseRequirements(tangent=False)),
                                         R.initialize_rve_state(m, c), c, retain_micro_fields=False)
            expected.append("ok")
        except R.RVEEquilibriumError as e:
            expected.append(e.reason)
    assert list(resp.reasons) == expected
    assert "maximum_iterations_exceeded" in expected and "ok" in expected
    tr = B.solve_rve_batch(batch, eps, np.full(len(models), 1e-5), _settings(),
                           batch.initial_state(), raise_on_failure=False, need_tangent=True)
    assert list(tr.reasons) == expected
    for i, why in enumerate(expected):
        assert np.all(np.isfinite(tr.effective_tangent[i])) == (why == "ok"), i
    with pytest.raises(R.RVEEquilibriumError) as err:
        B.solve_rve_batch(batch, eps, np.full(len(models), 1e-5), _settings(), batch.initial_state())
    assert err.value.reason == "maximum_iterations_exceeded"
    assert err.value.diagnostics["rve_index"] == expected.index("maximum_iterations_exceeded")


def test_batch_autotuned_grids_change_no_bit(gpu):
    from concrete_impact_ai_amd import _lib
    from concrete_impact_ai_amd import batch as B

    models, _ = _models(SIZES)
    eps = np.asarray([[2e-3, -5e-4, -5e-4, 1e-3, 0.0, 5e-4]] * len(models))
    out = []
    for tune in (True, False):
        batch = B.RVEBatch(models, autotune=tune)
        if not tune:
            batch.set_tuning({k: 0 for k in _lib.BATCH_KERNELS})  # one tile per workgroup
        r = B.solve_rve_batch(batch, eps, np.full(len(models), 1e-5), _settings(), batch.initial_state())
        out.append((r.macro_stress, r.state.packed.cpu().numpy(), r.state.w.cpu().numpy()))
        if tune:
            tuning = batch.tuning()
            assert set(tuning) == set(_lib.BATCH_KERNELS)
            assert all(v["us"] > 0 for v in tuning.values())
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


def test_batch_predictor_reaches_the_same_states(gpu):
    """Starting Newton from the extrapolated fluctuation converges to the states of
    the reference start within the Newton tolerance, with no more Newton iterations."""
    from concrete_impact_ai_amd import batch as B

    models, _ = _models(SIZES[:3], newton_rtol=1e-12)
    steps = 6
    times, strains = _paths(len(models), steps, seed=21)
    outs = []
    for pred in (False, True):
        batch = B.RVEBatch(models)
        sig, its = [], 0
        for step in B.iterate_prescribed_rve_paths(batch, times, strains, _settings(), predictor=pred):
            sig.append(step.response.macro_stress)
            its += int(np.sum(step.response.diagnostics["newton_iterations"]))
        outs.append((np.asarray(sig), its, step.response.state.w.cpu().numpy()))
    (s0, n0, w0), (s1, n1, w1) = outs
    assert np.max(np.abs(s1 - s0)) <= 1e-9 * np.max(np.abs(s0))
    assert np.max(np.abs(w1 - w0)) <= 1e-7 * np.max(np.abs(w0))
    assert n1 <= n0
