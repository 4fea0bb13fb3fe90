"""CPU tests of the YAML config drop-in (``src/fem/io/config.py`` and the settings
builders of the reference's benchmark runners) and of the device view of a
general load function (trapezoidal external work of ``material.py:221-229``)."""

from pathlib import Path

import numpy as np
import pytest
import torch

from concrete_impact_ai_amd import benchmarks as B
from concrete_impact_ai_amd import config as CFG
from concrete_impact_ai_amd import dynamics as D
from concrete_impact_ai_amd import materials as M
from concrete_impact_ai_amd import solvers as S

GOLD = Path(__file__).resolve().parent / "golden" / "configs"


def test_reference_configs_load_into_settings():
    rve = CFG.load_yaml_config(GOLD / "single_phase_j2_viscoplastic.yaml")
    mat = M.build_material(rve["rve"]["material"])
    assert isinstance(mat, M.J2ViscoplasticMaterial) and mat.rate_exponent == 2.0
    ns = CFG.newton_settings(rve["rve"]["newton"])
    assert ns == S.NonlinearNewtonSettings(20, 1e-10, 1e-10, 1e-12, 1e-10,
                                           S.ArmijoSettings(True, 1e-4, 0.5, 20))
    assert CFG.material_update_settings(rve["rve"]["material_update"]) == M.MaterialUpdateSettings(
        50, 1e-10, 1e-12, 1e-10)
    dyn = CFG.load_yaml_config(GOLD / "dynamic_j2_viscoplastic_pressure_pulse_3d.yaml")
    te = CFG.time_integration_settings(dyn["analysis"]["explicit"])
    ti = CFG.time_integration_settings(dyn["analysis"]["implicit"])
    assert (te.time_step, te.num_steps, te.cfl_safety_factor) == (5e-4, 80, 0.8)
    assert (ti.beta, ti.gamma, ti.adaptive.enabled) == (0.25, 0.5, False)
    assert CFG.newton_settings(dyn["analysis"]["implicit"]["newton"]).residual_relative_tolerance == 1e-8
    mp = CFG.load_yaml_config(GOLD / "plastic_j2_viscoplastic_material_point.yaml")
    assert M.build_material(mp["model"]["material"]).time_scale == 30.0


def test_yaml_round_trip(tmp_path):
    data = CFG.load_yaml_config(GOLD / "single_phase_j2_viscoplastic.yaml")
    out = CFG.write_yaml_config(data, tmp_path / "sub" / "c.yaml")
    assert CFG.load_yaml_config(out) == data


def test_unknown_material_and_config_are_rejected():
    with pytest.raises(ValueError):
        M.build_material({"type": "drucker_prager_cap", "name": "x"})
    with pytest.raises(ValueError):
        B.run_benchmark({"case": {"name": "x"}})


def test_pure_shear_exact_matches_the_oracle_closed_form():
    """The runner's exact curve is the reference's recurrence (plasticity.py:642-679);
    the oracle's restatement of the same closed form agrees."""
    from oracle import j2vp as OJ

    cfg = CFG.load_yaml_config(GOLD / "plastic_j2_viscoplastic_material_point.yaml")
    mat = M.build_material(cfg["model"]["material"])
    g = mat.young_modulus / (2 * (1 + mat.poisson_ratio))
    shear = np.linspace(0.0, 0.023, 121)
    st = {"plastic_shear_strain": 0.0, "equivalent_plastic_strain": 0.0}
    ours = np.asarray([B._pure_shear_exact(mat, g, s, 2.0, st)["shear_stress"] for s in shear])
    ref = OJ.pure_shear_recurrence(OJ.J2VPParams(mat.young_modulus, mat.poisson_ratio, 10.0, 80.0, 30.0,
                                                 10.0, 1.0), shear, 2.0)
    np.testing.assert_allclose(ours, ref["shear_stress"], rtol=1e-14, atol=1e-12)


def test_device_load_work_for_separable_and_general_loads():
    n = 12
    unit = np.linspace(-1.0, 2.0, n)
    sep = D.SeparableLoad(unit, lambda t: 3.0 * t)
    gen = lambda t: 3.0 * t * unit  # noqa: E731  same load as a plain callable
    du = torch.linspace(0.5, -0.25, n, dtype=torch.float64)
    a = D._DeviceLoad(sep, n, torch.device("cpu"))
    b = D._DeviceLoad(gen, n, torch.device("cpu"))
    assert a.separable and not b.separable
    want = 0.5 * (3.0 * 0.1 + 3.0 * 0.3) * float(unit @ du.numpy())
    assert a.work(0.1, 0.3, du) == pytest.approx(want, rel=1e-14)
    assert b.work(0.1, 0.3, du) == pytest.approx(want, rel=1e-14)
    vec, s = b.at(0.3)
    np.testing.assert_allclose(s * vec.numpy(), gen(0.3), rtol=1e-15)
    with pytest.raises(ValueError):
        D._DeviceLoad(lambda t: np.zeros(3), n, torch.device("cpu")).host(0.0)


def test_direct_linear_solver_maps_onto_the_pcg():
    from concrete_impact_ai_amd import rve as R

    assert R.linear_method(S.LinearSolverSettings(backend="scipy", method="direct")) == "spectral"
    assert R.linear_method(S.LinearSolverSettings(method="block_jacobi")) == "block_jacobi"
    with pytest.raises(ValueError):
        R.linear_method(S.LinearSolverSettings(method="cholmod"))
