"""Device production of the reference's fixed-cell plan paths
(concrete_impact_ai_amd.data_plan.produce_plan_responses): tasks of the c48 plan
grouped by increment count and advanced together by the batched FE² engine on an
8^3 voxel cell with the plan's phases; two of the paths (one per group, both
yielding) against the oracle's SuperLU path solve on the same cell (macro stress
1e-8, free energy and dissipation 1e-7 relative)."""

from pathlib import Path

import numpy as np
import pytest

from concrete_impact_ai_amd import data_plan as DP
from oracle import j2vp as OJ
from oracle import rve as OR

pytestmark = pytest.mark.gpu

PLAN = Path(__file__).resolve().parent / "golden" / "configs" / "rve_rno_c48_v2_2000.yaml"


def _rel(a, b):
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300)


def test_plan_paths_produced_in_batches_equal_the_oracle(gpu):
    plan = DP.load_rve_data_plan(PLAN)
    src = plan.sources[0]
    tasks = DP.build_rve_data_task_manifest(plan)
    pick = [t for t in tasks if t.accepted_increment_count == 64 and t.amplitude_band == "A5"
            and t.rate_band == "R2"][:4]
    pick += [t for t in tasks if t.accepted_increment_count == 96 and t.amplitude_band == "A4"
             and t.family == "nonproportional"][:2]
    model, eph = DP.fixed_cell_voxel_model(src, 8)
    out = DP.produce_plan_responses(src, pick, model)
    assert set(out) == {t.task_id for t in pick}
    for t in pick:
        r = out[t.task_id]
        assert r["macro_stress"].shape == (t.accepted_increment_count, 6)
        assert np.all(np.isfinite(r["macro_stress"]))
    mp = {k: v["value"] for k, v in src.material_parameters.items()}
    omodel = OR.Model(
        (1.0, 1.0, 1.0), (8, 8, 8),
        [("matrix", OJ.J2VPParams(mp["matrix_young_modulus"], mp["matrix_poisson_ratio"],
                                  mp["matrix_yield_stress"], mp["matrix_hardening_modulus"],
                                  mp["matrix_time_scale"], mp["matrix_yield_stress"],
                                  mp["matrix_rate_exponent"])),
         ("inclusion", OJ.ElasticParams(mp["inclusion_young_modulus"], mp["inclusion_poisson_ratio"]))],
        OR.Newton(25, 1e-10 * mp["matrix_yield_stress"], 1e-10, 1e-14, 1e-10,
                  OR.Armijo(True, 1e-4, 0.5, 20)),
        element_phase_ids=eph)
    settings = OJ.UpdateSettings(50, 1e-10, 1e-12, 1e-10)
    for t in (pick[0], pick[-1]):
        path = DP.generate_data_control_path(src, t)
        ref = OR.solve_path(omodel, path.times, path.values, settings, need_tangent=False)
        got = out[t.task_id]
        want = np.asarray([x["macro_stress"] for x in ref])
        assert _rel(got["macro_stress"], want) < 1e-8, t.task_id
        psi = np.asarray([x["free_energy_density"] for x in ref])
        dis = np.asarray([x["dissipation_density"] for x in ref])
        assert _rel(got["free_energy_density"], psi) < 1e-7
        assert _rel(got["dissipation_density"], dis) < 1e-7
        act = np.asarray([x["diagnostics"]["viscoplastic_active_volume_fraction"] for x in ref])
        assert np.max(act) > 0.3, t.task_id  # the path yields
        np.testing.assert_allclose(got["viscoplastic_active_volume_fraction"], act, atol=1e-12)
