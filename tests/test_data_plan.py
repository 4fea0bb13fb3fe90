"""The reference's fixed-cell data plan (configs/experiments/rve_rno_c48_v2_2000.yaml,
copied to tests/golden/configs/) read and expanded by concrete_impact_ai_amd.data_plan.

Pinned by
* the reference's own plan tests (tests/unit/concrete_impact/test_fixed_cell_rve_plan.py:
  113-160, 206-218, 248-251): 2000 unique tasks in 100 strata of 20, 200 snapshots, 100
  tangent checks, exact 1400 / 300 / 300 and per-stratum 14 / 3 / 3 splits; the
  production path's peak norm equals its amplitude, its peak rate the target rate
  (from below, 2e-3) and its largest turning angle the task's angle; the 1000-path
  supplement with its 700 / 150 / 150 split and no reused task ids; the parallel budget;
* the reference's sample shard (tests/golden/rve_rno_c48_sample_response.h5), whose
  path metadata (seed 72204, amplitude 4.575e-4, target rate 0.6175, train split, no
  snapshot or tangent check) and whose stored macro strain history must be the
  manifest's task and the generated path, to rounding;
* the validators' error messages on broken designs.
"""

import dataclasses
import json
from pathlib import Path

import numpy as np
import pytest

from concrete_impact_ai_amd import data_plan as DP
from concrete_impact_ai_amd import hdf5 as H
from concrete_impact_ai_amd.config import load_yaml_config

GOLDEN = Path(__file__).resolve().parent / "golden"
PLAN = GOLDEN / "configs" / "rve_rno_c48_v2_2000.yaml"
SAMPLE = GOLDEN / "rve_rno_c48_sample_response.h5"
W = np.asarray([1.0, 1.0, 1.0, 0.5, 0.5, 0.5])


@pytest.fixture(scope="module")
def plan():
    return DP.load_rve_data_plan(PLAN)


@pytest.fixture(scope="module")
def tasks(plan):
    return DP.build_rve_data_task_manifest(plan)


def _count(items, key):
    out = {}
    for t in items:
        out[getattr(t, key)] = out.get(getattr(t, key), 0) + 1
    return out


def test_factorial_counts_and_frozen_splits(tasks):
    assert len(tasks) == 2000
    assert len({t.task_id for t in tasks}) == 2000
    assert len({t.design_stratum for t in tasks}) == 100
    assert sum(t.snapshot_required for t in tasks) == 200
    assert sum(t.tangent_check_required for t in tasks) == 100
    assert _count(tasks, "preassigned_split") == {"train": 1400, "validation": 300, "test": 300}
    for stratum in {t.design_stratum for t in tasks}:
        sel = [t for t in tasks if t.design_stratum == stratum]
        assert len(sel) == 20
        assert _count(sel, "preassigned_split") == {"train": 14, "validation": 3, "test": 3}
        # amplitudes fill their band's 20 cell midpoints, rates a permutation of them
        band = {t.amplitude_band for t in sel}
        assert len(band) == 1
    assert [t.seed for t in tasks] == list(range(71000, 73000))
    assert [t.path_index for t in tasks] == list(range(2000))
    assert [t.shard_index for t in tasks] == list(range(2000))  # one path per shard


def test_sample_shard_task_and_path_are_the_manifests(plan, tasks):
    shard = H.File(SAMPLE)
    (pid,) = shard["paths"].keys()
    p = shard["paths"][pid]
    meta = json.loads(p.attrs["metadata_json"])
    task = next(t for t in tasks if t.task_id == pid)
    assert task.seed == meta["seed"] == 72204
    assert task.family == meta["family"] and task.tensor_direction == meta["tensor_direction"]
    assert task.amplitude_band == meta["amplitude_band"] and task.rate_band == meta["rate_band"]
    assert task.repeat_index == meta["repeat_index"]
    assert task.design_stratum == meta["design_stratum"]
    assert task.preassigned_split == meta["preassigned_split"]
    assert task.snapshot_required == meta["snapshot_required"]
    assert task.tangent_check_required == meta["tangent_check_required"]
    assert task.accepted_increment_count == meta["accepted_increment_count"]
    assert task.expected_response_regime == meta["expected_response"]["regime"]
    assert task.amplitude == pytest.approx(meta["target_amplitude"], rel=1e-15)
    assert task.target_strain_rate == pytest.approx(meta["target_maximum_strain_rate"], rel=1e-15)
    assert task.lineage_id == meta["lineage_id"]
    path = DP.generate_data_control_path(plan.sources[0], task)
    t_ref = p["macro/time"].read()
    e_ref = p["macro/strain"].read()
    np.testing.assert_allclose(path.times[1:], t_ref, rtol=1e-14, atol=0.0)
    np.testing.assert_allclose(path.values[1:], e_ref, rtol=0.0, atol=1e-14 * np.max(np.abs(e_ref)))
    np.testing.assert_allclose(path.values[0], 0.0)


def test_production_path_amplitude_rate_and_turning_targets(plan, tasks):
    task = next(t for t in tasks if t.family == "nonproportional" and t.turning_angle == 1.9
                and t.amplitude_band == "A3" and t.rate_band == "R3")
    path = DP.generate_data_control_path(plan.sources[0], task)
    norms = np.sqrt(np.sum(path.values ** 2 * W, axis=-1))
    assert abs(np.max(norms) - task.amplitude) <= 1e-14
    rv = np.diff(path.values, axis=0) / np.diff(path.times)[:, None]
    rates = np.sqrt(np.sum(rv ** 2 * W, axis=-1))
    assert np.max(rates) <= task.target_strain_rate
    assert np.max(rates) == pytest.approx(task.target_strain_rate, rel=2e-3)
    inc = np.diff(path.values, axis=0)
    n = np.sqrt(np.sum(inc ** 2 * W, axis=-1))
    u = inc[n > 1e-14] / n[n > 1e-14][:, None]
    turn = np.arccos(np.clip(np.sum(u[1:] * u[:-1] * W, axis=1), -1.0, 1.0))
    assert abs(np.max(turn) - task.turning_angle) <= 1e-10


def test_every_family_reaches_its_amplitude_in_its_increments(plan, tasks):
    src = plan.sources[0]
    for fam in ("monotonic", "load_unload_reload", "reverse", "hold", "nonproportional"):
        for t in [x for x in tasks if x.family == fam][::37]:
            path = DP.generate_data_control_path(src, t)
            assert path.values.shape == (t.accepted_increment_count + 1, 6)
            norms = np.sqrt(np.sum(path.values ** 2 * W, axis=-1))
            # the peak knot lies on the time grid or between two samples of the
            # smoothstep's flat top
            assert 0.999 * t.amplitude <= np.max(norms) <= (1 + 1e-12) * t.amplitude
    # the seeded stream decides the path: a different seed moves it
    t0 = tasks[5]
    a = DP.generate_data_control_path(src, t0).values
    b = DP.generate_data_control_path(src, dataclasses.replace(t0, seed=t0.seed + 1)).values
    assert np.max(np.abs(a - b)) > 1e-6 * t0.amplitude


def test_supplement_has_1000_new_paths_and_a_700_150_150_split(plan, tasks):
    sup = DP.build_rve_supplemental_task_manifest(plan)
    assert len(sup) == 1000
    assert len({t.design_stratum for t in sup}) == 100
    assert not ({t.task_id for t in tasks} & {t.task_id for t in sup})
    assert _count(sup, "preassigned_split") == {"train": 700, "validation": 150, "test": 150}
    assert [t.path_index for t in sup] == list(range(2000, 3000))


def test_parallel_budget(plan):
    DP.validate_parallel_resources(plan, 32)  # 5 workers x 3 threads <= min(16, 16)
    with pytest.raises(DP.PlanError, match="exceed half of available CPUs"):
        DP.validate_parallel_resources(plan, 16)


@pytest.mark.parametrize("edit,message", [
    (lambda d: d["production_design"]["rate_bands"].pop(), "four rate bands"),
    (lambda d: d["production_design"].update(paths_per_stratum=21), "directions times resolutions"),
    (lambda d: d["production_design"].update(test_per_stratum=4), "cover every path"),
    (lambda d: d["production_design"]["amplitude_bands"][1].update(lower=5.0e-4), "must not overlap"),
    (lambda d: d["load_cases"][0].update(path_count=1999), "does not match its factorial design"),
    (lambda d: d.update(kind="single_phase_rve"), "requires a heterogeneous RVE"),
    (lambda d: d["load_cases"][0].update(pulse_duration_fraction=0.5), "Only impact data"),
    (lambda d: d.update(unexpected=1), "undeclared fields"),
])
def test_validators_reject_broken_plans(edit, message):
    data = load_yaml_config(PLAN)
    edit(data["sources"][0])
    with pytest.raises(DP.PlanError, match=message):
        DP.parse_rve_data_plan(data)


def test_fixed_cell_purpose_requires_schema_4_and_mesh_report():
    data = load_yaml_config(PLAN)
    data["schema_version"] = "3.0"
    with pytest.raises(DP.PlanError, match="schema_version=4.0"):
        DP.parse_rve_data_plan(data)
    data = load_yaml_config(PLAN)
    del data["mesh_selection_report"]
    with pytest.raises(DP.PlanError, match="mesh_selection_report"):
        DP.parse_rve_data_plan(data)


def test_voxel_cell_has_the_plan_phases_and_inclusion_fraction(plan):
    model, eph = DP.fixed_cell_voxel_model(plan.sources[0], 24)
    assert eph.shape == (24 ** 3,)
    assert abs(np.mean(eph == 1) - 0.05) < 0.01
    # a prism along z: every z layer holds the same disc
    layers = eph.reshape(24, 24, 24)
    assert np.all(layers == layers[:1])
    phases = model.material.phases
    assert phases[0].material.yield_stress == 10.0 and phases[0].material.rate_exponent == 2.0
    assert phases[1].material.young_modulus == pytest.approx(6666.666666666667)


def test_fixed_cell_v1_plan_has_eighty_balanced_unique_paths():
    """``test_fixed_cell_rve_plan.py:93-110``: 40 elastic and 40 viscoplastic paths,
    five families balanced per load case, unique task and lineage ids; its parallel
    budget fits 32 CPUs."""
    plan = DP.load_rve_data_plan(GOLDEN / "configs" / "rve_rno_fixed_cell_v1.yaml")
    tasks = DP.build_rve_data_task_manifest(plan)
    assert len(tasks) == 80
    assert len({t.task_id for t in tasks}) == 80 and len({t.lineage_id for t in tasks}) == 80
    for lc in plan.sources[0].load_cases:
        assert lc.path_count == 40
        counts = {}
        for i in range(lc.path_count):
            f = lc.path_families[i % len(lc.path_families)]
            counts[f] = counts.get(f, 0) + 1
        assert set(counts.values()) == {8}
    DP.validate_parallel_resources(plan, 32)
    # legacy strain families along a seeded unit direction, peak_scale at the control peak
    src = plan.sources[0]
    for t in tasks[:10]:
        path = DP.generate_data_control_path(src, t)
        lc = next(c for c in src.load_cases if c.name == t.load_case_name)
        assert path.values.shape == (lc.time_points, 6) and path.times[-1] == lc.duration
        np.testing.assert_allclose(path.values[0], 0.0, atol=1e-15)


@pytest.mark.parametrize("name,purpose,count", [
    ("rve_rno_data_plan.yaml", "complete_four_source", 3392),
    ("rve_rno_data_pilot.yaml", "complete_four_source", 40),
    ("rve_rno_fe2_preflight.yaml", "direct_fe2_preflight", 8),
])
def test_four_source_and_preflight_plans(name, purpose, count):
    """The reference's other data plans load and expand: every canonical source kind,
    impact tasks sharing one macro-run lineage per load case, the impact pressure a
    half-sine of the load case's peak over its duration."""
    plan = DP.load_rve_data_plan(GOLDEN / "configs" / name)
    assert plan.purpose == purpose
    tasks = DP.build_rve_data_task_manifest(plan)
    assert len(tasks) == count and len({t.task_id for t in tasks}) == count
    for src in plan.sources:
        mine = [t for t in tasks if t.source_name == src.name]
        assert len(mine) == sum(lc.path_count for lc in src.load_cases)
        t = mine[0]
        path = DP.generate_data_control_path(src, t)
        lc = next(c for c in src.load_cases if c.name == t.load_case_name)
        if src.kind == "impact_path":
            assert path.control_kind == "pressure"
            assert {x.lineage_id for x in mine if x.load_case_name == lc.name} == {
                f"{src.name}__{lc.name}__macro_run"}
            np.testing.assert_allclose(path.values[:, 0],
                                       lc.peak_scale * np.sin(np.pi * path.times / lc.duration),
                                       rtol=0, atol=1e-12 * lc.peak_scale)
        else:
            assert path.control_kind == "macro_strain" and path.values.shape == (lc.time_points, 6)
