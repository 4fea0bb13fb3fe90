"""RVE data shards (``src/fem/rve/io.py``) through this package's HDF5 reader /
writer: the reference's sample shard validates and reads as training data; a
schema-3.0 path and a schema-4.0 compact shard written here read back exactly
(the same arrays, attributes as str, bool availability masks as bool), carry the
reference's layout and publish atomically from ``.partial``."""

import json
from pathlib import Path
from types import SimpleNamespace

import numpy as np
import pytest

from concrete_impact_ai_amd import hdf5 as H
from concrete_impact_ai_amd import rve_io as IO

SAMPLE = Path(__file__).resolve().parent / "golden" / "rve_rno_c48_sample_response.h5"


def test_reference_sample_validates_and_reads_as_training_data():
    assert IO.validate_rve_data_shard(SAMPLE) == {"model_count": 1, "path_count": 1}
    paths = IO.read_rve_macro_paths(SAMPLE)
    (pid, rec), = paths.items()
    assert rec["strain"].shape == (64, 6) and rec["metadata"]["family"] == "hold"
    f = H.File(SAMPLE)
    assert IO.model_phase_names(f, f.attrs["model_id"]) == ("matrix", "inclusion")


def _model(n=(2, 2, 2)):
    from concrete_impact_ai_amd import rve as R

    nodes, elements = R.build_structured_hex8_box((1.0, 1.0, 1.0), n)
    eph = np.asarray([0, 1] * (n[0] * n[1] * n[2] // 2), dtype=np.int64)
    return SimpleNamespace(nodes=nodes, elements=elements, divisions=n, element_phase_ids=eph,
                           phase_names=("matrix", "inclusion"),
                           constraint=R.build_structured_hex8_periodic_constraint(nodes, n))


def _response(k, n_gp, rng):
    diag = {"newton_iterations": float(k), "micro_residual_norm": 1e-12 * k, "armijo_backtracks": 0.0,
            "macro_work_increment": 0.1 * k, "positive_work_density": 0.2, "negative_work_density": -0.1,
            "viscoplastic_active_volume_fraction": 0.5, "unloading_volume_fraction": 0.0,
            "hill_mandel_residual": 1e-15, "hill_mandel_error": 1e-14, "periodic_fluctuation_error": 0.0,
            "reaction_antiperiodicity_error": 1e-13, "minimum_yield_activation_margin": -1.0,
            "maximum_yield_activation_margin": 2.0, "yield_activation_margin_available": 1.0,
            "volume": 1.0}
    phase = {p: {"volume": 0.5, "volume_fraction": 0.5, "average_stress": rng.standard_normal(6),
                 "free_energy_density": 1.0 + k, "dissipation_density": 0.1 * k, "state_fields": ["q"]}
             for p in ("matrix", "inclusion")}
    act = rng.random(n_gp)
    act[:3] = np.nan  # unavailable entries
    return SimpleNamespace(
        macro_stress=rng.standard_normal(6), effective_tangent=rng.standard_normal((6, 6)),
        free_energy_density=float(k), dissipation_density=0.5 * k,
        micro_strains=rng.standard_normal((n_gp, 6)), micro_stresses=rng.standard_normal((n_gp, 6)),
        micro_diagnostics={"viscoplastic_active": act},
        state=SimpleNamespace(macro_strain=rng.standard_normal(6), material_state=SimpleNamespace(
            variables={"phase_matrix__equivalent_plastic_strain": rng.random(n_gp // 2)})),
        diagnostics=diag, phase_diagnostics=phase)


def test_schema3_path_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    model = _model()
    resp = tuple(_response(k, 64, rng) for k in range(1, 5))
    result = SimpleNamespace(times=np.asarray([0.1, 0.2, 0.35, 0.5]), responses=resp,
                             validation_diagnostics={"fd_tangent_error": np.asarray([1e-8])})
    meta = {"model_id": "box2", "material_parameters": {"b": 2.0, "a": 1.0}}
    out = IO.write_rve_path_hdf5(tmp_path / "s.h5", model, "path_0", result, meta)
    assert out.exists() and not (tmp_path / "s.h5.partial").exists()
    assert IO.validate_rve_data_shard(out) == {"model_count": 1, "path_count": 1}
    arr = IO.read_rve_path_arrays(out, "path_0")
    np.testing.assert_array_equal(arr["macro_stress"], np.stack([r.macro_stress for r in resp]))
    np.testing.assert_array_equal(arr["effective_tangent"], np.stack([r.effective_tangent for r in resp]))
    np.testing.assert_array_equal(arr["micro_stress"], np.stack([r.micro_stresses for r in resp]))
    np.testing.assert_allclose(arr["time_step"], [0.1, 0.1, 0.15, 0.15])
    f = H.File(out)
    assert f.attrs["schema_version"] == "3.0" and f.attrs["status"] == "complete"
    assert json.loads(f.attrs["metadata_json"]) == meta
    g = f["paths/path_0"]
    assert g.attrs["model_id"] == "box2"
    assert list(g["features/material_parameters_names"].read()) == [b"a", b"b"]
    np.testing.assert_array_equal(g["features/material_parameters"].read(), [1.0, 2.0])
    av = g["micro/diagnostics/viscoplastic_active__available"].read()
    assert av.dtype == bool and not av[:, :3].any() and av[:, 3:].all()
    mesh = f["models/box2/mesh"]
    assert mesh["quadrature_weights"].read().sum() == pytest.approx(1.0)
    np.testing.assert_array_equal(mesh["periodic_equivalence_class"].read(),
                                  model.constraint.equivalence_class_ids)
    assert json.loads(g["phases/matrix"].attrs["state_fields_json"]) == ["q"]
    # a non-finite accepted value is rejected by the validator
    resp[0].macro_stress[0] = np.nan
    bad = IO.write_rve_path_hdf5(tmp_path / "bad.h5", model, "p", result, meta)
    with pytest.raises(ValueError, match="non-finite"):
        IO.validate_rve_data_shard(bad)


def test_schema4_compact_shard_streams_paths(tmp_path):
    rng = np.random.default_rng(2)
    model = _model()
    w = IO.RVECompactDataShardWriter(tmp_path / "c.h5", model, {"model_id": "box2"})
    assert (tmp_path / "c.h5.partial").exists()
    with pytest.raises(ValueError):
        w.close_complete()
    kept = {}
    for p in range(3):
        w.start_path(f"p{p}", {"family": "monotonic", "material_parameters": {"E": 1.0 + p}})
        rs = [_response(k, 64, rng) for k in range(1, 6)]
        for k, r in enumerate(rs):
            w.append_step(0.1 * (k + 1), 0.1, r)
        w.finish_path()
        kept[f"p{p}"] = rs
    w.write_failure("t7", {"reason": "maximum_iterations_exceeded"})
    out = w.close_complete()
    assert IO.validate_rve_data_shard(out)["path_count"] == 3
    paths = IO.read_rve_macro_paths(out)
    for pid, rs in kept.items():
        np.testing.assert_array_equal(paths[pid]["stress"], np.stack([r.macro_stress for r in rs]))
        assert paths[pid]["metadata"]["family"] == "monotonic"
    f = H.File(out)
    assert f.attrs["schema_version"] == "4.0"
    assert json.loads(f["failures/t7"].attrs["diagnostics_json"])["reason"] == "maximum_iterations_exceeded"
    q = f["paths/p1/macro/matrix_maximum_q"].read()
    np.testing.assert_allclose(q, [np.max(r.state.material_state.variables[
        "phase_matrix__equivalent_plastic_strain"]) for r in kept["p1"]])
    np.testing.assert_array_equal(f["paths/p0/solver/incremental_work_density"].read(),
                                  [r.diagnostics["macro_work_increment"] for r in kept["p0"]])
