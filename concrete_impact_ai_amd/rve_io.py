"""RVE data shards in the reference's HDF5 layout (``src/fem/rve/io.py:24-620``).

Two schemas, written and read through :mod:`concrete_impact_ai_amd.hdf5`:

* 3.0 (``write_rve_path_hdf5`` / ``RVEDataShardWriter``, ``io.py:32-117,427-582``):
  ``models/<id>/mesh/{nodes, elements, element_phase_id, phase_names, divisions,
  periodic_equivalence_class, periodic_representative_nodes, quadrature_weights}``
  and per path ``macro/`` (time, time_step, strain, stress, effective_tangent,
  energies), ``micro/`` (strain, stress, state, diagnostics with ``__available``
  masks), ``solver/``, ``phases/<name>/``, ``features/`` and ``validation/``;
* 4.0 (``RVECompactDataShardWriter``, ``io.py:120-309``): the same mesh, per path
  only the accepted macro, solver and phase histories, appended step by step
  (the layout of the reference's production training shards, e.g.
  ``data/rve_rno_c48_v2_2000/sample/response.h5``).

Root attributes as the reference writes them (schema_version, status,
voigt_order, strain_convention, units, dataset_kind, metadata_json, model_id).
A shard is assembled in memory, written to ``<name>.partial`` and published by
``os.replace`` when it is closed complete (or failed), so a reader never sees a
half-written ``<name>``.  Arrays are stored uncompressed (the reference's LZF
compression is a storage choice; the reader decodes it).
"""

from __future__ import annotations

import json
import os
from pathlib import Path
from typing import Any

import numpy as np

from concrete_impact_ai_amd import hdf5 as H

RVE_DATA_SCHEMA_VERSION = "3.0"
RVE_COMPACT_DATA_SCHEMA_VERSION = "4.0"
SUPPORTED_RVE_DATA_SCHEMA_VERSIONS = (RVE_DATA_SCHEMA_VERSION, RVE_COMPACT_DATA_SCHEMA_VERSION)

_SOLVER_STREAMS = (
    "micro_residual_norm", "newton_iterations", "armijo_backtracks", "incremental_work_density",
    "positive_work_density", "negative_work_density", "viscoplastic_active_volume_fraction",
    "unloading_volume_fraction", "hill_mandel_residual", "hill_mandel_error",
    "periodic_fluctuation_error", "reaction_antiperiodicity_error",
    "minimum_yield_activation_margin", "maximum_yield_activation_margin", "yield_activation_margin",
    "yield_activation_margin_available",
)


def _quadrature_weights(model) -> np.ndarray:
    """w detJ per (element, Gauss point) of a structured box: every element has the
    volume h_x h_y h_z and the 2x2x2 rule weights 1 (``io.py:585-595``)."""
    n_el = model.elements.shape[0]
    h = (np.max(model.nodes, axis=0) - np.min(model.nodes, axis=0)) / np.asarray(model.divisions)
    return np.full((n_el, 8), float(np.prod(h)) / 8.0)


def _root_attrs(w: H.Writer, schema: str, metadata: dict, kind_default: str):
    a = w.attrs
    a["schema_version"] = schema
    a["status"] = "incomplete"
    a["voigt_order"] = "xx,yy,zz,yz,xz,xy"
    a["strain_convention"] = "engineering_shear"
    a["stress_unit"] = "project_material_unit"
    a["strain_unit"] = "1"
    a["dataset_kind"] = str(metadata.get("dataset_kind", kind_default))
    a["metadata_json"] = json.dumps(metadata, sort_keys=True)
    a["model_id"] = str(metadata.get("model_id", "default"))


def _write_mesh_group(w: H.Writer, model, model_id: str):
    """``_write_mesh_group`` (``io.py:427-457``)."""
    base = f"models/{model_id}/mesh"
    w.dataset(f"{base}/nodes", np.asarray(model.nodes, dtype=np.float64))
    w.dataset(f"{base}/elements", np.asarray(model.elements, dtype=np.int64))
    phase_ids = (np.zeros(model.elements.shape[0], dtype=np.int64) if model.element_phase_ids is None
                 else np.asarray(model.element_phase_ids, dtype=np.int64))
    names = tuple(model.phase_names) or ("domain",)
    w.dataset(f"{base}/element_phase_id", phase_ids)
    w.dataset(f"{base}/phase_names", np.asarray(names, dtype="S"))
    w.dataset(f"{base}/divisions", np.asarray(model.divisions, dtype=np.int64))
    w.dataset(f"{base}/periodic_equivalence_class",
              np.asarray(model.constraint.equivalence_class_ids, dtype=np.int64))
    w.dataset(f"{base}/periodic_representative_nodes",
              np.asarray(model.constraint.representative_nodes, dtype=np.int64))
    w.dataset(f"{base}/quadrature_weights", _quadrature_weights(model))


def _write_feature_group(w: H.Writer, base: str, metadata: dict):
    """``_write_feature_group`` (``io.py:343-353``)."""
    for kind in ("material_parameters", "microstructure_features"):
        mapping = metadata.get(kind, {})
        names = tuple(sorted(mapping))
        w.dataset(f"{base}/features/{kind}_names", np.asarray(names, dtype="S") if names
                  else np.zeros(0, dtype="S1"))
        w.dataset(f"{base}/features/{kind}", np.asarray([mapping[n] for n in names], dtype=np.float64))


class _ShardBase:
    def __init__(self, output_path, model, metadata, schema, kind_default):
        self.output_path = Path(output_path)
        self.partial_path = self.output_path.with_suffix(self.output_path.suffix + ".partial")
        self.output_path.parent.mkdir(parents=True, exist_ok=True)
        self.w = H.Writer()
        _root_attrs(self.w, schema, metadata, kind_default)
        self.model_id = self.w.attrs["model_id"]
        _write_mesh_group(self.w, model, self.model_id)
        self.phase_names = tuple(model.phase_names) or ("domain",)
        self.w.save(self.partial_path)  # the incomplete shard exists from the start

    def _publish(self, status):
        self.w.attrs["status"] = status
        self.w.save(self.partial_path)
        os.replace(self.partial_path, self.output_path)
        return self.output_path

    def write_failure(self, task_id: str, diagnostics: dict):
        """A numerical failure record (``io.py:102-106,288-291``)."""
        self.w.group(f"failures/{task_id}").attrs["diagnostics_json"] = json.dumps(diagnostics, sort_keys=True)

    def close_failed(self):
        if "failures" not in self.w or not self.w.get("failures").children:
            raise ValueError("Failed RVE shard requires at least one failure record.")
        return self._publish("failed")

    def close_complete(self):
        if "paths" not in self.w or not self.w.get("paths").children:
            raise ValueError("RVE data shard must contain at least one accepted path.")
        return self._publish("complete")


class RVEDataShardWriter(_ShardBase):
    """One schema-3.0 shard (``RVEDataShardWriter``, ``io.py:53-117``)."""

    def __init__(self, output_path, model, metadata: dict):
        super().__init__(output_path, model, metadata, RVE_DATA_SCHEMA_VERSION, "rve_validation")

    def write_path(self, path_name: str, result, metadata: dict):
        """``_write_path_group`` (``io.py:460-582``) plus features and attributes."""
        if f"paths/{path_name}" in self.w:
            raise ValueError(f"Duplicate RVE path id in shard: {path_name}.")
        resp = result.responses
        times = np.asarray(result.times, dtype=np.float64)
        if len(resp) != times.size:
            raise ValueError("RVE path times and responses must have identical lengths.")
        base = f"paths/{path_name}"
        w = self.w
        steps = np.zeros(times.size)
        steps[0] = times[0]
        steps[1:] = np.diff(times)
        w.dataset(f"{base}/macro/time", times)
        w.dataset(f"{base}/macro/time_step", steps)
        w.dataset(f"{base}/macro/strain", np.stack([r.state.macro_strain for r in resp]))
        w.dataset(f"{base}/macro/stress", np.stack([r.macro_stress for r in resp]))
        if any(r.effective_tangent is None for r in resp):
            raise ValueError("RVE HDF5 paths require an effective tangent at every accepted step.")
        w.dataset(f"{base}/macro/effective_tangent", np.stack([r.effective_tangent for r in resp]))
        w.dataset(f"{base}/macro/free_energy_density", np.asarray([r.free_energy_density for r in resp]))
        w.dataset(f"{base}/macro/dissipation_density", np.asarray([r.dissipation_density for r in resp]))
        w.dataset(f"{base}/micro/strain", np.stack([r.micro_strains for r in resp]))
        w.dataset(f"{base}/micro/stress", np.stack([r.micro_stresses for r in resp]))
        for name in resp[0].state.material_state.variables:
            w.dataset(f"{base}/micro/state/{name}",
                      np.stack([np.asarray(r.state.material_state.variables[name]) for r in resp]))
        names = sorted(set().union(*(r.micro_diagnostics.keys() for r in resp)))
        npt = resp[0].micro_stresses.shape[0]
        for name in names:
            vals = np.full((len(resp), npt), np.nan)
            avail = np.zeros((len(resp), npt), dtype=bool)
            for k, r in enumerate(resp):
                if name in r.micro_diagnostics:
                    vals[k] = r.micro_diagnostics[name]
                    avail[k] = np.isfinite(r.micro_diagnostics[name])
            w.dataset(f"{base}/micro/diagnostics/{name}", vals)
            w.dataset(f"{base}/micro/diagnostics/{name}__available", avail)
        w.dataset(f"{base}/solver/accepted_time_step", steps)
        w.dataset(f"{base}/solver/rejected_step_count", np.zeros(times.size, dtype=np.int64))
        w.dataset(f"{base}/solver/state_commit_count", np.ones(times.size, dtype=np.int64))
        w.dataset(f"{base}/solver/failure_status", np.zeros(times.size, dtype=np.int8))
        for name in resp[0].diagnostics:
            w.dataset(f"{base}/solver/{name}", np.asarray([r.diagnostics[name] for r in resp]))
        for pname in sorted(set().union(*(r.phase_diagnostics.keys() for r in resp))):
            diag = [r.phase_diagnostics[pname] for r in resp]
            pb = f"{base}/phases/{pname}"
            w.dataset(f"{pb}/volume_fraction", np.asarray([d["volume_fraction"] for d in diag]))
            w.dataset(f"{pb}/average_stress", np.stack([d["average_stress"] for d in diag]))
            w.dataset(f"{pb}/free_energy_density", np.asarray([d["free_energy_density"] for d in diag]))
            w.dataset(f"{pb}/dissipation_density", np.asarray([d["dissipation_density"] for d in diag]))
            w.get(pb).attrs["state_fields_json"] = json.dumps(diag[0].get("state_fields", []), sort_keys=True)
        w.group(f"{base}/validation")
        for name, values in getattr(result, "validation_diagnostics", {}).items():
            w.dataset(f"{base}/validation/{name}", np.asarray(values))
        g = w.get(base)
        g.attrs["model_id"] = self.model_id
        g.attrs["metadata_json"] = json.dumps(metadata, sort_keys=True)
        _write_feature_group(w, base, metadata)


def write_rve_path_hdf5(output_path, model, path_name: str, result, metadata: dict) -> Path:
    """``write_rve_path_hdf5`` (``io.py:32-50``): one accepted path with its micro
    histories in a schema-3.0 shard."""
    writer = RVEDataShardWriter(output_path, model, metadata)
    writer.write_path(path_name, result, metadata)
    return writer.close_complete()


class RVECompactDataShardWriter(_ShardBase):
    """A schema-4.0 shard (``RVECompactDataShardWriter``, ``io.py:120-309``):
    ``start_path`` / ``append_step`` / ``finish_path`` per accepted path."""

    def __init__(self, output_path, model, metadata: dict, compression: str = "lzf"):
        super().__init__(output_path, model, metadata, RVE_COMPACT_DATA_SCHEMA_VERSION, "rve_training")
        self.compression = compression
        self.active = None

    def start_path(self, path_name: str, metadata: dict):
        if self.active is not None:
            raise RuntimeError("Compact RVE writer already has an active path.")
        if f"paths/{path_name}" in self.w:
            raise ValueError(f"Duplicate RVE path id in shard: {path_name}.")
        self.active = {"name": path_name, "metadata": metadata,
                       "macro": {k: [] for k in ("time", "time_step", "strain", "stress", "effective_tangent",
                                                 "free_energy_density", "dissipation_density",
                                                 "matrix_average_q", "matrix_maximum_q")},
                       "solver": {k: [] for k in _SOLVER_STREAMS},
                       "phases": {p: {k: [] for k in ("volume_fraction", "average_stress",
                                                      "free_energy_density", "dissipation_density")}
                                  for p in self.phase_names}}

    def append_step(self, time: float, time_step: float, response):
        """``append_step`` (``io.py:196-236``): one converged, accepted response."""
        a = self.active
        if a is None:
            raise RuntimeError("Compact RVE writer has no active path.")
        if response.effective_tangent is None:
            raise ValueError("RVE HDF5 paths require an effective tangent at every accepted step.")
        q = np.asarray(response.state.material_state.variables["phase_matrix__equivalent_plastic_strain"])
        m = a["macro"]
        m["time"].append(time)
        m["time_step"].append(time_step)
        m["strain"].append(np.asarray(response.state.macro_strain))
        m["stress"].append(np.asarray(response.macro_stress))
        m["effective_tangent"].append(np.asarray(response.effective_tangent))
        m["free_energy_density"].append(response.free_energy_density)
        m["dissipation_density"].append(response.dissipation_density)
        m["matrix_average_q"].append(float(np.mean(q)))
        m["matrix_maximum_q"].append(float(np.max(q)))
        for name in _SOLVER_STREAMS:
            src = ("macro_work_increment" if name == "incremental_work_density" else
                   "minimum_yield_activation_margin" if name == "yield_activation_margin" else name)
            a["solver"][name].append(float(response.diagnostics[src]))
        for pname, rec in a["phases"].items():
            d = response.phase_diagnostics[pname]
            for k in rec:
                rec[k].append(np.asarray(d[k]))

    def finish_path(self):
        a = self.active
        if a is None:
            raise RuntimeError("Compact RVE writer has no active path.")
        t = np.asarray(a["macro"]["time"], dtype=np.float64)
        if t.size == 0 or (t.size > 1 and np.any(np.diff(t) <= 0.0)):
            raise ValueError("Compact RVE path requires strictly increasing accepted times.")
        base = f"paths/{a['name']}"
        w = self.w
        for k, v in a["macro"].items():
            w.dataset(f"{base}/macro/{k}", np.asarray(v, dtype=np.float64))
        for k, v in a["solver"].items():
            w.dataset(f"{base}/solver/{k}", np.asarray(v, dtype=np.float64))
        for pname, rec in a["phases"].items():
            for k, v in rec.items():
                w.dataset(f"{base}/phases/{pname}/{k}", np.asarray(v, dtype=np.float64))
        g = w.get(base)
        g.attrs["model_id"] = self.model_id
        g.attrs["metadata_json"] = json.dumps(a["metadata"], sort_keys=True)
        _write_feature_group(w, base, a["metadata"])
        self.active = None

    def close_complete(self):
        if self.active is not None:
            raise RuntimeError("Cannot close compact shard while a path is active.")
        return super().close_complete()


def model_phase_names(shard: H.File, model_id: str) -> tuple:
    """``model_phase_names`` (``io.py:336-340``)."""
    return tuple(v.decode() for v in shard[f"models/{model_id}/mesh/phase_names"].read())


def read_rve_path_arrays(input_path, path_name: str) -> dict:
    """``read_rve_path_arrays`` (``io.py:356-374``): the principal macro and micro
    arrays of one schema-3.0 path."""
    f = H.File(input_path)
    if f.attrs["schema_version"] != RVE_DATA_SCHEMA_VERSION:
        raise ValueError("Unsupported RVE data schema version.")
    g = f[f"paths/{path_name}"]
    return {"time": g["macro/time"].read(), "time_step": g["macro/time_step"].read(),
            "macro_strain": g["macro/strain"].read(), "macro_stress": g["macro/stress"].read(),
            "effective_tangent": g["macro/effective_tangent"].read(),
            "micro_strain": g["micro/strain"].read(), "micro_stress": g["micro/stress"].read()}


def read_rve_macro_paths(input_path) -> dict:
    """Every path's accepted macro histories of a shard (either schema):
    {path id: {"time", "time_step", "strain", "stress", "effective_tangent",
    "metadata"}} -- the surrogate's training data in the reference's format."""
    f = H.File(input_path)
    out = {}
    for pid in f["paths"].keys():
        g = f[f"paths/{pid}"]
        rec = {k: g[f"macro/{k}"].read() for k in ("time", "time_step", "strain", "stress",
                                                   "effective_tangent")}
        rec["metadata"] = json.loads(g.attrs.get("metadata_json", "{}"))
        out[pid] = rec
    return out


def validate_rve_data_shard(input_path) -> dict:
    """``validate_rve_data_shard`` (``io.py:377-415``): schema, status, groups,
    strictly increasing times and finite values (masked by the availability
    datasets and, for the activation margins, by yield_activation_margin_available)."""
    f = H.File(input_path)
    if f.attrs["schema_version"] not in SUPPORTED_RVE_DATA_SCHEMA_VERSIONS:
        raise ValueError("Unsupported RVE data schema version.")
    if f.attrs["status"] != "complete":
        raise ValueError("RVE data shard is not complete.")
    if "models" not in f or "paths" not in f:
        raise ValueError("RVE data shard requires models and paths groups.")
    for pid in f["paths"].keys():
        g = f[f"paths/{pid}"]
        if "features" not in g or "macro" not in g:
            raise ValueError(f"Incomplete RVE data path: {pid}.")
        t = g["macro/time"].read()
        if t.size == 0 or np.any(np.diff(t) <= 0.0):
            raise ValueError(f"RVE data path time must be strictly increasing: {pid}.")
        datasets = []
        g.visit(lambda name, obj: datasets.append((name, obj)) if isinstance(obj, H.Dataset) else None)
        names = {n for n, _ in datasets}
        for name, ds in datasets:
            if name.endswith("__available"):
                continue
            values = ds.read()
            if f"{name}__available" in names:
                values = values[g[f"{name}__available"].read()]
            if name.endswith(("minimum_yield_activation_margin", "maximum_yield_activation_margin")):
                values = values[g["solver/yield_activation_margin_available"].read().astype(bool)]
            if np.issubdtype(np.asarray(values).dtype, np.floating) and not np.all(np.isfinite(values)):
                raise ValueError(f"RVE data path contains non-finite values: {pid}/{name}.")
    return {"model_count": len(f["models"].keys()), "path_count": len(f["paths"].keys())}
