"""RVE training-data plans of the reference and their path production.

A plan (YAML, ``src/concrete_impact/experiments/rve_data_plan.py:36-433``) names
the data sources, their load cases and, for the fixed-cell training set, a
factorial production design: 5 loading families x 5 amplitude bands x 4 rate
bands = 100 strata of 20 paths (5 tensor directions x 4 time resolutions), with
frozen per-stratum train / validation / test splits.  This module

* reads and validates a plan (``load_rve_data_plan``: strict keys, the design's
  canonical shape and the same error messages as the reference's validators),
* expands it into the deterministic task manifest (``build_rve_data_task_manifest``,
  ``rve_data_plan.py:478-517, 596-716``: task ids, seeds, strata, amplitudes and
  target rates placed inside their bands, sha256-ranked splits, snapshot and
  tangent-check selection) and the optional 1000-path supplement
  (``build_rve_supplemental_task_manifest``, ``rve_data_plan.py:518-594``),
* generates every task's control history (``generate_data_control_path``,
  ``rve_path_generation.py:34-177``: the production strain path with analytic
  timing, or the legacy families and the impact pressure pulse), and
* produces the macro responses of a task set on the device
  (``produce_plan_responses``): tasks grouped by increment count, each group one
  batched FE² solve per increment on a structured voxel cell built from the plan's
  material and microstructure parameters (``fixed_cell_voxel_model``).

The reference's own fixed cell is an O-grid mesh with a cylindrical inclusion
(unstructured Hex8, ``fem/mesh/structured.py``); the device solver takes
structured boxes, so the voxel cell is a stand-in with the same phases, material
parameters and inclusion volume fraction, and its responses are not the
reference's labels.  The paths and task metadata are the reference's exactly.
"""

from __future__ import annotations

import dataclasses
import hashlib
from pathlib import Path
from typing import Any

import numpy as np

from concrete_impact_ai_amd.config import load_yaml_config
from concrete_impact_ai_amd.microstructure import (FAMILY_CONTROLS, TENSOR_DIRECTIONS,
                                                   production_strain_path)

__all__ = ["AmplitudeBand", "RateBand", "ProductionDesign", "LoadCase", "DataSource",
           "ParallelPlan", "RVEDataPlan", "RVEDataTask", "GeneratedControlPath",
           "load_rve_data_plan", "parse_rve_data_plan", "validate_parallel_resources",
           "build_rve_data_task_manifest", "build_rve_supplemental_task_manifest",
           "generate_data_control_path", "fixed_cell_voxel_model", "produce_plan_responses"]

SOURCE_KINDS = ("material_point", "single_phase_rve", "heterogeneous_rve", "impact_path")
PATH_ORIGINS = ("prescribed_macro_strain", "direct_j2_explicit", "direct_fe2_explicit",
                "direct_fe2_implicit")
FAMILIES = ("monotonic", "load_unload_reload", "reverse", "hold", "nonproportional")
REGIMES = ("elastic", "transition", "viscoplastic")
SPLITS = ("train", "validation", "test")


class PlanError(ValueError):
    """A plan that violates the reference's plan schema."""


def _block(data: Any, where: str, required: tuple[str, ...], optional: tuple[str, ...] = ()):
    """A mapping with exactly the required keys and a subset of the optional ones
    (undeclared fields are rejected, as the reference's strict plan models do)."""
    if not isinstance(data, dict):
        raise PlanError(f"{where} must be a mapping.")
    missing = [k for k in required if k not in data]
    extra = [k for k in data if k not in required and k not in optional]
    if missing:
        raise PlanError(f"{where} is missing fields {missing}.")
    if extra:
        raise PlanError(f"{where} has undeclared fields {extra}.")
    return data


def _choice(value, allowed, where):
    if value not in allowed:
        raise PlanError(f"{where} must be one of {list(allowed)}, got {value!r}.")
    return value


def _positive(value, where, integer=False):
    if isinstance(value, bool) or not isinstance(value, (int, float)) or value <= 0:
        raise PlanError(f"{where} must be positive, got {value!r}.")
    if integer and int(value) != value:
        raise PlanError(f"{where} must be an integer, got {value!r}.")
    return int(value) if integer else float(value)


@dataclasses.dataclass(frozen=True)
class AmplitudeBand:
    name: str
    lower: float
    upper: float
    expected_response: str


@dataclasses.dataclass(frozen=True)
class RateBand:
    name: str
    lower: float
    upper: float


@dataclasses.dataclass(frozen=True)
class ProductionDesign:
    """The fixed-cell factorial design (``rve_data_plan.py:97-143``)."""

    amplitude_bands: tuple[AmplitudeBand, ...]
    rate_bands: tuple[RateBand, ...]
    tensor_directions: tuple[str, ...]
    accepted_increment_counts: tuple[int, ...]
    nonproportional_turning_angles: tuple[float, ...]
    paths_per_stratum: int
    train_per_stratum: int
    validation_per_stratum: int
    test_per_stratum: int
    snapshots_per_stratum: int
    tangent_checks_per_stratum: int
    direction_perturbation: float = 0.0


@dataclasses.dataclass(frozen=True)
class LoadCase:
    name: str
    path_count: int
    time_points: int
    duration: float
    excitation: str
    seed: int
    peak_scale: float
    path_families: tuple[str, ...]
    expected_response: dict
    pulse_duration_fraction: float | None = None
    axial_region_count: int | None = None


@dataclasses.dataclass(frozen=True)
class DataSource:
    name: str
    kind: str
    model: str
    boundary: dict
    load_cases: tuple[LoadCase, ...]
    fields: dict
    training_role: str
    macro_path_origin: str
    material_parameters: dict
    microstructure_parameters: dict
    production_design: ProductionDesign | None = None


@dataclasses.dataclass(frozen=True)
class ParallelPlan:
    workers: int
    paths_per_task: int
    shard_path_count: int
    compression: str
    thread_count_per_worker: int
    logical_cpu_budget: int = 24


@dataclasses.dataclass(frozen=True)
class RVEDataPlan:
    schema_version: str
    purpose: str
    storage_mode: str
    output_directory: Path
    mesh_selection_report: Path | None
    sources: tuple[DataSource, ...]
    parallel: ParallelPlan
    audit: dict


@dataclasses.dataclass(frozen=True)
class RVEDataTask:
    """One deterministic path task (the fields of ``rve_data_plan.py:405-430``)."""

    task_id: str
    source_name: str
    source_kind: str
    load_case_name: str
    path_index: int
    seed: int
    shard_index: int
    lineage_id: str
    macro_path_origin: str
    family: str | None = None
    amplitude_band: str | None = None
    rate_band: str | None = None
    tensor_direction: str | None = None
    repeat_index: int | None = None
    amplitude: float | None = None
    target_strain_rate: float | None = None
    accepted_increment_count: int | None = None
    turning_angle: float | None = None
    design_stratum: str | None = None
    preassigned_split: str | None = None
    expected_response_regime: str | None = None
    snapshot_required: bool = False
    tangent_check_required: bool = False


# ---------------------------------------------------------------------------
# parsing and validation

def _bands(items, where, amplitude):
    out = []
    for i, b in enumerate(items):
        w = f"{where}[{i}]"
        if amplitude:
            _block(b, w, ("name", "lower", "upper", "expected_response"))
            band = AmplitudeBand(str(b["name"]), _positive(b["lower"], w + ".lower"),
                                 _positive(b["upper"], w + ".upper"),
                                 _choice(b["expected_response"], REGIMES, w + ".expected_response"))
        else:
            _block(b, w, ("name", "lower", "upper"))
            band = RateBand(str(b["name"]), _positive(b["lower"], w + ".lower"),
                            _positive(b["upper"], w + ".upper"))
        if band.upper <= band.lower:
            kind = "amplitude" if amplitude else "rate"
            raise PlanError(f"Production {kind} bands require upper > lower.")
        out.append(band)
    label = "amplitude" if amplitude else "rate"
    names = [b.name for b in out]
    if len(names) != len(set(names)):
        raise PlanError(f"Production {label} band names must be unique.")
    for prev, cur in zip(out[:-1], out[1:]):
        if cur.lower < prev.upper:
            raise PlanError(f"Production {label} bands must not overlap.")
    return tuple(out)


def _design(d, where) -> ProductionDesign:
    keys = ("amplitude_bands", "rate_bands", "tensor_directions", "accepted_increment_counts",
            "nonproportional_turning_angles", "paths_per_stratum", "train_per_stratum",
            "validation_per_stratum", "test_per_stratum", "snapshots_per_stratum",
            "tangent_checks_per_stratum")
    _block(d, where, keys, ("direction_perturbation",))
    amp = _bands(d["amplitude_bands"], where + ".amplitude_bands", True)
    rate = _bands(d["rate_bands"], where + ".rate_bands", False)
    dirs = tuple(_choice(x, tuple(TENSOR_DIRECTIONS), where + ".tensor_directions")
                 for x in d["tensor_directions"])
    counts = tuple(_positive(x, where + ".accepted_increment_counts", True)
                   for x in d["accepted_increment_counts"])
    angles = tuple(_positive(x, where + ".nonproportional_turning_angles")
                   for x in d["nonproportional_turning_angles"])
    ints = {k: _positive(d[k], f"{where}.{k}", True) for k in keys[5:]}
    pert = float(d.get("direction_perturbation", 0.0))
    if not 0.0 <= pert < 0.25:
        raise PlanError(f"{where}.direction_perturbation must lie in [0, 0.25).")
    design = ProductionDesign(amp, rate, dirs, counts, angles, **ints, direction_perturbation=pert)
    if len(amp) != 5 or len(rate) != 4:
        raise PlanError("Fixed-cell production requires five amplitude and four rate bands.")
    if len(dirs) != 5 or len(set(dirs)) != 5:
        raise PlanError("Fixed-cell production requires five unique tensor directions.")
    if len(counts) != 4:
        raise PlanError("Fixed-cell production requires four time resolutions.")
    if len(angles) != 4:
        raise PlanError("Fixed-cell production requires four turning-angle levels.")
    if design.paths_per_stratum != len(dirs) * len(counts):
        raise PlanError("Production paths_per_stratum must equal directions times resolutions.")
    if (design.train_per_stratum + design.validation_per_stratum + design.test_per_stratum
            != design.paths_per_stratum):
        raise PlanError("Production per-stratum split counts must cover every path.")
    if design.snapshots_per_stratum > design.paths_per_stratum:
        raise PlanError("Production snapshot count exceeds paths_per_stratum.")
    if design.tangent_checks_per_stratum > design.paths_per_stratum:
        raise PlanError("Production tangent-check count exceeds paths_per_stratum.")
    return design


def _load_case(c, where) -> LoadCase:
    _block(c, where, ("name", "path_count", "time_points", "duration", "excitation", "seed",
                      "peak_scale", "path_families", "expected_response"),
           ("pulse_duration_fraction", "axial_region_count"))
    exc = _choice(c["excitation"], ("prescribed_six_component_strain",
                                    "cyclic_nonproportional_strain", "half_sine_pressure_pulse"),
                  where + ".excitation")
    fams = tuple(_choice(f, FAMILIES, where + ".path_families") for f in c["path_families"])
    er = _block(c["expected_response"], where + ".expected_response",
                ("regime", "require_unloading", "require_reverse_loading"))
    _choice(er["regime"], REGIMES, where + ".expected_response.regime")
    seed = c["seed"]
    if isinstance(seed, bool) or not isinstance(seed, int) or seed < 0:
        raise PlanError(f"{where}.seed must be a non-negative integer.")
    pdf = c.get("pulse_duration_fraction")
    arc = c.get("axial_region_count")
    lc = LoadCase(str(c["name"]), _positive(c["path_count"], where + ".path_count", True),
                  _positive(c["time_points"], where + ".time_points", True),
                  _positive(c["duration"], where + ".duration"), exc, int(seed),
                  _positive(c["peak_scale"], where + ".peak_scale"), fams, dict(er),
                  None if pdf is None else float(pdf),
                  None if arc is None else _positive(arc, where + ".axial_region_count", True))
    if not lc.path_families:
        raise PlanError("RVE data load case requires at least one path family.")
    impact = exc == "half_sine_pressure_pulse"
    if impact and lc.pulse_duration_fraction is None:
        raise PlanError("Impact data load case requires pulse_duration_fraction.")
    if impact and lc.axial_region_count is None:
        raise PlanError("Impact data load case requires axial_region_count.")
    if not impact and lc.pulse_duration_fraction is not None:
        raise PlanError("Only impact data may define pulse_duration_fraction.")
    if not impact and lc.axial_region_count is not None:
        raise PlanError("Only impact data may define axial_region_count.")
    if lc.pulse_duration_fraction is not None and not 0.0 < lc.pulse_duration_fraction < 1.0:
        raise PlanError(f"{where}.pulse_duration_fraction must lie in (0, 1).")
    if lc.axial_region_count is not None and lc.path_count % lc.axial_region_count != 0:
        raise PlanError("Impact path_count must be divisible by axial_region_count.")
    return lc


def _parameters(block, where):
    out = {}
    for name, spec in dict(block).items():
        w = f"{where}.{name}"
        kind = spec.get("kind") if isinstance(spec, dict) else None
        if kind == "fixed":
            _block(spec, w, ("kind", "value"))
        elif kind == "grid":
            _block(spec, w, ("kind", "values"))
            if not spec["values"]:
                raise PlanError("Grid parameter sampling requires at least one value.")
        elif kind in ("uniform", "log_uniform"):
            _block(spec, w, ("kind", "lower", "upper"))
            if kind == "log_uniform":
                _positive(spec["lower"], w + ".lower")
            if spec["upper"] <= spec["lower"]:
                raise PlanError("Uniform parameter sampling requires upper > lower." if kind == "uniform"
                                else "Log-uniform parameter sampling requires upper > lower.")
        else:
            raise PlanError(f"{w} has an unknown parameter kind {kind!r}.")
        out[name] = dict(spec)
    return out


def _source(s, where) -> DataSource:
    _block(s, where, ("name", "kind", "model", "boundary", "load_cases", "fields", "training_role",
                      "macro_path_origin", "material_parameters", "microstructure_parameters"),
           ("production_design",))
    kind = _choice(s["kind"], SOURCE_KINDS, where + ".kind")
    origin = _choice(s["macro_path_origin"], PATH_ORIGINS, where + ".macro_path_origin")
    b = _block(s["boundary"], where + ".boundary", ("kind", "description"))
    _choice(b["kind"], ("none", "periodic_macro_strain", "macro_impact_structure"),
            where + ".boundary.kind")
    f = _block(s["fields"], where + ".fields", ("inputs", "labels", "state_fields", "diagnostics"))
    _choice(s["training_role"], ("pretraining", "homogenization", "coverage", "validation"),
            where + ".training_role")
    cases = tuple(_load_case(c, f"{where}.load_cases[{i}]") for i, c in enumerate(s["load_cases"]))
    design = (None if s.get("production_design") is None
              else _design(s["production_design"], where + ".production_design"))
    src = DataSource(str(s["name"]), kind, str(s["model"]), dict(b), cases, dict(f),
                     str(s["training_role"]), origin,
                     _parameters(s["material_parameters"], where + ".material_parameters"),
                     _parameters(s["microstructure_parameters"], where + ".microstructure_parameters"),
                     design)
    if kind == "impact_path" and origin == "prescribed_macro_strain":
        raise PlanError("Impact paths require a direct macro-dynamics origin.")
    if kind != "impact_path" and origin != "prescribed_macro_strain":
        raise PlanError("Non-impact sources require prescribed_macro_strain origin.")
    if design is not None:
        if kind != "heterogeneous_rve":
            raise PlanError("Fixed-cell production design requires a heterogeneous RVE.")
        if len(cases) != 1:
            raise PlanError("Fixed-cell production design requires exactly one load case.")
        expected = (len(cases[0].path_families) * len(design.amplitude_bands)
                    * len(design.rate_bands) * design.paths_per_stratum)
        if cases[0].path_count != expected:
            raise PlanError("Fixed-cell production path_count does not match its factorial design: "
                            f"configured={cases[0].path_count}, expected={expected}.")
    return src


def parse_rve_data_plan(data: dict) -> RVEDataPlan:
    """Validate a plan mapping (``rve_data_plan.py:325-402`` and the models it holds)."""
    _block(data, "plan", ("schema_version", "output_directory", "sources", "parallel", "audit"),
           ("purpose", "storage_mode", "mesh_selection_report"))
    version = _choice(str(data["schema_version"]), ("3.0", "4.0"), "plan.schema_version")
    purpose = _choice(data.get("purpose", "complete_four_source"),
                      ("complete_four_source", "direct_fe2_preflight", "fixed_cell_rve_training"),
                      "plan.purpose")
    storage = _choice(data.get("storage_mode", "full_history"),
                      ("full_history", "compact_streaming"), "plan.storage_mode")
    p = _block(data["parallel"], "plan.parallel",
               ("workers", "paths_per_task", "shard_path_count", "compression",
                "thread_count_per_worker"), ("logical_cpu_budget",))
    parallel = ParallelPlan(_positive(p["workers"], "workers", True),
                            _positive(p["paths_per_task"], "paths_per_task", True),
                            _positive(p["shard_path_count"], "shard_path_count", True),
                            _choice(p["compression"], ("lzf", "gzip"), "plan.parallel.compression"),
                            _positive(p["thread_count_per_worker"], "thread_count_per_worker", True),
                            _positive(p.get("logical_cpu_budget", 24), "logical_cpu_budget", True))
    audit_keys = ("q_tolerance", "plastic_q_threshold", "dissipation_tolerance",
                  "positive_cumulative_dissipation_threshold", "work_tolerance",
                  "stress_drop_threshold", "strain_increment_tolerance", "hold_strain_tolerance",
                  "hold_stress_relaxation_threshold", "tangent_perturbation",
                  "tangent_stress_scale", "tangent_direction_tolerance")
    audit = _block(data["audit"], "plan.audit", audit_keys)
    for k in audit_keys:
        _positive(audit[k], "plan.audit." + k)
    sources = tuple(_source(s, f"plan.sources[{i}]") for i, s in enumerate(data["sources"]))
    report = data.get("mesh_selection_report")
    plan = RVEDataPlan(version, purpose, storage, Path(data["output_directory"]),
                       None if report is None else Path(report), sources, parallel, dict(audit))
    kinds = [s.kind for s in sources]
    names = [s.name for s in sources]
    if len(names) != len(set(names)):
        raise PlanError("RVE-RNO data source names must be unique.")
    if purpose == "direct_fe2_preflight":
        if len(sources) != 1 or kinds != ["impact_path"]:
            raise PlanError("Direct FE2 preflight requires exactly one impact_path source.")
        if sources[0].macro_path_origin != "direct_fe2_explicit":
            raise PlanError("Direct FE2 preflight requires macro_path_origin=direct_fe2_explicit.")
    elif purpose == "fixed_cell_rve_training":
        if version != "4.0" or storage != "compact_streaming":
            raise PlanError("Fixed-cell RVE training requires schema_version=4.0 and "
                            "storage_mode=compact_streaming.")
        if len(sources) != 1 or kinds != ["heterogeneous_rve"]:
            raise PlanError("Fixed-cell RVE training requires exactly one heterogeneous_rve source.")
        missing = {"circumferential_divisions", "matrix_radial_divisions",
                   "axial_divisions"} - set(sources[0].microstructure_parameters)
        if missing:
            raise PlanError("Fixed-cell RVE training requires explicit mesh parameters: "
                            f"missing={sorted(missing)}.")
        if report is None:
            raise PlanError("Fixed-cell RVE training requires an explicit mesh_selection_report.")
    else:
        if not set(SOURCE_KINDS).issubset(kinds):
            raise PlanError("RVE-RNO data plan requires every canonical source kind.")
        for k in ("material_point", "single_phase_rve", "heterogeneous_rve"):
            if kinds.count(k) != 1:
                raise PlanError(f"RVE-RNO data plan requires exactly one {k} source.")
    return plan


def load_rve_data_plan(path: str | Path) -> RVEDataPlan:
    """Load and validate a plan (``rve_data_plan.py:433-435``)."""
    return parse_rve_data_plan(load_yaml_config(path))


def validate_parallel_resources(plan: RVEDataPlan, available_cpus: int) -> None:
    """Reject workers x threads above half the CPUs or the plan's logical-CPU budget
    (``rve_data_plan.py:461-475``)."""
    requested = plan.parallel.workers * plan.parallel.thread_count_per_worker
    allowed = min(available_cpus // 2, plan.parallel.logical_cpu_budget)
    if requested > allowed:
        raise PlanError("Configured RVE path processes exceed half of available CPUs or the "
                        f"explicit logical CPU budget: workers={plan.parallel.workers}, "
                        f"threads_per_worker={plan.parallel.thread_count_per_worker}, "
                        f"requested={requested}, allowed={allowed}.")


# ---------------------------------------------------------------------------
# task manifests

def _split_ranks(stratum: str, n: int) -> list[int]:
    """Rank of every candidate of a stratum: candidates ordered by the sha256 digest
    of "<stratum>|candidate_<index:02d>" (list order does not decide the split)."""
    digests = [hashlib.sha256(f"{stratum}|candidate_{i:02d}".encode()).hexdigest() for i in range(n)]
    ranks = [0] * n
    for rank, i in enumerate(sorted(range(n), key=digests.__getitem__)):
        ranks[i] = rank
    return ranks


def _split_of(rank: int, d: ProductionDesign) -> str:
    if rank < d.train_per_stratum:
        return "train"
    return "validation" if rank < d.train_per_stratum + d.validation_per_stratum else "test"


def _diagnostic_ranks(count: int, total: int) -> set[int]:
    """Ranks selected for a per-stratum diagnostic: both ends and evenly between."""
    if count == 1:
        return {0}
    return {round(i * (total - 1) / (count - 1)) for i in range(count)}


def _production_tasks(src: DataSource, lc: LoadCase, offset: int, shard: int) -> list[RVEDataTask]:
    d = src.production_design
    n = d.paths_per_stratum
    snaps = _diagnostic_ranks(d.snapshots_per_stratum, n)
    tasks = []
    for family in lc.path_families:
        for ab in d.amplitude_bands:
            for rb in d.rate_bands:
                stratum = f"{family}|{ab.name}|{rb.name}"
                ranks = _split_ranks(stratum, n)
                for di, direction in enumerate(d.tensor_directions):
                    for ri, increments in enumerate(d.accepted_increment_counts):
                        w = di * len(d.accepted_increment_counts) + ri  # place in the stratum
                        # amplitude at the w-th of n cell midpoints of its band, rate at
                        # the midpoint of cell (7 w + 3) mod n (decorrelated from amplitude)
                        amp = ab.lower + (w + 0.5) / n * (ab.upper - ab.lower)
                        rate = rb.lower + (((7 * w + 3) % n) + 0.5) / n * (rb.upper - rb.lower)
                        local = len(tasks)
                        tid = f"{src.name}__{family}__{ab.name}__{rb.name}__{direction}__rep_{ri:02d}"
                        tasks.append(RVEDataTask(
                            task_id=tid, source_name=src.name, source_kind=src.kind,
                            load_case_name=lc.name, path_index=local, seed=lc.seed + local,
                            shard_index=(offset + local) // shard, lineage_id=tid,
                            macro_path_origin=src.macro_path_origin, family=family,
                            amplitude_band=ab.name, rate_band=rb.name, tensor_direction=direction,
                            repeat_index=ri, amplitude=amp, target_strain_rate=rate,
                            accepted_increment_count=increments,
                            turning_angle=(d.nonproportional_turning_angles[ri]
                                           if family == "nonproportional" else None),
                            design_stratum=stratum, preassigned_split=_split_of(ranks[w], d),
                            expected_response_regime=ab.expected_response,
                            snapshot_required=ranks[w] in snaps,
                            tangent_check_required=ranks[w] < d.tangent_checks_per_stratum))
    return tasks


def build_rve_data_task_manifest(plan: RVEDataPlan) -> tuple[RVEDataTask, ...]:
    """Every task of the plan in source / load-case order, global shard indices."""
    tasks: list[RVEDataTask] = []
    shard = plan.parallel.shard_path_count
    for src in plan.sources:
        for lc in src.load_cases:
            if src.production_design is not None:
                tasks.extend(_production_tasks(src, lc, len(tasks), shard))
                continue
            for i in range(lc.path_count):
                tid = f"{src.name}__{lc.name}__path_{i:06d}"
                lineage = f"{src.name}__{lc.name}__macro_run" if src.kind == "impact_path" else tid
                tasks.append(RVEDataTask(
                    task_id=tid, source_name=src.name, source_kind=src.kind, load_case_name=lc.name,
                    path_index=i, seed=lc.seed + i, shard_index=len(tasks) // shard,
                    lineage_id=lineage, macro_path_origin=src.macro_path_origin))
    return tuple(tasks)


def build_rve_supplemental_task_manifest(plan: RVEDataPlan) -> tuple[RVEDataTask, ...]:
    """The optional 10-paths-per-stratum extension of the fixed-cell plan, split
    700 / 150 / 150 by the sha256 rank of its task ids."""
    src = plan.sources[0]
    d = src.production_design
    if plan.purpose != "fixed_cell_rve_training" or d is None:
        raise PlanError("RVE supplemental paths require the c48 production design.")
    lc = src.load_cases[0]
    base = len(build_rve_data_task_manifest(plan))
    shard = plan.parallel.shard_path_count
    tasks = []
    for family in lc.path_families:
        for ab in d.amplitude_bands:
            for rb in d.rate_bands:
                stratum = f"{family}|{ab.name}|{rb.name}"
                for s in range(10):
                    ri = s % 4
                    idx = base + len(tasks)
                    tid = (f"{src.name}__supplement__{family}__{ab.name}__{rb.name}"
                           f"__sample_{s:02d}")
                    tasks.append(RVEDataTask(
                        task_id=tid, source_name=src.name, source_kind=src.kind,
                        load_case_name=lc.name, path_index=idx, seed=lc.seed + idx,
                        shard_index=idx // shard, lineage_id=tid,
                        macro_path_origin=src.macro_path_origin, family=family,
                        amplitude_band=ab.name, rate_band=rb.name,
                        tensor_direction=d.tensor_directions[s % 5], repeat_index=ri,
                        amplitude=ab.lower + (s + 1.0) / 11.0 * (ab.upper - ab.lower),
                        target_strain_rate=rb.lower + (((3 * s + 1) % 10) + 1.0) / 11.0
                        * (rb.upper - rb.lower),
                        accepted_increment_count=d.accepted_increment_counts[ri],
                        turning_angle=(d.nonproportional_turning_angles[ri]
                                       if family == "nonproportional" else None),
                        design_stratum=f"supplement|{stratum}",
                        expected_response_regime=ab.expected_response))
    order = sorted(range(len(tasks)),
                   key=lambda i: hashlib.sha256(tasks[i].task_id.encode("utf-8")).hexdigest())
    split = {}
    for rank, i in enumerate(order):
        split[i] = "train" if rank < 700 else ("validation" if rank < 850 else "test")
    return tuple(dataclasses.replace(t, preassigned_split=split[i]) for i, t in enumerate(tasks))


# ---------------------------------------------------------------------------
# control histories

@dataclasses.dataclass(frozen=True)
class GeneratedControlPath:
    task_id: str
    control_kind: str  # "macro_strain" or "pressure"
    times: np.ndarray
    values: np.ndarray
    family: str


def _smooth_controls(u, controls):
    """Scalar controls joined by a cubic smoothstep on each segment of u in [0, 1]."""
    c = np.asarray(controls, dtype=np.float64)
    x = u * (c.size - 1)
    seg = np.minimum(x.astype(np.int64), c.size - 2)
    loc = x - seg
    return c[seg] + loc * loc * (3.0 - 2.0 * loc) * (c[seg + 1] - c[seg])


def generate_data_control_path(source: DataSource, task: RVEDataTask) -> GeneratedControlPath:
    """The control history of a task: the production strain path of a fixed-cell
    task (analytic timing, ``rve_path_generation.py:55-103``), the impact pressure
    half-sine, or a legacy strain family along a seeded unit direction
    (``rve_path_generation.py:34-53, 198-234``)."""
    lc = next(c for c in source.load_cases if c.name == task.load_case_name)
    if source.production_design is not None:
        d = source.production_design
        t, e = production_strain_path(task.family, task.tensor_direction, float(task.amplitude),
                                      float(task.target_strain_rate),
                                      int(task.accepted_increment_count), seed=task.seed,
                                      perturbation=d.direction_perturbation,
                                      turning_angle=task.turning_angle)
        return GeneratedControlPath(task.task_id, "macro_strain", t, e, str(task.family))
    family = lc.path_families[task.path_index % len(lc.path_families)]
    times = np.linspace(0.0, lc.duration, lc.time_points)
    if source.kind == "impact_path":
        p = lc.peak_scale * np.sin(np.pi * times / lc.duration)
        return GeneratedControlPath(task.task_id, "pressure", times, p[:, None], family)
    rng = np.random.default_rng(task.seed)
    direction = rng.normal(size=6)
    direction /= np.linalg.norm(direction)
    u = (times - times[0]) / (times[-1] - times[0])
    controls = dict(FAMILY_CONTROLS, nonproportional=(0.0, 1.0, 0.5, 1.1))[family]
    strain = lc.peak_scale * _smooth_controls(u, controls)[:, None] * direction[None, :]
    if family == "nonproportional":
        second = rng.normal(size=6)
        second -= direction * float(second @ direction)
        second /= np.linalg.norm(second)
        strain = strain + 0.5 * lc.peak_scale * np.sin(np.pi * u)[:, None] * second[None, :]
    return GeneratedControlPath(task.task_id, "macro_strain", times, strain, family)


# ---------------------------------------------------------------------------
# device production

def _fixed(params: dict, name: str) -> float:
    spec = params[name]
    if spec["kind"] != "fixed":
        raise PlanError(f"{name} must be a fixed parameter for production.")
    return float(spec["value"])


def fixed_cell_voxel_model(source: DataSource, divisions: int, *, newton_rtol: float = 1e-10,
                           precision: str = "fp64"):
    """A structured periodic cell of ``divisions``^3 Hex8 (unit cube) with the
    source's phases: a J2 Perzyna matrix and an elastic cylindrical inclusion along
    z, centred, whose voxel volume fraction is the closest to the plan's
    ``inclusion_volume_fraction``.  Material parameters from the plan (density 1, as
    the reference's c48 fields).  Returns (model, element phase ids)."""
    from concrete_impact_ai_amd import materials as M
    from concrete_impact_ai_amd import rve as R
    from concrete_impact_ai_amd.solvers import (ArmijoSettings, LinearSolverSettings,
                                                NonlinearNewtonSettings)

    mp, sp = source.material_parameters, source.microstructure_parameters
    n = int(divisions)
    c = (np.arange(n) + 0.5) / n - 0.5
    r2 = c[:, None] ** 2 + c[None, :] ** 2  # (x, y) of the element centres
    fraction = _fixed(sp, "inclusion_volume_fraction")
    # the radius whose voxelised disc area is the closest to the target fraction
    cand = np.unique(r2)
    frac = np.asarray([np.mean(r2 <= v) for v in cand])
    radius2 = cand[int(np.argmin(np.abs(frac - fraction)))]
    disc = (r2 <= radius2).astype(np.int32)  # [ix, iy]
    eph = np.broadcast_to(disc.T[None, :, :], (n, n, n)).reshape(-1)  # e = ix + n (iy + n iz)
    matrix = M.J2ViscoplasticMaterial(
        "matrix", 1.0, _fixed(mp, "matrix_young_modulus"), _fixed(mp, "matrix_poisson_ratio"),
        _fixed(mp, "matrix_yield_stress"), _fixed(mp, "matrix_hardening_modulus"),
        _fixed(mp, "matrix_time_scale"), _fixed(mp, "matrix_yield_stress"),
        _fixed(mp, "matrix_rate_exponent"))
    inclusion = M.LinearElasticMaterial("inclusion", 1.0, _fixed(mp, "inclusion_young_modulus"),
                                        _fixed(mp, "inclusion_poisson_ratio"))
    nodes, elements = R.build_structured_hex8_box((1.0, 1.0, 1.0), (n, n, n))
    sig = _fixed(mp, "matrix_yield_stress")
    newton = NonlinearNewtonSettings(25, newton_rtol * sig, newton_rtol, 1e-14, newton_rtol,
                                     ArmijoSettings(True, 1e-4, 0.5, 20))
    model = R.StructuredHex8RVE(
        nodes, elements, (n, n, n),
        M.PhasedMaterialModel("fixed_cell", (M.RVEPhase("matrix", matrix),
                                             M.RVEPhase("inclusion", inclusion)),
                              np.repeat(eph, 8), 1.0),
        R.build_structured_hex8_periodic_constraint(nodes, (n, n, n)), newton,
        element_phase_ids=np.ascontiguousarray(eph), phase_names=("matrix", "inclusion"),
        linear_settings=LinearSolverSettings(method="spectral", relative_tolerance=1e-10,
                                             adaptive_forcing=True),
        precision=precision)
    return model, np.ascontiguousarray(eph)


def produce_plan_responses(source: DataSource, tasks, model, *, progress=None) -> dict:
    """Macro responses of the tasks' strain paths on ``model``, on the device: the
    tasks are grouped by increment count and every group is advanced together, one
    batched FE² solve per increment (each path with its own time step).  Returns
    {task_id: {"time", "time_step", "macro_strain", "macro_stress",
    "free_energy_density", "dissipation_density", "viscoplastic_active_volume_fraction"}}
    with one row per increment."""
    from concrete_impact_ai_amd import batch as B
    from concrete_impact_ai_amd.materials import MaterialUpdateSettings

    settings = MaterialUpdateSettings(50, 1e-10, 1e-12, 1e-10)
    paths = {t.task_id: generate_data_control_path(source, t) for t in tasks}
    groups: dict[int, list] = {}
    for t in tasks:
        groups.setdefault(len(paths[t.task_id].times) - 1, []).append(t)
    out = {}
    for steps, group in sorted(groups.items()):
        batch = B.RVEBatch([model] * len(group))
        state = batch.initial_state()
        times = np.stack([paths[t.task_id].times for t in group])
        strains = np.stack([paths[t.task_id].values for t in group])
        rows = {k: [] for k in ("macro_stress", "free_energy_density", "dissipation_density",
                                "viscoplastic_active_volume_fraction")}
        for k in range(1, steps + 1):
            r = B.solve_rve_batch(batch, strains[:, k], times[:, k] - times[:, k - 1], settings,
                                  state)
            state = r.state
            rows["macro_stress"].append(np.asarray(r.macro_stress))
            rows["free_energy_density"].append(np.asarray(r.free_energy_density))
            rows["dissipation_density"].append(np.asarray(r.dissipation_density))
            rows["viscoplastic_active_volume_fraction"].append(
                np.asarray(r.diagnostics["viscoplastic_active_volume_fraction"]))
            if progress is not None and (k % 16 == 0 or k == steps):
                progress(steps, k)
        batch.close()
        for i, t in enumerate(group):
            out[t.task_id] = {"time": times[i, 1:], "time_step": np.diff(times[i]),
                              "macro_strain": strains[i, 1:],
                              **{k: np.stack(v)[:, i] for k, v in rows.items()}}
    return out

