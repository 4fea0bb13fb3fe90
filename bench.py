"""Benchmark: periodic two-phase RVE increments (FE² micro solve) on one MI355X.

Workload (the north-star metric's configuration that fits one GPU): periodic
RVE of 48 x 48 x 48 Hex8 elements (110,592 elements, 884,736 Gauss points),
spherical elastic aggregates (E 50 GPa, nu 0.2) at 40 % volume fraction with
radii U[2, 6] elements in a J2 Perzyna mortar (E 25 GPa, nu 0.2, sigma_y 20 MPa,
H 0.5 GPa, eta 1e-3 s, m 1), driven along a seeded random 6-component
piecewise-linear macro strain path (200 steps, |eps| <= 3e-3, rates 10-1000 1/s).
A step is one converged RVE increment (Newton + Armijo + PCG) with inputs
resident in HBM; the effective tangent is not requested (explicit FE² use).

Prints one JSON line (see the repository's measurement notes for the
roofline and CPU-baseline definitions).
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Gauss-pt updates/s & RVE solves/s; surrogate vs plain FE² speedup at ≤2% err"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6

KERNEL_NAMES = {"pcg_elem": "k_sp_elem", "pcg_node": "k_sp_node", "pcg_update": "k_sp_xfwd",
                "assemble": "k_assemble",
                "precond": "spectral preconditioner slot: k_dft_strided (forward y) + k_sp_zcol + "
                           "k_dft_strided (inverse y) + k_sp_xinv",
                "gather": "k_gather + k_finalize"}
# RVE edges of 16, 32, 48 and 64 run the row-lane spectral passes (same bytes per class)
KERNEL_NAMES_ROW_LANE = dict(KERNEL_NAMES, pcg_update="k_sp_xlines_fwd",
                             precond="spectral preconditioner slot: k_sp_ylines (forward y) + "
                                     "k_sp_zlines + k_sp_ylines (inverse y) + k_sp_xlines_inv")

# HBM bytes per launch of a kernel from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of
# this script (--copy-calibration), each counter scaled by the ratio of the bytes a plain
# copy kernel of known size moves to what the same counter reads for it.  Keyed by
# (slot, RVE edge, precision); absent keys report null.  Source summary:
# profiles/k_sp_elem_rve48_fp64_hbm_traffic.md, profiles/spectral_precond_rve48_fp64_hbm_traffic.md
MEASURED_TRAFFIC: dict = {("pcg_elem", 48, "fp64"): 77_042_165, ("precond", 48, "fp64"): 45_494_701}


def copy_calibration(torch) -> None:
    """A plain device copy of 256 MiB, run 4 times: the calibration kernel for
    FETCH_SIZE / WRITE_SIZE (each launch reads and writes exactly 256 MiB)."""
    src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(4):
        dst.copy_(src)
    torch.cuda.synchronize()
    print("copy calibration: 4 launches x 268435456 B read + 268435456 B written",
          file=sys.stderr, flush=True)


def build_model(n: int, seed: int, precision: str, pcg_rtol: float, forcing: bool = True,
                newton_rtol: float = 1.0e-10):
    from concrete_impact_ai_amd import materials as M
    from concrete_impact_ai_amd import rve as R
    from concrete_impact_ai_amd.microstructure import spherical_aggregates
    from concrete_impact_ai_amd.solvers import (ArmijoSettings, LinearSolverSettings,
                                                NonlinearNewtonSettings)

    eph = spherical_aggregates((n, n, n), 0.40, (2.0, 6.0), seed=seed)
    mortar = M.J2ViscoplasticMaterial("mortar", 2400.0, 25.0e9, 0.2, 20.0e6, 0.5e9, 1.0e-3, 20.0e6, 1.0)
    agg = M.LinearElasticMaterial("aggregate", 2400.0, 50.0e9, 0.2)
    nodes, elements = R.build_structured_hex8_box((1.0, 1.0, 1.0), (n, n, n))
    # Newton tolerance 1e-10 (relative; 1e-6 for FP32 element work, its residual floor);
    # absolute tolerances force-scaled for SI units: rtol * sigma_y * L^2 for the
    # residual, 1e-12 m for the increment.
    newton = NonlinearNewtonSettings(25, newton_rtol * 20.0e6, newton_rtol, 1.0e-12, newton_rtol,
                                     ArmijoSettings(True, 1.0e-4, 0.5, 20))
    lin = LinearSolverSettings(method="spectral", max_iterations=5000, relative_tolerance=pcg_rtol,
                               check_interval=8, adaptive_forcing=forcing)
    model = R.StructuredHex8RVE(
        nodes, elements, (n, n, n),
        M.PhasedMaterialModel("c2_rve", (M.RVEPhase("mortar", mortar), M.RVEPhase("aggregate", agg)),
                              np.repeat(eph, 8), 2400.0),
        R.build_structured_hex8_periodic_constraint(nodes, (n, n, n)), newton,
        element_phase_ids=eph, phase_names=("mortar", "aggregate"), linear_settings=lin,
        precision=precision)
    return model, eph


def cpu_baseline(n: int, seed: int, n_cpu: int = 12, k: int = 3) -> dict:
    """The reference algorithm on the host, timed in full: the oracle (FP64 NumPy port
    of the reference's point updates, Hex8 assembly and Newton / Armijo with SciPy
    ``spsolve`` = SuperLU, ``src/fem/rve/solver.py:62-227``, ``src/fem/solvers/
    linear.py:15-22``) on one core (BLAS limited to one thread; SuperLU is serial).

    SuperLU's factorisation of the 3-D periodic stiffness grows roughly as DOF^2
    (2.1 s per solve at 12^3 with 5 181 DOFs, ~10 s at 16^3 on the build host), so one
    increment of the 48^3 RVE (331 773 DOFs) would take hours.  The timed sample is
    the same workload -- microstructure generator, seed and volume fraction, phases,
    the seeded 200-step path, Newton and Armijo settings -- on an ``n_cpu``^3 RVE: ``k``
    increments (3 .. 2 + k of the path), each timed in full, median and MAD.  Beside it:
    SuperLU solve times at 8^3 and ``n_cpu``^3 and the sample's assembly time
    (measured), and the 48^3 increment extrapolated from them (marked as such)."""
    import os

    try:
        from threadpoolctl import threadpool_limits
    except ImportError:  # pragma: no cover
        threadpool_limits = None
    from concrete_impact_ai_amd.microstructure import random_strain_path, spherical_aggregates
    from oracle import j2vp as OJ
    from oracle import rve as OR

    def model_for(edge):
        eph = spherical_aggregates((edge, edge, edge), 0.40, (2.0, 6.0), seed=seed)
        return OR.Model((1.0, 1.0, 1.0), (edge, edge, edge),
                        [("mortar", OJ.J2VPParams(25e9, 0.2, 20e6, 0.5e9, 1e-3, 20e6, 1.0)),
                         ("aggregate", OJ.ElasticParams(50e9, 0.2))],
                        OR.Newton(25, 2e-3, 1e-10, 1e-12, 1e-10, OR.Armijo(True)),
                        element_phase_ids=eph)

    times, strains = random_strain_path(200, 8, 3.0e-3, (10.0, 1000.0), seed=seed)
    solve_times = []
    real_spsolve = OR.spsolve

    def timed_spsolve(a, b):
        t0 = time.perf_counter()
        x = real_spsolve(a, b)
        solve_times.append(time.perf_counter() - t0)
        return x

    ctx = threadpool_limits(limits=1) if threadpool_limits else None
    OR.spsolve = timed_spsolve
    try:
        def run(edge, incs):
            model = model_for(edge)
            st = model.initial_state()
            out = []
            for j in range(1, 3 + incs):
                del solve_times[:]
                t0 = time.perf_counter()
                r = OR.solve_increment(model, strains[j], times[j] - times[j - 1], OJ.UpdateSettings(),
                                       st, need_tangent=False)
                dt_inc = time.perf_counter() - t0
                st = r["state"]
                if j >= 3:
                    out.append((dt_inc, r["diagnostics"]["newton_iterations"],
                                r["diagnostics"]["armijo_backtracks"], list(solve_times)))
            return out

        small = run(8, 2)
        sample = run(n_cpu, k)
    finally:
        OR.spsolve = real_spsolve
        if ctx is not None:
            ctx.__exit__(None, None, None)
    inc = np.asarray([s[0] for s in sample])
    med = float(np.median(inc))
    mad = float(np.median(np.abs(inc - med)))
    lu8 = float(np.median([t for s in small for t in s[3]]))
    lu = float(np.median([t for s in sample for t in s[3]]))
    dof = lambda e: 3 * e ** 3 - 3  # noqa: E731  (reduced periodic DOFs)
    expo = np.log(lu / lu8) / np.log(dof(n_cpu) / dof(8))
    lu48 = lu * (dof(n) / dof(n_cpu)) ** expo
    newton = float(np.mean([s[1] for s in sample]))
    backtracks = float(np.mean([s[2] for s in sample]))
    # one full assembly of the sample (the increment less its SuperLU solves, over its
    # 1 + 2 newton + backtracks assemblies), scaled by the number of Gauss points
    per_asm = float(np.median([(s[0] - sum(s[3])) / (1.0 + 2.0 * s[1] + s[2]) for s in sample]))
    t_asm48 = per_asm * (n / n_cpu) ** 3
    inc48 = (1.0 + 2.0 * newton + backtracks) * t_asm48 + newton * lu48
    return {
        "value": 1.0 / med,
        "unit": f"RVE solves/s ({n_cpu}^3 sample of the {n}^3 workload)",
        "cores": 1,
        "host_cores_available": len(os.sched_getaffinity(0)),
        "kind": "port",
        "sample": (f"oracle FP64 (reference algorithm, SuperLU) on the {n_cpu}^3 RVE of the same "
                   f"generator, seed, phases, path and Newton/Armijo settings: increments 3-{2 + k} "
                   f"timed in full, median {med:.2f} s +- MAD {mad:.2f} s, one core"),
        "increment_seconds": inc.tolist(),
        "newton_iterations_per_increment": newton,
        "superlu_seconds_per_solve": {"8": lu8, str(n_cpu): lu},
        "assembly_seconds": {str(n_cpu): per_asm},
        "extrapolated_48": {"increment_seconds": inc48, "rve_solves_per_s": 1.0 / inc48,
                            "assembly_seconds": t_asm48,
                            "superlu_seconds_per_solve": lu48,
                            "superlu_exponent_in_dofs": float(expo),
                            "note": f"{n_cpu}^3 assembly time x Gauss-point ratio, at the sample's "
                                    "Newton/Armijo counts, + Newton solves x the SuperLU time "
                                    f"extrapolated from 8^3 and {n_cpu}^3 in the DOF count "
                                    "(extrapolated, not timed)"},
    }


def plan_section(args) -> dict:
    """The reference's fixed-cell training-data plan produced on the device: the 2000
    tasks of ``configs/experiments/rve_rno_c48_v2_2000.yaml`` (the copy under
    tests/golden/configs), their production strain paths (64-192 increments), the macro
    responses from the batched FP64 FE² engine on a ``--plan-rve``^3 voxel cell with
    the plan's phases (tasks grouped by increment count, one batched solve per
    increment; a structured stand-in for the reference's O-grid cell)."""
    import torch

    from concrete_impact_ai_amd import data_plan as DP

    plan = DP.load_rve_data_plan(Path(__file__).resolve().parent / "tests" / "golden" / "configs"
                                 / "rve_rno_c48_v2_2000.yaml")
    src = plan.sources[0]
    tasks = DP.build_rve_data_task_manifest(plan)
    model, eph = DP.fixed_cell_voxel_model(src, args.plan_rve)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = DP.produce_plan_responses(src, tasks, model)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    solves = sum(t.accepted_increment_count for t in tasks)
    act = np.asarray([np.max(out[t.task_id]["viscoplastic_active_volume_fraction"]) for t in tasks])
    return {"workload": f"{len(tasks)} plan tasks (5 families x 5 amplitude x 4 rate bands x 20), "
                        f"64-192 increments, FP64 batched FE2 on a {args.plan_rve}^3 voxel cell",
            "paths": len(tasks), "rve_solves": solves, "seconds": el, "rve_solves_per_s": solves / el,
            "inclusion_fraction": float(np.mean(eph == 1)),
            "yielding_paths": int(np.sum(act > 0.0)),
            "splits": {k: sum(t.preassigned_split == k for t in tasks)
                       for k in ("train", "validation", "test")}}


def c1_section() -> dict:
    """The plain-method RVE on the device: the 8^3 homogeneous J2 Perzyna RVE in SI
    units (E 30 GPa, nu 0.2, sigma_y 30 MPa, H 1 GPa, eta 1e-3 s, m 1), uniaxial strain
    ramped to 2e-3 at 100 1/s over 100 steps through ``iterate_prescribed_rve_path``,
    Newton relative tolerance 1e-10 with the absolute tolerance scaled to the forces.
    One untimed pass, then the whole path timed (median of three)."""
    import torch

    from concrete_impact_ai_amd import materials as M
    from concrete_impact_ai_amd import rve as R
    from concrete_impact_ai_amd.materials import MaterialResponseRequirements
    from concrete_impact_ai_amd.solvers import ArmijoSettings, NonlinearNewtonSettings

    E, nu, sy, H, tau = 30.0e9, 0.2, 30.0e6, 1.0e9, 1.0e-3
    mat = M.J2ViscoplasticMaterial("c1_matrix", 2400.0, E, nu, sy, H, tau, sy, 1.0)
    div = (8, 8, 8)
    nodes, elements = R.build_structured_hex8_box((1.0, 1.0, 1.0), div)
    newton = NonlinearNewtonSettings(20, 1e-10 * sy, 1e-10, 1e-12, 1e-10, ArmijoSettings(True))
    model = R.StructuredHex8RVE(nodes, elements, div, mat,
                                R.build_structured_hex8_periodic_constraint(nodes, div), newton)
    steps, dt = 100, 2.0e-7
    times = dt * np.arange(steps + 1)
    strains = np.zeros((steps + 1, 6))
    strains[:, 0] = 2.0e-3 * np.arange(steps + 1) / steps
    settings = M.MaterialUpdateSettings(50, 1.0e-10, 1.0e-12, 1.0e-10)
    req = MaterialResponseRequirements(tangent=False)
    runs = []
    last = None
    for rep in range(4):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for s_ in R.iterate_prescribed_rve_path(model, times, strains, settings, requirements=req):
            last = s_
        torch.cuda.synchronize()
        if rep:
            runs.append(time.perf_counter() - t0)
    med = float(np.median(runs))
    return {"workload": "8^3 homogeneous J2 Perzyna RVE (SI), 100-step uniaxial ramp to 2e-3 at "
                        "100 1/s, FP64, Newton 1e-10",
            "path_seconds": runs, "rve_solves_per_s": steps / med,
            "final_stress_xx": float(last.response.macro_stress[0])}


def cpu_baseline_paths(seed: int) -> dict:
    """More of the reference algorithm on the host, one core (BLAS limited to one
    thread): the plain-method RVE (8^3 homogeneous J2 Perzyna RVE in SI units,
    uniaxial strain ramped to 2e-3 at 100 1/s over 100 steps, Newton 1e-10) run in
    full with the vectorised oracle, three times (median and MAD), and its first
    three increments with the reference's literal per-point and per-element loops
    (``oracle/loops.py``, the form of ``plasticity.py:343-431`` and
    ``nonlinear_solid.py:460-493``); and five explicit velocity-Verlet steps of a
    32 x 32 x 16 sample of the macro slab with the oracle's dynamics (plain J2,
    ``src/fem/dynamics/material.py:89-297``), with the 128 x 128 x 64 slab step
    scaled from it by the element count (marked as extrapolated); and the surrogate
    network's forward pass in torch on the host (FP32, every host core)."""
    import os

    from threadpoolctl import threadpool_limits

    from oracle import dynamics as OD
    from oracle import j2vp as OJ
    from oracle import loops as OLP
    from oracle import rve as OR

    out = {"cores": 1, "blas_threads": 1, "host_cores_available": len(os.sched_getaffinity(0))}
    with threadpool_limits(limits=1):
        E, nu, sy, H, tau = 30.0e9, 0.2, 30.0e6, 1.0e9, 1.0e-3
        mat = OJ.J2VPParams(E, nu, sy, H, tau, sy, 1.0)
        kw = dict(lengths=(1.0, 1.0, 1.0), divisions=(8, 8, 8), phases=[("c1", mat)],
                  newton=OR.Newton(20, 1e-10 * sy, 1e-10, 1e-12, 1e-10, OR.Armijo(True)))
        steps, dt = 100, 2.0e-7
        times = dt * np.arange(steps + 1)
        strains = np.zeros((steps + 1, 6))
        strains[:, 0] = 2.0e-3 * np.arange(steps + 1) / steps
        runs = []
        for _ in range(3):
            t0 = time.perf_counter()
            OR.solve_path(OR.Model(**kw), times, strains, OJ.UpdateSettings(), need_tangent=False)
            runs.append(time.perf_counter() - t0)
        med = float(np.median(runs))
        out["c1_vectorised"] = {
            "path_seconds": runs, "median_seconds": med,
            "mad_seconds": float(np.median(np.abs(np.asarray(runs) - med))),
            "rve_solves_per_s": steps / med,
            "sample": "8^3 SI RVE, 100-step uniaxial ramp, vectorised oracle + SuperLU, whole path"}
        model = OLP.LoopModel(**kw)
        st = model.initial_state()
        per = []
        for k in range(1, 4):
            t0 = time.perf_counter()
            st = OR.solve_increment(model, strains[k], dt, OJ.UpdateSettings(), st,
                                    need_tangent=False)["state"]
            per.append(time.perf_counter() - t0)
        mp = float(np.median(per))
        out["c1_literal_loops"] = {
            "increment_seconds": per, "rve_solves_per_s": 1.0 / mp,
            "path_seconds_extrapolated": steps * mp,
            "sample": "the same RVE, increments 1-3, per-point and per-element Python loops "
                      "(the 100-step path time is 100 x the median increment: extrapolated)"}
        nx, ny, nz, h = 32, 32, 16, 0.01
        slab = OD.Box((nx * h, ny * h, nz * h), (nx, ny, nz),
                      [("slab", OJ.J2VPParams(30e9, 0.2, 30e6, 1e9, 1e-3, 30e6, 1.0, 2400.0))])
        f_unit = OD.top_patch_unit_pressure(slab.lengths, (nx, ny, nz), ((11, 21), (11, 21)))
        bottom = np.flatnonzero(np.isclose(slab.nodes[:, 2], 0.0))
        fixed = (3 * bottom[:, None] + np.arange(3)).ravel()
        dt_m = 0.5 * h / float(np.sqrt(30e9 * 0.8 / (1.2 * 0.6) / 2400.0))
        t0 = time.perf_counter()
        OD.explicit(slab, dt_m, 5, fixed, f_unit, OD.half_sine(50e6, 2e-4), OJ.UpdateSettings())
        per_step = (time.perf_counter() - t0) / 5.0
        out["macro_explicit_sample"] = {
            "seconds_per_step": per_step, "elements": nx * ny * nz,
            "slab_128x128x64_seconds_per_step_extrapolated": per_step * 64.0,
            "slab_128x128x64_steps_per_s_extrapolated": 1.0 / (per_step * 64.0),
            "sample": "32x32x16 slab, 5 explicit steps (initial assembly included), vectorised "
                      "oracle; the 128x128x64 figure scales by the element count"}
    # the surrogate network (13 -> 256^4 -> 12, GELU) in torch on the host, FP32, all cores
    import torch

    n_thr = len(os.sched_getaffinity(0))
    prev = torch.get_num_threads()
    torch.set_num_threads(n_thr)
    try:
        g = torch.Generator().manual_seed(seed)
        dims = [13, 256, 256, 256, 256, 12]
        ws = [torch.randn(dims[i + 1], dims[i], generator=g) / dims[i] ** 0.5 for i in range(5)]
        bs = [torch.zeros(dims[i + 1]) for i in range(5)]
        x = torch.randn(1 << 16, 13, generator=g)

        def net(v):
            for i in range(5):
                v = torch.nn.functional.linear(v, ws[i], bs[i])
                if i < 4:
                    v = torch.nn.functional.gelu(v, approximate="tanh")
            return v

        with torch.no_grad():
            net(x)
            reps = []
            for _ in range(3):
                t0 = time.perf_counter()
                net(x)
                reps.append(time.perf_counter() - t0)
        med = float(np.median(reps))
        out["surrogate_inference_torch_cpu"] = {
            "points_per_s": x.shape[0] / med, "seconds_per_2^16_points": reps, "threads": n_thr,
            "sample": "the surrogate network's forward pass at 2^16 points, torch FP32 on the "
                      "host, every host core (median of three after one warm-up)"}
    finally:
        torch.set_num_threads(prev)
    return out


BF16_PEAK_TFLOPS = 2500.0
MLP_FLOP_PER_POINT = 2 * (13 * 256 + 3 * 256 * 256 + 256 * 12)


def surrogate_section(args) -> dict:
    """Surrogate vs plain FE² (the recurrent material-point surrogate at its stated
    scale): FE² data of ``--surrogate-paths`` training + ``--surrogate-test-paths``
    held-out paths of ``--surrogate-steps`` steps on a ``--surrogate-rve``^3 two-phase
    RVE (FP64, the batched engine), training by the fused BPTT kernels, the held-out
    stress error, the fused-inference rate at 2^20 points with its matrix-core
    roofline, and one BPTT iteration at 4096 paths x 32 steps with its roofline.

    Training schedule (every training path in every iteration): the first 64 steps
    (``--surrogate-iters``, lr 2e-3), then whole paths (``--surrogate-full-iters``, lr
    3e-4).  Training stages on 32-step windows overfit the first steps of the 200-step
    paths (held-out error 6.7 % RMS against 0.9 % on the same budget), so the 4096 x 32
    shape is the measured BPTT step, not a training stage."""
    import torch

    from concrete_impact_ai_amd import surrogate as S
    from concrete_impact_ai_amd import surrogate_data as SD

    n, T = args.surrogate_rve, args.surrogate_steps
    p_train, p_test = args.surrogate_paths, args.surrogate_test_paths
    model = SD.two_phase_rve(n, seed=args.seed)
    times, strains = SD.training_paths(p_train + p_test, T, seed=1000 + args.seed,
                                       production_fraction=args.surrogate_production_fraction)
    print(f"surrogate: FE² data {p_train + p_test} paths x {T} steps on {n}^3 RVEs",
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    d_eps, rate, stress, fe2_rate = SD.rve_responses_batched(model, times, strains, progress=sys.stderr)
    data_s = time.perf_counter() - t0
    tr = slice(0, p_train)
    te = slice(p_train, p_train + p_test)
    scales = S.SurrogateScales.fit(d_eps[tr], rate[tr], stress[tr], SD.homogenised_stiffness(model))
    params = S.SurrogateParams(seed=args.seed)
    data = S.BpttData.from_paths(scales, d_eps[tr], rate[tr], stress[tr])
    stages = [(args.surrogate_iters, min(64, T), p_train, 2e-3),
              (args.surrogate_full_iters, T, p_train, 3e-4)]
    print(f"surrogate: training {stages[0][0]} + {stages[1][0]} BPTT iterations", file=sys.stderr,
          flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist, skipped = [], 0
    for k, (iters, window, batch, lr) in enumerate(stages):
        if iters <= 0:
            continue
        h = S.train(params, scales, None, None, None, iterations=iters, batch_paths=batch,
                    window=window, lr=lr, seed=args.seed + k, log_every=max(1, iters // 5),
                    progress=sys.stderr, data=data)
        hist += [(it + sum(s[0] for s in stages[:k]), loss) for it, loss in h]
        skipped += h.skipped_steps
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    S.envelope(params, scales, d_eps[tr], rate[tr])
    dev = torch.device("cuda")
    pred, _ = S.rollout(params, scales, torch.as_tensor(d_eps[te], dtype=torch.float32, device=dev),
                        torch.as_tensor(rate[te], dtype=torch.float32, device=dev), fused=True)
    errs = SD.stress_errors(pred.double().cpu().numpy(), stress[te])
    rms = SD.stress_rms_errors(pred.double().cpu().numpy(), stress[te])
    train_errs = SD.stress_errors(
        S.rollout(params, scales, torch.as_tensor(d_eps[tr][:64], dtype=torch.float32, device=dev),
                  torch.as_tensor(rate[tr][:64], dtype=torch.float32, device=dev),
                  fused=True)[0].double().cpu().numpy(), stress[tr][:64])
    # fused inference at 2^20 points
    M = 1 << 20
    X = torch.randn((M, 13), dtype=torch.float32, device=dev) * 0.5
    for _ in range(3):
        S.forward_fused(X, params)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 20
    e0.record()
    for _ in range(reps):
        S.forward_fused(X, params)
    e1.record()
    torch.cuda.synchronize()
    inf_ms = e0.elapsed_time(e1) / reps
    tflops = MLP_FLOP_PER_POINT * M / (inf_ms * 1e-3) / 1e12
    # one BPTT iteration (cia_surrogate_bptt: forward and backward through time, weight
    # gradients) at 4096 paths x 32 steps of the training data, HIP events on its stream
    Bp, W = 4096, 32
    tmp = S.SurrogateParams(seed=1)
    bstep = S.BpttStep(tmp, Bp, W)
    idx = torch.as_tensor(np.random.default_rng(3).choice(p_train, Bp), dtype=torch.int32, device=dev)
    pw = data.path_weights(W)
    for _ in range(3):
        bstep(tmp, data, idx, pw)
    reps = 20
    e0.record()
    for _ in range(reps):
        bstep(tmp, data, idx, pw)
    e1.record()
    torch.cuda.synchronize()
    bptt_ms = e0.elapsed_time(e1) / reps
    bptt_tf = 3 * MLP_FLOP_PER_POINT * Bp * W / (bptt_ms * 1e-3) / 1e12
    pts_per_s = M / (inf_ms * 1e-3)
    return {
        "rve_edge": n,
        "fe2_rve_solves_per_s": fe2_rate,
        "fe2_data": f"{p_train + p_test} paths x {T} steps, one batched FP64 FE² solve per step",
        "fe2_data_seconds": data_s,
        "train_paths": p_train,
        "test_paths": p_test,
        "train_schedule": [{"iterations": i, "window": w, "batch_paths": b, "lr": lr}
                           for i, w, b, lr in stages],
        "train_iterations": sum(s[0] for s in stages),
        "train_seconds": train_s,
        "train_skipped_steps": skipped,
        "loss_history": hist,
        "test_stress_error_median": float(np.median(errs)),
        "test_stress_error_p95": float(np.percentile(errs, 95)),
        "test_stress_error_max": float(np.max(errs)),
        "train_stress_error_median": float(np.median(train_errs)),
        "test_stress_rms_error_median": float(np.median(rms)),
        "test_stress_rms_error_p95": float(np.percentile(rms, 95)),
        "error_definition": "per held-out path: max error = max_t |sigma_sur - sigma_FE2|_inf / "
                            "max_t |sigma_FE2|_inf; rms error = sqrt(mean (sigma_sur - sigma_FE2)^2) "
                            "/ max_t |sigma_FE2|_inf",
        "inference_points_per_s": pts_per_s,
        "inference_ms_2^20": inf_ms,
        "inference_roofline": {"bound": "mfma", "achieved": tflops, "peak": BF16_PEAK_TFLOPS,
                               "unit": "TFLOP/s", "frac": tflops / BF16_PEAK_TFLOPS},
        "bptt_ms_4096x32": bptt_ms,
        "bptt_roofline": {"kernel": "cia_surrogate_bptt (k_bptt_prep, k_bptt_fwd, k_bptt_bwd, "
                                    "k_wgrad_nt x5, k_bptt_reduce, k_grad_norm)",
                          "bound": "mfma", "achieved": bptt_tf, "peak": BF16_PEAK_TFLOPS,
                          "unit": "TFLOP/s", "frac": bptt_tf / BF16_PEAK_TFLOPS,
                          "flop_per_iteration": 3 * MLP_FLOP_PER_POINT * Bp * W},
        "speedup_vs_fe2_per_point_update": pts_per_s / fe2_rate,
        "meets_2pct_rms_median": bool(np.median(rms) <= 0.02),
        "meets_2pct_max_p95": bool(np.percentile(errs, 95) <= 0.02),
    }, {"model": model, "params": params, "scales": scales, "fe2_rate": fe2_rate}


def _impact_slab(divisions, material, patch_elems, h=0.01):
    """Hex8 slab (spacing h) with the bottom fixed and the 50 MPa / 0.2 ms half-sine
    pulse on a centred patch of the top face."""
    from concrete_impact_ai_amd import dynamics as D
    from concrete_impact_ai_amd.solvers import DirichletDofSet

    nx, ny, nz = divisions
    bundle = D.build_structured_box_bundle((nx * h, ny * h, nz * h), (nx, ny, nz), material)
    c0 = (nx - patch_elems) // 2, (ny - patch_elems) // 2
    load = D.build_half_sine_pressure_load(bundle, 50e6, 2e-4, patch=(
        (c0[0], c0[0] + patch_elems), (c0[1], c0[1] + patch_elems)))
    bottom = bundle.mesh_info.boundary_groups["bottom"].nodes
    dofs = (3 * bottom[:, None] + np.arange(3)).ravel()
    return bundle, load, DirichletDofSet(dofs, np.zeros(dofs.size)), c0


def macro_surrogate_section(args, ctx) -> dict:
    """Surrogate-driven explicit impact step against the plain FE² step (config 5).

    The macro slab of the plain-J2 line (128 x 128 x 64 Hex8, h = 10 mm, 50 MPa
    half-sine over 0.2 ms on the centred 20 x 20-element top patch, bottom fixed)
    with the trained surrogate at every Gauss point, ``--macro-surrogate-steps`` steps,
    and the gated FE² fallback on ``--fallback-rve``^3 two-phase RVEs (FP64, batched
    device solves): ``--macro-monitors`` monitor points under the patch carry their own
    fallback RVEs and trip the error gate when their surrogate stress leaves the FE²
    stress by more than 5 % of the stress scale.  The time step is the CFL-0.5 step of
    the stiffest phase's P-wave speed.  Reported:
    * surrogate macro steps/s (host clock between material calls, after warm-up);
    * the plain FE² step: end-to-end FE² macro steps of the 8 x 8 x 4 sample slab
      (2048 fallback-RVE solves per step) measured, and its Gauss-point rate scaled to
      the full slab's Gauss points;
    * the stress error at probe Gauss points under the patch against the FP64 FE²
      oracle on the training RVE, driven along each probe's strain history;
    * a gate-trip leg: the sample slab with every point monitored and a 1 % error
      bound, so that points trip and continue on their FE² RVEs."""
    import torch

    from concrete_impact_ai_amd import dynamics as D
    from concrete_impact_ai_amd import fe2
    from concrete_impact_ai_amd import materials as M
    from concrete_impact_ai_amd import surrogate as S
    from concrete_impact_ai_amd import surrogate_data as SD
    from concrete_impact_ai_amd.solvers import TimeIntegrationSettings

    model, params, scales = ctx["model"], ctx["params"], ctx["scales"]
    settings = M.MaterialUpdateSettings(50, 1e-10, 1e-12, 1e-10)
    nx, ny, nz = args.macro_divisions
    h = 0.01
    fb_model = SD.two_phase_rve(args.fallback_rve, seed=args.seed + 1)
    batched = fe2.RVEExecutionSettings("batched")
    fb = fe2.RVEHomogenizedMaterial(fb_model, "fallback", batched)
    cp = fb.maximum_wave_speed  # the stiffest phase's P-wave speed, as the FE² material states it
    dt = 0.5 * h / cp
    n_gp = 8 * nx * ny * nz
    rng = np.random.default_rng(args.seed)
    c0 = ((nx - 20) // 2, (ny - 20) // 2)

    def under_patch(count):
        pi = rng.integers(c0[0] - 4, c0[0] + 24, count)
        pj = rng.integers(c0[1] - 4, c0[1] + 24, count)
        pk = rng.integers(nz // 2, nz, count)
        return 8 * (pi + nx * (pj + ny * pk)) + rng.integers(0, 8, count)

    P = args.macro_probes
    probes = under_patch(P)
    monitors = under_patch(args.macro_monitors)

    def surrogate_material(probe_points=None, monitor_points=None, error_bound=0.05):
        return S.SurrogateMaterial("concrete_surrogate", params, scales, 2400.0, cp, fallback=fb,
                                   probe_points=probe_points, monitor_points=monitor_points,
                                   check_interval=20, error_bound=error_bound)

    try:
        sm = surrogate_material(probes, monitors)
        bundle, load, dirichlet, _ = _impact_slab((nx, ny, nz), sm, 20, h)
        n = bundle.mesh_info.dof_map.size
        steps = args.macro_surrogate_steps
        print(f"macro surrogate: {nx}x{ny}x{nz} slab, {steps} steps", file=sys.stderr, flush=True)
        ts = TimeIntegrationSettings("explicit", dt, steps, cfl_safety_factor=0.5)
        sol = D.solve_material_explicit(bundle, ts, np.zeros(n), np.zeros(n), dirichlet, load,
                                        "three_dimensional", settings, record_fields=False)
        torch.cuda.synchronize()
        w = min(20, steps // 4)
        ct = sm.call_times
        per_step = (ct[-1] - ct[w]) / (len(ct) - 1 - w)
        strain, stress, dts = sm.probe_history()      # (T, P, 6): calls 0..steps
        stats = dict(sm.stats)
        # FP64 FE² oracle on the training RVE along the probe strain histories
        times = np.concatenate([[0.0], np.cumsum(dts[1:])])
        hist = np.concatenate([np.zeros((1, P, 6)), strain[1:]], axis=0).transpose(1, 0, 2)
        print(f"macro surrogate: FP64 FE² oracle on {P} probe paths", file=sys.stderr, flush=True)
        _, _, ref, oracle_rate = SD.rve_responses_batched(model, np.repeat(times[None], P, 0), hist)
        sur = stress[1:].transpose(1, 0, 2)
        peak = np.max(np.abs(ref), axis=(1, 2))
        keep = peak >= 0.05 * np.max(peak)
        rms = SD.stress_rms_errors(sur[keep], ref[keep])
        mx = SD.stress_errors(sur[keep], ref[keep])
        # the same error over the first K steps of the run (the network's drift grows with
        # the number of its steps per point beyond the 200-step training paths)
        horizon = {}
        for K in (300, 1000):
            if K < steps:
                e_k = SD.stress_rms_errors(sur[keep][:, :K], ref[keep][:, :K])
                horizon[str(K)] = {"rms_median": float(np.median(e_k)),
                                   "rms_p95": float(np.percentile(e_k, 95))}
        print(f"macro surrogate: {1e3 * per_step:.2f} ms/step, {stats['gate_trips']} gate trips, "
              f"stress rms error median {np.median(rms):.4f} p95 {np.percentile(rms, 95):.4f}, "
              f"max error median {np.median(mx):.4f}", file=sys.stderr, flush=True)
        # end-to-end FE² macro steps on the sample slab, and the surrogate on the same slab
        small = tuple(args.macro_small)
        small_steps = 6
        n_small = 8 * small[0] * small[1] * small[2]

        def run_small(material, k_steps):
            b, ld, dr, _ = _impact_slab(small, material, max(2, small[0] // 2), h)
            nn = b.mesh_info.dof_map.size
            t = TimeIntegrationSettings("explicit", dt, k_steps, cfl_safety_factor=0.5)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            D.solve_material_explicit(b, t, np.zeros(nn), np.zeros(nn), dr, ld, "three_dimensional",
                                      settings, record_fields=False)
            torch.cuda.synchronize()
            return time.perf_counter() - t0

        print(f"macro surrogate: end-to-end FE² on a {small} slab", file=sys.stderr, flush=True)
        fe2_small = fe2.RVEHomogenizedMaterial(fb_model, "plain_fe2", batched)
        try:
            run_small(fe2_small, 2)                     # batch set-up and autotuning
            ta = run_small(fe2_small, 2)
            tb = run_small(fe2_small, 2 + small_steps)
        finally:
            fe2_small.close()
        fe2_small_step = (tb - ta) / small_steps
        sm_small = surrogate_material()
        run_small(sm_small, 2)
        ta = run_small(sm_small, 2)
        tb = run_small(sm_small, 2 + small_steps)
        sur_small_step = (tb - ta) / small_steps
        # gate-trip leg: every point of the sample slab monitored, 1 % error bound
        print("macro surrogate: gate-trip leg", file=sys.stderr, flush=True)
        sm_trip = surrogate_material(None, np.arange(n_small), error_bound=0.01)
        trip_s = run_small(sm_trip, 40)
        trip_stats = dict(sm_trip.stats)
    finally:
        fb.close()
    fe2_pts = n_small / fe2_small_step
    fe2_full_step = n_gp / fe2_pts
    return {
        "workload": f"explicit velocity Verlet, {nx}x{ny}x{nz} Hex8 slab, surrogate at every Gauss point "
                    f"(trained on {args.surrogate_rve}^3 FE² data) with FE²-judged error gate and "
                    f"fallback on {args.fallback_rve}^3 two-phase RVEs (FP64), 50 MPa / 0.2 ms "
                    f"half-sine on a 20x20 top patch, dt {dt:.4e} s (CFL 0.5)",
        "gauss_points": n_gp,
        "steps": steps,
        "surrogate_macro_steps_per_s": 1.0 / per_step,
        "surrogate_ms_per_step": per_step * 1e3,
        "surrogate_gp_updates_per_s": n_gp / per_step,
        "monitor_points": int(args.macro_monitors),
        "error_checks": stats["error_checks"],
        "gate_trips": stats["gate_trips"],
        "error_trips": stats["error_trips"],
        "judged_error_max": stats["judged_error_max"],
        "fallback_point_steps": stats["fallback_points"],
        "plain_fe2_sample": {"divisions": list(small), "gauss_points": n_small,
                             "steps_timed": small_steps,
                             "ms_per_step": fe2_small_step * 1e3,
                             "rve_solves_per_s": fe2_pts},
        "plain_fe2_ms_per_step": fe2_full_step * 1e3,
        "speedup_vs_plain_fe2_step": fe2_full_step / per_step,
        "fe2_step_note": "plain FE² step of this slab = its Gauss points at the RVE-solve rate measured "
                         "end to end on the sample slab (one fallback RVE per Gauss point, batched); the "
                         "full-slab FE² step (8.4 M RVE solves) is not run",
        "small_slab": {"divisions": list(small), "steps_timed": small_steps,
                       "plain_fe2_ms_per_step": fe2_small_step * 1e3,
                       "surrogate_ms_per_step": sur_small_step * 1e3,
                       "speedup": fe2_small_step / sur_small_step},
        "gate_trip_leg": {"divisions": list(small), "steps": 40, "monitor_points": n_small,
                          "error_bound": 0.01, "seconds": trip_s,
                          "error_checks": trip_stats["error_checks"],
                          "gate_trips": trip_stats["gate_trips"],
                          "error_trips": trip_stats["error_trips"],
                          "fallback_point_steps": trip_stats["fallback_points"]},
        "probes": P,
        "probes_scored": int(np.sum(keep)),
        "probe_peak_stress_max": float(np.max(peak)),
        "stress_rms_error_median": float(np.median(rms)),
        "stress_rms_error_p95": float(np.percentile(rms, 95)),
        "stress_max_error_median": float(np.median(mx)),
        "stress_max_error_p95": float(np.percentile(mx, 95)),
        "stress_rms_error_first_steps": horizon,
        "oracle_rve_solves_per_s": oracle_rate,
        "error_definition": "per probe Gauss point under the patch (peak stress >= 5 % of the largest): "
                            "rms / max over the run of |sigma_sur - sigma_FE2|_inf, divided by the probe's "
                            f"peak |sigma_FE2|_inf; FE2 = FP64 {args.surrogate_rve}^3 training RVE driven "
                            "along the probe's strain history",
        "final_dissipated_energy": float(sol.dissipated_energy[-1]),
        "final_free_energy": float(sol.free_energy[-1]),
        "meets_50x": bool(fe2_full_step / per_step >= 50.0),
        "meets_2pct_rms_median": bool(np.median(rms) <= 0.02),
    }


def batched_models(n_rves: int, seed: int, precision: str, newton_rtol: float):
    """The batched FE² workload: RVEs with edges drawn uniformly from 16-32
    elements, each with its own seeded spherical-aggregate microstructure (40 %
    aggregates, radii U[2, 6] elements) and its own seeded random 6-component
    strain path (100 steps, |eps| <= 3e-3, rates 10-1000 1/s)."""
    from concrete_impact_ai_amd import materials as M
    from concrete_impact_ai_amd import rve as R
    from concrete_impact_ai_amd.microstructure import random_strain_path, spherical_aggregates
    from concrete_impact_ai_amd.solvers import (ArmijoSettings, LinearSolverSettings,
                                                NonlinearNewtonSettings)

    rng = np.random.default_rng(seed)
    edges = rng.integers(16, 33, size=n_rves)
    mortar = M.J2ViscoplasticMaterial("mortar", 2400.0, 25.0e9, 0.2, 20.0e6, 0.5e9, 1.0e-3, 20.0e6, 1.0)
    agg = M.LinearElasticMaterial("aggregate", 2400.0, 50.0e9, 0.2)
    newton = NonlinearNewtonSettings(25, newton_rtol * 20.0e6, newton_rtol, 1.0e-12, newton_rtol,
                                     ArmijoSettings(True, 1.0e-4, 0.5, 20))
    lin = LinearSolverSettings(method="spectral", max_iterations=5000, relative_tolerance=1e-10,
                               check_interval=8, adaptive_forcing=True)
    models, times, strains = [], [], []
    for i, n in enumerate(int(v) for v in edges):
        eph = spherical_aggregates((n, n, n), 0.40, (2.0, 6.0), seed=seed + 1 + i)
        nodes, elements = R.build_structured_hex8_box((1.0, 1.0, 1.0), (n, n, n))
        models.append(R.StructuredHex8RVE(
            nodes, elements, (n, n, n),
            M.PhasedMaterialModel("rve", (M.RVEPhase("mortar", mortar), M.RVEPhase("aggregate", agg)),
                                  np.repeat(eph, 8), 2400.0),
            R.build_structured_hex8_periodic_constraint(nodes, (n, n, n)), newton,
            element_phase_ids=eph, phase_names=("mortar", "aggregate"), linear_settings=lin,
            precision=precision))
        t, e = random_strain_path(100, 6, 3.0e-3, (10.0, 1000.0), seed=seed + 1000 + i)
        times.append(t)
        strains.append(e)
    return models, np.asarray(times), np.asarray(strains), edges


def assemble_roofline(models, tuned) -> dict:
    """Roofline of the batched Newton assembly (k_b_assemble, FP32 point data) from
    the autotuner's HIP-event time of its all-active launch (every RVE forced active,
    at the first increment).  Algorithmic bytes: per Gauss point its stress (6) and
    tangent coefficients d, r (2) written; per mortar (J2) point its committed state
    (8) read and its next state (8) written; per element its corner forces (24)
    written, its corner fluctuation (3 FP64, one node per element on the periodic
    grid) read and its phase byte.  The plastic-flow rows of the tangent (6 more per
    plastic point) are not counted: the first increment is elastic."""
    n_el = sum(int(np.asarray(m.element_phase_ids).size) for m in models)
    n_j2 = sum(int(np.sum(np.asarray(m.element_phase_ids) == 0)) for m in models)
    alg = 8 * n_el * 8 * 4 + 8 * n_j2 * 16 * 4 + n_el * (24 * 4 + 3 * 8 + 1)
    us = float(tuned["assemble"]["us"])
    gbs = alg / (us * 1e-6) / 1e9
    return {"kernel": "k_b_assemble<float> (all RVEs active, autotuned grid)", "bound": "hbm",
            "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
            "alg_bytes_per_launch": alg, "us": us, "mortar_element_fraction": n_j2 / n_el}


def batched_section(args) -> dict:
    """Batched FE²: 256 uneven RVEs advanced together (FP32 element work, FP64
    residuals / dots / multiplier solve, Newton relative tolerance 1e-6; once from
    the committed fluctuation and once from the extrapolated one) next to the same
    batch in FP64 (Newton 1e-10) and the plain FP64 single-RVE solver on a sample
    of the RVEs.  Rates are RVE increments (solves) per second over the
    timed increments; the FP32 error is the homogenised-stress difference to the
    FP64 batch on the same increments."""
    import dataclasses

    import torch

    from concrete_impact_ai_amd import batch as B
    from concrete_impact_ai_amd import materials as M
    from concrete_impact_ai_amd import rve as R
    from concrete_impact_ai_amd.materials import MaterialResponseRequirements

    nr, W, K = args.batch_rves, args.batch_warmup, args.batch_steps
    models32, times, strains, edges = batched_models(nr, 3000 + args.seed, "fp32", 1e-6)
    newton64 = dataclasses.replace(models32[0].newton_settings,
                                   residual_absolute_tolerance=1e-10 * 20.0e6,
                                   residual_relative_tolerance=1e-10,
                                   increment_relative_tolerance=1e-10)
    models64 = [dataclasses.replace(m, precision="fp64", newton_settings=newton64) for m in models32]
    settings = M.MaterialUpdateSettings(50, 1.0e-10, 1.0e-12, 1.0e-10)
    n_el = int(np.sum(edges.astype(np.int64) ** 3))
    out = {"rves": nr, "edges_min": int(edges.min()), "edges_max": int(edges.max()),
           "elements": n_el, "gauss_points": 8 * n_el, "steps_warmup": W, "steps_timed": K}
    stresses = {}
    for tag, models, use_pred in (("fp32", models32, False), ("fp32_predictor", models32, True),
                                  ("fp64", models64, False)):
        print(f"batched {tag}: building {nr} RVEs", file=sys.stderr, flush=True)
        batch = B.RVEBatch(models)
        st = batch.initial_state()
        pred = B.FluctuationPredictor(batch) if use_pred else None

        def solve(k, st):
            w0 = pred.initial_w(st, strains[:, k]) if pred is not None else None
            return B.solve_rve_batch(batch, strains[:, k], times[:, k] - times[:, k - 1], settings,
                                     st, initial_w=w0)

        for k in range(1, W + 1):
            st = solve(k, st).state
            print(f"batched {tag}: warm-up increment {k} done", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        cnt0 = batch.stats()
        t0 = time.perf_counter()
        gp = newton = lin = 0
        sig = []
        for k in range(W + 1, W + K + 1):
            r = solve(k, st)
            st = r.state
            sig.append(r.macro_stress)
            gp += int(np.sum(r.diagnostics["gauss_point_updates"]))
            newton += int(np.sum(r.diagnostics["newton_iterations"]))
            lin += int(np.max(r.diagnostics["linear_iterations"]))
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        print(f"batched {tag}: {K} increments in {el:.3f} s", file=sys.stderr, flush=True)
        stresses[tag] = np.asarray(sig)
        out[tag] = {"rve_solves_per_s": nr * K / el, "ms_per_batch_increment": 1e3 * el / K,
                    "gp_updates_per_s": gp / el, "newton_iterations_per_rve_step": newton / (nr * K),
                    "max_pcg_iterations_per_step": lin / K,
                    "active_fraction_last": float(np.mean(
                        r.diagnostics["viscoplastic_active_volume_fraction"]))}
        cnt1 = batch.stats()
        out[tag]["counters_timed"] = {k: cnt1[k] - cnt0[k] for k in cnt1}
        if tag == "fp32":
            out["autotuned_grids"] = batch.tuning()
            out["assemble_roofline"] = assemble_roofline(models32, out["autotuned_grids"])
        del batch, st
        torch.cuda.empty_cache()
    den = np.maximum(np.max(np.abs(stresses["fp64"]), axis=2), 1e-300)
    for tag in ("fp32", "fp32_predictor"):
        err = np.max(np.abs(stresses[tag] - stresses["fp64"]), axis=2) / den
        out[f"{tag}_stress_error_max"] = float(np.max(err))
        out[f"{tag}_stress_error_median"] = float(np.median(err))
    # plain FP64 single-RVE solver, one RVE at a time: every RVE of the batch (the
    # default), or a sample of --batch-sample RVEs scaled to the batch's mean size
    if args.batch_sample <= 0 or args.batch_sample >= nr:
        sample = list(range(nr))
    else:
        sample = list(range(0, nr, max(1, nr // args.batch_sample)))[: args.batch_sample]
    t_single = 0.0
    n_single = 0
    req = MaterialResponseRequirements(tangent=False)
    for i in sample:
        cache = R.build_rve_operator_cache(models64[i])
        st = R.initialize_rve_state(models64[i], cache)
        for k in range(1, W + K + 1):
            if k == W + 1:
                torch.cuda.synchronize()
                t0 = time.perf_counter()
            st = R.solve_rve_microequilibrium(
                models64[i], R.RVERequest(strains[i, k], float(times[i, k] - times[i, k - 1]), settings,
                                          req), st, cache, retain_micro_fields=False).state
        torch.cuda.synchronize()
        t_single += time.perf_counter() - t0
        n_single += K
        cache.close()
    if len(sample) == nr:
        per_solve = t_single / n_single
        note = ("every RVE of the batch, one at a time (the reference's serial loop) on the GPU, "
                "over the same timed increments")
    else:
        # the sample's mean time per solve, scaled to the batch's mean RVE size
        sample_el = float(np.mean(edges[sample].astype(np.float64) ** 3))
        per_solve = t_single / n_single * (n_el / nr) / sample_el
        note = ("a sample of the batch, one RVE at a time (the reference's serial loop) on the GPU; "
                "time per solve scaled by the batch's mean RVE size")
    out["fp64_single_rve"] = {"rve_solves_per_s": 1.0 / per_solve, "sample_rves": len(sample),
                              "note": note}
    out["speedup_fp32_batched_vs_fp64_single"] = out["fp32"]["rve_solves_per_s"] * per_solve
    out["speedup_fp32_predictor_batched_vs_fp64_single"] = (out["fp32_predictor"]["rve_solves_per_s"]
                                                            * per_solve)
    out["speedup_fp32_batched_vs_fp64_batched"] = (out["fp32"]["rve_solves_per_s"] /
                                                   out["fp64"]["rve_solves_per_s"])
    return out


def macro_explicit_section(args) -> dict:
    """Explicit velocity-Verlet macro slab, plain J2 (the 128 x 128 x 64 Hex8 slab, h =
    10 mm, rho 2400, E 30 GPa, nu 0.2, sigma_y 30 MPa, H 1 GPa, tau 1e-3 s, m 1; 50 MPa
    half-sine over 0.2 ms on a centred 20 x 20-element top patch, bottom fixed, CFL 0.5).
    Steps/s from the difference of two runs (warm-up steps + timed steps), so set-up and
    the initial assembly cancel."""
    import torch

    from concrete_impact_ai_amd import dynamics as D
    from concrete_impact_ai_amd import materials as M
    from concrete_impact_ai_amd.solvers import DirichletDofSet, TimeIntegrationSettings

    nx, ny, nz = args.macro_divisions
    h = 0.01
    lengths = (nx * h, ny * h, nz * h)
    j2 = M.J2ViscoplasticMaterial("slab", 2400.0, 30e9, 0.2, 30e6, 1e9, 1e-3, 30e6, 1.0)
    bundle = D.build_structured_box_bundle(lengths, (nx, ny, nz), j2)
    c0 = (nx - 20) // 2, (ny - 20) // 2
    load = D.build_half_sine_pressure_load(bundle, 50e6, 2e-4, patch=((c0[0], c0[0] + 20),
                                                                     (c0[1], c0[1] + 20)))
    bottom = bundle.mesh_info.boundary_groups["bottom"].nodes
    dofs = (3 * bottom[:, None] + np.arange(3)).ravel()
    cp = float(j2.maximum_wave_speed)
    dt = 0.5 * h / cp
    n = bundle.mesh_info.dof_map.size
    settings = M.MaterialUpdateSettings(50, 1e-10, 1e-12, 1e-10)

    def run(steps):
        ts = TimeIntegrationSettings("explicit", dt, steps, cfl_safety_factor=0.5)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sol = D.solve_material_explicit(bundle, ts, np.zeros(n), np.zeros(n),
                                        DirichletDofSet(dofs, np.zeros(dofs.size)), load,
                                        "three_dimensional", settings, record_fields=False)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, sol

    w_steps, t_steps = 20, args.macro_steps
    run(w_steps)
    ta, _ = run(w_steps)
    tb, sol = run(w_steps + t_steps)
    per_step = (tb - ta) / t_steps
    n_el = nx * ny * nz
    n_gp = 8 * n_el
    alg = n_gp * 128 + n * 64  # state 8 values read + written per point; u, v, a, m, f_ext per dof
    return {
        "workload": f"explicit velocity Verlet, {nx}x{ny}x{nz} Hex8 slab, plain J2, FP64, "
                    f"dt {dt:.4e} s (CFL 0.5)",
        "gauss_points": n_gp,
        "dofs": n,
        "steps_timed": t_steps,
        "macro_steps_per_s": 1.0 / per_step,
        "ms_per_step": per_step * 1e3,
        "gp_updates_per_s": n_gp / per_step,
        "roofline": {"bound": "hbm", "achieved": alg / per_step / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": alg / per_step / 1e9 / HBM_PEAK_GBS,
                     "alg_bytes_per_step": alg},
        "final_dissipated_energy": float(sol.dissipated_energy[-1]),
        "final_active_fraction": float(sol.viscoplastic_active_volume_fraction[-1]),
    }


FP32_PEAK_TFLOPS = 157.3
RNO_FLOP_PER_POINT = 708_864  # algorithmic: 3 energy-gradient chains + 1 evolution forward


def rve_rno_section(args) -> dict:
    """The reference's trained RVE-RNO (artifacts/rve_rno/c48_v2_liu_direct, committed
    as tests/golden/rve_rno_c48_v2_liu_direct.npz) through cia_rno_update, FP32:

    * frozen-test evaluation shape of the reference (training_summary.json: one padded
      batch, 38 400 material slots = 200 paths x 192 steps, 24 064 valid steps,
      159 156 valid steps/s on an RTX 5060 Ti): 200 synthetic load / unload paths
      rolled out for 192 sequential steps, slots/s and valid steps/s at the same
      62.67 % padding efficiency;
    * batch throughput: one update of ``--rno-batch`` points (HIP events around K
      updates), without and with the tangent; roofline on the FP32 matrix-core peak
      with 708 864 algorithmic flop per point (the two energies' value-and-gradient
      chains at 6 and twice at 12 inputs, the evolution forward; 8 x that with the
      tangent's six directions and one padding column)."""
    import torch

    from concrete_impact_ai_amd import rve_rno as RR

    model = RR.load_rve_rno_artifact(ROOT / "tests" / "golden" / "rve_rno_c48_v2_liu_direct.npz")
    dev = model.device
    meta_scale = torch.tensor([0.0107, 0.0049, 0.0059, 0.0082, 0.0024, 0.0014],
                              dtype=torch.float32, device=dev)  # training strain rms
    gen = torch.Generator(device=dev).manual_seed(5)

    # frozen-test shape: 200 paths x 192 steps, sequential
    paths, steps = 200, 192
    amp = torch.randn((paths, 6), generator=gen, device=dev) * meta_scale * 2.0
    t = torch.arange(1, steps + 1, device=dev, dtype=torch.float32)
    shape = torch.where(t <= steps // 2, t / (steps // 2), (steps - t) / (steps // 2))
    strain = shape[:, None, None] * amp[None]
    dt = torch.full((paths,), 1.0e-4, device=dev)

    def rollout():
        z = model.initial_state_batch(paths)
        for k in range(steps):
            z = model.update_batch(strain[k], z, dt).latent_state
        return z

    rollout()
    torch.cuda.synchronize()
    reps = []
    for _ in range(5):
        t0 = time.perf_counter()
        rollout()
        torch.cuda.synchronize()
        reps.append(time.perf_counter() - t0)
    el = float(np.median(reps))
    slots = paths * steps
    frozen = {"paths": paths, "steps": steps, "seconds_median_of_5": el,
              "slots_per_s": slots / el, "valid_steps_per_s_at_reference_padding": 0.6267 * slots / el,
              "reference_valid_steps_per_s": 159156.0,
              "reference_hardware": "NVIDIA GeForce RTX 5060 Ti, torch inductor, FP32"}

    def batch_rate(n, tangent, reps=10):
        s = torch.randn((n, 6), generator=gen, device=dev) * meta_scale * 2.0
        z = torch.randn((n, 6), generator=gen, device=dev) * 0.5
        h = torch.full((n,), 1.0e-4, device=dev)
        model.update_batch(s, z, h, compute_tangent=tangent)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            model.update_batch(s, z, h, compute_tangent=tangent)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        flop = RNO_FLOP_PER_POINT * n * (8 if tangent else 1)
        return {"points": n, "ms_per_update": ms, "point_updates_per_s": n / ms * 1e3,
                "tflops": flop / ms / 1e9, "frac_fp32_mfma_peak": flop / ms / 1e9 / FP32_PEAK_TFLOPS}

    n = args.rno_batch
    plain = batch_rate(n, False)
    tangent = batch_rate(max(n // 8, 64), True)
    return {"workload": "reference RVE-RNO artifact c48_v2_liu_direct (6 -> 128^4 SiLU energies, "
                        "direct-rate evolution, latent 6), FP32",
            "frozen_test_shape": frozen, "batch": plain, "batch_with_tangent": tangent,
            "roofline": {"bound": "mfma", "achieved": plain["tflops"], "peak": FP32_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": plain["frac_fp32_mfma_peak"],
                         "flop_per_point": RNO_FLOP_PER_POINT}}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=48)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--pcg-rtol", type=float, default=1.0e-10)
    ap.add_argument("--fixed-forcing", action="store_true",
                    help="solve every Newton step to --pcg-rtol instead of the inexact-Newton forcing")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32 side-by-side run")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a 256 MiB device copy (HBM-counter calibration kernel)")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the surrogate section")
    ap.add_argument("--no-macro", action="store_true", help="skip the macro explicit slab section")
    ap.add_argument("--macro-only", action="store_true")
    ap.add_argument("--macro-divisions", type=int, nargs=3, default=(128, 128, 64))
    ap.add_argument("--macro-steps", type=int, default=200)
    ap.add_argument("--surrogate-only", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched FE² section")
    ap.add_argument("--batched-only", action="store_true")
    ap.add_argument("--no-rno", action="store_true", help="skip the reference RVE-RNO section")
    ap.add_argument("--rno-only", action="store_true")
    ap.add_argument("--rno-batch", type=int, default=1 << 20)
    ap.add_argument("--batch-rves", type=int, default=256)
    ap.add_argument("--batch-warmup", type=int, default=2)
    ap.add_argument("--batch-steps", type=int, default=8)
    ap.add_argument("--batch-sample", type=int, default=0,
                    help="RVEs of the single-RVE comparison (0: all)")
    ap.add_argument("--surrogate-rve", type=int, default=24)
    ap.add_argument("--surrogate-steps", type=int, default=200)
    ap.add_argument("--surrogate-paths", type=int, default=2000)
    ap.add_argument("--surrogate-test-paths", type=int, default=400)
    ap.add_argument("--surrogate-iters", type=int, default=3000,
                    help="BPTT iterations on 4096 x 64-step windows")
    ap.add_argument("--surrogate-full-iters", type=int, default=6000,
                    help="BPTT iterations on whole paths (every training path)")
    ap.add_argument("--fallback-rve", type=int, default=16,
                    help="edge of the FE² fallback / monitor RVE of the macro impact run")
    ap.add_argument("--surrogate-production-fraction", type=float, default=0.5,
                    help="share of training paths from the production path design")
    ap.add_argument("--no-macro-surrogate", action="store_true",
                    help="skip the surrogate-driven macro impact run")
    ap.add_argument("--macro-surrogate-steps", type=int, default=2000)
    ap.add_argument("--macro-probes", type=int, default=64)
    ap.add_argument("--macro-monitors", type=int, default=64)
    ap.add_argument("--macro-small", type=int, nargs=3, default=(8, 8, 4))
    ap.add_argument("--no-plan", action="store_true", help="skip the data-plan production section")
    ap.add_argument("--plan-rve", type=int, default=16)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding of this path is design-only).", file=sys.stderr)
        return 2

    import torch

    from concrete_impact_ai_amd import _lib
    from concrete_impact_ai_amd import materials as M
    from concrete_impact_ai_amd import rve as R
    from concrete_impact_ai_amd.materials import MaterialResponseRequirements
    from concrete_impact_ai_amd.microstructure import random_strain_path

    torch.cuda.set_device(0)
    lib = _lib.lib()
    if args.copy_calibration:
        copy_calibration(torch)
    if args.surrogate_only:
        sec, ctx = surrogate_section(args)
        if not args.no_macro_surrogate:
            sec["macro_impact"] = macro_surrogate_section(args, ctx)
        print(json.dumps({"surrogate": sec}))
        return 0
    if args.macro_only:
        print(json.dumps({"macro_explicit": macro_explicit_section(args)}))
        return 0
    if args.batched_only:
        print(json.dumps({"batched": batched_section(args)}))
        return 0
    if args.rno_only:
        print(json.dumps({"rve_rno": rve_rno_section(args)}))
        return 0
    n = args.size
    model, eph = build_model(n, args.seed, args.precision, args.pcg_rtol, not args.fixed_forcing)
    times, strains = random_strain_path(200, 8, 3.0e-3, (10.0, 1000.0), seed=args.seed)
    settings = M.MaterialUpdateSettings(50, 1.0e-10, 1.0e-12, 1.0e-10)
    req_no_tan = MaterialResponseRequirements(tangent=False)
    cache = R.build_rve_operator_cache(model)
    state = R.initialize_rve_state(model, cache)
    total = args.warmup + args.steps
    if total > 200:
        raise SystemExit("the path has 200 increments")

    def step(k, st):
        return R.solve_rve_microequilibrium(
            model, R.RVERequest(strains[k], float(times[k] - times[k - 1]), settings, req_no_tan),
            st, cache, retain_micro_fields=False)

    for k in range(1, args.warmup + 1):
        state = step(k, state).state
        print(f"warmup step {k} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    start_state = state  # committed states are never modified by a solve
    newton = backtracks = lin_its = 0
    gp_updates = 0
    sig64 = []
    t0 = time.perf_counter()
    for k in range(args.warmup + 1, total + 1):
        resp = step(k, state)
        state = resp.state
        sig64.append(resp.macro_stress)
        d = resp.diagnostics
        newton += d["newton_iterations"]
        backtracks += d["armijo_backtracks"]
        lin_its += d["linear_iterations"]
        gp_updates += d["gauss_point_updates"]
        print(f"step {k}: newton {d['newton_iterations']} pcg {d['linear_iterations']} "
              f"active {d['viscoplastic_active_volume_fraction']:.3f}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    # Kernel averages come from a replay of the same increments with HIP-event timing
    # on (the events add host API calls, so the timed pass above runs without them).
    lib.cia_profile_enable(1)
    tot_ms = (C.c_double * 6)()
    cnt = (C.c_int64 * 6)()
    lib.cia_profile_read(tot_ms, cnt, 6)  # clear
    state = start_state
    for k in range(args.warmup + 1, total + 1):
        state = step(k, state).state
    torch.cuda.synchronize()
    lib.cia_profile_read(tot_ms, cnt, 6)
    lib.cia_profile_enable(0)
    print("profiled replay done", file=sys.stderr, flush=True)

    K = args.steps
    n_el = n**3
    n_gp = 8 * n_el
    n_dof = 3 * n_el
    plastic_frac = float(d["viscoplastic_active_volume_fraction"])
    bpe = 8 if args.precision == "fp64" else 4
    # Algorithmic bytes per launch (unique HBM bytes the operation must move):
    #  pcg_elem: compact tangent (d, r per point; flow direction n[6] on plastic
    #            points), p_old + z read and p_new written once per class, the
    #            24 element force rows written.
    #  pcg_update: p, y, x, r read and x, r written once per class, then the
    #            x-transformed spectra of the three residual components written.
    #  assemble: a Newton pass (all but one launch per increment): committed state
    #            read (J2 points), compact tangent written, w read once per class,
    #            force rows written (the final pass also writes stresses, next state
    #            and response rows).
    #  precond: one application of the spectral preconditioner after the x-DFT
    #            (forward y pass, z pass with the symbol solve, inverse y pass,
    #            inverse x pass): the three components' spectra (16 B each) read and
    #            written by each y pass (96 B per class), by the z-column pass
    #            (96 B per mode) together with the FP32 symbol (48 B per mode),
    #            then read once more by the inverse x pass, which writes the real z
    #            (48 + 24 B per class): 408 B per class.
    #  gather:  the element force rows read, the residual written per class.
    mortar_gp = float(np.mean(eph == 0)) * n_gp
    alg = {
        "pcg_elem": n_gp * bpe * (2 + 6 * plastic_frac) + n_dof * 8 * 3 + n_el * 24 * bpe,
        "pcg_node": n_el * 24 * bpe + n_dof * 8 * 2,
        "pcg_update": n_dof * 8 * 6 + n_el * 48,
        "assemble": mortar_gp * bpe * 8 + n_gp * bpe * (2 + 6 * plastic_frac) + n_dof * 8
                    + n_el * 24 * bpe,
        "precond": n_el * 408,
        "gather": n_el * 24 * bpe + n_dof * 8,
    }
    kernels = {}
    for i, name in enumerate(_lib.PROFILE_SLOTS):
        if cnt[i] == 0:
            continue
        avg_s = tot_ms[i] / cnt[i] / 1e3
        entry = {"launches": int(cnt[i]), "avg_us": avg_s * 1e6, "total_ms": tot_ms[i]}
        if name in alg:
            entry["alg_bytes"] = alg[name]
            entry["achieved_GBs"] = alg[name] / avg_s / 1e9
        kernels[name] = entry
    # The roofline object is the slot with the largest total time in the replay of
    # the timed increments (every slot carries its algorithmic bytes).
    dom = max(kernels, key=lambda k: kernels[k]["total_ms"])
    ach = kernels[dom]["achieved_GBs"]
    traffic = MEASURED_TRAFFIC.get((dom, args.size, args.precision))
    roofline = {
        "kernel": (KERNEL_NAMES_ROW_LANE if n % 16 == 0 and n <= 64 else KERNEL_NAMES)[dom],
        "bound": "hbm",
        "achieved": ach,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": ach / HBM_PEAK_GBS,
        "traffic": traffic,
        "alg_bytes_per_launch": kernels[dom]["alg_bytes"],
    }
    value = K / elapsed
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "RVE solves/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 element work, f64 residual/dots",
        "data": "synthetic (seeded sphere microstructure and strain path)",
        "config": {
            "workload": f"periodic two-phase RVE {n}^3 Hex8, increments {args.warmup + 1}-{total} "
                        "of a 200-step random 6-component path, no effective tangent",
            "elements": n_el,
            "gauss_points": n_gp,
            "reduced_dofs": n_dof - 3,
            "aggregate_fraction": float(np.mean(eph == 1)),
            "seed": args.seed,
            "precision": args.precision,
            "pcg_rtol": args.pcg_rtol,
            "inexact_newton_forcing": not args.fixed_forcing,
        },
        "gp_updates_per_s": gp_updates / elapsed,
        "newton_iterations_per_step": newton / K,
        "armijo_backtracks_per_step": backtracks / K,
        "pcg_iterations_per_step": lin_its / K,
        "kernels": kernels,
        "roofline": roofline,
        "cpu_baseline": None,
    }
    if args.precision == "fp64" and not args.no_fp32:
        # the same increments with FP32 element work and FP64 residuals, side by side
        print("fp32 + refinement side by side ...", file=sys.stderr, flush=True)
        m32, _ = build_model(n, args.seed, "fp32", args.pcg_rtol, not args.fixed_forcing,
                             newton_rtol=1.0e-6)
        c32 = R.build_rve_operator_cache(m32)
        st32 = R.initialize_rve_state(m32, c32)

        def step32(k, st):
            return R.solve_rve_microequilibrium(
                m32, R.RVERequest(strains[k], float(times[k] - times[k - 1]), settings, req_no_tan),
                st, c32, retain_micro_fields=False)

        for k in range(1, args.warmup + 1):
            st32 = step32(k, st32).state
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        sig32, n32, l32 = [], 0, 0
        for k in range(args.warmup + 1, total + 1):
            r32 = step32(k, st32)
            st32 = r32.state
            sig32.append(r32.macro_stress)
            n32 += r32.diagnostics["newton_iterations"]
            l32 += r32.diagnostics["linear_iterations"]
        torch.cuda.synchronize()
        el32 = time.perf_counter() - t1
        a32, a64 = np.asarray(sig32), np.asarray(sig64)
        err = np.max(np.abs(a32 - a64), axis=1) / np.max(np.abs(a64), axis=1)
        line["fp32_refined"] = {
            "rve_solves_per_s": K / el32, "ms_per_step": 1e3 * el32 / K,
            "newton_tolerance": 1e-6, "newton_iterations_per_step": n32 / K,
            "pcg_iterations_per_step": l32 / K,
            "stress_error_vs_fp64_max": float(np.max(err)),
            "stress_error_vs_fp64_median": float(np.median(err)),
            "error_definition": "per increment |sigma_fp32 - sigma_fp64|_inf / |sigma_fp64|_inf",
        }
        c32.close()
    print("plain-method RVE (8^3, 100 steps) ...", file=sys.stderr, flush=True)
    line["c1_rve"] = c1_section()
    if not args.no_plan:
        print("data plan production ...", file=sys.stderr, flush=True)
        line["plan_production"] = plan_section(args)
    if not args.no_cpu_baseline:
        print("cpu baseline (oracle) ...", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(n, args.seed)
        line["cpu_baseline"]["other_paths"] = cpu_baseline_paths(args.seed)
    if not args.no_macro:
        print("macro explicit slab ...", file=sys.stderr, flush=True)
        line["macro_explicit"] = macro_explicit_section(args)
    if not args.no_batched:
        print("batched FE² ...", file=sys.stderr, flush=True)
        line["batched"] = batched_section(args)
    if not args.no_rno:
        print("reference RVE-RNO ...", file=sys.stderr, flush=True)
        line["rve_rno"] = rve_rno_section(args)
    if not args.no_surrogate:
        line["surrogate"], ctx = surrogate_section(args)
        if not args.no_macro_surrogate:
            line["surrogate"]["macro_impact"] = macro_surrogate_section(args, ctx)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
