"""FE² macro material adapter: one independent periodic RVE per macro integration
point behind the material-point protocol.

Reference: ``RVEHomogenizedMaterial`` (``src/fem/rve/material.py:41-191``), the
per-point state codec ``decode_rve_point_state`` / ``_encode_rve_point_state``
(``:194-217``), the macro diagnostics ``_macro_diagnostics`` (``:220-264``) and
``RVEExecutionSettings`` (``src/fem/rve/data.py:116-138``).

The reference runs the per-point RVE solves serially or in a process pool.  Here
every RVE's state stays on the device in one batch allocation (``[n][8][n_micro]``
material state, ``[n][3 C]`` fluctuation field) and each solve is the device
Newton/PCG of ``solve_rve_microequilibrium``.  The ``streams`` backend (and the
reference's ``process_pool``, which maps to it) runs ``workers`` solves at a time,
each on its own HIP stream with its own operator workspace, from host threads:
the ctypes call releases the GIL, so small RVEs overlap on the GPU.
"""

from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass

import ctypes as C
import numpy as np
import torch

from concrete_impact_ai_amd import _lib
from concrete_impact_ai_amd.materials import (
    DEFAULT_RESPONSE_REQUIREMENTS,
    DeviceResponse,
    J2ViscoplasticMaterial,
    MaterialPointResponse,
    MaterialResponseRequirements,
    MaterialState,
    _pack_state,
    device,
    stream_handle,
    unpack_state,
)
from concrete_impact_ai_amd.rve import (
    DeviceMaterialState,
    RVEOperatorCache,
    RVEState,
    StructuredHex8RVE,
    solve_increment_raw,
)

VOIGT_COMPONENTS = ("xx", "yy", "zz", "yz", "xz", "xy")


@dataclass(frozen=True)
class RVEExecutionSettings:
    """Serial, process-pool, device-stream or batched execution of the per-point RVE
    solves (``src/fem/rve/data.py:116-138``; ``streams`` and ``batched`` are this
    build's backends and ``process_pool`` is executed by ``streams``).  ``batched``
    advances every macro point's RVE in one batched device solve
    (``concrete_impact_ai_amd.batch``); updates that request the effective tangent
    add the batched Schur-complement solves of the converged RVEs."""

    backend: str = "serial"
    workers: int = 1
    threads_per_worker: int = 1

    def __post_init__(self) -> None:
        if self.backend not in ("serial", "process_pool", "streams", "batched"):
            raise ValueError(f"Unknown RVE execution backend {self.backend!r}.")
        if self.workers <= 0:
            raise ValueError("RVE execution workers must be positive.")
        if self.threads_per_worker <= 0:
            raise ValueError("RVE execution threads_per_worker must be positive.")
        if self.backend == "serial" and self.workers != 1:
            raise ValueError("Serial RVE execution requires exactly one worker.")
        if self.backend == "process_pool":
            logical_cpu_limit = max((os.cpu_count() or 1) // 2, 1)
            requested = self.workers * self.threads_per_worker
            if requested > logical_cpu_limit:
                raise ValueError(
                    "RVE process CPU budget exceeds half of available CPUs: "
                    f"workers={self.workers}, threads_per_worker={self.threads_per_worker}, "
                    f"requested={requested}, limit={logical_cpu_limit}.")
        if self.backend == "streams" and self.workers > 16:
            raise ValueError("At most 16 concurrent RVE streams are supported.")


class RVEBatchState(MaterialState):
    """Committed RVE states of all macro points, device-resident.

    ``variables`` (read lazily) carries the reference keys: ``rve_macro_strain``
    (n, 6), ``rve_fluctuation_dofs`` (n, 3(C-1)) and ``rve_material__<var>``
    (n, n_micro_points, ...)."""

    def __init__(self, packed: torch.Tensor, w: torch.Tensor, macro: np.ndarray, cache):
        object.__setattr__(self, "packed", packed)
        object.__setattr__(self, "w", w)
        object.__setattr__(self, "macro", np.asarray(macro, dtype=np.float64))
        object.__setattr__(self, "_cache", cache)
        object.__setattr__(self, "_vars", None)

    @property
    def variables(self) -> dict:
        if self._vars is None:
            c = self._cache
            n = self.macro.shape[0]
            out = {"rve_macro_strain": self.macro.copy(),
                   "rve_fluctuation_dofs": self.w.cpu().numpy()[:, 3:].copy()}
            packed = self.packed.double().cpu().numpy()
            ppid = c.point_phase_ids if c.phased else None
            per_point = [unpack_state(c.phases, ppid, packed[i], c.phased) for i in range(n)]
            names = per_point[0].keys() if n else unpack_state(
                c.phases, ppid, np.zeros((8, c.n_points)), c.phased).keys()
            for name in names:
                out[f"rve_material__{name}"] = (np.stack([p[name] for p in per_point]) if n else
                                                np.zeros((0,)))
            object.__setattr__(self, "_vars", out)
        return self._vars

    def __len__(self) -> int:
        return self.macro.shape[0]


def decode_rve_point_state(state: MaterialState, point_id: int) -> RVEState:
    """One macro point's RVE state from the batch (``material.py:194-205``)."""
    if isinstance(state, RVEBatchState):
        c = state._cache
        ms = DeviceMaterialState(state.packed[point_id].clone(), c.phases,
                                 c.point_phase_ids if c.phased else None, c.phased)
        return RVEState(ms, state.macro[point_id].copy(), None,
                        w_device=state.w[point_id].clone())
    material_variables = {name.removeprefix("rve_material__"): values[point_id].copy()
                          for name, values in state.variables.items()
                          if name.startswith("rve_material__")}
    return RVEState(MaterialState(material_variables),
                    state.variables["rve_macro_strain"][point_id].copy(),
                    state.variables["rve_fluctuation_dofs"][point_id].copy())


def _encode_rve_point_state(variables: dict, point_id: int, state: RVEState) -> None:
    """Write one accepted RVE state into host batch variables (``material.py:208-217``)."""
    variables["rve_macro_strain"][point_id] = state.macro_strain
    variables["rve_fluctuation_dofs"][point_id] = state.fluctuation_dofs
    for name, values in state.material_state.variables.items():
        variables[f"rve_material__{name}"][point_id] = values


class RVEHomogenizedMaterial:
    """Expose a fixed quasi-static RVE through the material-point protocol
    (``src/fem/rve/material.py:41-191``)."""

    def __init__(self, model: StructuredHex8RVE, name: str,
                 execution_settings: RVEExecutionSettings | None = None):
        self.model = model
        self.name = name
        self.execution_settings = execution_settings or RVEExecutionSettings()
        es = self.execution_settings
        self._batched = es.backend == "batched"
        self._batches = {}  # RVEs per batched solve -> batched device operator
        self._workers = 1 if es.backend in ("serial", "batched") else es.workers
        self.operator_cache = RVEOperatorCache(model)
        self._caches = [self.operator_cache] + [RVEOperatorCache(model)
                                                for _ in range(self._workers - 1)]
        self._streams = ([torch.cuda.Stream(device=device()) for _ in range(self._workers)]
                         if self._workers > 1 else None)
        self._executor = (ThreadPoolExecutor(max_workers=self._workers)
                          if self._workers > 1 else None)
        c = self.operator_cache
        self._dtype = c.dtype
        self._q_masks = None

    # -- material-point protocol ------------------------------------------------
    @property
    def density(self) -> float:
        return float(self.model.material.density)

    @property
    def maximum_wave_speed(self) -> float:
        return float(self.model.material.maximum_wave_speed)

    @property
    def young_modulus(self) -> float:
        return float(self.model.material.young_modulus)

    @property
    def poisson_ratio(self) -> float:
        return float(self.model.material.poisson_ratio)

    def initialize_state(self, n_points: int) -> RVEBatchState:
        """One independent committed RVE state per macro point (``material.py:95-116``)."""
        c = self.operator_cache
        dev = device()
        packed = torch.zeros((n_points, 8, c.n_points), dtype=self._dtype, device=dev)
        w = torch.zeros((n_points, 3 * c.n_classes), dtype=torch.float64, device=dev)
        return RVEBatchState(packed, w, np.zeros((n_points, 6)), c)

    def _device_batch(self, state: MaterialState, n: int) -> RVEBatchState:
        if isinstance(state, RVEBatchState):
            if len(state) != n:
                raise ValueError("RVE batch state size does not match the number of points.")
            return state
        c = self.operator_cache
        dev = device()
        ppid = c.point_phase_ids if c.phased else None
        packed = np.zeros((n, 8, c.n_points))
        for i in range(n):
            rs = decode_rve_point_state(state, i)
            packed[i] = _pack_state(c.phases, ppid, rs.material_state, c.n_points, c.phased)
        w = np.zeros((n, 3 * c.n_classes))
        w[:, 3:] = np.asarray(state.variables["rve_fluctuation_dofs"])
        return RVEBatchState(torch.as_tensor(packed, device=dev).to(self._dtype).contiguous(),
                             torch.as_tensor(w, device=dev),
                             np.asarray(state.variables["rve_macro_strain"]), c)

    # Gauss points of all RVEs of one batched solve: larger point sets are solved in
    # chunks, so that a fallback set of any size fits device memory
    BATCH_GAUSS_POINTS = 1 << 28

    def _batch_for(self, m: int):
        from concrete_impact_ai_amd import batch as B

        bt = self._batches.get(m)
        if bt is None:
            if len(self._batches) >= 2:  # keep the operators of the last two chunk sizes
                self._batches.pop(next(iter(self._batches))).close()
            bt = self._batches[m] = B.RVEBatch([self.model] * m)
        return bt

    def _solve_batched(self, strains: np.ndarray, time_step, settings, committed: RVEBatchState,
                       need_tangent: bool = False):
        """All macro points' RVEs in batched solves of at most BATCH_GAUSS_POINTS Gauss
        points: the per-point state layout [n][8][n_micro] is permuted to the batch
        arena's [8][m n_micro] and back.  Effective tangents, when asked for, come from
        the batched Schur-complement solves of each chunk.  The last chunk is padded to a
        power of two up to 64 RVEs, beyond that to a multiple of 64 (copies of its last
        point, results discarded), so that few distinct batch sizes are ever built."""
        from concrete_impact_ai_amd import batch as B

        n = strains.shape[0]
        c = self.operator_cache
        cap = max(1, self.BATCH_GAUSS_POINTS // c.n_points)
        steps_all = np.broadcast_to(np.asarray(time_step, dtype=np.float64), (n,))
        results: list = []
        packs, ws = [], []
        for s0 in range(0, n, cap):
            m = min(cap, n - s0)
            mb = cap if m == cap else min(cap, 1 << (m - 1).bit_length() if m <= 64 else -(-m // 64) * 64)
            rows = np.minimum(np.arange(s0, s0 + mb), s0 + m - 1)
            rows_t = torch.as_tensor(rows, device=committed.packed.device)
            bt = self._batch_for(mb)
            packed = committed.packed[rows_t].permute(1, 0, 2).reshape(8, mb * c.n_points)
            status, res, nxt_p, nxt_w = B.solve_batch_raw(
                bt, packed.contiguous(), committed.w[rows_t].reshape(-1).contiguous(),
                committed.macro[rows], strains[rows], steps_all[rows], settings)
            if need_tangent and status == _lib.CIA_OK:
                status, _ = B.effective_tangents_raw(bt, res)
            if status == _lib.CIA_ERR_RVE_FAILURE:
                i = next(k for k, r in enumerate(res[:m]) if r.reason != 0)
                B.raise_failure(bt, i, res[i], strains[s0 + i], float(steps_all[s0 + i]), settings)
            results.extend(list(res)[:m])
            packs.append(nxt_p.reshape(8, mb, c.n_points)[:, :m].permute(1, 0, 2))
            ws.append(nxt_w.reshape(mb, 3 * c.n_classes)[:m])
        nxt_packed = torch.cat(packs).contiguous() if len(packs) > 1 else packs[0].contiguous()
        nxt_w = torch.cat(ws) if len(ws) > 1 else ws[0]
        return results, RVEBatchState(nxt_packed, nxt_w.contiguous(), strains.copy(), c)

    def _solve_all(self, strains: np.ndarray, time_step: float, settings, need_tangent: bool,
                   committed: RVEBatchState):
        n = strains.shape[0]
        if self._batched and n > 0:
            return self._solve_batched(strains, time_step, settings, committed, need_tangent)
        nxt_packed = torch.empty_like(committed.packed)
        nxt_w = torch.empty_like(committed.w)
        results: list = [None] * n

        steps = np.broadcast_to(np.asarray(time_step, dtype=np.float64), (n,))

        def work(wid: int, stream_ptr):
            cache = self._caches[wid]
            for i in range(wid, n, self._workers):
                try:
                    results[i] = solve_increment_raw(
                        cache, stream_ptr, committed.packed[i].data_ptr(), committed.macro[i],
                        committed.w[i].data_ptr(), strains[i], float(steps[i]), settings,
                        need_tangent, nxt_packed[i].data_ptr(), nxt_w[i].data_ptr())
                except Exception as exc:  # first failure of this worker, by point order
                    return i, exc
            return None

        if self._workers == 1:
            fail = work(0, stream_handle())
            failures = [] if fail is None else [fail]
        else:
            torch.cuda.current_stream().synchronize()  # inputs were written on this stream
            futures = [self._executor.submit(work, k, C.c_void_p(self._streams[k].cuda_stream))
                       for k in range(self._workers)]
            failures = [f for f in (fu.result() for fu in futures) if f is not None]
            for s in self._streams:
                s.synchronize()
        if failures:
            raise min(failures, key=lambda t: t[0])[1]
        return results, RVEBatchState(nxt_packed, nxt_w, strains.copy(), self.operator_cache)

    def _phase_q(self, packed: torch.Tensor):
        """Volume-average and maximum equivalent plastic strain per J2 phase."""
        c = self.operator_cache
        out = {}
        q = packed[:, 6, :].double()
        for pid, (pname, mat) in enumerate(c.phases):
            if not isinstance(mat, J2ViscoplasticMaterial):
                continue
            if c.phased:
                mask = torch.as_tensor(c.point_phase_ids == pid, device=q.device)
                qp = q[:, mask]
            else:
                qp = q
            out[pid] = (qp.mean(dim=1).cpu().numpy(), qp.max(dim=1).values.cpu().numpy())
        return out

    def _diagnostics(self, results, next_state: RVEBatchState) -> dict:
        """``_macro_diagnostics`` (``material.py:220-264``) for every point."""
        c = self.operator_cache
        n = len(results)
        d = {
            "incremental_work_density": np.asarray(
                [r.positive_work_density + r.negative_work_density for r in results]),
            "viscoplastic_active": np.asarray([r.viscoplastic_active_volume_fraction for r in results]),
            "yield_activation_margin": np.asarray([r.minimum_yield_activation_margin for r in results]),
            "micro_newton_iterations": np.asarray([float(r.newton_iterations) for r in results]),
            "micro_armijo_backtracks": np.asarray([float(r.armijo_backtracks) for r in results]),
            "hill_mandel_error": np.asarray([r.hill_mandel_error for r in results]),
        }
        qstats = self._phase_q(next_state.packed) if n else {}
        if not c.phased and 0 in qstats:
            d["equivalent_plastic_strain"] = qstats[0][0]
        for pid, pname in enumerate(self.model.phase_names):
            vol = np.asarray([r.phase_volume[pid] for r in results])
            d[f"phase_{pname}__volume_fraction"] = vol / c.volume
            for k, comp in enumerate(VOIGT_COMPONENTS):
                d[f"phase_{pname}__stress_{comp}"] = np.asarray(
                    [r.phase_average_stress[pid][k] for r in results])
            if pid in qstats:
                d[f"phase_{pname}__equivalent_plastic_strain"] = qstats[pid][0]
                d[f"phase_{pname}__maximum_equivalent_plastic_strain"] = qstats[pid][1]
        return {k: np.asarray(v, dtype=np.float64) for k, v in sorted(d.items())}

    def update(self, request, state: MaterialState,
               requirements: MaterialResponseRequirements = DEFAULT_RESPONSE_REQUIREMENTS
               ) -> MaterialPointResponse:
        """Solve independent RVEs for every macro point strain (``material.py:118-184``)."""
        strains = np.ascontiguousarray(np.asarray(request.strains, dtype=np.float64))
        n = strains.shape[0]
        committed = self._device_batch(state, n)
        results, nxt = self._solve_all(strains, float(request.time_step), request.update_settings,
                                       bool(requirements.tangent), committed)
        stresses = np.asarray([r.macro_stress[:] for r in results]).reshape(n, 6)
        tangents = (np.asarray([r.effective_tangent[:] for r in results]).reshape(n, 6, 6)
                    if requirements.tangent else None)
        return MaterialPointResponse(
            stresses=stresses,
            state=nxt,
            tangents=tangents,
            free_energy=(np.asarray([r.free_energy_density for r in results])
                         if requirements.free_energy else None),
            dissipation=(np.asarray([r.dissipation_density for r in results])
                         if requirements.dissipation else None),
            diagnostics=self._diagnostics(results, nxt),
        )

    def solve_batch(self, strains: np.ndarray, time_steps, settings, state: RVEBatchState,
                    need_tangent: bool = False):
        """Per-point time steps (the material protocol has one step for all points):
        returns the per-point solver results and the next batch state."""
        strains = np.ascontiguousarray(np.asarray(strains, dtype=np.float64))
        return self._solve_all(strains, np.asarray(time_steps, dtype=np.float64), settings,
                               bool(need_tangent), self._device_batch(state, strains.shape[0]))

    def device_kernel(self, point_phase_ids=None) -> "_RVEKernel":
        return _RVEKernel(self)

    def close(self) -> None:
        for bt in self._batches.values():
            bt.close()
        self._batches = {}
        if self._executor is not None:
            self._executor.shutdown(wait=True, cancel_futures=False)
            self._executor = None
        for c in self._caches:
            c.close()


class _RVEKernel:
    """Device-protocol view of the adapter for the generic macro drivers: macro
    strains come from the device strain pass, the stresses and the point arrays
    of the energy histories go back as device tensors."""

    def __init__(self, material: RVEHomogenizedMaterial):
        self.material = material

    def initial_state(self, n: int) -> RVEBatchState:
        return self.material.initialize_state(n)

    def update(self, strain, rate, state, dt, settings, tangent=False) -> DeviceResponse:
        m = self.material
        strains = strain.t().contiguous().cpu().numpy()
        results, nxt = m._solve_all(strains, float(dt), settings, bool(tangent), state)
        dev = strain.device

        def col(vals):
            return torch.as_tensor(np.asarray(vals, dtype=np.float64), device=dev)

        stress = torch.as_tensor(np.asarray([r.macro_stress[:] for r in results]).T.copy(),
                                 device=dev)
        tan = None
        if tangent:
            tan = torch.as_tensor(np.asarray([r.effective_tangent[:] for r in results]).T.copy(),
                                  device=dev)
        return DeviceResponse(
            stress=stress.contiguous(), state=nxt, tangent=tan,
            free_energy=col([r.free_energy_density for r in results]),
            dissipation=col([r.dissipation_density for r in results]),
            work=col([r.positive_work_density + r.negative_work_density for r in results]),
            active=col([r.viscoplastic_active_volume_fraction for r in results]),
            margin=col([r.minimum_yield_activation_margin for r in results]))

    def material_state(self, state) -> MaterialState:
        return state

    def equivalent_plastic_strain(self, state):
        """Per macro point RVE-average equivalent plastic strain (the macro
        diagnostic the reference records, ``material.py:242-246``); None for
        phased RVEs."""
        c = self.material.operator_cache
        if c.phased or not isinstance(c.phases[0][1], J2ViscoplasticMaterial):
            return None
        return state.packed[:, 6, :].double().mean(dim=1)
