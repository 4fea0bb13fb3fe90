"""Batched periodic RVEs: many independent structured RVEs of different sizes
solved together on one device.

Reference: the reference solves RVEs one at a time — the per-point loop of
``RVEHomogenizedMaterial.update`` (``src/fem/rve/material.py:104-191``, process
pool workers ``:267-291``) and one ``iterate_prescribed_rve_path`` per RVE
(``src/fem/rve/path.py:47-100``); every increment is
``solve_rve_microequilibrium`` (``src/fem/rve/solver.py:62-227``).  Here one
``cia_rve_batch_solve`` call advances every RVE of the batch by one increment:
each kernel launch covers all RVEs, each RVE runs the reference's Newton /
Armijo state machine on the device, and the grids of the batched kernels are
chosen by an autotuner on the first solve.

Device layout: RVE ``r`` owns elements ``[offsets[r], offsets[r + 1])``; the
state rows are ``[8][8 offsets[-1]]`` (component-major over the whole batch) and
the fluctuations ``[3 offsets[-1]]`` (class 0 of each RVE holds zeros).
"""

from __future__ import annotations

import ctypes as C
from collections.abc import Iterator, Sequence
from dataclasses import dataclass

import numpy as np
import torch

from concrete_impact_ai_amd import _lib
from concrete_impact_ai_amd.materials import (
    MaterialUpdateSettings,
    cia_settings,
    device,
    point_failure_error,
    stream_handle,
)
from concrete_impact_ai_amd.rve import (
    DeviceMaterialState,
    RVEEquilibriumError,
    RVEState,
    StructuredHex8RVE,
    _box_lengths,
    _phases_of,
    linear_method,
    rve_descriptor,
)

# Homogenised diagnostics the batched solve fills (per RVE); the field
# diagnostics of the single-RVE response (periodic, reaction and virtual-work
# errors, activation margins) need the full node set and are not computed here.
BATCH_DIAGNOSTICS = (
    "micro_residual_norm", "micro_residual_tolerance", "newton_iterations", "armijo_backtracks",
    "volume", "macro_work_increment", "micro_work_increment", "hill_mandel_residual",
    "hill_mandel_error", "positive_work_density", "negative_work_density",
    "unloading_volume_fraction", "viscoplastic_active_volume_fraction", "linear_iterations",
    "gauss_point_updates",
)


class RVEBatch:
    """Device operator of a batch of structured periodic RVEs.

    All models share precision, Newton and linear-solver settings (spectral
    preconditioner); divisions, lengths, phases and microstructure are free.
    """

    def __init__(self, models: Sequence[StructuredHex8RVE], autotune: bool = True):
        models = list(models)
        if not models:
            raise ValueError("An RVE batch needs at least one model.")
        p0 = models[0].precision
        if any(m.precision != p0 for m in models):
            raise ValueError("All RVEs of a batch share one precision.")
        if any(m.newton_settings != models[0].newton_settings or
               m.linear_settings != models[0].linear_settings for m in models):
            raise ValueError("All RVEs of a batch share Newton and linear-solver settings.")
        if linear_method(models[0].linear_settings) != "spectral":
            raise ValueError("The batched RVE solver uses the spectral preconditioner.")
        self.models = models
        self.precision = p0
        self.dtype = torch.float64 if p0 == "fp64" else torch.float32
        self.phases, self.phased, self.point_phase_ids = [], [], []
        descs = (_lib.RVEDesc * len(models))()
        self._keep = []
        for i, m in enumerate(models):
            phases, phased = _phases_of(m)
            div = tuple(int(d) for d in m.divisions)
            if max(div) > 64 or min(div) < 2:
                raise ValueError("Batched RVE edges must lie in [2, 64] elements.")
            n_el = div[0] * div[1] * div[2]
            ep = (np.zeros(n_el, dtype=np.int64) if m.element_phase_ids is None
                  else np.asarray(m.element_phase_ids, dtype=np.int64))
            desc, keep = rve_descriptor(m, phases, div, _box_lengths(m), ep)
            descs[i] = desc
            self._keep.append(keep)
            self.phases.append(phases)
            self.phased.append(phased)
            self.point_phase_ids.append(np.repeat(ep, 8))
        device()
        handle = C.c_void_p()
        _lib.check(_lib.lib().cia_rve_batch_create(descs, len(models), int(bool(autotune)),
                                                   C.byref(handle)), "cia_rve_batch_create")
        self.handle = handle
        off = (C.c_int64 * (len(models) + 1))()
        _lib.check(_lib.lib().cia_rve_batch_layout(handle, off), "cia_rve_batch_layout")
        self.element_offsets = np.asarray(off[:], dtype=np.int64)
        self.n_rves = len(models)
        self.n_elements = int(self.element_offsets[-1])
        self.n_points = 8 * self.n_elements

    def __len__(self) -> int:
        return self.n_rves

    def tuning(self) -> dict:
        """Autotuned grid (workgroups; 0 = one tile per workgroup) and time (us)
        of every batched kernel."""
        n = len(_lib.BATCH_KERNELS)
        grids = (C.c_int32 * n)()
        us = (C.c_double * n)()
        _lib.check(_lib.lib().cia_rve_batch_tuning(self.handle, grids, us, n), "cia_rve_batch_tuning")
        return {k: {"grid": int(grids[i]), "us": float(us[i])}
                for i, k in enumerate(_lib.BATCH_KERNELS)}

    def stats(self) -> dict:
        """Counters since creation: solves, Newton rounds, PCG iterations launched for
        the whole batch, PCG iterations summed over the RVEs, host synchronisations."""
        c = (C.c_int64 * 5)()
        _lib.check(_lib.lib().cia_rve_batch_stats(self.handle, c, 5), "cia_rve_batch_stats")
        return dict(zip(("solves", "newton_rounds", "pcg_iterations_launched",
                         "pcg_iterations_rve_sum", "host_syncs"), (int(v) for v in c)))

    def set_tuning(self, grids: dict) -> None:
        """Fix the kernel grids (before the first solve) instead of autotuning."""
        n = len(_lib.BATCH_KERNELS)
        arr = (C.c_int32 * n)(*[int(grids.get(k, 0)) for k in _lib.BATCH_KERNELS])
        _lib.check(_lib.lib().cia_rve_batch_set_tuning(self.handle, arr, n), "cia_rve_batch_set_tuning")

    def close(self) -> None:
        if getattr(self, "handle", None) is not None and self.handle.value:
            _lib.lib().cia_rve_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state ---------------------------------------------------------------

    def initial_state(self) -> "RVEBatchState":
        """Virgin states of every RVE (``initialize_rve_state``, solver.py:46-59)."""
        dev = device()
        packed = torch.zeros((8, self.n_points), dtype=self.dtype, device=dev)
        w = torch.zeros(3 * self.n_elements, dtype=torch.float64, device=dev)
        return RVEBatchState(self, packed, w, np.zeros((self.n_rves, 6)))

    def point_slice(self, i: int) -> slice:
        return slice(8 * int(self.element_offsets[i]), 8 * int(self.element_offsets[i + 1]))

    def class_slice(self, i: int) -> slice:
        return slice(3 * int(self.element_offsets[i]), 3 * int(self.element_offsets[i + 1]))


class RVEBatchState:
    """Committed states of all RVEs of a batch, device-resident."""

    def __init__(self, batch: RVEBatch, packed: torch.Tensor, w: torch.Tensor, macro: np.ndarray):
        self.batch = batch
        self.packed = packed
        self.w = w
        self.macro = np.asarray(macro, dtype=np.float64)

    def rve(self, i: int) -> RVEState:
        """One RVE's committed state (``RVEState``) as device copies."""
        b = self.batch
        packed = self.packed[:, b.point_slice(i)].contiguous()
        ms = DeviceMaterialState(packed, b.phases[i],
                                 b.point_phase_ids[i] if b.phased[i] else None, b.phased[i])
        return RVEState(ms, self.macro[i].copy(), None,
                        w_device=self.w[b.class_slice(i)].contiguous())


@dataclass(frozen=True)
class RVEBatchResponse:
    """Homogenised responses of one batched increment."""

    macro_stress: np.ndarray          # (n, 6)
    free_energy_density: np.ndarray   # (n,)
    dissipation_density: np.ndarray   # (n,)
    diagnostics: dict                 # name -> (n,) array (BATCH_DIAGNOSTICS)
    reasons: tuple                    # per RVE: "ok" or the reference failure reason
    state: RVEBatchState
    effective_tangent: np.ndarray | None = None  # (n, 6, 6) when asked for, NaN where failed

    @property
    def converged(self) -> np.ndarray:
        return np.asarray([r == "ok" for r in self.reasons])


def solve_rve_batch(batch: RVEBatch, macro_strains, time_steps,
                    update_settings: MaterialUpdateSettings, committed: RVEBatchState,
                    raise_on_failure: bool = True,
                    initial_w: torch.Tensor | None = None,
                    need_tangent: bool = False) -> RVEBatchResponse:
    """One strain-driven increment of every RVE of the batch.

    ``macro_strains`` (n, 6), ``time_steps`` (n,).  ``initial_w`` (device, shaped
    like ``committed.w``) starts Newton from a predicted fluctuation instead of the
    committed one (the reference's start); the converged state meets the same
    tolerances.  ``need_tangent`` also returns every converged RVE's effective
    tangent (``solver.py:438-461``, batched by ``effective_tangents_raw``); its
    six linear solves count in ``linear_iterations``.  With ``raise_on_failure`` the
    lowest-index failed RVE raises the reference's exception
    (``RVEEquilibriumError`` or ``MaterialPointConvergenceError``) with
    ``rve_index`` in its diagnostics; otherwise failures are reported in
    ``reasons`` and the failed RVEs' next states are not valid.
    """
    n = batch.n_rves
    macro = np.ascontiguousarray(macro_strains, dtype=np.float64)
    dts = np.ascontiguousarray(np.broadcast_to(np.asarray(time_steps, dtype=np.float64), (n,)))
    if macro.shape != (n, 6):
        raise ValueError("Batched RVE macro strains must have shape (n_rves, 6).")
    if np.any(dts <= 0.0):
        raise ValueError("RVE time step must be positive.")
    if committed.packed.dtype != batch.dtype or committed.packed.shape != (8, batch.n_points):
        raise ValueError("committed state does not belong to this batch")
    status, res, next_packed, next_w = solve_batch_raw(
        batch, committed.packed, committed.w, committed.macro, macro, dts, update_settings,
        initial_w)
    tangents = None
    if need_tangent:
        tstatus, tangents = effective_tangents_raw(batch, res)
        if tstatus == _lib.CIA_ERR_RVE_FAILURE:
            status = tstatus
    a = np.frombuffer(res, dtype=_RESULT_DTYPE)  # the cia_rve_result array as columns
    reasons = tuple(_REASONS[a["reason"]])
    if status == _lib.CIA_ERR_RVE_FAILURE and raise_on_failure:
        i = next(k for k, r in enumerate(reasons) if r != "ok")
        raise_failure(batch, i, res[i], macro[i], float(dts[i]), update_settings)
    diag = {key: a[field].copy() for key, field in _DIAG_FIELDS}
    return RVEBatchResponse(
        macro_stress=a["macro_stress"].astype(np.float64),
        free_energy_density=a["free_energy_density"].copy(),
        dissipation_density=a["dissipation_density"].copy(),
        diagnostics=diag,
        reasons=reasons,
        state=RVEBatchState(batch, next_packed, next_w, macro.copy()),
        effective_tangent=tangents,
    )


_RESULT_DTYPE = np.dtype(_lib.RVEResult)
_REASONS = np.asarray(_lib.RVE_REASONS, dtype=object)
_DIAG_FIELDS = tuple((k, k) for k in (
    "micro_residual_norm", "micro_residual_tolerance", "newton_iterations", "armijo_backtracks",
    "volume", "macro_work_increment", "micro_work_increment", "hill_mandel_residual",
    "hill_mandel_error", "positive_work_density", "negative_work_density",
    "unloading_volume_fraction", "viscoplastic_active_volume_fraction",
    "linear_iterations")) + (("gauss_point_updates", "gp_updates"),)


def solve_batch_raw(batch: RVEBatch, packed: torch.Tensor, w: torch.Tensor, committed_macro,
                    macro, time_steps, update_settings, initial_w: torch.Tensor | None = None):
    """One increment on raw device arrays (state rows [8][n_points], fluctuations
    [3 n_elements]); returns the status, the per-RVE ``cia_rve_result`` array and
    the next state rows and fluctuations."""
    packed = packed.contiguous()
    w = w.contiguous()
    w0 = None
    if initial_w is not None:
        if initial_w.shape != w.shape or initial_w.dtype != torch.float64:
            raise ValueError("initial_w must be a float64 tensor shaped like the fluctuations")
        w0 = initial_w.contiguous()
    n = batch.n_rves
    macro = np.ascontiguousarray(macro, dtype=np.float64)
    dts = np.ascontiguousarray(np.broadcast_to(np.asarray(time_steps, dtype=np.float64), (n,)))
    next_packed = torch.empty_like(packed)
    next_w = torch.empty_like(w)
    res = (_lib.RVEResult * n)()
    cmacro = np.ascontiguousarray(committed_macro, dtype=np.float64)
    settings = cia_settings(update_settings)
    dptr = C.POINTER(C.c_double)
    status = _lib.lib().cia_rve_batch_solve(
        batch.handle, stream_handle(), C.c_void_p(packed.data_ptr()), C.c_void_p(w.data_ptr()),
        C.c_void_p(None if w0 is None else w0.data_ptr()), cmacro.ctypes.data_as(dptr),
        macro.ctypes.data_as(dptr), dts.ctypes.data_as(dptr), C.byref(settings),
        C.c_void_p(next_packed.data_ptr()), C.c_void_p(next_w.data_ptr()), res)
    _lib.check(status, "cia_rve_batch_solve")
    return status, res, next_packed, next_w


_NON_FINITE_TANGENT = _lib.RVE_REASONS.index("non_finite_effective_tangent_sensitivity")


def effective_tangents_raw(batch: RVEBatch, res):
    """Effective tangents of the RVEs that converged in the batch's last solve
    (``_compute_effective_tangent``, ``src/fem/rve/solver.py:438-461``: the Schur
    complement of the converged tangent, six linear solves with the Newton systems'
    preconditioned CG, all RVEs together).  Fills ``effective_tangent`` of each
    converged entry of ``res`` (the ``cia_rve_result`` array of that solve), adds the
    six solves' iterations to its ``linear_iterations`` and marks an RVE whose
    sensitivity came out non-finite with that failure reason; returns the status and
    the (n, 6, 6) tangents (NaN for RVEs that did not converge)."""
    n = batch.n_rves
    tan = np.empty((n, 36), dtype=np.float64)
    its = np.zeros(n, dtype=np.int32)
    status = _lib.lib().cia_rve_batch_effective_tangent(
        batch.handle, stream_handle(), tan.ctypes.data_as(C.c_void_p),
        its.ctypes.data_as(C.c_void_p))
    _lib.check(status, "cia_rve_batch_effective_tangent")
    for i in range(n):
        r = res[i]
        if r.reason != 0:
            continue
        if not np.all(np.isfinite(tan[i])):
            r.reason = _NON_FINITE_TANGENT
            continue
        C.memmove(r.effective_tangent, tan[i].ctypes.data, 36 * 8)
        r.linear_iterations += int(its[i])
    return status, tan.reshape(n, 6, 6)


def raise_failure(batch: RVEBatch, i: int, res, macro, dt, update_settings):
    if res.reason == 5:  # material point failure
        fail = res.point_failure
        p = int(fail.point_id)
        phases = batch.phases[i]
        mats = [m for _, m in phases]
        ppid = batch.point_phase_ids[i]
        phase = int(ppid[p])
        local = int(np.count_nonzero(ppid[:p] == phase)) if batch.phased[i] else p
        mreq = type("Req", (), {"time_step": dt, "kinematics": "three_dimensional",
                                "update_settings": update_settings})
        err = point_failure_error(fail, mats[phase], mreq, local, mats[phase])
        raise err.with_context(element_id=local // 8, quadrature_id=local % 8, rve_index=i)
    raise RVEEquilibriumError(_lib.RVE_REASONS[res.reason], {
        "algorithm": "periodic_rve_microequilibrium_newton",
        "rve_index": i,
        "macro_strain": np.asarray(macro).tolist(),
        "time_step": dt,
        "iteration": int(res.failure_iteration),
        "residual_norm": float(res.failure_residual_norm),
        "residual_tolerance": float(res.failure_residual_tolerance),
        "increment_norm": float(res.failure_increment_norm),
        "increment_tolerance": float(res.failure_increment_tolerance),
    })


@dataclass(frozen=True)
class AcceptedRVEBatchStep:
    step_id: int
    times: np.ndarray        # (n,) physical time of each RVE after the step
    time_steps: np.ndarray   # (n,)
    response: RVEBatchResponse


class FluctuationPredictor:
    """Newton start for the next increment of every RVE: the committed fluctuation
    extrapolated along the last increment, w_k + s (w_k - w_{k-1}) with
    s = clamp((d_eps_k . d_eps_{k-1}) / |d_eps_{k-1}|^2, 0, 2) per RVE (the
    fluctuation responds to the macro strain to first order, so on a straight
    path segment this is its linear response)."""

    def __init__(self, batch: RVEBatch):
        counts = np.diff(batch.element_offsets) * 3
        # RVE index of every fluctuation entry (expanded once: no host sync later)
        self._owner = torch.as_tensor(np.repeat(np.arange(len(counts)), counts), device=device())
        self._w_prev = None
        self._de_prev = None

    def initial_w(self, committed: RVEBatchState, macro_next) -> torch.Tensor | None:
        de = np.asarray(macro_next, dtype=np.float64) - committed.macro
        w0 = None
        if self._w_prev is not None:
            den = np.einsum("ij,ij->i", self._de_prev, self._de_prev)
            s = np.where(den > 0.0, np.einsum("ij,ij->i", de, self._de_prev) / np.maximum(den, 1e-300),
                         0.0)
            s = np.clip(s, 0.0, 2.0)
            scale = torch.as_tensor(s, device=self._owner.device)[self._owner]
            w0 = committed.w + scale * (committed.w - self._w_prev)
        self._w_prev = committed.w
        self._de_prev = de
        return w0


def iterate_prescribed_rve_paths(batch: RVEBatch, times, macro_strains,
                                 update_settings: MaterialUpdateSettings,
                                 state: RVEBatchState | None = None,
                                 predictor: bool = False) -> Iterator[AcceptedRVEBatchStep]:
    """``iterate_prescribed_rve_path`` (``src/fem/rve/path.py:47-100``) for every
    RVE of the batch at once: ``times`` (n, m + 1) and ``macro_strains``
    (n, m + 1, 6), one path per RVE, all with m increments.  ``predictor`` starts
    each increment's Newton from :class:`FluctuationPredictor` instead of the
    committed fluctuation."""
    times = np.asarray(times, dtype=np.float64)
    strains = np.asarray(macro_strains, dtype=np.float64)
    n = batch.n_rves
    if times.ndim != 2 or times.shape[0] != n or strains.shape != times.shape + (6,):
        raise ValueError("Prescribed RVE paths require times (n, m) and strains (n, m, 6).")
    dts = np.diff(times, axis=1)
    if times.shape[1] < 2 or np.any(dts <= 0.0):
        raise ValueError("Prescribed RVE path requires at least two increasing times.")
    st = state if state is not None else batch.initial_state()
    pred = FluctuationPredictor(batch) if predictor else None
    for k in range(1, times.shape[1]):
        w0 = pred.initial_w(st, strains[:, k]) if pred is not None else None
        resp = solve_rve_batch(batch, strains[:, k], dts[:, k - 1], update_settings, st,
                               initial_w=w0)
        st = resp.state
        yield AcceptedRVEBatchStep(k, times[:, k].copy(), dts[:, k - 1].copy(), resp)
