"""Oracle: explicit velocity-Verlet material dynamics (FP64).

Test infrastructure only (see ``oracle/__init__.py``).

Restates ``solve_material_explicit`` (``src/fem/dynamics/material.py:89-297``):
initial assembly at u0 with zero rate and state commit (``:124-148``), the step
loop (``:164-284``: predictor, assembly from the committed state, acceleration,
velocity, Dirichlet zeroing), dissipation and external-work bookkeeping
(``:221-229``), ``_record_step`` energies (``:781-848``) and the lumped kinetic
energy (``:851-856``); CFL check ``check_explicit_stability``
(``src/fem/dynamics/stability.py:34-67``); uniform boundary pressure
(``src/fem/assembly/load.py:17-58``) and the half-sine pulse of
``src/concrete_impact/benchmarks/dynamic_viscoplastic.py:147-170``.

Pinning (tests/test_oracle_dynamics.py): an elastic bar's energy balance
(kinetic + free - external work = 0 to the Verlet truncation order), momentum
balance of the free-free bar, the reference's explicit acceptance bound
(energy residual < 0.05 of the external work, ``test_plasticity_benchmarks.py:58-75``),
dissipation >= 0 and the CFL rejection (``test_beam_response_benchmarks.py:70``).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from oracle import hex8, j2vp, rve_rno


def box_faces(divisions):
    """Node ids of the six faces of a structured box (node id = i + (nx+1)(j + (ny+1)k))."""
    nx, ny, nz = divisions
    k, j, i = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    ids = np.arange(i.size)
    return {"left": ids[i == 0], "right": ids[i == nx], "front": ids[j == 0], "back": ids[j == ny],
            "bottom": ids[k == 0], "top": ids[k == nz]}


def top_patch_unit_pressure(lengths, divisions, patch):
    """Equivalent nodal force of a unit pressure on the z = top face quads whose
    element indices lie in ``patch = ((i0, i1), (j0, j1))`` (outward normal +z,
    traction = -p n, area / 4 to each corner: ``load.py:17-58``)."""
    nx, ny, nz = divisions
    hx, hy = lengths[0] / nx, lengths[1] / ny
    f = np.zeros(3 * (nx + 1) * (ny + 1) * (nz + 1))
    (i0, i1), (j0, j1) = patch
    for j in range(j0, j1):
        for i in range(i0, i1):
            for di, dj in ((0, 0), (1, 0), (0, 1), (1, 1)):
                node = (i + di) + (nx + 1) * ((j + dj) + (ny + 1) * nz)
                f[3 * node + 2] += -1.0 * hx * hy / 4.0
    return f


def half_sine(amplitude, duration):
    def scale(t):
        return amplitude * np.sin(np.pi * t / duration) if 0.0 <= t <= duration else 0.0
    return scale


class Box:
    def __init__(self, lengths, divisions, phases, element_phase_ids=None, density=None):
        """``density``: the bundle material's density used by the mass matrices for
        every element (``elasticity.py:109-126``); defaults to the single phase's."""
        self.lengths, self.divisions = lengths, divisions
        self.nodes, self.elements = hex8.structured_box(lengths, divisions)
        self.b, self.jw = hex8.geometry_cache(self.nodes, self.elements)
        self.phases = phases
        ne = self.elements.shape[0]
        self.eph = np.zeros(ne, dtype=np.int64) if element_phase_ids is None else np.asarray(element_phase_ids)
        self.ppid = np.repeat(self.eph, 8)
        self.phased = element_phase_ids is not None
        self.density = float(phases[0][1].density if density is None else density)
        rho = np.full(ne, self.density)
        self.mass = hex8.lumped_mass(self.nodes, self.elements, rho)
        self.n_dof = 3 * self.nodes.shape[0]

    def init_state(self):
        st = {}
        mat0 = self.phases[0][1]
        if isinstance(mat0, rve_rno.RVERNO):
            return {"latent_state": np.zeros((self.ppid.size, mat0.config.latent_dimension))}
        for pid, (name, mat) in enumerate(self.phases):
            if isinstance(mat, j2vp.J2VPParams):
                cnt = int(np.sum(self.ppid == pid))
                pre = f"phase_{name}__" if self.phased else ""
                st[pre + "plastic_strain"] = np.zeros((cnt, 6))
                st[pre + "equivalent_plastic_strain"] = np.zeros(cnt)
                st[pre + "viscoplastic_multiplier"] = np.zeros(cnt)
        return st

    def assemble(self, u, u_prev, dt, settings, state, need_tangent=False):
        eps, rate = hex8.strain_fields(self.elements, self.b, u, u_prev, dt)
        if self.phased:
            # The reference's phased model raises for elastic phases without a tangent
            # (phased.py:178-179), so explicit dynamics of phased media needs the tangent
            # for the elastic-phase diagnostics; the GPU build evaluates them directly.
            r = j2vp.phased_update(self.phases, self.ppid, eps, state, dt, settings, rate, True)
        else:
            mat = self.phases[0][1]
            if isinstance(mat, rve_rno.RVERNO):
                # the RVE-RNO deployment adapter (nn/deployment.py:97-165): one explicit
                # update of every point from its committed latent state
                r = mat.update_batch(eps, state["latent_state"], np.full(eps.shape[0], dt),
                                     tangent=need_tangent)
                r["state"] = {"latent_state": r["latent_state"]}
            elif isinstance(mat, j2vp.J2VPParams):
                r = j2vp.j2vp_update(mat, eps, state["plastic_strain"], state["equivalent_plastic_strain"],
                                     state["viscoplastic_multiplier"], dt, settings, rate, need_tangent)
                r["state"] = {k: r[k] for k in ("plastic_strain", "equivalent_plastic_strain",
                                                "viscoplastic_multiplier")}
            else:
                r = j2vp.elastic_update(mat, eps, dt, rate, need_tangent, False)
                r["state"] = {}
        r["internal_force"] = hex8.internal_force(self.elements, self.b, self.jw, r["stress"], self.n_dof)
        if need_tangent:
            r["tangent_matrix"] = hex8.tangent_matrix(self.elements, self.b, self.jw, r["tangent"],
                                                      self.n_dof)
        return r


def explicit(box: Box, dt, n_steps, fixed_dofs, f_unit, scale, settings, cfl=None, wave_speed=None):
    """Velocity Verlet; returns histories like ``MaterialDynamicSolution``."""
    if cfl is not None:
        h = min(box.lengths[i] / box.divisions[i] for i in range(3))
        if dt > cfl * h / wave_speed:
            raise ValueError("Explicit time step violates the CFL stability bound")
    n = box.n_dof
    u = np.zeros(n)
    v = np.zeros(n)
    state = box.init_state()
    asm = box.assemble(u, u, dt, settings, state)
    state = asm["state"]
    ext = scale(0.0) * f_unit
    a = (ext - asm["internal_force"]) / box.mass
    a[fixed_dofs] = 0.0
    w = box.jw.reshape(-1)
    hist = {"u": [u.copy()], "v": [v.copy()], "a": [a.copy()], "kinetic": [0.5 * np.sum(box.mass * v**2)],
            "free": [np.sum(asm["free_energy"] * w)], "dissipated": [0.0], "external_work": [0.0],
            "stress": [asm["stress"].copy()], "internal_force": [asm["internal_force"].copy()]}
    for s in range(n_steps):
        t1 = (s + 1) * dt
        ext1 = scale(t1) * f_unit
        un = u + dt * v + 0.5 * dt**2 * a
        un[fixed_dofs] = 0.0
        asm = box.assemble(un, u, dt, settings, state)
        an = (ext1 - asm["internal_force"]) / box.mass
        an[fixed_dofs] = 0.0
        vn = v + 0.5 * dt * (a + an)
        vn[fixed_dofs] = 0.0
        hist["dissipated"].append(hist["dissipated"][-1] + dt * np.sum(asm["dissipation"] * w))
        hist["external_work"].append(hist["external_work"][-1] + 0.5 * (ext + ext1) @ (un - u))
        u, v, a, ext, state = un, vn, an, ext1, asm["state"]
        hist["u"].append(u.copy())
        hist["v"].append(v.copy())
        hist["a"].append(a.copy())
        hist["kinetic"].append(0.5 * np.sum(box.mass * v**2))
        hist["free"].append(np.sum(asm["free_energy"] * w))
        hist["stress"].append(asm["stress"].copy())
        hist["internal_force"].append(asm["internal_force"].copy())
    out = {k: np.asarray(v) for k, v in hist.items()}
    out["energy_residual"] = (out["kinetic"] + out["free"] + out["dissipated"] - out["external_work"]
                              - (out["kinetic"][0] + out["free"][0]))
    out["state"] = state
    return out


class NewmarkStepError(RuntimeError):
    """A failed Newton step (``NonlinearStepConvergenceError``, ``material.py:75-82``);
    ``recoverable`` marks the reasons eligible for a cut-back
    (``RecoverableNonlinearStepError``, ``material.py:85-86``)."""

    def __init__(self, reason, diagnostics, recoverable=True):
        super().__init__(f"Implicit Newmark step failed: {reason}.")
        self.reason = reason
        self.diagnostics = diagnostics
        self.recoverable = recoverable


@dataclass(frozen=True)
class Adaptive:
    """``AdaptiveTimeStepSettings`` (``src/fem/solvers/data.py:51-74``)."""

    enabled: bool = False
    minimum_factor: float = 1.0 / 16.0
    maximum_factor: float = 2.0
    cutback_factor: float = 0.5
    growth_factor: float = 2.0
    easy_iteration_limit: int = 4
    easy_step_count: int = 3


def cut_back(current, initial, adaptive: Adaptive):
    """``cut_back_time_step`` (``src/fem/solvers/time_step.py:12-21``)."""
    return max(initial * adaptive.minimum_factor, current * adaptive.cutback_factor)


def grow(current, initial, adaptive: Adaptive):
    """``grow_time_step`` (``src/fem/solvers/time_step.py:24-31``)."""
    return min(initial * adaptive.maximum_factor, current * adaptive.growth_factor)


def newmark(box: Box, dt, n_steps, fixed_dofs, f_unit, scale, settings, newton, beta=0.25,
            gamma=0.5, adaptive: Adaptive | None = None):
    """Implicit Newmark with full-step Newton and Armijo (``solve_material_implicit_newmark``,
    ``material.py:300-443``; ``_solve_implicit_step`` ``:446-662``; ``_evaluate_implicit_trial``
    ``:665-701``), consistent mass, SuperLU solves, and the adaptive cut-back / growth
    of the step (``material.py:377-430``, ``time_step.py:12-31``) when
    ``adaptive.enabled``.

    ``newton`` is an ``oracle.rve.Newton``.  Returns histories of the accepted
    states like ``MaterialDynamicSolution`` (consistent kinetic energy ``:859-864``),
    with ``times``, ``time_steps``, ``rejected_counts`` and ``rejected_reasons``."""
    from scipy.sparse.linalg import spsolve

    from oracle.rve import ArmijoExhausted, armijo_search

    adaptive = adaptive or Adaptive()
    n = box.n_dof
    M = hex8.consistent_mass(box.nodes, box.elements, box.density)
    fixed = np.zeros(n, dtype=bool)
    fixed[np.asarray(fixed_dofs)] = True
    free = np.flatnonzero(~fixed)
    Mff = M[free][:, free].tocsc()
    w = box.jw.reshape(-1)
    u = np.zeros(n)
    v = np.zeros(n)
    a = np.zeros(n)
    state = box.init_state()
    asm = box.assemble(u, u, dt, settings, state)
    state = asm["state"]
    a[free] = spsolve(Mff, (scale(0.0) * f_unit - asm["internal_force"])[free])
    hist = {"u": [u.copy()], "v": [v.copy()], "a": [a.copy()], "kinetic": [0.5 * v @ (M @ v)],
            "free": [np.sum(asm["free_energy"] * w)], "dissipated": [0.0], "external_work": [0.0],
            "stress": [asm["stress"].copy()], "internal_force": [asm["internal_force"].copy()],
            "newton_iterations": [0], "armijo_backtracks": [0], "times": [0.0], "time_steps": [0.0],
            "rejected_counts": [0]}
    rejected_reasons = [()]

    def trial(disp, up, ustar, vstar, t_dt, ext, committed):
        acc = (disp - ustar) / (beta * t_dt**2)
        vel = vstar + gamma * t_dt * acc
        r = box.assemble(disp, up, t_dt, settings, committed, need_tangent=True)
        res = M @ acc + r["internal_force"] - ext
        return acc, vel, r, res

    def step(t0, u, v, a, state, h):
        """``_solve_implicit_step`` (``material.py:446-662``)."""
        t1 = t0 + h
        ustar = u + h * v + h**2 * (0.5 - beta) * a
        vstar = v + h * (1.0 - gamma) * a
        disp = ustar.copy()
        disp[fixed] = 0.0
        ext = scale(t1) * f_unit
        r0 = None
        last_inc = np.inf
        backtracks = 0
        for it in range(newton.max_iterations + 1):
            acc, vel, r, res = trial(disp, u, ustar, vstar, h, ext, state)
            rn = float(np.linalg.norm(res[free]))
            r0 = rn if r0 is None else r0
            rtol = max(newton.residual_absolute_tolerance, newton.residual_relative_tolerance * r0)
            itol = max(newton.increment_absolute_tolerance,
                       newton.increment_relative_tolerance * max(float(np.linalg.norm(disp[free])), 1.0))
            if rn <= rtol and (it == 0 or last_inc <= itol):
                break
            if it == newton.max_iterations:
                raise NewmarkStepError("maximum_iterations_exceeded",
                                       {"iteration": it, "time_step": h})
            keff = (r["tangent_matrix"] + M / (beta * h**2))[free][:, free].tocsc()
            inc = spsolve(keff, -res[free])
            if not np.all(np.isfinite(inc)):
                raise NewmarkStepError("non_finite_increment", {"iteration": it, "time_step": h},
                                       recoverable=False)
            inc_n = float(np.linalg.norm(inc))
            if rn <= rtol and inc_n <= itol:
                break
            if newton.armijo.enabled:
                def evaluate(alpha, base=disp, d=inc):
                    cand = base.copy()
                    cand[free] += alpha * d
                    return trial(cand, u, ustar, vstar, h, ext, state)[3][free]
                try:
                    alpha, bt, _ = armijo_search(res[free], evaluate, newton.armijo)
                except ArmijoExhausted as exc:
                    raise NewmarkStepError("armijo_exhausted",
                                           {"iteration": it, "time_step": h,
                                            "armijo": exc.diagnostics}) from exc
                disp[free] += alpha * inc
                last_inc = float(np.linalg.norm(alpha * inc))
                backtracks += bt
            else:
                disp[free] += inc
                last_inc = inc_n
        return {"t": t1, "u": disp, "v": vel, "a": acc, "r": r, "it": it, "bt": backtracks,
                "diss": h * np.sum(r["dissipation"] * w),
                "work": 0.5 * (scale(t0) * f_unit + ext) @ (disp - u)}

    final_time = dt * n_steps
    minimum = dt * adaptive.minimum_factor
    cur_dt = dt
    easy = 0
    pending = []
    t = 0.0
    while t < final_time:
        h = min(cur_dt, final_time - t)
        try:
            out = step(t, u, v, a, state, h)
        except NewmarkStepError as err:
            if not (adaptive.enabled and err.recoverable):
                raise
            pending.append(err.reason)
            if h <= minimum:
                err.diagnostics["minimum_time_step"] = minimum
                err.diagnostics["rejected_reasons"] = tuple(pending)
                raise
            cur_dt = cut_back(h, dt, adaptive)
            easy = 0
            continue
        t, u, v, a, state = out["t"], out["u"], out["v"], out["a"], out["r"]["state"]
        hist["dissipated"].append(hist["dissipated"][-1] + out["diss"])
        hist["external_work"].append(hist["external_work"][-1] + out["work"])
        hist["u"].append(u.copy())
        hist["v"].append(v.copy())
        hist["a"].append(a.copy())
        hist["kinetic"].append(0.5 * v @ (M @ v))
        hist["free"].append(np.sum(out["r"]["free_energy"] * w))
        hist["stress"].append(out["r"]["stress"].copy())
        hist["internal_force"].append(out["r"]["internal_force"].copy())
        hist["newton_iterations"].append(out["it"])
        hist["armijo_backtracks"].append(out["bt"])
        hist["times"].append(t)
        hist["time_steps"].append(h)
        hist["rejected_counts"].append(len(pending))
        rejected_reasons.append(tuple(pending))
        is_easy = (not pending and out["bt"] == 0 and out["it"] <= adaptive.easy_iteration_limit)
        easy = easy + 1 if is_easy else 0
        pending = []
        if adaptive.enabled and easy >= adaptive.easy_step_count:
            cur_dt = grow(h, dt, adaptive)
            easy = 0
        else:
            cur_dt = h
    out = {k: np.asarray(v) for k, v in hist.items()}
    out["energy_residual"] = (out["kinetic"] + out["free"] + out["dissipated"] - out["external_work"]
                              - (out["kinetic"][0] + out["free"][0]))
    out["state"] = state
    out["rejected_reasons"] = tuple(rejected_reasons)
    return out
