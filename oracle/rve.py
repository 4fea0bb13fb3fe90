"""Oracle: periodic-RVE Newton microequilibrium and homogenisation (FP64, SuperLU).

Test infrastructure only (see ``oracle/__init__.py``).

Restates:
* ``initialize_rve_state``            ``src/fem/rve/solver.py:46-59``
* ``solve_rve_microequilibrium``      ``solver.py:62-227`` (Newton on T^T f, SciPy
  ``spsolve`` as ``src/fem/solvers/linear.py:15-22``, Armijo as
  ``src/fem/solvers/line_search.py:40-67``)
* ``_build_rve_response``             ``solver.py:253-364``
* ``build_rve_operator_cache``        ``solver.py:367-388``
* ``_compute_phase_diagnostics``      ``solver.py:391-435``
* ``_compute_effective_tangent``      ``solver.py:438-461`` (Schur complement)
* ``iterate_prescribed_rve_path``     ``src/fem/rve/path.py:47-100``

Pinning (tests/test_oracle_rve.py): a homogeneous RVE equals the single
material point (stress 1e-10 abs, tangent rtol 1e-10 / atol 1e-9, state
1e-12; periodic error <= 1e-14; Hill-Mandel <= 1e-12; virtual work <= 1e-10)
as ``tests/unit/fem/test_rve.py:66-130`` and
``configs/benchmarks/rve/single_phase_j2_viscoplastic.yaml``; directional FD of
the effective tangent (1e-7, ``src/concrete_impact/benchmarks/rve.py:239-278``);
an equal-material two-phase RVE equals the homogeneous one; elastic two-phase
effective stiffness lies within the Voigt-Reuss bounds; Armijo acceptance,
backtracking and exhaustion (``tests/unit/fem/test_nonlinear_controls.py:22-58``).
The heterogeneous plastic RVE is pinned by the laminate closed form
(``oracle/laminate.py``, ``tests/test_oracle_laminate.py``): on RVEs layered
along x the Newton solution equals the per-phase recurrence at ~1e-14 for
the macro stress, every Gauss point's strain and the plastic state.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
from scipy.sparse.linalg import spsolve

from oracle import hex8, j2vp, periodic


@dataclass(frozen=True)
class Armijo:
    enabled: bool = False
    sufficient_decrease: float = 1.0e-4
    reduction_factor: float = 0.5
    max_backtracks: int = 20


@dataclass(frozen=True)
class Newton:
    max_iterations: int = 20
    residual_absolute_tolerance: float = 1.0e-10
    residual_relative_tolerance: float = 1.0e-10
    increment_absolute_tolerance: float = 1.0e-12
    increment_relative_tolerance: float = 1.0e-10
    armijo: Armijo = field(default_factory=Armijo)


class OracleRVEError(RuntimeError):
    def __init__(self, reason: str, diagnostics: dict):
        super().__init__(f"RVE microequilibrium failed: {reason}.")
        self.reason = reason
        self.diagnostics = diagnostics


class ArmijoExhausted(RuntimeError):
    def __init__(self, diagnostics):
        super().__init__("Armijo line search exhausted all configured candidates.")
        self.diagnostics = diagnostics


def armijo_search(r0, evaluate, settings: Armijo):
    """``perform_armijo_search`` (``line_search.py:40-67``)."""
    rn = float(np.linalg.norm(r0))
    merit = 0.5 * rn**2
    last = np.inf
    for bt in range(settings.max_backtracks + 1):
        alpha = settings.reduction_factor**bt
        cand = evaluate(alpha)
        last = float(np.linalg.norm(cand))
        if 0.5 * last**2 <= merit - settings.sufficient_decrease * alpha * rn**2:
            return alpha, bt, last
    raise ArmijoExhausted({"current_residual_norm": rn, "current_merit": merit,
                           "last_candidate_residual_norm": last,
                           "max_backtracks": settings.max_backtracks})


@dataclass
class Model:
    """Structured periodic RVE (``StructuredHex8RVE``, ``src/fem/rve/data.py:88-102``).

    ``phases`` is a list of ``(name, J2VPParams | ElasticParams)``; with
    ``element_phase_ids`` None the single phase is unprefixed.
    """

    lengths: tuple
    divisions: tuple
    phases: list
    newton: Newton
    element_phase_ids: np.ndarray | None = None
    hill_mandel_power_scale: float = 1.0
    mesh: tuple | None = None  # (nodes, elements, class_ids, faces) of an unstructured periodic RVE

    def __post_init__(self):
        if self.mesh is None:
            self.nodes, self.elements = hex8.structured_box(self.lengths, self.divisions)
            self.constraint = periodic.structured_constraint(self.nodes, self.divisions)
        else:
            nodes, elements, class_ids, faces = self.mesh
            self.nodes = np.asarray(nodes, dtype=np.float64)
            self.elements = np.asarray(elements, dtype=np.int64)
            self.constraint = periodic.class_constraint(self.nodes, class_ids, self.divisions, faces)
        self.b, self.jw = hex8.geometry_cache(self.nodes, self.elements)
        self.weights = self.jw.reshape(-1)
        self.volume = float(np.sum(self.weights))
        ep = (np.zeros(self.elements.shape[0], dtype=np.int64) if self.element_phase_ids is None
              else np.asarray(self.element_phase_ids, dtype=np.int64))
        self.point_phase_ids = np.repeat(ep, 8)
        self.n_dof = 3 * self.nodes.shape[0]

    @property
    def phased(self) -> bool:
        return self.element_phase_ids is not None

    def initial_material_state(self) -> dict:
        st = {}
        for pid, (name, mat) in enumerate(self.phases):
            if isinstance(mat, j2vp.J2VPParams):
                cnt = int(np.count_nonzero(self.point_phase_ids == pid))
                pre = f"phase_{name}__" if self.phased else ""
                st[pre + "plastic_strain"] = np.zeros((cnt, 6))
                st[pre + "equivalent_plastic_strain"] = np.zeros(cnt)
                st[pre + "viscoplastic_multiplier"] = np.zeros(cnt)
        return st

    def initial_state(self) -> dict:
        return {"material_state": self.initial_material_state(),
                "macro_strain": np.zeros(6),
                "fluctuation_dofs": np.zeros(self.constraint.transformation.shape[1])}

    def material_update(self, strains, rates, dt, settings, state):
        if self.phased:
            r = j2vp.phased_update(self.phases, self.point_phase_ids, strains, state, dt,
                                   settings, rates, True)
            return r
        name, mat = self.phases[0]
        if isinstance(mat, j2vp.J2VPParams):
            r = j2vp.j2vp_update(mat, strains, state["plastic_strain"],
                                 state["equivalent_plastic_strain"],
                                 state["viscoplastic_multiplier"], dt, settings, rates, True)
            r["state"] = {k: r[k] for k in ("plastic_strain", "equivalent_plastic_strain",
                                            "viscoplastic_multiplier")}
            return r
        r = j2vp.elastic_update(mat, strains, dt, rates, True, False)
        r["state"] = {}
        return r

    def assemble(self, u, u_prev, dt, settings, state):
        eps, rate = hex8.strain_fields(self.elements, self.b, u, u_prev, dt)
        r = self.material_update(eps, rate, dt, settings, state)
        f = hex8.internal_force(self.elements, self.b, self.jw, r["stress"], self.n_dof)
        k = hex8.tangent_matrix(self.elements, self.b, self.jw, r["tangent"], self.n_dof)
        r.update({"strains": eps, "strain_rates": rate, "internal_force": f, "K": k})
        return r


def solve_increment(model: Model, macro_strain, dt, settings, committed, need_tangent=True):
    """``solve_rve_microequilibrium`` (``src/fem/rve/solver.py:62-227``)."""
    macro_strain = np.asarray(macro_strain, dtype=np.float64)
    t = model.constraint.transformation
    a = model.constraint.affine
    ns = model.newton
    u_prev = t @ committed["fluctuation_dofs"] + a @ committed["macro_strain"]
    w = committed["fluctuation_dofs"].copy()
    r0 = None
    last_inc = np.inf
    backtracks = 0
    for it in range(ns.max_iterations + 1):
        u = t @ w + a @ macro_strain
        asm = model.assemble(u, u_prev, dt, settings, committed["material_state"])
        rr = np.asarray(t.T @ asm["internal_force"])
        rn = float(np.linalg.norm(rr))
        if r0 is None:
            r0 = rn
        rtol = max(ns.residual_absolute_tolerance, ns.residual_relative_tolerance * r0)
        itol = max(ns.increment_absolute_tolerance,
                   ns.increment_relative_tolerance * max(float(np.linalg.norm(w)), 1.0))
        if rn <= rtol and (it == 0 or last_inc <= itol):
            return _response(model, macro_strain, dt, committed, w, u, asm, it, backtracks, rn,
                             rtol, need_tangent)
        if it == ns.max_iterations:
            raise OracleRVEError("maximum_iterations_exceeded",
                                 {"iteration": it, "residual_norm": rn,
                                  "residual_tolerance": rtol})
        kred = (t.T @ asm["K"] @ t).tocsc()
        inc = np.asarray(spsolve(kred, -rr), dtype=np.float64)
        if not np.all(np.isfinite(inc)):
            raise OracleRVEError("non_finite_increment", {"iteration": it})
        inc_n = float(np.linalg.norm(inc))
        if rn <= rtol and inc_n <= itol:
            return _response(model, macro_strain, dt, committed, w, u, asm, it, backtracks, rn,
                             rtol, need_tangent)
        if ns.armijo.enabled:
            def evaluate(alpha, w0=w, d=inc):
                uc = t @ (w0 + alpha * d) + a @ macro_strain
                ac = model.assemble(uc, u_prev, dt, settings, committed["material_state"])
                return np.asarray(t.T @ ac["internal_force"])
            try:
                alpha, bt, _ = armijo_search(rr, evaluate, ns.armijo)
            except ArmijoExhausted as err:
                raise OracleRVEError("armijo_exhausted", {"armijo": err.diagnostics,
                                                          "iteration": it}) from err
            w = w + alpha * inc
            last_inc = float(np.linalg.norm(alpha * inc))
            backtracks += bt
        else:
            w = w + inc
            last_inc = float(np.linalg.norm(inc))
    raise AssertionError("unreachable")


def effective_tangent(model: Model, k, volume):
    """``_compute_effective_tangent`` (``solver.py:438-461``)."""
    t = model.constraint.transformation
    a = model.constraint.affine
    ka = np.asarray(k @ a)
    kee = a.T @ ka
    if t.shape[1] == 0:
        return kee / volume
    kww = (t.T @ k @ t).tocsc()
    kwe = np.asarray(t.T @ ka)
    sens = np.asarray(spsolve(kww, kwe)).reshape(kwe.shape)
    kew = np.asarray((a.T @ k) @ t)
    return (kee - kew @ sens) / volume


def _response(model, macro_strain, dt, committed, w, u, asm, iters, backtracks, rn, rtol,
              need_tangent):
    """``_build_rve_response`` (``solver.py:253-364``)."""
    wts = model.weights
    vol = model.volume
    stress = asm["stress"]
    macro_stress = np.sum(stress * wts[:, None], axis=0) / vol
    virtual = model.constraint.affine.T @ asm["internal_force"] / vol
    c_eff = effective_tangent(model, asm["K"], vol) if need_tangent else None
    prev = asm["strains"] - dt * asm["strain_rates"]
    dstrain = asm["strains"] - prev
    micro_work = float(np.sum(np.einsum("ij,ij->i", stress, dstrain) * wts) / vol)
    macro_work = float(macro_stress @ (macro_strain - committed["macro_strain"]))
    hm_res = macro_work - micro_work
    hm_err = abs(hm_res) / (abs(macro_work) + abs(micro_work)
                            + model.hill_mandel_power_scale * dt)
    diag = asm["diagnostics"]
    pw = diag["incremental_work_density"]
    act = diag["viscoplastic_active"]
    margin = diag.get("yield_activation_margin", np.full(wts.size, np.nan))
    fin = margin[np.isfinite(margin)]
    phases = {}
    names = [p[0] for p in model.phases] if model.phased else ["domain"]
    for pid, name in enumerate(names):
        mask = model.point_phase_ids == pid
        pv = float(np.sum(wts[mask]))
        phases[name] = {
            "volume": pv,
            "volume_fraction": pv / vol,
            "average_stress": np.sum(stress[mask] * wts[mask, None], axis=0) / pv,
            "free_energy_density": float(np.sum(asm["free_energy"][mask] * wts[mask]) / pv),
            "dissipation_density": float(np.sum(asm["dissipation"][mask] * wts[mask]) / pv),
        }
    return {
        "macro_stress": macro_stress,
        "effective_tangent": c_eff,
        "state": {"material_state": asm["state"], "macro_strain": macro_strain.copy(),
                  "fluctuation_dofs": w.copy()},
        "displacement": u,
        "micro_strains": asm["strains"],
        "micro_stresses": stress,
        "micro_diagnostics": diag,
        "free_energy_density": float(np.sum(asm["free_energy"] * wts) / vol),
        "dissipation_density": float(np.sum(asm["dissipation"] * wts) / vol),
        "phase_diagnostics": phases,
        "internal_force": asm["internal_force"],
        "diagnostics": {
            "micro_residual_norm": rn,
            "micro_residual_tolerance": rtol,
            "newton_iterations": iters,
            "armijo_backtracks": backtracks,
            "volume": vol,
            "virtual_work_stress_error": float(np.linalg.norm(virtual - macro_stress)),
            "periodic_fluctuation_error": periodic.fluctuation_error(u, macro_strain,
                                                                     model.constraint),
            "reaction_antiperiodicity_error": periodic.face_reaction_error(
                asm["internal_force"], model.constraint),
            "periodic_class_reaction_error": periodic.class_reaction_error(
                asm["internal_force"], model.constraint),
            "macro_work_increment": macro_work,
            "micro_work_increment": micro_work,
            "hill_mandel_residual": hm_res,
            "hill_mandel_error": hm_err,
            "positive_work_density": float(np.sum(np.maximum(pw, 0.0) * wts) / vol),
            "negative_work_density": float(np.sum(np.minimum(pw, 0.0) * wts) / vol),
            "unloading_volume_fraction": float(np.sum(wts[pw < 0.0]) / vol),
            "viscoplastic_active_volume_fraction": float(np.sum(wts[act > 0.5]) / vol),
            "yield_activation_margin_available": int(fin.size > 0),
            "minimum_yield_activation_margin": float(np.min(fin)) if fin.size else np.nan,
            "maximum_yield_activation_margin": float(np.max(fin)) if fin.size else np.nan,
        },
    }


def solve_path(model: Model, times, macro_strains, settings, need_tangent=True):
    """``iterate_prescribed_rve_path`` collected (``src/fem/rve/path.py:47-123``)."""
    times = np.asarray(times, dtype=np.float64)
    dts = np.diff(times)
    state = model.initial_state()
    out = []
    for step, dt in enumerate(dts, start=1):
        r = solve_increment(model, macro_strains[step], float(dt), settings, state, need_tangent)
        state = r["state"]
        out.append(r)
    return out
