"""Oracle: J2 Perzyna viscoplastic, linear-elastic and phased material points (FP64).

Test infrastructure only (see ``oracle/__init__.py``).

Restates, vectorised over points but with the reference's per-point formulas,
branch logic and tolerance semantics:

* Voigt helpers             ``src/fem/materials/plasticity.py:30-37,1408-1475``
* batch driver              ``plasticity.py:343-431`` (``_update_points``)
* point update              ``plasticity.py:526-681`` (``_update_j2_viscoplastic_point``)
* increment diagnostics     ``plasticity.py:684-708``
* multiplier solve          ``plasticity.py:711-854``
* local modulus / tangent   ``plasticity.py:907-963``
* payload / energies        ``plasticity.py:1300-1328``
* diagnostics lazy fill     ``plasticity.py:1684-1694`` (a key first seen on a
                            later point reads 0.0 on earlier points)
* linear elastic points     ``src/fem/materials/linear_elastic.py:74-100,139-159``
* phased dispatch           ``src/fem/materials/phased.py:67-143,171-199``

Pinning (tests/test_oracle_*.py):
* m = 1 closed form         -- pure-shear recurrence of
                               ``src/concrete_impact/benchmarks/plasticity.py:642-679``
                               at 1e-10 rel over the 121-step reference config, and
                               the uniaxial-strain scalar recurrence.
* m != 1 Newton             -- bracketed bisection of r(dg) = 0 in long double.
* tangent                   -- central finite differences (rtol 1e-5, atol 1e-6 as
                               ``tests/unit/fem/materials/test_plasticity.py:179-256``),
                               symmetry, reduction to C on the elastic branch.
* energies                  -- psi = 1/2 sigma.(eps - eps_p) + 1/2 H q^2 and
                               D = plastic power - hardening storage (1e-12).
* activation / failures     -- reference test cases for activation threshold and
                               reason strings (test_plasticity.py:89-138,259-345).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

VOIGT = 6
IDENTITY = np.asarray([1.0, 1.0, 1.0, 0.0, 0.0, 0.0])
STRESS_WEIGHTS = np.asarray([1.0, 1.0, 1.0, 2.0, 2.0, 2.0])
T2E = np.diag([1.0, 1.0, 1.0, 2.0, 2.0, 2.0])
E2T = np.diag([1.0, 1.0, 1.0, 0.5, 0.5, 0.5])
P_VOL = np.outer(IDENTITY, IDENTITY) / 3.0
P_DEV = np.eye(VOIGT) - P_VOL


class OracleMaterialError(RuntimeError):
    """Mirror of ``MaterialPointConvergenceError`` (``src/fem/materials/errors.py:15-34``)."""

    def __init__(self, reason: str, point_id: int, diagnostics: dict | None = None):
        super().__init__(f"{reason} at point {point_id}")
        self.reason = reason
        self.point_id = point_id
        self.diagnostics = dict(diagnostics or {})
        self.diagnostics.setdefault("point_id", point_id)


@dataclass(frozen=True)
class J2VPParams:
    """``J2ViscoplasticMaterial`` fields (``plasticity.py:79-91``)."""

    young_modulus: float
    poisson_ratio: float
    yield_stress: float
    hardening_modulus: float
    time_scale: float
    reference_stress: float
    rate_exponent: float
    density: float = 1.0


@dataclass(frozen=True)
class ElasticParams:
    """``LinearElasticMaterial`` fields (``linear_elastic.py:25-32``)."""

    young_modulus: float
    poisson_ratio: float
    density: float = 1.0


@dataclass(frozen=True)
class UpdateSettings:
    """``MaterialUpdateSettings`` (``src/fem/materials/data.py:16-24``)."""

    max_iterations: int = 50
    yield_relative_tolerance: float = 1.0e-10
    residual_absolute_tolerance: float = 1.0e-12
    residual_relative_tolerance: float = 1.0e-10


def shear_modulus(e: float, nu: float) -> float:
    return e / (2.0 * (1.0 + nu))


def bulk_modulus(e: float, nu: float) -> float:
    return e / (3.0 * (1.0 - 2.0 * nu))


def elasticity_matrix(e: float, nu: float) -> np.ndarray:
    """``_elasticity_matrix`` (``plasticity.py:1408-1423``)."""
    g = shear_modulus(e, nu)
    lam = e * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    c = np.zeros((VOIGT, VOIGT))
    c[:3, :3] = lam
    c[0, 0] += 2.0 * g
    c[1, 1] += 2.0 * g
    c[2, 2] += 2.0 * g
    c[3, 3] = g
    c[4, 4] = g
    c[5, 5] = g
    return c


def deviatoric(stress: np.ndarray) -> np.ndarray:
    """``_deviatoric`` row-wise (``plasticity.py:1473-1475``)."""
    return stress - (np.sum(stress[:, :3], axis=1) / 3.0)[:, None] * IDENTITY[None, :]


def j2_radicand(dev: np.ndarray) -> np.ndarray:
    """``1.5 * _stress_inner(s, s)`` row-wise (``plasticity.py:1436-1445``)."""
    return 1.5 * np.sum(STRESS_WEIGHTS[None, :] * dev * dev, axis=1)


def _first_failure(status: np.ndarray, reasons: np.ndarray) -> tuple[int, str] | None:
    bad = np.flatnonzero(status)
    if bad.size == 0:
        return None
    p = int(bad[0])
    return p, str(reasons[p])


def j2vp_update(
    params: J2VPParams,
    strains: np.ndarray,
    plastic_strain: np.ndarray,
    equivalent_plastic_strain: np.ndarray,
    viscoplastic_multiplier: np.ndarray,
    time_step: float,
    settings: UpdateSettings,
    strain_rates: np.ndarray | None = None,
    need_tangent: bool = True,
) -> dict:
    """Vectorised restatement of ``_update_points`` for ``J2ViscoplasticMaterial``.

    Inputs are never mutated (``test_plasticity.py:472-494``).  Returns a dict
    with ``stress``, ``tangent``, ``free_energy``, ``dissipation``, next state
    arrays and a ``diagnostics`` dict keyed like the reference.  On failure,
    raises ``OracleMaterialError`` for the first failing point in point order.
    """
    strains = np.asarray(strains, dtype=np.float64)
    n = strains.shape[0]
    e, nu = params.young_modulus, params.poisson_ratio
    c = elasticity_matrix(e, nu)
    g = shear_modulus(e, nu)
    k = bulk_modulus(e, nu)
    h_mod = params.hardening_modulus
    sy, sref, tau, m = (
        params.yield_stress,
        params.reference_stress,
        params.time_scale,
        params.rate_exponent,
    )
    dt = float(time_step)
    eps_p = np.asarray(plastic_strain, dtype=np.float64)
    qbar = np.asarray(equivalent_plastic_strain, dtype=np.float64)
    gam = np.asarray(viscoplastic_multiplier, dtype=np.float64)

    status = np.zeros(n, dtype=bool)
    reasons = np.empty(n, dtype=object)
    fail_diag: dict[int, dict] = {}

    def fail(mask: np.ndarray, reason: str) -> None:
        new = mask & ~status
        reasons[new] = reason
        status[new] = True

    # plasticity.py:537-541 -- trial state
    trial = (strains - eps_p) @ c.T
    tdev = deviatoric(trial)
    rad = j2_radicand(tdev)
    fail(~np.isfinite(rad), "non_finite_j2_invariant")
    fail(np.isfinite(rad) & (rad < 0.0), "negative_j2_invariant")
    with np.errstate(invalid="ignore"):
        q_trial = np.sqrt(rad)
    radius = sy + h_mod * qbar
    f_trial = q_trial - radius
    yield_scale = np.maximum(np.maximum(sy, sref), q_trial)
    # np.maximum propagates NaN just like max() would not; NaN points already failed.
    tol_rel = settings.yield_relative_tolerance
    threshold = tol_rel * yield_scale
    with np.errstate(invalid="ignore"):
        margin = (f_trial - threshold) / yield_scale
    if not np.isfinite(tol_rel) or tol_rel < 0.0:
        fail(np.ones(n, dtype=bool), "invalid_yield_activation_tolerance")
    fail(~np.isfinite(threshold), "invalid_yield_activation_tolerance")

    with np.errstate(invalid="ignore"):
        elastic = f_trial <= threshold
    plastic = ~elastic & ~status

    # multiplier solve, plasticity.py:711-854
    hs = 3.0 * g + h_mod
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        upper = f_trial / hs
        tol_newton = np.maximum(
            settings.residual_absolute_tolerance,
            settings.residual_relative_tolerance * upper,
        )
    dgam = np.zeros(n)
    iters = np.zeros(n, dtype=np.int64)
    local_res = np.zeros(n)
    if m == 1.0:
        with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
            den = tau * sref / dt + hs
            mult = f_trial / den
            corr = f_trial - hs * mult
            res = mult - dt * corr / (tau * sref)
            der = 1.0 + dt * hs / (tau * sref)
        finite = (
            np.isfinite(den) & np.isfinite(mult) & np.isfinite(corr) & np.isfinite(res)
            & np.isfinite(der)
        )
        fail(plastic & ~finite, "non_finite_closed_form")
        oob = ~((mult > 0.0) & (mult < upper)) | (corr <= 0.0)
        fail(plastic & finite & oob, "closed_form_solution_out_of_bounds")
        dgam = np.where(plastic, mult, 0.0)
        local_res = np.where(plastic, res, 0.0)
    else:
        if settings.max_iterations <= 0:
            fail(plastic, "invalid_iteration_limit")
        fail(plastic & ~(np.isfinite(upper) & np.isfinite(tol_newton)), "non_finite_problem_scale")
        active = plastic & ~status
        x = 0.5 * upper
        for it in range(1, max(settings.max_iterations, 0) + 1):
            if not np.any(active):
                break
            with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
                over = f_trial - hs * x
                nover = over / sref
            bad = active & (~np.isfinite(nover) | (nover <= 0.0))
            fail(bad, "invalid_corrected_overstress")
            for p in np.flatnonzero(bad):
                fail_diag.setdefault(int(p), {"iteration": it, "plastic_multiplier": float(x[p]),
                                              "corrected_overstress": float(over[p])})
            active &= ~bad
            with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
                rate = np.power(np.where(active, nover, 1.0), m) / tau
                res = x - dt * rate
                der = 1.0 + dt * m * hs * np.power(np.where(active, nover, 1.0), m - 1.0) / (
                    tau * sref
                )
            bad = active & ~(np.isfinite(rate) & np.isfinite(res) & np.isfinite(der))
            fail(bad, "non_finite_iteration_value")
            for p in np.flatnonzero(bad):
                fail_diag.setdefault(int(p), {"iteration": it, "residual": float(res[p]),
                                              "plastic_multiplier": float(x[p])})
            active &= ~bad
            bad = active & (der <= 0.0)
            fail(bad, "non_positive_derivative")
            active &= ~bad
            with np.errstate(invalid="ignore"):
                conv = active & (np.abs(res) <= tol_newton)
            oob = conv & (~((x > 0.0) & (x < upper)) | (over <= 0.0))
            fail(oob, "converged_solution_out_of_bounds")
            good = conv & ~oob
            dgam = np.where(good, x, dgam)
            iters = np.where(good, it, iters)
            local_res = np.where(good, res, local_res)
            active &= ~conv
            with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
                xn = x - res / der
            bad = active & ~np.isfinite(xn)
            fail(bad, "non_finite_newton_candidate")
            active &= ~bad
            bad = active & ~((xn > 0.0) & (xn < upper))
            fail(bad, "newton_candidate_out_of_bounds")
            active &= ~bad
            if it == settings.max_iterations:
                fail(active, "maximum_iterations_exceeded")
                for p in np.flatnonzero(active):
                    fail_diag.setdefault(int(p), {"iteration": it, "newton_candidate": float(xn[p]),
                                                  "physical_upper_bound": float(upper[p]),
                                                  "plastic_multiplier": float(x[p])})
                active[:] = False
            x = np.where(active, xn, x)

    plastic = plastic & ~status
    with np.errstate(divide="ignore", invalid="ignore"):
        flow = np.where(plastic[:, None], 1.5 * tdev / q_trial[:, None], 0.0)
    dg = np.where(plastic, dgam, 0.0)
    stress = trial - 2.0 * g * dg[:, None] * flow
    dep = dg[:, None] * (flow @ T2E.T)
    eps_p_new = eps_p + dep
    qbar_new = qbar + dg
    gam_new = gam + dg

    fdev = deviatoric(stress)
    frad = j2_radicand(fdev)
    q_final = np.sqrt(np.maximum(frad, 0.0))
    final_yield = q_final - (sy + h_mod * qbar_new)
    plastic_work = np.sum(stress * dep, axis=1)

    # hv (plasticity.py:907-926)
    if m == 1.0:
        hv = np.full(n, tau * sref / dt + hs)
    else:
        bad_hv = plastic & (final_yield <= 0.0)
        fail(bad_hv, "non_positive_tangent_overstress")
        with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
            fder = m * np.power(np.maximum(final_yield, 0.0) / sref, m - 1.0) / (tau * sref)
            hv = 1.0 / (dt * fder) + hs

    first = _first_failure(status, reasons)

    # increment diagnostics (plasticity.py:684-708) -- may also raise
    incr_work = eq_incr = None
    if strain_rates is not None:
        rates = np.asarray(strain_rates, dtype=np.float64)
        dstrain = dt * rates
        prev_stress = (strains - dstrain - eps_p) @ c.T
        prad = j2_radicand(deviatoric(prev_stress))
        bad = ~np.isfinite(prad) & ~status
        fail(bad, "non_finite_j2_invariant")
        first = _first_failure(status, reasons)
        with np.errstate(invalid="ignore"):
            incr_work = 0.5 * np.sum((prev_stress + stress) * dstrain, axis=1)
            eq_incr = q_final - np.sqrt(np.maximum(prad, 0.0))

    if first is not None:
        p, reason = first
        raise OracleMaterialError(reason, p, fail_diag.get(p, {}))

    # free energy (plasticity.py:1312-1315 plus 1/2 H q^2)
    ee = strains - eps_p_new
    free_energy = 0.5 * np.sum((ee @ c.T) * ee, axis=1) + 0.5 * h_mod * qbar_new**2
    with np.errstate(divide="ignore", invalid="ignore"):
        plastic_power = np.where(plastic, plastic_work / dt, 0.0)
        hard_rate = np.where(plastic, h_mod * qbar_new * dg / dt, 0.0)
    dissipation = np.where(plastic, plastic_power - hard_rate, 0.0)

    tangent = None
    if need_tangent:
        tangent = np.broadcast_to(c, (n, VOIGT, VOIGT)).copy()
        if np.any(plastic):
            idx = np.flatnonzero(plastic)
            vol = 3.0 * k * (E2T @ P_VOL)
            devf = 2.0 * g - 6.0 * g**2 * dg[idx] / q_trial[idx]
            rad_c = 4.0 * g**2 * (dg[idx] / q_trial[idx] - 1.0 / hv[idx])
            nn = flow[idx, :, None] * flow[idx, None, :]
            tangent[idx] = (
                vol[None, :, :]
                + devf[:, None, None] * (E2T @ P_DEV)[None, :, :]
                + rad_c[:, None, None] * nn
            )

    diag = {
        "yield_value": np.where(plastic, final_yield, f_trial),
        "plastic_work": np.where(plastic, plastic_work, 0.0),
        "local_iterations": iters.astype(np.float64),
        "active_surface": np.where(plastic, 4.0, 0.0),
        "plastic_multiplier": dg.copy(),
        "viscoplastic_multiplier": dg.copy(),
        "equivalent_stress": np.where(plastic, q_final, q_trial),
        "trial_yield_value": f_trial,
        "yield_scale": yield_scale,
        "yield_activation_threshold": threshold,
        "yield_activation_margin": margin,
        "viscoplastic_active": np.where(plastic, 1.0, 0.0),
        "time_step": np.full(n, dt),
        "equivalent_plastic_strain": qbar_new.copy(),
    }
    if np.any(plastic):
        diag["viscoplastic_hv"] = np.where(plastic, hv, 0.0)
    diag["plastic_power"] = plastic_power
    diag["hardening_storage_rate"] = hard_rate
    diag["local_residual"] = np.where(plastic, local_res, 0.0)
    diag["root_lower_bound"] = np.zeros(n)
    diag["root_upper_bound"] = np.where(plastic, upper, 0.0)
    if incr_work is not None:
        diag["incremental_work_density"] = incr_work
        diag["equivalent_stress_increment"] = eq_incr

    return {
        "stress": stress,
        "tangent": tangent,
        "free_energy": free_energy,
        "dissipation": dissipation,
        "plastic_strain": eps_p_new,
        "equivalent_plastic_strain": qbar_new,
        "viscoplastic_multiplier": gam_new,
        "diagnostics": diag,
        "flow": flow,
        "q_trial": q_trial,
        "plastic_mask": plastic,
        "hv": hv,
    }


def elastic_update(
    params: ElasticParams,
    strains: np.ndarray,
    time_step: float,
    strain_rates: np.ndarray | None,
    need_tangent: bool = True,
    phased_diagnostics: bool = False,
) -> dict:
    """``evaluate_linear_elastic_points`` (``linear_elastic.py:74-100``).

    With ``phased_diagnostics`` the elastic-phase increment diagnostics of
    ``phased.py:171-190`` are added (they need the tangent and strain rates).
    """
    c = elasticity_matrix(params.young_modulus, params.poisson_ratio)
    strains = np.asarray(strains, dtype=np.float64)
    stress = strains @ c.T
    n = strains.shape[0]
    out = {
        "stress": stress,
        "tangent": np.broadcast_to(c, (n, VOIGT, VOIGT)).copy() if need_tangent else None,
        "free_energy": 0.5 * np.einsum("qi,qi->q", strains, stress),
        "dissipation": np.zeros(n),
        "diagnostics": {},
    }
    if phased_diagnostics:
        if strain_rates is None or not need_tangent:
            raise ValueError("Elastic RVE phase diagnostics require strain rates and tangents.")
        dstrain = time_step * np.asarray(strain_rates, dtype=np.float64)
        dstress = np.einsum("qij,qj->qi", out["tangent"], dstrain)
        prev = stress - dstress
        out["diagnostics"] = {
            "incremental_work_density": 0.5 * np.einsum("qi,qi->q", prev + stress, dstrain),
            "equivalent_stress_increment": _eq_rows(stress) - _eq_rows(prev),
            "viscoplastic_active": np.zeros(n),
        }
    return out


def _eq_rows(stress: np.ndarray) -> np.ndarray:
    """``_equivalent_stress`` (``phased.py:193-199``)."""
    dev = stress.copy()
    p = np.mean(stress[:, :3], axis=1)
    dev[:, :3] -= p[:, None]
    return np.sqrt(1.5 * np.einsum("qi,i,qi->q", dev, STRESS_WEIGHTS, dev))


def phased_update(
    phases: list[tuple[str, object]],
    point_phase_ids: np.ndarray,
    strains: np.ndarray,
    state: dict[str, np.ndarray],
    time_step: float,
    settings: UpdateSettings,
    strain_rates: np.ndarray | None,
    need_tangent: bool = True,
) -> dict:
    """``PhasedMaterialModel.update`` (``phased.py:67-143``).

    ``state`` holds phase-compact arrays keyed ``phase_<name>__<var>``.  The
    combined diagnostics are the sorted union of phase keys, NaN where a phase
    does not define a key (``phased.py:125-134``).
    """
    n = strains.shape[0]
    stress = np.zeros((n, VOIGT))
    tangent = np.zeros((n, VOIGT, VOIGT)) if need_tangent else None
    free_energy = np.zeros(n)
    dissipation = np.zeros(n)
    next_state: dict[str, np.ndarray] = {}
    parts = []
    for pid, (name, mat) in enumerate(phases):
        ids = np.flatnonzero(point_phase_ids == pid)
        rates = None if strain_rates is None else strain_rates[ids]
        if isinstance(mat, J2VPParams):
            pre = f"phase_{name}__"
            try:
                r = j2vp_update(
                    mat, strains[ids], state[pre + "plastic_strain"],
                    state[pre + "equivalent_plastic_strain"],
                    state[pre + "viscoplastic_multiplier"], time_step, settings, rates,
                    need_tangent,
                )
            except OracleMaterialError as err:
                raise OracleMaterialError(err.reason, int(ids[err.point_id]), err.diagnostics)
            for var in ("plastic_strain", "equivalent_plastic_strain", "viscoplastic_multiplier"):
                next_state[pre + var] = r[var]
        else:
            r = elastic_update(mat, strains[ids], time_step, rates, need_tangent, True)
        stress[ids] = r["stress"]
        if need_tangent:
            tangent[ids] = r["tangent"]
        free_energy[ids] = r["free_energy"]
        dissipation[ids] = r["dissipation"]
        parts.append((ids, r["diagnostics"]))
    names = sorted(set().union(*(d.keys() for _, d in parts)))
    diag = {nm: np.full(n, np.nan) for nm in names}
    for ids, d in parts:
        for nm, v in d.items():
            diag[nm][ids] = v
    return {
        "stress": stress,
        "tangent": tangent,
        "free_energy": free_energy,
        "dissipation": dissipation,
        "state": next_state,
        "diagnostics": diag,
    }


def pure_shear_recurrence(params: J2VPParams, shear_strains: np.ndarray, time_step: float) -> dict:
    """Closed-form linear-Perzyna pure-shear recurrence.

    Restates ``_evaluate_j2_viscoplastic_pure_shear_exact``
    (``src/concrete_impact/benchmarks/plasticity.py:642-679``); m = 1 only.
    """
    g = shear_modulus(params.young_modulus, params.poisson_ratio)
    gp, qb = 0.0, 0.0
    out = {"shear_stress": [], "plastic_multiplier": [], "equivalent_plastic_strain": []}
    for gam in shear_strains:
        trial = g * (gam - gp)
        sign = 1.0 if trial >= 0.0 else -1.0
        tq = np.sqrt(3.0) * abs(trial)
        fy = tq - (params.yield_stress + params.hardening_modulus * qb)
        if fy <= 0.0:
            dg, tau_s = 0.0, trial
        else:
            den = (params.time_scale * params.reference_stress / time_step
                   + 3.0 * g + params.hardening_modulus)
            dg = fy / den
            tau_s = trial - sign * np.sqrt(3.0) * g * dg
            gp += sign * np.sqrt(3.0) * dg
            qb += dg
        out["shear_stress"].append(tau_s)
        out["plastic_multiplier"].append(dg)
        out["equivalent_plastic_strain"].append(qb)
    return {key: np.asarray(v) for key, v in out.items()}
