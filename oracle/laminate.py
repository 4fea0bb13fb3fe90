"""Oracle check: the layered periodic RVE (laminate) by a per-phase recurrence (FP64).

Test infrastructure only (see ``oracle/__init__.py``).

A periodic RVE whose element phases depend only on the element's x index is a
laminate with layers normal to x.  Its exact incremental solution is piecewise
uniform: every Gauss point of phase p carries one strain e_p whose in-plane
components (yy, zz, yz) equal the macro strain and whose out-of-plane
components o = (xx, xz, xy) satisfy

* sum_p f_p e_p[o] = macro[o]  (the fluctuation u(x), piecewise linear in x,
  returns to itself across the periodic cell; f_p = layer fraction of phase p),
* sigma_p[o] equal for all p   (traction continuity across every interface, so
  div sigma = 0 holds in the weak form for every periodic test function).

A trilinear Hex8 element represents a fluctuation that is linear in x
exactly, and the layer interfaces lie on element faces, so this piecewise
uniform field lies in the finite-element space of the structured periodic
RVE: the Galerkin solution of ``solve_rve_microequilibrium``
(``src/fem/rve/solver.py:62-227``) equals it up to the Newton tolerance, for
any path-dependent material, as long as the increment's solution is unique
(hardening H > 0).  The recurrence below solves the 3 (P - 1) equations of each
increment by Newton with the point update's consistent tangent
(``oracle.j2vp``, itself pinned by closed forms); it involves no assembly, no
periodic constraint, no sparse solve and no Gauss-point bookkeeping, so it pins
exactly the parts of the RVE path that the homogeneous closed forms leave
open: the phase and state mapping of the micro points along a heterogeneous
plastic history, the assembled residual and the homogenisation.

``swap_states_after`` plants a mistake for the check's own test: after that
step the committed states of phases 0 and 1 are exchanged.
"""

from __future__ import annotations

import numpy as np

from oracle import j2vp

OUT_OF_PLANE = np.asarray([0, 4, 5])  # xx, xz, xy in the Voigt order xx, yy, zz, yz, xz, xy


def layer_element_phases(layer_phases, divisions) -> np.ndarray:
    """Element phase ids of a structured box (element order i fastest,
    ``src/fem/mesh/structured.py:58-80``) whose phase is ``layer_phases[i]``."""
    nx, ny, nz = divisions
    layer_phases = np.asarray(layer_phases, dtype=np.int64)
    if layer_phases.shape != (nx,):
        raise ValueError("one phase id per x layer is required")
    return np.tile(layer_phases, ny * nz)


def _initial_state(params):
    if isinstance(params, j2vp.J2VPParams):
        return {"plastic_strain": np.zeros((1, 6)), "equivalent_plastic_strain": np.zeros(1),
                "viscoplastic_multiplier": np.zeros(1)}
    return {}


def _update(params, strain, state, dt, settings, rate):
    e = strain[None, :]
    if isinstance(params, j2vp.J2VPParams):
        r = j2vp.j2vp_update(params, e, state["plastic_strain"], state["equivalent_plastic_strain"],
                             state["viscoplastic_multiplier"], dt, settings, rate[None, :], True)
        nxt = {k: r[k] for k in ("plastic_strain", "equivalent_plastic_strain",
                                 "viscoplastic_multiplier")}
        return r["stress"][0], r["tangent"][0], nxt, r
    r = j2vp.elastic_update(params, e, dt, None, True)
    return r["stress"][0], r["tangent"][0], {}, r


def laminate_path(phases, fractions, times, macro_strains, settings=None, tolerance=1e-14,
                  max_iterations=50, swap_states_after=None) -> list[dict]:
    """Per-phase recurrence of a laminate RVE along a prescribed macro path.

    ``phases``: list of ``J2VPParams | ElasticParams`` (phase id = list index);
    ``fractions``: volume fraction of every phase (sums to 1); ``times`` (n+1),
    ``macro_strains`` (n+1, 6) with the first row the committed start (zero
    state).  Returns one dict per increment: ``phase_strain`` (P, 6),
    ``phase_stress`` (P, 6), ``phase_state`` (list of dicts of (1, ...)
    arrays), ``macro_stress`` (6), ``iterations``.
    """
    settings = settings or j2vp.UpdateSettings()
    f = np.asarray(fractions, dtype=np.float64)
    n_ph = len(phases)
    if f.shape != (n_ph,) or abs(f.sum() - 1.0) > 1e-12 or np.any(f <= 0):
        raise ValueError("fractions must be positive and sum to one")
    times = np.asarray(times, dtype=np.float64)
    macro = np.asarray(macro_strains, dtype=np.float64)
    o = OUT_OF_PLANE
    states = [_initial_state(p) for p in phases]
    strain = np.repeat(macro[0][None, :], n_ph, axis=0)
    out = []
    last = n_ph - 1
    for k in range(1, macro.shape[0]):
        dt = float(times[k] - times[k - 1])
        prev = strain.copy()
        # start from the previous phase strains moved by the macro increment
        strain = prev + (macro[k] - macro[k - 1])[None, :]
        x = strain[:, o].copy()

        def evaluate(xv):
            e = np.repeat(macro[k][None, :], n_ph, axis=0)
            xl = (macro[k][o] - np.einsum("p,pi->i", f[:last], xv[:last])) / f[last]
            xs = xv.copy()
            xs[last] = xl
            e[:, o] = xs
            res = [_update(phases[p], e[p], states[p], dt, settings, (e[p] - prev[p]) / dt)
                   for p in range(n_ph)]
            return e, res

        scale = 1.0
        for it in range(max_iterations + 1):
            e, res = evaluate(x)
            sig = np.stack([r[0] for r in res])
            scale = max(scale, float(np.max(np.abs(sig))))
            rvec = np.concatenate([sig[p, o] - sig[last, o] for p in range(last)])
            if np.max(np.abs(rvec)) <= tolerance * scale:
                break
            if it == max_iterations:
                raise RuntimeError(f"laminate recurrence did not converge at step {k}")
            cl = res[last][1][np.ix_(o, o)]
            jac = np.zeros((3 * last, 3 * last))
            for p in range(last):
                cp = res[p][1][np.ix_(o, o)]
                for q in range(last):
                    blk = (f[q] / f[last]) * cl
                    if p == q:
                        blk = blk + cp
                    jac[3 * p:3 * p + 3, 3 * q:3 * q + 3] = blk
            dx = np.linalg.solve(jac, -rvec)
            x = x.copy()
            x[:last] += dx.reshape(last, 3)
        strain = e
        states = [r[2] for r in res]
        if swap_states_after is not None and k == swap_states_after:
            states[0], states[1] = states[1], states[0]
        sig = np.stack([r[0] for r in res])
        out.append({"phase_strain": e.copy(), "phase_stress": sig,
                    "phase_state": [dict(s) for s in states],
                    "macro_stress": np.einsum("p,pi->i", f, sig), "iterations": it})
    return out
