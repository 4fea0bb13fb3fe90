"""Oracle: the reference's literal per-point and per-element loops (FP64).

Test infrastructure only (see ``oracle/__init__.py``): used by ``bench.py``'s CPU
baseline to time the reference algorithm in the form the reference runs it, beside
the vectorised restatement of ``oracle/rve.py``.

Restates:
* ``_update_points``                  ``src/fem/materials/plasticity.py:343-431`` -- a
  Python loop over the points, each a one-point update (``j2vp.j2vp_update`` on a
  single row, the reference's ``update_point``)
* ``_assemble_force_and_tangent``     ``src/fem/assembly/nonlinear_solid.py:460-493`` --
  a Python loop over the elements, each summing B^T sigma w and B^T C B w over its
  Gauss points and appending its 24 x 24 block to COO lists

``LoopModel`` is ``rve.Model`` with this assembly, so ``rve.solve_increment`` runs the
reference's Newton / Armijo / SuperLU on it unchanged.  Its results equal the
vectorised model's to round-off (``tests/test_oracle_rve.py``).
"""

from __future__ import annotations

import numpy as np
from scipy.sparse import coo_matrix

from oracle import hex8, j2vp
from oracle.rve import Model


def update_points_loop(params, strains, state, dt, settings, rates):
    """One J2 Perzyna point at a time, in point order."""
    n = strains.shape[0]
    stress = np.zeros((n, 6))
    tangent = np.zeros((n, 6, 6))
    free = np.zeros(n)
    diss = np.zeros(n)
    nxt = {k: np.array(v, copy=True) for k, v in state.items()}
    diagnostics: dict = {}
    for p in range(n):
        r = j2vp.j2vp_update(params, strains[p:p + 1], state["plastic_strain"][p:p + 1],
                             state["equivalent_plastic_strain"][p:p + 1],
                             state["viscoplastic_multiplier"][p:p + 1], dt, settings,
                             None if rates is None else rates[p:p + 1], True)
        stress[p] = r["stress"][0]
        tangent[p] = r["tangent"][0]
        free[p] = r["free_energy"][0]
        diss[p] = r["dissipation"][0]
        nxt["plastic_strain"][p] = r["plastic_strain"][0]
        nxt["equivalent_plastic_strain"][p] = r["equivalent_plastic_strain"][0]
        nxt["viscoplastic_multiplier"][p] = r["viscoplastic_multiplier"][0]
        for key, val in r["diagnostics"].items():  # _append_diagnostics, row by row
            diagnostics.setdefault(key, np.full(n, np.nan))[p] = np.asarray(val).reshape(-1)[0]
    return {"stress": stress, "tangent": tangent, "free_energy": free, "dissipation": diss,
            "diagnostics": diagnostics, "state": nxt, **nxt}


def assemble_loop(elements, b, jw, stress, tangents, n_dof):
    """Element by element: force vector and COO tangent -> CSR."""
    f = np.zeros(n_dof)
    rows: list[int] = []
    cols: list[int] = []
    vals: list[float] = []
    nq = jw.shape[1]
    dofs_all = hex8.element_dofs(elements)
    for e in range(elements.shape[0]):
        dofs = dofs_all[e]
        fe = np.zeros(24)
        ke = np.zeros((24, 24))
        for q in range(nq):
            p = e * nq + q
            bq = b[e, q]
            fe += bq.T @ stress[p] * jw[e, q]
            ke += bq.T @ tangents[p] @ bq * jw[e, q]
        f[dofs] += fe
        r_ids, c_ids = np.meshgrid(dofs, dofs, indexing="ij")
        rows.extend(r_ids.reshape(-1).tolist())
        cols.extend(c_ids.reshape(-1).tolist())
        vals.extend(ke.reshape(-1).tolist())
    k = coo_matrix((vals, (rows, cols)), shape=(n_dof, n_dof)).tocsr()
    return f, k


class LoopModel(Model):
    """A single-phase J2 Perzyna ``rve.Model`` assembled by the literal loops."""

    def assemble(self, u, u_prev, dt, settings, state):
        if self.phased or not isinstance(self.phases[0][1], j2vp.J2VPParams):
            raise ValueError("LoopModel times the single-phase J2 Perzyna RVE")
        eps, rate = hex8.strain_fields(self.elements, self.b, u, u_prev, dt)
        r = update_points_loop(self.phases[0][1], eps, state, dt, settings, rate)
        f, k = assemble_loop(self.elements, self.b, self.jw, r["stress"], r["tangent"], self.n_dof)
        r.update({"strains": eps, "strain_rates": rate, "internal_force": f, "K": k})
        return r
