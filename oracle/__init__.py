"""FP64 CPU oracle: a NumPy/SciPy restatement of the reference FEM hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this
package.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker
(or, for ``cpu_baseline``, as the timed CPU port of the reference algorithm).

Each module restates one part of ``zhenhao2049/concrete_impact_ai`` and cites
the reference ``file:line`` it follows:

* ``oracle.j2vp``      -- J2 Perzyna return map, elastic and phased points
                          (``src/fem/materials/plasticity.py``, ``linear_elastic.py``,
                          ``phased.py``)
* ``oracle.hex8``      -- Hex8 shape functions, quadrature, B matrices, strain
                          fields, internal force and tangent assembly, lumped
                          mass (``src/fem/assembly/*``, ``src/fem/elements``,
                          ``src/fem/quadrature``)
* ``oracle.periodic``  -- structured periodic constraint (``src/fem/rve/periodic.py``)
* ``oracle.rve``       -- periodic RVE Newton/Armijo microequilibrium,
                          homogenisation, Schur-complement tangent, path driver
                          (``src/fem/rve/solver.py``, ``path.py``,
                          ``src/fem/solvers/line_search.py``)
* ``oracle.dynamics``  -- explicit velocity-Verlet and Newmark updates
                          (``src/fem/dynamics/material.py``, ``nonlinear.py``)
* ``oracle.mlp``       -- FP64 reference of the material-point surrogate MLP
* ``oracle.laminate``  -- closed-form check of the layered two-phase RVE (per-phase
                          recurrence on the point update) that pins ``oracle.rve``
                          on heterogeneous plastic paths

Pinning status of every function is recorded in each module header.
"""
