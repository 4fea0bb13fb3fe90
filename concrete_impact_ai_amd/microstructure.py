"""Seeded two-phase voxel microstructures and macro-strain paths.

The reference builds only a cylindrical O-grid inclusion
(``src/fem/rve/builders.py:22-78``); the spherical-aggregate voxel generator
below is new.  It produces ``element_phase_ids`` for ``StructuredHex8RVE``
(``src/fem/rve/data.py:88-102``): 0 = mortar (matrix), 1 = aggregate.

Path generation follows the seeded-knot approach of
``src/concrete_impact/experiments/rve_path_generation.py:34-234``
(``numpy.random.default_rng(seed)``, piecewise-linear segments).
"""

from __future__ import annotations

import numpy as np


def spherical_aggregates(divisions, volume_fraction=0.40, radius_range=(2.0, 6.0), seed=0):
    """Element phase ids of a periodic box filled with spheres (radii in
    element units, centres uniform, periodic minimum-image distance) until the
    aggregate voxel fraction reaches ``volume_fraction``.

    A voxel (element) belongs to a sphere when its centre lies inside it.
    Element order is i fastest (``src/fem/mesh/structured.py:58-80``).
    """
    nx, ny, nz = (int(d) for d in divisions)
    rng = np.random.default_rng(seed)
    phase = np.zeros((nz, ny, nx), dtype=bool)
    zc, yc, xc = np.meshgrid(np.arange(nz) + 0.5, np.arange(ny) + 0.5, np.arange(nx) + 0.5,
                             indexing="ij")
    total = nx * ny * nz
    target = volume_fraction * total
    guard = 0
    while phase.sum() < target:
        guard += 1
        if guard > 100000:
            raise RuntimeError("sphere placement did not reach the requested volume fraction")
        r = rng.uniform(*radius_range)
        c = rng.uniform(0.0, 1.0, size=3) * np.asarray([nx, ny, nz])
        dx = np.abs(xc - c[0])
        dy = np.abs(yc - c[1])
        dz = np.abs(zc - c[2])
        dx = np.minimum(dx, nx - dx)
        dy = np.minimum(dy, ny - dy)
        dz = np.minimum(dz, nz - dz)
        phase |= dx * dx + dy * dy + dz * dz <= r * r
    return phase.ravel().astype(np.int64)


def random_strain_path(n_steps=200, n_segments=8, max_strain=3.0e-3, rate_range=(10.0, 1000.0),
                       seed=0):
    """Random 6-component piecewise-linear macro strain path.

    Knots are uniform in [-max_strain, max_strain] per component; each
    segment has a strain rate drawn log-uniformly from ``rate_range`` and
    its step time is ``||d eps||_inf / rate``.  Returns ``times`` (n+1) and
    ``strains`` (n+1, 6) starting from zero.
    """
    rng = np.random.default_rng(seed)
    knots = np.vstack([np.zeros(6), rng.uniform(-max_strain, max_strain, size=(n_segments, 6))])
    rates = np.exp(rng.uniform(np.log(rate_range[0]), np.log(rate_range[1]), size=n_segments))
    per = np.full(n_segments, n_steps // n_segments)
    per[: n_steps - per.sum()] += 1
    strains = [np.zeros(6)]
    times = [0.0]
    for s in range(n_segments):
        a, b = knots[s], knots[s + 1]
        for k in range(1, per[s] + 1):
            e = a + (b - a) * (k / per[s])
            de = np.max(np.abs(e - strains[-1]))
            strains.append(e)
            times.append(times[-1] + max(de, 1e-12) / rates[s])
    return np.asarray(times), np.asarray(strains)


def uniaxial_ramp(n_steps=100, final_strain=2.0e-3, rate=100.0):
    """C1 load: eps_xx ramped linearly to ``final_strain`` at ``rate`` (1/s)."""
    strains = np.zeros((n_steps + 1, 6))
    strains[:, 0] = final_strain * np.arange(n_steps + 1) / n_steps
    dt = final_strain / n_steps / rate
    return dt * np.arange(n_steps + 1), strains
