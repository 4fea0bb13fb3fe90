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


# Production path design of the reference (``rve_path_generation.py:55-195``):
# canonical engineering-Voigt tensor directions, loading families given as scalar
# control knots along the direction, cubic smoothstep in time on every segment, and a
# duration chosen so the peak strain rate equals the target rate.
TENSOR_WEIGHTS = np.asarray([1.0, 1.0, 1.0, 0.5, 0.5, 0.5])
TENSOR_DIRECTIONS = {
    "deviatoric_uniaxial": (1.0, -0.5, -0.5, 0.0, 0.0, 0.0),
    "deviatoric_biaxial": (0.5, 0.5, -1.0, 0.0, 0.0, 0.0),
    "shear_dominant": (0.1, -0.1, 0.0, 1.0, 0.25, 0.0),
    "positive_triaxial": (1.0, 0.35, 0.35, 0.15, 0.0, 0.0),
    "negative_triaxial": (-1.0, -0.35, -0.35, 0.15, 0.0, 0.0),
}
FAMILY_CONTROLS = {
    "monotonic": (0.0, 1.0),
    "load_unload_reload": (0.0, 1.0, 0.25, 1.2),
    "reverse": (0.0, 1.0, -1.0, 0.5),
    "hold": (0.0, 1.0, 1.0),
}


def tensor_norm(v):
    """Engineering-tensor norm (shear components weighted 1/2)."""
    v = np.asarray(v, dtype=np.float64)
    return np.sqrt(np.sum(v * v * TENSOR_WEIGHTS, axis=-1))


def production_strain_path(family, direction, amplitude, target_rate, n_steps, seed=0,
                           perturbation=0.0, turning_angle=None):
    """One path of the reference's production design: times (n+1), strains (n+1, 6).

    ``direction`` is a name of ``TENSOR_DIRECTIONS`` or a 6-vector; it is normalised
    and perturbed by ``perturbation`` times a seeded unit direction orthogonal to it
    (in the engineering-tensor inner product; drawn first, whatever the magnitude).  ``family`` is one of
    ``FAMILY_CONTROLS`` (scalar knots along the direction) or ``nonproportional``
    (knots 0, d, d + 0.75 r with r the direction turned by ``turning_angle`` towards
    a seeded orthogonal one).  The knots are scaled to unit peak norm, interpolated
    by a cubic smoothstep on each segment in normalised time, and the duration is
    ``amplitude * (peak normalised rate) / target_rate`` with the peak rate of
    the smoothstep, 1.5 * segments * max segment norm."""
    rng = np.random.default_rng(seed)
    d = np.asarray(TENSOR_DIRECTIONS[direction] if isinstance(direction, str) else direction,
                   dtype=np.float64)
    d = d / tensor_norm(d)

    def orthogonal():
        o = rng.normal(size=6)
        o -= np.sum(o * d * TENSOR_WEIGHTS) * d
        return o / tensor_norm(o)

    # the perturbation direction is drawn even at zero magnitude, so the seeded
    # stream (and a nonproportional path's turn) does not depend on it
    d = d + perturbation * orthogonal()
    d = d / tensor_norm(d)
    if family == "nonproportional":
        if turning_angle is None:
            raise ValueError("a nonproportional path needs a turning angle")
        r = np.cos(turning_angle) * d + np.sin(turning_angle) * orthogonal()
        knots = np.stack([np.zeros(6), d, d + 0.75 * r])
    else:
        knots = np.asarray(FAMILY_CONTROLS[family])[:, None] * d[None, :]
    knots = knots / float(np.max(tensor_norm(knots)))
    segments = knots.shape[0] - 1
    peak_rate = 1.5 * segments * float(np.max(tensor_norm(np.diff(knots, axis=0))))
    duration = amplitude * peak_rate / target_rate
    times = np.linspace(0.0, duration, int(n_steps) + 1)
    u = times / duration * segments
    seg = np.minimum(u.astype(np.int64), segments - 1)
    loc = u - seg
    s = loc * loc * (3.0 - 2.0 * loc)
    strains = amplitude * (knots[seg] + s[:, None] * (knots[seg + 1] - knots[seg]))
    return times, strains
