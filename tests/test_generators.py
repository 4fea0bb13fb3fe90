"""CPU tests of the microstructure and macro-path generators.

* ``spherical_aggregates``: deterministic per seed, reaches the requested
  voxel fraction without overshooting by more than the voxel volume of one
  sphere, wraps periodically; 48^3 at seed 2 (the benchmark's default seed)
  gives the two-phase workload's 40.2 % aggregate fraction.
* ``random_strain_path``: deterministic per seed, |eps|_inf <= 3e-3 and every
  increment's rate |d eps|_inf / dt within [10, 1000] 1/s (the bounds the
  48^3 and batched workloads state).
* ``production_strain_path`` is pinned in ``tests/test_oracle_mlp.py``.
"""

import numpy as np
import pytest

from concrete_impact_ai_amd import microstructure as MS


@pytest.mark.parametrize("div,seed", [((16, 16, 16), 1), ((20, 17, 23), 2), ((32, 24, 16), 3)])
def test_spherical_aggregates_fraction_determinism_and_bound(div, seed):
    a = MS.spherical_aggregates(div, 0.40, (2.0, 6.0), seed)
    b = MS.spherical_aggregates(div, 0.40, (2.0, 6.0), seed)
    c = MS.spherical_aggregates(div, 0.40, (2.0, 6.0), seed + 100)
    assert a.dtype == np.int64 and a.shape == (div[0] * div[1] * div[2],)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c)
    assert set(np.unique(a)) <= {0, 1}
    total = a.size
    frac = a.sum() / total
    # a voxel is inside when its centre is, so one sphere of radius r <= 6 covers
    # at most the voxels whose centres lie within r: fewer than 4/3 pi (r + sqrt(3)/2)^3
    one_sphere = 4.0 / 3.0 * np.pi * (6.0 + np.sqrt(3.0) / 2.0) ** 3
    assert 0.40 <= frac < 0.40 + one_sphere / total


def test_spherical_aggregates_wrap_periodically():
    """Spheres are placed with the minimum-image distance: the phase field of a
    box is the same as the field generated on the box shifted by whole periods,
    i.e. aggregates cut by a face reappear on the opposite face; on a small box
    both faces of every axis carry aggregate voxels."""
    div = (12, 12, 12)
    ph = MS.spherical_aggregates(div, 0.40, (2.0, 6.0), 5).reshape(div[2], div[1], div[0])
    for axis in range(3):
        lo = np.take(ph, 0, axis=axis)
        hi = np.take(ph, -1, axis=axis)
        assert lo.sum() > 0 and hi.sum() > 0
        # aggregate voxels of one face have aggregate neighbours across the periodic face
        assert np.sum(lo & hi) > 0


def test_two_phase_workload_fraction():
    ph = MS.spherical_aggregates((48, 48, 48), 0.40, (2.0, 6.0), 2)
    assert ph.sum() / ph.size == pytest.approx(0.40158, abs=1e-5)


@pytest.mark.parametrize("seed", [0, 1, 7, 123])
def test_random_strain_path_bounds_and_determinism(seed):
    t, e = MS.random_strain_path(200, 8, 3.0e-3, (10.0, 1000.0), seed)
    t2, e2 = MS.random_strain_path(200, 8, 3.0e-3, (10.0, 1000.0), seed)
    assert np.array_equal(t, t2) and np.array_equal(e, e2)
    assert t.shape == (201,) and e.shape == (201, 6)
    assert np.all(e[0] == 0.0) and t[0] == 0.0
    assert np.max(np.abs(e)) <= 3.0e-3
    dt = np.diff(t)
    assert np.all(dt > 0.0)
    rate = np.max(np.abs(np.diff(e, axis=0)), axis=1) / dt
    assert np.all(rate >= 10.0 * (1 - 1e-12)) and np.all(rate <= 1000.0 * (1 + 1e-12))
    # distinct seeds give distinct paths; all six components move
    t3, e3 = MS.random_strain_path(200, 8, 3.0e-3, (10.0, 1000.0), seed + 1)
    assert not np.array_equal(e, e3)
    assert np.all(np.max(np.abs(e), axis=0) > 0.0)


def test_uniaxial_ramp_is_the_plain_method_load():
    t, e = MS.uniaxial_ramp(100, 2.0e-3, 100.0)
    assert e[-1, 0] == pytest.approx(2.0e-3) and np.all(e[:, 1:] == 0.0)
    np.testing.assert_allclose(np.diff(t), 2.0e-7, rtol=1e-12)
