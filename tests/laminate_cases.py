"""Laminate RVE cases shared by the oracle and GPU laminate tests.

Two-phase periodic RVEs whose element phases depend only on the x layer, with
the mortar and aggregate of the 48^3 two-phase workload (mortar E 25 GPa,
nu 0.2, sigma_y 20 MPa, H 0.5 GPa, eta 1e-3 s; aggregate E 50 GPa elastic),
and a second J2 phase for the cases that need two plastic phases.  The path is
20 load / unload increments in a random 6-component direction, |eps|_inf up
to 3e-3 at about 100 1/s.
"""

from __future__ import annotations

import numpy as np

from oracle import j2vp as OJ
from oracle import laminate as OL


def mortar(m=1.0):
    return OJ.J2VPParams(25.0e9, 0.2, 20.0e6, 0.5e9, 1.0e-3, 20.0e6, m)


def aggregate():
    return OJ.ElasticParams(50.0e9, 0.2)


def soft(m=1.0):
    """A weaker J2 phase (an interfacial-zone-like layer)."""
    return OJ.J2VPParams(18.0e9, 0.22, 8.0e6, 0.2e9, 2.0e-3, 8.0e6, m)


CASES = {
    # name: (divisions, lengths, layer phases, phase list factory)
    "mortar_aggregate_m1": ((5, 2, 3), (1.0, 0.8, 1.2), [0, 1, 1, 0, 0], lambda: [mortar(1.0), aggregate()]),
    "mortar_aggregate_m2": ((5, 3, 2), (0.9, 1.0, 1.1), [0, 0, 1, 0, 1], lambda: [mortar(2.0), aggregate()]),
    "mortar_soft_m12": ((4, 3, 3), (1.0, 1.0, 1.0), [0, 1, 0, 0], lambda: [mortar(1.0), soft(2.0)]),
}

# Device-only cases whose y and z edges (16, 32, 48, 64) take the single-RVE solver's
# row-lane spectral passes (all three edges for the last: the x passes too); held to
# the laminate recurrence alone (their SuperLU oracle path would take minutes).
LANE_CASES = {
    "lanes_mortar_aggregate_m1_16x48": ((5, 16, 48), (1.0, 0.8, 1.2), [0, 1, 1, 0, 0],
                                        lambda: [mortar(1.0), aggregate()]),
    "lanes_mortar_soft_m12_32x64": ((4, 32, 64), (1.0, 1.0, 1.0), [0, 1, 0, 0],
                                    lambda: [mortar(1.0), soft(2.0)]),
    "lanes_mortar_aggregate_m2_16x16x32": ((16, 16, 32), (1.0, 0.9, 1.1),
                                           [0, 0, 1, 0, 1, 1, 0, 1, 0, 0, 1, 0, 1, 1, 0, 1],
                                           lambda: [mortar(2.0), aggregate()]),
}


def load_unload_path(seed=11, steps=20, peak=3.0e-3, rate=100.0):
    """Times (steps+1) and macro strains (steps+1, 6): loading over 12 steps to
    ``peak`` along a random direction, then unloading to 20 % of it."""
    rng = np.random.default_rng(seed)
    d = rng.normal(size=6)
    d /= np.max(np.abs(d))
    up = 12
    scale = np.concatenate([np.arange(up + 1) / up, 1.0 - 0.8 * np.arange(1, steps - up + 1) / (steps - up)])
    strains = peak * scale[:, None] * d[None, :]
    dt = peak / up / rate
    return dt * np.arange(steps + 1), strains


def layer_setup(name):
    div, lengths, layers, factory = CASES[name] if name in CASES else LANE_CASES[name]
    phases = factory()
    layers = np.asarray(layers)
    fractions = np.asarray([np.count_nonzero(layers == p) for p in range(len(phases))], float) / div[0]
    eph = OL.layer_element_phases(layers, div)
    return div, lengths, phases, fractions, eph


def point_phase_rows(eph):
    """Gauss-point phase ids (point_id = 8 element_id + q)."""
    return np.repeat(eph, 8)
