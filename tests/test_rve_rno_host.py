"""Host side of the device RVE-RNO (no GPU): the artifact-metadata contract of
``load_rve_rno_state_dict`` (``src/concrete_impact/nn/deployment.py:30-65``), the
config checks, and the packed weight layout that ``cia_rno_update`` reads
(``include/cia.h``) against the state dict it was packed from."""

from __future__ import annotations

import ctypes as C
import json

import numpy as np
import pytest

from rve_rno_cases import ARTIFACT, artifact_state, random_state


def _unpack_network(buf, n_in, n_out):
    """Inverse of the documented per-network layout; returns arrays and the size used."""
    H, K = 128, 16
    o = 0

    def take(*shape):
        nonlocal o
        size = int(np.prod(shape))
        a = buf[o:o + size].reshape(shape)
        o += size
        return a
    w1p, b1 = take(H, K), take(H)
    hidden = [(take(H, H), take(H)) for _ in range(3)]
    transposes = [take(H, H) for _ in range(3)]
    w1t = take(K, H)
    rows = 16 * ((n_out + 15) // 16)
    wo, bo = take(rows, H), take(rows)
    assert np.all(w1p[:, n_in:] == 0) and np.array_equal(w1t, w1p.T)
    for (w, _), wt in zip(hidden, transposes):
        assert np.array_equal(wt, w.T)
    assert np.all(wo[n_out:] == 0) and np.all(bo[n_out:] == 0)
    return (w1p[:, :n_in], b1, hidden, wo[:n_out], bo[:n_out]), o


@pytest.mark.parametrize("evolution,latent", [("direct_rate", 6), ("mobility_gradient", 6),
                                              ("mobility_gradient", 3)])
def test_packed_layout_matches_state_dict(evolution, latent):
    from concrete_impact_ai_amd import _lib
    from concrete_impact_ai_amd import rve_rno as RR

    if evolution == "direct_rate" and latent == 6:
        state, _ = artifact_state()
    else:
        state = random_state(evolution, latent, seed=3)
    cfg = RR.RVERNOModelConfig(evolution=evolution, latent_dimension=latent)
    buf = RR.pack_state_dict(cfg, state)
    size = C.c_int64()
    _lib.check(_lib.lib().cia_rno_packed_size(latent, RR.EVOLUTIONS[evolution], C.byref(size)),
               "cia_rno_packed_size")
    assert buf.size == size.value and buf.dtype == np.float32
    n_evo = latent if evolution == "direct_rate" else latent * (latent + 1) // 2
    evo = "state_evolution.network" if evolution == "direct_rate" else "state_evolution.network.network"
    o = 0
    for prefix, n_in, n_out in (("equilibrium_energy.network.network", 6, 1),
                                ("nonequilibrium_energy.network.network", 6 + latent, 1),
                                (evo, 6 + latent, n_evo)):
        (w1, b1, hidden, wo, bo), used = _unpack_network(buf[o:], n_in, n_out)
        o += used
        np.testing.assert_array_equal(w1, state[f"{prefix}.0.weight"])
        np.testing.assert_array_equal(b1, state[f"{prefix}.0.bias"])
        for j, (w, b) in enumerate(hidden):
            np.testing.assert_array_equal(w, state[f"{prefix}.{2 * j + 2}.weight"])
            np.testing.assert_array_equal(b, state[f"{prefix}.{2 * j + 2}.bias"])
        np.testing.assert_array_equal(wo, state[f"{prefix}.8.weight"])
        if f"{prefix}.8.bias" in state:
            np.testing.assert_array_equal(bo, state[f"{prefix}.8.bias"])
        else:
            assert np.all(bo == 0)
    assert o == buf.size


def test_config_rejects_what_the_kernel_does_not_compute():
    from concrete_impact_ai_amd import rve_rno as RR

    with pytest.raises(ValueError, match="width"):
        RR.RVERNOModelConfig(hidden_size=64)
    with pytest.raises(ValueError, match="float32"):
        RR.RVERNOModelConfig(dtype="float64")
    with pytest.raises(ValueError, match="six positive"):
        RR.RVERNOModelConfig(strain_input_scale=(1.0, 1.0))
    with pytest.raises(ValueError, match="integrator"):
        RR.RVERNOModelConfig(integrator="implicit")
    with pytest.raises(ValueError, match="latent"):
        RR.RVERNOModelConfig(latent_dimension=9)


def test_metadata_contract_is_strict(tmp_path):
    from concrete_impact_ai_amd import rve_rno as RR

    _, meta = artifact_state()
    good = tmp_path / "artifact_metadata.json"
    good.write_text(json.dumps(meta))
    for change in ({"latent_dimension": 5}, {"integrator": "cs_semi_implicit_forward_euler"},
                   {"reference_time_scale": 0.01}):
        with pytest.raises(ValueError, match="does not match"):
            RR.load_rve_rno_state_dict(RR.RVERNOModelConfig(**change), ARTIFACT, good)
    bad = dict(meta, deployment_integrator="implicit")
    path = tmp_path / "bad.json"
    path.write_text(json.dumps(bad))
    with pytest.raises(ValueError, match="does not match"):
        RR.load_rve_rno_state_dict(RR.RVERNOModelConfig(), ARTIFACT, path)
