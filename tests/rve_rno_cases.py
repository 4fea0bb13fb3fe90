"""Shared cases of the RVE-RNO tests: torch modules under the reference's parameter
names (``src/concrete_impact/nn/models/rve_rno.py:63-186``), the artifact fixture
and seeded random networks of the artifact's architecture."""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import torch
from torch import nn

GOLDEN = Path(__file__).resolve().parent / "golden"
ARTIFACT = GOLDEN / "rve_rno_c48_v2_liu_direct.npz"
SAMPLE = GOLDEN / "rve_rno_c48_sample_response.h5"


class _Dense(nn.Module):
    def __init__(self, n_in, n_out, width, layers, bias=True):
        super().__init__()
        mods, w = [], n_in
        for _ in range(layers):
            mods += [nn.Linear(w, width), nn.SiLU()]
            w = width
        mods.append(nn.Linear(w, n_out, bias=bias))
        self.network = nn.Sequential(*mods)

    def forward(self, x):
        return self.network(x)


class _Wrap(nn.Module):
    def __init__(self, inner):
        super().__init__()
        self.network = inner

    def forward(self, x):
        return self.network(x)


class Tree(nn.Module):
    """Torch modules under the reference's parameter names (``rve_rno.py:167-186``)."""

    def __init__(self, latent, width, layers, evolution):
        super().__init__()
        self.equilibrium_energy = _Wrap(_Dense(6, 1, width, layers, bias=False))
        self.nonequilibrium_energy = _Wrap(_Dense(6 + latent, 1, width, layers, bias=False))
        if evolution == "direct_rate":
            self.state_evolution = _Dense(6 + latent, latent, width, layers)
        else:
            self.state_evolution = _Wrap(
                _Dense(6 + latent, latent * (latent + 1) // 2, width, layers))


def artifact_state():
    with np.load(ARTIFACT, allow_pickle=False) as z:
        return {k: z[k] for k in z.files if k != "metadata"}, json.loads(str(z["metadata"]))


def artifact_tree():
    state, _ = artifact_state()
    tree = Tree(6, 128, 4, "direct_rate").double()
    tree.load_state_dict({k: torch.tensor(v, dtype=torch.float64) for k, v in state.items()},
                         strict=True)
    return tree


def random_state(evolution: str, latent: int, seed: int, width: int = 128, layers: int = 4,
                 gain: float = 1.0) -> dict:
    """A seeded network set (Xavier-uniform weights times ``gain``, small random biases)
    as float32 arrays under the reference's names."""
    torch.manual_seed(seed)
    tree = Tree(latent, width, layers, evolution)
    with torch.no_grad():
        for name, p in tree.named_parameters():
            if p.dim() == 2:
                nn.init.xavier_uniform_(p, gain=gain)
            else:
                p.uniform_(-0.1, 0.1)
    return {k: v.detach().numpy().astype(np.float32) for k, v in tree.state_dict().items()}


def artifact_inputs(n: int, seed: int):
    """Strains at twice the artifact's training strain scale, latents of the size seen
    in its test metrics, time steps log-uniform over its accepted range."""
    _, meta = artifact_state()
    scale = np.asarray(meta["normalization"]["macro_strain"]["rms_scale"])
    lo, hi = meta["accepted_time_step_range"]
    rng = np.random.default_rng(seed)
    strain = rng.standard_normal((n, 6)) * scale * 2.0
    latent = rng.standard_normal((n, 6)) * 0.5
    dt = np.exp(rng.uniform(np.log(lo), np.log(hi), n))
    return strain, latent, dt
