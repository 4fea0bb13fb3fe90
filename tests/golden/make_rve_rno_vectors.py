"""Write ``rve_rno_c48_v2_liu_direct.npz``: the weights of the reference's trained
RVE-RNO artifact ``artifacts/rve_rno/c48_v2_liu_direct/best_model.pt`` (read with
``torch.load(weights_only=True)``; its ``model_state_dict`` stored as float32
arrays under their state-dict names) and its ``artifact_metadata.json`` as
text.  Data only: the GPU box has no reference checkout.

The reference's model module cannot be imported in this container (its config
module imports ``fem.solvers``, which imports ``gmsh``; gmsh is not installed),
so no reference-computed vectors are written: the tests pin the oracle by the
reference tests' properties, by torch's own ``nn.Linear``/``SiLU`` modules loaded
strictly with these state-dict names, and by the reference-held FE² sample.

Run in the build container: ``python tests/golden/make_rve_rno_vectors.py``.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import torch

ARTIFACT = Path("/root/reference/artifacts/rve_rno/c48_v2_liu_direct")
OUT = Path(__file__).resolve().parent


def main() -> None:
    checkpoint = torch.load(ARTIFACT / "best_model.pt", map_location="cpu", weights_only=True)
    state = checkpoint["model_state_dict"]
    weights = {k: v.detach().cpu().numpy().astype(np.float32) for k, v in state.items()}
    meta = (ARTIFACT / "artifact_metadata.json").read_text(encoding="utf-8")
    np.savez_compressed(OUT / "rve_rno_c48_v2_liu_direct.npz",
                        metadata=np.array(meta), **weights)
    for k, v in weights.items():
        print(k, v.shape)


if __name__ == "__main__":
    main()
