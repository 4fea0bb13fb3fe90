"""Copy the reference's benchmark configs that the GPU config tests run into
tests/golden/configs/ (input data: the GPU box has no reference checkout).

Run in the build container: ``python tests/golden/copy_configs.py``.
"""

import shutil
from pathlib import Path

SRC = Path("/root/reference/configs/benchmarks")
DST = Path(__file__).resolve().parent / "configs"
FILES = [
    "rve/single_phase_j2_viscoplastic.yaml",
    "plasticity/plastic_j2_viscoplastic_material_point.yaml",
    "plasticity/dynamic_j2_viscoplastic_pressure_pulse_3d.yaml",
]

if __name__ == "__main__":
    DST.mkdir(parents=True, exist_ok=True)
    for f in FILES:
        shutil.copyfile(SRC / f, DST / Path(f).name)
        print("copied", f)
