"""Copy the reference's data files that the tests read into tests/golden/: the
benchmark configs that the GPU config tests run and the data plans
(tests/golden/configs/), and the
reference's sample RVE response shard (tests/golden/rve_rno_c48_sample_response.h5,
from data/rve_rno_c48_v2_2000/sample/response.h5).  Input data only: the GPU box
has no reference checkout.

Run in the build container: ``python tests/golden/copy_reference_data.py``.
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

PLANS = ["rve_rno_c48_v2_2000.yaml", "rve_rno_fixed_cell_v1.yaml", "rve_rno_data_plan.yaml",
         "rve_rno_data_pilot.yaml", "rve_rno_fe2_preflight.yaml"]  # configs/experiments: data plans

SAMPLE = Path("/root/reference/data/rve_rno_c48_v2_2000/sample/response.h5")

if __name__ == "__main__":
    DST.mkdir(parents=True, exist_ok=True)
    for f in FILES:
        shutil.copyfile(SRC / f, DST / Path(f).name)
        print("copied", f)
    for f in PLANS:
        shutil.copyfile(SRC.parent / "experiments" / f, DST / f)
        print("copied", f)
    shutil.copyfile(SAMPLE, DST.parent / "rve_rno_c48_sample_response.h5")
    print("copied", SAMPLE)
