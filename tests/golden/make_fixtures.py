"""Copies the reference's shipped example data into tests/golden/.

    python tests/golden/make_fixtures.py /path/to/PyCD

The files are the inputs and outputs of the reference's own example runs
(examples/{Hematite,BVO}/): POSCAR and YAML inputs, the Ewald table
precomputed_array.npy with the minimum-image vectors it was built from, and
traj1/time_data.npy + unwrapped_traj.npy of one seeded trajectory.
initial_rnd_state.dump is kept as raw bytes: tests compare it byte for byte
with a state pickled by the test itself and never unpickle it.
hop_neighbor_list.npy is not copied (it can only be read by unpickling);
the hop tables are rebuilt from POSCAR and pinned by the trajectory replay.
"""
import shutil
import sys
from pathlib import Path

FILES = {
    "InputFiles": ["POSCAR", "sys_config.yml", "precomputed_array.npy",
                   "pairwise_min_image_vector_data.npy"],
    "traj1": ["time_data.npy", "unwrapped_traj.npy",
              "initial_rnd_state.dump"],
    ".": ["simulation_parameters.yml"],
}


def main(ref_root):
    here = Path(__file__).resolve().parent
    for material in ("Hematite", "BVO"):
        src = Path(ref_root) / "examples" / material
        dst = here / material.lower()
        dst.mkdir(exist_ok=True)
        for sub, names in FILES.items():
            for name in names:
                shutil.copyfile(src / sub / name, dst / name)
    print("fixtures written to", here)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
