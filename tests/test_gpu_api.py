"""Drop-in use of Run.do_kmc_steps and Analysis.compute_msd."""
import pickle

import numpy as np
import pytest

from oracle import msd as omsd, seeding
from pycd_amd import Analysis, constants

pytestmark = pytest.mark.gpu


def test_do_kmc_steps_reproduces_shipped_files(tmp_path, hematite):
    ex = hematite
    ex.run.do_kmc_steps(tmp_path, ex.sim["output_data"], ex.sim["random_seed"],
                        ex.sim["compute_mode"])
    traj = tmp_path / "traj1"
    # preproduction wrote the seeded state with this Python's default pickle
    # protocol (the shipped dump is protocol 3; its state is pinned by
    # test_seeding_protocol_bytes)
    state = seeding.trajectory_states(ex.sim["random_seed"], 1)[0]
    assert (traj / "initial_rnd_state.dump").read_bytes() == pickle.dumps(state)
    td = np.load(traj / "time_data.npy")
    assert len(td) == len(ex.time_data)
    assert np.max(np.abs(td[1:] - ex.time_data[1:]) / ex.time_data[1:]) < 1e-12
    ut = np.load(traj / "unwrapped_traj.npy")
    assert np.max(np.abs(ut - ex.unwrapped)) < 1e-9
    assert (tmp_path / "Run.log").exists()

    sim = ex.sim
    an = Analysis(ex.material, sim["n_dim"], np.array(sim["species_count"]),
                  sim["n_traj"], sim["t_final"], sim["time_interval"],
                  sim["msd_t_final"], sim["trim_length"], sim["temp"],
                  sim["repr_time"], sim["repr_dist"])
    out = an.compute_msd(tmp_path, sim["output_data"])
    P = an.num_path_steps_per_traj
    pos = ut.reshape(P, 1, 3) * (1 / constants.ANG2BOHR)
    ref = omsd.squared_displacement(pos, 1, P, an.num_msd_steps_per_traj)
    assert np.allclose(out.msd_data[:, 1], ref[0, :, 0], rtol=1e-12, atol=0)
    assert out.msd_data.shape == (501, 2)
    log = (tmp_path / f"MSD_Analysis_{out.file_name}.log").read_text()
    assert "electron diffusivity" in log
