"""GPU hop neighbour lists and minimum-image vectors vs the oracle."""
import numpy as np
import pytest

from oracle import ewald
from oracle.neighbors import hop_neighbor_list
from pycd_amd import _native, workloads

pytestmark = pytest.mark.gpu


def _compare_lists(gpu, ref, cell):
    assert gpu.keys() == ref.keys()
    inv = np.linalg.inv(cell)
    for key in ref:
        for cg, cr in zip(gpu[key], ref[key]):
            for sg, sr in zip(cg, cr):
                assert np.array_equal(sg.num_neighbors, sr.num_neighbors)
                for k in range(len(sr.neighbor_system_element_indices)):
                    assert np.array_equal(sg.neighbor_system_element_indices[k],
                                          sr.neighbor_system_element_indices[k])
                    a = np.asarray(sg.displacement_vector_list[k])
                    b = np.asarray(sr.displacement_vector_list[k])
                    if len(b):
                        # equal, or equal images of an exact half-box tie
                        f = (a - b) @ inv
                        assert np.max(np.abs(f - np.round(f))) < 1e-9
                        assert np.allclose(np.linalg.norm(a, axis=1),
                                           np.linalg.norm(b, axis=1),
                                           rtol=0, atol=1e-12)


@pytest.mark.parametrize("name", ["hematite", "bvo"])
def test_examples(name, request):
    ex = request.getfixturevalue(name)
    gpu = ex.neighbors.build_hop_neighbor_list()
    ref = hop_neighbor_list(ex.material, ex.neighbors)
    _compare_lists(gpu, ref, ex.neighbors.simulation_cell_matrix)


@pytest.mark.parametrize("builder,size", [(workloads.corundum, 4),
                                          (workloads.monoclinic, 5)])
def test_workload_lattices(builder, size):
    w = builder(size)
    gpu = w.neighbors.build_hop_neighbor_list()
    ref = hop_neighbor_list(w.material, w.neighbors)
    _compare_lists(gpu, ref, w.neighbors.simulation_cell_matrix)


def test_min_image_vectors(hematite):
    pos = hematite.shipped_pos
    cell = hematite.neighbors.simulation_cell_matrix
    gpu = _native.min_image_vectors(pos, cell, hematite.sim["pbc"])
    ref = ewald.pairwise_vectors(pos, cell, hematite.sim["pbc"])
    assert np.allclose(np.linalg.norm(gpu, axis=2),
                       np.linalg.norm(hematite.pairwise, axis=2),
                       rtol=0, atol=1e-12)
    f = (gpu - ref).reshape(-1, 3) @ np.linalg.inv(cell)
    assert np.max(np.abs(f - np.round(f))) < 1e-9


def test_neighbor_list_file_round_trip(tmp_path, hematite):
    """Neighbors.generate_neighbor_list writes hop_neighbor_list.npy; loaded
    the way the reference's drivers load it (np.load with allow_pickle,
    then [()], PyCD/material_run.py:68-69, a file this package wrote), it
    holds the oracle's lists, and a Run built from the loaded lists replays
    the shipped trajectory's first hops."""
    from pycd_amd.core import Run, System
    from tests.kmc_helpers import run_gpu
    ex = hematite
    ex.neighbors.generate_neighbor_list(tmp_path, [3, 3, 3])
    path = tmp_path / "hop_neighbor_list.npy"
    assert (tmp_path / "neighbor_list.log").exists()
    loaded = np.load(path, allow_pickle=True)[()]
    ref = hop_neighbor_list(ex.material, ex.neighbors)
    _compare_lists(loaded, ref, ex.neighbors.simulation_cell_matrix)
    system = System(ex.material, ex.neighbors, loaded, ex.pairwise,
                    ex.cfg["alpha"], ex.cfg["r_cut"], ex.cfg["k_cut"],
                    ex.cfg["precision_parameters"], [], [])
    run = Run(system, ex.table, ex.sim["temp"], ex.sim["ion_charge_type"],
              ex.sim["species_charge_type"], ex.sim["n_traj"],
              ex.sim["t_final"], ex.sim["time_interval"],
              np.array(ex.sim["species_count"]), ex.sim["initial_occupancy"],
              ex.sim["relative_energies"], ex.sim["external_field"],
              ex.sim["doping"])
    res, _ = run_gpu(run, ex.table, 2, 1, max_hops=2000, record=False)
    td = ex.time_data[:res["hops"] + 1]
    assert res["hops"] == 2000
    rel = np.abs(res["time_data"][1:] - td[1:]) / td[1:]
    assert rel.max() < 1e-12
