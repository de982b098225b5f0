"""PyCD's `Material` / `Neighbors` / `System` / `Run` / `Analysis` API on MI355X.

The constructors and methods keep the reference's names, arguments and side
effects (files written, logs, return values) so that a PyCD user script or
driver works unchanged:

  System(...)                    PyCD/core.py:678-797
  System.get_precomputed_array   PyCD/core.py:1611-1679  -> HIP Ewald table
  Run(...)                       PyCD/core.py:1685-1933
  Run.do_kmc_steps               PyCD/core.py:2647-2917  -> HIP KMC kernel
  Analysis(...)                  PyCD/core.py:2922-2972
  Analysis.compute_msd           PyCD/core.py:2974-3081  -> HIP MSD kernel

The three hot loops run in `libpycd_hip.so` (see `pycd_amd._native`); there
is no CPU fallback: a missing library raises.  Host code here only prepares
inputs (lattice, hop tables, seeded uniform stream) and writes outputs.

Doping (manual and random insertion) changes per-trajectory lattice charges,
site energies and initial occupancy (`pycd_amd.doping`); the kernel takes
them through `pycd_kmc_set_sites`.

Scope limits (raise NotImplementedError): automatic Ewald parameter
optimisation ('optimal'/'simulation_cell'/'converge' values), pairwise and
gradient dopant insertion, and more than one carrier species type with
non-zero count.
"""
from datetime import datetime
from collections import defaultdict
from pathlib import Path
import itertools
import pickle
import random as rnd

import numpy as np

from . import constants
from .io import read_poscar, generate_report
from .lattice import z_sorted_unit_cell, supercell_sites, MinimumImage
from . import rng
from . import _native
from . import doping as doping_mod


class ReturnValues(object):
    """Attribute bag (PyCD/core.py:3457-3462)."""

    def __init__(self, **kwargs):
        for key, value in kwargs.items():
            setattr(self, key, value)


def _real(value):
    return isinstance(value, (int, float, np.integer, np.floating)) \
        and not isinstance(value, bool)


class Material(object):
    """Material description (PyCD/core.py:37-152)."""

    def __init__(self, material_parameters):
        info = read_poscar(material_parameters.input_coord_file_location)
        self.lattice_matrix = info["lattice_matrix"]
        self.element_types = info["element_types"]
        self.n_elements_per_unit_cell = info["num_elements"]
        self.total_elements_per_unit_cell = info["total_elements"]
        coords = info["coordinates"]
        if info["coordinate_type"] == "Direct":
            frac = coords
        elif info["coordinate_type"] == "Cartesian":
            frac = np.dot(coords, np.linalg.inv(self.lattice_matrix))
        else:
            raise ValueError(f"unknown POSCAR coordinate type "
                             f"{info['coordinate_type']!r}")
        self.lattice_matrix *= constants.ANG2BOHR
        self.num_element_types = len(self.element_types)
        self.element_type_index_list = np.repeat(
            np.arange(self.num_element_types), self.n_elements_per_unit_cell)
        self.name = material_parameters.name
        self.species_types = material_parameters.species_types[:]
        self.num_species_types = len(self.species_types)
        self.species_charge_list = material_parameters.species_charge_list
        self.species_to_element_type_map = dict(
            material_parameters.species_to_element_type_map.items())
        (self.fractional_unit_cell_coords,
         self.unit_cell_class_list) = z_sorted_unit_cell(
            frac, self.element_type_index_list, self.n_elements_per_unit_cell,
            material_parameters.class_list, self.element_types)
        self.cartesian_unit_cell_coords = np.dot(
            self.fractional_unit_cell_coords, self.lattice_matrix)
        self.charge_types = material_parameters.charge_types
        self.vn = material_parameters.vn / constants.SEC2AUTIME

        def per_class(table, scale):
            return {key: [[v * scale for v in vals] for vals in cls]
                    for key, cls in table.items()}

        self.lambda_values = per_class(material_parameters.lambda_values,
                                       constants.EV2HARTREE)
        self.v_ab = per_class(material_parameters.v_ab, constants.EV2HARTREE)
        self.neighbor_cutoff_dist = {
            key: [[(v * constants.ANG2BOHR) if v else None for v in vals]
                  for vals in cls]
            for key, cls in material_parameters.neighbor_cutoff_dist.items()}
        self.neighbor_cutoff_dist_tol = {
            key: [[(v * constants.ANG2BOHR) if v else None for v in vals]
                  for vals in cls]
            for key, cls in
            material_parameters.neighbor_cutoff_dist_tol.items()}
        self.num_unique_hopping_distances = {
            key: [len(vals) for vals in cls]
            for key, cls in self.neighbor_cutoff_dist.items()}
        self.element_type_delimiter = \
            material_parameters.element_type_delimiter
        self.dielectric_constant = material_parameters.dielectric_constant
        self.num_classes = [len(set(material_parameters.class_list[e]))
                            for e in self.element_types]
        self.element_type_to_species_map = defaultdict(list)
        for etype in self.element_types:
            for stype in self.species_types:
                if etype in self.species_to_element_type_map[stype]:
                    self.element_type_to_species_map[etype].append(stype)
        self.hop_element_types = {
            stype: [self.element_type_delimiter.join(comb)
                    for comb in itertools.product(
                        self.species_to_element_type_map[stype], repeat=2)]
            for stype in self.species_types}

    def generate_sites(self, element_type_indices, cell_size):
        """Supercell sites (PyCD/core.py:154-241)."""
        cell_size = np.asarray(cell_size)
        assert all(size > 0 for size in cell_size), \
            "Input size should always be greater than 0"
        coords, quantum, system_index = supercell_sites(
            self.cartesian_unit_cell_coords, self.lattice_matrix, cell_size,
            list(element_type_indices), self.element_type_index_list,
            self.n_elements_per_unit_cell)
        return ReturnValues(cell_coordinates=coords,
                            quantum_index_list=quantum,
                            system_element_index_list=system_index)


class Neighbors(object):
    """Sites, minimum image and hop neighbour lists (PyCD/core.py:244-668)."""

    def __init__(self, material, system_size, pbc):
        self.start_time = datetime.now()
        self.material = material
        self.system_size = np.asarray(system_size)
        self.n_dim = len(system_size)
        self.pbc = np.asarray(pbc)
        self.num_cells = self.system_size.prod()
        self.num_system_elements = (
            self.num_cells * self.material.total_elements_per_unit_cell)
        self.element_type_indices = range(self.material.num_element_types)
        self.bulk_sites = self.material.generate_sites(
            self.element_type_indices, self.system_size)
        self._setup_efficient_pbc()

    def _setup_efficient_pbc(self):
        self._mic = MinimumImage(self.material.lattice_matrix,
                                 self.system_size, self.pbc)
        self.simulation_cell_matrix = self._mic.cell
        self.simulation_cell_matrix_inv = self._mic.cell_inv
        self.pbc_mask = self._mic.mask

    def apply_minimum_image_convention(self, displacement_vector):
        return self._mic.apply(np.asarray(displacement_vector))

    def apply_minimum_image_convention_vectorized(self, displacement_vectors):
        return self._mic.apply(np.asarray(displacement_vectors))

    def get_system_element_index(self, system_size, quantum_indices):
        """Quantum indices -> system index (PyCD/core.py:379-418)."""
        cell = np.asarray(quantum_indices[:3])
        etype, elem = quantum_indices[-2:]
        tot = self.material.total_elements_per_unit_cell
        index = elem + self.material.n_elements_per_unit_cell[:etype].sum()
        size = np.asarray(system_size)
        n = len(size)
        for k in range(n):
            stride = tot if k == 0 else tot * size[-k:].prod()
            index += stride * cell[n - 1 - k]
        return index

    def get_quantum_indices(self, system_size, system_element_index):
        """System index -> quantum indices (PyCD/core.py:420-449)."""
        tot = self.material.total_elements_per_unit_cell
        size = np.asarray(system_size)
        out = np.zeros(5, dtype=int)
        in_cell = system_element_index % tot
        cum = self.material.n_elements_per_unit_cell.cumsum()
        out[3] = np.nonzero(cum >= in_cell + 1)[0][0]
        out[4] = in_cell - self.material.n_elements_per_unit_cell[:out[3]].sum()
        cells = (system_element_index - in_cell) // tot
        for k in range(3):
            stride = size[k + 1:].prod()
            out[k] = cells // stride
            cells -= out[k] * stride
        return out

    def get_coordinates(self, system_size, system_element_index):
        q = self.get_quantum_indices(system_size, system_element_index)
        shift = np.dot(q[:3], self.material.lattice_matrix)
        k = q[4] + self.material.n_elements_per_unit_cell[:q[3]].sum()
        return shift + self.material.cartesian_unit_cell_coords[k]

    def compute_distance(self, system_size, system_element_index_1,
                         system_element_index_2):
        a = self.get_coordinates(system_size, system_element_index_1)
        b = self.get_coordinates(system_size, system_element_index_2)
        return float(np.linalg.norm(self._mic.apply(b - a)))

    def hop_neighbor_sites(self, bulk_sites, center_site_indices,
                           neighbor_site_indices, cutoff_dist_limits,
                           cutoff_dist_key=None):
        """Neighbours in the window (lo, hi] (PyCD/core.py:485-569), GPU."""
        return self._hop_shells(bulk_sites, center_site_indices,
                                neighbor_site_indices,
                                [cutoff_dist_limits])[0]

    def _hop_shells(self, bulk_sites, centers, targets, limits):
        """All windows in one pass of the GPU neighbour kernel."""
        centers = np.asarray(centers, dtype=np.int64)
        targets = np.asarray(targets, dtype=np.int64)
        lo = [float(a) for a, _ in limits]
        hi = [float(b) for _, b in limits]
        count, index, vec = _native.hop_neighbors(
            bulk_sites.cell_coordinates, centers, targets,
            self.simulation_cell_matrix, self.pbc, lo, hi)
        sys_index = bulk_sites.system_element_index_list[targets]
        out = []
        for s in range(len(limits)):
            nbr = np.empty(len(centers), dtype=object)
            disp = np.empty(len(centers), dtype=object)
            for k in range(len(centers)):
                m = int(count[k, s])
                nbr[k] = sys_index[index[k, s, :m]]
                disp[k] = vec[k, s, :m].copy()
            out.append(ReturnValues(neighbor_system_element_indices=nbr,
                                    displacement_vector_list=disp,
                                    num_neighbors=count[:, s].astype(int)))
        return out

    def get_pairwise_min_image_vector_data(self, dst_path):
        """N x N x 3 minimum-image vectors to disk (PyCD/core.py:571-594)."""
        out = _native.min_image_vectors(self.bulk_sites.cell_coordinates,
                                        self.simulation_cell_matrix, self.pbc)
        np.save(Path(dst_path).joinpath("pairwise_min_image_vector_data.npy"),
                out)
        return None

    def hopping_site_indices(self, element_type_index):
        """System indices of one element type, cell-major (core.py:621-632)."""
        tot = self.material.total_elements_per_unit_cell
        n_e = self.material.n_elements_per_unit_cell[element_type_index]
        head = self.material.n_elements_per_unit_cell[:element_type_index].sum()
        offsets = np.repeat(np.arange(0, tot * self.num_cells, tot), n_e)
        return np.tile(head + np.arange(n_e), self.num_cells) + offsets

    def build_hop_neighbor_list(self):
        """The hop neighbour dictionary without writing it to disk."""
        out = {}
        tol = self.material.neighbor_cutoff_dist_tol
        delim = self.material.element_type_delimiter
        for key, per_class in self.material.neighbor_cutoff_dist.items():
            center_type = key.split(delim)[0]
            t_index = self.material.element_types.index(center_type)
            sites = self.hopping_site_indices(t_index)
            classes = []
            for c, cutoffs in enumerate(per_class):
                limits = [[cut - tol[key][c][t], cut + tol[key][c][t]]
                          for t, cut in enumerate(cutoffs)]
                classes.append(self._hop_shells(self.bulk_sites, sites, sites,
                                                limits))
            out[key] = classes
        return out

    def generate_neighbor_list(self, dst_path, local_system_size):
        """Hop neighbour list to disk (PyCD/core.py:596-668)."""
        assert dst_path, ("Please provide the path to the parent directory "
                          "of neighbor list files")
        assert all(size >= 3 for size in local_system_size), \
            ("Local system size in all dimensions should always be greater "
             "than or equal to 3")
        dst_path = Path(dst_path)
        Path.mkdir(dst_path, parents=True, exist_ok=True)
        np.save(dst_path.joinpath("hop_neighbor_list.npy"),
                self.build_hop_neighbor_list())
        generate_report(self.start_time, dst_path, "neighbor_list", 1)
        return None


class System(object):
    """Supercell, Ewald parameters and the pair table (PyCD/core.py:671-1679).

    `pairwise_min_image_vector_data` may be None: the HIP table kernel forms
    minimum-image distances from the site coordinates itself.  The keyword
    `table_precision` selects the reciprocal-space arithmetic: 'f64'
    (default, drop-in), 'f32' (FP32 MFMA) or 'bf16x3' (three BF16 MFMA
    products with FP32 accumulation).
    """

    def __init__(self, material_info, material_neighbors, hop_neighbor_list,
                 pairwise_min_image_vector_data, alpha, r_cut, k_cut,
                 precision_parameters, step_system_size_array,
                 step_hop_neighbor_master_list, table_precision="f64"):
        self.start_time = datetime.now()
        self.material = material_info
        self.neighbors = material_neighbors
        if hop_neighbor_list is None:
            hop_neighbor_list = material_neighbors.build_hop_neighbor_list()
        self.hop_neighbor_list = hop_neighbor_list
        self.step_system_size_array = step_system_size_array
        self.step_hop_neighbor_master_list = step_hop_neighbor_master_list
        self.num_unique_step_systems = len(step_hop_neighbor_master_list)
        self.pbc = self.neighbors.pbc
        self.system_size = self.neighbors.system_size
        self.num_cells = self.system_size.prod()
        self.pairwise_min_image_vector_data = pairwise_min_image_vector_data
        self.table_precision = table_precision

        # Column-broadcast translational matrix (reference quirk,
        # PyCD/core.py:709-710) drives volume and reciprocal lattice.
        self.translational_matrix = np.multiply(self.system_size,
                                                self.material.lattice_matrix)
        t = self.translational_matrix
        self.system_volume = abs(np.dot(t[0], np.cross(t[1], t[2])))
        self.reciprocal_lattice_matrix = (
            2 * np.pi / self.system_volume
            * np.array([np.cross(t[1], t[2]), np.cross(t[2], t[0]),
                        np.cross(t[0], t[1])]))
        self.translational_vector_length = np.linalg.norm(t, axis=1)
        self.reciprocal_lattice_vector_length = np.linalg.norm(
            self.reciprocal_lattice_matrix, axis=1)
        self.system_class_index_list = np.tile(
            self.material.unit_cell_class_list, self.num_cells)

        self.num_neighbors = np.zeros(self.material.num_species_types, int)
        for s_index, stype in enumerate(self.material.species_types):
            hop_type = self.material.hop_element_types[stype][0]
            if hop_type in self.material.neighbor_cutoff_dist:
                n_types = len(self.material.neighbor_cutoff_dist[hop_type][0])
            else:
                n_types = 0
            for t_index in range(n_types):
                self.num_neighbors[s_index] += (
                    self.hop_neighbor_list[hop_type][0][t_index]
                    .num_neighbors[0])

        # alpha in 1/Angstrom -> divided by ANG2BOHR and used as eta^2
        # (PyCD/core.py:768-784).
        self.alpha = alpha / constants.ANG2BOHR if _real(alpha) else alpha
        self.r_cut = r_cut * constants.ANG2BOHR if _real(r_cut) else r_cut
        if isinstance(k_cut, list):
            self.k_cut = k_cut
        elif _real(k_cut):
            self.k_cut = k_cut / constants.ANG2BOHR
        else:
            self.k_cut = k_cut
        pp = precision_parameters or {}
        self.precision_parameters = pp
        self.err_tol = pp.get("err_tol", 0.0) * constants.EV2HARTREE

    # -- helpers ---------------------------------------------------------
    def base_charge_config_for_accuracy_analysis(self, ion_charge_type):
        charges = np.array([
            self.material.charge_types[ion_charge_type][
                self.material.element_types[t]]
            for t in self.material.element_type_index_list])
        return np.tile(charges, self.num_cells)[:, np.newaxis]

    def _ewald_parameters(self):
        if not (_real(self.alpha) and _real(self.r_cut)):
            raise NotImplementedError(
                "automatic Ewald parameter selection is not part of this "
                "implementation; give alpha, r_cut and k_cut as numbers")
        if isinstance(self.k_cut, list):
            k_cut = 1.10 * max(np.asarray(self.k_cut)
                               * self.reciprocal_lattice_vector_length)
        elif _real(self.k_cut):
            k_cut = self.k_cut
        else:
            raise NotImplementedError(
                "automatic k_cut selection is not part of this "
                "implementation; give k_cut as a number or k_max list")
        return self.alpha, self.r_cut, k_cut

    def k_max(self, k_cut):
        return np.ceil(k_cut / self.reciprocal_lattice_vector_length
                       ).astype(int)

    def get_precomputed_array(self, dst_path, compute_energy_contributions,
                              return_k_vector_data):
        """Ewald pair table on the GPU (PyCD/core.py:1611-1679).

        Returns (A, dst_path) with A the N x N FP64 table, and writes
        `precomputed_array.log` with alpha on its fourth line.  With
        `compute_energy_contributions` the log also gets the real-space,
        Fourier-space and self-interaction energies of the lattice charges
        (PyCD/core.py:1662-1675); with `return_k_vector_data` the energy of
        each half-space k vector goes, largest first, to
        `k_vector_energy_contribution_within_user_specified_k_cut.log`
        (PyCD/core.py:1628-1649).
        """
        alpha, r_cut, k_cut = self._ewald_parameters()
        pos = self.neighbors.bulk_sites.cell_coordinates
        table, report = _native.pair_table(
            pos, self.neighbors.simulation_cell_matrix,
            self.translational_matrix, self.pbc, alpha, r_cut, k_cut,
            self.material.dielectric_constant, self.table_precision)
        self.last_table_report = report
        k_max = self.k_max(k_cut)
        n = len(pos)
        tau_r = report["ms_real"] / max(report["n_real_pairs"], 1)
        tau_f = report["ms_recip"] / (n * n * max(report["n_kvec"], 1))
        lines = []
        lines.append(f"tau_ratio, (tau_r/tau_f): "
                     f"{(tau_r / tau_f if tau_f > 0 else 0.0):.3e}\n")
        lines.append(f"time_ratio, (time_r/time_f): "
                     f"{(report['ms_real'] / report['ms_recip'] if report['ms_recip'] > 0 else 0.0):.3e}\n\n")
        choice = "user-specified"
        lines.append(f"alpha: {alpha * constants.ANG2BOHR:.3e} / angstrom "
                     f"({choice})\n")
        lines.append(f"r_cut: {r_cut / constants.ANG2BOHR:.3e} angstrom "
                     f"({choice})\n")
        lines.append(f"k_cut: {k_cut * constants.ANG2BOHR:.3e} / angstrom "
                     f"({choice if not isinstance(self.k_cut, list) else 'user-specified (k_max)'})\n")
        q = self.base_charge_config_for_accuracy_analysis("full")
        q2 = float(np.einsum("ii", q * q.T))
        vl = np.power(self.system_volume, 1 / 3)
        err_r = q2 * np.sqrt(r_cut / (2 * self.system_volume)) * (
            np.exp(-(alpha * r_cut) ** 2) / (alpha * r_cut) ** 2)
        n_cut = k_cut * vl / (2 * np.pi)
        x = np.pi * n_cut / (alpha * vl)
        err_k = q2 * np.sqrt(n_cut) / (alpha * vl ** 2) * (np.exp(-x ** 2)
                                                         / x ** 2)
        lines.append(f"Real-space cutoff error: {err_r:.3e}\n")
        lines.append(f"Fourier-space cutoff error: {err_k:.3e}\n\n")
        n_est = np.ceil(np.prod(2 * k_max + 1) * np.pi / 6 - 1).astype(int)
        lines.append(f"k_max: [{k_max[0]}, {k_max[1]}, {k_max[2]}]\n")
        lines.append(f"number of k-vectors: {n_est}\n\n")
        if return_k_vector_data:
            self._k_vector_report(q[:, 0], alpha, k_cut, k_max, dst_path)
        if compute_energy_contributions:
            parts = self.energy_parts(table, q[:, 0], alpha, r_cut)
            ev = constants.EV2HARTREE
            lines.append(f"Energy contribution from Real space: "
                         f"{parts['real'] / ev} eV\n")
            lines.append(f"Energy contribution from Fourier-space: "
                         f"{parts['fourier'] / ev} eV\n")
            lines.append(f"Energy contribution from self-interactions: "
                         f"{parts['self'] / ev} eV\n")
            lines.append(f"Total system energy (neutral): "
                         f"{parts['total'] / ev} eV\n\n")
        if dst_path is not None:
            generate_report(self.start_time, Path(dst_path),
                            "precomputed_array", 1, "".join(lines))
        return (table, dst_path)

    def energy_parts(self, table, q, alpha, r_cut):
        """q^T A q split into its real-space, Fourier-space and self parts
        (PyCD/core.py:1662-1675) for charges q.  The real and self parts
        come from a second table without reciprocal sum (k_cut = 0) on the
        GPU; the self part is -sqrt(alpha / pi) / eps sum q^2."""
        eps = self.material.dielectric_constant
        pos = self.neighbors.bulk_sites.cell_coordinates
        real_self, _ = _native.pair_table(
            pos, self.neighbors.simulation_cell_matrix,
            self.translational_matrix, self.pbc, alpha, r_cut, 0.0, eps,
            "f64")
        q = np.asarray(q, dtype=np.float64)
        e_self = -np.sqrt(alpha / np.pi) / eps * float(q @ q)
        e_rs = float(q @ real_self @ q)
        e_all = float(q @ table @ q)
        real, fourier = e_rs - e_self, e_all - e_rs
        return {"real": real, "fourier": fourier, "self": e_self,
                "total": real + fourier + e_self}

    @staticmethod
    def half_space(k_max):
        """Integer k triples of [-k_max, k_max]^3 without 0, one of each
        +/- pair, in the order of get_effective_k_vectors
        (PyCD/core.py:816-838): mgrid order, keeping v unless -v came
        earlier, i.e. v whose first non-zero component is negative."""
        g = np.mgrid[-k_max[0]:k_max[0] + 1, -k_max[1]:k_max[1] + 1,
                     -k_max[2]:k_max[2] + 1].reshape(3, -1).T
        a, b, c = g[:, 0], g[:, 1], g[:, 2]
        keep = (a < 0) | ((a == 0) & (b < 0)) | ((a == 0) & (b == 0) & (c < 0))
        return g[keep]

    def k_vector_energies(self, q, alpha, k_cut, k_max):
        """Energy of charges q per half-space k vector (PyCD/core.py:880-911,
        before the factor 2 and the dielectric constant):
        (2 pi / V) e^{-k^2/4 alpha} / k^2 sum_ij q_i q_j cos(k . r_ij) for
        |k|^2 < k_cut^2, 0 beyond.  k is a reciprocal vector of the
        supercell, so cos(k . r_ij) is the same for any image of r_ij and
        the double sum is |S(k)|^2 with S(k) = sum_i q_i e^{i k . r_i}."""
        kv = self.half_space(k_max)
        k = kv @ self.reciprocal_lattice_matrix
        k2 = np.einsum("ij,ij->i", k, k)
        inside = k2 < k_cut ** 2
        pos = self.neighbors.bulk_sites.cell_coordinates
        q = np.asarray(q, dtype=np.float64)
        out = np.zeros(len(kv))
        idx = np.nonzero(inside)[0]
        for lo in range(0, len(idx), 256):
            sel = idx[lo:lo + 256]
            phase = pos @ k[sel].T
            s2 = (q @ np.cos(phase)) ** 2 + (q @ np.sin(phase)) ** 2
            out[sel] = ((2 * np.pi) / self.system_volume
                        * np.exp(-k2[sel] / (4 * alpha)) * s2 / k2[sel])
        return kv, out

    def _k_vector_report(self, q, alpha, k_cut, k_max, dst_path):
        kv, energy = self.k_vector_energies(q, alpha, k_cut, k_max)
        order = np.argsort(energy)[::-1]
        lines = ["k-vectors sorted in the decreasing order of their energy "
                 "contributions\n"]
        for i in order:
            v = kv[i]
            lines.append(f"{v[0]:4d} {v[1]:4d} {v[2]:4d}: "
                         f"{energy[i] / constants.EV2HARTREE:.3e} eV\n")
        if dst_path is not None:
            generate_report(
                self.start_time, Path(dst_path),
                "k_vector_energy_contribution_within_user_specified_k_cut", 0,
                "".join(lines))
        self.k_vector_data = kv
        self.k_vector_energy = energy


class Run(object):
    """KMC driver (PyCD/core.py:1682-2917) over the HIP plain-path kernel."""

    def __init__(self, system, precomputed_array, temp, ion_charge_type,
                 species_charge_type, n_traj, t_final, time_interval,
                 species_count, initial_occupancy, relative_energies,
                 external_field, doping):
        self.start_time = datetime.now()
        self.system = system
        self.material = system.material
        self.neighbors = system.neighbors
        self.precomputed_array = precomputed_array
        self.temp = temp * constants.K2AUTEMP
        self.ion_charge_type = ion_charge_type
        self.species_charge_type = species_charge_type
        self.n_traj = int(n_traj)
        self.t_final = t_final * constants.SEC2AUTIME
        self.time_interval = time_interval * constants.SEC2AUTIME
        self.species_count = np.asarray(species_count)
        self.initial_occupancy = initial_occupancy or {}
        self.system_size = system.system_size
        self.relative_energies = relative_energies or {"class_index": {},
                                                       "doping": {}}
        self.doping = doping if doping is not None else {"num_dopants": []}
        spec = doping_mod.DopingSpec(self.doping, self.material)
        self.doping_spec = spec
        self.doping_active = int(spec.active)
        self.dopant_element_types = spec.dopant_element_types
        self.substitution_element_types = spec.substitution_element_types
        self.dopant_species_types = spec.dopant_species_types
        self.dopant_to_substitution_element_type_map = \
            spec.dopant_to_substitution_element_type_map
        self.num_dopant_element_types = len(spec.dopant_element_types)
        # shells with an energy shift, per substitution element type
        # (PyCD/core.py:1731-1736)
        self.num_shells = {
            etype: [len(shifts) - 1 for shifts in per_map]
            for etype, per_map in
            (self.relative_energies.get("doping", {}) or {}).items()}
        # distance -> shell lookups, built on first use by doping
        # (PyCD/core.py:1783-1821)
        self._shell_lookups = {}
        self._hop_targets = {}
        if np.count_nonzero(self.species_count) > 1:
            raise NotImplementedError(
                "more than one carrier species type with non-zero count is "
                "not supported (the reference indexes per-carrier lists by "
                "the per-type species index, PyCD/core.py:1964,2002,2011)")

        # per-site relative energies (PyCD/core.py:1716-1730)
        unit = np.zeros(self.material.total_elements_per_unit_cell)
        start = 0
        class_energy = self.relative_energies.get("class_index", {}) or {}
        for e in range(self.material.num_element_types):
            end = start + self.material.n_elements_per_unit_cell[e]
            if self.material.num_classes[e] != 1:
                etype = self.material.element_types[e]
                unit[start:end] += [
                    class_energy[etype][c] * constants.EV2HARTREE
                    for c in self.material.unit_cell_class_list[start:end]]
            start = end
        self.undoped_system_relative_energies = np.tile(unit,
                                                        system.num_cells)
        self.system_relative_energies = np.copy(
            self.undoped_system_relative_energies)

        self.n_proc = int(np.dot(self.species_count, system.num_neighbors))
        self.head_start_n_elements_per_unit_cell_cum_sum = [
            self.material.n_elements_per_unit_cell[:t].sum()
            for t in self.neighbors.element_type_indices]

        # electric field in atomic units, no conversion (core.py:1749-1761)
        field = external_field["electric"]
        self.electric_field_ld = field.get("ld", 0)
        self.electric_field_mag = field["mag"]
        if field["active"]:
            self.electric_field_active = 1
            direction = np.asarray(field["dir"], dtype=np.float64)
            if self.electric_field_ld == 1:
                direction = np.dot(direction, self.material.lattice_matrix)
            self.electric_field = (self.electric_field_mag
                                   * (direction / np.linalg.norm(direction)))
        else:
            self.electric_field_active = 0
            self.electric_field = np.zeros(self.neighbors.n_dim)

        self.species_type_list = [self.material.species_types[i]
                                  for i, v in enumerate(self.species_count)
                                  for _ in range(v)]
        self.species_type_index_list = [i for i, v in
                                        enumerate(self.species_count)
                                        for _ in range(v)]
        self.species_charge_list = [
            self.material.species_charge_list[self.species_charge_type][i]
            for i in self.species_type_index_list]
        self.total_species = int(self.species_count.sum())
        self.system_coordinates = self.neighbors.bulk_sites.cell_coordinates
        self._build_site_tables()

    # -- tables ------------------------------------------------------------
    def get_element_type_element_index(self, site_element_type_index,
                                       system_element_index):
        """PyCD/core.py:1935-1944."""
        tot = self.material.total_elements_per_unit_cell
        element_index = (system_element_index % tot
                         - self.head_start_n_elements_per_unit_cell_cum_sum[
                             site_element_type_index])
        ete = (system_element_index // tot
               * self.material.n_elements_per_unit_cell[
                   site_element_type_index] + element_index)
        return (ete, element_index)

    def _build_site_tables(self):
        """Per-site hop slots: target, vector, lambda, V_ab.

        Slot order per site = hop-distance type, then ascending neighbour
        index, which is the process order of get_process_attributes
        (PyCD/core.py:1946-1987); lambda and V_ab per (class, distance type)
        as in PyCD/core.py:1884-1927.
        """
        n = self.neighbors.num_system_elements
        hop = self.system.hop_neighbor_list
        delim = self.material.element_type_delimiter
        slots = {}
        max_z = 0
        self.hop_site_lists = {}
        for s_index, stype in enumerate(self.material.species_types):
            key = self.material.hop_element_types[stype][0]
            etype = self.material.species_to_element_type_map[stype][0]
            e_index = self.material.element_types.index(etype)
            sites = self.neighbors.hopping_site_indices(e_index)
            self.hop_site_lists[stype] = sites
            if self.species_count[s_index] == 0 or key not in hop:
                continue
            expected = self.system.num_neighbors[s_index]
            for s in sites:
                cls = self.system.system_class_index_list[s]
                ete, _ = self.get_element_type_element_index(e_index, s)
                entries = []
                for t, shell in enumerate(hop[key][cls]):
                    nb = shell.neighbor_system_element_indices[ete]
                    vv = shell.displacement_vector_list[ete]
                    lam = self.material.lambda_values[key][cls][t]
                    vab = self.material.v_ab[key][cls][t]
                    for m in range(len(nb)):
                        entries.append((int(nb[m]), vv[m], lam, vab))
                if len(entries) != expected:
                    raise NotImplementedError(
                        f"site {s} has {len(entries)} hop targets but its "
                        f"species has {expected}; non-uniform neighbour "
                        f"counts are outside the reference's process model")
                slots[int(s)] = entries
                max_z = max(max_z, len(entries))
        self.max_z = max(max_z, 1)
        self.nbr_count = np.zeros(n, dtype=np.int32)
        self.nbr_site = np.full((n, self.max_z), -1, dtype=np.int32)
        self.nbr_vec = np.zeros((n, self.max_z, 3))
        self.nbr_lambda = np.ones((n, self.max_z))
        self.nbr_vab = np.zeros((n, self.max_z))
        for s, entries in slots.items():
            self.nbr_count[s] = len(entries)
            for j, (t, v, lam, vab) in enumerate(entries):
                self.nbr_site[s, j] = t
                self.nbr_vec[s, j] = v
                self.nbr_lambda[s, j] = lam
                self.nbr_vab[s, j] = vab

    def get_process_attributes(self, occupancy):
        """(old sites, new sites, element-type indices) (core.py:1946-1987)."""
        old, new, ete_list = [], [], []
        for c, s in enumerate(occupancy):
            e_index = self.material.element_types.index(
                self.material.species_to_element_type_map[
                    self.species_type_list[c]][0])
            ete, _ = self.get_element_type_element_index(e_index, s)
            for j in range(self.nbr_count[s]):
                old.append(s)
                new.append(int(self.nbr_site[s, j]))
                ete_list.append(ete)
        return (np.asarray(old, dtype=int), np.asarray(new, dtype=int),
                np.asarray(ete_list, dtype=int))

    # -- initial state -------------------------------------------------------
    def generate_initial_occupancy(self, dopant_site_indices=None):
        """Dopant sites, preset sites, then rnd.sample over hopping sites.

        With site_charge_initiation 'yes' a species whose type a dopant maps
        to is first sampled from that dopant's sites (PyCD/core.py:2478-2528).
        """
        dopant_site_indices = dopant_site_indices or {}
        spec = self.doping_spec
        occupancy = []
        for s_index, count in enumerate(self.species_count):
            stype = self.material.species_types[s_index]
            remaining = int(count)
            if self.doping_active:
                for m, dspecies in enumerate(spec.dopant_species_types):
                    n_sites = self.doping["num_dopants"][m]
                    if (self.doping["site_charge_initiation"][m] == "yes"
                            and dspecies == stype and n_sites and remaining):
                        sites = dopant_site_indices[
                            spec.dopant_element_types[m]]
                        occupancy.extend(rnd.sample(sites, remaining)[:])
                        remaining -= len(sites[:remaining])
            if stype in self.initial_occupancy:
                occupancy.extend(list(self.initial_occupancy[stype]))
                remaining -= len(self.initial_occupancy[stype])
            if remaining:
                candidates = []
                for etype in self.material.species_to_element_type_map[stype]:
                    e_index = self.material.element_types.index(etype)
                    candidates.extend(
                        list(self.neighbors.hopping_site_indices(e_index)))
                    candidates = [i for i in candidates if i not in occupancy]
                occupancy.extend(rnd.sample(candidates, remaining)[:])
        return occupancy

    def base_charge_config(self):
        charges = np.array([
            self.material.charge_types[self.ion_charge_type][
                self.material.element_types[t]]
            for t in self.material.element_type_index_list])
        return np.tile(charges, self.system.num_cells)[:, np.newaxis]

    def lattice_charges(self, dopant_site_indices=None):
        """Lattice charges with each dopant site at its dopant's charge
        (PyCD/core.py:2540-2551)."""
        q = self.base_charge_config()
        if self.doping_active:
            for dtype, sites in (dopant_site_indices or {}).items():
                q[sites] = self.doping["charge"][self.ion_charge_type][dtype]
        return q

    def charge_config(self, occupancy, dopant_site_indices=None):
        """Lattice charges plus carrier charges (PyCD/core.py:2540-2564)."""
        q = self.lattice_charges(dopant_site_indices)
        start = 0
        for s_index in range(self.material.num_species_types):
            end = start + self.species_count[s_index]
            q[occupancy[start:end]] += self.material.species_charge_list[
                self.species_charge_type][s_index]
            start = end
        return q

    # -- doping ---------------------------------------------------------------
    def hop_targets(self, element_type_index):
        """Hop targets of a site of one element type, in the reference's
        slot order (hop-distance type, then neighbour list order), as used
        by the shell search (PyCD/core.py:2366-2380)."""
        if element_type_index not in self._hop_targets:
            etype = self.material.element_types[element_type_index]
            key = self.material.element_type_delimiter.join([etype] * 2)
            per_class = self.system.hop_neighbor_list[key]
            classes = self.system.system_class_index_list
            cache = {}

            def targets(s):
                out = cache.get(s)
                if out is None:
                    ete, _ = self.get_element_type_element_index(
                        element_type_index, s)
                    out = []
                    for shell in per_class[classes[s]]:
                        out.extend(
                            shell.neighbor_system_element_indices[ete]
                            .tolist())
                    cache[s] = out
                return out
            self._hop_targets[element_type_index] = targets
        return self._hop_targets[element_type_index]

    def _element_type_of(self, site):
        e_index = int(self.neighbors.get_quantum_indices(self.system_size,
                                                         site)[3])
        return e_index, self.material.element_types[e_index]

    def get_shell_based_neighbors(self, site_system_element_index, num_shells,
                                  system_size=None,
                                  system_class_index_list=None,
                                  hop_neighbor_list=None):
        """Shells 0..num_shells of hop distance (PyCD/core.py:2352-2389)."""
        e_index, _ = self._element_type_of(site_system_element_index)
        return doping_mod.shells(int(site_system_element_index), num_shells,
                                 self.hop_targets(e_index))

    def shell_lookup(self, element_type):
        """Distance -> shell lookup of a sublattice (core.py:1783-1821)."""
        if element_type not in self._shell_lookups:
            e_index = self.material.element_types.index(element_type)
            sample = int(self.neighbors.get_system_element_index(
                self.system_size, [0, 0, 0, e_index, 0]))
            self._shell_lookups[element_type] = doping_mod.ShellLookup(
                sample, self.hop_targets(e_index),
                lambda a, b: self.neighbors.compute_distance(
                    self.system_size, a, b))
        return self._shell_lookups[element_type]

    @property
    def max_neighbor_shells(self):
        spec = self.doping_spec
        return {sub: self.shell_lookup(sub).max_shells
                for sub in spec.substitution_element_types}

    def get_doping_distribution(self):
        """Dopant sites per dopant type (PyCD/core.py:2109-2303), manual and
        random insertion."""
        spec = self.doping_spec
        dopant_site_indices = {}
        inserted = 0
        available = None
        used_types = None
        for m, n_dopants in enumerate(self.doping["num_dopants"]):
            kind = self.doping["insertion_type"][m]
            dtype = spec.dopant_element_types[m]
            if kind == "gradient" or not n_dopants:
                continue
            if kind == "manual":
                dopant_site_indices[dtype] = (
                    self.doping["dopant_site_indices"][m][:n_dopants])
            elif kind == "random":
                if inserted == 0:
                    used_types, available = [], []
                elif used_types is None:
                    raise ValueError(
                        "random insertion after another insertion type: the "
                        "reference has no list of available sites here")
                sub = spec.substitution_element_types[m]
                e_index = self.material.element_types.index(sub)
                if e_index not in used_types:
                    available.extend(doping_mod.sublattice_sites(
                        self.material, self.system.num_cells,
                        e_index).tolist())
                used_types.append(e_index)
                chosen, available = doping_mod.random_insertion(
                    available, n_dopants,
                    self.doping["min_shell_separation"][m],
                    lambda s, k: self.get_shell_based_neighbors(s, k))
                dopant_site_indices[dtype] = chosen
            else:
                raise NotImplementedError(f"{kind!r} dopant insertion")
            inserted += 1
        return dopant_site_indices

    def get_doping_analysis(self, dopant_site_indices, prefix_list):
        """Report lines and minimum shell separations (core.py:2305-2350)."""
        spec = self.doping_spec
        lookups = {sub: self.shell_lookup(sub)
                   for sub in set(spec.substitution_element_types)}
        lines, min_sep = doping_mod.analysis_report(
            spec, dopant_site_indices, lookups,
            lambda a, b: self.neighbors.compute_distance(self.system_size,
                                                         a, b),
            lambda s: self.neighbors.get_coordinates(self.system_size, s))
        prefix_list.extend(lines)
        return prefix_list, min_sep

    def get_site_wise_shell_indices_table(self, dopant_site_indices):
        """Shells of every dopant and the per-site shell table
        (PyCD/core.py:2391-2476).  Returns (table, independent)."""
        spec = self.doping_spec
        max_shells = self.max_neighbor_shells
        dopant_types, dopant_shells = {}, {}
        for dtype, sites in dopant_site_indices.items():
            m = spec.dopant_element_types.index(dtype)
            sub = spec.substitution_element_types[m]
            depth = 0 if dtype == "X" else max_shells[sub]
            for s in sites:
                dopant_shells[s] = self.get_shell_based_neighbors(s, depth)
                dopant_types[s] = dtype
        return doping_mod.site_wise_shell_indices(
            dopant_types, dopant_shells, self._element_type_of,
            lambda e: doping_mod.sublattice_sites(
                self.material, self.system.num_cells, e),
            max_shells, spec.dopant_element_types)

    def preproduction(self, dst_path, random_seed):
        """Seed each trajectory and dump its generator state; with doping,
        place the dopants of each trajectory and write its shell table
        `site_indices.npy` and `PreProduction.log` (PyCD/core.py:2566-2645).
        """
        assert dst_path, ("Please provide the destination path where "
                          "simulation output files needs to be saved")
        dst_path = Path(dst_path)
        rnd.seed(random_seed)
        seeds = [rnd.random() for _ in range(self.n_traj)]
        for k in range(self.n_traj):
            traj_dir = dst_path.joinpath(f"traj{k + 1}")
            Path.mkdir(traj_dir, parents=True, exist_ok=True)
            rnd.seed(seeds[k])
            with open(traj_dir.joinpath("initial_rnd_state.dump"), "wb") as fh:
                pickle.dump(rnd.getstate(), fh)
        if not self.doping_active:
            return None
        required = list(self.doping["min_shell_separation"])
        repo = {}
        sub_prefix = None
        dopant_site_indices = None
        for k in range(self.n_traj):
            traj_dir = dst_path.joinpath(f"traj{k + 1}")
            prefix = [f"Trajectory {k + 1}:\n"]
            repo[k] = {}
            old = [-1] * len(self.doping["num_dopants"])
            attempt = 1
            # the reference compares these lists as Python lists
            while (list(old) < required
                   and attempt <= self.doping["max_attempts"]):
                trial = self.get_doping_distribution()
                trial_prefix, new = self.get_doping_analysis(trial, [])
                if list(new) > list(old):
                    unique = all(
                        not (set(sites) == set(trial[t]) and t != "X")
                        for prev in repo.values()
                        for t, sites in prev.items())
                    if unique:
                        sub_prefix = list(trial_prefix)
                        old = new
                        dopant_site_indices = {t: list(v)
                                               for t, v in trial.items()}
                attempt += 1
            if dopant_site_indices is None:
                raise RuntimeError("no dopant distribution met the minimum "
                                   "shell separation")
            prefix.extend(sub_prefix)
            repo[k] = {t: list(v) for t, v in dopant_site_indices.items()}
            table, overlap = self.get_site_wise_shell_indices_table(
                dopant_site_indices)
            prefix.append("All shell based neighbor sites are independent\n\n"
                          if overlap else
                          "All shell based neighbor sites are NOT "
                          "independent\n\n")
            np.save(traj_dir / "site_indices", table)
            generate_report(self.start_time, traj_dir, "PreProduction", 0,
                            "".join(prefix))
        return None

    def doped_inputs(self, traj_dir):
        """Site energies, lattice charges and dopant sites of one doped
        trajectory from its `site_indices.npy` (PyCD/core.py:2709-2764)."""
        table = np.load(Path(traj_dir) / "site_indices.npy")
        energy, dopant_sites = doping_mod.doped_sites(
            table, self.doping_spec, self.undoped_system_relative_energies,
            self.relative_energies["doping"])
        return energy, self.lattice_charges(dopant_sites)[:, 0], dopant_sites

    # -- main loop ----------------------------------------------------------
    def _kmc_problem(self):
        return _native.KmcProblem(
            table=self.precomputed_array,
            lattice_q=self.base_charge_config()[:, 0],
            nbr_count=self.nbr_count, nbr_site=self.nbr_site,
            nbr_vec=self.nbr_vec, nbr_lambda=self.nbr_lambda,
            nbr_vab=self.nbr_vab, site_energy=self.system_relative_energies,
            carrier_q=np.asarray(self.species_charge_list, dtype=np.float64),
            field=self.electric_field if self.electric_field_active
            else np.zeros(3),
            nu=self.material.vn, kT=self.temp,
            time_interval=self.time_interval,
            n_path=int(self.t_final / self.time_interval) + 1,
            # a fully periodic supercell: the loop may gather from the table
            # rows of cell 0 (checked against the full table on the device)
            cells=self.system_size if np.all(self.neighbors.pbc) else None)

    def do_kmc_steps(self, dst_path, output_data, random_seed, compute_mode,
                     max_hops=None, record=False):
        """Run every trajectory on the GPU (PyCD/core.py:2647-2917).

        Writes traj{i}/ .npy outputs with the reference's names and append
        mode and `Run.log`.  `max_hops`/`record` are extensions used by
        tests and benchmarks: stop after that many hops, and keep the
        chosen process index of each hop in `self.last_records`.
        """
        assert dst_path, ("Please provide the destination path where "
                          "simulation output files needs to be saved")
        dst_path = Path(dst_path)
        if compute_mode != "parallel":
            self.preproduction(dst_path, random_seed)
        n_path = int(self.t_final / self.time_interval) + 1
        drift = np.zeros((self.n_traj, self.total_species, 3))
        system_charge = np.dot(self.species_count,
                               self.material.species_charge_list[
                                   self.species_charge_type])
        ewald_neut = -(np.pi * (system_charge ** 2)
                       / (2 * self.system.system_volume * self.system.alpha))
        problem = self._kmc_problem()
        session = _native.KmcSession(problem)
        self.last_records = []
        try:
            for k in range(self.n_traj):
                traj_dir = (dst_path.joinpath(f"traj{k + 1}")
                            if compute_mode != "parallel" else dst_path)
                Path.mkdir(traj_dir, parents=True, exist_ok=True)
                with open(traj_dir.joinpath("initial_rnd_state.dump"),
                          "rb") as fh:
                    rnd.setstate(pickle.load(fh))
                dopant_sites = {}
                if self.doping_active:
                    energy, lattice_q, dopant_sites = self.doped_inputs(
                        traj_dir)
                    self.system_relative_energies = energy
                    session.set_sites(lattice_q, energy)
                occupancy = self.generate_initial_occupancy(dopant_sites)
                stream = rng.UniformStream(rnd.getstate())
                # a doped run saves the occupancy after every hop, which is
                # rebuilt from the chosen processes
                res = session.run(occupancy, stream, ewald_neut=ewald_neut,
                                  max_hops=max_hops,
                                  record=record or self.doping_active)
                rnd.setstate(stream.state_after(2 * res["hops"]))
                drift[k] = res["drift"]
                if self.doping_active:
                    np.save(traj_dir / "occupancy.npy",
                            self.occupancy_rows(occupancy, res["procs"]))
                if record:
                    self.last_records.append(res)
                self._write_outputs(traj_dir, output_data, res, n_path)
        finally:
            session.close()
        prefix = []
        if self.electric_field_active:
            prefix = self.compute_drift_mobility(drift, dst_path, prefix)
        generate_report(self.start_time, dst_path, "Run", 1, "".join(prefix))
        return None

    def occupancy_rows(self, occupancy, procs):
        """Occupancy before the first hop and after every hop, (hops + 1)
        x n_c, as a doped run saves it (PyCD/core.py:2772, 2818, 2903-2906).
        Process p is slot p % z of carrier p // z (every hopping site has
        z slots, PyCD/core.py:1957-1978)."""
        occ = np.asarray(occupancy, dtype=np.int64)
        procs = np.asarray(procs, dtype=np.int64)
        nc, hops = len(occ), len(procs)
        z = int(self.nbr_count[occ[0]])
        owner = procs // z
        slot = procs - owner * z
        nbr = self.nbr_site
        cur = occ.tolist()
        new_site = np.empty(hops, dtype=np.int64)
        for h, (c, j) in enumerate(zip(owner.tolist(), slot.tolist())):
            cur[c] = int(nbr[cur[c], j])
            new_site[h] = cur[c]
        rows = np.empty((hops + 1, nc), dtype=int)
        step = np.arange(hops + 1)
        for c in range(nc):
            # forward fill of carrier c's site from the hops that move it
            col = np.empty(hops + 1, dtype=np.int64)
            col[0] = occ[c]
            moved = np.zeros(hops + 1, dtype=bool)
            moved[0] = True
            hit = np.nonzero(owner == c)[0]
            col[hit + 1] = new_site[hit]
            moved[hit + 1] = True
            last = np.maximum.accumulate(np.where(moved, step, 0))
            rows[:, c] = col[last]
        return rows

    def _write_outputs(self, traj_dir, output_data, res, n_path):
        arrays = {"unwrapped_traj": res["unwrapped"],
                  "time": res["time_data"],
                  "wrapped_traj": np.zeros((n_path, 3 * self.total_species)),
                  "energy": res["energy"],
                  "delg_0": res["delg_0"],
                  "potential": np.zeros((n_path, self.total_species))}
        for kind, attrs in output_data.items():
            if kind not in arrays or not attrs.get("write"):
                continue
            with open(traj_dir.joinpath(attrs["file_name"]), "ab") as fh:
                np.save(fh, arrays[kind])

    def compute_drift_mobility(self, drift_velocity_array, dst_path,
                               prefix_list):
        """Drift mobility from sum(hop * k) (PyCD/core.py:2052-2082)."""
        mob_au = (np.dot(drift_velocity_array, self.electric_field)
                  / self.electric_field_mag ** 2)
        mob = mob_au * (constants.BOHR2CM ** 2 * constants.SEC2AUTIME
                        * constants.V2AUPOT)
        with open(Path(dst_path).joinpath("drift_mobility.dat"), "wb") as fh:
            np.savetxt(fh, mob)
        start = 0
        for s_index, count in enumerate(self.species_count):
            end = start + count
            if count != 0:
                stype = self.material.species_types[s_index]
                per_traj = np.mean(mob[:, start:end], axis=1)
                prefix_list.append(
                    f"Estimated value of {stype} drift mobility is: "
                    f"{np.mean(per_traj):4.3e} cm2/V.s.\n")
                prefix_list.append(
                    f"Standard error of mean in {stype} drift mobility is: "
                    f"{np.std(per_traj) / np.sqrt(self.n_traj):4.3e} "
                    f"cm2/V.s.\n")
            start = end
        return prefix_list


class Analysis(object):
    """MSD analysis (PyCD/core.py:2920-3081) over the HIP MSD kernel."""

    def __init__(self, material_info, n_dim, species_count, n_traj, t_final,
                 time_interval, msd_t_final, trim_length, temp, repr_time="ns",
                 repr_dist="Angstrom"):
        self.start_time = datetime.now()
        self.material = material_info
        self.n_dim = n_dim
        self.species_count = np.asarray(species_count)
        self.total_species = int(self.species_count.sum())
        self.n_traj = int(n_traj)
        self.t_final = t_final * constants.SEC2AUTIME
        self.time_interval = time_interval * constants.SEC2AUTIME
        self.trim_length = trim_length
        self.temp = temp
        self.num_path_steps_per_traj = (int(self.t_final / self.time_interval)
                                        + 1)
        self.repr_time = repr_time
        self.repr_dist = repr_dist
        time_conv = {"ns": constants.AUTIME2NS, "ps": constants.AUTIME2PS,
                     "fs": constants.AUTIME2FS,
                     "s": 1E+00 / constants.SEC2AUTIME}
        if repr_time in time_conv:
            self.time_conversion = time_conv[repr_time]
        # Only the lower-case spellings are recognised, as in
        # PyCD/core.py:2961-2966 (the default 'Angstrom' sets nothing).
        dist_conv = {"m": constants.BOHR, "um": constants.BOHR2UM,
                     "angstrom": 1E+00 / constants.ANG2BOHR}
        if repr_dist in dist_conv:
            self.dist_conversion = dist_conv[repr_dist]
        self.kBT = constants.KB * self.temp / constants.EV2J
        self.msd_t_final = msd_t_final / self.time_conversion
        self.num_msd_steps_per_traj = int(self.msd_t_final
                                          / self.time_interval) + 1

    def msd_from_positions(self, position_array):
        """Per-trajectory, per-carrier squared displacement on the GPU.

        `position_array` is (n_traj*P, n_c, 3) in output units; returns
        sd of shape (n_traj, M, n_c) with sd[:, 0] = 0 (core.py:2996-3007).
        """
        P = self.num_path_steps_per_traj
        pos = np.ascontiguousarray(position_array).reshape(
            self.n_traj, P, self.total_species, 3)
        return _native.msd(pos, self.num_msd_steps_per_traj)

    def compute_msd(self, dst_path, output_data):
        """MSD, SEM and D/kBT report (PyCD/core.py:2974-3081)."""
        assert dst_path, ("Please provide the destination path where MSD "
                          "output files needs to be saved")
        dst_path = Path(dst_path)
        name = output_data["unwrapped_traj"]["file_name"]
        parts = [np.load(dst_path / f"traj{k + 1}" / name)
                 for k in range(self.n_traj)]
        position_array = np.vstack(parts)
        P = self.num_path_steps_per_traj
        position_array = (
            position_array[:self.n_traj * P + 1].reshape(
                (self.n_traj * P, self.total_species, 3))
            * self.dist_conversion)
        sd_array = self.msd_from_positions(position_array)
        return self.reduce_msd(sd_array, dst_path)

    def reduce_msd(self, sd_array, dst_path):
        """Species/trajectory means, SEM, slopes and the report."""
        n_species = (self.material.num_species_types
                     - list(self.species_count).count(0))
        M = self.num_msd_steps_per_traj
        avg = np.zeros((self.n_traj, M, n_species))
        start = 0
        skipped = 0
        missing = []
        for s_index in range(self.material.num_species_types):
            if self.species_count[s_index] != 0:
                end = start + self.species_count[s_index]
                avg[:, :, s_index - skipped] = np.mean(
                    sd_array[:, :, start:end], axis=2)
                start = end
            else:
                skipped += 1
                missing.append(s_index)
        msd_data = np.zeros((M, n_species + 1))
        msd_data[:, 0] = (np.arange(M) * self.time_interval
                          * self.time_conversion)
        msd_data[:, 1:] = np.mean(avg, axis=0)
        sem_data = np.std(avg, axis=0) / np.sqrt(self.n_traj)
        # The reference's n_traj clause compares n_traj with itself and is
        # therefore always empty (PyCD/core.py:3037-3041).
        file_name = (("%1.2E" % (self.msd_t_final * self.time_conversion))
                     + str(self.repr_time) + "_trim=" + str(self.trim_length))
        np.save(dst_path.joinpath("".join(["MSD_Data_", file_name, ".npy"])),
                msd_data)
        species_types = [t for i, t in enumerate(self.material.species_types)
                         if i not in missing]
        trim = self.trim_length
        slopes = np.zeros((self.n_traj, n_species))
        lines = []
        conv = constants.ANG2CM ** 2 * constants.SEC2NS / (2 * self.n_dim)
        for s_index, stype in enumerate(species_types):
            x = msd_data[trim:-trim, 0]
            for k in range(self.n_traj):
                slopes[k, s_index] = least_squares_slope(
                    x, avg[k, trim:-trim, s_index])
            slope = np.mean(slopes[:, s_index])
            lines.append(f"Estimated value of {stype} diffusivity is: "
                         f"{slope * conv / self.kBT:.3e} cm2/Vs\n")
            sem = np.std(slopes[:, s_index]) / np.sqrt(self.n_traj)
            lines.append(f"Standard error of mean in {stype} diffusivity is: "
                         f"{sem * conv / self.kBT:.3e} cm2/Vs\n")
        self.slope_data = slopes
        report = "".join(["MSD_Analysis", ("_" if file_name else ""),
                          file_name])
        generate_report(self.start_time, dst_path, report, 1, "".join(lines))
        return ReturnValues(msd_data=msd_data, sem_data=sem_data,
                            species_types=species_types, file_name=file_name)


def least_squares_slope(x, y):
    """Ordinary least-squares slope (what scipy.stats.linregress returns as
    `slope`, used at PyCD/core.py:3055-3060): cov(x, y) / var(x)."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    xm = x - x.mean()
    ym = y - y.mean()
    return float(np.dot(xm, ym) / np.dot(xm, xm))
