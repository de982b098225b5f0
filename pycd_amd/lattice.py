"""Host-side lattice preparation for the KMC path (runs once per system).

This is the input producer of the hot path: it turns a POSCAR and the
material YAML into site coordinates (bohr), the minimum-image cell and the
per-site hop tables that the HIP kernels consume.  It follows

  * Material.__init__ / generate_sites   PyCD/core.py:42-241
  * minimum image                       PyCD/core.py:292-361

The O(N^2) neighbour search and minimum-image tables run on the GPU
(pycd_hop_neighbors / pycd_min_image_vectors); the class below is the
single-vector form used by the API's small helpers.
"""
import numpy as np


def z_sorted_unit_cell(fractional, element_index_list, n_per_type,
                       class_list, element_types):
    """Per-element ascending-z order of the unit cell (core.py:85-97).

    Uses NumPy's default argsort, as the reference does; z ties keep
    whatever order that sort gives (reference behaviour).
    """
    out = np.zeros(fractional.shape)
    classes = []
    start = 0
    for t, etype in enumerate(element_types):
        block = fractional[element_index_list == t]
        order = block[:, 2].argsort()
        end = start + n_per_type[t]
        out[start:end] = block[order]
        classes.extend(class_list[etype][k] - 1 for k in order)
        start = end
    return out, classes


def supercell_sites(cart_unit, lattice, cell_size, element_type_indices,
                    element_index_list, n_per_type):
    """Site coordinates, quantum indices and system indices (core.py:154-241).

    Sites are ordered cell-major (x slowest, z fastest over cells) and, within
    a cell, by the unit-cell order of the selected element types.
    """
    cell_size = np.asarray(cell_size)
    sel = np.nonzero(np.isin(element_index_list, element_type_indices))[0]
    unit = cart_unit[sel]
    n_unit = int(n_per_type[list(element_type_indices)].sum())
    gx, gy, gz = np.meshgrid(np.arange(cell_size[0]), np.arange(cell_size[1]),
                             np.arange(cell_size[2]), indexing="ij")
    cells = np.column_stack((gx.ravel(), gy.ravel(), gz.ravel()))
    shift = cells @ lattice
    coords = (unit[np.newaxis, :, :] + shift[:, np.newaxis, :]).reshape(-1, 3)
    n_cells = len(cells)
    type_col = np.concatenate([[t] * n_per_type[t]
                               for t in element_type_indices])
    elem_col = np.concatenate([np.arange(n_per_type[t])
                               for t in element_type_indices])
    quantum = np.zeros((n_cells * n_unit, 5), dtype=int)
    quantum[:, :3] = np.repeat(cells, n_unit, axis=0)
    quantum[:, 3] = np.tile(type_col, n_cells)
    quantum[:, 4] = np.tile(elem_col, n_cells)
    system_index = np.arange(n_cells * n_unit)
    return coords, quantum, system_index


class MinimumImage:
    """Minimum-image convention of a (possibly skewed) periodic supercell.

    frac = d S^-1; base offset -round(frac) on periodic axes (round half to
    even); then the 26 neighbouring offsets are tried in (dx, dy, dz) loop
    order and a candidate wins only with a strictly smaller norm
    (PyCD/core.py:292-361).  S has rows system_size[i] * lattice[i].
    """

    def __init__(self, lattice_bohr, system_size, pbc):
        self.cell = lattice_bohr * np.asarray(system_size)[:, np.newaxis]
        self.cell_inv = np.linalg.inv(self.cell)
        self.mask = np.asarray(pbc, dtype=bool)
        deltas = []
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    if dx == 0 and dy == 0 and dz == 0:
                        continue
                    dlt = [dx if self.mask[0] else 0, dy if self.mask[1] else 0,
                           dz if self.mask[2] else 0]
                    if any(dlt):
                        deltas.append(dlt)
        self.deltas = np.array(deltas, dtype=np.float64).reshape(-1, 3)

    def apply(self, d):
        d = np.asarray(d, dtype=np.float64)
        shape = d.shape
        d = d.reshape(-1, 3)
        frac = d @ self.cell_inv
        base = np.zeros_like(frac)
        base[:, self.mask] = -np.round(frac[:, self.mask])
        best = (frac + base) @ self.cell
        best_n = np.sqrt(np.einsum("ij,ij->i", best, best))
        for delta in self.deltas:
            cand = (frac + (base + delta)) @ self.cell
            cand_n = np.sqrt(np.einsum("ij,ij->i", cand, cand))
            better = cand_n < best_n
            best[better] = cand[better]
            best_n[better] = cand_n[better]
        return best.reshape(shape)
