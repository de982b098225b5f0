"""Oracle: hop neighbour lists (test infrastructure only).

Restates Neighbors.hop_neighbor_sites / generate_neighbor_list
(PyCD/core.py:485-569, 596-668): for every centre, the targets whose
minimum-image distance lies in (cutoff - tol, cutoff + tol], in ascending
target order, with their minimum-image vectors; one such list per
(element-pair key, class, hop-distance type).  Pinned by the trajectory
replays of the shipped examples (tests/test_oracle_golden.py), which use
these lists.
"""
import types

import numpy as np

from oracle.ewald import mic_rows, simulation_cell


def hop_window(coords, centers, targets, lo, hi, cell, pbc):
    """Per centre: (ascending positions into targets, minimum-image vectors)."""
    idx, vecs = [], []
    tpos = coords[targets]
    for c in centers:
        v = mic_rows(tpos - coords[c], cell, pbc)
        r = np.linalg.norm(v, axis=1)
        hit = np.nonzero((lo < r) & (r <= hi))[0]
        idx.append(hit)
        vecs.append(v[hit])
    return idx, vecs


def hop_neighbor_list(material, neighbors):
    """The reference's hop_neighbor_list dictionary for a Neighbors object."""
    out = {}
    cell = simulation_cell(material.lattice_matrix, neighbors.system_size)
    coords = neighbors.bulk_sites.cell_coordinates
    tol = material.neighbor_cutoff_dist_tol
    for key, per_class in material.neighbor_cutoff_dist.items():
        center_type = key.split(material.element_type_delimiter)[0]
        t_index = material.element_types.index(center_type)
        sites = neighbors.hopping_site_indices(t_index)
        classes = []
        for c, cutoffs in enumerate(per_class):
            shells = []
            for t, cut in enumerate(cutoffs):
                lo, hi = cut - tol[key][c][t], cut + tol[key][c][t]
                idx, vecs = hop_window(coords, sites, sites, lo, hi, cell,
                                       neighbors.pbc)
                nbr = np.empty(len(idx), dtype=object)
                disp = np.empty(len(idx), dtype=object)
                for k in range(len(idx)):
                    nbr[k] = sites[idx[k]]
                    disp[k] = vecs[k]
                shells.append(types.SimpleNamespace(
                    neighbor_system_element_indices=nbr,
                    displacement_vector_list=disp,
                    num_neighbors=np.array([len(v) for v in idx])))
            classes.append(shells)
        out[key] = classes
    return out
