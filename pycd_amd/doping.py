"""Dopants: where they go, and what they change in the hop loop (host side).

A doped trajectory differs from an undoped one in three per-site inputs of
the hop loop, which the kernels take as arrays:

* lattice charges: a dopant site carries the dopant's charge
  (PyCD/core.py:2540-2551);
* site energies: a site k hop shells away from its nearest dopant is
  shifted by shifts[k] eV when k < len(shifts) (PyCD/core.py:2723-2761);
* initial occupancy: with site_charge_initiation 'yes' the carriers are
  first sampled from the dopant sites (PyCD/core.py:2478-2497).

Placement follows PyCD/core.py:2084-2476 and 2584-2645: manual or random
insertion, shells by hop distance, and the per-site shell table
(`site_indices.npy`: site, dopant type, dopant, shell).  Random insertion
draws from the module-global `random` generator, as the reference does,
so the same seed places the same dopants.
"""
import random as rnd

import numpy as np

from . import constants


class DopingSpec(object):
    """Doping parameters of a run (PyCD/core.py:1763-1781)."""

    def __init__(self, doping, material):
        self.params = doping or {}
        self.active = bool(np.any(self.params.get("num_dopants", [0])))
        self.dopant_element_types = []
        self.substitution_element_types = []
        self.dopant_species_types = []
        self.dopant_to_substitution_element_type_map = {}
        if not self.active:
            return
        for entry in self.params["doping_element_map"]:
            sub, dop = entry.split(material.element_type_delimiter)
            self.dopant_element_types.append(dop)
            self.substitution_element_types.append(sub)
            self.dopant_species_types.append(
                material.element_type_to_species_map[sub][0])
            self.dopant_to_substitution_element_type_map[dop] = sub
        for k, kind in enumerate(self.params["insertion_type"]):
            if self.params["num_dopants"][k] and kind not in ("manual",
                                                              "random"):
                raise NotImplementedError(
                    f"{kind!r} dopant insertion is not implemented (manual "
                    f"and random are)")


def shells(center, n_shells, neighbors):
    """Sites by hop distance from `center`: shells 0..n_shells.

    Shell 0 is [center]; shell k holds the hop targets of shell k-1 that are
    in no earlier shell, each once (PyCD/core.py:2352-2389; the reference
    walks every inner site, but targets of shells before k-1 all lie in
    earlier shells, so only shell k-1 contributes).  Within a shell the
    order is that of `list(set(...))` over the targets in walk order, the
    reference's own.
    """
    out = [list({center})]
    seen = {center}
    for _ in range(n_shells):
        walk = []
        for s in out[-1]:
            walk.extend(neighbors(s))
        shell = list(set([x for x in walk if x not in seen]))
        seen.update(shell)
        out.append(shell)
    return out


def all_shells(center, neighbors):
    """Shells from `center` up to the last non-empty one."""
    out = [[center]]
    seen = {center}
    while True:
        walk = []
        for s in out[-1]:
            walk.extend(neighbors(s))
        shell = list(set([x for x in walk if x not in seen]))
        if not shell:
            return out
        seen.update(shell)
        out.append(shell)


class ShellLookup(object):
    """Distance -> shell index of one sublattice (PyCD/core.py:1783-1821).

    From a sample site, the distance in Angstrom of every site of every
    shell, sorted; bins at the midpoints (0.5 A beyond the ends).  A
    distance d falls in bin digitize(d) - 1, whose site's shell is the
    lookup's answer.
    """

    def __init__(self, sample, neighbors, distance, tol_dist=0.5):
        dist, shell_of = [], []
        found = all_shells(sample, neighbors)
        for k, sh in enumerate(found):
            for s in sh:
                dist.append(distance(sample, s) / constants.ANG2BOHR)
                shell_of.append(k)
        order = np.argsort(dist)
        d = np.asarray(dist)[order]
        self.shell_index_list = np.asarray(shell_of)[order]
        mid = d[:-1] + (d[1:] - d[:-1]) / 2
        self.dist_bins = np.concatenate([[d[0] - tol_dist], mid,
                                         [d[-1] + tol_dist]])
        self.max_shells = len(found) - 1

    def shell(self, dist_ang):
        return self.shell_index_list[np.digitize(dist_ang,
                                                 self.dist_bins) - 1]


def sublattice_sites(material, n_cells, element_type_index):
    """System indices of one element type, cell-major (core.py:2129-2141)."""
    tot = material.total_elements_per_unit_cell
    n_e = material.n_elements_per_unit_cell[element_type_index]
    head = material.n_elements_per_unit_cell[:element_type_index].sum()
    offsets = np.repeat(np.arange(0, tot * n_cells, tot), n_e)
    return np.tile(head + np.arange(0, n_e), n_cells) + offsets


def random_insertion(available, n_dopants, n_discard, shells_of):
    """Draw dopant sites one by one; each takes its shells 0..n_discard out
    of the available sites (PyCD/core.py:2084-2107).  Returns (sites,
    available left)."""
    chosen = []
    while (n_dopants - len(chosen)) and available:
        site = rnd.choice(available)
        chosen.append(site)
        gone = set(s for sh in shells_of(site, n_discard) for s in sh)
        available = [s for s in available if s not in gone]
    return chosen, available


def site_wise_shell_indices(dopant_types, dopant_shells, element_type_of,
                            sublattice, max_shells, dopant_element_types):
    """Per-site (site, dopant type index, dopant index, shell) table.

    `dopant_shells` maps each dopant site (in insertion order) to its shells;
    `dopant_types` maps it to its dopant element type.  Every site of a
    dopant's sublattice starts at shell max_shells + 1 and takes the
    smallest shell any dopant reaches it in, the earlier dopant on ties
    (PyCD/core.py:2414-2476).  A site no dopant reaches has no dopant type,
    which the reference reports as a ValueError; so does this function.
    Returns (array, independent): `independent` is the flag the reference
    reports (it is set when a shell site changes element type).
    """
    sites, dop_index, shell, shell_type = [], [], [], []
    position = {}
    seen_types = []
    overlap = 0
    for d_index, d_site in enumerate(dopant_shells):
        etype_index, etype = element_type_of(d_site)
        if etype_index not in seen_types:
            seen_types.append(etype_index)
            block = sublattice(etype_index)
            for s in block:
                position[int(s)] = len(sites)
                sites.append(int(s))
            shell.extend([max_shells[etype] + 1] * len(block))
            dop_index.extend([-1] * len(block))
            shell_type.extend([etype] * len(block))
        for k, members in enumerate(dopant_shells[d_site]):
            for s in members:
                i = position[s]
                if k < shell[i]:
                    shell[i] = k
                    dop_index[i] = d_index
                    if shell_type[i] != etype:
                        overlap = 1
                    shell_type[i] = dopant_types[d_site]
    type_index = []
    for t in shell_type:
        if t not in dopant_element_types:
            raise ValueError(f"a site of element type {t!r} lies in no "
                             f"dopant's shells")
        type_index.append(dopant_element_types.index(t))
    table = np.hstack((np.asarray(sites)[:, None],
                       np.asarray(type_index)[:, None],
                       np.asarray(dop_index)[:, None],
                       np.asarray(shell)[:, None]))
    return table, overlap


def analysis_report(spec, dopant_sites, lookups, distance, coordinates):
    """Inter-dopant distances and shell separations (PyCD/core.py:2305-2350).

    Returns (report lines, min shell separation per dopant type).  As in the
    reference, a dopant type with one site returns the configured list
    itself, dropping what earlier types set.
    """
    lines = []
    min_sep = spec.params["min_shell_separation"][:]
    names = ["site1", "site2", "pairwise distance (ang.)", "shells apart"]
    widths = [len(x) for x in names]
    for dtype, members in dopant_sites.items():
        m = spec.dopant_element_types.index(dtype)
        sub = spec.substitution_element_types[m]
        n = len(members)
        tail = "s" if n > 1 else ""
        lines.append(f"Inserted {n} site{tail} of dopant element type "
                     f"{dtype}\n")
        ip = int(np.ceil(np.log10(n + 1)))
        tp = len(dtype)
        if n > 1:
            lines.append("\t".join(names) + "\n")
        best = None
        for i1, s1 in enumerate(members):
            for i2, s2 in enumerate(members[i1 + 1:]):
                dist = distance(s1, s2) / constants.ANG2BOHR
                sep = lookups[sub].shell(dist) - 1
                lines.append(
                    f"{dtype:>{widths[0] - ip}}{i1 + 1:0{tp}}\t"
                    f"{dtype:>{widths[1] - ip}}{i1 + i2 + 2:0{tp}}\t"
                    f"{dist:>{widths[2]}.2f}\t{sep:{widths[3]}}\n")
                if best is None or sep < best:
                    best = sep
        if n > 1:
            min_sep[m] = best
            lines.append(f"All dopant sites of element type '{dtype}' are "
                         f"separated by at least {min_sep[m]} shells\n")
            xyz = np.array([coordinates(s) for s in members])
            mean = np.mean(xyz, axis=0) / constants.ANG2BOHR
            std = np.std(xyz, axis=0) / constants.ANG2BOHR
            lines.append(f"Mean center of dopant sites of element type "
                         f"'{dtype}' is: ["
                         + "".join(f"{v:7.3f}" for v in mean) + "]\n")
            lines.append(f"Standard distance deviation of dopant sites of "
                         f"element type '{dtype}' is: ["
                         + "".join(f"{v:7.3f}" for v in std) + "]\n")
        else:
            min_sep = spec.params["min_shell_separation"]
    return lines, min_sep


def doped_sites(table, spec, undoped_energy, shifts_ev):
    """Site energies and dopant sites from a shell table (core.py:2723-2761).

    Returns (site energies in Hartree, {dopant type: [sites at shell 0]}),
    the dopant sites in table order.
    """
    energy = np.array(undoped_energy, dtype=np.float64, copy=True)
    dopant_sites = {}
    for row in np.asarray(table):
        site, t_index, shell = int(row[0]), int(row[1]), int(row[3])
        dtype = spec.dopant_element_types[t_index]
        if shell == 0:
            dopant_sites.setdefault(dtype, []).append(site)
        m = spec.dopant_element_types.index(dtype)
        sub = spec.substitution_element_types[m]
        shifts = shifts_ev[sub][m]
        if shell < len(shifts):
            energy[site] += shifts[shell] * constants.EV2HARTREE
    return energy, dopant_sites
