"""Helpers for the KMC parity tests: seeded runs of the HIP path and oracle."""
import random

import numpy as np

from oracle import kmc, seeding
from pycd_amd import _native, rng
from tests.systems import oracle_problem


def native_problem(run, table, compact=True):
    """The plain path's problem; `compact`: give the supercell's shape, so
    that the hop loop gathers from the table rows of cell 0."""
    return _native.KmcProblem(
        table=table, lattice_q=run.base_charge_config()[:, 0],
        nbr_count=run.nbr_count, nbr_site=run.nbr_site, nbr_vec=run.nbr_vec,
        nbr_lambda=run.nbr_lambda, nbr_vab=run.nbr_vab,
        site_energy=run.system_relative_energies,
        carrier_q=np.asarray(run.species_charge_list, dtype=np.float64),
        field=run.electric_field if run.electric_field_active else np.zeros(3),
        nu=run.material.vn, kT=run.temp, time_interval=run.time_interval,
        n_path=int(run.t_final / run.time_interval) + 1,
        cells=run.system_size if compact else None)


def seeded_start(run, seed, n_carriers, traj=0, n_traj=1):
    """(generator after the occupancy draw, occupancy) for one trajectory."""
    state = seeding.trajectory_states(seed, n_traj)[traj]
    gen = random.Random()
    gen.setstate(state)
    occ = seeding.initial_occupancy(gen, run.hop_site_lists["electron"],
                                    n_carriers)
    return gen, occ


def run_gpu(run, table, seed, n_carriers, max_hops=None, record=True,
            session=None, chunk=1 << 16, compact=True):
    gen, occ = seeded_start(run, seed, n_carriers)
    sess = session or _native.KmcSession(native_problem(run, table, compact),
                                         chunk=chunk)
    stream = rng.UniformStream(gen.getstate())
    res = sess.run(occ, stream, max_hops=max_hops, record=record)
    if session is None:
        sess.close()
    return res, occ


def run_oracle(run, table, seed, n_carriers, max_hops=None):
    gen, occ = seeded_start(run, seed, n_carriers)
    return kmc.run(oracle_problem(run, table=table), occ, gen.random,
                   max_hops=max_hops, record=True)
