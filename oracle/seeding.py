"""Oracle: per-trajectory seeding and initial occupancy (test infrastructure).

Restates PyCD/core.py:2566-2580 (`preproduction`): seed the generator with
`random_seed`, draw one float per trajectory, and reseed with that float.
The initial carrier sites are `sample(hopping sites in system order, n)`
(PyCD/core.py:2478-2528, no dopants, no preset occupancy).
"""
import random


def trajectory_states(random_seed, n_traj):
    gen = random.Random()
    gen.seed(random_seed)
    floats = [gen.random() for _ in range(n_traj)]
    states = []
    for f in floats:
        gen.seed(f)
        states.append(gen.getstate())
    return states


def initial_occupancy(gen, hop_sites, n_carriers):
    return gen.sample(list(hop_sites), n_carriers)
