"""CPU oracle of the KMC hot path -- TEST INFRASTRUCTURE ONLY.

This package restates, in plain NumPy, the reference algorithm of PyCD's
hot path so that the HIP implementation in
`pycd_amd` can be checked against it.  It is imported only by `tests/`, by
`__graft_entry__.smoke()` and by the CPU-baseline leg of `bench.py`; the
product path never imports, links or calls anything here.

Modules and what pins each of them (reference file:line in each docstring):

  ewald   -- minimum-image vectors, real/reciprocal Ewald sums and the table
             assembly (PyCD/core.py:304-361, 708-726, 799-878, 1594-1661).
             Pinned by the shipped tables examples/{Hematite,BVO}/InputFiles/
             precomputed_array.npy (tests/golden), by a 27-image brute force
             of the minimum image, and by the closed-form Ewald identity
             E(lattice sum) = E(table) for a uniformly shifted charge set.
  kmc     -- the `Run.do_kmc_steps` loop (PyCD/core.py:1946-2050, 2647-2861)
             in NumPy, the drift mobility (:2052-2082) and the per-hop
             occupancy of doped runs.  Pinned by replaying the shipped
             trajectories examples/{Hematite,BVO}/traj1/{time_data,
             unwrapped_traj}.npy hop for hop, by the identity
             dE = E(after) - E(before), by detailed balance, and (mobility)
             by the CODATA atomic unit of mobility and hand arithmetic.
  msd     -- `Analysis.compute_msd` (PyCD/core.py:2974-3081): the lag loop,
             pinned by a brute-force double loop on tiny inputs and the
             ballistic closed form MSD(t) = v^2 t^2; the species/trajectory
             reduction, slopes and D/kBT, pinned by hand arithmetic on
             linear inputs, scipy.stats.linregress and the exact corundum
             diffusivity.
  diffusion -- exact D of one carrier on equivalent sites (sublattice
             corrector form), pinned by the one-sublattice closed form.
  neighbors -- hop neighbour lists, pinned by the shipped trajectory
             replays that use them.
  seeding -- the trajectory seeding protocol (PyCD/core.py:2566-2580).
             Pinned by byte equality with the shipped initial_rnd_state.dump.
"""
