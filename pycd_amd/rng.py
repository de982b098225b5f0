"""The reference's uniform stream, generated in bulk on the host.

PyCD draws its KMC uniforms from the module-global stdlib `random`
(MT19937; `random()` = ((a >> 5) * 2**26 + (b >> 6)) / 2**53 from two
consecutive 32-bit outputs a, b), seeded per trajectory by `preproduction`
(PyCD/core.py:2566-2580) and consumed two per hop: u1 selects the process,
u2 advances the clock (PyCD/core.py:2799-2803).

`UniformStream` moves a stdlib generator state into NumPy's MT19937 (same
recurrence and tempering, same 624-word key and position) so the same
stream comes out a million draws at a time, and can hand the stdlib state
back after any number of draws.
"""
import numpy as np

_N = 624
_SCALE_HI = 67108864.0                 # 2**26
_INV_2_53 = 1.0 / 9007199254740992.0   # 2**-53


def _to_numpy_state(py_state):
    version, internal, _ = py_state
    if version != 3 or len(internal) != _N + 1:
        raise ValueError("expected a CPython random state (version 3)")
    key = np.array(internal[:_N], dtype=np.uint32)
    return {"bit_generator": "MT19937",
            "state": {"key": key, "pos": int(internal[_N])}}


def _to_python_state(np_state):
    st = np_state["state"]
    return (3, tuple(int(v) for v in st["key"]) + (int(st["pos"]),), None)


def words_to_uniforms(words):
    """Pairs of 32-bit outputs -> the 53-bit doubles of random.random()."""
    w = words.reshape(-1, 2)
    a = (w[:, 0] >> np.uint64(5)).astype(np.float64)
    b = (w[:, 1] >> np.uint64(6)).astype(np.float64)
    return (a * _SCALE_HI + b) * _INV_2_53


class UniformStream:
    """Bulk uniforms that continue a stdlib `random` state exactly."""

    def __init__(self, py_state):
        self._start = _to_numpy_state(py_state)
        self._bg = np.random.MT19937()
        self._bg.state = self._start
        self.drawn = 0

    def take(self, n):
        """The next `n` uniforms as float64."""
        words = self._bg.random_raw(2 * int(n))
        self.drawn += int(n)
        return words_to_uniforms(np.asarray(words, dtype=np.uint64))

    def state_after(self, n):
        """The stdlib state after exactly `n` uniforms from the start."""
        bg = np.random.MT19937()
        bg.state = self._start
        left = 2 * int(n)
        while left > 0:
            step = min(left, 1 << 24)
            bg.random_raw(step)
            left -= step
        return _to_python_state(bg.state)

    @staticmethod
    def mt_state_words(py_state):
        """(key as uint32[624], pos) of a stdlib state, for device MT."""
        st = _to_numpy_state(py_state)["state"]
        return st["key"], st["pos"]


def seeded_starts(random_seed, n_traj, hop_sites, n_carriers):
    """Per-trajectory (occupancy, MT words, position) for a batch.

    `hop_sites[t]` and `n_carriers[t]` are trajectory t's hopping sites (in
    system order) and carrier count.

    Follows the reference's seeding (PyCD/core.py:2573-2580): seed the
    stdlib generator with `random_seed`, draw one float per trajectory,
    reseed with it; then the initial carriers are
    `sample(hop_sites[t], n_carriers[t])` (PyCD/core.py:2527).  The state
    after the sample is what the trajectory's hops consume.
    """
    import random as _random
    gen = _random.Random()
    gen.seed(random_seed)
    floats = [gen.random() for _ in range(n_traj)]
    occs, words = [], np.zeros((n_traj, _N + 1), dtype=np.uint32)
    for t in range(n_traj):
        gen.seed(floats[t])
        occs.append([int(v) for v in gen.sample(list(hop_sites[t]),
                                                int(n_carriers[t]))])
        key, pos = UniformStream.mt_state_words(gen.getstate())
        words[t, :_N] = key
        words[t, _N] = pos
    return occs, words
