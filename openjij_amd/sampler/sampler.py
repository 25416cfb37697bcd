"""Shared sampler front end (openjij/sampler/sampler.py:166-248)."""
from __future__ import annotations

import numpy as np

from openjij_amd.model.model import BinaryQuadraticModel
from openjij_amd.variable_type import BINARY, SPIN


class BaseSampler:
    parameters: dict = {}

    def _set_params(self, **kwargs):
        for key, value in kwargs.items():
            self._params[key] = self._default_params[key] if value is None else value

    def sample_ising(self, h, J, **kwargs):
        """Ising sampling; ``sparse`` defaults to True as in the reference
        (openjij/sampler/sampler.py:206-222)."""
        sparse = kwargs.pop("sparse", True)
        bqm = BinaryQuadraticModel.from_ising(h, J, sparse=sparse)
        return self.sample(bqm, sparse=sparse, **kwargs)

    def sample_qubo(self, Q, **kwargs):
        sparse = kwargs.pop("sparse", True)
        if isinstance(Q, np.ndarray):
            bqm = BinaryQuadraticModel.from_numpy_matrix(Q, vartype=BINARY, sparse=sparse)
        else:
            bqm = BinaryQuadraticModel.from_qubo(Q, sparse=sparse)
        return self.sample(bqm, sparse=sparse, **kwargs)


def resolve_seed(seed):
    if seed is None:
        return int(np.random.SeedSequence().generate_state(1, dtype=np.uint64)[0])
    return int(seed) & 0xFFFFFFFFFFFFFFFF


def initial_states(num_reads, n, seed, rng_tag):
    """num_reads x n random +-1 states.  With a seed they are reproducible and
    differ between reads (the reference repeats gen_spin(seed) for every read,
    openjij/sampler/sa_sampler.py:251-258, so its seeded reads coincide)."""
    rng = np.random.default_rng([rng_tag, seed])
    return (2 * rng.integers(0, 2, size=(num_reads, n)) - 1).astype(np.int8)


def to_vartype_states(spins, vartype):
    spins = np.asarray(spins)
    if vartype == BINARY:
        return ((spins + 1) // 2).astype(int)
    return spins.astype(int)


def parse_initial_state(initial_state, bqm, n):
    vals = []
    if isinstance(initial_state, dict):
        seq = [initial_state[k] for k in bqm.variables]
    elif isinstance(initial_state, (list, tuple, np.ndarray)):
        seq = list(initial_state)
    else:
        raise RuntimeError("Unsupported type of initial_state.")
    if len(seq) != n:
        raise ValueError("the size of the initial state should be {}".format(n))
    for v in seq:
        if bqm.vartype == BINARY:
            if v != 0 and v != 1:
                raise RuntimeError("The initial variables must be 0 or 1.")
            vals.append(2 * v - 1)
        else:
            if v != -1 and v != 1:
                raise RuntimeError("The initial variables must be -1 or +1.")
            vals.append(v)
    return np.array(vals, dtype=np.int8)


__all__ = ["BaseSampler", "SPIN", "BINARY"]
