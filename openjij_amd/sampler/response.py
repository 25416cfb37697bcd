"""Sample container returned by the samplers (openjij/sampler/response.py:22-37).

The reference wraps dimod's SampleSet; this keeps the attributes the
single-spin-flip API hands back: ``states``, ``energies``, ``record``,
``first``, ``variables``, ``vartype``, ``info`` and ``samples()``.
"""
from __future__ import annotations

from collections import namedtuple

import numpy as np

from openjij_amd.variable_type import cast_vartype

Sample = namedtuple("Sample", ["sample", "energy", "num_occurrences"])


class Response:
    def __init__(self, states, energies, variables, vartype, info=None):
        self.states = np.asarray(states)
        self.energies = np.asarray(energies, dtype=float)
        self.variables = list(variables)
        self.vartype = cast_vartype(vartype)
        self.info = dict(info or {})
        n = len(self.energies)
        rec = np.rec.fromarrays(
            [self.states.reshape(n, -1), self.energies, np.ones(n, dtype=int)],
            dtype=[("sample", self.states.dtype, (len(self.variables),)),
                   ("energy", float), ("num_occurrences", int)],
        ) if n else np.rec.array(np.zeros(0, dtype=[("sample", int, (len(self.variables),)),
                                                     ("energy", float),
                                                     ("num_occurrences", int)]))
        self.record = rec

    @classmethod
    def from_samples(cls, states, vartype, energies, variables=None, info=None):
        states = np.asarray(states)
        if variables is None:
            variables = list(range(states.shape[1] if states.ndim == 2 else 0))
        return cls(states, energies, variables, vartype, info)

    def __len__(self):
        return len(self.energies)

    @property
    def data_vectors(self):
        return {"energy": self.energies, "num_occurrences": np.ones(len(self), dtype=int)}

    def samples(self):
        return [dict(zip(self.variables, (int(x) for x in row))) for row in self.states]

    @property
    def first(self) -> Sample:
        if len(self) == 0:
            raise ValueError("empty response")
        k = int(np.argmin(self.energies))
        return Sample(dict(zip(self.variables, (int(x) for x in self.states[k]))),
                      float(self.energies[k]), 1)

    def lowest(self):
        k = np.flatnonzero(self.energies == self.energies.min())
        return Response(self.states[k], self.energies[k], self.variables, self.vartype, self.info)
