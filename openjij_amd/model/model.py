"""Binary quadratic model and its conversion to the engine's coupling layout.

Covers the ingestion format of the single-spin-flip path
(openjij/model/model.py:72-109): a SPIN-domain (N+1) x (N+1) upper-triangular
``interaction_matrix`` with h in column N and 1 in the corner, handed to
``cxxjij.graph.Dense`` or ``cxxjij.graph.CSRSparse``.  Variable order is the
sorted order of the labels, as the reference model keeps it.
"""
from __future__ import annotations

from typing import Dict, Hashable, Iterable, Tuple

import numpy as np

from openjij_amd._native import cxxjij
from openjij_amd.variable_type import BINARY, SPIN, Vartype, cast_vartype


def _sort_labels(labels: Iterable[Hashable]) -> list:
    labels = list(labels)
    try:
        return sorted(labels)
    except TypeError:
        return sorted(labels, key=lambda v: (type(v).__name__, str(v)))


class BinaryQuadraticModel:
    """E(x) = sum_i linear_i x_i + sum_{(i,j)} quadratic_ij x_i x_j + offset."""

    def __init__(self, linear: Dict, quadratic: Dict, *args, **kwargs):
        offset = kwargs.pop("offset", None)
        vartype = kwargs.pop("vartype", None)
        self.sparse = bool(kwargs.pop("sparse", False))
        if kwargs:
            raise TypeError(f"unexpected keyword arguments {sorted(kwargs)}")
        rest = list(args)
        if rest and not isinstance(rest[0], (str, Vartype, set, frozenset)):
            offset = rest.pop(0)
        if rest:
            vartype = rest.pop(0)
        if rest:
            raise TypeError("too many positional arguments")
        if vartype is None:
            raise ValueError("vartype must be given")
        self.vartype = cast_vartype(vartype)
        self.offset = float(offset) if offset is not None else 0.0
        self.linear: Dict = {}
        self.quadratic: Dict = {}
        for v, b in dict(linear).items():
            self.linear[v] = self.linear.get(v, 0.0) + float(b)
        for (u, v), b in dict(quadratic).items():
            if u == v:
                self.linear[u] = self.linear.get(u, 0.0) + float(b)
                continue
            self.linear.setdefault(u, 0.0)
            self.linear.setdefault(v, 0.0)
            key = (u, v) if (v, u) not in self.quadratic else (v, u)
            self.quadratic[key] = self.quadratic.get(key, 0.0) + float(b)
        self.variables = _sort_labels(self.linear.keys())
        self._index = {v: k for k, v in enumerate(self.variables)}
        self.model_type = "openjij.BinaryQuadraticModel"

    # -- constructors -------------------------------------------------------
    @classmethod
    def from_ising(cls, linear: Dict, quadratic: Dict, offset: float = 0.0, **kw):
        return cls(linear, quadratic, offset, SPIN, **kw)

    @classmethod
    def from_qubo(cls, Q: Dict, offset: float = 0.0, **kw):
        return cls({}, Q, offset, BINARY, **kw)

    @classmethod
    def from_numpy_matrix(cls, mat, variables=None, offset=0.0, vartype="BINARY", **kw):
        mat = np.asarray(mat, dtype=float)
        n = mat.shape[0]
        labels = list(range(n)) if variables is None else list(variables)
        lin = {labels[i]: mat[i, i] for i in range(n)}
        quad = {}
        for i in range(n):
            for j in range(i + 1, n):
                v = mat[i, j] + mat[j, i]
                if v != 0:
                    quad[(labels[i], labels[j])] = v
        return cls(lin, quad, offset, vartype, **kw)

    # -- accessors ----------------------------------------------------------
    @property
    def num_variables(self) -> int:
        return len(self.variables)

    @property
    def indices(self) -> list:
        return self.variables

    def get_linear(self) -> Dict:
        return dict(self.linear)

    def get_quadratic(self) -> Dict:
        return dict(self.quadratic)

    def get_offset(self) -> float:
        return self.offset

    def get_vartype(self) -> Vartype:
        return self.vartype

    def copy(self):
        return BinaryQuadraticModel(self.linear, self.quadratic, self.offset, self.vartype,
                                    sparse=self.sparse)

    # -- domain change ------------------------------------------------------
    def _converted(self, vartype: Vartype) -> Tuple[Dict, Dict, float]:
        if vartype == self.vartype:
            return dict(self.linear), dict(self.quadratic), self.offset
        lin = {v: 0.0 for v in self.linear}
        quad = {}
        off = self.offset
        if self.vartype == BINARY and vartype == SPIN:
            # x = (s + 1) / 2
            for v, b in self.linear.items():
                lin[v] += b / 2.0
                off += b / 2.0
            for (u, v), b in self.quadratic.items():
                quad[(u, v)] = b / 4.0
                lin[u] += b / 4.0
                lin[v] += b / 4.0
                off += b / 4.0
        else:
            # s = 2x - 1
            for v, b in self.linear.items():
                lin[v] += 2.0 * b
                off -= b
            for (u, v), b in self.quadratic.items():
                quad[(u, v)] = 4.0 * b
                lin[u] -= 2.0 * b
                lin[v] -= 2.0 * b
                off += b
        return lin, quad, off

    def change_vartype(self, vartype, inplace: bool = True):
        vt = cast_vartype(vartype)
        lin, quad, off = self._converted(vt)
        if not inplace:
            return BinaryQuadraticModel(lin, quad, off, vt, sparse=self.sparse)
        self.linear, self.quadratic, self.offset, self.vartype = lin, quad, off, vt
        return self

    # -- matrices and energies ----------------------------------------------
    def interaction_matrix(self) -> np.ndarray:
        """(N+1) x (N+1) upper-triangular matrix of the current domain, linear
        terms in column N, 1 in the corner (tests/test_model.py:57-68)."""
        n = self.num_variables
        m = np.zeros((n + 1, n + 1))
        for v, b in self.linear.items():
            m[self._index[v], n] += b
        for (u, v), b in self.quadratic.items():
            i, j = self._index[u], self._index[v]
            if i > j:
                i, j = j, i
            m[i, j] += b
        m[n, n] = 1.0
        return m

    def _vector(self, sample) -> np.ndarray:
        if isinstance(sample, dict):
            return np.array([sample[v] for v in self.variables], dtype=float)
        return np.asarray(sample, dtype=float)

    def energy(self, sample) -> float:
        x = self._vector(sample)
        e = self.offset
        for v, b in self.linear.items():
            e += b * x[self._index[v]]
        for (u, v), b in self.quadratic.items():
            e += b * x[self._index[u]] * x[self._index[v]]
        return e

    def calc_energy(self, sample, **kwargs) -> float:
        return self.energy(sample)

    def energies(self, samples) -> np.ndarray:
        """Vectorised energies of an (R, N) array in the current domain."""
        X = np.asarray(samples, dtype=float)
        if X.ndim == 1:
            X = X[None, :]
        n = self.num_variables
        h = np.zeros(n)
        for v, b in self.linear.items():
            h[self._index[v]] += b
        e = X @ h + self.offset
        if self.quadratic:
            ii = np.array([self._index[u] for (u, _) in self.quadratic])
            jj = np.array([self._index[v] for (_, v) in self.quadratic])
            bb = np.array(list(self.quadratic.values()))
            e = e + (X[:, ii] * X[:, jj]) @ bb
        return e

    def ising_arrays(self):
        """(h[N], J_upper[N,N], offset) of the SPIN form."""
        lin, quad, off = self._converted(SPIN)
        n = self.num_variables
        h = np.zeros(n)
        J = np.zeros((n, n))
        for v, b in lin.items():
            h[self._index[v]] += b
        for (u, v), b in quad.items():
            i, j = self._index[u], self._index[v]
            if i > j:
                i, j = j, i
            J[i, j] += b
        return h, J, off

    def get_cxxjij_ising_graph(self):
        """SPIN graph for the engine: ``CSRSparse`` when ``sparse`` else
        ``Dense`` (openjij/model/model.py:72-109); returns (graph, offset)."""
        h, J, off = self.ising_arrays()
        n = self.num_variables
        mat = np.zeros((n + 1, n + 1))
        mat[:n, :n] = J
        mat[:n, n] = h
        mat[n, n] = 1.0
        if self.sparse:
            import scipy.sparse as sp

            return cxxjij.graph.CSRSparse(sp.csr_matrix(mat)), off
        g = cxxjij.graph.Dense(n)
        g.set_interaction_matrix(mat)
        return g, off


BQM = BinaryQuadraticModel
