"""Polynomial-model simulated annealing, quadratic case, on the GPU engine.

Mirrors ``base_sample_hubo`` (openjij/sampler/base_sa_sample_hubo.py:30-140)
and the C++ ``sampler::SASampler`` it drives
(include/openjij/sampler/sa_sampler.hpp:27-293): the model is built as the
reference's ``IsingPolynomialModel`` / ``BinaryPolynomialModel`` constructors
build it (include/openjij/graph/ising_polynomial_model.hpp:40-146,
include/openjij/graph/binary_polynomial_model.hpp:45-146), the inverse
temperatures come from ``GenerateBetaList``
(include/openjij/utility/schedule_list.hpp:275-336) with the automatic bounds
of ``SetBetaMinAuto`` / ``SetBetaMaxAuto`` (sa_sampler.hpp:86-93), and every
read makes ``num_sweeps`` sequential single-spin-flip sweeps with the
METROPOLIS or HEAT_BATH rule (include/openjij/updater/single_spin_flip.hpp:340-376).

After reduction (s_i^2 = 1, x_i^2 = x_i) every term must have degree <= 2:
such a model is an Ising problem, and all reads anneal in one launch of the
dense or CSR kernels.  The random stream is the engine's counter-based one,
so ``seed`` reproduces GPU results, not the reference's Xorshift / MT stream.
"""
from __future__ import annotations

import math
import time
from typing import Optional

import numpy as np

from openjij_amd.sampler.response import Response
from openjij_amd.sampler.sampler import initial_states, resolve_seed
from openjij_amd.variable_type import BINARY, SPIN

UPDATE_METHODS = {"METROPOLIS": 0, "HEAT_BATH": 1}
_OTHER_METHODS = ("SUWA_TODO", "OPT_METROPOLIS")  # valid names, not single-flip rules
RANDOM_NUMBER_ENGINES = ("XORSHIFT", "MT", "MT_64")
TEMPERATURE_SCHEDULES = ("GEOMETRIC", "LINEAR")
_EPS = np.finfo(np.float64).eps
_RATIO = 1e-8  # min_max_energy_difference_ratio_ of the polynomial models


def _vartype(v):
    if v in ("SPIN", SPIN):
        return SPIN
    if v in ("BINARY", BINARY):
        return BINARY
    raise ValueError("vartype must `BINARY` or `SPIN`")


def _sort_key(x):
    return (isinstance(x, str), x)


class QuadraticPolynomialModel:
    """Terms of a polynomial model of degree <= 2 over sorted variables.

    ``terms`` maps sorted integer-position tuples ((), (i,), (i, j)) to their
    merged coefficients, as the reference's key_value_list."""

    def __init__(self, J: dict, vartype):
        self.vartype = _vartype(vartype)
        kept = [(tuple(k), float(v)) for k, v in J.items() if abs(float(v)) > _EPS]
        self.abs_max = max((abs(v) for _, v in kept), default=-1.0)
        reduced = []
        for key, v in kept:
            if self.vartype == SPIN:  # s^2 = 1: indices of even multiplicity drop out
                cnt = {}
                for x in key:
                    cnt[x] = cnt.get(x, 0) + 1
                red = [x for x, c in cnt.items() if c % 2 == 1]
            else:  # x^2 = x
                red = list(dict.fromkeys(key))
            reduced.append((red, v))
        self.variables = sorted({x for red, _ in reduced for x in red}, key=_sort_key)
        pos = {x: i for i, x in enumerate(self.variables)}
        terms = {}
        for red, v in reduced:
            ik = tuple(sorted(pos[x] for x in red))
            if len(ik) > 2:
                raise NotImplementedError(
                    "terms of degree > 2 after reduction need the k-local updater, "
                    "which this engine does not provide")
            terms[ik] = terms.get(ik, 0.0) + v
        self.terms = terms
        self.n = len(self.variables)
        self.degree = max((len(k) for k in terms), default=0)

    def energy_difference_bounds(self):
        """(estimated min, estimated max) energy difference of one flip
        (ising_polynomial_model.hpp:119-145; the binary model has no factor 2)."""
        f = 2.0 if self.vartype == SPIN else 1.0
        rel = self.abs_max * _RATIO
        row = np.zeros(self.n)
        dmin = np.finfo(np.float64).max
        for key, v in self.terms.items():
            if abs(v) >= rel:
                for i in key:
                    row[i] += abs(v)
                if key:
                    dmin = min(dmin, f * abs(v))
        dmax = float(f * row.max()) if self.n else -1.0
        if self.degree == 0:
            return 0.0, 0.0
        return float(dmin), dmax

    def ising_ext(self):
        """(n+1)^2 engine layout (J_ij upper triangle, h_i in column n, corner 1)
        of the same energy differences, and the energy offset it drops."""
        n = self.n
        ext = np.zeros((n + 1, n + 1))
        const = 0.0
        for key, v in self.terms.items():
            if self.vartype == SPIN:
                if len(key) == 0:
                    const += v
                elif len(key) == 1:
                    ext[key[0], n] += v
                else:
                    ext[key[0], key[1]] += v
            else:  # x = (1 + s) / 2
                if len(key) == 0:
                    const += v
                elif len(key) == 1:
                    ext[key[0], n] += v / 2
                    const += v / 2
                else:
                    i, j = key
                    ext[i, j] += v / 4
                    ext[i, n] += v / 4
                    ext[j, n] += v / 4
                    const += v / 4
        ext[n, n] = 1.0
        return ext, const

    def energies(self, states: np.ndarray) -> np.ndarray:
        """sum_k value_k * prod of the key's variables (CalculateEnergy,
        ising_polynomial_model.hpp:198-211), per row of ``states``."""
        x = np.asarray(states, dtype=np.float64)
        e = np.zeros(x.shape[0])
        for key, v in self.terms.items():
            if len(key) == 0:
                e += v
            elif len(key) == 1:
                e += v * x[:, key[0]]
            else:
                e += v * x[:, key[0]] * x[:, key[1]]
        return e


def generate_beta_list(schedule: str, beta_min: float, beta_max: float, num_sweeps: int):
    """GenerateBetaList (include/openjij/utility/schedule_list.hpp:275-336):
    LINEAR beta_min + i (beta_max - beta_min) / (L - 1); GEOMETRIC by repeated
    multiplication with alpha = (beta_max / beta_min)^(1 / (L - 1))."""
    if num_sweeps == 1:
        return np.array([beta_min], dtype=np.float64)
    out = np.empty(num_sweeps)
    if schedule == "LINEAR":
        for i in range(num_sweeps):
            out[i] = beta_min + i * (beta_max - beta_min) / (num_sweeps - 1)
    else:
        alpha = math.pow(beta_max / beta_min, 1.0 / float(num_sweeps - 1))
        b = beta_min
        for i in range(num_sweeps):
            out[i] = b
            b = b * alpha
    return out


def _engine_context(ext, R, precision):
    from openjij_amd import capi

    n = ext.shape[0] - 1
    sym = np.triu(ext) + np.triu(ext, 1).T
    offdiag = np.count_nonzero(np.triu(ext[:n, :n], 1))
    if n > 4096 and 2 * offdiag <= 16 * n:
        # sparse: (n+1)-row symmetric CSR with h in column n
        rows, cols = np.nonzero(sym)
        rp = np.zeros(n + 2, np.int64)
        np.add.at(rp, rows + 1, 1)
        return capi.Context.csr(n, np.cumsum(rp), cols, sym[rows, cols], R, 1, 0.0, precision)
    return capi.Context.dense(sym, R, 1, 0.0, precision)


def sample_hubo_quadratic(J: dict, vartype=None, num_sweeps: int = 1000, num_reads: int = 1,
                          num_threads: int = 1, beta_min: Optional[float] = None,
                          beta_max: Optional[float] = None, update_method: str = "METROPOLIS",
                          random_number_engine: str = "XORSHIFT", seed: Optional[int] = None,
                          temperature_schedule: str = "GEOMETRIC", precision=None) -> Response:
    from openjij_amd import capi
    from openjij_amd.sampler.sa_sampler import _precision_code

    start = time.time()
    if update_method in _OTHER_METHODS:
        raise RuntimeError("Unknown UpdateMethod")  # single_spin_flip.hpp:377-379
    if update_method not in UPDATE_METHODS:
        raise RuntimeError(f"Invalid update_method={update_method}")
    if random_number_engine not in RANDOM_NUMBER_ENGINES:
        raise RuntimeError(f"Invalid random_number_engine={random_number_engine}")
    if temperature_schedule not in TEMPERATURE_SCHEDULES:
        raise RuntimeError(f"Invalid temperature_schedule={temperature_schedule}")
    if num_sweeps <= 0:
        raise RuntimeError("num_sweeps must be larger than zero.")
    if num_reads <= 0:
        raise RuntimeError("num_reads must be larger than zero.")
    if num_threads <= 0:
        raise RuntimeError("num_threads must be non-negative integer.")
    if beta_min is not None and beta_min < 0:
        raise RuntimeError("beta_min must be positive number")
    if beta_max is not None and beta_max < 0:
        raise RuntimeError("beta_max must be positive number")

    t_def = time.time()
    model = QuadraticPolynomialModel(J, vartype)
    dmin, dmax = model.energy_difference_bounds()
    with np.errstate(divide="ignore"):
        bmin = float(beta_min) if beta_min is not None else float(np.log(2.0) / np.float64(dmax))
        bmax = float(beta_max) if beta_max is not None else float(np.log(100.0) / np.float64(dmin))
    betas = generate_beta_list(temperature_schedule, bmin, bmax, int(num_sweeps))
    define_time = time.time() - t_def

    R, n = int(num_reads), model.n
    seed_v = resolve_seed(seed)
    t_s = time.time()
    exec_ms = 0.0
    if n == 0:
        spins = np.zeros((R, 0), np.int8)
    else:
        ext, _ = model.ising_ext()
        scale = 1.0
        if not np.all(ext == np.round(ext)) and np.all(4 * ext == np.round(4 * ext)):
            scale = 4.0  # quarter-integer Ising from an integer binary model: exact path
        ext = ext * scale
        ext[n, n] = 1.0
        ctx = _engine_context(ext, R, _precision_code(precision))
        try:
            ctx.set_option("update_method", UPDATE_METHODS[update_method])
            ctx.set_spins(initial_states(R, n, seed_v, 0x5D))
            ctx.run(betas / scale, np.ones(betas.size, np.int64), seed_v, 0)
            spins = ctx.solution()
            exec_ms = ctx.info()["last_kernel_ms"]
        finally:
            ctx.close()
    sample_time = time.time() - t_s

    t_r = time.time()
    states = spins.astype(int) if model.vartype == SPIN else ((spins.astype(int) + 1) // 2)
    energies = model.energies(states)
    resp = Response.from_samples(states, model.vartype, energies, variables=model.variables)
    resp_time = time.time() - t_r
    resp.info["schedule"] = {
        "num_sweeps": num_sweeps, "num_reads": num_reads, "num_threads": num_threads,
        "beta_min": bmin, "beta_max": bmax, "update_method": update_method,
        "random_number_engine": random_number_engine,
        "temperature_schedule": temperature_schedule, "seed": seed_v,
    }
    resp.info["sampling_time"] = (sample_time + define_time) * 1e6
    resp.info["execution_time"] = exec_ms * 1e3 / R
    resp.info["time"] = {"define_cxx_sampler": define_time, "sample": sample_time,
                         "make_oj_response": resp_time, "total": time.time() - start}
    return resp
