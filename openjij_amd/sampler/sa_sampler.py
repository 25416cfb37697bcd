"""Simulated annealing sampler on the GPU engine.

Same call surface and defaults as the reference SASampler.sample
(openjij/sampler/sa_sampler.py:137-329): num_sweeps 1000, num_reads 1,
geometric beta schedule with automatic bounds, ``sparse=True`` by default.
The per-read loop of the reference (openjij/sampler/sampler.py:115-140) is
replaced by one batched launch over all ``num_reads`` replicas.
"""
from __future__ import annotations

import time
from typing import Optional

import numpy as np

from openjij_amd._native import cxxjij
from openjij_amd.model.model import BinaryQuadraticModel
from openjij_amd.sampler.response import Response
from openjij_amd.sampler.sampler import (
    BaseSampler,
    initial_states,
    parse_initial_state,
    resolve_seed,
    to_vartype_states,
)

FP32_REFRESH_EVERY = 16  # sweeps between matrix-core field refreshes on the FP32 path

_PRECISIONS = {
    None: 0, "auto": 0, "fp64": 1, "float64": 1, "fp32": 2, "mixed": 2, "float32": 2,
    "int32": 3, "exact": 3,
}


def _precision_code(p) -> int:
    if isinstance(p, int):
        return p
    key = p.lower() if isinstance(p, str) else p
    if key not in _PRECISIONS:
        raise ValueError(f"unknown precision {p!r}")
    return _PRECISIONS[key]


_TINY = 1e-8


def beta_stats_host(mat, spins):
    """dE statistics of the automatic bounds (openjij/sampler/sa_sampler.py:882-911)
    for the (n+1)^2 interaction matrix `mat` (dense or scipy) and a random spin
    matrix `spins` of shape (n+1, 2) whose last row is 1.  Returns
    (dE_max, dE_min, rate_dE)."""
    dense = isinstance(mat, np.ndarray)
    dE_max, dE_min, rate = 1.0, 1.0, 0.5
    quad = mat[:-1, :-1] @ spins[:-1, :] * (-2 * spins[:-1, :])
    lin_col = mat[:-1, -1][:, np.newaxis] if dense else mat[:-1, -1].toarray()
    lin = lin_col * spins[:-1]
    quad_abs = np.abs(quad)
    big = quad_abs > _TINY
    if np.any(big):
        rate = np.max(np.abs(lin[big]) / (quad_abs[big].mean() + _TINY))
    pos = quad[quad > _TINY]
    if len(pos) == 0:
        dE_min, dE_max = 1, 1.1
    else:
        dE_max = np.median(pos, axis=0).mean()
        dE_min = np.min(pos, axis=0).mean()
    return dE_max, dE_min, rate


def int_min_host(mat):
    """Smallest |entry| above the threshold of the (n+1)^2 layout (sa_sampler.py:937-938)."""
    absm = np.abs(mat)
    if not isinstance(mat, np.ndarray):
        absm = absm.toarray()
    return np.min(absm[absm > _TINY])


def beta_bounds(dE_max, dE_min, rate, int_min, n1, num_sweeps, beta_min=None, beta_max=None):
    """The scalar formulas of the automatic bounds (sa_sampler.py:914-968):
    beta_min by a five-step bisection of 2 beta p_init = exp(-beta dE_min) -
    exp(-beta dE_max), beta_max from int_min / dE_min; `int_min` is a value or a
    callable (evaluated only when beta_max is None).  Returns (beta_min, beta_max)."""
    n = n1
    p_init = 9 / 10 - 8 / 10 * np.tanh(rate / 50) + _TINY
    p_final = 1 / 1000
    if beta_min is None:
        lo, hi = -np.log(p_init) / dE_max, -np.log(p_init) / dE_min
        for _ in range(5):
            mid = (lo + hi) / 2
            f = 2 * mid * p_init - np.exp(-mid * dE_min) + np.exp(-mid * dE_max)
            if abs(f) < 1e-8:
                break
            if f > 0:
                hi = mid
            else:
                lo = mid
        beta_min = lo
        bmm = beta_min * 0.8
        beta_min = bmm + (beta_min * 1.5 - bmm) * np.tanh(num_sweeps / n)
    if beta_max is None:
        im = int_min() if callable(int_min) else int_min
        im = max(im, dE_min / 10.0)
        b_int = -np.log(p_final) / im
        b_dE = -np.log(p_final) / dE_min
        if rate < 2:
            lo_, hi_ = min(b_dE, b_int), max(b_dE, b_int)
            beta_max = lo_ + (hi_ - lo_) * np.tanh(num_sweeps / (2 * n))
        else:
            last = min(100, num_sweeps // 10 + 1)
            beta_max = -np.log(p_final) / max(dE_min, im) / last
        if beta_max < beta_min:
            beta_max = beta_min * 10.0
    if rate > 1:
        beta_min *= rate
        beta_max *= rate
    return beta_min, beta_max


def device_beta_stats(row, scale=1.0):
    """(dE_max, dE_min, rate_dE, int_min) of the reference's analysis from one
    row of ojmi_batch_beta_stats computed on couplings scaled by `scale`
    (1 or 4, so the divisions are exact)."""
    dE_max, dE_min, lin_max, mean_abs, im = (float(x) for x in row)
    dE_max, dE_min = dE_max / scale, dE_min / scale
    rate = 0.5 if mean_abs < 0 else (lin_max / scale) / (mean_abs / scale + _TINY)
    return dE_max, dE_min, rate, min(im / scale, 1.0)


def random_spin_pair(n1, seed=None):
    """The reference's random spin matrix: np.random.choice([-1, 1], size=(n1, 2))
    after np.random.seed(seed), last row set to 1 (sa_sampler.py:876-890)."""
    if seed is not None:
        np.random.seed(seed)
    spins = np.random.choice([-1, 1], size=(n1, 2))
    spins[-1, :] = 1
    return spins


def geometric_schedule(beta_min, beta_max, num_sweeps):
    per_beta = max(1, num_sweeps // 1000)
    return cxxjij.utility.make_classical_schedule_list(
        beta_min=beta_min, beta_max=beta_max, one_mc_step=per_beta,
        num_call_updater=num_sweeps // per_beta)


def geometric_ising_beta_schedule(cxxgraph, beta_max=None, beta_min=None, num_sweeps=1000,
                                  seed=None):
    """Geometric beta schedule with automatic bounds, as the reference builds it
    (openjij/sampler/sa_sampler.py:853-976): the bounds come from local-field
    statistics of two random spin vectors drawn with NumPy's legacy global
    stream after ``np.random.seed(seed)``; one_mc_step = max(1, num_sweeps // 1000).
    Returns (schedule list, [beta_max, beta_min])."""
    if seed is not None:
        np.random.seed(seed)
    mat = cxxgraph.get_interactions()
    n1 = mat.shape[0]
    dE_max, dE_min, rate = 1.0, 1.0, 0.5
    if beta_min is None or beta_max is None:
        if n1 <= 1:
            dE_max = dE_min = 1
        else:
            dE_max, dE_min, rate = beta_stats_host(mat, random_spin_pair(n1))
    beta_min, beta_max = beta_bounds(dE_max, dE_min, rate, lambda: int_min_host(mat), n1,
                                     num_sweeps, beta_min, beta_max)
    return geometric_schedule(beta_min, beta_max, num_sweeps), [beta_max, beta_min]


def schedule_arrays(schedule):
    beta = np.array([e.updater_parameter.beta for e in schedule], dtype=float)
    mcs = np.array([e.one_mc_step for e in schedule], dtype=np.int64)
    return beta, mcs


class SASampler(BaseSampler):
    """Simulated annealing over classical Ising / QUBO models, batched on the
    GPU: all ``num_reads`` replicas anneal in one kernel launch."""

    def __init__(self):
        self._default_params = {
            "beta_min": None, "beta_max": None, "num_sweeps": 1000, "schedule": None,
            "num_reads": 1,
        }
        self._params = dict(self._default_params)
        self.schedule_info = {}

    def _convert_validation_schedule(self, schedule):
        """openjij/sampler/sa_sampler.py:98-135"""
        if not isinstance(schedule, (list, np.ndarray)):
            raise ValueError("schedule should be list or numpy.array")
        if isinstance(schedule[0], cxxjij.utility.ClassicalSchedule):
            return list(schedule)
        if len(schedule[0]) != 2:
            raise ValueError(
                "schedule is list of tuple or list (beta : float, step_length : int)")
        beta = np.array(schedule, dtype=float).T[0]
        if not np.all(0 <= beta):
            raise ValueError("schedule beta range is '0 <= beta'.")
        out = []
        for b, steps in schedule:
            e = cxxjij.utility.ClassicalSchedule()
            e.one_mc_step = int(steps)
            e.updater_parameter.beta = float(b)
            out.append(e)
        return out

    def sample(self, bqm, beta_min: Optional[float] = None, beta_max: Optional[float] = None,
               num_sweeps: Optional[int] = None, num_reads: Optional[int] = None,
               schedule: Optional[list] = None, initial_state=None,
               updater: Optional[str] = None, sparse: Optional[bool] = None,
               reinitialize_state: Optional[bool] = None, seed: Optional[int] = None,
               precision=None) -> Response:
        if updater is None:
            updater = "single spin flip"
        if sparse is None:
            sparse = True
        if reinitialize_state is None:
            reinitialize_state = True
        name = updater.lower().replace("_", "").replace(" ", "")
        if name != "singlespinflip":
            raise ValueError('updater is one of "single spin flip"')
        if bqm.sparse != bool(sparse):
            bqm = BinaryQuadraticModel(bqm.linear, bqm.quadratic, bqm.offset, bqm.vartype,
                                       sparse=bool(sparse))
        model = bqm
        graph, offset = model.get_cxxjij_ising_graph()
        self._set_params(beta_min=beta_min, beta_max=beta_max, num_sweeps=num_sweeps,
                         num_reads=num_reads, schedule=schedule)
        if self._params["schedule"] is None:
            sched, rng = geometric_ising_beta_schedule(
                graph, beta_max=self._params["beta_max"], beta_min=self._params["beta_min"],
                num_sweeps=self._params["num_sweeps"], seed=seed)
            self.schedule_info = {"beta_max": rng[0], "beta_min": rng[1],
                                  "num_sweeps": self._params["num_sweeps"]}
        else:
            sched = self._convert_validation_schedule(self._params["schedule"])
            self.schedule_info = {"schedule": "custom schedule"}
        beta, mcs = schedule_arrays(sched)
        R = int(self._params["num_reads"])
        n = graph.size()
        seed_v = resolve_seed(seed)
        if initial_state is None:
            init = initial_states(R, n, seed_v, 0x5A)
        else:
            init = np.tile(parse_initial_state(initial_state, model, n), (R, 1))

        t0 = time.perf_counter()
        info = {}
        if n == 0:
            spins = np.zeros((R, 0), dtype=np.int8)
            exec_us = np.zeros(R)
        elif reinitialize_state:
            eng = cxxjij.engine.Engine(graph, R, 1, 0.0, _precision_code(precision))
            if _precision_code(precision) == 2:
                # FP32 fields refreshed on the matrix cores every 16 sweeps
                # (dense systems; sparse ones keep their carried fields)
                try:
                    eng.set_option("refresh_every", FP32_REFRESH_EVERY)
                except RuntimeError:
                    pass
            eng.set_spins(init)
            eng.run(beta, None, mcs, seed_v, 0)
            spins = eng.solution()
            inf = eng.info()
            exec_us = np.full(R, inf["last_kernel_ms"] * 1e3 / R)
            info["engine"] = inf
        else:
            # reads continue from the previous read's final state
            eng = cxxjij.engine.Engine(graph, 1, 1, 0.0, _precision_code(precision))
            cur = init[:1]
            spins = np.empty((R, n), dtype=np.int8)
            exec_us = np.empty(R)
            sweeps = int(mcs.sum())
            for r in range(R):
                eng.set_spins(cur)
                eng.run(beta, None, mcs, seed_v, r * sweeps)
                cur = eng.spins()
                spins[r] = cur[0]
                exec_us[r] = eng.info()["last_kernel_ms"] * 1e3
        states = to_vartype_states(spins, model.vartype)
        energies = model.energies(states) if n else np.full(R, model.offset)
        sampling_us = (time.perf_counter() - t0) * 1e6
        resp = Response(states, energies, model.variables, model.vartype, info)
        resp.info["sampling_time"] = sampling_us
        resp.info["execution_time"] = float(np.mean(exec_us))
        resp.info["list_exec_times"] = exec_us
        resp.info["schedule"] = self.schedule_info
        return resp

    def sample_hubo(self, J, vartype: Optional[str] = None, num_sweeps: int = 1000,
                    num_reads: int = 1, num_threads: int = 1,
                    beta_min: Optional[float] = None, beta_max: Optional[float] = None,
                    updater: str = "METROPOLIS", random_number_engine: str = "XORSHIFT",
                    seed: Optional[int] = None, temperature_schedule: str = "GEOMETRIC",
                    precision=None) -> Response:
        """Polynomial (dict) models whose terms reduce to degree <= 2, with the
        METROPOLIS or HEAT_BATH rule and a GEOMETRIC or LINEAR schedule
        (openjij/sampler/sa_sampler.py:487-567).  The reference's "k-local"
        updater and non-dict inputs take its older path, which is not part of
        this engine."""
        from openjij_amd.sampler.hubo import sample_hubo_quadratic

        if updater == "k-local" or not isinstance(J, dict):
            raise NotImplementedError(
                "the k-local updater (higher-order models) is not part of this engine")
        if updater == "single spin flip":
            updater = "METROPOLIS"
        return sample_hubo_quadratic(
            J, vartype=vartype, num_sweeps=num_sweeps, num_reads=num_reads,
            num_threads=num_threads, beta_min=beta_min, beta_max=beta_max,
            update_method=updater, random_number_engine=random_number_engine, seed=seed,
            temperature_schedule=temperature_schedule, precision=precision)

    def sample_batch(self, models, beta_min: Optional[float] = None,
                     beta_max: Optional[float] = None, num_sweeps: Optional[int] = None,
                     num_reads: Optional[int] = None, seed: Optional[int] = None,
                     precision=None) -> list:
        """Anneal independent models of different sizes in one launch.

        Replaces a loop of ``sample`` calls, one per model (each a loop over
        reads, openjij/sampler/sampler.py:115-140), by one batch context: every
        model gets its own automatic geometric schedule (or the given bounds)
        over the same number of sweeps, and ``num_reads`` replicas.  A BINARY
        model with integer coefficients has quarter-integer Ising couplings;
        they are scaled by 4 (and beta by 1/4, so beta * E is unchanged) to run
        on the exact integer path.  Each model's random stream carries its
        position in ``models``, so results differ from separate ``sample``
        calls with the same seed.  Returns one Response per model."""
        from openjij_amd import capi

        self._set_params(beta_min=beta_min, beta_max=beta_max, num_sweeps=num_sweeps,
                         num_reads=num_reads)
        R = int(self._params["num_reads"])
        sweeps = int(self._params["num_sweeps"])
        seed_v = resolve_seed(seed)
        models = list(models)
        idx, J_exts, scales = [], [], []
        for k, bqm in enumerate(models):
            if bqm.num_variables == 0:
                continue
            dense = BinaryQuadraticModel(bqm.linear, bqm.quadratic, bqm.offset, bqm.vartype,
                                         sparse=False)
            h, J, _ = dense.ising_arrays()
            n = dense.num_variables
            ext = np.zeros((n + 1, n + 1))
            ext[:n, :n] = J
            ext[:n, n] = h
            scale = 1.0
            if not (np.all(ext == np.round(ext))) and np.all(4 * ext == np.round(4 * ext)):
                scale = 4.0
            ext *= scale
            ext[n, n] = 1.0
            idx.append(k)
            J_exts.append(ext)
            scales.append(scale)
        t0 = time.perf_counter()
        spins = {}
        exec_us = 0.0
        self.schedule_info = {}
        if idx:
            ctx = capi.BatchContext.create(J_exts, R, _precision_code(precision))
            try:
                bmin, bmax = self._params["beta_min"], self._params["beta_max"]
                stats = None
                if bmin is None or bmax is None:
                    # the reference's random spin pair of each model (the same
                    # draws as one sample() call per model with this seed), and
                    # the dE statistics of all models at once on the GPU
                    pairs = [random_spin_pair(E.shape[0], seed)[:-1].T for E in J_exts]
                    stats = ctx.beta_stats(pairs)
                betas, bounds, mcs = [], [], None
                for i, E in enumerate(J_exts):
                    if stats is None:
                        lo, hi = bmin, bmax
                    else:
                        lo, hi = beta_bounds(*device_beta_stats(stats[i], scales[i]),
                                             E.shape[0], sweeps, bmin, bmax)
                    beta, m = schedule_arrays(geometric_schedule(lo, hi, sweeps))
                    betas.append(beta / scales[i])
                    bounds.append((float(lo), float(hi)))
                    mcs = m
                self.schedule_info = {"beta_min": [b[0] for b in bounds],
                                      "beta_max": [b[1] for b in bounds],
                                      "num_sweeps": sweeps, "auto_bounds_on_gpu": stats is not None}
                init = [initial_states(R, E.shape[0] - 1, seed_v, 0x5C + i)
                        for i, E in enumerate(J_exts)]
                ctx.set_spins(init)
                ctx.run_batch(np.stack(betas), mcs, seed_v, 0)
                out = ctx.solution()
                exec_us = ctx.info()["last_kernel_ms"] * 1e3 / (R * len(idx))
            finally:
                ctx.close()
            spins = dict(zip(idx, out))
        sampling_us = (time.perf_counter() - t0) * 1e6
        responses = []
        for k, bqm in enumerate(models):
            n = bqm.num_variables
            sp = spins.get(k, np.zeros((R, 0), np.int8))
            states = to_vartype_states(sp, bqm.vartype)
            energies = bqm.energies(states) if n else np.full(R, bqm.offset)
            resp = Response(states, energies, bqm.variables, bqm.vartype, {})
            resp.info["sampling_time"] = sampling_us
            resp.info["execution_time"] = exec_us
            resp.info["list_exec_times"] = np.full(R, exec_us)
            responses.append(resp)
        return responses
