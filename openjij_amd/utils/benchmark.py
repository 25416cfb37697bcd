"""Metric definitions of the reference benchmark helpers
(openjij/utils/benchmark.py:25-223): residual energy, success probability,
time to solution and their standard errors, plus ``solver_benchmark``.

TTS follows tau * log(1 - p_r) / log(1 - p_s) with the reference's edge cases
(p_s = 1 -> 0, p_s = 0 -> inf).
"""
from __future__ import annotations

import logging
import time

import numpy as np

logger = logging.getLogger("openjij_amd.utils.benchmark")


def residual_energy(response, ref_energy):
    return np.mean(response.energies) - ref_energy


def se_residual_energy(response, ref_energy):
    return np.std(response.energies, ddof=1)


def _hits(response, solutions):
    if isinstance(solutions[0], dict):
        return [1 if dict(s) in solutions else 0 for s in response.samples()]
    return [1 if list(s) in [list(x) for x in solutions] else 0 for s in response.states]


def success_probability(response, solutions, ref_energy=0, measure_with_energy=False):
    if measure_with_energy:
        e = np.asarray(response.energies)
        return np.count_nonzero(e <= ref_energy) / len(e)
    return float(np.mean(_hits(response, solutions)))


def se_success_probability(response, solutions, ref_energy=0, measure_with_energy=False):
    if measure_with_energy:
        e = np.asarray(response.energies)
        return np.sqrt(np.count_nonzero(e <= ref_energy) / (len(e) - 1))
    h = _hits(response, solutions)
    return float(np.std(h) / np.sqrt(len(h)))


def time_to_solution(success_prob, computation_time, p_r):
    if success_prob == 1.0:
        return 0.0
    if success_prob == 0.0:
        return np.inf
    return computation_time * np.log(1 - p_r) / np.log(1 - success_prob)


def se_lower_tts(tts, success_prob, computation_time, p_r, se_success_prob):
    if 1 - (success_prob + se_success_prob) <= 0.0 or success_prob == 0.0:
        low = 0.0
    else:
        low = computation_time * np.log(1 - p_r) / np.log(1 - (success_prob + se_success_prob))
    return abs(low - tts)


def se_upper_tts(tts, success_prob, computation_time, p_r, se_success_prob):
    if success_prob in (0.0, 1.0):
        up = 0.0
    else:
        up = computation_time * np.log(1 - p_r) / np.log(1 - (success_prob - se_success_prob))
    return abs(up - tts)


def solver_benchmark(solver, time_list, solutions=None, p_r=0.99, ref_energy=0,
                     measure_with_energy=False, time_name="execution_time"):
    """For each time parameter t, call ``solver(t)`` and collect TTS, success
    probability and residual energy with their standard errors
    (openjij/utils/benchmark.py:25-111)."""
    if not (solutions or measure_with_energy):
        raise ValueError("input solutions or ref_energy")
    out = {k: [] for k in ("time", "success_prob", "tts", "residual_energy", "se_success_prob",
                            "se_upper_tts", "se_lower_tts", "se_residual_energy", "info")}
    for t in time_list:
        start = time.perf_counter()
        resp = solver(t)
        wall = time.perf_counter() - start
        tau = resp.info.get(time_name, wall * 1e6) if hasattr(resp, "info") else wall * 1e6
        ps = success_probability(resp, solutions, ref_energy, measure_with_energy)
        se_ps = se_success_probability(resp, solutions, ref_energy, measure_with_energy)
        tts = time_to_solution(ps, tau, p_r)
        out["time"].append(tau)
        out["success_prob"].append(ps)
        out["se_success_prob"].append(se_ps)
        out["tts"].append(tts)
        out["se_upper_tts"].append(se_upper_tts(tts, ps, tau, p_r, se_ps))
        out["se_lower_tts"].append(se_lower_tts(tts, ps, tau, p_r, se_ps))
        out["residual_energy"].append(residual_energy(resp, ref_energy))
        out["se_residual_energy"].append(se_residual_energy(resp, ref_energy))
        out["info"].append(getattr(resp, "info", {}))
    return out
