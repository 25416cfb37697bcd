"""Simulated quantum annealing (discrete-time Suzuki-Trotter) on the GPU engine.

Same call surface and defaults as the reference SQASampler
(openjij/sampler/sqa_sampler.py:53-345): beta 5, gamma 1, trotter 4,
num_sweeps 1000 and the quartic s-schedule of Morita & Nishimori by default;
custom schedules as (s, one_mc_step) or (s, beta, one_mc_step).  All reads
anneal in one launch (one workgroup per read, one wave per Trotter slice).
"""
from __future__ import annotations

import time
from typing import Optional

import numpy as np

from openjij_amd._native import cxxjij
from openjij_amd.model.model import BinaryQuadraticModel
from openjij_amd.sampler.response import Response
from openjij_amd.sampler.sa_sampler import _precision_code
from openjij_amd.sampler.sampler import (
    BaseSampler,
    initial_states,
    parse_initial_state,
    resolve_seed,
    to_vartype_states,
)


def linear_ising_schedule(model, beta, gamma, num_sweeps):
    sched = cxxjij.utility.make_transverse_field_schedule_list(
        beta=beta, one_mc_step=1, num_call_updater=num_sweeps)
    return sched, [beta, gamma]


def quartic_ising_schedule(model, beta, gamma, num_sweeps):
    """s(t) = t^4 (35 - 84 t + 70 t^2 - 20 t^3) on a uniform grid
    (openjij/sampler/sqa_sampler.py:327-345)."""
    t = np.linspace(0, 1, num_sweeps)
    fs = t ** 4 * (35 - 84 * t + 70 * t ** 2 - 20 * t ** 3)
    return [((beta, float(v)), 1) for v in fs], [beta, gamma]


class SQASampler(BaseSampler):
    def __init__(self):
        self._default_params = {
            "beta": 5.0, "gamma": 1.0, "num_sweeps": 1000, "schedule": None, "trotter": 4,
            "num_reads": 1,
        }
        self._params = dict(self._default_params)
        self.schedule_info = {}

    def _convert_validation_schedule(self, schedule, beta):
        """openjij/sampler/sqa_sampler.py:86-126 -> (beta[], s[], mcs[])"""
        if not isinstance(schedule, (list, np.ndarray)):
            raise ValueError("schedule should be list or numpy.array")
        if isinstance(schedule[0], cxxjij.utility.TransverseFieldSchedule):
            return (np.array([e.updater_parameter.beta for e in schedule], float),
                    np.array([e.updater_parameter.s for e in schedule], float),
                    np.array([e.one_mc_step for e in schedule], np.int64))
        if isinstance(schedule[0], tuple) and len(schedule[0]) == 2 and isinstance(
                schedule[0][0], tuple):
            # ((beta, s), mcs) tuple list, the quartic default's form
            return (np.array([e[0][0] for e in schedule], float),
                    np.array([e[0][1] for e in schedule], float),
                    np.array([e[1] for e in schedule], np.int64))
        sch = np.array(schedule, dtype=float).T[0]
        if not np.all((0 <= sch) & (sch <= 1)):
            raise ValueError("schedule range is '0 <= s <= 1'.")
        if len(schedule[0]) == 2:
            return (np.full(len(schedule), float(beta)),
                    np.array([e[0] for e in schedule], float),
                    np.array([e[1] for e in schedule], np.int64))
        if len(schedule[0]) == 3:
            return (np.array([e[1] for e in schedule], float),
                    np.array([e[0] for e in schedule], float),
                    np.array([e[2] for e in schedule], np.int64))
        raise ValueError(
            "schedule is list of tuple or list (annealing parameter s : float, step_length : "
            "int) or (annealing parameter s : float, beta: float, step_length : int)")

    def sample(self, bqm, beta: Optional[float] = None, gamma: Optional[float] = None,
               num_sweeps: Optional[int] = None, schedule: Optional[list] = None,
               trotter: Optional[int] = None, num_reads: Optional[int] = None,
               initial_state=None, updater: Optional[str] = None,
               sparse: Optional[bool] = None, reinitialize_state: Optional[bool] = None,
               seed: Optional[int] = None, precision=None) -> Response:
        if sparse is None:
            sparse = True
        if reinitialize_state is None:
            reinitialize_state = True
        if updater is None:
            updater = "single spin flip"
        if updater.lower().replace("_", "").replace(" ", "") != "singlespinflip":
            raise ValueError('updater is one of "single spin flip"')
        if bqm.sparse != bool(sparse):
            bqm = BinaryQuadraticModel(bqm.linear, bqm.quadratic, bqm.offset, bqm.vartype,
                                       sparse=bool(sparse))
        graph, offset = bqm.get_cxxjij_ising_graph()
        self._set_params(beta=beta, gamma=gamma, num_sweeps=num_sweeps, num_reads=num_reads,
                         trotter=trotter, schedule=schedule)
        p = self._params
        if p["schedule"]:
            b_arr, s_arr, mcs = self._convert_validation_schedule(p["schedule"], p["beta"])
            self.schedule_info = {"schedule": "custom schedule"}
        else:
            sched, bg = quartic_ising_schedule(bqm, p["beta"], p["gamma"], p["num_sweeps"])
            b_arr, s_arr, mcs = self._convert_validation_schedule(sched, p["beta"])
            self.schedule_info = {"beta": bg[0], "gamma": bg[1], "num_sweeps": p["num_sweeps"]}
        M = int(p["trotter"])
        if M < 2:
            raise ValueError("trotter slices must be equal or larger than 2.")
        R = int(p["num_reads"])
        n = graph.size()
        seed_v = resolve_seed(seed)
        if initial_state is None:
            base = initial_states(R, n, seed_v, 0x5B)
        else:
            base = np.tile(parse_initial_state(initial_state, bqm, n), (R, 1))
        init = np.repeat(base[:, None, :], M, axis=1)  # every slice starts equal

        t0 = time.perf_counter()
        info = {"system": []}
        if n == 0:
            sol = np.zeros((R, 0), dtype=np.int8)
            tspins = np.zeros((R, M, 0), dtype=np.int8)
            exec_us = np.zeros(R)
        else:
            if reinitialize_state:
                eng = cxxjij.engine.Engine(graph, R, M, float(p["gamma"]),
                                           _precision_code(precision))
                eng.set_spins(init)
                eng.run(b_arr, s_arr, mcs, seed_v, 0)
                exec_us = np.full(R, eng.info()["last_kernel_ms"] * 1e3 / R)
                tspins = eng.spins().reshape(R, M, n)
                sol = eng.solution()
            else:
                eng1 = cxxjij.engine.Engine(graph, 1, M, float(p["gamma"]),
                                            _precision_code(precision))
                cur = init[:1]
                tspins = np.empty((R, M, n), dtype=np.int8)
                sol = np.empty((R, n), dtype=np.int8)
                exec_us = np.empty(R)
                sweeps = int(mcs.sum())
                for r in range(R):
                    eng1.set_spins(cur)
                    eng1.run(b_arr, s_arr, mcs, seed_v, r * sweeps)
                    cur = eng1.spins().reshape(1, M, n)
                    tspins[r] = cur[0]
                    sol[r] = eng1.solution()[0]
                    exec_us[r] = eng1.info()["last_kernel_ms"] * 1e3
        states = to_vartype_states(sol, bqm.vartype)
        energies = bqm.energies(states) if n else np.full(R, bqm.offset)
        for r in range(R):
            q = to_vartype_states(tspins[r], bqm.vartype)
            info["system"].append({
                "trotter_state": q,
                "trotter_energies": list(bqm.energies(q)) if n else [bqm.offset] * M,
            })
        resp = Response(states, energies, bqm.variables, bqm.vartype, info)
        resp.info["sampling_time"] = (time.perf_counter() - t0) * 1e6
        resp.info["execution_time"] = float(np.mean(exec_us))
        resp.info["list_exec_times"] = exec_us
        resp.info["schedule"] = self.schedule_info
        return resp
