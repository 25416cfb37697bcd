"""openjij_amd: OpenJij's single-spin-flip annealing hot path on AMD Instinct
MI355X (gfx950).

User-facing names follow OpenJij: ``SASampler``, ``SQASampler``,
``BinaryQuadraticModel``, ``Response`` and the ``cxxjij`` binding module.
"""
from openjij_amd._native import cxxjij
from openjij_amd.variable_type import BINARY, SPIN, Vartype, cast_vartype
from openjij_amd.model.model import BinaryQuadraticModel, BQM
from openjij_amd.sampler.response import Response
from openjij_amd.sampler.sa_sampler import SASampler
from openjij_amd.sampler.sqa_sampler import SQASampler
from openjij_amd.utils.benchmark import (
    residual_energy,
    se_lower_tts,
    se_residual_energy,
    se_success_probability,
    se_upper_tts,
    solver_benchmark,
    success_probability,
    time_to_solution,
)

__all__ = [
    "cxxjij",
    "SPIN",
    "BINARY",
    "Vartype",
    "cast_vartype",
    "BinaryQuadraticModel",
    "BQM",
    "Response",
    "SASampler",
    "SQASampler",
    "solver_benchmark",
    "residual_energy",
    "time_to_solution",
    "success_probability",
    "se_residual_energy",
    "se_success_probability",
    "se_upper_tts",
    "se_lower_tts",
]
