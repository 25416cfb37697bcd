"""SPIN / BINARY variable types (openjij/variable_type.py)."""
from __future__ import annotations

import enum


class Vartype(enum.Enum):
    SPIN = frozenset({-1, 1})
    BINARY = frozenset({0, 1})


SPIN = Vartype.SPIN
BINARY = Vartype.BINARY


def cast_vartype(vartype) -> Vartype:
    if isinstance(vartype, Vartype):
        return vartype
    if isinstance(vartype, str):
        v = vartype.upper()
        if v == "SPIN":
            return SPIN
        if v == "BINARY":
            return BINARY
    if isinstance(vartype, (set, frozenset)):
        return Vartype(frozenset(vartype))
    raise TypeError("expected vartype to be one of: 'SPIN', 'BINARY', SPIN, BINARY")
