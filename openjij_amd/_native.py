"""Loads the compiled extension.  There is no Python or CPU fallback: a missing
or unloadable extension is an ImportError naming the build command."""
from __future__ import annotations

import os

_HERE = os.path.dirname(os.path.abspath(__file__))

try:
    from openjij_amd import cxxjij  # noqa: F401  (pybind11 module on libojmi.so)
except ImportError as exc:  # pragma: no cover - exercised only on a broken install
    raise ImportError(
        "openjij_amd: the native extension is missing or failed to load "
        f"({exc}); build it with `python -c 'import __graft_entry__ as g; g.build()'` "
        "or `python openjij_amd/_build.py`"
    ) from exc

LIBOJMI = os.path.join(_HERE, "libojmi.so")
