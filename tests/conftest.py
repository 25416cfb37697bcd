"""Test configuration: the `gpu` marker and an in-tree build of the native
libraries when they are missing (the engine, the cxxjij module, the oracle)."""
import glob
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def _ensure_built():
    pkg = os.path.join(ROOT, "openjij_amd")
    need_engine = not os.path.exists(os.path.join(pkg, "libojmi.so"))
    need_mod = not glob.glob(os.path.join(pkg, "cxxjij*.so"))
    need_oracle = not os.path.exists(os.path.join(ROOT, "oracle", "liboj_oracle.so"))
    if need_engine or need_mod or need_oracle:
        from openjij_amd import _build

        if need_engine:
            _build.build_engine()
        if need_engine or need_mod:
            _build.build_cxxjij()
        if need_oracle:
            _build.build_oracle()


_ensure_built()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (runs on the GPU box)")


@pytest.fixture(scope="session")
def fixtures():
    import json

    with open(os.path.join(ROOT, "tests", "golden", "reference_fixtures.json")) as f:
        return json.load(f)
