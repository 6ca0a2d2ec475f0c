"""Shared fixtures.  Tests marked `gpu` need an MI355X and run the HIP
library through its C ABI; everything else runs on the CPU."""
import importlib.util
import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "cross-dpc-episdet_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X; runs the HIP library through its C ABI")


def load_episdet():
    spec = importlib.util.spec_from_file_location("episdet", os.path.join(PKG, "episdet.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["episdet"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def epi():
    return load_episdet()


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle


@pytest.fixture(scope="session")
def golden():
    def read(name):
        with open(os.path.join(GOLDEN, name)) as f:
            return json.load(f)
    return read


@pytest.fixture(scope="session")
def ctx(epi):
    """One device context for all GPU tests (fails loudly without a GPU)."""
    c = epi.Context(0)
    yield c
    c.close()
