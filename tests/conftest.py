"""Shared fixtures.  GPU tests are marked ``gpu``; everything else runs on CPU.

``oracle`` is the numpy restatement in ``oracle/sv_oracle.py``; ``ref`` is the
reference C++ core compiled from its sources into ``oracle/_ref`` (skipped when
that build is absent).  Both are checkers only.
"""

import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


def _load_oracle():
    path = os.path.join(ROOT, "oracle", "sv_oracle.py")
    spec = importlib.util.spec_from_file_location("sv_oracle", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def load_reference():
    """The compiled reference module, or None if oracle/_ref was not built."""
    d = os.path.join(ROOT, "oracle", "_ref")
    if not os.path.isdir(d):
        return None
    for f in os.listdir(d):
        if f.startswith("qforte.") and f.endswith(".so"):
            spec = importlib.util.spec_from_file_location("qforte", os.path.join(d, f))
            mod = importlib.util.module_from_spec(spec)
            spec.loader.exec_module(mod)
            return mod
    return None


@pytest.fixture(scope="session")
def oracle():
    return _load_oracle()


@pytest.fixture(scope="session")
def ref():
    mod = load_reference()
    if mod is None:
        pytest.skip("compiled reference (oracle/_ref) not built")
    return mod


@pytest.fixture(scope="session")
def qf():
    import qforte_amd

    return qforte_amd


@pytest.fixture(scope="session")
def gpu(qf):
    if qf.device_count() < 1:
        pytest.fail("GPU test selected but no GPU is visible")
    return True
