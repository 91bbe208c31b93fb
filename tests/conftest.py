"""Shared test configuration: the ``gpu`` marker, the repository root on sys.path and the
default in-call binning threshold."""

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(autouse=True, scope="session")
def _default_binning_threshold():
    """Tests that choose a batch size to land on one side of the in-call binning threshold
    (2^18 markers) need that threshold: pin it, whatever tuned value the library loaded.
    The launch variants a tuned table sets change no result, so they stay."""
    try:
        from mpifield_amd import _lib
        from mpifield_amd import kernels as K

        lib = _lib.load()
    except Exception:  # no built library: the tests that need it say so themselves
        yield
        return
    old = lib.mpf_set_binned_min(K.BINNED_MIN_MARKERS)
    yield
    lib.mpf_set_binned_min(old)
