"""Shared fixtures. Tests marked `gpu` need a HIP device and the built libfhp.so; the rest run on CPU."""
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X) and the built libfhp.so")


def _has_gpu() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.build_oracle()
    return oracle


@pytest.fixture(scope="session")
def gpu():
    """Skips GPU tests on machines without a HIP device; fails loudly if the library is missing."""
    if not _has_gpu():
        pytest.skip("no HIP device")
    from falcor_amd import _lib
    _lib.load()  # raises if libfhp.so is missing: GPU tests never run without the native library
    import torch
    return torch.device("cuda", 0)
