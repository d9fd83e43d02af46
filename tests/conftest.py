"""Test configuration: import paths and the `gpu` marker.

Tests marked ``gpu`` run the HIP kernels and need an MI355X; everything else
runs on the CPU (the oracle, host logic, the C ABI symbol table).
"""

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / "taichi-mpm-lbm_amd", ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: longer CPU-side oracle runs")


@pytest.fixture(scope="session")
def gpu_available():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test selected but torch.cuda.is_available() is False")
    from fsi import _native

    _native.load_library()
    return True
