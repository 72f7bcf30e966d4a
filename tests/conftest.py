import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "bsam1.2_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libbsam_gpu.so)")


@pytest.fixture(scope="session")
def gpu_solver_cls():
    """The GPU Solver class; fails loudly when the library or GPU is missing."""
    from bsam_amd import gpu
    gpu.load()
    return gpu.Solver
