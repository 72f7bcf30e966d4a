"""MI355X-native FAS multigrid V-cycle for BSAM 1.2 block-structured AMR.

The compute path is libbsam_gpu.so (HIP, gfx950), reached through the C ABI
of include/bsam_gpu.h; this package is its host-side mirror (hierarchy
construction, workload definitions, ctypes binding and build recipe).
"""
from . import abi, configs, hierarchy  # noqa: F401
