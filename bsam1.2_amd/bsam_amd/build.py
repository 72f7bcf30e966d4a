"""Build recipe of libbsam_gpu.so (hipcc, gfx950) and of the Fortran host
driver that calls it through the ISO_C_BINDING module (amdflang).

The HIP sources are compiled with -ffp-contract=off so that the FP64 path
keeps the reference's operation order (no fused multiply-adds), and with
packed FP32 instructions disabled.
"""
import os
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(ROOT, "csrc")
OUT = os.path.join(ROOT, "libbsam_gpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SOURCES = ["bsam_gpu.hip", "bsam_grid.cpp"]
FLAGS = ["-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
         "-Wno-unused-value", "-Wno-unused-result",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


FORTRAN = os.path.join(ROOT, "fortran")
FC = os.environ.get("FC", "/opt/rocm/bin/amdflang")
DRIVER = os.path.join(FORTRAN, "bsam_gpu_driver")


def build_fortran(verbose=False):
    """bsam_gpu_mod.f90 + bsam_gpu_driver.f90 -> fortran/bsam_gpu_driver,
    linked against libbsam_gpu.so (found next to the package at run time)."""
    if not os.path.exists(FC):
        if verbose:
            print("amdflang not found; Fortran driver not built")
        return None
    mod = os.path.join(FORTRAN, "_build")
    os.makedirs(mod, exist_ok=True)
    cmd = [FC, "-O2", "-module-dir", mod, "-o", DRIVER + ".tmp",
           os.path.join(FORTRAN, "bsam_gpu_mod.f90"), os.path.join(FORTRAN, "bsam_gpu_driver.f90"),
           "-L" + ROOT, "-lbsam_gpu", "-Wl,-rpath,$ORIGIN/.."]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("amdflang failed:\n" + p.stdout[-3000:] + p.stderr[-3000:])
    os.replace(DRIVER + ".tmp", DRIVER)
    if verbose:
        print("built", DRIVER)
    return DRIVER


def build(verbose=False):
    cmd = [HIPCC] + FLAGS + ["-o", OUT + ".tmp"] + [os.path.join(CSRC, s) for s in SOURCES]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + p.stderr[-6000:])
    os.replace(OUT + ".tmp", OUT)
    if verbose:
        print("built", OUT)
    build_fortran(verbose)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
