"""The Fortran host binding (bsam1.2_amd/fortran): the ISO_C_BINDING types
have the C layouts of include/bsam_gpu.h, the driver links against the
library and reports a missing GPU through the status convention, and on a
GPU its time steps equal the oracle's."""
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from bsam_amd import build, configs, hierarchy

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bsam_gpu.h")
MODDIR = os.path.join(build.FORTRAN, "_build")


def _need_driver():
    if not os.path.exists(build.DRIVER):
        pytest.skip("Fortran driver not built (amdflang absent)")


def test_bind_c_layouts_match_header():
    """C_SIZEOF and the byte offsets (via C_LOC) of bsam_patch_desc and
    bsam_params seen from Fortran equal sizeof/offsetof from C."""
    _need_driver()
    if not os.path.exists(build.FC):
        pytest.skip("amdflang absent")
    fsrc = """
PROGRAM layout
USE, INTRINSIC:: ISO_C_BINDING
USE BSAMGPU
IMPLICIT NONE
TYPE(bsam_patch_desc), TARGET:: d
TYPE(bsam_params), TARGET:: p
INTEGER(C_INTPTR_T):: b
b = TRANSFER(C_LOC(d), b)
PRINT '(6I8)', C_SIZEOF(d), TRANSFER(C_LOC(d%dx), b)-b, TRANSFER(C_LOC(d%xlower), b)-b, &
  C_SIZEOF(p), TRANSFER(C_LOC(p%dt), 0_C_INTPTR_T)-TRANSFER(C_LOC(p), 0_C_INTPTR_T), &
  TRANSFER(C_LOC(p%model_params), 0_C_INTPTR_T)-TRANSFER(C_LOC(p), 0_C_INTPTR_T)
END PROGRAM layout
"""
    csrc = r"""
#include <stdio.h>
#include <stddef.h>
#include "bsam_gpu.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(bsam_patch_desc), offsetof(bsam_patch_desc, dx),
         offsetof(bsam_patch_desc, xlower), sizeof(bsam_params), offsetof(bsam_params, dt),
         offsetof(bsam_params, model_params));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "l.f90"), "w").write(fsrc)
        open(os.path.join(d, "l.c"), "w").write(csrc)
        subprocess.run([build.FC, "-module-dir", d, "-I", MODDIR, "-o", os.path.join(d, "lf"),
                        os.path.join(build.FORTRAN, "bsam_gpu_mod.f90"), os.path.join(d, "l.f90"),
                        "-L" + build.ROOT, "-lbsam_gpu", "-Wl,-rpath," + build.ROOT],
                       check=True, capture_output=True)
        subprocess.run(["gcc", "-I", os.path.dirname(HEADER), os.path.join(d, "l.c"), "-o", os.path.join(d, "lc")],
                       check=True)
        f = list(map(int, subprocess.run([os.path.join(d, "lf")], capture_output=True, text=True).stdout.split()))
        c = list(map(int, subprocess.run([os.path.join(d, "lc")], capture_output=True, text=True).stdout.split()))
    assert f == c, (f, c)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is visible; the GPU test covers the driver")
def test_driver_reports_missing_gpu():
    """Without a GPU, bsam_create fails with BSAM_ENODEV and the driver
    stops with its status (3) after printing the library's message; the
    library itself never exits."""
    _need_driver()
    with tempfile.TemporaryDirectory() as d:
        p = subprocess.run([build.DRIVER, "3", "8", "1", "5e-5", "0.015", "1e-8", "1003", "0.1", "0.05", "0",
                            os.path.join(d, "q.bin")], capture_output=True, text=True, timeout=120)
    assert p.returncode == 3, (p.returncode, p.stdout, p.stderr)
    assert "bsam_create failed: status  -6" in p.stdout


def _run_driver(nd, n, steps, prec, d):
    out = os.path.join(d, f"q_{prec}.bin")
    p = subprocess.run([build.DRIVER, str(nd), str(n), str(steps), "5e-5", "0.015", "1e-8", "1003", "0.1", "0.05",
                        str(prec), out], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, (p.stdout, p.stderr)
    rows = re.findall(r"step\s+(\d+)\s+vcycles\s+(\d+)\s+residual\s+([-+0-9.Ee]+)", p.stdout)
    return [(int(a), int(b), float(c)) for a, b, c in rows], np.fromfile(out, dtype="<f8")


@pytest.mark.gpu
@pytest.mark.parametrize("nd,n", [(3, 32), (2, 64)])
def test_driver_matches_oracle(nd, n):
    """The Fortran host's FP64 time steps through the C ABI are bit-identical
    to the oracle's (V-cycle counts, state) with residuals to 1e-10
    relative; its mixed-precision run is within 1e-7 of it."""
    import oracle
    _need_driver()
    print(f"[gpu] fortran driver {nd}D {n}", flush=True)
    steps = 2
    with tempfile.TemporaryDirectory() as d:
        rows, qg = _run_driver(nd, n, steps, 0, d)
        rows_m, qm = _run_driver(nd, n, steps, 1, d)
    minlevel = -(int(np.log2(n)) - 1)
    hier = hierarchy.uniform(nd, (n,) * nd, minlevel, [2] * (2 * nd))
    root = len(hier) - 1
    O = oracle.Oracle(configs.params(nd, 5e-5, 0.015, 1e-8), hier)
    q0 = np.zeros(hierarchy.field_shape(hier[root], nd, 2, "q"))
    q0[(0,) + (slice(1, -1),) * nd] = configs.init_phi(hier[root], nd, 1, 1003, 0.1, 0.05)
    O.put(root, "q", q0)
    O.set_ghost(0, 0)
    O.copy_q_to_qold(0)
    for s in range(steps):
        rc, ncyc, r = O.time_step(s == 0)
        assert rows[s][1] == ncyc, (rows, s, ncyc)
        assert abs(rows[s][2] - r) <= 1e-10 * abs(r), (rows[s], r)
    qo = O.get(root, "q").ravel()
    assert np.array_equal(qg, qo), float(np.max(np.abs(qg - qo)))
    assert float(np.max(np.abs(qm - qo))) < 1e-7
