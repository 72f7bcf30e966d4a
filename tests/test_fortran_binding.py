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
    """C_SIZEOF and the byte offsets (via C_LOC) of bsam_patch_desc,
    bsam_params and bsam_grid_params seen from Fortran equal sizeof/offsetof from C."""
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
TYPE(bsam_grid_params), TARGET:: g
INTEGER(C_INTPTR_T):: b
b = TRANSFER(C_LOC(d), b)
PRINT '(6I8)', C_SIZEOF(d), TRANSFER(C_LOC(d%dx), b)-b, TRANSFER(C_LOC(d%xlower), b)-b, &
  C_SIZEOF(p), TRANSFER(C_LOC(p%dt), 0_C_INTPTR_T)-TRANSFER(C_LOC(p), 0_C_INTPTR_T), &
  TRANSFER(C_LOC(p%model_params), 0_C_INTPTR_T)-TRANSFER(C_LOC(p), 0_C_INTPTR_T)
PRINT '(3I8)', C_SIZEOF(g), TRANSFER(C_LOC(g%qtolerance), 0_C_INTPTR_T)-TRANSFER(C_LOC(g), 0_C_INTPTR_T), &
  TRANSFER(C_LOC(g%desiredfillratios), 0_C_INTPTR_T)-TRANSFER(C_LOC(g), 0_C_INTPTR_T)
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
  printf("%zu %zu %zu\n", sizeof(bsam_grid_params), offsetof(bsam_grid_params, qtolerance),
         offsetof(bsam_grid_params, desiredfillratios));
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


def _run_driver_amr(nd, n, steps, prec, maxlevel, tagtol, d):
    out = os.path.join(d, f"qa_{prec}.bin")
    p = subprocess.run([build.DRIVER, str(nd), str(n), str(steps), "5e-5", "0.015", "1e-8", "1003", "0.1", "0.05",
                        str(prec), out, str(maxlevel), repr(tagtol)], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, (p.stdout, p.stderr)
    rows = re.findall(r"step\s+(\d+)\s+vcycles\s+(\d+)\s+residual\s+([-+0-9.Ee]+)\s+patches\s+(\d+)\s+finest\s+(\d+)"
                      r"\s+finest-cells\s+(\d+)", p.stdout)
    assert len(rows) == steps, p.stdout
    return [(int(a), int(b), float(c), int(e), int(f), int(g)) for a, b, c, e, f, g in rows], \
        np.fromfile(out, dtype="<f8")


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("nd,n,tagtol", [(2, 64, 0.03), (3, 32, 0.04)])
def test_driver_amr_matches_python_host(nd, n, tagtol, prec, gpu_solver_cls):
    """The Fortran host's time steps with the regrid (bsam_time_step_amr
    with a bsam_grid_params filled in Fortran, the mesh read back with
    bsam_get_hierarchy) equal the same sequence driven from Python through
    ctypes: V-cycle counts and residuals bit for bit, the same mesh (patch
    count, finest level, cells on it) every step, the same root state."""
    from bsam_amd import abi
    _need_driver()
    print(f"[gpu] fortran driver AMR {nd}D {n} precision {prec}", flush=True)
    steps, maxlevel = 3, 2
    with tempfile.TemporaryDirectory() as d:
        rows, qf = _run_driver_amr(nd, n, steps, prec, maxlevel, tagtol, d)
    minlevel = -(int(np.log2(n)) - 1)
    hier = hierarchy.uniform(nd, (n,) * nd, minlevel, [2] * (2 * nd))
    root = len(hier) - 1
    S = gpu_solver_cls(configs.params(nd, 5e-5, 0.015, 1e-8, precision=prec, use_graph=1), hier)
    q0 = np.zeros(hierarchy.field_shape(hier[root], nd, 2, "q"))
    q0[(0,) + (slice(1, -1),) * nd] = configs.init_phi(hier[root], nd, 1, 1003, 0.1, 0.05)
    S.put(root, "q", q0)
    S.set_ghost(0, 0)
    S.copy_q_to_qold(0)
    S.autotune()
    grid = abi.grid_params(maxlevel, 11, tagtol, ibuffer=1, minimumgridpoints=8, desiredfillratio=0.8)
    refined = 0
    for s in range(steps):
        rc, ncyc, r = S.time_step_amr(grid, s == 0)
        h = S.hier
        fin = max(x["level"] for x in h)
        cells = sum(int(np.prod(x["mx"][:nd])) for x in h if x["level"] == fin)
        assert rows[s][1:] == (ncyc, r, len(h), fin, cells), (s, rows[s], (ncyc, r, len(h), fin, cells))
        refined = max(refined, fin)
    assert refined >= 1, "the tagging refined nothing"
    qp = S.get(root, "q").ravel()
    assert np.array_equal(qf, qp), float(np.max(np.abs(qf - qp)))
    S.close()
