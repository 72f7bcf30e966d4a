"""bsam_amr and bsam_time_step_amr: the reference's regrid pass
(AMR(rootlevel), bsamroutines.f90:966-991: SetGhost, tagging, clustering,
InitFields, TransferValues, level by level) and its time step
(:672-758) on the device, against the reference's own runs.

From the end state of a fixture's first time step, with the probe module's
tagging (errflagopt 11, tagtol) and the run's grid parameters (ibuffer 1,
minimumgridpoints 4, desiredfillratio 0.8):

* bsam_amr builds the reference's mesh of the second step, descriptor for
  descriptor in the reference's visiting order, and the fields the
  reference held right after its AMR (its qold of the second step), every
  patch bit for bit;
* bsam_time_step_amr (AMR, CopyQToQold, MultigridIterations) ends on the
  reference's end state of the second step, bit for bit, with the same
  V-cycle count;
* a second pass from that mesh keeps it valid and leaves the levels at and
  below the root as they were.
"""
import numpy as np
import pytest

import fixture_util as fu
from bsam_amd import abi
from bsam_amd.gpu import BsamError

pytestmark = pytest.mark.gpu
# maxlevel and tagtol of the two-step AMR fixtures (tests/golden/make_golden.py)
RUNS = {"ch2d_amr": (2, 0.05), "ch3d_amr2_phys": (2, 0.3), "tm2d_amr_phys": (2, 0.05)}
KEYS = ("level", "parent", "mx", "mglobal", "mbounds", "mthbc", "dx", "xlower")


def _start(make, name):
    meta, z = fu.load(name)
    s0 = meta["steps"][0]
    S = make(fu.step_params(meta), s0["hierarchy"])
    for i in s0["end"]:
        S.put(i, "q", fu.end_state(z, 0, i).reshape(S.get(i, "q").shape))
    # below the root, q is the FillDown that ended the step's last iteration
    S.fill_down(abi.SOLUTIONFIELD)
    maxlevel, tagtol = RUNS[name]
    return S, meta, z, abi.grid_params(maxlevel, 11, tagtol, ibuffer=1, minimumgridpoints=4, desiredfillratio=0.8)


def _same_mesh(got, want):
    assert len(got) == len(want), (len(got), len(want))
    for k, (g, w) in enumerate(zip(got, want)):
        for key in KEYS:
            assert g[key] == w[key], (k, key, g[key], w[key])


@pytest.mark.parametrize("name", list(RUNS))
def test_amr_builds_the_reference_mesh(name, gpu_solver_cls):
    print(f"[gpu] amr {name}", flush=True)
    S, meta, z, grid = _start(gpu_solver_cls, name)
    s1 = meta["steps"][1]
    S.amr(grid)
    _same_mesh(S.hier, s1["hierarchy"])
    for i in s1["src"]:
        got = S.get(i, "q").ravel()
        ref = z[f"s1_qold_p{i}"]
        assert np.array_equal(got, ref), ("after AMR", i, float(np.max(np.abs(got - ref))))
    # a second pass from the new mesh: the root levels stay, the mesh is valid
    root = [S.get(i, "q") for i, d in enumerate(S.hier) if d["level"] <= 0]
    S.amr(grid)
    again = [S.get(i, "q") for i, d in enumerate(S.hier) if d["level"] <= 0]
    assert all(np.array_equal(a, b) for a, b in zip(root, again)) and len(root) == len(again)
    assert S.finestlevel <= grid.maxlevel
    S.close()


@pytest.mark.parametrize("tables", [1, 0])
@pytest.mark.parametrize("name", list(RUNS))
def test_time_step_amr_matches_reference(name, tables, gpu_solver_cls):
    """tables 0: the levels AMR builds keep the staged ghost fill
    (BSAM_KNOB_REGRID_TABLES), which must give the same bits."""
    print(f"[gpu] time step with AMR {name} tables {tables}", flush=True)
    S, meta, z, grid = _start(gpu_solver_cls, name)
    S.set_knob(abi.KNOB_REGRID_TABLES, 0, tables)
    assert S.get_knob(abi.KNOB_REGRID_TABLES, 0) == tables
    s1 = meta["steps"][1]
    rc, n, r = S.time_step_amr(grid, False)
    assert rc == abi.BSAM_OK, (rc, n, r)
    assert n == len(s1["residuals"]), (n, s1["residuals"])
    assert abs(r - s1["residuals"][-1]) <= 1e-10 * abs(s1["residuals"][-1]), (r, s1["residuals"][-1])
    _same_mesh(S.hier, s1["hierarchy"])
    for i in s1["end"]:
        got = S.get(i, "q").ravel()
        ref = fu.end_state(z, 1, i)
        assert np.array_equal(got, ref), ("step end", i, float(np.max(np.abs(got - ref))))
    S.close()


def test_amr_refuses_bad_grid_parameters(gpu_solver_cls):
    S, meta, z, grid = _start(gpu_solver_cls, "ch2d_amr")
    grid.minimumgridpoints[0] = 5
    with pytest.raises(BsamError) as ei:
        S.amr(grid)
    assert ei.value.code == abi.BSAM_EINVAL and "minimumgridpoints" in str(ei.value)
    # refused before anything changed
    s0 = meta["steps"][0]
    i = s0["end"][0]
    assert np.array_equal(S.get(i, "q").ravel(), fu.end_state(z, 0, i))
    S.close()


# The bench workloads' meshes and grid parameters (bench.py REGRID): levels
# of up to 256^3 cells whose tags AMR buffers on the device.
BENCH_GRIDS = {"ch3d_amr_64_l3": 0.05, "tm3d_amr_64_l3": 0.05, "ch2d_amr_128_l2": 0.05}


@pytest.mark.parametrize("name", list(BENCH_GRIDS))
def test_amr_device_buffering_matches_host_clustering(name, gpu_solver_cls):
    """bsam_amr buffers a level's tags on the device (a dilation over the
    level's bit map) and clusters without buffering; bsam_new_level
    buffers on the host (BufferAndList, InflateEdgeTags).  After a pass,
    each level is in the state it was tagged on (AMR builds upward and
    SetGhost(level, 0) is idempotent), so the host path on the same flags
    must rebuild every new level descriptor for descriptor."""
    from bsam_amd.grid import new_level
    from bsam_amd.workload import Workload
    print(f"[gpu] device tag buffering {name}", flush=True)
    wl = Workload(name, 256)
    S = gpu_solver_cls(wl.params("fp64", 0, 1), wl.hier)
    wl.load(S)
    grid = abi.grid_params(wl.finest, 11, BENCH_GRIDS[name], ibuffer=1, minimumgridpoints=8, desiredfillratio=0.8)
    S.amr(grid)
    hier = S.hier
    for lv in range(0, S.finestlevel):
        S.set_ghost(lv, 0)
        flags, ntag = S.estimate_level_errors(lv, 11, BENCH_GRIDS[name])
        assert ntag > 0
        want = new_level(wl.nd, hier, lv, flags, 1, 8, 0.8)
        got = [d for d in hier if d["level"] == lv + 1]
        _same_mesh(got, want)
    S.close()
