"""bsam_regrid (MakeNewGrid's field set-up, InitFields with the P1MC mass
correction, TransferValues, SetGhost level by level) on the GPU.

* The reference's own regrids: from the end state of the first time step on
  the first step's mesh onto the second step's mesh, the state right after
  the regrid (the reference's qold of the second step) and after FillDown
  (its q at MultigridIterations entry) bit for bit; then CopyQToQold and
  MultigridIterations give the reference's end state of the second step.
* The 1883-patch 3D AMR bench mesh against the oracle: a regrid that drops
  every other patch of the finest level and one that brings them back (so
  those patches take InitFields' values), every patch's q and qold bit for
  bit.
"""
import numpy as np
import pytest

import fixture_util as fu
import oracle
from bsam_amd import abi

pytestmark = pytest.mark.gpu
TWO_STEP = [n for n in fu.NAMES + fu.TUMOUR_NAMES if len(fu.load(n)[0]["steps"]) >= 2]


def _from_first_step(make, meta, z):
    s0 = meta["steps"][0]
    S = make(fu.step_params(meta), s0["hierarchy"])
    for i in s0["end"]:
        S.put(i, "q", fu.end_state(z, 0, i).reshape(S.get(i, "q").shape))
    return S


@pytest.mark.parametrize("tables", [1, 0])
@pytest.mark.parametrize("name", TWO_STEP)
def test_regrid_matches_reference(name, tables, gpu_solver_cls):
    """tables 0: the rebuilt levels keep the staged fill, with interpolation
    rules built on the device (BSAM_KNOB_REGRID_TABLES)."""
    meta, z = fu.load(name)
    s1 = meta["steps"][1]
    print(f"[gpu] regrid {name} tables {tables}", flush=True)
    S = _from_first_step(gpu_solver_cls, meta, z)
    S.set_knob(abi.KNOB_REGRID_TABLES, 0, tables)
    S.regrid(s1["hierarchy"])
    for i in s1["src"]:
        got = S.get(i, "q").ravel()
        ref = z[f"s1_qold_p{i}"]
        assert np.array_equal(got, ref), ("after regrid", i, float(np.max(np.abs(got - ref))))
    # the start of the reference's step: CopyQToQold on every level, then
    # MultigridIterations, which starts with FillDown
    for lv in range(S.minlevel, S.finestlevel + 1):
        S.copy_q_to_qold(lv)
    if meta.get("tumour"):
        fu.below_root_qold(S, meta, z, 1)
    S.fill_down(abi.SOLUTIONFIELD)
    for i in s1["src"]:
        got = S.get(i, "q").ravel()
        ref = z[f"s1_q_p{i}"]
        assert np.array_equal(got, ref), ("after FillDown", i, float(np.max(np.abs(got - ref))))
    rc, n, r, hist = S.multigrid_iterations()
    assert rc == abi.BSAM_OK, (rc, n, r)
    assert len(hist) == len(s1["residuals"]), (hist, s1["residuals"])
    for i in s1["end"]:
        got = S.get(i, "q").ravel()
        ref = fu.end_state(z, 1, i)
        assert np.array_equal(got, ref), ("step end", i, float(np.max(np.abs(got - ref))))
    S.close()


def test_regrid_keeps_the_root_and_refuses_bad_meshes(gpu_solver_cls):
    """A new mesh whose root extent differs is refused before anything
    changes: the context keeps its mesh and fields."""
    from bsam_amd.gpu import BsamError
    meta, z = fu.load("ch2d_amr")
    s0 = meta["steps"][0]
    S = _from_first_step(gpu_solver_cls, meta, z)
    bad = [dict(d) for d in s0["hierarchy"]]
    root = [i for i, d in enumerate(bad) if d["level"] == 0][0]
    bad[root] = dict(bad[root], mx=[bad[root]["mx"][0] * 2] + bad[root]["mx"][1:])
    with pytest.raises(BsamError) as ei:
        S.regrid(bad)
    assert ei.value.code == abi.BSAM_EINVAL and "extent" in str(ei.value)
    # refused before anything changed: the first step's state is still there
    i = s0["end"][0]
    assert np.array_equal(S.get(i, "q").ravel(), fu.end_state(z, 0, i))
    S.close()


_BENCH = {}


def _bench_mesh(make, tables=None):
    from bsam_amd.workload import Workload
    if "wl" not in _BENCH:
        print("[gpu] regrid bench mesh: building the workload", flush=True)
        wl = Workload("ch3d_amr_64_l3", 256)
        full = wl.hier
        finest = max(d["level"] for d in full)
        # drop every other patch of the finest level (they have no children)
        keep = [i for i, d in enumerate(full) if d["level"] < finest or i % 2 == 0]
        remap = {old: new for new, old in enumerate(keep)}
        thin = [dict(full[i], parent=remap.get(full[i]["parent"], -1) if full[i]["parent"] >= 0 else -1)
                for i in keep]
        _BENCH.update(wl=wl, full=full, thin=thin)
    wl, full, thin = _BENCH["wl"], _BENCH["full"], _BENCH["thin"]
    S = make(wl.params("fp64", 1, 1), full)
    wl.load(S)
    if tables is not None:
        S.set_knob(abi.KNOB_REGRID_TABLES, 0, tables)
    print(f"[gpu] regrid bench mesh on {type(S).__name__}", flush=True)
    S.regrid(thin)
    a = [(S.get(k, "q"), S.get(k, "qold")) for k in range(len(thin))]
    S.regrid(full)
    b = [(S.get(k, "q"), S.get(k, "qold")) for k in range(len(full))]
    return a, b


@pytest.mark.parametrize("tables", [1, 0])
def test_regrid_bench_mesh_vs_oracle(tables, gpu_solver_cls):
    if "oracle" not in _BENCH:
        _BENCH["oracle"] = _bench_mesh(oracle.Oracle)
    (ga, gb), (oa, ob) = _bench_mesh(gpu_solver_cls, tables), _BENCH["oracle"]
    for tag, g, o in (("thin", ga, oa), ("full", gb, ob)):
        for k in range(len(g)):
            for f in range(2):
                assert np.array_equal(g[k][f], o[k][f]), (tag, k, f, float(np.max(np.abs(g[k][f] - o[k][f]))))
