"""The HIP path on meshes shaped like the 3D AMR workload: levels whose
parents are refined patches themselves (the reference's two-refined-level
fixtures ch3d_amr2 and ch3d_amr2_phys), in every launch mode the bench uses
(graph capture, autotuned and hand-set launch shapes), the frozen-mesh time
step on a refined mesh against the oracle, and one V-cycle on the 1883-patch
bench mesh itself against the oracle.  States bit for bit; residual norms to
1e-10 relative (their summation order differs)."""
import numpy as np
import pytest

import fixture_util as fu
import oracle
from bsam_amd import abi

pytestmark = pytest.mark.gpu
RTOL_NORM = 1e-10
AMR2 = ["ch3d_amr2", "ch3d_amr2_phys"]


def _hist_close(a, b):
    assert len(a) == len(b), (a, b)
    for x, y in zip(a, b):
        assert abs(x - y) <= RTOL_NORM * abs(y), (x, y)


def _check_end(S, meta, z, name, mode):
    st = meta["steps"][0]
    for i in st["end"]:
        got = S.get(i, "q").ravel()
        ref = fu.end_state(z, 0, i)
        assert np.array_equal(got, ref), (name, mode, i, float(np.max(np.abs(got - ref))))


@pytest.mark.parametrize("name", AMR2)
@pytest.mark.parametrize("mode", ["graph", "autotune", "knobs"])
def test_two_refined_levels_modes(name, mode, gpu_solver_cls):
    """MultigridIterations from the reference's state with hipGraph capture,
    after bsam_autotune, and with hand-set launch shapes (round-robin block
    order, runs of 256 blocks per XCD for the gathers, no coarse chain) on
    every level: the reference's end state bit for bit."""
    meta, z = fu.load(name)
    st = meta["steps"][0]
    print(f"[gpu] {name} {mode}", flush=True)
    S = gpu_solver_cls(fu.step_params(meta, use_graph=1 if mode != "knobs" else 0), st["hierarchy"])
    fu.load_state(S, meta, z, 0)
    if mode == "autotune":
        rep = S.autotune()
        # the tuner takes levels of at least 32768 cells
        big = {lv for lv in range(1, S.finestlevel + 1)
               if sum(int(np.prod(d["mx"])) for d in st["hierarchy"] if d["level"] == lv) >= 32768}
        assert {L["level"] for L in rep["levels"]} >= big, rep
    elif mode == "knobs":
        for lv in range(S.minlevel, S.finestlevel + 1):
            S.set_knob(abi.KNOB_ORDER, lv, 1)
            S.set_knob(abi.KNOB_GATHER_ORDER_FP64, lv, 256)
            S.set_knob(abi.KNOB_GATHER_BLOCK_FP64, lv, 1024 if lv % 2 else 128)
            assert S.get_knob(abi.KNOB_ORDER, lv) == 1
        S.set_knob(abi.KNOB_CHAIN_CELLS, 0, 0)
    rc, n, r, hist = S.multigrid_iterations()
    _hist_close(hist, st["residuals"])
    _check_end(S, meta, z, name, mode)
    S.close()


@pytest.mark.parametrize("name", AMR2)
def test_time_step_on_refined_mesh_vs_oracle(name, gpu_solver_cls):
    """bsam_time_step (the frozen-mesh time step: SetGhost from the root up,
    FillDown on the first step, CopyQToQold, MultigridIterations) against the
    oracle's, two steps from the reference's state, every patch of every
    level bit for bit."""
    meta, z = fu.load(name)
    st = meta["steps"][0]
    print(f"[gpu] time steps {name}", flush=True)
    out = []
    for S in (gpu_solver_cls(fu.step_params(meta), st["hierarchy"]),
              oracle.Oracle(fu.step_params(meta), st["hierarchy"])):
        fu.load_state(S, meta, z, 0)
        res = []
        for s in range(2):
            rc, n, r = S.time_step(s == 0)
            assert rc == abi.BSAM_OK, (s, rc, r)
            res.append((n, r))
        out.append((res, [S.get(k, "q") for k in range(len(st["hierarchy"]))]))
    (rg, qg), (ro, qo) = out
    for (ng, a), (no, b) in zip(rg, ro):
        assert ng == no and abs(a - b) <= RTOL_NORM * abs(b), (ng, no, a, b)
    for k in range(len(qg)):
        assert np.array_equal(qg[k], qo[k]), (name, k, float(np.max(np.abs(qg[k] - qo[k]))))


def test_bench_mesh_one_vcycle_vs_oracle(gpu_solver_cls):
    """One V-cycle of the first time step on the 3D AMR bench mesh (64^3
    root, three refined levels, 1883 patches, the reference's own regrid of
    the workload's initial data), batched with graph capture, against the
    oracle: every patch bit for bit, the residual norm to 1e-10."""
    from bsam_amd.workload import Workload
    print("[gpu] bench mesh: building the workload", flush=True)
    wl = Workload("ch3d_amr_64_l3", 256)
    res = []
    for make in (gpu_solver_cls, oracle.Oracle):
        p = wl.params("fp64", 1, 1)
        p.maxvcycles = 1
        S = make(p, wl.hier)
        wl.load(S)
        print(f"[gpu] bench mesh: one V-cycle on {type(S).__name__}", flush=True)
        rc, n, r = S.time_step(True)
        assert rc == abi.BSAM_NOT_CONVERGED and n == 1, (rc, n, r)
        res.append((r, [S.get(k, "q") for k in range(len(wl.hier))]))
    (rg, qg), (ro, qo) = res
    assert abs(rg - ro) <= RTOL_NORM * abs(ro), (rg, ro)
    for k in range(len(qg)):
        assert np.array_equal(qg[k], qo[k]), (k, float(np.max(np.abs(qg[k] - qo[k]))))


@pytest.mark.parametrize("name", AMR2)
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_sweep_images_match_gather(name, precision, gpu_solver_cls):
    """The colour sweeps write the copies of the reduced fill after them
    (BSAM_KNOB_SWEEP_IMAGES 1, the default) or leave them to the gather
    (0): two time steps from the reference's state give the same bits
    either way, in FP64 and in the mixed path."""
    meta, z = fu.load(name)
    st = meta["steps"][0]
    print(f"[gpu] sweep images {name} {precision}", flush=True)
    kw = {"precision": abi.PREC_MIXED} if precision == "mixed" else {}
    out = []
    for images in (1, 0):
        S = gpu_solver_cls(fu.step_params(meta, **kw), st["hierarchy"])
        S.set_knob(abi.KNOB_SWEEP_IMAGES, 0, images)
        assert S.get_knob(abi.KNOB_SWEEP_IMAGES) == images
        fu.load_state(S, meta, z, 0)
        res = [S.time_step(s == 0)[1:] for s in range(2)]
        out.append((res, [S.get(k, "q") for k in range(len(st["hierarchy"]))]))
        S.close()
    (ra, qa), (rb, qb) = out
    assert ra == rb, (ra, rb)
    for k in range(len(qa)):
        assert np.array_equal(qa[k], qb[k]), (name, precision, k, float(np.max(np.abs(qa[k] - qb[k]))))
