"""HIP path against the reference's golden vectors and the CPU oracle.

Every kernel except the residual-norm reduction keeps the reference's
floating-point operation order (compiled with -ffp-contract=off), so states
must agree bit for bit.  Residual norms are sums of n = nrvars x cells
squares taken in a different order (the reference sums sequentially, whose
rounding error is bounded by n*u, u = 1.1e-16); they must agree to 1e-10
relative, which covers n up to ~1e6 with margin and is far below the 1e-8
stopping tolerance.
"""
import numpy as np
import pytest

import fixture_util as fu
import oracle
from bsam_amd import abi, configs, hierarchy

pytestmark = pytest.mark.gpu
RTOL_NORM = 1e-10


def _hist_close(a, b):
    assert len(a) == len(b), (a, b)
    for x, y in zip(a, b):
        assert abs(x - y) <= RTOL_NORM * abs(y), (x, y)


@pytest.mark.parametrize("name", fu.NAMES + fu.TUMOUR_NAMES)
@pytest.mark.parametrize("batching", [1, 0])
def test_golden_multigrid_iterations(name, batching, gpu_solver_cls):
    """MultigridIterations on the reference's states: batched per-level
    launches and one launch per patch give the reference's end states bit
    for bit."""
    meta, z = fu.load(name)
    for s, st in enumerate(meta["steps"]):
        print(f"[gpu] {name} step {s} batching {batching}", flush=True)
        S, hist = fu.run_step(gpu_solver_cls, meta, z, s, batching=batching)
        _hist_close(hist, st["residuals"])
        for i in st["end"]:
            got = S.get(i, "q").ravel()
            ref = fu.end_state(z, s, i)
            assert np.array_equal(got, ref), (name, s, i, float(np.max(np.abs(got - ref))))
        S.close()


@pytest.mark.parametrize("name", fu.NAMES + fu.TUMOUR_NAMES)
def test_golden_sweeps(name, gpu_solver_cls):
    meta, z = fu.load(name)
    print(f"[gpu] sweeps {name}", flush=True)

    def cmp(S, pid, golden, rec):
        got = S.get(pid, "q").ravel()
        assert np.array_equal(got, golden), (name, rec, float(np.max(np.abs(got - golden))))

    n = fu.replay_sweeps(gpu_solver_cls, meta, z, 0, cmp)
    assert n == len(meta["steps"][0]["sweeps"]) > 0


def _uniform_pair(nd, n, eps, dt, seed, base, gpu_cls, **kw):
    minlevel = -(int(np.log2(n)) - 1)
    hier = hierarchy.uniform(nd, (n,) * nd, minlevel, [2] * (2 * nd))
    p = configs.params(nd, dt, eps, 1e-9, **kw)
    root = len(hier) - 1
    phi = configs.init_phi(hier[root], nd, 1, seed, base, 0.05)
    q0 = np.zeros((2,) + (n + 2,) * nd)
    q0[(0,) + (slice(1, -1),) * nd] = phi
    out = []
    for S in (gpu_cls(p, hier), oracle.Oracle(configs.params(nd, dt, eps, 1e-9), hier)):
        S.put(root, "q", q0)
        S.set_ghost(0, 0)
        S.copy_q_to_qold(0)
        hists = []
        for step in range(2):
            S.set_ghost(0, 0)
            if step == 0:
                S.fill_down(abi.SOLUTIONFIELD)
            for lv in range(minlevel, 1):
                S.copy_q_to_qold(lv)
            rc, ncyc, r, h = S.multigrid_iterations()
            hists.append(h)
        out.append((hists, S.get(root, "q")))
    return out


@pytest.mark.parametrize("nd,n", [(2, 128), (3, 32)])
def test_uniform_two_steps_vs_oracle(nd, n, gpu_solver_cls):
    print(f"[gpu] uniform {nd}D {n}", flush=True)
    (hg, qg), (ho, qo) = _uniform_pair(nd, n, 0.03 if nd == 2 else 0.05, 1e-4, 1001, 0.0, gpu_solver_cls)
    for a, b in zip(hg, ho):
        _hist_close(a, b)
    assert np.array_equal(qg, qo), float(np.max(np.abs(qg - qo)))


def test_graph_replay_matches_eager(gpu_solver_cls):
    """The hipGraph of the V-cycle loop body, replayed once per V-cycle with
    host synchronisation in between, gives the eager result bit for bit."""
    print("[gpu] graph vs eager", flush=True)
    (hg, qg), _ = _uniform_pair(3, 32, 0.05, 1e-4, 1003, 0.1, gpu_solver_cls, use_graph=1)
    (he, qe), _ = _uniform_pair(3, 32, 0.05, 1e-4, 1003, 0.1, gpu_solver_cls, use_graph=0)
    assert hg == he
    assert np.array_equal(qg, qe)


def test_graph_replay_amr(gpu_solver_cls):
    meta, z = fu.load("ch2d_amr")
    st = meta["steps"][0]
    print("[gpu] graph amr", flush=True)
    S, hist = fu.run_step(gpu_solver_cls, meta, z, 0, use_graph=1)
    _hist_close(hist, st["residuals"])
    for i in st["end"]:
        assert np.array_equal(S.get(i, "q").ravel(), fu.end_state(z, 0, i))


@pytest.mark.parametrize("nd,n,steps", [(3, 32, 3), (2, 64, 3)])
def test_mixed_precision_matches_fp64(nd, n, steps, gpu_solver_cls):
    """Mixed mode (FP32 correction V-cycles, FP64 residual and update)
    converges to the FP64 path's solution: after each time step both stop at
    RMS residual < 1e-9 and the max-norm difference of phi and mu is within
    a small multiple of the tolerance."""
    print(f"[gpu] mixed {nd}D {n}", flush=True)
    minlevel = -(int(np.log2(n)) - 1)
    hier = hierarchy.uniform(nd, (n,) * nd, minlevel, [2] * (2 * nd))
    root = len(hier) - 1
    phi = configs.init_phi(hier[root], nd, 1, 1003, 0.1, 0.05)
    q0 = np.zeros((2,) + (n + 2,) * nd)
    q0[(0,) + (slice(1, -1),) * nd] = phi
    outs = {}
    for prec in (abi.PREC_FP64, abi.PREC_MIXED):
        S = gpu_solver_cls(configs.params(nd, 1e-4, 0.05, 1e-9, precision=prec), hier)
        S.put(root, "q", q0)
        S.set_ghost(0, 0)
        S.copy_q_to_qold(0)
        cyc = []
        for s in range(steps):
            rc, ncyc, r = S.time_step(s == 0)
            assert rc == abi.BSAM_OK and r < 1e-9, (prec, s, rc, r)
            cyc.append(ncyc)
        outs[prec] = (S.get(root, "q"), cyc)
        S.close()
    qa, ca = outs[abi.PREC_FP64]
    qb, cb = outs[abi.PREC_MIXED]
    err = float(np.max(np.abs(qa - qb)))
    print(f"[gpu] mixed max-norm diff {err:.3e}, cycles fp64 {ca} mixed {cb}", flush=True)
    assert err < 1e-7
    assert max(cb) <= max(ca) + 2


@pytest.mark.parametrize("name", ["ch2d_amr", "ch3d_amr", "ch2d_amr_phys", "ch3d_amr_phys", "ch3d_amr2",
                                  "ch3d_amr2_phys"] + fu.TUMOUR_NAMES)
def test_mixed_precision_on_amr_matches_fp64(name, gpu_solver_cls):
    """Mixed mode on the reference's AMR hierarchies (flattened ghost fill
    of the correction, q += e fused into the finest level's last sweep,
    staged fill where a level cannot be flattened): from the reference's
    state at the first step, both paths stop below the fixture's tolerance
    and agree on every level to a small multiple of it.  The tumour cases
    run the four-field correction cycle (linearised phi^3, lagged t0)."""
    meta, z = fu.load(name)
    print(f"[gpu] mixed amr {name}", flush=True)
    outs = {}
    for prec in (abi.PREC_FP64, abi.PREC_MIXED):
        st = meta["steps"][0]
        S = gpu_solver_cls(fu.step_params(meta, precision=prec), st["hierarchy"])
        fu.load_state(S, meta, z, 0)
        rc, ncyc, r, hist = S.multigrid_iterations()
        assert rc == abi.BSAM_OK and r < meta["tol"], (prec, rc, r)
        outs[prec] = ({i: S.get(i, "q") for i in st["end"]}, ncyc)
        S.close()
    qa, ca = outs[abi.PREC_FP64]
    qb, cb = outs[abi.PREC_MIXED]
    err = max(float(np.max(np.abs(qa[i] - qb[i]))) for i in qa)
    print(f"[gpu] mixed amr {name} max-norm diff {err:.3e}, cycles fp64 {ca} mixed {cb}", flush=True)
    assert err < 1e-7
    assert cb <= ca + 2


@pytest.mark.parametrize("shape,mthbc,chain", [((128, 128, 128), (2, 2, 2, 2, 2, 2), None),
                                                ((128, 128, 128), (2, 2, 2, 2, 10, 10), None),
                                                ((200, 36, 16), (2, 2, 2, 2, 2, 2), None),
                                                ((200, 36, 16), (2, 2, 2, 2, 2, 2), 32768),
                                                ((200, 36, 16), (2, 2, 10, 10, 2, 2), None)])
def test_large_patch_sweep_vs_oracle(shape, mthbc, chain, gpu_solver_cls):
    """A root large enough for the marching 3D sweep and residual on the
    finest level (with the periodic ghost images written by the sweep when
    the patch is periodic in every direction).  200 x 36 x 16 leaves partial
    tiles in x (100 = 64 + 36 half-rows) and y (36 = 4 x 8 + 4 rows); its
    coarse levels 100 x 18 x 8 and 50 x 9 x 4 end on an odd extent, whose
    sweeps wrap after the sweep instead of writing images.  With the coarse
    chain starting at 32^3 cells (chain=32768) both levels run in the
    one-workgroup chain.  One time step, bit-identical to the oracle."""
    nd = 3
    print(f"[gpu] {shape} sweep mthbc={mthbc} chain={chain}", flush=True)
    minlevel = -(int(np.log2(min(shape))) - 2)
    hier = hierarchy.uniform(nd, shape, minlevel, list(mthbc))
    root = len(hier) - 1
    phi = configs.init_phi(hier[root], nd, 1, 1003, 0.1, 0.05)
    q0 = np.zeros(hierarchy.field_shape(hier[root], nd, 2, "q"))
    q0[0, 1:-1, 1:-1, 1:-1] = phi
    res = []
    for S in (gpu_solver_cls(configs.params(nd, 5e-5, 0.015, 1e-8), hier),
              oracle.Oracle(configs.params(nd, 5e-5, 0.015, 1e-8), hier)):
        if chain is not None and hasattr(S, "set_knob"):
            S.set_knob(abi.KNOB_CHAIN_CELLS, 0, chain)
        S.put(root, "q", q0)
        S.set_ghost(0, 0)
        S.copy_q_to_qold(0)
        S.set_ghost(0, 0)
        S.fill_down(abi.SOLUTIONFIELD)
        for lv in range(minlevel, 1):
            S.copy_q_to_qold(lv)
        rc, ncyc, r, h = S.multigrid_iterations()
        S.residual(0)
        res.append((h, S.get(root, "q"), S.get(root, "rf")))
    (hg, qg, rg), (ho, qo, ro) = res
    assert np.array_equal(qg, qo), float(np.max(np.abs(qg - qo)))
    assert np.array_equal(rg, ro), float(np.max(np.abs(rg - ro)))
    _hist_close(hg, ho)


@pytest.mark.parametrize("shape,mthbc", [((128, 128, 128), (2, 2, 2, 2, 2, 2)),
                                          ((96, 64, 48), (2, 2, 10, 10, 2, 2))])
def test_autotuned_shapes_vs_oracle(shape, mthbc, gpu_solver_cls):
    """bsam_autotune changes launch shapes only: after tuning (block orders
    of the level kernels and ghost gather, planes per workgroup of the
    marching sweeps and residuals) a time step is still bit-identical to the
    oracle, in FP64; the report lists the tuned levels and their choices."""
    nd = 3
    print(f"[gpu] autotune {shape} mthbc={mthbc}", flush=True)
    minlevel = -(int(np.log2(min(shape))) - 2)
    hier = hierarchy.uniform(nd, shape, minlevel, list(mthbc))
    root = len(hier) - 1
    phi = configs.init_phi(hier[root], nd, 1, 1007, 0.1, 0.05)
    q0 = np.zeros(hierarchy.field_shape(hier[root], nd, 2, "q"))
    q0[0, 1:-1, 1:-1, 1:-1] = phi
    res = []
    for S in (gpu_solver_cls(configs.params(nd, 5e-5, 0.015, 1e-8), hier),
              oracle.Oracle(configs.params(nd, 5e-5, 0.015, 1e-8), hier)):
        S.put(root, "q", q0)
        S.set_ghost(0, 0)
        S.copy_q_to_qold(0)
        S.fill_down(abi.SOLUTIONFIELD)
        for lv in range(minlevel, 1):
            S.copy_q_to_qold(lv)
        if hasattr(S, "autotune"):
            rep = S.autotune()
            assert rep["levels"] and rep["levels"][0]["cells"] >= 32768
            assert all(any(k.endswith("_choice") for k in L) for L in rep["levels"]), rep
            # the fields are untouched by the tuning
            assert np.array_equal(S.get(root, "q")[(slice(None),) + (slice(1, -1),) * nd],
                                  q0[(slice(None),) + (slice(1, -1),) * nd])
        rc, ncyc, r, h = S.multigrid_iterations()
        res.append((h, S.get(root, "q")))
    (hg, qg), (ho, qo) = res
    assert np.array_equal(qg, qo), float(np.max(np.abs(qg - qo)))
    _hist_close(hg, ho)


@pytest.mark.parametrize("shape,chain", [((32, 32, 32), None), ((64, 32, 16), None), ((48, 24, 20), None),
                                         ((48, 24, 20), 32768), ((96, 32), None), ((40, 24), None),
                                         ((40, 24), 0)])
def test_coarse_chain_vs_oracle(shape, chain, gpu_solver_cls):
    """Roots whose coarse levels run in the one-workgroup V-cycle chain:
    cubic, non-cubic (64 x 32 x 16: levels 32 x 16 x 8 and 16 x 8 x 4), an
    odd coarsest extent (48 x 24 x 20 ends at 12 x 6 x 5, whose sweeps wrap
    after the sweep), the whole 48 x 24 x 20 hierarchy in one chain (cap
    32^3 cells), 2D, and no chain at all (cap 0).  One time step,
    bit-identical to the oracle."""
    nd = len(shape)
    print(f"[gpu] chain {shape} cap={chain}", flush=True)
    minlevel = -(int(np.log2(min(shape))) - 2)
    hier = hierarchy.uniform(nd, shape, minlevel, [2] * (2 * nd))
    root = len(hier) - 1
    phi = configs.init_phi(hier[root], nd, 1, 1005, 0.1, 0.05)
    q0 = np.zeros(hierarchy.field_shape(hier[root], nd, 2, "q"))
    q0[(0,) + (slice(1, -1),) * nd] = phi
    res = []
    for S in (gpu_solver_cls(configs.params(nd, 5e-5, 0.015, 1e-9), hier),
              oracle.Oracle(configs.params(nd, 5e-5, 0.015, 1e-9), hier)):
        if chain is not None and hasattr(S, "set_knob"):
            S.set_knob(abi.KNOB_CHAIN_CELLS, 0, chain)
            assert S.get_knob(abi.KNOB_CHAIN_CELLS) == chain
        S.put(root, "q", q0)
        S.set_ghost(0, 0)
        S.copy_q_to_qold(0)
        S.fill_down(abi.SOLUTIONFIELD)
        for lv in range(minlevel, 1):
            S.copy_q_to_qold(lv)
        rc, ncyc, r, h = S.multigrid_iterations()
        res.append((h, [S.get(k, "q") for k in range(len(hier))]))
    (hg, qg), (ho, qo) = res
    for k in range(len(hier)):
        assert np.array_equal(qg[k], qo[k]), (k, float(np.max(np.abs(qg[k] - qo[k]))))
    _hist_close(hg, ho)


@pytest.mark.parametrize("shape", [(128, 128, 128), (200, 36, 16)])
def test_large_patch_mixed_vs_fp64(shape, gpu_solver_cls):
    """Marching FP32 correction sweep and FP32 residual output on the finest
    level, full and partial tiles: the mixed path reaches the FP64 path's
    solution to within 1e-7 in max norm after two time steps."""
    nd = 3
    print(f"[gpu] mixed {shape}", flush=True)
    minlevel = -(int(np.log2(min(shape))) - 2)
    hier = hierarchy.uniform(nd, shape, minlevel, [2] * 6)
    root = len(hier) - 1
    phi = configs.init_phi(hier[root], nd, 1, 1003, 0.1, 0.05)
    q0 = np.zeros(hierarchy.field_shape(hier[root], nd, 2, "q"))
    q0[0, 1:-1, 1:-1, 1:-1] = phi
    outs = []
    for prec in (abi.PREC_FP64, abi.PREC_MIXED):
        S = gpu_solver_cls(configs.params(nd, 5e-5, 0.015, 1e-9, precision=prec), hier)
        S.put(root, "q", q0)
        S.set_ghost(0, 0)
        S.copy_q_to_qold(0)
        for s in range(2):
            rc, ncyc, r = S.time_step(s == 0)
            assert r < 1e-9, (prec, s, rc, r)
        outs.append(S.get(root, "q"))
        S.close()
    err = float(np.max(np.abs(outs[0] - outs[1])))
    print(f"[gpu] mixed {shape} max-norm diff {err:.3e}", flush=True)
    assert err < 1e-7


@pytest.mark.parametrize("name", ["stock_bg2d_u32_phys", "stock_bg2d_u32_per"])
def test_bg2d_vs_oracle(name, gpu_solver_cls):
    """BioGrowth 2D (decoupled Gauss-Seidel, BioGrowth/problem.f90:255-353)
    on the GPU: two time steps from the pinned restart state, bit-identical
    to the oracle, which matches the reference's own module."""
    import json
    import os
    from test_oracle_pins import bg_setup, bg_steps
    print(f"[gpu] BioGrowth {name}", flush=True)
    p = json.load(open(os.path.join(fu.GOLDEN, "stock_pins.json")))[name]
    G, root = bg_setup(p, gpu_solver_cls)
    O, _ = bg_setup(p, oracle.Oracle)
    hg = bg_steps(G, p)
    ho = bg_steps(O, p)
    assert len(hg) == len(ho)
    for a, b in zip(hg, ho):
        _hist_close(a, b)
    assert np.array_equal(G.get(root, "q"), O.get(root, "q"))


@pytest.mark.parametrize("via", ["put", "frame"])
@pytest.mark.parametrize("name", ["stock_ch2d_u32", "stock_ch2d_u16_phys", "stock_ch3d_u8_phys",
                                  "stock_ch2d_u32_err2"])
def test_reference_ch_module_histories(name, via, gpu_solver_cls, tmp_path):
    """The GPU path against the residual histories of the reference's own
    PROBLEM/CH module (ErrorAFAS errortype 1 and 2), from the same
    12-decimal restart state, and bit-identical to the oracle's state.
    via "frame": the GPU run starts from the restart frame the reference
    itself read (written by tests/golden/refrun.py), through
    bsam_read_frame (ReadQ)."""
    import json
    import os
    import sys
    print(f"[gpu] reference CH module {name} via {via}", flush=True)
    p = json.load(open(os.path.join(fu.GOLDEN, "stock_pins.json")))[name]
    nd, n = p["ndims"], p["n"]
    hier = hierarchy.uniform(nd, (n,) * nd, p["minlevel"], p["mthbc"])
    root = len(hier) - 1
    phi = configs.init_phi(hier[root], nd, 1, p["seed"], 0.0, 0.05)
    phi = np.array([float("%25.12f" % v) for v in phi.ravel()]).reshape(phi.shape)
    q0 = np.zeros(hierarchy.field_shape(hier[root], nd, 2, "q"))
    q0[(0,) + (slice(1, -1),) * nd] = phi
    out = []
    for S in (gpu_solver_cls(configs.params(nd, p["dt"], p["eps"], 1e-9, errortype=p.get("errortype", 1)), hier),
              oracle.Oracle(configs.params(nd, p["dt"], p["eps"], 1e-9, errortype=p.get("errortype", 1)), hier)):
        if via == "frame" and not isinstance(S, oracle.Oracle):
            sys.path.insert(0, fu.GOLDEN)
            import refrun
            os.makedirs(tmp_path / "OUT", exist_ok=True)
            path = refrun.write_restart_frame(str(tmp_path), nd, (n,) * nd, 1.0 / n, q0)
            t, fin = S.read_frame(path)
            assert t == 0.0 and fin == 0 and np.array_equal(S.get(root, "q"), q0)
        else:
            S.put(root, "q", q0)
        S.set_ghost(0, 0)
        S.copy_q_to_qold(0)
        hists = []
        for s in range(p["steps"]):
            S.set_ghost(0, 0)
            if s == 0:
                S.fill_down(abi.SOLUTIONFIELD)
            for lv in range(p["minlevel"], 1):
                S.copy_q_to_qold(lv)
            rc, ncyc, r, h = S.multigrid_iterations()
            hists.append(h)
        out.append((hists, S.get(root, "q")))
    (hg, qg), (ho, qo) = out
    assert len(hg) == len(p["stock"])
    for a, b in zip(hg, p["stock"]):
        _hist_close(a, b)
    assert np.array_equal(qg, qo)
