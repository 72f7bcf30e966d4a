"""The CPU oracle against BSAM 1.2's own LIBRARY.

Golden vectors were written by the unmodified reference (tests/golden/
make_golden.py); the oracle must reproduce them bit for bit: residual
histories (afasroutines.f90:105), the state after each time step, and the
state after every recorded relaxation sweep.  stock_pins.json pins the
probe PROBLEM module used for the golden runs to the reference's own
PROBLEM/CH hooks.  The tumour fixtures (tm*) are the reference LIBRARY
running the probe module's 4-field tumour model (model = 3), which has no
counterpart among the reference's PROBLEM modules.
"""
import json
import os

import numpy as np
import pytest

import fixture_util as fu
import oracle
from bsam_amd import configs, hierarchy


@pytest.mark.parametrize("name", fu.NAMES + fu.TUMOUR_NAMES)
def test_multigrid_iterations_bitwise(name):
    meta, z = fu.load(name)
    for s, st in enumerate(meta["steps"]):
        S, hist = fu.run_step(oracle.Oracle, meta, z, s)
        assert hist == st["residuals"], (name, s, hist, st["residuals"])
        for i in st["end"]:
            got = S.get(i, "q").ravel()
            assert np.array_equal(got, fu.end_state(z, s, i)), (name, s, i)


@pytest.mark.parametrize("name", fu.NAMES + fu.TUMOUR_NAMES)
def test_sweeps_bitwise(name):
    meta, z = fu.load(name)

    def cmp(S, pid, golden, rec):
        assert np.array_equal(S.get(pid, "q").ravel(), golden), (name, rec)

    n = fu.replay_sweeps(oracle.Oracle, meta, z, 0, cmp)
    assert n == len(meta["steps"][0]["sweeps"]) > 0


def test_probe_hooks_match_reference_ch_module():
    pins = json.load(open(os.path.join(fu.GOLDEN, "stock_pins.json")))
    assert pins
    ch = {k: v for k, v in pins.items() if "probe" in v}
    assert len(ch) == 4
    for name, p in ch.items():
        assert p["stock"] == p["probe"], name


@pytest.mark.parametrize("name", ["stock_ch2d_u32", "stock_ch2d_u16_phys", "stock_ch3d_u8_phys",
                                  "stock_ch2d_u32_err2"])
def test_oracle_matches_reference_ch_module(name):
    """The oracle, started from the same 12-decimal restart frame, reproduces
    the residual histories of the reference's own PROBLEM/CH module."""
    p = json.load(open(os.path.join(fu.GOLDEN, "stock_pins.json")))[name]
    nd, n = p["ndims"], p["n"]
    hier = hierarchy.uniform(nd, (n,) * nd, p["minlevel"], p["mthbc"])
    root = len(hier) - 1
    O = oracle.Oracle(configs.params(nd, p["dt"], p["eps"], 1e-9, errortype=p.get("errortype", 1)), hier)
    phi = configs.init_phi(hier[root], nd, 1, p["seed"], 0.0, 0.05)
    phi = np.array([float("%25.12f" % v) for v in phi.ravel()]).reshape(phi.shape)
    q = O.field(root, "q")
    q[...] = 0.0
    q[(0,) + (slice(1, -1),) * nd] = phi
    O.set_ghost(0, 0)
    O.copy_q_to_qold(0)
    hists = []
    for s in range(p["steps"]):
        rc, ncyc, r = O.time_step(s == 0)
        hists.append(None)
    # rerun collecting histories through the explicit path
    O2 = oracle.Oracle(configs.params(nd, p["dt"], p["eps"], 1e-9, errortype=p.get("errortype", 1)), hier)
    q2 = O2.field(root, "q")
    q2[...] = 0.0
    q2[(0,) + (slice(1, -1),) * nd] = phi
    O2.set_ghost(0, 0)
    O2.copy_q_to_qold(0)
    hists = []
    for s in range(p["steps"]):
        O2.set_ghost(0, 0)
        if s == 0:
            O2.fill_down(2)
        for lv in range(p["minlevel"], 1):
            O2.copy_q_to_qold(lv)
        rc, ncyc, r, h = O2.multigrid_iterations()
        hists.append(h)
    assert hists == p["stock"]
    assert np.array_equal(O.get(root, "q"), O2.get(root, "q"))


def bg_setup(p, cls):
    """A solver of class `cls` holding the BioGrowth pin's initial state."""
    from bsam_amd import abi
    nd, n = p["ndims"], p["n"]
    hier = hierarchy.uniform(nd, (n,) * nd, p["minlevel"], p["mthbc"], xupper=(p["h"] * n,) * 3)
    root = len(hier) - 1
    prm = configs.params(nd, p["dt"], p["eps"], 1e-9, model=abi.MODEL_BG,
                         extra=(p["gamma"], p["lam"], p["czero"]))
    S = cls(prm, hier)
    q = np.zeros((2, n + 2, n + 2))
    for v, (sd, base, amp) in enumerate(((p["seed"], -0.2, 0.5), (p["seed"] + 7, 1.0, 0.1))):
        x = configs.init_phi(hier[root], nd, 1, sd, base, amp)
        q[v, 1:-1, 1:-1] = np.array([float("%25.12f" % t) for t in x.ravel()]).reshape(x.shape)
    S.put(root, "q", q)
    S.set_ghost(0, 0)
    S.copy_q_to_qold(0)
    return S, root


def bg_steps(S, p):
    hists = []
    for s in range(p["steps"]):
        S.set_ghost(0, 0)
        if s == 0:
            S.fill_down(2)
        for lv in range(p["minlevel"], 1):
            S.copy_q_to_qold(lv)
        rc, ncyc, r, h = S.multigrid_iterations()
        hists.append(h)
    return hists


@pytest.mark.parametrize("name", ["stock_bg2d_u32_phys", "stock_bg2d_u32_per"])
def test_oracle_matches_reference_bg_module(name):
    """The oracle's BioGrowth model (BioGrowth/problem.f90:255-405, 623-736),
    started from the same 12-decimal restart frame, reproduces the residual
    histories of the reference's own PROBLEM/BioGrowth module bit for bit,
    and its final q agrees with the reference's output frame to the frame's
    12 decimals."""
    p = json.load(open(os.path.join(fu.GOLDEN, "stock_pins.json")))[name]
    S, root = bg_setup(p, oracle.Oracle)
    hists = bg_steps(S, p)
    assert hists == p["stock"], (hists, p["stock"])
    end = np.array(p["end_q"])
    got = S.get(root, "q").ravel()
    assert float(np.max(np.abs(got - end))) <= 5.0e-13 + 1e-15
