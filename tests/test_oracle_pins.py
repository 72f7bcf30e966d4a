"""The CPU oracle against BSAM 1.2's own LIBRARY.

Golden vectors were written by the unmodified reference (tests/golden/
make_golden.py); the oracle must reproduce them bit for bit: residual
histories (afasroutines.f90:105), the state after each time step, and the
state after every recorded relaxation sweep.  stock_pins.json pins the
probe PROBLEM module used for the golden runs to the reference's own
PROBLEM/CH hooks.
"""
import json
import os

import numpy as np
import pytest

import fixture_util as fu
import oracle
from bsam_amd import configs, hierarchy


@pytest.mark.parametrize("name", fu.NAMES)
def test_multigrid_iterations_bitwise(name):
    meta, z = fu.load(name)
    for s, st in enumerate(meta["steps"]):
        S, hist = fu.run_step(oracle.Oracle, meta, z, s)
        assert hist == st["residuals"], (name, s, hist, st["residuals"])
        for i in st["end"]:
            got = S.get(i, "q").ravel()
            assert np.array_equal(got, fu.end_state(z, s, i)), (name, s, i)


@pytest.mark.parametrize("name", fu.NAMES)
def test_sweeps_bitwise(name):
    meta, z = fu.load(name)

    def cmp(S, pid, golden, rec):
        assert np.array_equal(S.get(pid, "q").ravel(), golden), (name, rec)

    n = fu.replay_sweeps(oracle.Oracle, meta, z, 0, cmp)
    assert n == len(meta["steps"][0]["sweeps"]) > 0


def test_probe_hooks_match_reference_ch_module():
    pins = json.load(open(os.path.join(fu.GOLDEN, "stock_pins.json")))
    assert pins
    for name, p in pins.items():
        assert p["stock"] == p["probe"], name


@pytest.mark.parametrize("name", ["stock_ch2d_u32", "stock_ch2d_u16_phys", "stock_ch3d_u8_phys"])
def test_oracle_matches_reference_ch_module(name):
    """The oracle, started from the same 12-decimal restart frame, reproduces
    the residual histories of the reference's own PROBLEM/CH module."""
    p = json.load(open(os.path.join(fu.GOLDEN, "stock_pins.json")))[name]
    nd, n = p["ndims"], p["n"]
    hier = hierarchy.uniform(nd, (n,) * nd, p["minlevel"], p["mthbc"])
    root = len(hier) - 1
    O = oracle.Oracle(configs.params(nd, p["dt"], p["eps"], 1e-9), hier)
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
    O2 = oracle.Oracle(configs.params(nd, p["dt"], p["eps"], 1e-9), hier)
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
