"""Replay the reference's golden vectors (tests/golden/*.npz) on a solver.

A solver is either the CPU oracle (oracle.Oracle) or the GPU library
(bsam_amd.gpu.Solver); both expose the reference's per-level operations, so
one replay serves both.  The golden vectors come from BSAM 1.2's own LIBRARY
(tests/golden/make_golden.py).
"""
import json
import os

import numpy as np

from bsam_amd import abi, configs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ["ch2d_u32", "ch3d_u16", "ch2d_u16_phys", "ch3d_u8_phys", "ch2d_amr", "ch2d_amr_phys",
         "ch3d_amr", "ch3d_amr_phys", "ch3d_amr2", "ch3d_amr2_phys"]
# the tumour-growth model (4 fields) run by the reference LIBRARY through the
# probe PROBLEM module (model = 3)
TUMOUR_NAMES = ["tm2d_u16_phys", "tm3d_u8_phys", "tm2d_amr_phys", "tm3d_amr_phys"]


def load(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    meta = json.loads(bytes(z["meta"]).decode())
    return meta, z


def step_params(meta, **kw):
    tm = meta.get("tumour")
    if tm:
        kw = dict(kw, model=abi.MODEL_TUMOUR, extra=(tm["lamp"], tm["lama"], tm["kappa"]))
    return configs.params(meta["ndims"], meta["dt"], meta["eps"], meta["tol"], **kw)


def load_state(S, meta, z, s):
    st = meta["steps"][s]
    for i in st["src"]:
        shp = S.get(i, "q").shape
        S.put(i, "q", z[f"s{s}_q_p{i}"].reshape(shp))
        S.put(i, "qold", z[f"s{s}_qold_p{i}"].reshape(shp))
    if meta.get("tumour"):
        below_root_qold(S, meta, z, s)


def below_root_qold(S, meta, z, s):
    """qold of the levels below the root as the reference holds it at
    MultigridIterations entry: CopyQToQold on every level copies q there,
    and q below the root is the FillDown of the root's q at the end of the
    previous step (or of the initial data, bsamroutines.f90:693-702), which
    is this step's root qold.  Only hooks that read qold on coarse levels
    (the tumour model's lagged tumour fraction) see it."""
    st = meta["steps"][s]
    shp = {i: S.get(i, "q").shape for i in st["src"]}
    for i in st["src"]:
        S.put(i, "q", z[f"s{s}_qold_p{i}"].reshape(shp[i]))
    S.fill_down(abi.SOLUTIONFIELD)
    for lv in range(meta["minlevel"], 0):
        S.copy_q_to_qold(lv)
    for i in st["src"]:
        S.put(i, "q", z[f"s{s}_q_p{i}"].reshape(shp[i]))


def run_step(make_solver, meta, z, s, **kw):
    """MultigridIterations on the state of step s; returns (solver, history)."""
    st = meta["steps"][s]
    S = make_solver(step_params(meta, **kw), st["hierarchy"])
    load_state(S, meta, z, s)
    rc, n, r, hist = S.multigrid_iterations()
    return S, hist


def end_state(z, s, i):
    return z[f"s{s}_end_p{i}"]


def replay_sweeps(make_solver, meta, z, s, compare):
    """Drive AFASVcycle (afasroutines.f90:495-557) by hand and call
    compare(solver, patch, golden_q, record) after every relaxation that has
    golden records.  Returns the number of records compared."""
    st = meta["steps"][s]
    recs = st["sweeps"]
    if not recs:
        return 0
    S = make_solver(step_params(meta), st["hierarchy"])
    load_state(S, meta, z, s)
    hier = st["hierarchy"]
    minlevel = min(d["level"] for d in hier)
    finest = st["finestlevel"]
    S.fill_down(abi.SOLUTIONFIELD)
    for lv in range(finest, -1, -1):
        S.get_source(lv)
    for lv in range(finest, -1, -1):
        S.update_source(lv)
    S.fill_down(abi.SOURCEFIELD)
    by_level = {}
    for i, d in enumerate(hier):
        by_level.setdefault(d["level"], []).append(i)
    pos = [0]

    class Done(Exception):
        pass

    def relax(level, ipass):
        S.relax_level(level, ipass)
        for pid in by_level[level]:
            if pos[0] >= len(recs):
                raise Done
            r = recs[pos[0]]
            assert r["patch"] == pid and r["ipass"] == ipass, (r, pid, ipass)
            compare(S, pid, z[f"s{s}_sweep{pos[0]}"], r)
            pos[0] += 1

    def level_relax(level):
        for _ in range(2):
            for rb in (1, 2):
                relax(level, rb)
                S.set_ghost(level, rb)

    def vcycle(level):
        level_relax(level)
        if level > minlevel:
            S.restrict_solution(level)
            S.set_ghost(level - 1, 0)
            S.coarse_ghost_points(level)
            S.coarse_loading(level)
            vcycle(level - 1)
            S.correct_fine(level)
            S.set_ghost(level, 0)
            level_relax(level)

    try:
        while True:
            vcycle(finest)
            S.fill_down(abi.SOLUTIONFIELD)
            for lv in range(finest, -1, -1):
                S.update_source(lv)
            S.fill_down(abi.SOURCEFIELD)
    except Done:
        pass
    return pos[0]
