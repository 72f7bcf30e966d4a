"""The flattened ghost-fill tables against the staged SetGhost, on the host.

Per-level launches replace SetGhost(level, ipass) (boundary.f90:67-97:
interpolation, sibling transfers, periodic copies, physical conditions,
later stages overwriting earlier ones) by one gather over precomputed
rules, one per ghost cell, and inside red-black sweep sequences by a
reduced table of the face ghosts the next sweep can read.  The library's
host self-test evaluates both forms on seeded random data and requires
bit-identical values (bsam_selftest_ghost_maps in include/bsam_gpu.h).
It runs here on the reference's own AMR hierarchies (periodic and with
physical boundaries, 2D and 3D) and on the two AMR bench meshes; no GPU is
used.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

import fixture_util as fu
from bsam_amd import abi, configs, gpu, hierarchy

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WORKLOADS = os.path.join(ROOT, "bsam1.2_amd", "bsam_amd", "workloads")


def _selftest(params, hier, seed=7):
    """(values compared, levels checked, lowest level, finest level)"""
    lib = gpu.load()
    lo, hi = hierarchy.levels(hier)
    descs = hierarchy.to_ctypes(hier)
    n, nl = C.c_int64(0), C.c_int32(0)
    rc = lib.bsam_selftest_ghost_maps(C.byref(params), len(hier), descs, lo, hi, seed, C.byref(n), C.byref(nl))
    assert rc == abi.BSAM_OK, lib.bsam_last_error().decode()
    return n.value, nl.value, lo, hi


@pytest.mark.parametrize("name", ["ch2d_amr", "ch2d_amr_phys", "ch3d_amr", "ch3d_amr_phys", "ch3d_amr2",
                                  "ch3d_amr2_phys", "tm2d_amr_phys", "tm3d_amr_phys"])
def test_golden_hierarchies(name):
    """Every refined level is flattened; with physical boundaries so are the
    root and the levels below it, whose fill is the physical condition
    face by face (a 3D face reads edge ghosts from before the call, which
    the later faces overwrite).  A root periodic in every dimension keeps
    its wrap kernel.  The tumour fixtures exercise four fields and their
    mixed boundary condition.  On ch3d_amr2_phys the finest level keeps the
    staged fill: a level-2 patch on a physical face next to a coarse-fine
    edge reads an edge ghost the call has not written, which no later stage
    overwrites, so its final value depends on the state before the call."""
    meta, _ = fu.load(name)
    for st in meta["steps"]:
        n, nl, lo, hi = _selftest(fu.step_params(meta), st["hierarchy"])
        expect = hi - lo + 1 if name.endswith("_phys") else hi
        if name == "ch3d_amr2_phys":
            expect -= 1
        assert nl == expect and n > 0


@pytest.mark.parametrize("name", ["ch2d_amr_128_l2", "ch3d_amr_64_l3", "tm3d_amr_64_l3"])
def test_bench_meshes(name):
    """The AMR bench meshes: the periodic CH meshes flatten every refined
    level, the tumour mesh (physical conditions on every face) every
    level."""
    w = json.load(open(os.path.join(WORKLOADS, name + ".json")))
    nd = w["ndims"]
    base = hierarchy.uniform(nd, (w["n"],) * nd, w["minlevel"], w["mthbc"])
    hier = hierarchy.from_patch_boxes(base, w["boxes"], nd)
    kw = {}
    if "tumour" in w:
        t = w["tumour"]
        kw = dict(model=abi.MODEL_TUMOUR, extra=(t["lamp"], t["lama"], t["kappa"]))
    p = configs.params(nd, w["dt"], w["eps"], w["tol"], **kw)
    n, nl, lo, hi = _selftest(p, hier)
    assert hi == w["maxlevel"] and lo == w["minlevel"] and n > 0
    assert nl == (hi - lo + 1 if "tumour" in w else hi)


def test_overlapping_patches_keep_the_staged_fill():
    """Two level-1 patches that overlap (not a valid BSAM mesh): a sibling
    transfer then writes cells another transfer of the same call reads, so
    the fill cannot be flattened and the level keeps the staged order."""
    base = hierarchy.uniform(2, (16, 16), -3, [2, 2, 2, 2])
    hier = hierarchy.from_patch_boxes(base, [[1, [[1, 8], [1, 8]]], [1, [[7, 14], [1, 8]]]], 2)
    n, nl, lo, hi = _selftest(configs.params(2, 1e-4, 0.03, 1e-9), hier)
    assert hi == 1 and nl == 0
