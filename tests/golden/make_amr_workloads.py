"""Export the AMR mesh hierarchies of the two AMR benchmark workloads from
the reference's own regrid.

Runs the reference (oracle/_ref/xbsam_probe: the unmodified LIBRARY with the
probe PROBLEM module) for one time step of each workload with one V-cycle,
reads the patches present when MultigridIterations starts (tag-1 dumps of
the probe's Source hook), and writes their boxes, in the reference's
visiting order, to bsam1.2_amd/bsam_amd/workloads/<name>.json.  The
benchmark rebuilds the hierarchy from these boxes (bsam_amd.hierarchy.
from_patch_boxes) and seeds its own fields on every patch; only the mesh
comes from the reference.

  2D: 128^2 root, maxlevel 2, eps = 0.01, circle of radius 0.25 with
      0.01 U(-1,1) noise (seed 1002), tagging |grad phi| h > 0.05.
  3D: 64^3 root, maxlevel 3, eps = 0.008, sphere of radius 0.3 perturbed by
      0.02 x a seeded combination of real spherical harmonics l <= 6
      (seed 1004), same tagging.
  3D tumour: the probe module's tumour-growth model (phi, mu, nutrient,
      pressure), 64^3 root, maxlevel 3, eps = 0.01, a tumour of radius 0.15
      perturbed by 0.01 x real spherical harmonics l <= 6 (seed 1005),
      physical boundaries on every face (nutrient 1, pressure 0), same
      tagging.

Build-container only: needs oracle/_ref (make -C oracle ref).
"""
import json
import math
import os
import shutil
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(REPO, "bsam1.2_amd"))

import refrun  # noqa: E402
from make_golden import build_step_hierarchy  # noqa: E402
from bsam_amd import hierarchy  # noqa: E402

OUT = os.path.join(REPO, "bsam1.2_amd", "bsam_amd", "workloads")

WORKLOADS = {
    "ch2d_amr_128_l2": dict(ndims=2, n=128, maxlevel=2, eps=0.01, dt=1e-5, tol=1e-9, steps=20,
                            init=dict(initmode=2, seed=1002, amp=0.01, radius=0.25)),
    "ch3d_amr_64_l3": dict(ndims=3, n=64, maxlevel=3, eps=0.008, dt=1e-5, tol=1e-8, steps=10,
                           init=dict(initmode=3, seed=1004, amp=0.0, radius=0.3, pamp=0.02, lmax=6)),
    "tm3d_amr_64_l3": dict(ndims=3, n=64, maxlevel=3, eps=0.01, dt=1e-3, tol=1e-8, steps=20,
                           init=dict(initmode=3, seed=1005, amp=0.0, radius=0.15, pamp=0.01, lmax=6),
                           mthbc=[10] * 6, tumour=dict(model=3, lamp=1.0, lama=0.1, kappa=10.0)),
}


def export(name, w):
    nd, n = w["ndims"], w["n"]
    mthbc = w.get("mthbc", [2] * (2 * nd))
    minlevel = -(int(round(math.log2(n))) - 1)
    work = tempfile.mkdtemp(prefix="bsam_amrw_")
    try:
        prob = [("alpha", 1.0), ("eps", w["eps"]), ("hroot", 1.0 / n)]
        prob += list(w["init"].items())
        prob += [("tagtol", 0.05), ("dumpsweeps", 0), ("dumpsrc", True), ("dumpend", False)]
        prob += list(w.get("tumour", {}).items())
        refrun.write_inputs(work, ndims=nd, mx=(n,) * nd, maxlevel=w["maxlevel"], minlevel=minlevel,
                            dt=w["dt"], steps=1, qerrortol=w["tol"], mthbc=mthbc, maxvcycles=1,
                            minimumgridpoints=8, problem=prob, getafterstepstats=False,
                            nrvars=4 if "tumour" in w else 2)
        refrun.run("xbsam_probe", work, timeout=3600)
        recs = [r for r in refrun.parse_dump(os.path.join(work, "dump.bin")) if r["tag"] == 1]
    finally:
        shutil.rmtree(work, ignore_errors=True)
    hroot = 1.0 / n
    base = hierarchy.uniform(nd, (n,) * nd, minlevel, mthbc)
    hier = build_step_hierarchy(base, recs, hroot, nd)
    boxes = [[d["level"], [d["mglobal"][a] for a in range(nd)]] for d in hier if d["level"] > 0]
    per = {}
    cells = {}
    for d in hier:
        if d["level"] >= 0:
            per[d["level"]] = per.get(d["level"], 0) + 1
            c = 1
            for a in range(nd):
                c *= d["mx"][a]
            cells[d["level"]] = cells.get(d["level"], 0) + c
    meta = dict(name=name, ndims=nd, n=n, minlevel=minlevel, maxlevel=w["maxlevel"], eps=w["eps"], dt=w["dt"],
                tol=w["tol"], steps=w["steps"], init=w["init"], mthbc=mthbc,
                patches_per_level={str(k): v for k, v in sorted(per.items())},
                cells_per_level={str(k): v for k, v in sorted(cells.items())}, boxes=boxes)
    if "tumour" in w:
        meta["tumour"] = dict(w["tumour"])
    os.makedirs(OUT, exist_ok=True)
    with open(os.path.join(OUT, name + ".json"), "w") as fh:
        json.dump(meta, fh, separators=(",", ":"))
    print(name, "patches per level", per, "cells per level", cells)


def main():
    only = set(sys.argv[1:])
    for name, w in WORKLOADS.items():
        if not only or name in only:
            export(name, w)


if __name__ == "__main__":
    main()
