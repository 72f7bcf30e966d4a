"""Generate the golden vectors of tests/golden/*.npz from the reference.

Runs BSAM 1.2's own LIBRARY (built from /root/reference by oracle/Makefile
into oracle/_ref) and records, per fixture:

  * the mesh hierarchy of every time step (from the patches the probe
    PROBLEM module sees at the source build),
  * q and qold of every patch on levels >= 0 at MultigridIterations entry
    (tag 1), q after the time step (tag 3),
  * the residual history printed by MultigridIterations,
  * q after each of the first relaxation sweeps (tag 2), with the patch and
    colour, for sweep-by-sweep parity.

It also runs the reference's unmodified PROBLEM/CH module (xbsam_stock) from
a seeded restart frame next to the probe module, so that the probe hooks are
pinned to the reference's own hooks.

Usage (build container only):  make -C oracle ref && python tests/golden/make_golden.py
"""
import json
import math
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(REPO, "bsam1.2_amd"))
import refrun  # noqa: E402
from bsam_amd import configs, hierarchy  # noqa: E402

FIXTURES = {
    # uniform periodic meshes (workloads 1 and 3 at reduced size)
    "ch2d_u32": dict(ndims=2, n=32, maxlevel=0, mthbc=(2, 2, 2, 2), eps=0.03, dt=1e-4, steps=2, tol=1e-9,
                     init=dict(initmode=1, seed=1001, base=0.0, amp=0.05), sweeps=40),
    "ch3d_u16": dict(ndims=3, n=16, maxlevel=0, mthbc=(2,) * 6, eps=0.05, dt=5e-5, steps=2, tol=1e-9,
                     init=dict(initmode=1, seed=1003, base=0.1, amp=0.05), sweeps=28),
    # physical boundaries (mthbc=10): 2D reflection, 3D 2-q
    "ch2d_u16_phys": dict(ndims=2, n=16, maxlevel=0, mthbc=(2, 2, 10, 10), eps=0.05, dt=1e-4, steps=2,
                          tol=1e-9, init=dict(initmode=1, seed=1011, base=0.0, amp=0.05), sweeps=20),
    "ch3d_u8_phys": dict(ndims=3, n=8, maxlevel=0, mthbc=(10, 10, 2, 2, 10, 10), eps=0.08, dt=1e-4,
                         steps=2, tol=1e-9, init=dict(initmode=1, seed=1012, base=0.2, amp=0.05), sweeps=12),
    # AMR (workloads 2 and 4 at reduced size)
    "ch2d_amr": dict(ndims=2, n=32, maxlevel=2, mthbc=(2, 2, 2, 2), eps=0.02, dt=1e-5, steps=2, tol=1e-9,
                     init=dict(initmode=2, seed=1002, amp=0.01, radius=0.25), sweeps=200, mgp=4),
    "ch2d_amr_phys": dict(ndims=2, n=32, maxlevel=2, mthbc=(10, 10, 2, 2), eps=0.02, dt=1e-5, steps=1,
                          tol=1e-9, init=dict(initmode=2, seed=1006, amp=0.01, radius=0.3,
                                              xc=(0.35, 0.5, 0.5)), sweeps=120, mgp=4),
    "ch3d_amr": dict(ndims=3, n=16, maxlevel=1, mthbc=(2,) * 6, eps=0.04, dt=1e-5, steps=1, tol=1e-9,
                     init=dict(initmode=2, seed=1004, amp=0.01, radius=0.25), sweeps=24, mgp=4),
    "ch3d_amr_phys": dict(ndims=3, n=16, maxlevel=1, mthbc=(2, 2, 2, 2, 10, 10), eps=0.04, dt=1e-5,
                          steps=1, tol=1e-9, init=dict(initmode=2, seed=1007, amp=0.01, radius=0.3,
                                                       xc=(0.5, 0.5, 0.3)), sweeps=24, mgp=4),
    # two refined levels in 3D: level-2 patches whose parents are refined
    # patches themselves, as on levels 2 and 3 of the 3D AMR workload
    # (a higher tagging threshold keeps the refined band, and the fixture,
    # small)
    "ch3d_amr2": dict(ndims=3, n=16, maxlevel=2, mthbc=(2,) * 6, eps=0.04, dt=1e-5, steps=1, tol=1e-9,
                      init=dict(initmode=2, seed=1008, amp=0.01, radius=0.22), sweeps=32, mgp=4, tagtol=0.3),
    "ch3d_amr2_phys": dict(ndims=3, n=8, maxlevel=2, mthbc=(2, 2, 2, 2, 10, 10), eps=0.04, dt=1e-5,
                           steps=2, tol=1e-9, init=dict(initmode=2, seed=1009, amp=0.01, radius=0.2,
                                                        xc=(0.4, 0.5, 0.45)), sweeps=32, mgp=4, tagtol=0.3),
}

# The tumour-growth model (model 3 of the probe module, nrvars = 4: phi, mu,
# nutrient, pressure) on physical boundaries, where the nutrient is held at
# 1 and the pressure at 0; proliferation 1.0, apoptosis 0.1, consumption 10.
TUMOUR = dict(model=3, lamp=1.0, lama=0.1, kappa=10.0)
FIXTURES.update({
    "tm2d_u16_phys": dict(ndims=2, n=16, maxlevel=0, mthbc=(10,) * 4, eps=0.06, dt=1e-3, steps=2, tol=1e-9,
                          init=dict(initmode=2, seed=1051, amp=0.01, radius=0.3), sweeps=20, tumour=TUMOUR),
    "tm3d_u8_phys": dict(ndims=3, n=8, maxlevel=0, mthbc=(10,) * 6, eps=0.1, dt=1e-3, steps=2, tol=1e-9,
                         init=dict(initmode=2, seed=1052, amp=0.01, radius=0.3), sweeps=12, tumour=TUMOUR),
    "tm2d_amr_phys": dict(ndims=2, n=32, maxlevel=2, mthbc=(10,) * 4, eps=0.02, dt=1e-4, steps=2, tol=1e-9,
                          init=dict(initmode=2, seed=1053, amp=0.01, radius=0.2), sweeps=120, mgp=4,
                          tumour=TUMOUR),
    "tm3d_amr_phys": dict(ndims=3, n=16, maxlevel=1, mthbc=(10,) * 6, eps=0.04, dt=1e-4, steps=1, tol=1e-9,
                          init=dict(initmode=2, seed=1054, amp=0.01, radius=0.2), sweeps=24, mgp=4,
                          tumour=TUMOUR),
})

# Probe module vs the reference's own PROBLEM/CH module from one restart frame.
STOCK_PINS = {
    "stock_ch2d_u32": dict(ndims=2, n=32, mthbc=(2, 2, 2, 2), eps=0.03, dt=1e-4, steps=2, seed=1021),
    "stock_ch2d_u16_phys": dict(ndims=2, n=16, mthbc=(2, 2, 10, 10), eps=0.05, dt=1e-4, steps=2, seed=1022),
    "stock_ch3d_u8_phys": dict(ndims=3, n=8, mthbc=(10, 10, 10, 10, 10, 10), eps=0.08, dt=1e-4, steps=2,
                               seed=1023),
    # ErrorAFAS errortype = 2 (max of the component RMS, afasroutines.f90:139-156)
    "stock_ch2d_u32_err2": dict(ndims=2, n=32, mthbc=(2, 2, 2, 2), eps=0.03, dt=1e-4, steps=2, seed=1024,
                                errortype=2),
}


# The reference's own PROBLEM/BioGrowth module (2D), run from a seeded
# restart frame of phi and the nutrient c; model constants as its shipped
# problemdata.dat (PROBLEM/BioGrowth/problemdata.dat), dx = 0.1 as its
# griddata.dat (64 cells on [0, 6.4]).
BG_PINS = {
    "stock_bg2d_u32_phys": dict(ndims=2, n=32, mthbc=(10, 10, 10, 10), dt=0.1, steps=2, seed=1031,
                                gamma=1.5, lam=2.0, czero=0.5, eps=0.05),
    "stock_bg2d_u32_per": dict(ndims=2, n=32, mthbc=(2, 2, 2, 2), dt=0.1, steps=2, seed=1032,
                               gamma=1.5, lam=2.0, czero=0.5, eps=0.05),
}


def bg_initial(n, seed):
    """phi = -0.2 + 0.5 U, c = 1 + 0.1 U' (U' keyed by seed + 7), both
    rounded to the 12 decimals of the frame format, zero ghosts."""
    hier = hierarchy.uniform(2, (n, n), -1, [2] * 4)
    root = hier[-1]
    q = np.zeros((2, n + 2, n + 2))
    for v, (sd, base, amp) in enumerate(((seed, -0.2, 0.5), (seed + 7, 1.0, 0.1))):
        x = configs.init_phi(root, 2, 1, sd, base, amp)
        q[v, 1:-1, 1:-1] = np.array([float("%25.12f" % t) for t in x.ravel()]).reshape(x.shape)
    return q


def make_bg_pin(name, fx, results):
    nd, n = fx["ndims"], fx["n"]
    minlevel = -(int(round(math.log2(n))) - 1)
    h = 0.1
    q = bg_initial(n, fx["seed"])
    work = tempfile.mkdtemp(prefix="bsam_bgpin_")
    try:
        prob = [("czero", fx["czero"]), ("gamma", fx["gamma"]), ("lambda", fx["lam"]), ("eps", fx["eps"])]
        refrun.write_inputs(work, ndims=nd, mx=(n,) * nd, maxlevel=0, minlevel=minlevel, dt=fx["dt"],
                            steps=fx["steps"], qerrortol=1e-9, mthbc=fx["mthbc"], restart=True,
                            problem=prob, getafterstepstats=False, xupper=h * n, minimumgridpoints=4)
        refrun.write_restart_frame(work, nd, (n,) * nd, h, q)
        out = refrun.run("xbsam_bg", work)
        hist = refrun.parse_residuals(out)
        end = refrun.parse_frame(os.path.join(work, "OUT", "m00001.dat"), nd)
    finally:
        shutil.rmtree(work, ignore_errors=True)
    assert len(end) == 1 and end[0]["level"] == 0
    print(f"{name}: BioGrowth histories {[len(x) for x in hist]}")
    results[name] = dict(fx, minlevel=minlevel, h=h, stock=hist, end_q=end[0]["q"].ravel().tolist())


def problem_items(fx, n, dumpsweeps):
    it = [("alpha", 1.0), ("eps", fx["eps"]), ("hroot", 1.0 / n)]
    init = dict(fx["init"])
    xc = init.pop("xc", None)
    for k, v in init.items():
        it.append((k, v))
    if xc is not None:
        for a in range(3):
            it.append((f"xc({a+1})", float(xc[a])))
    it += [("tagtol", fx.get("tagtol", 0.05)), ("dumpsweeps", dumpsweeps), ("dumpsrc", True), ("dumpend", True)]
    for k, v in fx.get("tumour", {}).items():
        it.append((k, v))
    return it


def patch_key(rec, hroot, nd):
    lev = int(round(math.log2(hroot / rec["h"])))
    lo = tuple(int(round(rec["xlower"][a] / rec["h"])) + 1 for a in range(nd))
    return lev, lo


def build_step_hierarchy(base, srcs, hroot, nd):
    boxes = []
    for r in srcs:
        lev, lo = patch_key(r, hroot, nd)
        if lev == 0:
            continue
        boxes.append((lev, [[lo[a], lo[a] + r["mx"][a] - 1] for a in range(nd)]))
    # levels in increasing order so parents exist first; keep visiting order
    boxes.sort(key=lambda b: b[0])
    return hierarchy.from_patch_boxes(base, boxes, nd)


def find_index(hier, lev, lo, nd):
    for i, d in enumerate(hier):
        if d["level"] == lev and all(d["mglobal"][a][0] == lo[a] for a in range(nd)):
            return i
    raise KeyError((lev, lo))


def make_fixture(name, fx, outdir):
    nd, n = fx["ndims"], fx["n"]
    minlevel = -(int(round(math.log2(n))) - 1)
    work = tempfile.mkdtemp(prefix="bsam_golden_")
    try:
        refrun.write_inputs(work, ndims=nd, mx=(n,) * nd, maxlevel=fx["maxlevel"], minlevel=minlevel,
                            dt=fx["dt"], steps=fx["steps"], qerrortol=fx["tol"], mthbc=fx["mthbc"],
                            minimumgridpoints=fx.get("mgp", 8),
                            problem=problem_items(fx, n, fx["sweeps"]), nrvars=4 if "tumour" in fx else 2)
        out = refrun.run("xbsam_probe", work)
        hist = refrun.parse_residuals(out)
        recs = refrun.parse_dump(os.path.join(work, "dump.bin"))
    finally:
        shutil.rmtree(work, ignore_errors=True)
    hroot = 1.0 / n
    base = hierarchy.uniform(nd, (n,) * nd, minlevel, fx["mthbc"])
    # split records into steps: a step starts with its first tag-1 record
    steps, cur = [], None
    pre = []
    for r in recs:
        if r["tag"] == 1:
            if cur is None or cur["phase"] != "src":
                cur = dict(src=[], end=[], sweeps=[], phase="src")
                steps.append(cur)
            cur["src"].append(r)
        elif r["tag"] == 2:
            if cur is None:
                pre.append(r)
            else:
                cur["phase"] = "run"
                cur["sweeps"].append(r)
        elif r["tag"] == 3:
            if cur is not None:
                cur["phase"] = "end"
                cur["end"].append(r)
    if len(steps) != len(hist):
        raise RuntimeError(f"{name}: {len(steps)} steps dumped, {len(hist)} residual histories")
    arrays = {}
    meta = dict(name=name, ndims=nd, n=n, minlevel=minlevel, eps=fx["eps"], dt=fx["dt"], tol=fx["tol"],
                mthbc=list(fx["mthbc"]), steps=[])
    if "tumour" in fx:
        meta["tumour"] = dict(fx["tumour"])
    for s, st in enumerate(steps):
        hier = build_step_hierarchy(base, st["src"], hroot, nd)
        finest = max(d["level"] for d in hier)
        sm = dict(hierarchy=hier, finestlevel=finest, residuals=hist[s], src=[], end=[], sweeps=[])
        for r in st["src"]:
            lev, lo = patch_key(r, hroot, nd)
            i = find_index(hier, lev, lo, nd)
            nq = r["data"].size // 2
            arrays[f"s{s}_q_p{i}"] = r["data"][:nq]
            arrays[f"s{s}_qold_p{i}"] = r["data"][nq:]
            sm["src"].append(i)
        for r in st["end"]:
            lev, lo = patch_key(r, hroot, nd)
            i = find_index(hier, lev, lo, nd)
            arrays[f"s{s}_end_p{i}"] = r["data"]
            sm["end"].append(i)
        for w, r in enumerate(st["sweeps"]):
            lev, lo = patch_key(r, hroot, nd)
            if lev < 0:
                # below-root patches: the single patch of that level
                i = [k for k, d in enumerate(hier) if d["level"] == lev][0]
            else:
                i = find_index(hier, lev, lo, nd)
            arrays[f"s{s}_sweep{w}"] = r["data"]
            sm["sweeps"].append(dict(patch=i, level=lev, ipass=r["ipass"]))
        meta["steps"].append(sm)
    path = os.path.join(outdir, name + ".npz")
    np.savez_compressed(path, meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8), **arrays)
    nb = os.path.getsize(path)
    print(f"{name}: {len(steps)} steps, patches/step {[len(st['src']) for st in steps]}, "
          f"V-cycles {[len(h) for h in hist]}, sweeps {[len(st['sweeps']) for st in steps]}, {nb/1e6:.2f} MB")


def seeded_root(nd, n, seed):
    """Seeded phi (base 0, amplitude 0.05) rounded to the 12 decimals of the
    frame format, as q(var, [k,] j, i) with zero ghosts."""
    hier = hierarchy.uniform(nd, (n,) * nd, -1, [2] * (2 * nd))
    root = hier[-1]
    phi = configs.init_phi(root, nd, 1, seed, 0.0, 0.05)
    phi = np.array([float("%25.12f" % v) for v in phi.ravel()]).reshape(phi.shape)
    q = np.zeros((2,) + tuple([n + 2] * nd))
    if nd == 2:
        q[0, 1:-1, 1:-1] = phi
    else:
        q[0, 1:-1, 1:-1, 1:-1] = phi
    return q


def make_stock_pin(name, fx, results):
    nd, n = fx["ndims"], fx["n"]
    minlevel = -(int(round(math.log2(n))) - 1)
    q = seeded_root(nd, n, fx["seed"])
    hists = {}
    for binary in ("xbsam_stock", "xbsam_probe"):
        work = tempfile.mkdtemp(prefix="bsam_pin_")
        try:
            if binary == "xbsam_stock":
                prob = [("alpha", 1.0), ("eps", fx["eps"])]
            else:
                prob = [("alpha", 1.0), ("eps", fx["eps"]), ("hroot", 1.0 / n)]
            refrun.write_inputs(work, ndims=nd, mx=(n,) * nd, maxlevel=0, minlevel=minlevel, dt=fx["dt"],
                                steps=fx["steps"], qerrortol=1e-9, mthbc=fx["mthbc"], restart=True,
                                problem=prob, getafterstepstats=False, errortype=fx.get("errortype", 1))
            refrun.write_restart_frame(work, nd, (n,) * nd, 1.0 / n, q)
            out = refrun.run(binary, work)
            hists[binary] = refrun.parse_residuals(out)
        finally:
            shutil.rmtree(work, ignore_errors=True)
    same = hists["xbsam_stock"] == hists["xbsam_probe"]
    print(f"{name}: stock == probe residual histories: {same}  {[len(h) for h in hists['xbsam_stock']]}")
    results[name] = dict(fx, minlevel=minlevel, stock=hists["xbsam_stock"], probe=hists["xbsam_probe"])


def main():
    only = set(sys.argv[1:])
    for name, fx in FIXTURES.items():
        if not only or name in only:
            make_fixture(name, fx, HERE)
    pins = {}
    for name, fx in STOCK_PINS.items():
        make_stock_pin(name, fx, pins)
    for name, fx in BG_PINS.items():
        make_bg_pin(name, fx, pins)
    with open(os.path.join(HERE, "stock_pins.json"), "w") as fh:
        json.dump(pins, fh, indent=1)


if __name__ == "__main__":
    main()
