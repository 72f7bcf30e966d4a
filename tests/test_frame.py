"""The output frame (WriteQ / OutputQ / WriteQ2D/3D, LIBRARY/
bsaminputoutput.f90:60-181) against the reference's own frames.

tests/golden/make_frames.py ran the reference for one time step and kept
the frame it wrote together with q of every patch as the probe's AfterStep
hook dumped it (nothing changes q in between), and the Fortran runtime's
F25.12 edit of a list of edge values (halfway cases at the twelfth
decimal, signed zeros, tiny negatives, overflowing and non-finite values).

  * the oracle's F25.12 edit equals the runtime's on every edge value, and
    the library's own edit (run on the host) equals both;
  * the oracle's frame (orc_write_frame) equals the reference's frame byte
    for byte, 2D and 3D, CH and the tumour model (four fields);
  * the HIP writer (bsam_write_frame) equals the reference's frame byte
    for byte, and the oracle's on fields made of the edge values.
"""
import json
import os

import numpy as np
import pytest

import fixture_util as fu
import oracle
from bsam_amd import hierarchy

CASES = ["frame_ch2d_amr_phys", "frame_tm3d_amr_phys"]


def _load(name):
    z = np.load(os.path.join(fu.GOLDEN, name + ".npz"))
    meta = json.loads(bytes(z["meta"]).decode())
    return meta, z


def _geom(hier, nd):
    """dx, xlower, xupper of every patch as the reference computes them:
    the root from the domain (bsamroutines.f90:378, 412-415), a child from
    its parent (MakeNewGrid :1749-1753), a level below the root from its
    child (InitSeed :477, 499)."""
    g = np.zeros((len(hier), 9))
    root = [k for k, d in enumerate(hier) if d["level"] == 0][0]
    mx = hier[root]["mx"]
    for a in range(3):
        g[root, a] = 1.0 / mx[a] if a < nd else 0.0
        g[root, 6 + a] = 1.0 if a < nd else 0.0
    for lv in range(1, max(d["level"] for d in hier) + 1):
        for k, d in enumerate(hier):
            if d["level"] != lv:
                continue
            p = d["parent"]
            for a in range(nd):
                g[k, 3 + a] = float(d["mbounds"][a][0] - 1) * g[p, a] + g[p, 3 + a]
                g[k, 6 + a] = float(d["mbounds"][a][1]) * g[p, a] + g[p, 3 + a]
                g[k, a] = g[p, a] / 2.0
    return g


def _fill(X, meta, z):
    hier = meta["hierarchy"]
    for k in meta["patches"]:
        X.put(k, "q", z[f"q_p{k}"].reshape(X.get(k, "q").shape))
    return hier


def _params(meta):
    fx = dict(meta)
    m = {"ch2d_amr_phys": dict(dt=1e-5, eps=0.02), "tm3d_amr_phys": dict(dt=1e-4, eps=0.04)}[meta["fixture"]]
    tm = dict(lamp=1.0, lama=0.1, kappa=10.0) if meta["nrvars"] == 4 else None
    return fu.step_params(dict(ndims=fx["ndims"], dt=m["dt"], eps=m["eps"], tol=1e-9, tumour=tm))


def _time(frame):
    return float(bytes(frame[:25]).decode())


def test_oracle_f25_matches_fortran_runtime():
    z = np.load(os.path.join(fu.GOLDEN, "f25_12.npz"))
    got = oracle.format_f25(z["values"])
    want = bytes(z["text"])
    bad = [(float(v), got[25 * i:25 * i + 25], want[25 * i:25 * i + 25]) for i, v in enumerate(z["values"])
           if got[25 * i:25 * i + 25] != want[25 * i:25 * i + 25]]
    assert not bad, bad[:5]


def test_library_f25_edit_on_the_host():
    """The library's own F25.12 edit (f25_edit, the function the frame
    kernel runs, here on the host through bsam_selftest_f25) against the
    Fortran runtime's on the edge values and against the oracle on 200 000
    values of every magnitude, halfway cases at the twelfth decimal and
    arbitrary bit patterns."""
    import ctypes as C
    from bsam_amd import abi, gpu
    lib = gpu.load()

    def edit(v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        buf = C.create_string_buffer(25 * len(v) + 1)
        assert lib.bsam_selftest_f25(v.ctypes.data_as(C.POINTER(C.c_double)), len(v), buf) == abi.BSAM_OK
        return buf.raw[:25 * len(v)]

    z = np.load(os.path.join(fu.GOLDEN, "f25_12.npz"))
    assert edit(z["values"]) == bytes(z["text"])
    rng = np.random.default_rng(17)
    v = np.concatenate([rng.uniform(-1, 1, 100000) * 10.0 ** rng.uniform(-20, 16, 100000),
                        np.arange(-20000, 20001) * 2.0 ** -13,
                        rng.integers(0, 2 ** 63, 60000, dtype=np.int64).view(np.float64)])
    assert edit(v) == oracle.format_f25(v)


@pytest.mark.parametrize("name", CASES)
def test_oracle_frame_matches_reference(name, tmp_path):
    meta, z = _load(name)
    hier = meta["hierarchy"]
    O = oracle.Oracle(_params(meta), hier)
    _fill(O, meta, z)
    path = str(tmp_path / "m00001.dat")
    O.write_frame(path, _time(z["frame"]), _geom(hier, meta["ndims"]))
    got, want = open(path, "rb").read(), bytes(z["frame"])
    assert got == want, next(i for i in range(min(len(got), len(want)) + 1) if got[i:i + 1] != want[i:i + 1])


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_gpu_frame_matches_reference(name, tmp_path, gpu_solver_cls):
    meta, z = _load(name)
    hier = meta["hierarchy"]
    print(f"[gpu] frame {name}", flush=True)
    S = gpu_solver_cls(_params(meta), hier)
    _fill(S, meta, z)
    path = str(tmp_path / "m00001.dat")
    S.write_frame(path, _time(z["frame"]), _geom(hier, meta["ndims"]))
    got, want = open(path, "rb").read(), bytes(z["frame"])
    S.close()
    assert got == want, next(i for i in range(min(len(got), len(want)) + 1) if got[i:i + 1] != want[i:i + 1])


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_gpu_frame_edge_values_vs_oracle(name, tmp_path, gpu_solver_cls):
    """Fields made of the edge values (and random magnitudes from 1e-14 to
    1e14): the HIP edit against the oracle's, default geometry."""
    meta, z = _load(name)
    hier = meta["hierarchy"]
    edge = np.load(os.path.join(fu.GOLDEN, "f25_12.npz"))["values"]
    rng = np.random.default_rng(3)
    print(f"[gpu] frame edge values {name}", flush=True)
    S = gpu_solver_cls(_params(meta), hier)
    O = oracle.Oracle(_params(meta), hier)
    for k in meta["patches"]:
        shp = S.get(k, "q").shape
        q = rng.standard_normal(shp) * 10.0 ** rng.uniform(-14, 14, shp)
        flat = q.reshape(-1)
        idx = rng.choice(flat.size, size=min(flat.size, len(edge)), replace=False)
        flat[idx] = edge[:len(idx)]
        S.put(k, "q", q)
        O.put(k, "q", q)
    p1, p2 = str(tmp_path / "gpu.dat"), str(tmp_path / "ora.dat")
    S.write_frame(p1, 0.125)
    O.write_frame(p2, 0.125)
    S.close()
    got, want = open(p1, "rb").read(), open(p2, "rb").read()
    assert got == want, next(i for i in range(min(len(got), len(want)) + 1) if got[i:i + 1] != want[i:i + 1])
