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


@pytest.mark.parametrize("name", CASES)
def test_parse_reference_frame(name, tmp_path):
    """ReadQ's parse (bsam_parse_frame, host only) of the reference's own
    frame: the time and finest level of the header; the patches of each
    level in the reverse of the file's order (the reference's tree visits
    the youngest node first after ReadQ), with the reference's extents,
    mglobal and xlower, and parents and mbounds as the reference's regrid
    made them; q equal to the frame's text read as doubles and within half a
    unit of the twelfth decimal of the q the reference held."""
    from bsam_amd.grid import parse_frame
    meta, z = _load(name)
    path = tmp_path / "m00001.dat"
    path.write_bytes(bytes(z["frame"]))
    nd, nv = meta["ndims"], meta["nrvars"]
    t, fin, hier, qs = parse_frame(str(path), nd, nv)
    ref = meta["hierarchy"]
    assert t == _time(z["frame"]) and fin == max(d["level"] for d in ref)
    want = []
    for lv in range(0, fin + 1):
        want += [k for k in meta["patches"] if ref[k]["level"] == lv][::-1]
    assert len(hier) == len(want), (len(hier), len(want))
    at = {k: n for n, k in enumerate(want)}
    for n, k in enumerate(want):
        g, w = hier[n], ref[k]
        for key in ("level", "mx", "mglobal"):
            assert g[key] == w[key], (n, key, g[key], w[key])
        assert all(abs(g["xlower"][a] - w["xlower"][a]) < 1e-12 for a in range(nd)), (g["xlower"], w["xlower"])
        if w["level"] > 0:
            assert g["parent"] == at[w["parent"]] and g["mbounds"] == w["mbounds"], (n, g, w)
        qref = z[f"q_p{k}"].reshape(qs[n].shape)
        # half a unit of the twelfth decimal, plus the rounding of the text
        # to the nearest double
        assert np.all(np.abs(qs[n] - qref) <= 5e-13 + 2 * np.spacing(np.abs(qref))), (
            n, float(np.max(np.abs(qs[n] - qref))))
    # the values are the text's, read as the nearest doubles
    lines = bytes(z["frame"]).decode().splitlines()
    body = [ln for ln in lines if len(ln) == 27 * nv - 2]
    first = np.array([float(x) for x in body[0].split()])
    last = np.array([float(x) for x in body[-1].split()])
    assert np.array_equal(first, qs[0].reshape(nv, -1)[:, 0]), (first, qs[0].reshape(nv, -1)[:, 0])
    # the file's last patch is the first of its level in the visiting order
    assert np.array_equal(last, qs[at[meta["patches"][-1]]].reshape(nv, -1)[:, -1])


def test_parse_frame_refuses_a_broken_file(tmp_path):
    from bsam_amd.gpu import BsamError
    from bsam_amd.grid import parse_frame
    meta, z = _load(CASES[0])
    text = bytes(z["frame"]).decode().splitlines()
    text[3] = text[3].replace("  2", "  3", 1)  # ndims of the root patch
    path = tmp_path / "m00002.dat"
    path.write_text("\n".join(text) + "\n")
    with pytest.raises(BsamError, match="ndims"):
        parse_frame(str(path), meta["ndims"], meta["nrvars"])


def _blocks(frame):
    """The patch records of a frame text (each from its blank record), and
    its two header lines."""
    lines = frame.decode().rstrip("\n").split("\n")
    head, blocks, cur = lines[:2], [], None
    for ln in lines[2:]:
        if ln.strip() == "" and ln != "":
            if cur is not None:
                blocks.append("\n".join(cur))
            cur = []
        elif cur is not None:
            cur.append(ln)
    if cur:
        blocks.append("\n".join(cur))
    return head, blocks


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_gpu_read_frame_round_trip(name, tmp_path, gpu_solver_cls):
    """bsam_read_frame (ReadQ) of the reference's own frame onto a run that
    holds the frame's root: the refined mesh and every patch's q as the
    parse gives them; written back with bsam_write_frame, the frame has the
    reference's header and the reference's patch records, the patches of a
    level in the reverse order (the reference's own order after ReadQ)."""
    from bsam_amd.grid import parse_frame
    meta, z = _load(name)
    nd, nv = meta["ndims"], meta["nrvars"]
    src = tmp_path / "m00001.dat"
    src.write_bytes(bytes(z["frame"]))
    base = [d for d in meta["hierarchy"] if d["level"] <= 0]
    S = gpu_solver_cls(_params(meta), base)
    print(f"[gpu] read frame {name}", flush=True)
    t, fin = S.read_frame(str(src))
    t2, fin2, fh, fq = parse_frame(str(src), nd, nv)
    assert (t, fin) == (t2, fin2)
    nb = len(base) - 1  # the root's row
    assert len(S.hier) == nb + len(fh)
    for n, d in enumerate(fh):
        g = S.hier[nb + n]
        # the root keeps its installed row (its mbounds are its box one
        # level down)
        for key in ("level", "mx", "mglobal") + (("mbounds",) if n else ()):
            assert g[key] == d[key], (n, key)
        assert np.array_equal(S.get(nb + n, "q"), fq[n]), n
    out = tmp_path / "m00002.dat"
    S.write_frame(str(out), t, _geom(S.hier, nd))
    h1, b1 = _blocks(bytes(z["frame"]))
    h2, b2 = _blocks(out.read_bytes())
    assert h1 == h2 and sorted(b1) == sorted(b2) and b1 != [] and len(b1) == len(b2)
    S.close()
