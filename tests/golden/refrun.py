"""Run the reference BSAM 1.2 build (oracle/_ref) and read what it wrote.

Test infrastructure: used only by tests/golden/make_golden.py in the build
container, where /root/reference and oracle/_ref exist.  It writes the
reference's three NAMELIST files (rundata, griddata, problemdata; read at
LIBRARY/bsamroutines.f90:99-105, :271-277 and by the PROBLEM module's
SetProb), optionally a restart frame in the format of
LIBRARY/bsaminputoutput.f90:60-181 (read by ReadQ :184-397), runs the
binary, and parses

  * the residual history printed by MultigridIterations
    (LIBRARY/afasroutines.f90:105), and
  * the binary dumps of the probe PROBLEM module (oracle/ref_problem):
    tag 1 = q and qold at the source build (state at MultigridIterations
    entry), tag 2 = q after a relaxation sweep, tag 3 = q after the step.
"""
import os
import re
import resource
import struct
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF_DIR = os.path.join(REPO, "oracle", "_ref")


def _nml(name, items):
    out = ["&" + name]
    for k, v in items:
        if isinstance(v, bool):
            s = ".true." if v else ".false."
        elif isinstance(v, int):
            s = str(v)
        elif isinstance(v, float):
            s = repr(float(v)).replace("e", "d")
            if "d" not in s:
                s += "d0"
        else:
            s = str(v)
        out.append(f"{k} = {s}")
    return "\n".join(out) + "/\n"


def write_inputs(d, *, ndims, mx, maxlevel, minlevel, dt, steps, qerrortol,
                 mthbc, xupper=1.0, maxvcycles=50, nsmoothingpasses=2,
                 restart=False, eps=0.03, problem=None, minimumgridpoints=8,
                 ibuffer=1, desiredfillratio=0.8, getafterstepstats=True, errortype=1, nrvars=2):
    os.makedirs(os.path.join(d, "OUT"), exist_ok=True)
    run = [("restart", restart), ("restartframe", 0), ("syncelliptic", False),
           ("outframes", 1), ("outputuniformmesh", False),
           ("getafterstepstats", getafterstepstats),
           ("timeiterations", steps), ("maxvcycles", maxvcycles),
           ("nsmoothingpasses", nsmoothingpasses), ("omega", 1.0),
           ("qerrortol", qerrortol), ("dt", dt), ("errortype", errortype)]
    grid = [("ndims", ndims), ("maxlevel", maxlevel), ("minlevel", minlevel)]
    for lv in range(0, 6):
        grid += [(f"qtolerance({lv})", 1.0e-2), (f"ibuffer({lv})", ibuffer),
                 (f"minimumgridpoints({lv})", minimumgridpoints),
                 (f"desiredfillratios({lv})", desiredfillratio),
                 (f"errflagopt({lv})", 10)]
    for a in range(ndims):
        grid += [(f"mx({a+1})", mx[a]), (f"mglobal({a+1},1)", 1),
                 (f"mglobal({a+1},2)", mx[a]), (f"xlower({a+1})", 0.0),
                 (f"xupper({a+1})", float(xupper))]
    grid += [("mbc", 1)]
    for f in range(2 * ndims):
        grid += [(f"mthbc({f+1})", mthbc[f])]
    grid += [("nrvars", nrvars), ("maux", 0)]
    with open(os.path.join(d, "rundata.dat"), "w") as fh:
        fh.write(_nml("rundata", run))
    with open(os.path.join(d, "griddata.dat"), "w") as fh:
        fh.write(_nml("griddata", grid))
    prob = [("eps", eps)] if problem is None else list(problem)
    if problem is None:
        prob = [("alpha", 1.0), ("eps", eps)]
    with open(os.path.join(d, "problemdata.dat"), "w") as fh:
        fh.write(_nml("problemdata", prob))


def write_restart_frame(d, ndims, mx, h, q):
    """Frame 0 with one root patch, q of shape (nv, mz+2?, my+2, mx+2) in
    reference order (i fastest); values are written F25.12 as WriteQ2D/3D
    do (bsaminputoutput.f90:132-181)."""
    path = os.path.join(d, "OUT", "m00000.dat")
    with open(path, "w") as fh:
        fh.write("%25.12f\n" % 0.0)
        fh.write("%3d\n" % 0)
        fh.write(" \n")
        fh.write("%3d %3d %3d %3d\n" % (0, ndims, 2, q.shape[0]))
        if ndims == 2:
            fh.write("%25.12f %25.12f\n" % (h, h))
            fh.write("%25.12f %25.12f\n" % (0.0, 0.0))
            fh.write("%25.12f %25.12f\n" % (h * mx[0], h * mx[1]))
            fh.write("%8d %8d\n" % (mx[0], mx[1]))
            fh.write("%8d %8d %8d %8d\n" % (1, mx[0], 1, mx[1]))
            for j in range(mx[1] + 2):
                for i in range(mx[0] + 2):
                    fh.write("  ".join("%25.12f" % q[v, j, i] for v in range(q.shape[0])) + "\n")
        else:
            fh.write("%25.12f %25.12f %25.12f\n" % (h, h, h))
            fh.write("%25.12f %25.12f %25.12f\n" % (0.0, 0.0, 0.0))
            fh.write("%25.12f %25.12f %25.12f\n" % (h * mx[0], h * mx[1], h * mx[2]))
            fh.write("%8d %8d %8d\n" % (mx[0], mx[1], mx[2]))
            fh.write("%8d %8d %8d %8d %8d %8d\n" % (1, mx[0], 1, mx[1], 1, mx[2]))
            for k in range(mx[2] + 2):
                for j in range(mx[1] + 2):
                    for i in range(mx[0] + 2):
                        fh.write("  ".join("%25.12f" % q[v, k, j, i] for v in range(q.shape[0])) + "\n")
    return path


def _unlimited_stack():
    # the reference keeps whole-patch array temporaries on the stack
    resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))


def run(binary, d, timeout=600):
    exe = os.path.join(REF_DIR, binary)
    p = subprocess.run([exe], cwd=d, capture_output=True, text=True, timeout=timeout,
                       preexec_fn=_unlimited_stack)
    if p.returncode != 0 or "Error" in p.stdout:
        raise RuntimeError(f"reference run failed in {d}:\n{p.stdout[-3000:]}\n{p.stderr[-2000:]}")
    return p.stdout


_RES = re.compile(r"it\s*=\s*(\d+)\s+residual\s*=\s*([-+0-9.EeDd]+)")


def parse_residuals(stdout):
    """List of per-MultigridIterations residual histories (list of floats)."""
    hist, cur, last = [], [], 0
    for m in _RES.finditer(stdout):
        it, r = int(m.group(1)), float(m.group(2).replace("D", "E").replace("d", "e"))
        if it <= last and cur:
            hist.append(cur)
            cur = []
        cur.append(r)
        last = it
    if cur:
        hist.append(cur)
    return hist


def parse_dump(path):
    """Records (tag, ipass, mx[3], h, xlower[3], data) from dump.bin."""
    recs = []
    with open(path, "rb") as fh:
        buf = fh.read()
    off = 0
    hsz = 5 * 4 + 8 + 24 + 8
    while off + hsz <= len(buf):
        tag, ipass, m1, m2, m3 = struct.unpack_from("<5i", buf, off)
        h = struct.unpack_from("<d", buf, off + 20)[0]
        xl = struct.unpack_from("<3d", buf, off + 28)
        cnt = struct.unpack_from("<q", buf, off + 52)[0]
        off += hsz
        data = np.frombuffer(buf, dtype="<f8", count=cnt, offset=off).copy()
        off += 8 * cnt
        recs.append(dict(tag=tag, ipass=ipass, mx=(m1, m2, m3), h=h, xlower=xl, data=data))
    return recs


def parse_frame(path, ndims):
    """Patches of an output frame written by WriteQ / OutputQ
    (LIBRARY/bsaminputoutput.f90:60-181): list of dicts with level, mx,
    mglobal and q as (nrvars, [mz+2,] my+2, mx+2), values as written
    (F25.12)."""
    toks = open(path).read().split()
    pos = 2  # time, finestlevel
    patches = []
    while pos < len(toks):
        level, nd, _two, nv = (int(t) for t in toks[pos:pos + 4])
        pos += 4
        pos += 3 * nd  # dx, xlower, xupper
        mx = [int(t) for t in toks[pos:pos + nd]]
        pos += nd
        mg = [int(t) for t in toks[pos:pos + 2 * nd]]
        pos += 2 * nd
        ext = [m + 2 for m in mx]
        cnt = int(np.prod(ext)) * nv
        vals = np.array([float(t) for t in toks[pos:pos + cnt]])
        pos += cnt
        # rows are (i fastest) with the nv values of one cell on a line
        q = vals.reshape(tuple(ext[::-1]) + (nv,))
        q = np.moveaxis(q, -1, 0)
        patches.append(dict(level=level, mx=mx, mglobal=mg, q=q))
    return patches

