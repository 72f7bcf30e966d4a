"""Host-side construction of BSAM patch hierarchies as bsam_patch_desc tables.

The reference keeps its hierarchy in a linked tree (LIBRARY/treeops.f90) and
fills each node in three places, all restated here so that a host that is
not the reference's Fortran can produce the same descriptor table:

  * RootInit (LIBRARY/bsamroutines.f90:226-440): the root patch, dx from
    xlower/xupper, mthbc padded with 10 (PeriodicSetup, boundary.f90:815-816);
  * InitSeed (bsamroutines.f90:443-510): one patch per level below the root,
    mx halved, dx doubled, mglobal = 1..mx, mbounds = 1..mx, and the child's
    mbounds reset to 1..mx/2;
  * MakeNewGrid (bsamroutines.f90:1663-1784): a refined child from a parent
    and the parent-local box `mbounds` -- mx, mglobal, mthbc inheritance of
    physical/periodic codes, xlower and dx.

A hierarchy is a list of dicts with the bsam_patch_desc fields; `to_ctypes`
packs it.  Order within a level is the visiting order.
"""
import ctypes as C

from .abi import PatchDesc

INTERNALBC = 999  # nodeinfodef.f90:62


def _desc(level, parent, mx, mglobal, mbounds, mthbc, dx, xlower):
    return dict(level=level, parent=parent, mx=list(mx), mglobal=[list(r) for r in mglobal],
                mbounds=[list(r) for r in mbounds], mthbc=list(mthbc), dx=float(dx),
                xlower=list(xlower))


def uniform(ndims, mx, minlevel, mthbc, xlower=(0.0, 0.0, 0.0), xupper=(1.0, 1.0, 1.0)):
    """Root patch plus the below-root multigrid levels down to `minlevel`.

    Returns the descriptor list ordered coarsest first (index 0 = minlevel).
    """
    mx3 = [mx[a] if a < ndims else 1 for a in range(3)]
    for a in range(ndims):
        if mx3[a] % 2:
            raise ValueError("root mx must be even (bsamroutines.f90:345-351)")
    dx = (xupper[0] - xlower[0]) / mx3[0]
    mth = [10] * 6
    mth[:2 * ndims] = list(mthbc)[:2 * ndims]
    root = _desc(0, -1, mx3, [[1, mx3[a]] if a < ndims else [1, 1] for a in range(3)],
                 [[1, mx3[a]] if a < ndims else [1, 1] for a in range(3)], mth, dx,
                 [xlower[a] if a < ndims else 0.0 for a in range(3)])
    chain = [root]
    for lv in range(-1, minlevel - 1, -1):
        child = chain[-1]
        if any(child["mx"][a] % 2 for a in range(ndims)):
            raise ValueError(f"grid on level {child['level']} will not coarsen (bsamroutines.f90:475-478)")
        smx = [child["mx"][a] // 2 if a < ndims else 1 for a in range(3)]
        seed = _desc(lv, -1, smx, [[1, child["mglobal"][a][1] // 2] if a < ndims else [1, 1] for a in range(3)],
                     [[1, smx[a]] if a < ndims else [1, 1] for a in range(3)], child["mthbc"],
                     child["dx"] * 2.0, child["xlower"])
        child["mbounds"] = [[1, smx[a]] if a < ndims else [1, 1] for a in range(3)]
        chain.append(seed)
    chain.reverse()  # coarsest first
    for i, d in enumerate(chain):
        d["parent"] = i - 1 if i > 0 else -1
    return chain


def refine(hier, parent_index, mbounds, ndims):
    """Append the child MakeNewGrid (bsamroutines.f90:1663-1784) would create
    inside patch `parent_index` over the parent-local box `mbounds`
    ([[lo, hi] per dim]).  Returns the new index."""
    par = hier[parent_index]
    mb = [list(mbounds[a]) if a < ndims else [1, 1] for a in range(3)]
    mx = [(mb[a][1] - mb[a][0] + 1) * 2 if a < ndims else 1 for a in range(3)]
    mg = []
    for a in range(3):
        if a < ndims:
            lo = par["mglobal"][a][0] + mb[a][0] - 1
            hi = lo + mb[a][1] - mb[a][0]
            mg.append([(lo - 1) * 2 + 1, hi * 2])
        else:
            mg.append([1, 1])
    mth = [INTERNALBC] * 6
    for n in range(ndims):
        lo, hi = 2 * n, 2 * n + 1
        if par["mthbc"][lo] < INTERNALBC and mb[n][0] == 1:
            mth[lo] = par["mthbc"][lo]
        if par["mthbc"][hi] < INTERNALBC and mb[n][1] == par["mx"][n]:
            mth[hi] = par["mthbc"][hi]
    for f in range(2 * ndims, 6):
        mth[f] = INTERNALBC
    xl = [(mb[a][0] - 1) * par["dx"] + par["xlower"][a] if a < ndims else 0.0 for a in range(3)]
    d = _desc(par["level"] + 1, parent_index, mx, mg, mb, mth, par["dx"] / 2.0, xl)
    hier.append(d)
    return len(hier) - 1


def from_patch_boxes(base, boxes, ndims):
    """Extend a uniform hierarchy with refined patches given as
    (level, mglobal) pairs in visiting order.  Each patch's parent is the
    patch one level down whose global box contains the patch's coarse box
    (every refined patch is created inside one parent, MakeNewGrid)."""
    hier = [dict(d) for d in base]
    for level, mg in boxes:
        cbox = [[(mg[a][0] + 1) // 2, mg[a][1] // 2] for a in range(ndims)]
        parent = None
        for i, d in enumerate(hier):
            if d["level"] != level - 1:
                continue
            if all(d["mglobal"][a][0] <= cbox[a][0] and cbox[a][1] <= d["mglobal"][a][1] for a in range(ndims)):
                parent = i
                break
        if parent is None:
            raise ValueError(f"no parent for level-{level} patch {mg}")
        pm = hier[parent]["mglobal"]
        mb = [[cbox[a][0] - pm[a][0] + 1, cbox[a][1] - pm[a][0] + 1] for a in range(ndims)]
        refine(hier, parent, mb, ndims)
    return hier


def levels(hier):
    lv = sorted({d["level"] for d in hier})
    return lv[0], lv[-1]


def to_ctypes(hier):
    arr = (PatchDesc * len(hier))()
    for i, d in enumerate(hier):
        p = arr[i]
        p.level = d["level"]
        p.parent = d["parent"]
        for a in range(3):
            p.mx[a] = d["mx"][a]
            p.mglobal[a][0], p.mglobal[a][1] = d["mglobal"][a]
            p.mbounds[a][0], p.mbounds[a][1] = d["mbounds"][a]
            p.xlower[a] = d["xlower"][a]
        for f in range(6):
            p.mthbc[f] = d["mthbc"][f]
        p.dx = d["dx"]
    return arr


def from_ctypes(p):
    """The descriptor dict of one PatchDesc."""
    return _desc(p.level, p.parent, [p.mx[a] for a in range(3)],
                 [[p.mglobal[a][0], p.mglobal[a][1]] for a in range(3)],
                 [[p.mbounds[a][0], p.mbounds[a][1]] for a in range(3)],
                 [p.mthbc[f] for f in range(6)], p.dx, [p.xlower[a] for a in range(3)])


def field_shape(d, ndims, nv, field):
    """numpy shape (reference order reversed: var, k, j, i) of one field."""
    mx = d["mx"]
    if field in ("q", "qold"):
        ext = [mx[a] + 2 for a in range(ndims)]
    elif field == "qc":
        ext = [(1 if mx[a] % 2 else mx[a] // 2) + 2 for a in range(ndims)]
    else:
        ext = [mx[a] for a in range(ndims)]
    return tuple([nv] + ext[::-1])
