"""Host side of a regrid: the new patches of a level from its tags
(bsam_new_level, LIBRARY/bsamroutines.f90:1366-1735).  Runs on the CPU; it
needs the library, not a GPU."""
import ctypes as C

import numpy as np

from . import abi, hierarchy
from .gpu import BsamError, load


def new_level(ndims, hier, level, flags, ibuffer, minimumgridpoints, desiredfillratio):
    """Descriptor dicts of the new patches of level+1, in the reference's
    visiting order, parents indexed into `hier`.  `flags` are the level's
    tags as estimate_level_errors returns them."""
    lib = load()
    descs = hierarchy.to_ctypes(hier)
    f = np.ascontiguousarray(flags, dtype=np.int32)
    cap = max(16, len(hier))
    while True:
        out = (abi.PatchDesc * cap)()
        n = C.c_int32()
        rc = lib.bsam_new_level(ndims, len(hier), descs, level, f.ctypes.data_as(C.POINTER(C.c_int32)),
                                ibuffer, minimumgridpoints, desiredfillratio, cap, out, C.byref(n))
        if rc == abi.BSAM_EINVAL and n.value > cap:
            cap = n.value
            continue
        if rc < 0:
            raise BsamError(rc, lib.bsam_last_error().decode())
        return [hierarchy.from_ctypes(out[i]) for i in range(n.value)]


def parse_frame(path, ndims, nrvars):
    """ReadQ's parse of a frame (bsam_parse_frame): (time, finestlevel,
    descriptor dicts of levels 0..finest in the reference's visiting order
    after ReadQ, q of each patch in the q field's numpy shape)."""
    lib = load()
    t, fin, n, nq = C.c_double(), C.c_int32(), C.c_int32(), C.c_int64()
    p = path.encode()
    rc = lib.bsam_parse_frame(p, ndims, nrvars, C.byref(t), C.byref(fin), 0, None, C.byref(n), None, 0,
                              C.byref(nq))
    if rc < 0:
        raise BsamError(rc, lib.bsam_last_error().decode())
    descs = (abi.PatchDesc * n.value)()
    q = np.zeros(nq.value)
    rc = lib.bsam_parse_frame(p, ndims, nrvars, C.byref(t), C.byref(fin), n.value, descs, C.byref(n),
                              q.ctypes.data_as(C.POINTER(C.c_double)), nq.value, C.byref(nq))
    if rc < 0:
        raise BsamError(rc, lib.bsam_last_error().decode())
    hier = [hierarchy.from_ctypes(descs[i]) for i in range(n.value)]
    out, off = [], 0
    for d in hier:
        shp = hierarchy.field_shape(d, ndims, nrvars, "q")
        cnt = int(np.prod(shp))
        out.append(q[off:off + cnt].reshape(shp))
        off += cnt
    return t.value, fin.value, hier, out
