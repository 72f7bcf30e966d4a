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
