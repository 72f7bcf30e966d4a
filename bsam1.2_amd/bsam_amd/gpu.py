"""Host-side mirror of the reference's V-cycle driver over the HIP library.

`Solver` owns one bsam_ctx of libbsam_gpu.so (built from csrc/ by
bsam_amd.build) and exposes the same per-level operations the reference
applies with ApplyOnLevel (LIBRARY/afasroutines.f90, LIBRARY/boundary.f90),
with the reference's names in snake case, plus field access in the
reference's array extents.  There is no CPU fallback: constructing a Solver
without the built library or without a visible GPU raises.
"""
import ctypes as C
import os

import numpy as np

from . import abi, hierarchy

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(PKG), "libbsam_gpu.so")
_lib = None


class BsamError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"bsam error {code}: {msg}")
        self.code = code


def load():
    """Load libbsam_gpu.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() (hipcc, gfx950)")
        _lib = abi.bind(C.CDLL(LIB_PATH))
    return _lib


class Solver:
    """One hierarchy on the GPU.  `params` is an abi.Params, `hier` a list of
    descriptor dicts (bsam_amd.hierarchy)."""

    def __init__(self, params, hier):
        self.lib = load()
        self.params = params
        self.ndims = params.ndims
        self.nv = params.nrvars
        self.ctx = C.c_void_p()
        self._chk(self.lib.bsam_create(C.byref(params), C.byref(self.ctx)))
        self.set_hierarchy(hier)

    def set_hierarchy(self, hier):
        self.hier = hier
        self.minlevel, self.finestlevel = hierarchy.levels(hier)
        self._descs = hierarchy.to_ctypes(hier)
        self._chk(self.lib.bsam_set_hierarchy(self.ctx, len(hier), self._descs, self.minlevel,
                                              self.finestlevel))

    def regrid(self, hier):
        """Carry the fields onto a new hierarchy (bsam_regrid)."""
        lo, hi = hierarchy.levels(hier)
        descs = hierarchy.to_ctypes(hier)
        self._chk(self.lib.bsam_regrid(self.ctx, len(hier), descs, lo, hi))
        self.hier, self._descs, self.minlevel, self.finestlevel = hier, descs, lo, hi

    def get_hierarchy(self):
        """The installed descriptors, as dicts."""
        n, lo, hi = C.c_int32(), C.c_int32(), C.c_int32()
        self._chk(self.lib.bsam_get_hierarchy(self.ctx, 0, None, C.byref(n), C.byref(lo), C.byref(hi)))
        out = (abi.PatchDesc * n.value)()
        self._chk(self.lib.bsam_get_hierarchy(self.ctx, n.value, out, C.byref(n), C.byref(lo), C.byref(hi)))
        return [hierarchy.from_ctypes(out[i]) for i in range(n.value)]

    def _took_mesh(self):
        # the new mesh is read from the library when it is first needed
        self._stale = True
        n, lo, hi = C.c_int32(), C.c_int32(), C.c_int32()
        self._chk(self.lib.bsam_get_hierarchy(self.ctx, 0, None, C.byref(n), C.byref(lo), C.byref(hi)))
        self.minlevel, self.finestlevel = lo.value, hi.value

    @property
    def hier(self):
        if getattr(self, "_stale", False):
            self._stale = False
            self._hier = self.get_hierarchy()
            self._descs = hierarchy.to_ctypes(self._hier)
        return self._hier

    @hier.setter
    def hier(self, h):
        self._stale = False
        self._hier = h

    def amr(self, grid):
        """AMR(rootlevel) with abi.GridParams `grid` (bsam_amr); the new mesh
        becomes self.hier."""
        self._chk(self.lib.bsam_amr(self.ctx, C.byref(grid)))
        self._took_mesh()

    def time_step_amr(self, grid, first):
        n = C.c_int32()
        r = C.c_double()
        rc = self._chk(self.lib.bsam_time_step_amr(self.ctx, C.byref(grid), 1 if first else 0, C.byref(n),
                                                   C.byref(r)))
        self._took_mesh()
        return rc, n.value, r.value

    def read_frame(self, path):
        """ReadQ onto this run (bsam_read_frame): returns (time, finestlevel);
        the frame's mesh becomes self.hier."""
        t, fin = C.c_double(), C.c_int32()
        self._chk(self.lib.bsam_read_frame(self.ctx, path.encode(), C.byref(t), C.byref(fin)))
        self._took_mesh()
        return t.value, fin.value

    def close(self):
        if getattr(self, "ctx", None) and self.ctx.value:
            self.lib.bsam_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc < 0:
            raise BsamError(rc, self.lib.bsam_last_error().decode())
        return rc

    # ---- fields -----------------------------------------------------------
    def get(self, patch, name):
        fid = abi.FIELD_NAMES[name]
        n = self.lib.bsam_field_size(self.ctx, patch, fid)
        out = np.empty(n, dtype=np.float64)
        self._chk(self.lib.bsam_download(self.ctx, patch, fid, out.ctypes.data_as(C.POINTER(C.c_double)), n))
        return out.reshape(hierarchy.field_shape(self.hier[patch], self.ndims, self.nv, name))

    def put(self, patch, name, arr):
        fid = abi.FIELD_NAMES[name]
        n = self.lib.bsam_field_size(self.ctx, patch, fid)
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        if a.size != n:
            raise ValueError(f"field {name} of patch {patch} holds {n} values, got {a.size}")
        self._chk(self.lib.bsam_upload(self.ctx, patch, fid, a.ctypes.data_as(C.POINTER(C.c_double)), n))

    # ---- per-level operations (one ApplyOnLevel each) -----------------------
    def relax_level(self, level, ipass):
        self._chk(self.lib.bsam_relax_level(self.ctx, level, ipass))

    def set_ghost(self, level, ipass):
        self._chk(self.lib.bsam_set_ghost(self.ctx, level, ipass))

    def level_relax(self, level):
        self._chk(self.lib.bsam_level_relax(self.ctx, level))

    def restrict_solution(self, level):
        self._chk(self.lib.bsam_restrict_solution(self.ctx, level))

    def coarse_ghost_points(self, level):
        self._chk(self.lib.bsam_coarse_ghost_points(self.ctx, level))

    def coarse_loading(self, level):
        self._chk(self.lib.bsam_coarse_loading(self.ctx, level))

    def correct_fine(self, level):
        self._chk(self.lib.bsam_correct_fine(self.ctx, level))

    def fill_down(self, field):
        self._chk(self.lib.bsam_fill_down(self.ctx, field))

    def get_source(self, level):
        self._chk(self.lib.bsam_get_source(self.ctx, level))

    def update_source(self, level):
        self._chk(self.lib.bsam_update_source(self.ctx, level))

    def residual(self, level):
        self._chk(self.lib.bsam_residual(self.ctx, level))

    def copy_q_to_qold(self, level):
        self._chk(self.lib.bsam_copy_q_to_qold(self.ctx, level))

    def write_frame(self, path, time, geom=None):
        """WriteQ: the frame of levels 0..finest to `path`; geom per patch
        (dx[3], xlower[3], xupper[3]) or None."""
        g = None if geom is None else np.ascontiguousarray(geom, dtype=np.float64).ctypes.data_as(C.POINTER(C.c_double))
        self._chk(self.lib.bsam_write_frame(self.ctx, path.encode(), time, g))

    def estimate_level_errors(self, level, errflagopt, qtolerance, ncells=None):
        """EstimateLevelErrors + ErrFlag tagging: the level's flags (patches
        in descriptor order, each column-major over 1..mx) and the number
        flagged.  The flag buffer is sized here from the installed
        hierarchy, since the library writes prod(mx) values per patch of the
        level; a caller's `ncells` must agree with that count."""
        n = sum(int(np.prod(d["mx"][:self.ndims])) for d in self.hier if d["level"] == level)
        if ncells is not None and ncells != n:
            raise ValueError(f"level {level} has {n} cells, not {ncells}")
        flags = np.zeros(n, dtype=np.int32)
        nt = C.c_int64()
        self._chk(self.lib.bsam_estimate_level_errors(self.ctx, level, errflagopt, qtolerance,
                                                      flags.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(nt)))
        return flags, nt.value

    def estimate_level_errors_noalloc(self, level, errflagopt, qtolerance):
        """The same tagging through the allocation-free entry point (flags
        staged in the library's fixed device scratch array)."""
        n = sum(int(np.prod(d["mx"][:self.ndims])) for d in self.hier if d["level"] == level)
        flags = np.zeros(n, dtype=np.int32)
        nt = C.c_int64()
        self._chk(self.lib.bsam_estimate_level_errors_noalloc(self.ctx, level, errflagopt, qtolerance,
                                                              flags.ctypes.data_as(C.POINTER(C.c_int32)),
                                                              C.byref(nt)))
        return flags, nt.value

    def vcycle(self, level):
        self._chk(self.lib.bsam_vcycle(self.ctx, level))

    def error_afas(self, level):
        r = C.c_double()
        self._chk(self.lib.bsam_error_afas(self.ctx, level, C.byref(r)))
        return r.value

    def multigrid_iterations(self, history_len=128):
        n = C.c_int32()
        r = C.c_double()
        h = (C.c_double * history_len)()
        rc = self._chk(self.lib.bsam_multigrid_iterations(self.ctx, C.byref(n), C.byref(r), h, history_len))
        return rc, n.value, r.value, [h[i] for i in range(min(n.value, history_len))]

    def time_step(self, first):
        n = C.c_int32()
        r = C.c_double()
        rc = self._chk(self.lib.bsam_time_step(self.ctx, 1 if first else 0, C.byref(n), C.byref(r)))
        return rc, n.value, r.value

    def set_dt(self, dt):
        self._chk(self.lib.bsam_set_dt(self.ctx, dt))

    def synchronize(self):
        self._chk(self.lib.bsam_synchronize(self.ctx))

    # ---- launch-shape autotuning --------------------------------------------
    def autotune(self):
        """Time the candidate launch shapes on scratch copies and keep the
        fastest (bit-identical fields either way); returns the report."""
        self._chk(self.lib.bsam_autotune(self.ctx))
        return self.tuning_report()

    def tuning_report(self):
        import json
        buf = C.create_string_buffer(1 << 16)
        self._chk(self.lib.bsam_tuning_report(self.ctx, buf, len(buf)))
        return json.loads(buf.value.decode())

    def set_knob(self, knob, level, value):
        """One launch-shape knob (abi.KNOB_*); bit-identical fields for
        every value."""
        self._chk(self.lib.bsam_set_knob(self.ctx, knob, level, value))

    def get_knob(self, knob, level=0):
        v = C.c_int32()
        self._chk(self.lib.bsam_get_knob(self.ctx, knob, level, C.byref(v)))
        return v.value

    def set_mixed_fallback(self, stall_ratio):
        """Mixed path: hand the step to FP64 V-cycles when an outer
        iteration reduces the residual by less than this ratio."""
        self._chk(self.lib.bsam_set_mixed_fallback(self.ctx, stall_ratio))

    def mixed_fallbacks(self):
        n = C.c_int64()
        self._chk(self.lib.bsam_mixed_fallbacks(self.ctx, C.byref(n)))
        return n.value

    # ---- instrumentation ----------------------------------------------------
    def kstat_enable(self, on=True):
        self._chk(self.lib.bsam_kstat_enable(self.ctx, 1 if on else 0))

    def ghost_table_digest(self, level):
        """Per flattened ghost table of the level (7): (copies, interpolation
        rules, first-launch copies, first-launch rules, hash of the bytes)."""
        out = (C.c_int64 * 35)()
        self._chk(self.lib.bsam_ghost_table_digest(self.ctx, level, out))
        return [tuple(out[5 * m:5 * m + 5]) for m in range(7)]

    def phase_enable(self, on=True):
        self._chk(self.lib.bsam_phase_enable(self.ctx, 1 if on else 0))

    def phase_read(self):
        """{phase name: (calls, ms)} of the wall-clock phase timers."""
        out = {}
        for i, name in enumerate(abi.PHASE_NAMES):
            n = C.c_int64()
            ms = C.c_double()
            self._chk(self.lib.bsam_phase_read(self.ctx, i, C.byref(n), C.byref(ms)))
            out[name] = (n.value, ms.value)
        return out

    def kstat_read(self, which):
        n = C.c_int64()
        ms = C.c_double()
        b = C.c_double()
        self._chk(self.lib.bsam_kstat_read(self.ctx, which, C.byref(n), C.byref(ms), C.byref(b)))
        return n.value, ms.value, b.value
