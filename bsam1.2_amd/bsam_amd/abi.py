"""ctypes mirror of the C ABI declared in include/bsam_gpu.h.

Struct layouts and constants here must match the header field for field;
tests/test_abi.py checks sizes and that the built library exports every
declared entry point.
"""
import ctypes as C

BSAM_OK = 0
BSAM_NOT_CONVERGED = 1
BSAM_EINVAL = -1
BSAM_EHIP = -2
BSAM_ENOMEM = -3
BSAM_ESTATE = -4
BSAM_EDIVERGED = -5
BSAM_ENODEV = -6

MODEL_CH = 1
MODEL_BG = 2
MODEL_TUMOUR = 3

FIELD_Q, FIELD_QOLD, FIELD_QC, FIELD_F, FIELD_RF, FIELD_FTMP = range(6)
FIELD_NAMES = {"q": FIELD_Q, "qold": FIELD_QOLD, "qc": FIELD_QC,
               "f": FIELD_F, "rf": FIELD_RF, "ftmp": FIELD_FTMP}

PREC_FP64 = 0
PREC_MIXED = 1

SOURCEFIELD = 1
SOLUTIONFIELD = 2

KSTAT_RELAX_FINEST = 0
KSTAT_RESID_FINEST = 1
KSTAT_FASRHS_FINEST = 2
KSTAT_COUNT = 3
PHASE_AMR_TAG = 0
PHASE_AMR_CLUSTER = 1
PHASE_AMR_INSTALL = 2
PHASE_AMR_FIELDS = 3
PHASE_SOLVE = 4
PHASE_GRAPH = 5
PHASE_COUNT = 6
PHASE_NAMES = ("amr_tag", "amr_cluster", "amr_install", "amr_fields", "solve", "graph")

KNOB_ORDER = 0
KNOB_GATHER_ORDER_FP64 = 1
KNOB_GATHER_ORDER_FP32 = 2
KNOB_SWEEP_PLANES_FP64 = 3
KNOB_SWEEP_PLANES_FP32 = 4
KNOB_RESID_PLANES = 5
KNOB_CHAIN_CELLS = 6
KNOB_REGRID_TABLES = 7
KNOB_GATHER_BLOCK_FP64 = 8
KNOB_GATHER_BLOCK_FP32 = 9
KNOB_TABLE_PLANNER = 10
KNOB_SWEEP_IMAGES = 11


class PatchDesc(C.Structure):
    """bsam_patch_desc (include/bsam_gpu.h)."""
    _fields_ = [
        ("level", C.c_int32),
        ("parent", C.c_int32),
        ("mx", C.c_int32 * 3),
        ("mglobal", (C.c_int32 * 2) * 3),
        ("mbounds", (C.c_int32 * 2) * 3),
        ("mthbc", C.c_int32 * 6),
        ("dx", C.c_double),
        ("xlower", C.c_double * 3),
    ]


MAX_AMR_LEVELS = 16


class GridParams(C.Structure):
    """bsam_grid_params (include/bsam_gpu.h)."""
    _fields_ = [
        ("maxlevel", C.c_int32),
        ("errflagopt", C.c_int32 * MAX_AMR_LEVELS),
        ("ibuffer", C.c_int32 * MAX_AMR_LEVELS),
        ("minimumgridpoints", C.c_int32 * MAX_AMR_LEVELS),
        ("qtolerance", C.c_double * MAX_AMR_LEVELS),
        ("desiredfillratios", C.c_double * MAX_AMR_LEVELS),
    ]


def grid_params(maxlevel, errflagopt, qtolerance, ibuffer=1, minimumgridpoints=8, desiredfillratio=0.8):
    """The same grid parameters on every level below maxlevel."""
    g = GridParams()
    g.maxlevel = maxlevel
    for lv in range(maxlevel):
        g.errflagopt[lv] = errflagopt
        g.qtolerance[lv] = qtolerance
        g.ibuffer[lv] = ibuffer
        g.minimumgridpoints[lv] = minimumgridpoints
        g.desiredfillratios[lv] = desiredfillratio
    return g


class Params(C.Structure):
    """bsam_params (include/bsam_gpu.h)."""
    _fields_ = [
        ("ndims", C.c_int32),
        ("nrvars", C.c_int32),
        ("mbc", C.c_int32),
        ("maux", C.c_int32),
        ("model", C.c_int32),
        ("nsmoothingpasses", C.c_int32),
        ("maxvcycles", C.c_int32),
        ("errortype", C.c_int32),
        ("updateauxfreq", C.c_int32),
        ("precision", C.c_int32),
        ("batching", C.c_int32),
        ("use_graph", C.c_int32),
        ("dt", C.c_double),
        ("qerrortol", C.c_double),
        ("model_params", C.c_double * 8),
    ]


# Every entry point declared in include/bsam_gpu.h, with its argtypes.
_P = C.c_void_p
_I = C.c_int32
_D = C.c_double
_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int32)
_LP = C.POINTER(C.c_int64)
ENTRY_POINTS = {
    "bsam_create": ([C.POINTER(Params), C.POINTER(_P)], _I),
    "bsam_destroy": ([_P], _I),
    "bsam_set_dt": ([_P, _D], _I),
    "bsam_set_hierarchy": ([_P, _I, C.POINTER(PatchDesc), _I, _I], _I),
    "bsam_upload": ([_P, _I, _I, _DP, C.c_int64], _I),
    "bsam_download": ([_P, _I, _I, _DP, C.c_int64], _I),
    "bsam_field_size": ([_P, _I, _I], C.c_int64),
    "bsam_relax_level": ([_P, _I, _I], _I),
    "bsam_set_ghost": ([_P, _I, _I], _I),
    "bsam_level_relax": ([_P, _I], _I),
    "bsam_restrict_solution": ([_P, _I], _I),
    "bsam_coarse_ghost_points": ([_P, _I], _I),
    "bsam_coarse_loading": ([_P, _I], _I),
    "bsam_correct_fine": ([_P, _I], _I),
    "bsam_fill_down": ([_P, _I], _I),
    "bsam_get_source": ([_P, _I], _I),
    "bsam_update_source": ([_P, _I], _I),
    "bsam_update_aux": ([_P, _I], _I),
    "bsam_residual": ([_P, _I], _I),
    "bsam_error_afas": ([_P, _I, _DP], _I),
    "bsam_copy_q_to_qold": ([_P, _I], _I),
    "bsam_estimate_level_errors": ([_P, _I, _I, _D, C.POINTER(_I), _LP], _I),
    "bsam_estimate_level_errors_noalloc": ([_P, _I, _I, _D, C.POINTER(_I), _LP], _I),
    "bsam_write_frame": ([_P, C.c_char_p, _D, _DP], _I),
    "bsam_selftest_f25": ([_DP, C.c_int64, C.c_char_p], _I),
    "bsam_selftest_thread_pool": ([_I, _I, _LP], _I),
    "bsam_vcycle": ([_P, _I], _I),
    "bsam_multigrid_iterations": ([_P, _IP, _DP, _DP, _I], _I),
    "bsam_time_step": ([_P, _I, _IP, _DP], _I),
    "bsam_regrid": ([_P, _I, C.POINTER(PatchDesc), _I, _I], _I),
    "bsam_amr": ([_P, C.POINTER(GridParams)], _I),
    "bsam_get_hierarchy": ([_P, _I, C.POINTER(PatchDesc), _IP, _IP, _IP], _I),
    "bsam_time_step_amr": ([_P, C.POINTER(GridParams), _I, _IP, _DP], _I),
    "bsam_parse_frame": ([C.c_char_p, _I, _I, _DP, _IP, _I, C.POINTER(PatchDesc), _IP, _DP, C.c_int64,
                          C.POINTER(C.c_int64)], _I),
    "bsam_read_frame": ([_P, C.c_char_p, _DP, _IP], _I),
    "bsam_new_level": ([_I, _I, C.POINTER(PatchDesc), _I, _IP, _I, _I, _D, _I, C.POINTER(PatchDesc), _IP], _I),
    "bsam_synchronize": ([_P], _I),
    "bsam_last_error": ([], C.c_char_p),
    "bsam_kstat_enable": ([_P, _I], _I),
    "bsam_kstat_read": ([_P, _I, _LP, _DP, _DP], _I),
    "bsam_phase_enable": ([_P, _I], _I),
    "bsam_ghost_table_digest": ([_P, _I, C.POINTER(C.c_int64)], _I),
    "bsam_phase_read": ([_P, _I, _LP, _DP], _I),
    "bsam_autotune": ([_P], _I),
    "bsam_tuning_report": ([_P, C.c_char_p, _I], _I),
    "bsam_set_knob": ([_P, _I, _I, _I], _I),
    "bsam_get_knob": ([_P, _I, _I, _IP], _I),
    "bsam_set_mixed_fallback": ([_P, _D], _I),
    "bsam_mixed_fallbacks": ([_P, _LP], _I),
    "bsam_selftest_ghost_maps": ([C.POINTER(Params), _I, C.POINTER(PatchDesc), _I, _I, C.c_uint64, _LP, _IP], _I),
}


def bind(lib):
    """Attach argtypes/restype for every ABI symbol of a loaded library."""
    for name, (args, res) in ENTRY_POINTS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib
