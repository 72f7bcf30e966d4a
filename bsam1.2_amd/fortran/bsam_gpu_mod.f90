! ISO_C_BINDING interface of libbsam_gpu.so (include/bsam_gpu.h) for a
! Fortran host such as BSAM 1.2's driver.
!
! The host keeps its patch tree (LIBRARY/treeops.f90, nodeinfodef.f90) and
! its time loop (LIBRARY/bsamroutines.f90:608-816); instead of the CPU
! V-cycle of LIBRARY/afasroutines.f90 it calls the library through the
! interfaces below.  Every function returns an INTEGER(C_INT32_T) status:
! BSAM_OK (0), BSAM_NOT_CONVERGED (1) or a negative error code whose text
! bsam_error_message() returns.  The library never stops the program; the
! host decides to STOP, as the reference does.
!
! Arrays cross the interface in the reference's own extents and order
! (column-major, i fastest, variable slowest, ghost layer included for q
! and qold), so `CALL bsam_put(ctx, p, BSAM_FIELD_Q, info%q)` hands over a
! patch's q(0:mx+1, 0:my+1[, 0:mz+1], 1:nrvars) as it is.
MODULE BSAMGPU
USE, INTRINSIC:: ISO_C_BINDING
IMPLICIT NONE
PRIVATE
!
INTEGER(C_INT32_T), PARAMETER, PUBLIC:: BSAM_OK = 0, BSAM_NOT_CONVERGED = 1, &
  BSAM_EINVAL = -1, BSAM_EHIP = -2, BSAM_ENOMEM = -3, BSAM_ESTATE = -4, &
  BSAM_EDIVERGED = -5, BSAM_ENODEV = -6
INTEGER(C_INT32_T), PARAMETER, PUBLIC:: BSAM_MODEL_CH = 1, BSAM_MODEL_BG = 2, &
  BSAM_MODEL_TUMOUR = 3
INTEGER(C_INT32_T), PARAMETER, PUBLIC:: BSAM_FIELD_Q = 0, BSAM_FIELD_QOLD = 1, &
  BSAM_FIELD_QC = 2, BSAM_FIELD_F = 3, BSAM_FIELD_RF = 4, BSAM_FIELD_FTMP = 5
INTEGER(C_INT32_T), PARAMETER, PUBLIC:: BSAM_PREC_FP64 = 0, BSAM_PREC_MIXED = 1
INTEGER(C_INT32_T), PARAMETER, PUBLIC:: BSAM_SOURCEFIELD = 1, BSAM_SOLUTIONFIELD = 2
INTEGER(C_INT32_T), PARAMETER, PUBLIC:: BSAM_KNOB_ORDER = 0, BSAM_KNOB_GATHER_ORDER_FP64 = 1, &
  BSAM_KNOB_GATHER_ORDER_FP32 = 2, BSAM_KNOB_SWEEP_PLANES_FP64 = 3, BSAM_KNOB_SWEEP_PLANES_FP32 = 4, &
  BSAM_KNOB_RESID_PLANES = 5, BSAM_KNOB_CHAIN_CELLS = 6, &
  BSAM_KNOB_REGRID_TABLES = 7, &
  BSAM_KNOB_GATHER_BLOCK_FP64 = 8, BSAM_KNOB_GATHER_BLOCK_FP32 = 9
!
! bsam_patch_desc.  C declares mglobal[3][2] and mbounds[3][2] (dimension
! first); in Fortran order that is (1:2, 1:3), the transpose of the
! reference's nodeinfo%mglobal(1:maxdims, 1:2).  bsam_make_desc converts.
TYPE, BIND(C), PUBLIC:: bsam_patch_desc
  INTEGER(C_INT32_T):: level
  INTEGER(C_INT32_T):: parent
  INTEGER(C_INT32_T):: mx(3)
  INTEGER(C_INT32_T):: mglobal(2,3)
  INTEGER(C_INT32_T):: mbounds(2,3)
  INTEGER(C_INT32_T):: mthbc(6)
  REAL(C_DOUBLE):: dx
  REAL(C_DOUBLE):: xlower(3)
END TYPE bsam_patch_desc

! bsam_grid_params: the griddata of AMR per level 0..maxlevel-1
INTEGER, PARAMETER:: BSAM_MAX_AMR_LEVELS = 16
TYPE, BIND(C):: bsam_grid_params
  INTEGER(C_INT32_T):: maxlevel
  INTEGER(C_INT32_T):: errflagopt(BSAM_MAX_AMR_LEVELS)
  INTEGER(C_INT32_T):: ibuffer(BSAM_MAX_AMR_LEVELS)
  INTEGER(C_INT32_T):: minimumgridpoints(BSAM_MAX_AMR_LEVELS)
  REAL(C_DOUBLE):: qtolerance(BSAM_MAX_AMR_LEVELS)
  REAL(C_DOUBLE):: desiredfillratios(BSAM_MAX_AMR_LEVELS)
END TYPE bsam_grid_params
!
TYPE, BIND(C), PUBLIC:: bsam_params
  INTEGER(C_INT32_T):: ndims = 2
  INTEGER(C_INT32_T):: nrvars = 2
  INTEGER(C_INT32_T):: mbc = 1
  INTEGER(C_INT32_T):: maux = 0
  INTEGER(C_INT32_T):: model = BSAM_MODEL_CH
  INTEGER(C_INT32_T):: nsmoothingpasses = 2
  INTEGER(C_INT32_T):: maxvcycles = 20
  INTEGER(C_INT32_T):: errortype = 1
  INTEGER(C_INT32_T):: updateauxfreq = 1
  INTEGER(C_INT32_T):: precision = BSAM_PREC_FP64
  INTEGER(C_INT32_T):: batching = 1
  INTEGER(C_INT32_T):: use_graph = 1
  REAL(C_DOUBLE):: dt = 0.0_C_DOUBLE
  REAL(C_DOUBLE):: qerrortol = 1.0E-06_C_DOUBLE
  REAL(C_DOUBLE):: model_params(8) = 0.0_C_DOUBLE
END TYPE bsam_params
!
PUBLIC:: bsam_create, bsam_destroy, bsam_set_dt, bsam_set_hierarchy, &
         bsam_upload, bsam_download, bsam_field_size, bsam_relax_level, &
         bsam_set_ghost, bsam_level_relax, bsam_restrict_solution, &
         bsam_coarse_ghost_points, bsam_coarse_loading, bsam_correct_fine, &
         bsam_fill_down, bsam_get_source, bsam_update_source, bsam_update_aux, &
         bsam_residual, bsam_error_afas, bsam_copy_q_to_qold, bsam_vcycle, &
         bsam_estimate_level_errors, bsam_estimate_level_errors_noalloc, bsam_write_frame, &
         bsam_multigrid_iterations, bsam_time_step, bsam_regrid, bsam_new_level, bsam_amr, &
         bsam_get_hierarchy, bsam_time_step_amr, bsam_grid_params, BSAM_MAX_AMR_LEVELS, bsam_synchronize, &
         bsam_parse_frame, bsam_read_frame, &
         bsam_autotune, bsam_set_knob, bsam_get_knob, bsam_set_mixed_fallback, &
         bsam_mixed_fallbacks, bsam_last_error
PUBLIC:: bsam_make_desc, bsam_put, bsam_get, bsam_error_message
!
INTERFACE
  INTEGER(C_INT32_T) FUNCTION bsam_create(params, ctx) BIND(C, NAME='bsam_create')
    IMPORT:: C_INT32_T, C_PTR, bsam_params
    TYPE(bsam_params), INTENT(IN):: params
    TYPE(C_PTR), INTENT(OUT):: ctx
  END FUNCTION
  INTEGER(C_INT32_T) FUNCTION bsam_destroy(ctx) BIND(C, NAME='bsam_destroy')
    IMPORT:: C_INT32_T, C_PTR
    TYPE(C_PTR), VALUE:: ctx
  END FUNCTION
  INTEGER(C_INT32_T) FUNCTION bsam_set_dt(ctx, dt) BIND(C, NAME='bsam_set_dt')
    IMPORT:: C_INT32_T, C_PTR, C_DOUBLE
    TYPE(C_PTR), VALUE:: ctx
    REAL(C_DOUBLE), VALUE:: dt
  END FUNCTION
  INTEGER(C_INT32_T) FUNCTION bsam_set_hierarchy(ctx, npatch, desc, minlevel, finestlevel) &
      BIND(C, NAME='bsam_set_hierarchy')
    IMPORT:: C_INT32_T, C_PTR, bsam_patch_desc
    TYPE(C_PTR), VALUE:: ctx
    INTEGER(C_INT32_T), VALUE:: npatch, minlevel, finestlevel
    TYPE(bsam_patch_desc), INTENT(IN):: desc(*)
  END FUNCTION
  INTEGER(C_INT32_T) FUNCTION bsam_upload(ctx, patch, field, host, count) BIND(C, NAME='bsam_upload')
    IMPORT:: C_INT32_T, C_INT64_T, C_PTR, C_DOUBLE
    TYPE(C_PTR), VALUE:: ctx
    INTEGER(C_INT32_T), VALUE:: patch, field
    REAL(C_DOUBLE), INTENT(IN):: host(*)
    INTEGER(C_INT64_T), VALUE:: count
  END FUNCTION
  INTEGER(C_INT32_T) FUNCTION bsam_download(ctx, patch, field, host, count) BIND(C, NAME='bsam_download')
    IMPORT:: C_INT32_T, C_INT64_T, C_PTR, C_DOUBLE
    TYPE(C_PTR), VALUE:: ctx
    INTEGER(C_INT32_T), VALUE:: patch, field
    REAL(C_DOUBLE), INTENT(OUT):: host(*)
    INTEGER(C_INT64_T), VALUE:: count
  END FUNCTION
  INTEGER(C_INT64_T) FUNCTION bsam_field_size(ctx, patch, field) BIND(C, NAME='bsam_field_size')
    IMPORT:: C_INT32_T, C_INT64_T, C_PTR
    TYPE(C_PTR), VALUE:: ctx
    INTEGER(C_INT32_T), VALUE:: patch, field
  END FUNCTION
  INTEGER(C_INT32_T) FUNCTION bsam_relax_level(ctx, level, ipass) BIND(C, NAME='bsam_relax_level')
    IMPORT:: C_INT32_T, C_PTR
    TYPE(C_PTR), VALUE:: ctx
    INTEGER(C_INT32_T), VALUE:: level, ipass
  END FUNCTION
  INTEGER(C_INT32_T) FUNCTION bsam_set_ghost(ctx, level, ipass) BIND(C, NAME='bsam_set_ghost')
    IMPORT:: C_INT32_T, C_PTR
    TYPE(C_PTR), VALUE:: ctx
    INTEGER(C_INT32_T), VALUE:: level, ipass
  END FUNCTION
  INTEGER(C_INT32_T) FUNCTION bsam_level_relax(ctx, level) BIND(C, NAME='bsam_level_relax')
    IMPORT:: C_INT32_T, C_PTR
    TYPE(C_PTR), VALUE:: ctx
    INTEGER(C_INT32_T), VALUE:: level
  END FUNCTION
  INTEGER(C_INT32_T) FUNCTION bsam_restrict_solution(ctx, level) BIND(C, NAME='bsam_restrict_solution')
    IMPORT:: C_INT32_T, C_PTR
    TYPE(C_PTR), VALUE:: ctx
    INTEGER(C_INT32_T), VALUE:: level
  END FUNCTION
  INTEGER(C_INT32_T) FUNCTION bsam_coarse_ghost_points(ctx, level) BIND(C, NAME='bsam_coarse_ghost_points')
    IMPORT:: C_INT32_T, C_PTR
    TYPE(C_PTR), VALUE:: ctx
    INTEGER(C_INT32_T), VALUE:: level
  END FUNCTION
  INTEGER(C_INT32_T) FUNCTION bsam_coarse_loading(ctx, level) BIND(C, NAME='bsam_coarse_loading')
    IMPORT:: C_INT32_T, C_PTR
    TYPE(C_PTR), VALUE:: ctx
    INTEGER(C_INT32_T), VALUE:: level
  END FUNCTION
  INTEGER(C_INT32_T) FUNCTION bsam_correct_fine(ctx, level) BIND(C, NAME='bsam_correct_fine')
    IMPORT:: C_INT32_T, C_PTR
    TYPE(C_PTR), VALUE:: ctx
    INTEGER(C_INT32_T), VALUE:: level
  END FUNCTION
  INTEGER(C_INT32_T) FUNCTION bsam_fill_down(ctx, field) BIND(C, NAME='bsam_fill_down')
    IMPORT:: C_INT32_T, C_PTR
    TYPE(C_PTR), VALUE:: ctx
    INTEGER(C_INT32_T), VALUE:: field
  END FUNCTION
  INTEGER(C_INT32_T) FUNCTION bsam_get_source(ctx, level) BIND(C, NAME='bsam_get_source')
    IMPORT:: C_INT32_T, C_PTR
    TYPE(C_PTR), VALUE:: ctx
    INTEGER(C_INT32_T), VALUE:: level
  END FUNCTION
  INTEGER(C_INT32_T) FUNCTION bsam_update_source(ctx, level) BIND(C, NAME='bsam_update_source')
    IMPORT:: C_INT32_T, C_PTR
    TYPE(C_PTR), VALUE:: ctx
    INTEGER(C_INT32_T), VALUE:: level
  END FUNCTION
  INTEGER(C_INT32_T) FUNCTION bsam_update_aux(ctx, level) BIND(C, NAME='bsam_update_aux')
    IMPORT:: C_INT32_T, C_PTR
    TYPE(C_PTR), VALUE:: ctx
    INTEGER(C_INT32_T), VALUE:: level
  END FUNCTION
  INTEGER(C_INT32_T) FUNCTION bsam_residual(ctx, level) BIND(C, NAME='bsam_residual')
    IMPORT:: C_INT32_T, C_PTR
    TYPE(C_PTR), VALUE:: ctx
    INTEGER(C_INT32_T), VALUE:: level
  END FUNCTION
  INTEGER(C_INT32_T) FUNCTION bsam_error_afas(ctx, level, norm) BIND(C, NAME='bsam_error_afas')
    IMPORT:: C_INT32_T, C_PTR, C_DOUBLE
    TYPE(C_PTR), VALUE:: ctx
    INTEGER(C_INT32_T), VALUE:: level
    REAL(C_DOUBLE), INTENT(OUT):: norm
  END FUNCTION
  INTEGER(C_INT32_T) FUNCTION bsam_copy_q_to_qold(ctx, level) BIND(C, NAME='bsam_copy_q_to_qold')
    IMPORT:: C_INT32_T, C_PTR
    TYPE(C_PTR), VALUE:: ctx
    INTEGER(C_INT32_T), VALUE:: level
  END FUNCTION
  ! EstimateLevelErrors + ErrFlag tagging (bsamroutines.f90:994-1131):
  ! flags receives errorflags(1:mx) of every patch of the level, patches in
  ! descriptor order, each column-major.
  INTEGER(C_INT32_T) FUNCTION bsam_estimate_level_errors(ctx, level, errflagopt, qtolerance, flags, ntagged) &
      BIND(C, NAME='bsam_estimate_level_errors')
    IMPORT:: C_INT32_T, C_INT64_T, C_DOUBLE, C_PTR
    TYPE(C_PTR), VALUE:: ctx
    INTEGER(C_INT32_T), VALUE:: level, errflagopt
    REAL(C_DOUBLE), VALUE:: qtolerance
    INTEGER(C_INT32_T), DIMENSION(*), INTENT(OUT):: flags
    INTEGER(C_INT64_T), INTENT(OUT):: ntagged
  END FUNCTION
  ! The same without any device allocation (flags staged in a fixed 16 MiB
  ! device array: levels of at most 4194304 cells); for hosts that cannot
  ! allocate inside their time loop.  Overlapping calls on one GPU race.
  INTEGER(C_INT32_T) FUNCTION bsam_estimate_level_errors_noalloc(ctx, level, errflagopt, qtolerance, flags, &
      ntagged) BIND(C, NAME='bsam_estimate_level_errors_noalloc')
    IMPORT:: C_INT32_T, C_INT64_T, C_DOUBLE, C_PTR
    TYPE(C_PTR), VALUE:: ctx
    INTEGER(C_INT32_T), VALUE:: level, errflagopt
    REAL(C_DOUBLE), VALUE:: qtolerance
    INTEGER(C_INT32_T), DIMENSION(*), INTENT(OUT):: flags
    INTEGER(C_INT64_T), INTENT(OUT):: ntagged
  END FUNCTION
  ! WriteQ (bsaminputoutput.f90:60-181) to path (a NUL-terminated string);
  ! geom: dx(1:3), xlower(1:3), xupper(1:3) per descriptor
  INTEGER(C_INT32_T) FUNCTION bsam_write_frame(ctx, path, time, geom) BIND(C, NAME='bsam_write_frame')
    IMPORT:: C_INT32_T, C_DOUBLE, C_PTR, C_CHAR
    TYPE(C_PTR), VALUE:: ctx
    CHARACTER(KIND=C_CHAR), DIMENSION(*), INTENT(IN):: path
    REAL(C_DOUBLE), VALUE:: time
    REAL(C_DOUBLE), DIMENSION(*), INTENT(IN):: geom
  END FUNCTION
  INTEGER(C_INT32_T) FUNCTION bsam_vcycle(ctx, level) BIND(C, NAME='bsam_vcycle')
    IMPORT:: C_INT32_T, C_PTR
    TYPE(C_PTR), VALUE:: ctx
    INTEGER(C_INT32_T), VALUE:: level
  END FUNCTION
  INTEGER(C_INT32_T) FUNCTION bsam_multigrid_iterations(ctx, ncycles, residual, history, history_len) &
      BIND(C, NAME='bsam_multigrid_iterations')
    IMPORT:: C_INT32_T, C_PTR, C_DOUBLE
    TYPE(C_PTR), VALUE:: ctx
    INTEGER(C_INT32_T), INTENT(OUT):: ncycles
    REAL(C_DOUBLE), INTENT(OUT):: residual
    REAL(C_DOUBLE), INTENT(OUT):: history(*)
    INTEGER(C_INT32_T), VALUE:: history_len
  END FUNCTION
  INTEGER(C_INT32_T) FUNCTION bsam_time_step(ctx, first_step, ncycles, residual) BIND(C, NAME='bsam_time_step')
    IMPORT:: C_INT32_T, C_PTR, C_DOUBLE
    TYPE(C_PTR), VALUE:: ctx
    INTEGER(C_INT32_T), VALUE:: first_step
    INTEGER(C_INT32_T), INTENT(OUT):: ncycles
    REAL(C_DOUBLE), INTENT(OUT):: residual
  END FUNCTION
  INTEGER(C_INT32_T) FUNCTION bsam_regrid(ctx, npatch, desc, minlevel, finestlevel) &
      BIND(C, NAME='bsam_regrid')
    IMPORT:: C_INT32_T, C_PTR, bsam_patch_desc
    TYPE(C_PTR), VALUE:: ctx
    INTEGER(C_INT32_T), VALUE:: npatch, minlevel, finestlevel
    TYPE(bsam_patch_desc), INTENT(IN):: desc(*)
  END FUNCTION
  INTEGER(C_INT32_T) FUNCTION bsam_new_level(ndims, npatch, desc, level, flags, ibuffer, &
      minimumgridpoints, desiredfillratio, maxout, out, nout) BIND(C, NAME='bsam_new_level')
    IMPORT:: C_INT32_T, C_DOUBLE, bsam_patch_desc
    INTEGER(C_INT32_T), VALUE:: ndims, npatch, level, ibuffer, minimumgridpoints, maxout
    TYPE(bsam_patch_desc), INTENT(IN):: desc(*)
    INTEGER(C_INT32_T), INTENT(IN):: flags(*)
    REAL(C_DOUBLE), VALUE:: desiredfillratio
    TYPE(bsam_patch_desc), INTENT(OUT):: out(*)
    INTEGER(C_INT32_T), INTENT(OUT):: nout
  END FUNCTION
  INTEGER(C_INT32_T) FUNCTION bsam_amr(ctx, gp) BIND(C, NAME='bsam_amr')
    IMPORT:: C_INT32_T, C_PTR, bsam_grid_params
    TYPE(C_PTR), VALUE:: ctx
    TYPE(bsam_grid_params), INTENT(IN):: gp
  END FUNCTION
  INTEGER(C_INT32_T) FUNCTION bsam_get_hierarchy(ctx, maxout, out, npatch, minlevel, finestlevel) &
      BIND(C, NAME='bsam_get_hierarchy')
    IMPORT:: C_INT32_T, C_PTR, bsam_patch_desc
    TYPE(C_PTR), VALUE:: ctx
    INTEGER(C_INT32_T), VALUE:: maxout
    TYPE(bsam_patch_desc), INTENT(OUT):: out(*)
    INTEGER(C_INT32_T), INTENT(OUT):: npatch, minlevel, finestlevel
  END FUNCTION
  INTEGER(C_INT32_T) FUNCTION bsam_time_step_amr(ctx, gp, first_step, ncycles, residual) &
      BIND(C, NAME='bsam_time_step_amr')
    IMPORT:: C_INT32_T, C_PTR, C_DOUBLE, bsam_grid_params
    TYPE(C_PTR), VALUE:: ctx
    TYPE(bsam_grid_params), INTENT(IN):: gp
    INTEGER(C_INT32_T), VALUE:: first_step
    INTEGER(C_INT32_T), INTENT(OUT):: ncycles
    REAL(C_DOUBLE), INTENT(OUT):: residual
  END FUNCTION
  INTEGER(C_INT32_T) FUNCTION bsam_parse_frame(path, ndims, nrvars, time, finestlevel, maxpatch, desc, &
      npatch, q, maxq, nq) BIND(C, NAME='bsam_parse_frame')
    IMPORT:: C_INT32_T, C_INT64_T, C_CHAR, C_DOUBLE, bsam_patch_desc
    CHARACTER(KIND=C_CHAR), INTENT(IN):: path(*)     ! NUL-terminated
    INTEGER(C_INT32_T), VALUE:: ndims, nrvars, maxpatch
    REAL(C_DOUBLE), INTENT(OUT):: time
    INTEGER(C_INT32_T), INTENT(OUT):: finestlevel, npatch
    TYPE(bsam_patch_desc), INTENT(OUT):: desc(*)
    REAL(C_DOUBLE), INTENT(OUT):: q(*)
    INTEGER(C_INT64_T), VALUE:: maxq
    INTEGER(C_INT64_T), INTENT(OUT):: nq
  END FUNCTION
  INTEGER(C_INT32_T) FUNCTION bsam_read_frame(ctx, path, time, finestlevel) BIND(C, NAME='bsam_read_frame')
    IMPORT:: C_INT32_T, C_PTR, C_CHAR, C_DOUBLE
    TYPE(C_PTR), VALUE:: ctx
    CHARACTER(KIND=C_CHAR), INTENT(IN):: path(*)     ! NUL-terminated
    REAL(C_DOUBLE), INTENT(OUT):: time
    INTEGER(C_INT32_T), INTENT(OUT):: finestlevel
  END FUNCTION
  INTEGER(C_INT32_T) FUNCTION bsam_synchronize(ctx) BIND(C, NAME='bsam_synchronize')
    IMPORT:: C_INT32_T, C_PTR
    TYPE(C_PTR), VALUE:: ctx
  END FUNCTION
  INTEGER(C_INT32_T) FUNCTION bsam_autotune(ctx) BIND(C, NAME='bsam_autotune')
    IMPORT:: C_INT32_T, C_PTR
    TYPE(C_PTR), VALUE:: ctx
  END FUNCTION
  INTEGER(C_INT32_T) FUNCTION bsam_set_knob(ctx, knob, level, value) BIND(C, NAME='bsam_set_knob')
    IMPORT:: C_INT32_T, C_PTR
    TYPE(C_PTR), VALUE:: ctx
    INTEGER(C_INT32_T), VALUE:: knob, level, value
  END FUNCTION
  INTEGER(C_INT32_T) FUNCTION bsam_get_knob(ctx, knob, level, value) BIND(C, NAME='bsam_get_knob')
    IMPORT:: C_INT32_T, C_PTR
    TYPE(C_PTR), VALUE:: ctx
    INTEGER(C_INT32_T), VALUE:: knob, level
    INTEGER(C_INT32_T), INTENT(OUT):: value
  END FUNCTION
  INTEGER(C_INT32_T) FUNCTION bsam_set_mixed_fallback(ctx, stall_ratio) BIND(C, NAME='bsam_set_mixed_fallback')
    IMPORT:: C_INT32_T, C_PTR, C_DOUBLE
    TYPE(C_PTR), VALUE:: ctx
    REAL(C_DOUBLE), VALUE:: stall_ratio
  END FUNCTION
  INTEGER(C_INT32_T) FUNCTION bsam_mixed_fallbacks(ctx, count) BIND(C, NAME='bsam_mixed_fallbacks')
    IMPORT:: C_INT32_T, C_INT64_T, C_PTR
    TYPE(C_PTR), VALUE:: ctx
    INTEGER(C_INT64_T), INTENT(OUT):: count
  END FUNCTION
  TYPE(C_PTR) FUNCTION bsam_last_error() BIND(C, NAME='bsam_last_error')
    IMPORT:: C_PTR
  END FUNCTION
END INTERFACE
!
CONTAINS
!
! A descriptor from the reference's nodeinfo fields (nodeinfodef.f90:
! 118-220): mglobal and mbounds as (1:3, 1:2), i.e. (dimension, lo/hi).
FUNCTION bsam_make_desc(level, parent, mx, mglobal, mbounds, mthbc, dx, xlower) RESULT(d)
INTEGER, INTENT(IN):: level, parent
INTEGER, DIMENSION(3), INTENT(IN):: mx
INTEGER, DIMENSION(3,2), INTENT(IN):: mglobal, mbounds
INTEGER, DIMENSION(6), INTENT(IN):: mthbc
REAL(KIND=C_DOUBLE), INTENT(IN):: dx
REAL(KIND=C_DOUBLE), DIMENSION(3), INTENT(IN):: xlower
TYPE(bsam_patch_desc):: d
d%level = level; d%parent = parent; d%mx = mx
d%mglobal = TRANSPOSE(mglobal); d%mbounds = TRANSPOSE(mbounds)
d%mthbc = mthbc; d%dx = dx; d%xlower = xlower
END FUNCTION bsam_make_desc
!
! Whole-array copies of one field of one patch; `a` is the patch's array
! in the reference's extents (any rank), passed as contiguous storage.
INTEGER(C_INT32_T) FUNCTION bsam_put(ctx, patch, field, a, n) RESULT(rc)
TYPE(C_PTR), INTENT(IN):: ctx
INTEGER, INTENT(IN):: patch, field
INTEGER(C_INT64_T), INTENT(IN):: n
REAL(C_DOUBLE), INTENT(IN):: a(n)
rc = bsam_upload(ctx, INT(patch,C_INT32_T), INT(field,C_INT32_T), a, n)
END FUNCTION bsam_put
!
INTEGER(C_INT32_T) FUNCTION bsam_get(ctx, patch, field, a, n) RESULT(rc)
TYPE(C_PTR), INTENT(IN):: ctx
INTEGER, INTENT(IN):: patch, field
INTEGER(C_INT64_T), INTENT(IN):: n
REAL(C_DOUBLE), INTENT(OUT):: a(n)
rc = bsam_download(ctx, INT(patch,C_INT32_T), INT(field,C_INT32_T), a, n)
END FUNCTION bsam_get
!
! The message of the last failed call, as a Fortran string.
FUNCTION bsam_error_message() RESULT(msg)
CHARACTER(LEN=512):: msg
CHARACTER(KIND=C_CHAR), DIMENSION(:), POINTER:: p
INTEGER:: n
TYPE(C_PTR):: cp
msg = ' '
cp = bsam_last_error()
IF(.NOT. C_ASSOCIATED(cp)) RETURN
CALL C_F_POINTER(cp, p, [512])
DO n = 1, 512
  IF(p(n) == C_NULL_CHAR) EXIT
  msg(n:n) = p(n)
END DO
END FUNCTION bsam_error_message
!
END MODULE BSAMGPU
