! A Fortran host of libbsam_gpu.so: the time loop of BSAM 1.2
! (LIBRARY/bsamroutines.f90:673-758) for Cahn-Hilliard on a periodic root
! grid, on the uniform mesh (maxlevel = 0) or with the reference's regrid
! every step, driving the library through the BSAMGPU module.
!
! It builds the root and the below-root multigrid levels as
! bsamroutines.f90:443-510 does (mx halved, dx doubled, mbounds covering the
! child, periodic mthbc inherited), seeds phi = base + amp*U(-1,1) with the
! counter-based generator keyed by (seed, level, global cell index) that the
! workloads of this repository use, and runs `steps` time steps of
! SetGhost / CopyQToQold / MultigridIterations.  One line per step reports
! the V-cycle count and the final residual as afasroutines.f90:105 prints it;
! the root's q is written at the end as raw little-endian doubles in the
! reference's order, q(0:mx+1, 0:my+1[, 0:mz+1], 1:2).
!
! Usage: bsam_gpu_driver ndims n steps dt eps tol seed base amp precision out.bin [maxlevel tagtol]
!
! With maxlevel > 0 every step is the reference's time step with its regrid
! (bsam_time_step_amr: AMR(rootlevel) with the gradient tagging at tagtol,
! ibuffer 1, minimumgridpoints 8, fill ratio 0.8, then CopyQToQold and
! MultigridIterations), and each step's line also reports the mesh the
! library built (bsam_get_hierarchy): its patch count and finest level.
PROGRAM bsam_gpu_driver
USE, INTRINSIC:: ISO_C_BINDING
USE BSAMGPU
IMPLICIT NONE
INTEGER, PARAMETER:: r8 = C_DOUBLE, i8 = C_INT64_T
INTEGER:: ndims, n, steps, seed, prec, nlev, l, a, s, i, j, k, minlevel, ng
REAL(KIND=r8):: dt, eps, tol, base, amp
CHARACTER(LEN=256):: arg, outfile
TYPE(bsam_params):: p
TYPE(bsam_patch_desc), ALLOCATABLE:: desc(:)
TYPE(C_PTR):: ctx
INTEGER(C_INT32_T):: rc, ncyc
REAL(C_DOUBLE):: res
REAL(KIND=r8), ALLOCATABLE:: q(:)
INTEGER(C_INT64_T):: nq
INTEGER:: mx(3), mxc(3), mg(3,2), mb(3,2), mth(6), kmax, maxlevel
REAL(KIND=r8):: h, tagtol
TYPE(bsam_grid_params):: grid
TYPE(bsam_patch_desc), ALLOCATABLE:: mesh(:)
INTEGER(C_INT32_T):: npatch, lo, fin
INTEGER(KIND=i8):: cells
!
IF(COMMAND_ARGUMENT_COUNT() < 11) THEN
  PRINT *, 'usage: bsam_gpu_driver ndims n steps dt eps tol seed base amp precision out.bin [maxlevel tagtol]'
  STOP 1
END IF
CALL GET_COMMAND_ARGUMENT(1, arg); READ(arg,*) ndims
CALL GET_COMMAND_ARGUMENT(2, arg); READ(arg,*) n
CALL GET_COMMAND_ARGUMENT(3, arg); READ(arg,*) steps
CALL GET_COMMAND_ARGUMENT(4, arg); READ(arg,*) dt
CALL GET_COMMAND_ARGUMENT(5, arg); READ(arg,*) eps
CALL GET_COMMAND_ARGUMENT(6, arg); READ(arg,*) tol
CALL GET_COMMAND_ARGUMENT(7, arg); READ(arg,*) seed
CALL GET_COMMAND_ARGUMENT(8, arg); READ(arg,*) base
CALL GET_COMMAND_ARGUMENT(9, arg); READ(arg,*) amp
CALL GET_COMMAND_ARGUMENT(10, arg); READ(arg,*) prec
CALL GET_COMMAND_ARGUMENT(11, outfile)
maxlevel = 0; tagtol = 0.0_r8
IF(COMMAND_ARGUMENT_COUNT() >= 13) THEN
  CALL GET_COMMAND_ARGUMENT(12, arg); READ(arg,*) maxlevel
  CALL GET_COMMAND_ARGUMENT(13, arg); READ(arg,*) tagtol
END IF
!
! levels 0 .. minlevel, coarsened to 2 cells per side
nlev = 0
i = n
DO WHILE(i >= 2 .AND. MOD(i,2) == 0)
  nlev = nlev+1; i = i/2
END DO
minlevel = -(nlev-1)
ALLOCATE(desc(nlev))
mth = 10
mth(1:2*ndims) = 2
mx = 1; mx(1:ndims) = n
h = 1.0_r8/REAL(n,r8)
DO l = 0, minlevel, -1
  mg = 1; mg(1:ndims,2) = mx(1:ndims)
  mxc = 1; mxc(1:ndims) = mx(1:ndims)/2
  mb = 1; mb(1:ndims,2) = mxc(1:ndims)
  IF(l == minlevel) THEN
    mb = 1; mb(1:ndims,2) = mx(1:ndims)
  END IF
  ! coarsest first; the parent of level l is the descriptor of level l-1
  desc(l-minlevel+1) = bsam_make_desc(l, l-minlevel-1, mx, mg, mb, mth, h, [0.0_r8, 0.0_r8, 0.0_r8])
  mx(1:ndims) = mxc(1:ndims)
  h = 2.0_r8*h
END DO
desc(1)%parent = -1
!
p%ndims = ndims; p%model = BSAM_MODEL_CH; p%maxvcycles = 50
p%precision = prec; p%dt = dt; p%qerrortol = tol
p%model_params(1) = eps*eps
rc = bsam_create(p, ctx)
CALL check('bsam_create')
rc = bsam_set_hierarchy(ctx, nlev, desc, minlevel, 0)
CALL check('bsam_set_hierarchy')
!
ng = n+2
kmax = 0
IF(ndims == 3) kmax = n+1
nq = INT(ng,i8)*INT(ng,i8)*INT(kmax+1,i8)*2_i8
ALLOCATE(q(nq))
q = 0.0_r8
DO k = MERGE(1,0,ndims==3), MERGE(n,0,ndims==3)
  DO j = 1, n
    DO i = 1, n
      q(1+i+ng*(j+ng*k)) = base+amp*seeded_u(seed, 0, i, j, MAX(k,1))
    END DO
  END DO
END DO
rc = bsam_put(ctx, nlev-1, BSAM_FIELD_Q, q, nq)
CALL check('bsam_upload')
rc = bsam_set_ghost(ctx, 0, 0)
CALL check('bsam_set_ghost')
rc = bsam_copy_q_to_qold(ctx, 0)
CALL check('bsam_copy_q_to_qold')
rc = bsam_autotune(ctx)
CALL check('bsam_autotune')
!
IF(maxlevel > 0) THEN
  grid%maxlevel = maxlevel
  grid%errflagopt = 11; grid%ibuffer = 1; grid%minimumgridpoints = 8
  grid%qtolerance = tagtol; grid%desiredfillratios = 0.8_r8
  ALLOCATE(mesh(100000))
END IF
DO s = 1, steps
  IF(maxlevel > 0) THEN
    rc = bsam_time_step_amr(ctx, grid, MERGE(1,0,s==1), ncyc, res)
    CALL check('bsam_time_step_amr')
    rc = bsam_get_hierarchy(ctx, SIZE(mesh), mesh, npatch, lo, fin)
    CALL check('bsam_get_hierarchy')
    cells = 0
    DO i = 1, npatch
      IF(mesh(i)%level == fin) cells = cells+PRODUCT(INT(mesh(i)%mx(1:ndims), i8))
    END DO
    PRINT '(A,I4,A,I4,A,ES24.16,A,I7,A,I3,A,I10)', ' step', s, ' vcycles', ncyc, ' residual', res, &
      ' patches', npatch, ' finest', fin, ' finest-cells', cells
  ELSE
    rc = bsam_time_step(ctx, MERGE(1,0,s==1), ncyc, res)
    CALL check('bsam_time_step')
    PRINT '(A,I4,A,I4,A,ES24.16)', ' step', s, ' vcycles', ncyc, ' residual', res
  END IF
END DO
!
rc = bsam_get(ctx, nlev-1, BSAM_FIELD_Q, q, nq)
CALL check('bsam_download')
OPEN(UNIT=31, FILE=TRIM(outfile), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
WRITE(31) q
CLOSE(31)
rc = bsam_destroy(ctx)
!
CONTAINS
!
SUBROUTINE check(what)
CHARACTER(LEN=*), INTENT(IN):: what
IF(rc < 0) THEN
  PRINT '(A,A,I4,A,A)', what, ' failed: status', rc, ': ', TRIM(bsam_error_message())
  IF(rc == BSAM_ENODEV) STOP 3
  STOP 2
END IF
END SUBROUTINE check
!
! 32-bit multiply modulo 2^32 in 16-bit halves, as the seeded generator of
! the workloads defines it.
FUNCTION mulmod32(x, m) RESULT(r)
INTEGER(KIND=i8), INTENT(IN):: x, m
INTEGER(KIND=i8):: r, xl, xh
INTEGER(KIND=i8), PARAMETER:: m16 = 65535_i8, m32 = 4294967295_i8
xl = IAND(x, m16); xh = IAND(ISHFT(x, -16), m16)
r = IAND(xl*m + ISHFT(IAND(xh*m, m16), 16), m32)
END FUNCTION mulmod32
!
FUNCTION mix32(x0) RESULT(x)
INTEGER(KIND=i8), INTENT(IN):: x0
INTEGER(KIND=i8):: x
INTEGER(KIND=i8), PARAMETER:: m32 = 4294967295_i8
x = IAND(x0, m32)
x = IEOR(x, ISHFT(x, -16)); x = mulmod32(x, 2146121005_i8)
x = IEOR(x, ISHFT(x, -15)); x = mulmod32(x, 2221713035_i8)
x = IEOR(x, ISHFT(x, -16))
END FUNCTION mix32
!
FUNCTION seeded_u(sd, lev, ig, jg, kg) RESULT(u)
INTEGER, INTENT(IN):: sd, lev, ig, jg, kg
REAL(KIND=r8):: u
INTEGER(KIND=i8):: x
x = mix32(INT(sd, i8))
x = mix32(IEOR(x, INT(lev+64, i8)))
x = mix32(IEOR(x, INT(ig, i8)))
x = mix32(IEOR(x, INT(jg, i8)))
x = mix32(IEOR(x, INT(kg, i8)))
u = 2.0_r8*(REAL(x, r8)/4294967296.0_r8)-1.0_r8
END FUNCTION seeded_u
!
END PROGRAM bsam_gpu_driver
