! Probe PROBLEM module: Cahn-Hilliard hooks with seeded initial data,
! gradient tagging and binary dumps, compiled together with the unmodified
! BSAM 1.2 LIBRARY (oracle/Makefile) to produce golden vectors.  TEST
! INFRASTRUCTURE ONLY -- it never ships with the product.
!
! The hook math (RelaxGrid, Operator, Source, SourceUpdate, UserBC) restates
! the Cahn-Hilliard convex-splitting step of PROBLEM/CH/problem.f90
! (RelaxGrid2D 318-369, Operator2D 393-422, Source2D 425-450, RelaxGrid3D
! 554-614, Operator3D 638-669, Source3D 672-695, UserBC2D 770-818,
! UserBC3D 821-881) in the same floating-point evaluation order; the
! residual histories of this module and of the unmodified CH module are
! compared bit for bit in tests/test_oracle_pins.py.
!
! With model = 3 the same hooks carry the tumour-growth system described in
! problemdef.f90 (nrvars = 4), a model the reference does not ship; it is
! written here with the hook signatures of PROBLEM/CH so that the
! reference's LIBRARY runs it unmodified.
MODULE Problem
USE NodeInfoDef, ONLY: r8
IMPLICIT NONE
INTEGER, PARAMETER, PRIVATE:: i8 = SELECTED_INT_KIND(18)
INTEGER(KIND=i8), PARAMETER, PRIVATE:: m32 = 4294967295_i8, m16 = 65535_i8
!
CONTAINS
!
SUBROUTINE SetProb
USE NodeInfoDef
USE ProblemDef
IMPLICIT NONE
INTEGER:: ierror
NAMELIST/problemdata/alpha, eps, initmode, seed, base, amp, radius, xc, &
                     hroot, tagtol, dumpsweeps, dumpsrc, dumpend, pamp, lmax, &
                     model, lamp, lama, kappa
model = 1; lamp = 1.0_r8; lama = 0.1_r8; kappa = 10.0_r8
alpha = 1.0_r8; eps = 0.03_r8; initmode = 1; seed = 1001; base = 0.0_r8
amp = 0.05_r8; radius = 0.25_r8; xc = 0.5_r8; hroot = 1.0_r8/128.0_r8
tagtol = 0.05_r8; dumpsweeps = 0; dumpsrc = .FALSE.; dumpend = .FALSE.
pamp = 0.02_r8; lmax = 6
OPEN(UNIT=75,FILE='problemdata.dat',STATUS='OLD',ACTION='READ',IOSTAT=ierror)
IF(ierror/=0) THEN
  PRINT *,'Error opening input file problemdata.dat. Program stop.'
  STOP
END IF
READ(75,NML=problemdata)
CLOSE(75)
eps2 = eps*eps
ndumpedsweeps = 0
dumpopen = .FALSE.
IF(dumpsweeps>0 .OR. dumpsrc .OR. dumpend) THEN
  OPEN(UNIT=91,FILE='dump.bin',ACCESS='STREAM',FORM='UNFORMATTED', &
       STATUS='REPLACE',ACTION='WRITE')
  dumpopen = .TRUE.
END IF
END SUBROUTINE SetProb
!
SUBROUTINE AfterRun
USE ProblemDef
IF(dumpopen) CLOSE(91)
END SUBROUTINE AfterRun
!
! ---- dumps -------------------------------------------------------------
SUBROUTINE DumpHeader(tag,ipass,mx,h,xlower,ndims,count)
INTEGER, INTENT(IN):: tag, ipass, ndims
INTEGER, DIMENSION(:), INTENT(IN):: mx
REAL(KIND=r8), INTENT(IN):: h
REAL(KIND=r8), DIMENSION(:), INTENT(IN):: xlower
INTEGER(KIND=i8), INTENT(IN):: count
INTEGER(KIND=4):: hdr(5)
REAL(KIND=r8):: xl(3)
hdr = 1
hdr(1) = tag; hdr(2) = ipass
hdr(3:2+ndims) = mx(1:ndims)
xl = 0.0_r8; xl(1:ndims) = xlower(1:ndims)
WRITE(91) hdr, h, xl, count
END SUBROUTINE DumpHeader
!
! ---- seeded counter-based generator ------------------------------------
! A 32-bit integer hash applied to (seed, level, i, j, k); every step uses
! only add, xor, shift and a 32x16-bit split multiply, so int64 Fortran,
! C and Python compute identical values.
FUNCTION MulMod32(x,m) RESULT(r)
INTEGER(KIND=i8), INTENT(IN):: x, m
INTEGER(KIND=i8):: r, xl, xh
xl = IAND(x,m16); xh = IAND(ISHFT(x,-16),m16)
r = IAND(xl*m + ISHFT(IAND(xh*m,m16),16),m32)
END FUNCTION MulMod32
!
FUNCTION Mix32(x0) RESULT(x)
INTEGER(KIND=i8), INTENT(IN):: x0
INTEGER(KIND=i8):: x
x = IAND(x0,m32)
x = IEOR(x,ISHFT(x,-16)); x = MulMod32(x,2146121005_i8)
x = IEOR(x,ISHFT(x,-15)); x = MulMod32(x,2221713035_i8)
x = IEOR(x,ISHFT(x,-16))
END FUNCTION Mix32
!
FUNCTION SeededU(lev,i,j,k) RESULT(u)
USE ProblemDef
INTEGER, INTENT(IN):: lev, i, j, k
REAL(KIND=r8):: u
INTEGER(KIND=i8):: x
x = Mix32(INT(seed,i8))
x = Mix32(IEOR(x,INT(lev+64,i8)))
x = Mix32(IEOR(x,INT(i,i8)))
x = Mix32(IEOR(x,INT(j,i8)))
x = Mix32(IEOR(x,INT(k,i8)))
u = 2.0_r8*(REAL(x,r8)/4294967296.0_r8)-1.0_r8
END FUNCTION SeededU
!
! Radius of the perturbed sphere in the direction (dx,dy,dz), |d| = r:
! radius + pamp * sum_{l=1..lmax} sum_{m=-l..l} c_lm Y_lm(theta,psi), with
! orthonormal real spherical harmonics, cos(theta) = dz/r, psi = atan2(dy,dx),
! and associated Legendre functions from the standard upward recursion
! (Condon-Shortley phase).
FUNCTION SphereRadius(dx,dy,dz,r) RESULT(rs)
USE ProblemDef
REAL(KIND=r8), INTENT(IN):: dx, dy, dz, r
REAL(KIND=r8):: rs, ct, st, psi, pmm, pa, pb, pl, fact, nrm, ylm, s
INTEGER:: l, mm, k
REAL(KIND=r8), PARAMETER:: pi = 3.14159265358979323846_r8
IF(r>0.0_r8) THEN
  ct = dz/r; psi = ATAN2(dy,dx)
ELSE
  ct = 1.0_r8; psi = 0.0_r8
END IF
st = SQRT(MAX(0.0_r8,1.0_r8-ct*ct))
s = 0.0_r8
DO mm = 0, lmax
  ! P_mm^mm = (-1)^mm (2mm-1)!! st^mm
  pmm = 1.0_r8; fact = 1.0_r8
  DO k = 1, mm
    pmm = -pmm*fact*st; fact = fact+2.0_r8
  END DO
  pa = 0.0_r8; pb = pmm
  DO l = mm, lmax
    IF(l==mm) THEN
      pl = pmm
    ELSE
      pl = (ct*REAL(2*l-1,r8)*pb-REAL(l+mm-1,r8)*pa)/REAL(l-mm,r8)
      pa = pb; pb = pl
    END IF
    IF(l>=1) THEN
      ! sqrt((2l+1)/(4 pi) (l-mm)!/(l+mm)!)
      nrm = 1.0_r8
      DO k = l-mm+1, l+mm
        nrm = nrm/REAL(k,r8)
      END DO
      nrm = SQRT(REAL(2*l+1,r8)/(4.0_r8*pi)*nrm)
      IF(mm==0) THEN
        s = s+SeededU(-32,l,8,1)*nrm*pl
      ELSE
        ylm = SQRT(2.0_r8)*nrm*pl
        s = s+SeededU(-32,l,mm+8,1)*ylm*COS(REAL(mm,r8)*psi)
        s = s+SeededU(-32,l,-mm+8,1)*ylm*SIN(REAL(mm,r8)*psi)
      END IF
    END IF
  END DO
END DO
rs = radius+pamp*s
END FUNCTION SphereRadius
!
FUNCTION InitPhi(lev,ig,jg,kg,x,y,z) RESULT(p)
USE ProblemDef
INTEGER, INTENT(IN):: lev, ig, jg, kg
REAL(KIND=r8), INTENT(IN):: x, y, z
REAL(KIND=r8):: p, r
SELECT CASE(initmode)
CASE(1)
  p = base+amp*SeededU(lev,ig,jg,kg)
CASE(2)
  r = SQRT((x-xc(1))**2+(y-xc(2))**2+(z-xc(3))**2)
  p = TANH((radius-r)/(SQRT(2.0_r8)*eps))+amp*SeededU(lev,ig,jg,kg)
CASE(3)
  r = SQRT((x-xc(1))**2+(y-xc(2))**2+(z-xc(3))**2)
  p = TANH((SphereRadius(x-xc(1),y-xc(2),z-xc(3),r)-r)/(SQRT(2.0_r8)*eps)) &
    + amp*SeededU(lev,ig,jg,kg)
CASE DEFAULT
  p = 0.0_r8
END SELECT
END FUNCTION InitPhi
!
SUBROUTINE QInit2D(q,mx,nrvars,h,xlower)
USE NodeInfoDef
USE ProblemDef
REAL(KIND=r8), DIMENSION(1:,1:,1:), INTENT(OUT):: q
INTEGER, DIMENSION(1:2), INTENT(IN):: mx
INTEGER, INTENT(IN):: nrvars
REAL(KIND=r8), INTENT(IN):: h
REAL(KIND=r8), DIMENSION(1:2), INTENT(IN):: xlower
INTEGER:: i, j, lev, i0, j0
REAL(KIND=r8):: x, y
lev = NINT(LOG(hroot/h)/LOG(2.0_r8))
i0 = NINT(xlower(1)/h); j0 = NINT(xlower(2)/h)
DO j = 1, mx(2)
  y = (REAL(j,r8)-0.5_r8)*h+xlower(2)
  DO i = 1, mx(1)
    x = (REAL(i,r8)-0.5_r8)*h+xlower(1)
    q(i,j,1) = InitPhi(lev,i0+i,j0+j,1,x,y,xc(3))
    q(i,j,2) = 0.0_r8
    IF(model==3) THEN
      q(i,j,3) = 1.0_r8; q(i,j,4) = 0.0_r8
    END IF
  END DO
END DO
END SUBROUTINE QInit2D
!
SUBROUTINE QInit3D(q,mx,nrvars,h,xlower)
USE NodeInfoDef
USE ProblemDef
REAL(KIND=r8), DIMENSION(1:,1:,1:,1:), INTENT(OUT):: q
INTEGER, DIMENSION(1:3), INTENT(IN):: mx
INTEGER, INTENT(IN):: nrvars
REAL(KIND=r8), INTENT(IN):: h
REAL(KIND=r8), DIMENSION(1:3), INTENT(IN):: xlower
INTEGER:: i, j, k, lev, i0, j0, k0
REAL(KIND=r8):: x, y, z
lev = NINT(LOG(hroot/h)/LOG(2.0_r8))
i0 = NINT(xlower(1)/h); j0 = NINT(xlower(2)/h); k0 = NINT(xlower(3)/h)
DO k = 1, mx(3)
  z = (REAL(k,r8)-0.5_r8)*h+xlower(3)
  DO j = 1, mx(2)
    y = (REAL(j,r8)-0.5_r8)*h+xlower(2)
    DO i = 1, mx(1)
      x = (REAL(i,r8)-0.5_r8)*h+xlower(1)
      q(i,j,k,1) = InitPhi(lev,i0+i,j0+j,k0+k,x,y,z)
      q(i,j,k,2) = 0.0_r8
      IF(model==3) THEN
        q(i,j,k,3) = 1.0_r8; q(i,j,k,4) = 0.0_r8
      END IF
    END DO
  END DO
END DO
END SUBROUTINE QInit3D
!
SUBROUTINE AfterStep2D(q,qc,mx,nrvars,h,xlower,level)
USE ProblemDef
REAL(KIND=r8), DIMENSION(0:,0:,1:), INTENT(IN):: q
REAL(KIND=r8), DIMENSION(0:,0:,1:), INTENT(IN):: qc
INTEGER, DIMENSION(1:2), INTENT(IN):: mx
INTEGER, INTENT(IN):: nrvars, level
REAL(KIND=r8), INTENT(IN):: h
REAL(KIND=r8), DIMENSION(1:2), INTENT(IN):: xlower
IF(dumpend) THEN
  CALL DumpHeader(3,level,mx,h,xlower,2,INT(SIZE(q),i8))
  WRITE(91) q
END IF
END SUBROUTINE AfterStep2D
!
SUBROUTINE AfterStep3D(q,qc,mx,nrvars,h,xlower,level)
USE ProblemDef
REAL(KIND=r8), DIMENSION(0:,0:,0:,1:), INTENT(IN):: q
REAL(KIND=r8), DIMENSION(0:,0:,0:,1:), INTENT(IN):: qc
INTEGER, DIMENSION(1:3), INTENT(IN):: mx
INTEGER, INTENT(IN):: nrvars, level
REAL(KIND=r8), INTENT(IN):: h
REAL(KIND=r8), DIMENSION(1:3), INTENT(IN):: xlower
IF(dumpend) THEN
  CALL DumpHeader(3,level,mx,h,xlower,3,INT(SIZE(q),i8))
  WRITE(91) q
END IF
END SUBROUTINE AfterStep3D
!
SUBROUTINE SetAux(info)
USE NodeInfoDef
TYPE(nodeinfo):: info
END SUBROUTINE SetAux
!
SUBROUTINE SetSrc(info)
USE NodeInfoDef
TYPE(nodeinfo):: info
END SUBROUTINE SetSrc
!
! Tagging |grad phi|*h > tagtol with central differences.
SUBROUTINE SetErrFlagsUser2D(qrte,qnew,errorflags,mx,cmx,nrvars,h,xlower,level)
USE ProblemDef
REAL(KIND=r8), DIMENSION(1:,1:,1:), INTENT(IN):: qrte
REAL(KIND=r8), DIMENSION(0:,0:,1:), INTENT(IN):: qnew
INTEGER, DIMENSION(1:,1:), INTENT(OUT):: errorflags
INTEGER, DIMENSION(1:2), INTENT(IN):: mx, cmx
INTEGER, INTENT(IN):: nrvars, level
REAL(KIND=r8), INTENT(IN):: h
REAL(KIND=r8), DIMENSION(1:2), INTENT(IN):: xlower
INTEGER:: i, j
REAL(KIND=r8):: gx, gy
DO j = 1, mx(2)
  DO i = 1, mx(1)
    gx = 0.5_r8*(qnew(i+1,j,1)-qnew(i-1,j,1))
    gy = 0.5_r8*(qnew(i,j+1,1)-qnew(i,j-1,1))
    errorflags(i,j) = 0
    IF(SQRT(gx*gx+gy*gy)>tagtol) errorflags(i,j) = 1
  END DO
END DO
END SUBROUTINE SetErrFlagsUser2D
!
SUBROUTINE SetErrFlagsUser3D(qrte,qnew,errorflags,mx,cmx,nrvars,h,xlower,level)
USE ProblemDef
REAL(KIND=r8), DIMENSION(1:,1:,1:,1:), INTENT(IN):: qrte
REAL(KIND=r8), DIMENSION(0:,0:,0:,1:), INTENT(IN):: qnew
INTEGER, DIMENSION(1:,1:,1:), INTENT(OUT):: errorflags
INTEGER, DIMENSION(1:3), INTENT(IN):: mx, cmx
INTEGER, INTENT(IN):: nrvars, level
REAL(KIND=r8), INTENT(IN):: h
REAL(KIND=r8), DIMENSION(1:3), INTENT(IN):: xlower
INTEGER:: i, j, k
REAL(KIND=r8):: gx, gy, gz
DO k = 1, mx(3)
  DO j = 1, mx(2)
    DO i = 1, mx(1)
      gx = 0.5_r8*(qnew(i+1,j,k,1)-qnew(i-1,j,k,1))
      gy = 0.5_r8*(qnew(i,j+1,k,1)-qnew(i,j-1,k,1))
      gz = 0.5_r8*(qnew(i,j,k+1,1)-qnew(i,j,k-1,1))
      errorflags(i,j,k) = 0
      IF(SQRT(gx*gx+gy*gy+gz*gz)>tagtol) errorflags(i,j,k) = 1
    END DO
  END DO
END DO
END SUBROUTINE SetErrFlagsUser3D
!
! ---- multigrid hooks ---------------------------------------------------
SUBROUTINE RelaxGrid2D(qnew,qold,aux,f,mx,nrvars,maux,h,xlower,ipass)
USE NodeInfoDef
USE ProblemDef
REAL(KIND=r8), DIMENSION(0:,0:,1:), INTENT(IN OUT):: qnew
REAL(KIND=r8), DIMENSION(0:,0:,1:), INTENT(IN):: qold
REAL(KIND=r8), DIMENSION(0:,0:,1:), INTENT(IN):: aux
REAL(KIND=r8), DIMENSION(1:,1:,1:), INTENT(IN):: f
INTEGER, DIMENSION(1:2), INTENT(IN):: mx
INTEGER, INTENT(IN):: nrvars, maux, ipass
REAL(KIND=r8), INTENT(IN):: h
REAL(KIND=r8), DIMENSION(1:2), INTENT(IN):: xlower
INTEGER:: i, j
REAL(KIND=r8):: cdt, ceps, cdiag, c12, c21, r1, r2, p, d
IF(model==3) THEN
  CALL TumourRelax2D(qnew,qold,f,mx,h,ipass)
ELSE
cdt = dt/(h*h); ceps = eps2/(h*h); cdiag = 4.0_r8*eps2/(h*h)
c12 = 4.0_r8*cdt
DO j = 1, mx(2)
  DO i = 1+MODULO(j+ipass,2), mx(1), 2
    p = qnew(i,j,1)
    c21 = -3.0_r8*(p*p)-cdiag
    r1 = f(i,j,1)+cdt*(qnew(i-1,j,2)+qnew(i+1,j,2)+qnew(i,j-1,2)+qnew(i,j+1,2))
    r2 = f(i,j,2)-ceps*(qnew(i+1,j,1)+qnew(i-1,j,1)+qnew(i,j+1,1)+qnew(i,j-1,1)) &
       - 2.0_r8*(p*p*p)
    d = 1.0_r8-c12*c21
    qnew(i,j,1) = (r1-c12*r2)/d
    qnew(i,j,2) = (r2-c21*r1)/d
  END DO
END DO
END IF
IF(ndumpedsweeps<dumpsweeps) THEN
  ndumpedsweeps = ndumpedsweeps+1
  CALL DumpHeader(2,ipass,mx,h,xlower,2,INT(SIZE(qnew),i8))
  WRITE(91) qnew
END IF
END SUBROUTINE RelaxGrid2D
!
SUBROUTINE RelaxGrid3D(qnew,qold,aux,f,mx,nrvars,maux,h,xlower,ipass)
USE NodeInfoDef
USE ProblemDef
REAL(KIND=r8), DIMENSION(0:,0:,0:,1:), INTENT(IN OUT):: qnew
REAL(KIND=r8), DIMENSION(0:,0:,0:,1:), INTENT(IN):: qold
REAL(KIND=r8), DIMENSION(0:,0:,0:,1:), INTENT(IN):: aux
REAL(KIND=r8), DIMENSION(1:,1:,1:,1:), INTENT(IN):: f
INTEGER, DIMENSION(1:3), INTENT(IN):: mx
INTEGER, INTENT(IN):: nrvars, maux, ipass
REAL(KIND=r8), INTENT(IN):: h
REAL(KIND=r8), DIMENSION(1:3), INTENT(IN):: xlower
INTEGER:: i, j, k
REAL(KIND=r8):: cdt, ceps, cdiag, c12, c21, r1, r2, p, d
IF(model==3) THEN
  CALL TumourRelax3D(qnew,qold,f,mx,h,ipass)
ELSE
cdt = dt/(h*h); ceps = eps2/(h*h); cdiag = 6.0_r8*eps2/(h*h)
c12 = 6.0_r8*cdt
DO k = 1, mx(3)
  DO j = 1, mx(2)
    DO i = 1+MODULO(k+j+ipass,2), mx(1), 2
      p = qnew(i,j,k,1)
      c21 = -3.0_r8*(p*p)-cdiag
      r1 = f(i,j,k,1)+cdt*(qnew(i-1,j,k,2)+qnew(i+1,j,k,2)+qnew(i,j-1,k,2) &
         + qnew(i,j+1,k,2)+qnew(i,j,k-1,2)+qnew(i,j,k+1,2))
      r2 = f(i,j,k,2)-ceps*(qnew(i+1,j,k,1)+qnew(i-1,j,k,1)+qnew(i,j+1,k,1) &
         + qnew(i,j-1,k,1)+qnew(i,j,k+1,1)+qnew(i,j,k-1,1))-2.0_r8*(p*p*p)
      d = 1.0_r8-c12*c21
      qnew(i,j,k,1) = (r1-c12*r2)/d
      qnew(i,j,k,2) = (r2-c21*r1)/d
    END DO
  END DO
END DO
END IF
IF(ndumpedsweeps<dumpsweeps) THEN
  ndumpedsweeps = ndumpedsweeps+1
  CALL DumpHeader(2,ipass,mx,h,xlower,3,INT(SIZE(qnew),i8))
  WRITE(91) qnew
END IF
END SUBROUTINE RelaxGrid3D
!
SUBROUTINE UpdateAuxVcycle2D(qnew,qold,aux,ll,mx,nrvars,maux,mbc,h,xlower)
INTEGER, INTENT(IN):: ll
REAL(KIND=r8), DIMENSION(ll:,ll:,1:), INTENT(IN):: qnew, qold
REAL(KIND=r8), DIMENSION(ll:,ll:,1:), INTENT(IN OUT):: aux
INTEGER, DIMENSION(1:2), INTENT(IN):: mx
INTEGER, INTENT(IN):: nrvars, maux, mbc
REAL(KIND=r8), INTENT(IN):: h
REAL(KIND=r8), DIMENSION(1:2), INTENT(IN):: xlower
END SUBROUTINE UpdateAuxVcycle2D
!
SUBROUTINE UpdateAuxVcycle3D(qnew,qold,aux,ll,mx,nrvars,maux,mbc,h,xlower)
INTEGER, INTENT(IN):: ll
REAL(KIND=r8), DIMENSION(ll:,ll:,ll:,1:), INTENT(IN):: qnew, qold
REAL(KIND=r8), DIMENSION(ll:,ll:,ll:,1:), INTENT(IN OUT):: aux
INTEGER, DIMENSION(1:3), INTENT(IN):: mx
INTEGER, INTENT(IN):: nrvars, maux, mbc
REAL(KIND=r8), INTENT(IN):: h
REAL(KIND=r8), DIMENSION(1:3), INTENT(IN):: xlower
END SUBROUTINE UpdateAuxVcycle3D
!
FUNCTION Operator2D(qnew,qold,aux,mx,nrvars,maux,h,xlower) RESULT(op)
USE NodeInfoDef
USE ProblemDef
REAL(KIND=r8), DIMENSION(0:,0:,1:), INTENT(IN):: qnew, qold, aux
INTEGER, DIMENSION(1:2), INTENT(IN):: mx
INTEGER, INTENT(IN):: nrvars, maux
REAL(KIND=r8), INTENT(IN):: h
REAL(KIND=r8), DIMENSION(1:2), INTENT(IN):: xlower
REAL(KIND=r8), DIMENSION(1:mx(1),1:mx(2),1:nrvars):: op
INTEGER:: i, j
REAL(KIND=r8):: cdt, ceps, l1, l2, l3, l4, p, ch, t0, n
cdt = dt/(h*h); ceps = eps2/(h*h); ch = 1.0_r8/(h*h)
DO j = 1, mx(2)
  DO i = 1, mx(1)
    l2 = qnew(i+1,j,2)+qnew(i-1,j,2)+qnew(i,j+1,2)+qnew(i,j-1,2)-4.0_r8*qnew(i,j,2)
    l1 = qnew(i+1,j,1)+qnew(i-1,j,1)+qnew(i,j+1,1)+qnew(i,j-1,1)-4.0_r8*qnew(i,j,1)
    p = qnew(i,j,1)
    IF(model==3) THEN
      l3 = qnew(i+1,j,3)+qnew(i-1,j,3)+qnew(i,j+1,3)+qnew(i,j-1,3)-4.0_r8*qnew(i,j,3)
      l4 = qnew(i+1,j,4)+qnew(i-1,j,4)+qnew(i,j+1,4)+qnew(i,j-1,4)-4.0_r8*qnew(i,j,4)
      t0 = 0.5_r8*(qold(i,j,1)+1.0_r8)
      n = qnew(i,j,3)
      op(i,j,1) = p-cdt*l2-dt*lamp*t0*n
      op(i,j,2) = qnew(i,j,2)-(p*p*p)+ceps*l1
      op(i,j,3) = -ch*l3+kappa*t0*n
      op(i,j,4) = -ch*l4-lamp*t0*n
    ELSE
      op(i,j,1) = p-cdt*l2
      op(i,j,2) = qnew(i,j,2)-(p*p*p)+ceps*l1
    END IF
  END DO
END DO
END FUNCTION Operator2D
!
FUNCTION Operator3D(qnew,qold,aux,mx,nrvars,maux,h,xlower) RESULT(op)
USE NodeInfoDef
USE ProblemDef
REAL(KIND=r8), DIMENSION(0:,0:,0:,1:), INTENT(IN):: qnew, qold, aux
INTEGER, DIMENSION(1:3), INTENT(IN):: mx
INTEGER, INTENT(IN):: nrvars, maux
REAL(KIND=r8), INTENT(IN):: h
REAL(KIND=r8), DIMENSION(1:3), INTENT(IN):: xlower
REAL(KIND=r8), DIMENSION(1:mx(1),1:mx(2),1:mx(3),1:nrvars):: op
INTEGER:: i, j, k
REAL(KIND=r8):: cdt, ceps, l1, l2, l3, l4, p, ch, t0, n
cdt = dt/(h*h); ceps = eps2/(h*h); ch = 1.0_r8/(h*h)
DO k = 1, mx(3)
  DO j = 1, mx(2)
    DO i = 1, mx(1)
      l2 = qnew(i+1,j,k,2)+qnew(i-1,j,k,2)+qnew(i,j+1,k,2)+qnew(i,j-1,k,2) &
         + qnew(i,j,k+1,2)+qnew(i,j,k-1,2)-6.0_r8*qnew(i,j,k,2)
      l1 = qnew(i+1,j,k,1)+qnew(i-1,j,k,1)+qnew(i,j+1,k,1)+qnew(i,j-1,k,1) &
         + qnew(i,j,k+1,1)+qnew(i,j,k-1,1)-6.0_r8*qnew(i,j,k,1)
      p = qnew(i,j,k,1)
      IF(model==3) THEN
        l3 = qnew(i+1,j,k,3)+qnew(i-1,j,k,3)+qnew(i,j+1,k,3)+qnew(i,j-1,k,3) &
           + qnew(i,j,k+1,3)+qnew(i,j,k-1,3)-6.0_r8*qnew(i,j,k,3)
        l4 = qnew(i+1,j,k,4)+qnew(i-1,j,k,4)+qnew(i,j+1,k,4)+qnew(i,j-1,k,4) &
           + qnew(i,j,k+1,4)+qnew(i,j,k-1,4)-6.0_r8*qnew(i,j,k,4)
        t0 = 0.5_r8*(qold(i,j,k,1)+1.0_r8)
        n = qnew(i,j,k,3)
        op(i,j,k,1) = p-cdt*l2-dt*lamp*t0*n
        op(i,j,k,2) = qnew(i,j,k,2)-(p*p*p)+ceps*l1
        op(i,j,k,3) = -ch*l3+kappa*t0*n
        op(i,j,k,4) = -ch*l4-lamp*t0*n
      ELSE
        op(i,j,k,1) = p-cdt*l2
        op(i,j,k,2) = qnew(i,j,k,2)-(p*p*p)+ceps*l1
      END IF
    END DO
  END DO
END DO
END FUNCTION Operator3D
!
FUNCTION Source2D(qnew,qold,aux,ll,mx,nrvars,maux,mbc,h,xlower) RESULT(s)
USE NodeInfoDef, ONLY: dt
USE ProblemDef
INTEGER, INTENT(IN):: ll
REAL(KIND=r8), DIMENSION(ll:,ll:,1:), INTENT(IN):: qnew, qold, aux
INTEGER, DIMENSION(1:2), INTENT(IN):: mx
INTEGER, INTENT(IN):: nrvars, maux, mbc
REAL(KIND=r8), INTENT(IN):: h
REAL(KIND=r8), DIMENSION(1:2), INTENT(IN):: xlower
REAL(KIND=r8), DIMENSION(1:mx(1),1:mx(2),1:nrvars):: s
IF(dumpsrc) THEN
  CALL DumpHeader(1,0,mx,h,xlower,2,INT(SIZE(qnew)+SIZE(qold),i8))
  WRITE(91) qnew, qold
END IF
IF(model==3) THEN
  s(:,:,1) = qold(1:mx(1),1:mx(2),1)-dt*lama*(0.5_r8*(qold(1:mx(1),1:mx(2),1)+1.0_r8))
  s(:,:,2) = -qold(1:mx(1),1:mx(2),1)
  s(:,:,3) = 0.0_r8
  s(:,:,4) = -lama*(0.5_r8*(qold(1:mx(1),1:mx(2),1)+1.0_r8))
ELSE
  s(:,:,1) = qold(1:mx(1),1:mx(2),1)
  s(:,:,2) = -qold(1:mx(1),1:mx(2),1)
END IF
END FUNCTION Source2D
!
FUNCTION Source3D(qnew,qold,aux,ll,mx,nrvars,maux,mbc,h,xlower) RESULT(s)
USE NodeInfoDef, ONLY: dt
USE ProblemDef
INTEGER, INTENT(IN):: ll
REAL(KIND=r8), DIMENSION(ll:,ll:,ll:,1:), INTENT(IN):: qnew, qold, aux
INTEGER, DIMENSION(1:3), INTENT(IN):: mx
INTEGER, INTENT(IN):: nrvars, maux, mbc
REAL(KIND=r8), INTENT(IN):: h
REAL(KIND=r8), DIMENSION(1:3), INTENT(IN):: xlower
REAL(KIND=r8), DIMENSION(1:mx(1),1:mx(2),1:mx(3),1:nrvars):: s
IF(dumpsrc) THEN
  CALL DumpHeader(1,0,mx,h,xlower,3,INT(SIZE(qnew)+SIZE(qold),i8))
  WRITE(91) qnew, qold
END IF
IF(model==3) THEN
  s(:,:,:,1) = qold(1:mx(1),1:mx(2),1:mx(3),1) &
             - dt*lama*(0.5_r8*(qold(1:mx(1),1:mx(2),1:mx(3),1)+1.0_r8))
  s(:,:,:,2) = -qold(1:mx(1),1:mx(2),1:mx(3),1)
  s(:,:,:,3) = 0.0_r8
  s(:,:,:,4) = -lama*(0.5_r8*(qold(1:mx(1),1:mx(2),1:mx(3),1)+1.0_r8))
ELSE
  s(:,:,:,1) = qold(1:mx(1),1:mx(2),1:mx(3),1)
  s(:,:,:,2) = -qold(1:mx(1),1:mx(2),1:mx(3),1)
END IF
END FUNCTION Source3D
!
FUNCTION SourceUpdate2D(qnew,qold,aux,ll,mx,nrvars,maux,mbc,h,xlower) RESULT(s)
INTEGER, INTENT(IN):: ll
REAL(KIND=r8), DIMENSION(ll:,ll:,1:), INTENT(IN):: qnew, qold, aux
INTEGER, DIMENSION(1:2), INTENT(IN):: mx
INTEGER, INTENT(IN):: nrvars, maux, mbc
REAL(KIND=r8), INTENT(IN):: h
REAL(KIND=r8), DIMENSION(1:2), INTENT(IN):: xlower
REAL(KIND=r8), DIMENSION(1:mx(1),1:mx(2),1:nrvars):: s
s = 0.0_r8
END FUNCTION SourceUpdate2D
!
FUNCTION SourceUpdate3D(qnew,qold,aux,ll,mx,nrvars,maux,mbc,h,xlower) RESULT(s)
INTEGER, INTENT(IN):: ll
REAL(KIND=r8), DIMENSION(ll:,ll:,ll:,1:), INTENT(IN):: qnew, qold, aux
INTEGER, DIMENSION(1:3), INTENT(IN):: mx
INTEGER, INTENT(IN):: nrvars, maux, mbc
REAL(KIND=r8), INTENT(IN):: h
REAL(KIND=r8), DIMENSION(1:3), INTENT(IN):: xlower
REAL(KIND=r8), DIMENSION(1:mx(1),1:mx(2),1:mx(3),1:nrvars):: s
s = 0.0_r8
END FUNCTION SourceUpdate3D
!
! Physical boundaries (mthbc=10): 2D reflection, 3D q_ghost = 2 - q_in.
SUBROUTINE UserBC2D(q,ll,mx,nrvars,mbc,bcnow)
USE ProblemDef, ONLY: model
INTEGER, INTENT(IN):: ll
REAL(KIND=r8), DIMENSION(ll:,ll:,1:), INTENT(IN OUT):: q
INTEGER, DIMENSION(1:2), INTENT(IN):: mx
INTEGER, INTENT(IN):: nrvars, mbc, bcnow
INTEGER:: b
DO b = 1, mbc
  SELECT CASE(bcnow)
  CASE(1); q(1-b,:,:) = q(b,:,:)
  CASE(2); q(mx(1)+b,:,:) = q(mx(1)-b+1,:,:)
  CASE(3); q(:,1-b,:) = q(:,b,:)
  CASE(4); q(:,mx(2)+b,:) = q(:,mx(2)-b+1,:)
  END SELECT
  IF(model==3) THEN
    ! the tumour system: phi, mu reflected; n = 1 and p = 0 on the face
    SELECT CASE(bcnow)
    CASE(1); q(1-b,:,3) = 2.0_r8-q(b,:,3); q(1-b,:,4) = -q(b,:,4)
    CASE(2); q(mx(1)+b,:,3) = 2.0_r8-q(mx(1)-b+1,:,3); q(mx(1)+b,:,4) = -q(mx(1)-b+1,:,4)
    CASE(3); q(:,1-b,3) = 2.0_r8-q(:,b,3); q(:,1-b,4) = -q(:,b,4)
    CASE(4); q(:,mx(2)+b,3) = 2.0_r8-q(:,mx(2)-b+1,3); q(:,mx(2)+b,4) = -q(:,mx(2)-b+1,4)
    END SELECT
  END IF
END DO
END SUBROUTINE UserBC2D
!
SUBROUTINE UserBC3D(q,ll,mx,nrvars,mbc,bcnow)
USE ProblemDef, ONLY: model
INTEGER, INTENT(IN):: ll
REAL(KIND=r8), DIMENSION(ll:,ll:,ll:,1:), INTENT(IN OUT):: q
INTEGER, DIMENSION(1:3), INTENT(IN):: mx
INTEGER, INTENT(IN):: nrvars, mbc, bcnow
INTEGER:: b
IF(model==3) THEN
  ! the tumour system: phi, mu reflected; n = 1 and p = 0 on the face
  DO b = 1, mbc
    SELECT CASE(bcnow)
    CASE(1)
      q(1-b,:,:,1:2) = q(b,:,:,1:2)
      q(1-b,:,:,3) = 2.0_r8-q(b,:,:,3); q(1-b,:,:,4) = -q(b,:,:,4)
    CASE(2)
      q(mx(1)+b,:,:,1:2) = q(mx(1)-b+1,:,:,1:2)
      q(mx(1)+b,:,:,3) = 2.0_r8-q(mx(1)-b+1,:,:,3); q(mx(1)+b,:,:,4) = -q(mx(1)-b+1,:,:,4)
    CASE(3)
      q(:,1-b,:,1:2) = q(:,b,:,1:2)
      q(:,1-b,:,3) = 2.0_r8-q(:,b,:,3); q(:,1-b,:,4) = -q(:,b,:,4)
    CASE(4)
      q(:,mx(2)+b,:,1:2) = q(:,mx(2)-b+1,:,1:2)
      q(:,mx(2)+b,:,3) = 2.0_r8-q(:,mx(2)-b+1,:,3); q(:,mx(2)+b,:,4) = -q(:,mx(2)-b+1,:,4)
    CASE(5)
      q(:,:,1-b,1:2) = q(:,:,b,1:2)
      q(:,:,1-b,3) = 2.0_r8-q(:,:,b,3); q(:,:,1-b,4) = -q(:,:,b,4)
    CASE(6)
      q(:,:,mx(3)+b,1:2) = q(:,:,mx(3)-b+1,1:2)
      q(:,:,mx(3)+b,3) = 2.0_r8-q(:,:,mx(3)-b+1,3); q(:,:,mx(3)+b,4) = -q(:,:,mx(3)-b+1,4)
    END SELECT
  END DO
  RETURN
END IF
DO b = 1, mbc
  SELECT CASE(bcnow)
  CASE(1); q(1-b,:,:,:) = 2.0_r8-q(b,:,:,:)
  CASE(2); q(mx(1)+b,:,:,:) = 2.0_r8-q(mx(1)-b+1,:,:,:)
  CASE(3); q(:,1-b,:,:) = 2.0_r8-q(:,b,:,:)
  CASE(4); q(:,mx(2)+b,:,:) = 2.0_r8-q(:,mx(2)-b+1,:,:)
  CASE(5); q(:,:,1-b,:) = 2.0_r8-q(:,:,b,:)
  CASE(6); q(:,:,mx(3)+b,:) = 2.0_r8-q(:,:,mx(3)-b+1,:)
  END SELECT
END DO
END SUBROUTINE UserBC3D
!
! ---- the tumour-growth system (model 3) ----------------------------------
! Pointwise Gauss-Seidel in the cell: the nutrient from its own equation,
! then the Cahn-Hilliard pair by the Newton step of RelaxGrid (with the new
! nutrient in the proliferation term), then the pressure.
SUBROUTINE TumourRelax2D(qnew,qold,f,mx,h,ipass)
USE NodeInfoDef
USE ProblemDef
REAL(KIND=r8), DIMENSION(0:,0:,1:), INTENT(IN OUT):: qnew
REAL(KIND=r8), DIMENSION(0:,0:,1:), INTENT(IN):: qold
REAL(KIND=r8), DIMENSION(1:,1:,1:), INTENT(IN):: f
INTEGER, DIMENSION(1:2), INTENT(IN):: mx
REAL(KIND=r8), INTENT(IN):: h
INTEGER, INTENT(IN):: ipass
INTEGER:: i, j
REAL(KIND=r8):: cdt, ceps, cdiag, c12, c21, r1, r2, p, d, ch, cd, t0, n
cdt = dt/(h*h); ceps = eps2/(h*h); cdiag = 4.0_r8*eps2/(h*h)
c12 = 4.0_r8*cdt; ch = 1.0_r8/(h*h); cd = 4.0_r8/(h*h)
DO j = 1, mx(2)
  DO i = 1+MODULO(j+ipass,2), mx(1), 2
    t0 = 0.5_r8*(qold(i,j,1)+1.0_r8)
    n = (f(i,j,3)+ch*(qnew(i-1,j,3)+qnew(i+1,j,3)+qnew(i,j-1,3)+qnew(i,j+1,3)))/(cd+kappa*t0)
    qnew(i,j,3) = n
    p = qnew(i,j,1)
    c21 = -3.0_r8*(p*p)-cdiag
    r1 = f(i,j,1)+cdt*(qnew(i-1,j,2)+qnew(i+1,j,2)+qnew(i,j-1,2)+qnew(i,j+1,2))+dt*lamp*t0*n
    r2 = f(i,j,2)-ceps*(qnew(i+1,j,1)+qnew(i-1,j,1)+qnew(i,j+1,1)+qnew(i,j-1,1)) &
       - 2.0_r8*(p*p*p)
    d = 1.0_r8-c12*c21
    qnew(i,j,1) = (r1-c12*r2)/d
    qnew(i,j,2) = (r2-c21*r1)/d
    qnew(i,j,4) = (f(i,j,4)+ch*(qnew(i-1,j,4)+qnew(i+1,j,4)+qnew(i,j-1,4)+qnew(i,j+1,4)) &
                + lamp*t0*n)/cd
  END DO
END DO
END SUBROUTINE TumourRelax2D
!
SUBROUTINE TumourRelax3D(qnew,qold,f,mx,h,ipass)
USE NodeInfoDef
USE ProblemDef
REAL(KIND=r8), DIMENSION(0:,0:,0:,1:), INTENT(IN OUT):: qnew
REAL(KIND=r8), DIMENSION(0:,0:,0:,1:), INTENT(IN):: qold
REAL(KIND=r8), DIMENSION(1:,1:,1:,1:), INTENT(IN):: f
INTEGER, DIMENSION(1:3), INTENT(IN):: mx
REAL(KIND=r8), INTENT(IN):: h
INTEGER, INTENT(IN):: ipass
INTEGER:: i, j, k
REAL(KIND=r8):: cdt, ceps, cdiag, c12, c21, r1, r2, p, d, ch, cd, t0, n
cdt = dt/(h*h); ceps = eps2/(h*h); cdiag = 6.0_r8*eps2/(h*h)
c12 = 6.0_r8*cdt; ch = 1.0_r8/(h*h); cd = 6.0_r8/(h*h)
DO k = 1, mx(3)
  DO j = 1, mx(2)
    DO i = 1+MODULO(k+j+ipass,2), mx(1), 2
      t0 = 0.5_r8*(qold(i,j,k,1)+1.0_r8)
      n = (f(i,j,k,3)+ch*(qnew(i-1,j,k,3)+qnew(i+1,j,k,3)+qnew(i,j-1,k,3) &
        + qnew(i,j+1,k,3)+qnew(i,j,k-1,3)+qnew(i,j,k+1,3)))/(cd+kappa*t0)
      qnew(i,j,k,3) = n
      p = qnew(i,j,k,1)
      c21 = -3.0_r8*(p*p)-cdiag
      r1 = f(i,j,k,1)+cdt*(qnew(i-1,j,k,2)+qnew(i+1,j,k,2)+qnew(i,j-1,k,2) &
         + qnew(i,j+1,k,2)+qnew(i,j,k-1,2)+qnew(i,j,k+1,2))+dt*lamp*t0*n
      r2 = f(i,j,k,2)-ceps*(qnew(i+1,j,k,1)+qnew(i-1,j,k,1)+qnew(i,j+1,k,1) &
         + qnew(i,j-1,k,1)+qnew(i,j,k+1,1)+qnew(i,j,k-1,1))-2.0_r8*(p*p*p)
      d = 1.0_r8-c12*c21
      qnew(i,j,k,1) = (r1-c12*r2)/d
      qnew(i,j,k,2) = (r2-c21*r1)/d
      qnew(i,j,k,4) = (f(i,j,k,4)+ch*(qnew(i-1,j,k,4)+qnew(i+1,j,k,4)+qnew(i,j-1,k,4) &
                    + qnew(i,j+1,k,4)+qnew(i,j,k-1,4)+qnew(i,j,k+1,4))+lamp*t0*n)/cd
    END DO
  END DO
END DO
END SUBROUTINE TumourRelax3D
!
END MODULE Problem
