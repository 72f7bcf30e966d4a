! Model constants and harness controls of the probe PROBLEM module used to
! run BSAM 1.2's own LIBRARY as the parity oracle (test infrastructure only).
MODULE ProblemDef
USE NodeInfoDef, ONLY: r8
IMPLICIT NONE
!
REAL(KIND=r8), PUBLIC:: alpha, eps, eps2
! Seeded initial data:
!   initmode 1: phi = base + amp*U(-1,1)
!   initmode 2: phi = tanh((radius-|x-xc|)/(sqrt(2)*eps)) + amp*U(-1,1)
!   initmode 3: phi = tanh((rs(theta,psi)-|x-xc|)/(sqrt(2)*eps)) + amp*U(-1,1),
!               rs = radius + pamp*sum_{l=1..lmax,|m|<=l} c_lm Y_lm(theta,psi)
!               (orthonormal real spherical harmonics about xc, polar axis z;
!               c_lm = U(-1,1) keyed by level -32, i = l, j = m+8, k = 1)
REAL(KIND=r8), PUBLIC:: base, amp, radius, hroot, tagtol, pamp
INTEGER, PUBLIC:: lmax
REAL(KIND=r8), DIMENSION(1:3), PUBLIC:: xc
INTEGER, PUBLIC:: initmode, seed
! Model: 1 Cahn-Hilliard (the hooks of PROBLEM/CH); 3 the tumour-growth
! system (phi, mu, nutrient n, pressure p; nrvars = 4): with the lagged
! tumour fraction t0 = (1+phi_old)/2 as BioGrowth's Growth/Uptake use it,
!   phi - dt Lap(mu) - dt lamp t0 n = phi_old - dt lama t0
!   mu - phi^3 + eps^2 Lap(phi)     = -phi_old
!   -Lap(n) + kappa t0 n            = 0,  n = 1 on physical boundaries
!   -Lap(p) - lamp t0 n             = -lama t0,  p = 0 on physical boundaries
! (proliferation lamp, apoptosis lama, nutrient consumption kappa).
INTEGER, PUBLIC:: model
REAL(KIND=r8), PUBLIC:: lamp, lama, kappa
! Dump controls: sweeps after relaxation (tag 2), state at source build
! (tag 1), state after each time step (tag 3).
INTEGER, PUBLIC:: dumpsweeps, ndumpedsweeps
LOGICAL, PUBLIC:: dumpsrc, dumpend, dumpopen
!
END MODULE ProblemDef
