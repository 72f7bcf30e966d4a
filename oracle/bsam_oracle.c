/*
 * bsam_oracle.c -- CPU restatement of BSAM 1.2's FAS multigrid V-cycle on a
 * block-structured AMR hierarchy.  TEST INFRASTRUCTURE ONLY: it is the
 * checker the parity tests, smoke() and bench.py's cpu_baseline use; the
 * product never links it.
 *
 * Plain C99, FP64, single thread, arrays in the reference's own layout
 * (bsamstorage.f90:54-144).  Arithmetic is written in the reference's
 * evaluation order (Fortran left-to-right, no reassociation) so that, built
 * with -ffp-contract=off, it reproduces the reference bit for bit.
 *
 * Restated reference routines (file:line under LIBRARY/ or PROBLEM/):
 *   afasroutines.f90   MultigridIterations 48-110, ErrorAFAS 113-160,
 *                      L2Error 163-189, L2ComponentErrors 192-219,
 *                      FillDown 236-254, FillDownLevel 257-327,
 *                      GetSourceFunction 330-381, UpdateSourceFunction
 *                      438-492, AFASVcycle 495-529, LevelRelax 532-557,
 *                      RelaxPatch 560-607, RestrictSolution 610-652,
 *                      CoarseLoadingFunction 673-736, Residual 817-857,
 *                      CorrectFine 860-980
 *   boundary.f90       SetGhost 67-97, Interpolate2D 143-333, GhostOverlap
 *                      359-477, SelfPeriodicBC 480-577, TransferPeriodicBC
 *                      580-611, SetPhysicalBC2D/3D 642-802, PeriodicSetup
 *                      805-872, GetPeriodicOffsets 875-954,
 *                      GetCoarseGhostPoints 957-1015, GetFaceIndex
 *                      1018-1045, Interpolate3D 1048-1550
 *   gridutilities.f90  ULap2D 50-72, Restriction2D 103-126,
 *                      BiLinProlongationP1 156-195, ULap3D 372-399,
 *                      Restriction3D 437-466, TriLinProlongationP1 510-566
 *   CH/problem.f90     RelaxGrid2D 318-369, Operator2D 393-422, Source2D
 *                      425-450, SourceUpdate2D 453-476, RelaxGrid3D 554-614,
 *                      Operator3D 638-669, Source3D 672-695, UserBC2D
 *                      770-818, UserBC3D 821-881
 *   BioGrowth/problem.f90  RelaxGrid2D 255-296, Operator2D 320-353,
 *                      Source2D 356-379, UserBC2D 623-673, Growth/Uptake
 *                      743-766
 *   bsamroutines.f90   CopyQToQold 592-605, time step 673-758, InitFields
 *                      1787-1856, TransferValues 1859-1951
 *   gridutilities.f90  BiLinProlongationP1MC 242-290, TriLinProlongationP1MC
 *                      628-705
 */
#include "bsam_oracle.h"

#include <float.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXLEVELS 40
#define LEVOFF 20 /* level L is stored at index L+LEVOFF */

/* Interpolation constants, boundary.f90:54-63. */
static const double A1 = -1.0 / 5.0, A2 = 2.0 / 3.0, A3 = 8.0 / 15.0;
static const double B1 = 5.0 / 32.0, B2 = 15.0 / 16.0, B3 = -3.0 / 32.0;
static const double C1 = 9.0 / 16.0, C2 = 3.0 / 16.0, C3 = 3.0 / 16.0, C4 = 1.0 / 16.0;

typedef struct {
  bsam_patch_desc d;
  int mx[3];         /* mx (1 in unused dims) */
  int cmx[3];        /* qc interior extents (AllocFields, bsamstorage.f90:66-73) */
  int nq[3];         /* q extents including ghosts */
  int nc[3];         /* qc extents including ghosts */
  double *q, *qold, *qc, *f, *rf, *ftmp;
} opatch;

struct orc_ctx {
  bsam_params p;
  int ndims, nv;
  int npatch;
  opatch* pt;
  int minlevel, finestlevel;
  int lvl_n[MAXLEVELS];
  int* lvl_list[MAXLEVELS];
  int periodic;              /* periodicboundaryconditions */
  int nper;                  /* nperiodicoffsets */
  int peridx[13];            /* periodicoffsetindex */
  int rootmx[3];             /* mxmax(0,:) */
};

static __thread char g_err[512];
static int fail(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return BSAM_EINVAL;
}
const char* orc_last_error(void) { return g_err; }

/* ---- indexing (reference index -> offset) ------------------------------ */
/* q(i,j,k,v) with i in 0..mx+1 etc.; in 2D k must be 1. */
static inline size_t QO(const opatch* P, int i, int j, int k, int v) {
  int kk = (P->nq[2] == 1) ? 0 : k;
  return (((size_t)v * P->nq[2] + kk) * P->nq[1] + j) * P->nq[0] + i;
}
static inline size_t CO(const opatch* P, int i, int j, int k, int v) {
  int kk = (P->nc[2] == 1) ? 0 : k;
  return (((size_t)v * P->nc[2] + kk) * P->nc[1] + j) * P->nc[0] + i;
}
/* f(i,j,k,v) with i in 1..mx */
static inline size_t FO(const opatch* P, int i, int j, int k, int v) {
  int nz = P->mx[2];
  int kk = (P->nq[2] == 1) ? 0 : k - 1;
  return (((size_t)v * nz + kk) * P->mx[1] + (j - 1)) * P->mx[0] + (i - 1);
}
#define Q(P, i, j, k, v) (P)->q[QO(P, i, j, k, v)]
#define QOLD(P, i, j, k, v) (P)->qold[QO(P, i, j, k, v)]
#define QC(P, i, j, k, v) (P)->qc[CO(P, i, j, k, v)]
#define F(P, i, j, k, v) (P)->f[FO(P, i, j, k, v)]
#define RF(P, i, j, k, v) (P)->rf[FO(P, i, j, k, v)]
#define FT(P, i, j, k, v) (P)->ftmp[FO(P, i, j, k, v)]

static inline int lidx(int level) { return level + LEVOFF; }

/* ---- context ------------------------------------------------------------ */
orc_ctx* orc_create(const bsam_params* p) {
  orc_ctx* c = (orc_ctx*)calloc(1, sizeof *c);
  c->p = *p;
  c->ndims = p->ndims;
  c->nv = p->nrvars;
  return c;
}

static void free_patches(orc_ctx* c) {
  for (int n = 0; n < c->npatch; n++) {
    opatch* P = &c->pt[n];
    free(P->q); free(P->qold); free(P->qc); free(P->f); free(P->rf); free(P->ftmp);
  }
  free(c->pt);
  c->pt = NULL;
  for (int l = 0; l < MAXLEVELS; l++) { free(c->lvl_list[l]); c->lvl_list[l] = NULL; c->lvl_n[l] = 0; }
  c->npatch = 0;
}

void orc_destroy(orc_ctx* c) {
  if (!c) return;
  free_patches(c);
  free(c);
}

void orc_set_dt(orc_ctx* c, double dt) { c->p.dt = dt; }

/* PeriodicSetup, boundary.f90:805-872 (offset index table). */
static void periodic_setup(orc_ctx* c, const int* mthbc_root) {
  int m[6];
  for (int i = 0; i < 6; i++) m[i] = 10;
  for (int i = 0; i < 2 * c->ndims; i++) m[i] = mthbc_root[i];
  c->periodic = 1;
  static const int t13[13] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13};
  if (m[0] == 2) {
    if (m[2] == 2) {
      if (m[4] == 2) { c->nper = 13; memcpy(c->peridx, t13, sizeof t13); }
      else { c->nper = 4; int t[4] = {1, 2, 4, 5}; memcpy(c->peridx, t, sizeof t); }
    } else {
      if (m[4] == 2) { c->nper = 4; int t[4] = {1, 3, 8, 9}; memcpy(c->peridx, t, sizeof t); }
      else { c->nper = 1; c->peridx[0] = 1; }
    }
  } else {
    if (m[2] == 2) {
      if (m[4] == 2) { c->nper = 4; int t[4] = {2, 3, 6, 7}; memcpy(c->peridx, t, sizeof t); }
      else { c->nper = 1; c->peridx[0] = 2; }
    } else {
      if (m[4] == 2) { c->nper = 1; c->peridx[0] = 3; }
      else { c->periodic = 0; c->nper = 0; }
    }
  }
}

int32_t orc_set_hierarchy(orc_ctx* c, int32_t npatch, const bsam_patch_desc* d,
                          int32_t minlevel, int32_t finestlevel) {
  if (c->p.mbc != 1) return fail("oracle: only mbc=1 is supported");
  if (c->ndims != 2 && c->ndims != 3) return fail("oracle: ndims must be 2 or 3");
  if (minlevel < -LEVOFF || finestlevel >= MAXLEVELS - LEVOFF || minlevel > 0 || finestlevel < 0)
    return fail("oracle: level range out of bounds");
  free_patches(c);
  c->npatch = npatch;
  c->minlevel = minlevel;
  c->finestlevel = finestlevel;
  c->pt = (opatch*)calloc((size_t)npatch, sizeof(opatch));
  int root = -1;
  for (int n = 0; n < npatch; n++) {
    opatch* P = &c->pt[n];
    P->d = d[n];
    for (int a = 0; a < 3; a++) {
      P->mx[a] = (a < c->ndims) ? d[n].mx[a] : 1;
      /* AllocFields, bsamstorage.f90:66-73: odd mx -> cmx = 1 */
      P->cmx[a] = (a < c->ndims) ? ((P->mx[a] % 2 == 1) ? 1 : P->mx[a] / 2) : 1;
      P->nq[a] = (a < c->ndims) ? P->mx[a] + 2 : 1;
      P->nc[a] = (a < c->ndims) ? P->cmx[a] + 2 : 1;
    }
    size_t nq = (size_t)P->nq[0] * P->nq[1] * P->nq[2] * c->nv;
    size_t nc = (size_t)P->nc[0] * P->nc[1] * P->nc[2] * c->nv;
    size_t nf = (size_t)P->mx[0] * P->mx[1] * P->mx[2] * c->nv;
    P->q = calloc(nq, 8); P->qold = calloc(nq, 8); P->qc = calloc(nc, 8);
    P->f = calloc(nf, 8); P->rf = calloc(nf, 8); P->ftmp = calloc(nf, 8);
    int L = d[n].level;
    if (L < minlevel || L > finestlevel) return fail("oracle: patch %d level %d out of range", n, L);
    if (L == 0) root = n;
    if (L > minlevel && (d[n].parent < 0 || d[n].parent >= npatch))
      return fail("oracle: patch %d has no parent", n);
  }
  for (int n = 0; n < npatch; n++) {
    int l = lidx(c->pt[n].d.level);
    c->lvl_list[l] = realloc(c->lvl_list[l], sizeof(int) * (size_t)(c->lvl_n[l] + 1));
    c->lvl_list[l][c->lvl_n[l]++] = n;
  }
  if (root < 0) return fail("oracle: no root-level patch");
  for (int a = 0; a < 3; a++) c->rootmx[a] = c->pt[root].d.mx[a];
  periodic_setup(c, c->pt[root].d.mthbc);
  return BSAM_OK;
}

int64_t orc_field_size(orc_ctx* c, int32_t n, int32_t field) {
  if (n < 0 || n >= c->npatch) return -1;
  opatch* P = &c->pt[n];
  switch (field) {
    case BSAM_FIELD_Q: case BSAM_FIELD_QOLD:
      return (int64_t)P->nq[0] * P->nq[1] * P->nq[2] * c->nv;
    case BSAM_FIELD_QC:
      return (int64_t)P->nc[0] * P->nc[1] * P->nc[2] * c->nv;
    case BSAM_FIELD_F: case BSAM_FIELD_RF: case BSAM_FIELD_FTMP:
      return (int64_t)P->mx[0] * P->mx[1] * P->mx[2] * c->nv;
  }
  return -1;
}

double* orc_field_ptr(orc_ctx* c, int32_t n, int32_t field) {
  if (n < 0 || n >= c->npatch) return NULL;
  opatch* P = &c->pt[n];
  switch (field) {
    case BSAM_FIELD_Q: return P->q;
    case BSAM_FIELD_QOLD: return P->qold;
    case BSAM_FIELD_QC: return P->qc;
    case BSAM_FIELD_F: return P->f;
    case BSAM_FIELD_RF: return P->rf;
    case BSAM_FIELD_FTMP: return P->ftmp;
  }
  return NULL;
}

/* ---- model hooks --------------------------------------------------------- */

/* The tumour model's relaxation (the probe PROBLEM module's TumourRelax2D/
 * 3D, oracle/ref_problem/problem.f90): per cell of the colour, the nutrient
 * n from -ULap(n)/h^2 + kappa t0 n = f3, the CH pair by RelaxGrid's Newton
 * step with dt*lamp*t0*n added to its first right-hand side, then the
 * pressure from -ULap(p)/h^2 - lamp t0 n = f4; t0 = (1 + phi_old)/2. */
static void relax_tumour(orc_ctx* c, opatch* P, int ipass) {
  const int nd = c->ndims;
  const double h = P->d.dx, dt = c->p.dt, eps2 = c->p.model_params[0];
  const double lamp = c->p.model_params[1], kappa = c->p.model_params[3];
  const double h2 = h * h, dd = nd == 3 ? 6.0 : 4.0;
  const double cdt = dt / h2, ceps = eps2 / h2, cdiag = dd * eps2 / h2, c12 = dd * cdt;
  const double chh = 1.0 / h2, cd = dd / h2;
  const int mz = nd == 3 ? P->mx[2] : 1;
  for (int k = 1; k <= mz; k++)
    for (int j = 1; j <= P->mx[1]; j++)
      for (int i = 1 + ((nd == 3 ? k : 0) + j + ipass) % 2; i <= P->mx[0]; i += 2) {
        const int kk = nd == 3 ? k : 1;
#define NB(v, op) ((nd == 3) ? (Q(P, i - 1, j, kk, v) op Q(P, i + 1, j, kk, v) op Q(P, i, j - 1, kk, v) op \
                                Q(P, i, j + 1, kk, v) op Q(P, i, j, kk - 1, v) op Q(P, i, j, kk + 1, v)) \
                             : (Q(P, i - 1, j, kk, v) op Q(P, i + 1, j, kk, v) op Q(P, i, j - 1, kk, v) op \
                                Q(P, i, j + 1, kk, v)))
#define NBP(v) ((nd == 3) ? (Q(P, i + 1, j, kk, v) + Q(P, i - 1, j, kk, v) + Q(P, i, j + 1, kk, v) + \
                             Q(P, i, j - 1, kk, v) + Q(P, i, j, kk + 1, v) + Q(P, i, j, kk - 1, v)) \
                          : (Q(P, i + 1, j, kk, v) + Q(P, i - 1, j, kk, v) + Q(P, i, j + 1, kk, v) + \
                             Q(P, i, j - 1, kk, v)))
        const double t0 = 0.5 * (QOLD(P, i, j, kk, 0) + 1.0);
        const double n = (F(P, i, j, kk, 2) + chh * NB(2, +)) / (cd + kappa * t0);
        Q(P, i, j, kk, 2) = n;
        const double p = Q(P, i, j, kk, 0);
        const double c21 = -3.0 * (p * p) - cdiag;
        const double r1 = F(P, i, j, kk, 0) + cdt * NB(1, +) + dt * lamp * t0 * n;
        const double r2 = F(P, i, j, kk, 1) - ceps * NBP(0) - 2.0 * (p * p * p);
        const double d = 1.0 - c12 * c21;
        Q(P, i, j, kk, 0) = (r1 - c12 * r2) / d;
        Q(P, i, j, kk, 1) = (r2 - c21 * r1) / d;
        Q(P, i, j, kk, 3) = (F(P, i, j, kk, 3) + chh * NB(3, +) + lamp * t0 * n) / cd;
#undef NB
#undef NBP
      }
}

/* CH RelaxGrid2D (CH/problem.f90:318-369) and BG RelaxGrid2D
 * (BioGrowth/problem.f90:255-296). */
static void relax_patch_2d(orc_ctx* c, opatch* P, int ipass) {
  const double h = P->d.dx, dt = c->p.dt, eps2 = c->p.model_params[0];
  const double h2 = h * h;
  const int mx = P->mx[0], my = P->mx[1];
  if (c->p.model == BSAM_MODEL_CH) {
    const double tmp1 = dt / h2, tmp2 = eps2 / h2, tmp3 = 4.0 * eps2 / h2;
    const double a12 = 4.0 * tmp1;
    for (int j = 1; j <= my; j++)
      for (int i = 1 + (j + ipass) % 2; i <= mx; i += 2) {
        double phi = Q(P, i, j, 1, 0);
        double a21 = -3.0 * (phi * phi) - tmp3;
        double b1 = F(P, i, j, 1, 0) + tmp1 * (Q(P, i - 1, j, 1, 1) + Q(P, i + 1, j, 1, 1) +
                                                Q(P, i, j - 1, 1, 1) + Q(P, i, j + 1, 1, 1));
        double b2 = F(P, i, j, 1, 1) - tmp2 * (Q(P, i + 1, j, 1, 0) + Q(P, i - 1, j, 1, 0) +
                                                Q(P, i, j + 1, 1, 0) + Q(P, i, j - 1, 1, 0)) -
                    2.0 * (phi * phi * phi);
        double det = 1.0 - a12 * a21;
        Q(P, i, j, 1, 0) = (b1 - a12 * b2) / det;
        Q(P, i, j, 1, 1) = (b2 - a21 * b1) / det;
      }
  } else if (c->p.model == BSAM_MODEL_TUMOUR) {
    relax_tumour(c, P, ipass);
  } else { /* BSAM_MODEL_BG */
    const double gam = c->p.model_params[1], lam = c->p.model_params[2], cz = c->p.model_params[3];
    const double tmp1 = dt * eps2 / h2, tmp3 = 1.0 / h2;
    for (int j = 1; j <= my; j++)
      for (int i = 1 + (j + ipass) % 2; i <= mx; i += 2) {
        double p1 = Q(P, i, j, 1, 0);
        double tmp2 = 4.0 * tmp1 + 1.0 + dt * (p1 * p1);
        double po = QOLD(P, i, j, 1, 0);
        double grow = gam * 0.5 * (po + 1.0);
        double upt = lam * 0.5 * (po + 1.0);
        Q(P, i, j, 1, 0) = (F(P, i, j, 1, 0) + dt * grow * (Q(P, i, j, 1, 1) - cz) +
                            tmp1 * (Q(P, i - 1, j, 1, 0) + Q(P, i + 1, j, 1, 0) +
                                    Q(P, i, j - 1, 1, 0) + Q(P, i, j + 1, 1, 0))) / tmp2;
        double tmp4 = 4.0 * tmp3 + upt;
        Q(P, i, j, 1, 1) = (F(P, i, j, 1, 1) + tmp3 * (Q(P, i + 1, j, 1, 1) + Q(P, i - 1, j, 1, 1) +
                                                        Q(P, i, j + 1, 1, 1) + Q(P, i, j - 1, 1, 1))) / tmp4;
      }
  }
}

/* CH RelaxGrid3D (CH/problem.f90:554-614). */
static void relax_patch_3d(orc_ctx* c, opatch* P, int ipass) {
  if (c->p.model == BSAM_MODEL_TUMOUR) {
    relax_tumour(c, P, ipass);
    return;
  }
  const double h = P->d.dx, dt = c->p.dt, eps2 = c->p.model_params[0];
  const double h2 = h * h;
  const double tmp1 = dt / h2, tmp2 = eps2 / h2, tmp3 = 6.0 * eps2 / h2;
  const double a12 = 6.0 * tmp1;
  const int mx = P->mx[0], my = P->mx[1], mz = P->mx[2];
  for (int k = 1; k <= mz; k++)
    for (int j = 1; j <= my; j++)
      for (int i = 1 + (k + j + ipass) % 2; i <= mx; i += 2) {
        double phi = Q(P, i, j, k, 0);
        double a21 = -3.0 * (phi * phi) - tmp3;
        double b1 = F(P, i, j, k, 0) +
                    tmp1 * (Q(P, i - 1, j, k, 1) + Q(P, i + 1, j, k, 1) + Q(P, i, j - 1, k, 1) +
                            Q(P, i, j + 1, k, 1) + Q(P, i, j, k - 1, 1) + Q(P, i, j, k + 1, 1));
        double b2 = F(P, i, j, k, 1) -
                    tmp2 * (Q(P, i + 1, j, k, 0) + Q(P, i - 1, j, k, 0) + Q(P, i, j + 1, k, 0) +
                            Q(P, i, j - 1, k, 0) + Q(P, i, j, k + 1, 0) + Q(P, i, j, k - 1, 0)) -
                    2.0 * (phi * phi * phi);
        double det = 1.0 - a12 * a21;
        Q(P, i, j, k, 0) = (b1 - a12 * b2) / det;
        Q(P, i, j, k, 1) = (b2 - a21 * b1) / det;
      }
}

/* Operator2D/3D applied to an array `a` with ghost layer (q-shaped, extents
 * n[]), interior mx[], spacing h; qold is q-shaped or NULL (CH ignores it).
 * CH: CH/problem.f90:393-422, 638-669 with ULap2D/3D gridutilities.f90:50-72,
 * 372-399.  BG: BioGrowth/problem.f90:320-353. out(i,j,k,v) for interior. */
typedef struct { const double* a; const int* n; int o[3]; } qview;
static inline double QV(qview v, int i, int j, int k, int var) {
  i += v.o[0]; j += v.o[1]; k += v.o[2];
  int kk = (v.n[2] == 1) ? 0 : k;
  return v.a[(((size_t)var * v.n[2] + kk) * v.n[1] + j) * v.n[0] + i];
}
static inline double ulap(orc_ctx* c, qview a, int i, int j, int k, int v) {
  if (c->ndims == 2)
    return QV(a, i + 1, j, k, v) + QV(a, i - 1, j, k, v) + QV(a, i, j + 1, k, v) +
           QV(a, i, j - 1, k, v) - 4.0 * QV(a, i, j, k, v);
  return QV(a, i + 1, j, k, v) + QV(a, i - 1, j, k, v) + QV(a, i, j + 1, k, v) +
         QV(a, i, j - 1, k, v) + QV(a, i, j, k + 1, v) + QV(a, i, j, k - 1, v) -
         6.0 * QV(a, i, j, k, v);
}
static inline void op_cell(orc_ctx* c, qview q, qview qo, double h, int i, int j, int k, double* o) {
  double* o1 = &o[0];
  double* o2 = &o[1];
  const double dt = c->p.dt, eps2 = c->p.model_params[0];
  const double h2 = h * h;
  if (c->p.model == BSAM_MODEL_CH) {
    const double tmp1 = dt / h2, tmp2 = eps2 / h2;
    double p = QV(q, i, j, k, 0);
    *o1 = QV(q, i, j, k, 0) - tmp1 * ulap(c, q, i, j, k, 1);
    *o2 = QV(q, i, j, k, 1) - (p * p * p) + tmp2 * ulap(c, q, i, j, k, 0);
  } else if (c->p.model == BSAM_MODEL_TUMOUR) {
    /* the probe module's Operator2D/3D with model = 3 */
    const double lamp = c->p.model_params[1], kappa = c->p.model_params[3];
    const double cdt = dt / h2, ceps = eps2 / h2, chh = 1.0 / h2;
    const double l2 = ulap(c, q, i, j, k, 1), l1 = ulap(c, q, i, j, k, 0);
    const double l3 = ulap(c, q, i, j, k, 2), l4 = ulap(c, q, i, j, k, 3);
    const double p = QV(q, i, j, k, 0), n = QV(q, i, j, k, 2);
    const double t0 = 0.5 * (QV(qo, i, j, k, 0) + 1.0);
    o[0] = p - cdt * l2 - dt * lamp * t0 * n;
    o[1] = QV(q, i, j, k, 1) - (p * p * p) + ceps * l1;
    o[2] = -chh * l3 + kappa * t0 * n;
    o[3] = -chh * l4 - lamp * t0 * n;
  } else {
    const double gam = c->p.model_params[1], lam = c->p.model_params[2], cz = c->p.model_params[3];
    const double tmp1 = dt * eps2 / h2, tmp3 = 1.0 / h2;
    double p = QV(q, i, j, k, 0);
    double po = QV(qo, i, j, k, 0);
    double grow = gam * 0.5 * (po + 1.0), upt = lam * 0.5 * (po + 1.0);
    *o1 = QV(q, i, j, k, 0) - tmp1 * ulap(c, q, i, j, k, 0) + dt * (p * p * p) -
          dt * grow * (QV(q, i, j, k, 1) - cz);
    *o2 = -tmp3 * ulap(c, q, i, j, k, 1) + upt * QV(q, i, j, k, 1);
  }
}

/* Variable v of Operator at fine cell (i,j,k) (the restriction of an
 * operator array reads it cell by cell). */
static inline double op_var(orc_ctx* c, qview q, qview qo, double h, int i, int j, int k, int v) {
  double o[4];
  op_cell(c, q, qo, h, i, j, k, o);
  return o[v];
}

/* Residual (afasroutines.f90:817-857): rf = f - Operator(q). */
static void residual_patch(orc_ctx* c, opatch* P) {
  qview q = {P->q, P->nq, {0, 0, 0}}, qo = {P->qold, P->nq, {0, 0, 0}};
  for (int k = 1; k <= P->mx[2]; k++)
    for (int j = 1; j <= P->mx[1]; j++)
      for (int i = 1; i <= P->mx[0]; i++) {
        double o[4];
        op_cell(c, q, qo, P->d.dx, i, j, k, o);
        for (int v = 0; v < c->nv; v++) RF(P, i, j, k, v) = F(P, i, j, k, v) - o[v];
      }
}

/* Restriction2D/3D (gridutilities.f90:103-126, 437-466) of one coarse cell of
 * an interior-indexed accessor; `get(i,j,k)` returns the fine value. */
#define RESTRICT_CELL(ndims, get, ic, jc, kc, out)                                          \
  do {                                                                                      \
    int i2_ = 2 * (ic), i1_ = 2 * (ic) - 1, j2_ = 2 * (jc), j1_ = 2 * (jc) - 1;             \
    if ((ndims) == 2) {                                                                     \
      (out) = 0.25 * (get(i2_, j2_, 1) + get(i1_, j2_, 1) + get(i2_, j1_, 1) + get(i1_, j1_, 1)); \
    } else {                                                                                \
      int k2_ = 2 * (kc), k1_ = 2 * (kc) - 1;                                               \
      (out) = 0.125 * (get(i2_, j2_, k2_) + get(i1_, j2_, k2_) + get(i2_, j1_, k2_) +       \
                       get(i1_, j1_, k2_) + get(i2_, j2_, k1_) + get(i1_, j2_, k1_) +       \
                       get(i2_, j1_, k1_) + get(i1_, j1_, k1_));                            \
    }                                                                                       \
  } while (0)

static inline int kc_hi(orc_ctx* c, const opatch* P) { return c->ndims == 3 ? P->cmx[2] : 1; }

/* RestrictSolution (afasroutines.f90:610-652). */
static void restrict_solution_patch(orc_ctx* c, opatch* P) {
  opatch* Par = &c->pt[P->d.parent];
  const int (*mb)[2] = P->d.mbounds;
  for (int v = 0; v < c->nv; v++)
    for (int kc = 1; kc <= kc_hi(c, P); kc++)
      for (int jc = 1; jc <= P->cmx[1]; jc++)
        for (int ic = 1; ic <= P->cmx[0]; ic++) {
          double r;
#define GQ(i, j, k) Q(P, i, j, k, v)
          RESTRICT_CELL(c->ndims, GQ, ic, jc, kc, r);
#undef GQ
          QC(P, ic, jc, kc, v) = r;
          int kp = c->ndims == 3 ? mb[2][0] + kc - 1 : 1;
          Q(Par, mb[0][0] + ic - 1, mb[1][0] + jc - 1, kp, v) = r;
        }
}

/* FillDownLevel (afasroutines.f90:257-327). */
static void filldown_patch(orc_ctx* c, opatch* P, int field) {
  opatch* Par = &c->pt[P->d.parent];
  const int (*mb)[2] = P->d.mbounds;
  for (int v = 0; v < c->nv; v++)
    for (int kc = 1; kc <= kc_hi(c, P); kc++)
      for (int jc = 1; jc <= P->cmx[1]; jc++)
        for (int ic = 1; ic <= P->cmx[0]; ic++) {
          double r;
          int ip = mb[0][0] + ic - 1, jp = mb[1][0] + jc - 1;
          int kp = c->ndims == 3 ? mb[2][0] + kc - 1 : 1;
          if (field == BSAM_SOURCEFIELD) {
#define GF(i, j, k) F(P, i, j, k, v)
            RESTRICT_CELL(c->ndims, GF, ic, jc, kc, r);
#undef GF
            F(Par, ip, jp, kp, v) = r;
          } else {
#define GQ(i, j, k) Q(P, i, j, k, v)
            RESTRICT_CELL(c->ndims, GQ, ic, jc, kc, r);
#undef GQ
            Q(Par, ip, jp, kp, v) = r;
          }
        }
}

/* CoarseLoadingFunction (afasroutines.f90:673-736): parent%f(mb) =
 * Restriction(rf) + Operator(qc(0:cmx+1), parent%qold(mb-1:mb+1), ch). */
static void coarse_loading_patch(orc_ctx* c, opatch* P) {
  opatch* Par = &c->pt[P->d.parent];
  const int (*mb)[2] = P->d.mbounds;
  residual_patch(c, P);
  qview qc = {P->qc, P->nc, {0, 0, 0}};
  qview qo = {Par->qold, Par->nq, {mb[0][0] - 1, mb[1][0] - 1, c->ndims == 3 ? mb[2][0] - 1 : 0}};
  const double ch = Par->d.dx;
  for (int kc = 1; kc <= kc_hi(c, P); kc++)
    for (int jc = 1; jc <= P->cmx[1]; jc++)
      for (int ic = 1; ic <= P->cmx[0]; ic++) {
        double o[4];
        op_cell(c, qc, qo, ch, ic, jc, kc, o);
        int ip = mb[0][0] + ic - 1, jp = mb[1][0] + jc - 1;
        int kp = c->ndims == 3 ? mb[2][0] + kc - 1 : 1;
        for (int v = 0; v < c->nv; v++) {
          double r;
#define GV(i, j, k) RF(P, i, j, k, v)
          RESTRICT_CELL(c->ndims, GV, ic, jc, kc, r);
#undef GV
          F(Par, ip, jp, kp, v) = r + o[v];
        }
      }
}

/* GetCoarseGhostPoints (boundary.f90:957-1015). */
static void coarse_ghost_patch(orc_ctx* c, opatch* P) {
  opatch* Par = &c->pt[P->d.parent];
  const int (*mb)[2] = P->d.mbounds;
  const int cx = P->cmx[0], cy = P->cmx[1], cz = P->cmx[2];
  for (int v = 0; v < c->nv; v++) {
    if (c->ndims == 2) {
      for (int jc = 0; jc <= cy + 1; jc++) {
        QC(P, 0, jc, 1, v) = Q(Par, mb[0][0] - 1, mb[1][0] - 1 + jc, 1, v);
        QC(P, cx + 1, jc, 1, v) = Q(Par, mb[0][1] + 1, mb[1][0] - 1 + jc, 1, v);
      }
      for (int ic = 1; ic <= cx; ic++) {
        QC(P, ic, 0, 1, v) = Q(Par, mb[0][0] + ic - 1, mb[1][0] - 1, 1, v);
        QC(P, ic, cy + 1, 1, v) = Q(Par, mb[0][0] + ic - 1, mb[1][1] + 1, 1, v);
      }
    } else {
      for (int kc = 0; kc <= cz + 1; kc++)
        for (int jc = 0; jc <= cy + 1; jc++) {
          QC(P, 0, jc, kc, v) = Q(Par, mb[0][0] - 1, mb[1][0] - 1 + jc, mb[2][0] - 1 + kc, v);
          QC(P, cx + 1, jc, kc, v) = Q(Par, mb[0][1] + 1, mb[1][0] - 1 + jc, mb[2][0] - 1 + kc, v);
        }
      for (int kc = 0; kc <= cz + 1; kc++)
        for (int ic = 1; ic <= cx; ic++) {
          QC(P, ic, 0, kc, v) = Q(Par, mb[0][0] + ic - 1, mb[1][0] - 1, mb[2][0] - 1 + kc, v);
          QC(P, ic, cy + 1, kc, v) = Q(Par, mb[0][0] + ic - 1, mb[1][1] + 1, mb[2][0] - 1 + kc, v);
        }
      for (int jc = 1; jc <= cy; jc++)
        for (int ic = 1; ic <= cx; ic++) {
          QC(P, ic, jc, 0, v) = Q(Par, mb[0][0] + ic - 1, mb[1][0] + jc - 1, mb[2][0] - 1, v);
          QC(P, ic, jc, cz + 1, v) = Q(Par, mb[0][0] + ic - 1, mb[1][0] + jc - 1, mb[2][1] + 1, v);
        }
    }
  }
}

/* CorrectFine with BiLinProlongationP1 / TriLinProlongationP1
 * (afasroutines.f90:860-980, gridutilities.f90:156-195, 510-566). */
static inline double lin_w(double lo, double hi, int fine) {
  /* even fine index: 3*a(lo)+a(lo+1); odd: a(lo)+3*a(lo+1) */
  return (fine % 2 == 0) ? 3.0 * lo + hi : lo + 3.0 * hi;
}
static void correct_fine_patch(orc_ctx* c, opatch* P) {
  opatch* Par = &c->pt[P->d.parent];
  const int (*mb)[2] = P->d.mbounds;
  const int cx = P->cmx[0], cy = P->cmx[1], cz = P->cmx[2];
  for (int v = 0; v < c->nv; v++) {
    if (c->ndims == 2) {
      for (int jc = 0; jc <= cy + 1; jc++)
        for (int ic = 0; ic <= cx + 1; ic++)
          QC(P, ic, jc, 1, v) = Q(Par, mb[0][0] - 1 + ic, mb[1][0] - 1 + jc, 1, v) - QC(P, ic, jc, 1, v);
      for (int j = 0; j <= P->mx[1] + 1; j++)
        for (int i = 0; i <= P->mx[0] + 1; i++) {
          int il = i / 2, jl = j / 2;
          double b0 = lin_w(QC(P, il, jl, 1, v), QC(P, il + 1, jl, 1, v), i);
          double b1 = lin_w(QC(P, il, jl + 1, 1, v), QC(P, il + 1, jl + 1, 1, v), i);
          double pr = lin_w(b0, b1, j) / 16.0;
          Q(P, i, j, 1, v) = Q(P, i, j, 1, v) + pr;
        }
    } else {
      for (int kc = 0; kc <= cz + 1; kc++)
        for (int jc = 0; jc <= cy + 1; jc++)
          for (int ic = 0; ic <= cx + 1; ic++)
            QC(P, ic, jc, kc, v) =
                Q(Par, mb[0][0] - 1 + ic, mb[1][0] - 1 + jc, mb[2][0] - 1 + kc, v) - QC(P, ic, jc, kc, v);
      for (int k = 0; k <= P->mx[2] + 1; k++)
        for (int j = 0; j <= P->mx[1] + 1; j++)
          for (int i = 0; i <= P->mx[0] + 1; i++) {
            int il = i / 2, jl = j / 2, kl = k / 2;
            double cc[2];
            for (int dk = 0; dk < 2; dk++) {
              double b0 = lin_w(QC(P, il, jl, kl + dk, v), QC(P, il + 1, jl, kl + dk, v), i);
              double b1 = lin_w(QC(P, il, jl + 1, kl + dk, v), QC(P, il + 1, jl + 1, kl + dk, v), i);
              cc[dk] = lin_w(b0, b1, j);
            }
            double pr = lin_w(cc[0], cc[1], k) / 64.0;
            Q(P, i, j, k, v) = Q(P, i, j, k, v) + pr;
          }
    }
  }
}

/* Source#D + SourceUpdate#D (CH/problem.f90:425-476, 672-721;
 * BioGrowth/problem.f90:356-405) via GetSourceFunction/UpdateSourceFunction
 * (afasroutines.f90:330-381, 438-492). */
static void get_source_patch(orc_ctx* c, opatch* P) {
  for (int k = 1; k <= P->mx[2]; k++)
    for (int j = 1; j <= P->mx[1]; j++)
      for (int i = 1; i <= P->mx[0]; i++) {
        double po = QOLD(P, i, j, k, 0);
        if (c->p.model == BSAM_MODEL_CH) {
          FT(P, i, j, k, 0) = po;
          FT(P, i, j, k, 1) = -po;
        } else if (c->p.model == BSAM_MODEL_TUMOUR) {
          /* the probe module's Source2D/3D with model = 3 */
          const double lama = c->p.model_params[2];
          FT(P, i, j, k, 0) = po - c->p.dt * lama * (0.5 * (po + 1.0));
          FT(P, i, j, k, 1) = -po;
          FT(P, i, j, k, 2) = 0.0;
          FT(P, i, j, k, 3) = -lama * (0.5 * (po + 1.0));
        } else {
          FT(P, i, j, k, 0) = (1.0 + c->p.dt) * po;
          FT(P, i, j, k, 1) = 0.0;
        }
      }
}
static void update_source_patch(orc_ctx* c, opatch* P) {
  for (int v = 0; v < c->nv; v++)
    for (int k = 1; k <= P->mx[2]; k++)
      for (int j = 1; j <= P->mx[1]; j++)
        for (int i = 1; i <= P->mx[0]; i++) F(P, i, j, k, v) = FT(P, i, j, k, v) + 0.0;
}

/* ---- ghost filling ------------------------------------------------------- */

/* Interpolate2D (boundary.f90:143-333). qc is the parent's q. */
static void interpolate_2d(orc_ctx* c, opatch* P, int ipass) {
  opatch* Par = &c->pt[P->d.parent];
  const int mx = P->mx[0], my = P->mx[1];
  const int (*mb)[2] = P->d.mbounds;
  const int* mthbc = P->d.mthbc;
#define QF(i, j) Q(P, i, j, 1, v)
#define PC(i, j) Q(Par, i, j, 1, v)
  for (int v = 0; v < c->nv; v++) {
    if (ipass == 1 || ipass == 0) {
      QF(mx + 1, my + 1) = A1 * QF(mx - 1, my - 1) + A2 * QF(mx, my) + A3 * PC(mb[0][1] + 1, mb[1][1] + 1);
      QF(0, 0) = A1 * QF(2, 2) + A2 * QF(1, 1) + A3 * PC(mb[0][0] - 1, mb[1][0] - 1);
    }
    if (ipass == 2 || ipass == 0) {
      QF(mx + 1, 0) = A1 * QF(mx - 1, 2) + A2 * QF(mx, 1) + A3 * PC(mb[0][1] + 1, mb[1][0] - 1);
      QF(0, my + 1) = A1 * QF(2, my - 1) + A2 * QF(1, my) + A3 * PC(mb[0][0] - 1, mb[1][1] + 1);
    }
    if (mthbc[0] == 2 || mthbc[0] == 999) {
      int ic = mb[0][0] - 1;
      if (ipass == 1 || ipass == 0) {
        for (int j = 2; j <= my - 2; j += 2)
          QF(0, j) = A1 * QF(2, j + 2) + A2 * QF(1, j + 1) + A3 * PC(ic, mb[1][0] + j / 2 - 1);
        QF(0, my) = A1 * QF(2, my) + A2 * QF(1, my) +
                    A3 * (B1 * PC(ic, mb[1][1] + 1) + B2 * PC(ic, mb[1][1]) + B3 * PC(ic, mb[1][1] - 1));
      }
      if (ipass == 2 || ipass == 0) {
        for (int j = 3; j <= my - 1; j += 2)
          QF(0, j) = A1 * QF(2, j - 2) + A2 * QF(1, j - 1) + A3 * PC(ic, mb[1][0] + (j - 1) / 2);
        QF(0, 1) = A1 * QF(2, 1) + A2 * QF(1, 1) +
                   A3 * (B1 * PC(ic, mb[1][0] - 1) + B2 * PC(ic, mb[1][0]) + B3 * PC(ic, mb[1][0] + 1));
      }
    }
    if (mthbc[1] == 2 || mthbc[1] == 999) {
      int ic = mb[0][1] + 1;
      if (ipass == 1 || ipass == 0) {
        for (int j = 3; j <= my - 1; j += 2)
          QF(mx + 1, j) = A1 * QF(mx - 1, j - 2) + A2 * QF(mx, j - 1) + A3 * PC(ic, mb[1][0] + (j - 1) / 2);
        QF(mx + 1, 1) = A1 * QF(mx - 1, 1) + A2 * QF(mx, 1) +
                        A3 * (B1 * PC(ic, mb[1][0] - 1) + B2 * PC(ic, mb[1][0]) + B3 * PC(ic, mb[1][0] + 1));
      }
      if (ipass == 2 || ipass == 0) {
        for (int j = 2; j <= my - 2; j += 2)
          QF(mx + 1, j) = A1 * QF(mx - 1, j + 2) + A2 * QF(mx, j + 1) + A3 * PC(ic, mb[1][0] + j / 2 - 1);
        QF(mx + 1, my) = A1 * QF(mx - 1, my) + A2 * QF(mx, my) +
                         A3 * (B1 * PC(ic, mb[1][1] + 1) + B2 * PC(ic, mb[1][1]) + B3 * PC(ic, mb[1][1] - 1));
      }
    }
    if (mthbc[2] == 2 || mthbc[2] == 999) {
      int jc = mb[1][0] - 1;
      if (ipass == 1 || ipass == 0) {
        for (int i = 2; i <= mx - 2; i += 2)
          QF(i, 0) = A1 * QF(i + 2, 2) + A2 * QF(i + 1, 1) + A3 * PC(mb[0][0] + i / 2 - 1, jc);
        QF(mx, 0) = A1 * QF(mx, 2) + A2 * QF(mx, 1) +
                    A3 * (B1 * PC(mb[0][1] + 1, jc) + B2 * PC(mb[0][1], jc) + B3 * PC(mb[0][1] - 1, jc));
      }
      if (ipass == 2 || ipass == 0) {
        for (int i = 3; i <= mx - 1; i += 2)
          QF(i, 0) = A1 * QF(i - 2, 2) + A2 * QF(i - 1, 1) + A3 * PC(mb[0][0] + (i - 1) / 2, jc);
        QF(1, 0) = A1 * QF(1, 2) + A2 * QF(1, 1) +
                   A3 * (B1 * PC(mb[0][0] - 1, jc) + B2 * PC(mb[0][0], jc) + B3 * PC(mb[0][0] + 1, jc));
      }
    }
    if (mthbc[3] == 2 || mthbc[3] == 999) {
      int jc = mb[1][1] + 1;
      if (ipass == 1 || ipass == 0) {
        for (int i = 3; i <= mx - 1; i += 2)
          QF(i, my + 1) = A1 * QF(i - 2, my - 1) + A2 * QF(i - 1, my) + A3 * PC(mb[0][0] + (i - 1) / 2, jc);
        QF(1, my + 1) = A1 * QF(1, my - 1) + A2 * QF(1, my) +
                        A3 * (B1 * PC(mb[0][0] - 1, jc) + B2 * PC(mb[0][0], jc) + B3 * PC(mb[0][0] + 1, jc));
      }
      if (ipass == 2 || ipass == 0) {
        for (int i = 2; i <= mx - 2; i += 2)
          QF(i, my + 1) = A1 * QF(i + 2, my - 1) + A2 * QF(i + 1, my) + A3 * PC(mb[0][0] + i / 2 - 1, jc);
        QF(mx, my + 1) = A1 * QF(mx, my - 1) + A2 * QF(mx, my) +
                         A3 * (B1 * PC(mb[0][1] + 1, jc) + B2 * PC(mb[0][1], jc) + B3 * PC(mb[0][1] - 1, jc));
      }
    }
  }
#undef QF
#undef PC
}

/* Face-index table of GetFaceIndex (boundary.f90:1018-1045), 0-based slot:
 * fi[0]=ghost, fi[1]=2nd interior, fi[2]=1st interior, fi[3]=coarse outside,
 * fi[4]=coarse inside. */
static void face_index(const opatch* P, int dim, int parity, int fi[5]) {
  const int (*mb)[2] = P->d.mbounds;
  if (parity == 1) {
    fi[0] = 0; fi[1] = 2; fi[2] = 1; fi[3] = mb[dim][0] - 1; fi[4] = mb[dim][0];
  } else {
    int m = P->mx[dim];
    fi[0] = m + 1; fi[1] = m - 1; fi[2] = m; fi[3] = mb[dim][1] + 1; fi[4] = mb[dim][1];
  }
}

/* Transverse index kinds of Interpolate3D's face/edge statements:
 * E: fine 2..mx-2 step 2, fine sources at +2/+1, coarse mb1+pos/2-1;
 * O: fine 3..mx-1 step 2, sources at -2/-1, coarse mb1+(pos-1)/2;
 * L: fine 1, sources at 0, coarse quadratic over mb1-1,mb1,mb1+1 (b1,b2,b3);
 * H: fine mx, sources at 0, coarse quadratic over mb2+1,mb2,mb2-1. */
enum { KE = 0, KO = 1, KL = 2, KH = 3 };

typedef struct {
  int pos, s1, s2;       /* ghost transverse position, offsets of a1 / a2 sources */
  int quad;              /* 1 for L/H */
  int cc[3];             /* coarse index (E/O: cc[0]); L/H: the b1,b2,b3 cells */
} tpos;

static int tkind_positions(const opatch* P, int dim, int kind, tpos* out) {
  const int m = P->mx[dim];
  const int mb1 = P->d.mbounds[dim][0], mb2 = P->d.mbounds[dim][1];
  int n = 0;
  switch (kind) {
    case KE:
      for (int p = 2; p <= m - 2; p += 2) { tpos t = {p, 2, 1, 0, {mb1 + p / 2 - 1, 0, 0}}; out[n++] = t; }
      break;
    case KO:
      for (int p = 3; p <= m - 1; p += 2) { tpos t = {p, -2, -1, 0, {mb1 + (p - 1) / 2, 0, 0}}; out[n++] = t; }
      break;
    case KL: { tpos t = {1, 0, 0, 1, {mb1 - 1, mb1, mb1 + 1}}; out[n++] = t; } break;
    case KH: { tpos t = {m, 0, 0, 1, {mb2 + 1, mb2, mb2 - 1}}; out[n++] = t; } break;
  }
  return n;
}

static inline int skip_face(const int* mthbc, int face, int ipass) {
  return mthbc[face] != 2 && mthbc[face] != 999 && ipass == 1;
}

/* Interpolate3D (boundary.f90:1048-1550). */
static void interpolate_3d(orc_ctx* c, opatch* P, int ipass) {
  opatch* Par = &c->pt[P->d.parent];
  const int* mthbc = P->d.mthbc;
  int fi[3][5];
  int maxm = P->mx[0];
  if (P->mx[1] > maxm) maxm = P->mx[1];
  if (P->mx[2] > maxm) maxm = P->mx[2];
  tpos* tb1 = malloc(sizeof(tpos) * (size_t)(maxm + 2));
  tpos* tb2 = malloc(sizeof(tpos) * (size_t)(maxm + 2));
  /* cell accessor with a permutation: coordinate along dims (d0,d1,d2) */
#define SETIJK(arr, x0, x1, x2) do { int t_[3]; t_[d0] = (x0); t_[d1] = (x1); t_[d2] = (x2); \
    arr[0] = t_[0]; arr[1] = t_[1]; arr[2] = t_[2]; } while (0)
  for (int v = 0; v < c->nv; v++) {
    /* (1) corners, boundary.f90:1075-1093 */
    for (int p1 = 1; p1 <= 2; p1++) {
      if (skip_face(mthbc, p1 - 1, ipass)) continue;
      face_index(P, 0, p1, fi[0]);
      for (int p2 = 1; p2 <= 2; p2++) {
        if (skip_face(mthbc, p2 + 1, ipass)) continue;
        face_index(P, 1, p2, fi[1]);
        for (int p3 = 1; p3 <= 2; p3++) {
          if (skip_face(mthbc, p3 + 3, ipass)) continue;
          face_index(P, 2, p3, fi[2]);
          Q(P, fi[0][0], fi[1][0], fi[2][0], v) = A1 * Q(P, fi[0][1], fi[1][1], fi[2][1], v) +
                                                 A2 * Q(P, fi[0][2], fi[1][2], fi[2][2], v) +
                                                 A3 * Q(Par, fi[0][3], fi[1][3], fi[2][3], v);
        }
      }
    }
    /* (2) edges, boundary.f90:1095-1207.  Each edge runs along dimension
     * `ed`; the other two dims (fa < fb) are fixed at their face indices. */
    for (int e = 0; e < 3; e++) {
      /* order in the reference: along dim 2 (1,3 fixed), along dim 3 (1,2
       * fixed), along dim 1 (2,3 fixed) */
      int ed = (e == 0) ? 1 : (e == 1) ? 2 : 0;
      int fa = (ed == 0) ? 1 : 0;
      int fb = (ed == 2) ? 1 : 2;
      for (int pa = 1; pa <= 2; pa++) {
        if (skip_face(mthbc, 2 * fa + pa - 1, ipass)) continue;
        face_index(P, fa, pa, fi[fa]);
        for (int pb = 1; pb <= 2; pb++) {
          int face_b = 2 * fb + pb - 1;
          if (e == 0) {
            /* boundary.f90:1104 tests mthbc(parity_3+4) against 99, not 999 */
            if (mthbc[face_b] != 2 && mthbc[face_b] != 99 && ipass == 1) continue;
          } else if (skip_face(mthbc, face_b, ipass)) {
            continue;
          }
          face_index(P, fb, pb, fi[fb]);
          /* statement order in each block: O, L, E, H */
          static const int korder[4] = {KO, KL, KE, KH};
          for (int kk = 0; kk < 4; kk++) {
            int nt = tkind_positions(P, ed, korder[kk], tb1);
            for (int t = 0; t < nt; t++) {
              int g[3], s1[3], s2[3], cq[3];
              g[fa] = fi[fa][0]; g[fb] = fi[fb][0]; g[ed] = tb1[t].pos;
              s1[fa] = fi[fa][1]; s1[fb] = fi[fb][1]; s1[ed] = tb1[t].pos + tb1[t].s1;
              s2[fa] = fi[fa][2]; s2[fb] = fi[fb][2]; s2[ed] = tb1[t].pos + tb1[t].s2;
              cq[fa] = fi[fa][3]; cq[fb] = fi[fb][3];
              double coarse;
              if (!tb1[t].quad) {
                cq[ed] = tb1[t].cc[0];
                coarse = Q(Par, cq[0], cq[1], cq[2], v);
              } else {
                double w[3];
                for (int r = 0; r < 3; r++) { cq[ed] = tb1[t].cc[r]; w[r] = Q(Par, cq[0], cq[1], cq[2], v); }
                coarse = B1 * w[0] + B2 * w[1] + B3 * w[2];
              }
              Q(P, g[0], g[1], g[2], v) = A1 * Q(P, s1[0], s1[1], s1[2], v) +
                                         A2 * Q(P, s2[0], s2[1], s2[2], v) + A3 * coarse;
            }
          }
        }
      }
    }
    /* (3) faces, boundary.f90:1209-1548 */
    for (int d = 0; d < 3; d++) {
      int t1 = (d == 0) ? 1 : 0;
      int t2 = (d == 2) ? 1 : 2;
      for (int p = 1; p <= 2; p++) {
        int face = 2 * d + p - 1;
        if (!(mthbc[face] == 2 || mthbc[face] == 999)) continue;
        face_index(P, d, p, fi[d]);
        int do_even = (ipass == 0) || ((p + ipass) % 2 == 0); /* Yellow, Blue */
        int do_odd = (ipass == 0) || ((p + ipass) % 2 == 1);  /* Green, Red */
        for (int k1 = 0; k1 < 4; k1++)
          for (int k2 = 0; k2 < 4; k2++) {
            int q1 = (k1 >= KL), q2 = (k2 >= KL);
            if (q1 && q2) continue; /* grey cells handled below */
            /* transverse parity: E,H even positions; O,L odd */
            int par1 = (k1 == KO || k1 == KL), par2 = (k2 == KO || k2 == KL);
            int odd = (par1 + par2) % 2;
            if (odd ? !do_odd : !do_even) continue;
            int n1 = tkind_positions(P, t1, k1, tb1);
            int n2 = tkind_positions(P, t2, k2, tb2);
            for (int b = 0; b < n2; b++)
              for (int a = 0; a < n1; a++) {
                int g[3], s1[3], s2[3], cq[3];
                g[d] = fi[d][0]; s1[d] = fi[d][1]; s2[d] = fi[d][2]; cq[d] = fi[d][3];
                g[t1] = tb1[a].pos; s1[t1] = tb1[a].pos + tb1[a].s1; s2[t1] = tb1[a].pos + tb1[a].s2;
                g[t2] = tb2[b].pos; s1[t2] = tb2[b].pos + tb2[b].s1; s2[t2] = tb2[b].pos + tb2[b].s2;
                double coarse;
                if (!q1 && !q2) {
                  cq[t1] = tb1[a].cc[0]; cq[t2] = tb2[b].cc[0];
                  coarse = Q(Par, cq[0], cq[1], cq[2], v);
                } else {
                  double w[3];
                  for (int r = 0; r < 3; r++) {
                    cq[t1] = q1 ? tb1[a].cc[r] : tb1[a].cc[0];
                    cq[t2] = q2 ? tb2[b].cc[r] : tb2[b].cc[0];
                    w[r] = Q(Par, cq[0], cq[1], cq[2], v);
                  }
                  coarse = B1 * w[0] + B2 * w[1] + B3 * w[2];
                }
                Q(P, g[0], g[1], g[2], v) = A1 * Q(P, s1[0], s1[1], s1[2], v) +
                                           A2 * Q(P, s2[0], s2[1], s2[2], v) + A3 * coarse;
              }
          }
        /* corner grey cells (always refreshed) */
        int ft1[5], ft2[5];
        for (int pa = 1; pa <= 2; pa++) {
          face_index(P, t1, pa, ft1);
          for (int pb = 1; pb <= 2; pb++) {
            face_index(P, t2, pb, ft2);
            int g[3], s1[3], s2[3], cq[3];
            g[d] = fi[d][0]; s1[d] = fi[d][1]; s2[d] = fi[d][2]; cq[d] = fi[d][3];
            g[t1] = s1[t1] = s2[t1] = ft1[2];
            g[t2] = s1[t2] = s2[t2] = ft2[2];
            double w[4];
            cq[t1] = ft1[4]; cq[t2] = ft2[4]; w[0] = Q(Par, cq[0], cq[1], cq[2], v);
            cq[t1] = ft1[3]; cq[t2] = ft2[4]; w[1] = Q(Par, cq[0], cq[1], cq[2], v);
            cq[t1] = ft1[4]; cq[t2] = ft2[3]; w[2] = Q(Par, cq[0], cq[1], cq[2], v);
            cq[t1] = ft1[3]; cq[t2] = ft2[3]; w[3] = Q(Par, cq[0], cq[1], cq[2], v);
            Q(P, g[0], g[1], g[2], v) = A1 * Q(P, s1[0], s1[1], s1[2], v) +
                                       A2 * Q(P, s2[0], s2[1], s2[2], v) +
                                       A3 * (C1 * w[0] + C2 * w[1] + C3 * w[2] + C4 * w[3]);
          }
        }
      }
    }
  }
#undef SETIJK
  free(tb1);
  free(tb2);
}

/* GhostOverlap (boundary.f90:359-431); off = grid2offset or NULL. */
static void ghost_overlap(orc_ctx* c, opatch* G1, opatch* G2, const int* off) {
  const int nd = c->ndims, mbc = 1;
  int mg1[3][2], mg2[3][2], ovg[3][2], o1[3][2], o2[3][2];
  for (int n = 0; n < 3; n++) {
    for (int s = 0; s < 2; s++) {
      mg1[n][s] = (n < nd) ? G1->d.mglobal[n][s] : 1;
      mg2[n][s] = (n < nd) ? G2->d.mglobal[n][s] + (off ? off[n] : 0) : 1;
    }
    ovg[n][0] = 1; ovg[n][1] = 1;
  }
  for (int n = 0; n < nd; n++) {
    ovg[n][0] = (mg1[n][0] - mbc > mg2[n][0] - mbc) ? mg1[n][0] - mbc : mg2[n][0] - mbc;
    ovg[n][1] = (mg1[n][1] + mbc < mg2[n][1] + mbc) ? mg1[n][1] + mbc : mg2[n][1] + mbc;
    if (ovg[n][1] - ovg[n][0] < 0) return;
  }
  for (int n = 0; n < 3; n++) {
    if (n < nd) {
      o1[n][0] = ovg[n][0] > mg2[n][0] ? ovg[n][0] : mg2[n][0];
      o1[n][1] = ovg[n][1] < mg2[n][1] ? ovg[n][1] : mg2[n][1];
      o2[n][0] = ovg[n][0] > mg1[n][0] ? ovg[n][0] : mg1[n][0];
      o2[n][1] = ovg[n][1] < mg1[n][1] ? ovg[n][1] : mg1[n][1];
    } else {
      o1[n][0] = o1[n][1] = o2[n][0] = o2[n][1] = 1;
    }
  }
  /* grid1 <- grid2 over o1 */
  for (int v = 0; v < c->nv; v++)
    for (int k = o1[2][0]; k <= o1[2][1]; k++)
      for (int j = o1[1][0]; j <= o1[1][1]; j++)
        for (int i = o1[0][0]; i <= o1[0][1]; i++) {
          int k1 = nd == 3 ? k - mg1[2][0] + 1 : 1, k2 = nd == 3 ? k - mg2[2][0] + 1 : 1;
          Q(G1, i - mg1[0][0] + 1, j - mg1[1][0] + 1, k1, v) = Q(G2, i - mg2[0][0] + 1, j - mg2[1][0] + 1, k2, v);
        }
  /* grid2 <- grid1 over o2 */
  for (int v = 0; v < c->nv; v++)
    for (int k = o2[2][0]; k <= o2[2][1]; k++)
      for (int j = o2[1][0]; j <= o2[1][1]; j++)
        for (int i = o2[0][0]; i <= o2[0][1]; i++) {
          int k1 = nd == 3 ? k - mg1[2][0] + 1 : 1, k2 = nd == 3 ? k - mg2[2][0] + 1 : 1;
          Q(G2, i - mg2[0][0] + 1, j - mg2[1][0] + 1, k2, v) = Q(G1, i - mg1[0][0] + 1, j - mg1[1][0] + 1, k1, v);
        }
}

/* SelfPeriodicBC (boundary.f90:480-577); qpo = 0 (griddata default). */
static void self_periodic(orc_ctx* c, opatch* P) {
  const int* m = P->d.mthbc;
  const int mx = P->mx[0], my = P->mx[1], mz = P->mx[2];
  const int nz0 = (c->ndims == 3) ? 0 : 1, nz1 = (c->ndims == 3) ? mz + 1 : 1;
  if (m[0] == 2 && m[1] == 2)
    for (int v = 0; v < c->nv; v++)
      for (int k = nz0; k <= nz1; k++)
        for (int j = 0; j <= my + 1; j++) {
          Q(P, mx + 1, j, k, v) = Q(P, 1, j, k, v) + 0.0;
          Q(P, 0, j, k, v) = Q(P, mx, j, k, v) - 0.0;
        }
  if (m[2] == 2 && m[3] == 2)
    for (int v = 0; v < c->nv; v++)
      for (int k = nz0; k <= nz1; k++)
        for (int i = 0; i <= mx + 1; i++) {
          Q(P, i, my + 1, k, v) = Q(P, i, 1, k, v);
          Q(P, i, 0, k, v) = Q(P, i, my, k, v);
        }
  if (c->ndims == 3 && m[4] == 2 && m[5] == 2)
    for (int v = 0; v < c->nv; v++)
      for (int j = 0; j <= my + 1; j++)
        for (int i = 0; i <= mx + 1; i++) {
          Q(P, i, j, mz + 1, v) = Q(P, i, j, 1, v);
          Q(P, i, j, 0, v) = Q(P, i, j, mz, v);
        }
}

/* GetPeriodicOffsets (boundary.f90:875-954). */
static void periodic_offset(orc_ctx* c, int level, int idx, int off[3]) {
  int s = 1 << level;
  int X = c->rootmx[0] * s, Y = c->rootmx[1] * s, Z = c->rootmx[2] * s;
  static const int tab[13][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {-1, 1, 0},
                                 {0, 1, 1}, {0, 1, -1}, {1, 0, 1}, {-1, 0, 1}, {1, 1, 1},
                                 {1, 1, -1}, {-1, 1, 1}, {1, -1, 1}};
  off[0] = tab[idx - 1][0] * X;
  off[1] = tab[idx - 1][1] * Y;
  off[2] = tab[idx - 1][2] * Z;
}

/* SetPhysicalBC2D/3D + UserBC (boundary.f90:642-802; CH/problem.f90:770-881;
 * BioGrowth/problem.f90:623-673).  Faces are applied in order 1..2*ndims,
 * each over the full ghost-inclusive extent of the other dimensions. */
static int physical_bc(orc_ctx* c, opatch* P) {
  const int nd = c->ndims;
  for (int face = 0; face < 2 * nd; face++) {
    int code = P->d.mthbc[face];
    if (code == 2 || code == 999) continue;
    if (code != 10) return fail("SetPhysicalBC: user boundary condition %d on boundary %d not coded", code, face + 1);
    int dim = face / 2, hi = face % 2;
    int ta = (dim == 0) ? 1 : 0, tb = (dim == 2) ? 1 : 2;
    int na = P->nq[ta], nb = P->nq[tb];
    int m = P->mx[dim];
    for (int v = 0; v < c->nv; v++) {
      /* CH 2D reflects both fields, CH 3D sets 2-q, BG reflects phi and sets
       * 2-c; the tumour model reflects phi and mu, sets 2-n and -p */
      int dirichlet = (c->p.model == BSAM_MODEL_CH) ? (nd == 3) : (v == 1);
      if (c->p.model == BSAM_MODEL_TUMOUR) dirichlet = (v == 2) ? 1 : (v == 3 ? 2 : 0);
      for (int b = 0; b < nb; b++)
        for (int a = 0; a < na; a++) {
          int g[3], s[3];
          g[ta] = s[ta] = a;
          g[tb] = s[tb] = (nd == 2 && tb == 2) ? 1 : b;
          g[dim] = hi ? m + 1 : 0;
          s[dim] = hi ? m : 1;
          double x = Q(P, s[0], s[1], s[2], v);
          Q(P, g[0], g[1], g[2], v) = dirichlet == 2 ? -x : (dirichlet ? 2.0 - x : x);
        }
    }
  }
  return BSAM_OK;
}

int32_t orc_set_ghost(orc_ctx* c, int32_t level, int32_t ipass) {
  int l = lidx(level), n = c->lvl_n[l];
  int* L = c->lvl_list[l];
  if (level > 0)
    for (int a = 0; a < n; a++) {
      if (c->ndims == 2) interpolate_2d(c, &c->pt[L[a]], ipass);
      else interpolate_3d(c, &c->pt[L[a]], ipass);
    }
  if (level > 0)
    for (int a = 0; a < n; a++)
      for (int b = a + 1; b < n; b++) ghost_overlap(c, &c->pt[L[a]], &c->pt[L[b]], NULL);
  if (c->periodic) {
    for (int a = 0; a < n; a++) self_periodic(c, &c->pt[L[a]]);
    if (level > 0)
      for (int a = 0; a < n; a++)
        for (int b = a + 1; b < n; b++)
          for (int pol = -1; pol <= 1; pol += 2)
            for (int o = 0; o < c->nper; o++) {
              int off[3];
              periodic_offset(c, level, c->peridx[o], off);
              for (int s = 0; s < 3; s++) off[s] *= pol;
              ghost_overlap(c, &c->pt[L[a]], &c->pt[L[b]], off);
            }
  }
  for (int a = 0; a < n; a++) {
    int r = physical_bc(c, &c->pt[L[a]]);
    if (r) return r;
  }
  return BSAM_OK;
}

/* ---- level drivers ------------------------------------------------------- */
#define FOR_LEVEL(c, level, P)                                       \
  for (int a_ = 0, l_ = lidx(level); a_ < (c)->lvl_n[l_]; a_++)      \
    for (opatch* P = &(c)->pt[(c)->lvl_list[l_][a_]]; P; P = NULL)

int32_t orc_relax_level(orc_ctx* c, int32_t level, int32_t ipass) {
  FOR_LEVEL(c, level, P) {
    if (c->ndims == 2) relax_patch_2d(c, P, ipass);
    else relax_patch_3d(c, P, ipass);
  }
  return BSAM_OK;
}

int32_t orc_level_relax(orc_ctx* c, int32_t level) {
  for (int s = 0; s < c->p.nsmoothingpasses; s++)
    for (int rb = 1; rb <= 2; rb++) {
      orc_relax_level(c, level, rb);
      int r = orc_set_ghost(c, level, rb);
      if (r) return r;
    }
  return BSAM_OK;
}

int32_t orc_restrict_solution(orc_ctx* c, int32_t level) {
  FOR_LEVEL(c, level, P) restrict_solution_patch(c, P);
  return BSAM_OK;
}
int32_t orc_coarse_ghost_points(orc_ctx* c, int32_t level) {
  FOR_LEVEL(c, level, P) coarse_ghost_patch(c, P);
  return BSAM_OK;
}
int32_t orc_coarse_loading(orc_ctx* c, int32_t level) {
  FOR_LEVEL(c, level, P) coarse_loading_patch(c, P);
  return BSAM_OK;
}
int32_t orc_correct_fine(orc_ctx* c, int32_t level) {
  FOR_LEVEL(c, level, P) correct_fine_patch(c, P);
  return BSAM_OK;
}
int32_t orc_get_source(orc_ctx* c, int32_t level) {
  FOR_LEVEL(c, level, P) get_source_patch(c, P);
  return BSAM_OK;
}
int32_t orc_update_source(orc_ctx* c, int32_t level) {
  FOR_LEVEL(c, level, P) update_source_patch(c, P);
  return BSAM_OK;
}
int32_t orc_residual(orc_ctx* c, int32_t level) {
  FOR_LEVEL(c, level, P) residual_patch(c, P);
  return BSAM_OK;
}
int32_t orc_copy_q_to_qold(orc_ctx* c, int32_t level) {
  FOR_LEVEL(c, level, P) {
    size_t n = (size_t)P->nq[0] * P->nq[1] * P->nq[2] * c->nv;
    memcpy(P->qold, P->q, n * 8);
  }
  return BSAM_OK;
}

/* Regrid tagging of one level: EstimateLevelErrors (bsamroutines.f90:
 * 994-1040) and ErrFlag (:1059-1131) up to BufferAndList, which stays with
 * the caller.  errflagopt 1 (errflagdefault): RestrictSolution,
 * SetGhost(level-1, 0), GetCoarseGhostPoints, CopyQToQold on the level and
 * its parent level, RelativeTruncationError (afasroutines.f90:739-814)
 *   qrte = Restriction(Operator(q, qold, h)) - Operator(qc, parent%qold, ch)
 * and SetErrFlags2D/3D (bsamroutines.f90:1134-1203): the 2^d children of a
 * coarse cell whose largest |qrte| over the variables exceeds
 * qtolerance/h/h (2D) or qtolerance/h/h/h (3D) are flagged.  errflagopt 10
 * (errflaguser): the second-difference test of the CH and BioGrowth
 * SetErrFlagsUser2D/3D (CH/problem.f90:248-307, BioGrowth/problem.f90:
 * 185-240) on variable 1 of q as it is:
 *   |q(i+1)-2q(i)+q(i-1)| + |..j..| (+ |..k..|) > ndims*qtolerance.
 * errflagopt 11: the gradient test of the probe PROBLEM module's
 * SetErrFlagsUser2D/3D (oracle/ref_problem/problem.f90), the user tagging
 * of the tumour model and of the AMR workloads' meshes, on variable 1:
 *   sqrt(gx^2 + gy^2 (+ gz^2)) > qtolerance, g = (q(+1) - q(-1))/2.
 * flags: the level's patches in level order, each column-major over 1..mx;
 * *ntagged counts the flagged cells. */
int32_t orc_estimate_level_errors(orc_ctx* c, int32_t level, int32_t errflagopt, double qtolerance,
                                  int32_t* flags, int64_t* ntagged) {
  if (!flags || !ntagged) return fail("null argument");
  if (level <= c->minlevel || level > c->finestlevel) return fail("bad level %d", level);
  if (errflagopt != 1 && errflagopt != 10 && errflagopt != 11) return fail("errflagopt must be 1, 10 or 11");
  if (errflagopt == 10 && c->p.model == BSAM_MODEL_TUMOUR)
    return fail("errflagopt 10 is not implemented for the tumour model: its user tagging is errflagopt 11");
  if (errflagopt == 1) {
    orc_restrict_solution(c, level);
    int r = orc_set_ghost(c, level - 1, 0);
    if (r) return r;
    orc_coarse_ghost_points(c, level);
    orc_copy_q_to_qold(c, level);
    orc_copy_q_to_qold(c, level - 1);
  }
  int64_t off = 0, nt = 0;
  FOR_LEVEL(c, level, P) {
    const int mx = P->mx[0], my = P->mx[1], mz = P->mx[2];
    int32_t* fl = flags + off;
#define FL(i, j, k) fl[((size_t)((k) - 1) * my + ((j) - 1)) * mx + ((i) - 1)]
    if (errflagopt == 1) {
      opatch* Par = &c->pt[P->d.parent];
      const int (*mb)[2] = P->d.mbounds;
      const double h = P->d.dx, ch = Par->d.dx;
      const double tol = c->ndims == 3 ? qtolerance / h / h / h : qtolerance / h / h;
      qview q = {P->q, P->nq, {0, 0, 0}}, qo = {P->qold, P->nq, {0, 0, 0}};
      qview qc = {P->qc, P->nc, {0, 0, 0}};
      qview pqo = {Par->qold, Par->nq, {mb[0][0] - 1, mb[1][0] - 1, c->ndims == 3 ? mb[2][0] - 1 : 0}};
      for (size_t n = 0; n < (size_t)mx * my * mz; n++) fl[n] = 0;
      for (int kc = 1; kc <= kc_hi(c, P); kc++)
        for (int jc = 1; jc <= P->cmx[1]; jc++)
          for (int ic = 1; ic <= P->cmx[0]; ic++) {
            double oc[4], mv = -DBL_MAX;
            op_cell(c, qc, pqo, ch, ic, jc, kc, oc);
            for (int v = 0; v < c->nv; v++) {
              double r;
#define GO(i, j, k) op_var(c, q, qo, h, i, j, k, v)
              RESTRICT_CELL(c->ndims, GO, ic, jc, kc, r);
#undef GO
              const double a = fabs(r - oc[v]);
              if (a > mv) mv = a;
            }
            if (mv > tol) {
              const int k0 = c->ndims == 3 ? 2 * kc - 1 : 1, k1 = c->ndims == 3 ? 2 * kc : 1;
              for (int k = k0; k <= k1; k++)
                for (int j = 2 * jc - 1; j <= 2 * jc; j++)
                  for (int i = 2 * ic - 1; i <= 2 * ic; i++) FL(i, j, k) = 1;
            }
          }
    } else if (errflagopt == 11) {
      for (int k = 1; k <= mz; k++)
        for (int j = 1; j <= my; j++)
          for (int i = 1; i <= mx; i++) {
            const double gx = 0.5 * (Q(P, i + 1, j, k, 0) - Q(P, i - 1, j, k, 0));
            const double gy = 0.5 * (Q(P, i, j + 1, k, 0) - Q(P, i, j - 1, k, 0));
            double g2 = gx * gx + gy * gy;
            if (c->ndims == 3) {
              const double gz = 0.5 * (Q(P, i, j, k + 1, 0) - Q(P, i, j, k - 1, 0));
              g2 = g2 + gz * gz;
            }
            FL(i, j, k) = sqrt(g2) > qtolerance ? 1 : 0;
          }
    } else {
      const double thr = (c->ndims == 3 ? 3.0 : 2.0) * qtolerance;
      for (int k = 1; k <= mz; k++)
        for (int j = 1; j <= my; j++)
          for (int i = 1; i <= mx; i++) {
            double d = fabs(Q(P, i + 1, j, k, 0) - 2.0 * Q(P, i, j, k, 0) + Q(P, i - 1, j, k, 0)) +
                       fabs(Q(P, i, j + 1, k, 0) - 2.0 * Q(P, i, j, k, 0) + Q(P, i, j - 1, k, 0));
            if (c->ndims == 3) d = d + fabs(Q(P, i, j, k + 1, 0) - 2.0 * Q(P, i, j, k, 0) + Q(P, i, j, k - 1, 0));
            FL(i, j, k) = d > thr ? 1 : 0;
          }
    }
#undef FL
    for (size_t n = 0; n < (size_t)mx * my * mz; n++) nt += fl[n];
    off += (int64_t)mx * my * mz;
  }
  *ntagged = nt;
  return BSAM_OK;
}

int32_t orc_fill_down(orc_ctx* c, int32_t field) {
  for (int level = c->finestlevel; level >= c->minlevel + 1; level--) {
    FOR_LEVEL(c, level, P) filldown_patch(c, P, field);
    if (field == BSAM_SOLUTIONFIELD) {
      int r = orc_set_ghost(c, level - 1, 0);
      if (r) return r;
    }
  }
  return BSAM_OK;
}

int32_t orc_error_afas(orc_ctx* c, int32_t level, double* norm) {
  double s = 0.0, cnt = 0.0;
  double comp[16] = {0};
  FOR_LEVEL(c, level, P) {
    residual_patch(c, P);
    size_t nf = (size_t)P->mx[0] * P->mx[1] * P->mx[2];
    if (c->p.errortype == 1) {
      double t = 0.0;
      for (size_t e = 0; e < nf * (size_t)c->nv; e++) t += P->rf[e] * P->rf[e];
      s = t + s;
      cnt = (double)(c->nv * (long)nf) + cnt;
    } else {
      for (int v = 0; v < c->nv; v++) {
        double t = 0.0;
        for (size_t e = 0; e < nf; e++) t += P->rf[v * nf + e] * P->rf[v * nf + e];
        comp[v] = t + comp[v];
      }
      cnt = (double)nf + cnt;
    }
  }
  if (c->p.errortype == 1) {
    *norm = sqrt(s / cnt);
  } else if (c->p.errortype == 2) {
    double m = -1.0;
    for (int v = 0; v < c->nv; v++) {
      double e = sqrt(comp[v] / cnt);
      if (e > m) m = e;
    }
    *norm = m;
  } else {
    return fail("ErrorAFAS: only errortype = 1,2 are supported");
  }
  return BSAM_OK;
}

int32_t orc_vcycle(orc_ctx* c, int32_t level) {
  int r = orc_level_relax(c, level);
  if (r) return r;
  if (level > c->minlevel) {
    orc_restrict_solution(c, level);
    if ((r = orc_set_ghost(c, level - 1, 0))) return r;
    orc_coarse_ghost_points(c, level);
    orc_coarse_loading(c, level);
    if ((r = orc_vcycle(c, level - 1))) return r;
    orc_correct_fine(c, level);
    if ((r = orc_set_ghost(c, level, 0))) return r;
    if ((r = orc_level_relax(c, level))) return r;
  }
  return BSAM_OK;
}

int32_t orc_multigrid_iterations(orc_ctx* c, int32_t* ncycles, double* residual,
                                 double* history, int32_t history_len) {
  int r;
  if ((r = orc_fill_down(c, BSAM_SOLUTIONFIELD))) return r;
  for (int level = c->finestlevel; level >= 0; level--) orc_get_source(c, level);
  for (int level = c->finestlevel; level >= 0; level--) orc_update_source(c, level);
  orc_fill_down(c, BSAM_SOURCEFIELD);
  double res = 0.0;
  int it;
  for (it = 1; it <= c->p.maxvcycles; it++) {
    if ((r = orc_vcycle(c, c->finestlevel))) return r;
    if ((r = orc_fill_down(c, BSAM_SOLUTIONFIELD))) return r;
    for (int level = c->finestlevel; level >= 0; level--) orc_update_source(c, level);
    orc_fill_down(c, BSAM_SOURCEFIELD);
    if ((r = orc_error_afas(c, c->finestlevel, &res))) return r;
    if (history && it - 1 < history_len) history[it - 1] = res;
    if (!isfinite(res)) { *ncycles = it; *residual = res; return BSAM_EDIVERGED; }
    if (res < c->p.qerrortol) break;
  }
  int done = it > c->p.maxvcycles ? c->p.maxvcycles : it;
  *ncycles = done;
  *residual = res;
  return (it > c->p.maxvcycles) ? BSAM_NOT_CONVERGED : BSAM_OK;
}

/* One time step on a frozen mesh, TakeTimeSteps (bsamroutines.f90:673-758)
 * with the regrid reduced to what AMR(rootlevel) (:966-991) does on an
 * unchanged mesh: SetGhost(level, 0) from the root up.  Then FillDown on the
 * first step (:693-697), CopyQToQold on every level (:700-702) and
 * MultigridIterations (:758). */
int32_t orc_time_step(orc_ctx* c, int32_t first_step, int32_t* ncycles, double* residual) {
  int r;
  for (int level = 0; level <= c->finestlevel; level++)
    if ((r = orc_set_ghost(c, level, 0))) return r;
  if (first_step && (r = orc_fill_down(c, BSAM_SOLUTIONFIELD))) return r;
  for (int level = c->minlevel; level <= c->finestlevel; level++) orc_copy_q_to_qold(c, level);
  return orc_multigrid_iterations(c, ncycles, residual, NULL, 0);
}

/* ---- regrid: fields onto a new hierarchy ---------------------------------
 * InitFields (bsamroutines.f90:1787-1856): a new patch's qc(0:cmx+1) is the
 * parent's q(mb-1:mb+1), and q(0:mx+1) the mass-corrected prolongation of
 * it, BiLinProlongationP1MC / TriLinProlongationP1MC (gridutilities.f90:
 * 242-290, 628-705): the P1 prolongation (as in CorrectFine), then each
 * interior coarse cell's value minus the mean of its 2^d children, summed
 * in the reference's order, added to those children.  MakeNewGrid
 * (:1773-1779) zeroes f and rf and sets qold = q.  TransferValues
 * (:1859-1951) then copies q over the interior overlap of every old patch
 * of the same level.  The levels are done from the root up, each followed
 * by its SetGhost(level, 0) (AMR, :978-989), since the next level's
 * InitFields reads the parent's ghosts.  Levels <= 0 keep their q and qold. */
static void init_fields_patch(orc_ctx* c, opatch* P) {
  const opatch* Par = &c->pt[P->d.parent];
  const int (*mb)[2] = P->d.mbounds;
  const int cx = P->cmx[0], cy = P->cmx[1], cz = P->cmx[2];
  const int nd = c->ndims;
  for (int v = 0; v < c->nv; v++) {
    if (nd == 2) {
      for (int jc = 0; jc <= cy + 1; jc++)
        for (int ic = 0; ic <= cx + 1; ic++)
          QC(P, ic, jc, 1, v) = Q(Par, mb[0][0] - 1 + ic, mb[1][0] - 1 + jc, 1, v);
      for (int j = 0; j <= P->mx[1] + 1; j++)
        for (int i = 0; i <= P->mx[0] + 1; i++) {
          int il = i / 2, jl = j / 2;
          double b0 = lin_w(QC(P, il, jl, 1, v), QC(P, il + 1, jl, 1, v), i);
          double b1 = lin_w(QC(P, il, jl + 1, 1, v), QC(P, il + 1, jl + 1, 1, v), i);
          Q(P, i, j, 1, v) = lin_w(b0, b1, j) / 16.0;
        }
      for (int jc = 1; jc <= cy; jc++)
        for (int ic = 1; ic <= cx; ic++) {
          const int i2 = 2 * ic, j2 = 2 * jc;
          double s = Q(P, i2, j2, 1, v);
          s = s + Q(P, i2 - 1, j2, 1, v);
          s = s + Q(P, i2, j2 - 1, 1, v);
          s = s + Q(P, i2 - 1, j2 - 1, 1, v);
          const double cor = QC(P, ic, jc, 1, v) - 0.25 * s;
          Q(P, i2, j2, 1, v) += cor;
          Q(P, i2 - 1, j2, 1, v) += cor;
          Q(P, i2, j2 - 1, 1, v) += cor;
          Q(P, i2 - 1, j2 - 1, 1, v) += cor;
        }
    } else {
      for (int kc = 0; kc <= cz + 1; kc++)
        for (int jc = 0; jc <= cy + 1; jc++)
          for (int ic = 0; ic <= cx + 1; ic++)
            QC(P, ic, jc, kc, v) = Q(Par, mb[0][0] - 1 + ic, mb[1][0] - 1 + jc, mb[2][0] - 1 + kc, v);
      for (int k = 0; k <= P->mx[2] + 1; k++)
        for (int j = 0; j <= P->mx[1] + 1; j++)
          for (int i = 0; i <= P->mx[0] + 1; i++) {
            int il = i / 2, jl = j / 2, kl = k / 2;
            double cc[2];
            for (int dk = 0; dk < 2; dk++) {
              double b0 = lin_w(QC(P, il, jl, kl + dk, v), QC(P, il + 1, jl, kl + dk, v), i);
              double b1 = lin_w(QC(P, il, jl + 1, kl + dk, v), QC(P, il + 1, jl + 1, kl + dk, v), i);
              cc[dk] = lin_w(b0, b1, j);
            }
            Q(P, i, j, k, v) = lin_w(cc[0], cc[1], k) / 64.0;
          }
      for (int kc = 1; kc <= cz; kc++)
        for (int jc = 1; jc <= cy; jc++)
          for (int ic = 1; ic <= cx; ic++) {
            const int i2 = 2 * ic, j2 = 2 * jc, k2 = 2 * kc;
            double s = Q(P, i2, j2, k2, v);
            s = s + Q(P, i2 - 1, j2, k2, v);
            s = s + Q(P, i2, j2 - 1, k2, v);
            s = s + Q(P, i2 - 1, j2 - 1, k2, v);
            s = s + Q(P, i2, j2, k2 - 1, v);
            s = s + Q(P, i2 - 1, j2, k2 - 1, v);
            s = s + Q(P, i2, j2 - 1, k2 - 1, v);
            s = s + Q(P, i2 - 1, j2 - 1, k2 - 1, v);
            const double cor = QC(P, ic, jc, kc, v) - 0.125 * s;
            for (int n = 0; n < 8; n++) Q(P, i2 - (n & 1), j2 - ((n >> 1) & 1), k2 - (n >> 2), v) += cor;
          }
    }
  }
  memcpy(P->qold, P->q, sizeof(double) * (size_t)P->nq[0] * P->nq[1] * P->nq[2] * c->nv);
}

/* Transferq / TransferOverlap: q over the interior overlap of mglobal. */
static void transfer_patch(orc_ctx* c, const opatch* S, opatch* T) {
  int lo[3] = {1, 1, 1}, hi[3] = {1, 1, 1};
  for (int a = 0; a < c->ndims; a++) {
    lo[a] = S->d.mglobal[a][0] > T->d.mglobal[a][0] ? S->d.mglobal[a][0] : T->d.mglobal[a][0];
    hi[a] = S->d.mglobal[a][1] < T->d.mglobal[a][1] ? S->d.mglobal[a][1] : T->d.mglobal[a][1];
    if (hi[a] < lo[a]) return;
  }
  for (int v = 0; v < c->nv; v++)
    for (int k = lo[2]; k <= hi[2]; k++)
      for (int j = lo[1]; j <= hi[1]; j++)
        for (int i = lo[0]; i <= hi[0]; i++) {
          const int ks = c->ndims == 3 ? k - S->d.mglobal[2][0] + 1 : 1;
          const int kt = c->ndims == 3 ? k - T->d.mglobal[2][0] + 1 : 1;
          Q(T, i - T->d.mglobal[0][0] + 1, j - T->d.mglobal[1][0] + 1, kt, v) =
              Q(S, i - S->d.mglobal[0][0] + 1, j - S->d.mglobal[1][0] + 1, ks, v);
        }
}

int32_t orc_regrid(orc_ctx* c, int32_t npatch, const bsam_patch_desc* d, int32_t minlevel, int32_t finestlevel) {
  /* detach the old hierarchy */
  opatch* old = c->pt;
  const int oldn = c->npatch, oldmin = c->minlevel;
  int* olist[MAXLEVELS];
  int on[MAXLEVELS];
  for (int l = 0; l < MAXLEVELS; l++) {
    olist[l] = c->lvl_list[l];
    on[l] = c->lvl_n[l];
    c->lvl_list[l] = NULL;
    c->lvl_n[l] = 0;
  }
  c->pt = NULL;
  c->npatch = 0;
  int r = orc_set_hierarchy(c, npatch, d, minlevel, finestlevel);
  if (r == BSAM_OK && minlevel != oldmin) r = fail("regrid: minlevel %d differs from %d", minlevel, oldmin);
  for (int L = minlevel; r == BSAM_OK && L <= 0; L++) {
    const int l = lidx(L);
    if (c->lvl_n[l] != 1 || on[l] != 1) { r = fail("regrid: level %d must hold one patch", L); break; }
    opatch* P = &c->pt[c->lvl_list[l][0]];
    const opatch* O = &old[olist[l][0]];
    for (int a = 0; a < 3; a++)
      if (P->mx[a] != O->mx[a]) r = fail("regrid: level %d changed its extent", L);
    if (r) break;
    const size_t n = (size_t)P->nq[0] * P->nq[1] * P->nq[2] * c->nv;
    memcpy(P->q, O->q, n * sizeof(double));
    memcpy(P->qold, O->qold, n * sizeof(double));
  }
  if (r == BSAM_OK) r = orc_set_ghost(c, 0, 0);
  for (int L = 1; r == BSAM_OK && L <= finestlevel; L++) {
    const int l = lidx(L);
    for (int a = 0; a < c->lvl_n[l]; a++) init_fields_patch(c, &c->pt[c->lvl_list[l][a]]);
    for (int a = 0; a < c->lvl_n[l]; a++)
      for (int b = 0; b < on[l]; b++) transfer_patch(c, &old[olist[l][b]], &c->pt[c->lvl_list[l][a]]);
    r = orc_set_ghost(c, L, 0);
  }
  for (int n = 0; n < oldn; n++) {
    free(old[n].q); free(old[n].qold); free(old[n].qc); free(old[n].f); free(old[n].rf); free(old[n].ftmp);
  }
  free(old);
  for (int l = 0; l < MAXLEVELS; l++) free(olist[l]);
  return r;
}

/* ---- output frame ------------------------------------------------------ */
/* Fortran F25.12 as the amdflang runtime edits it: correctly rounded
 * (glibc's %f is exact, halfway cases to even), the sign kept when a
 * negative value rounds to zero, 25 asterisks when the edited value needs
 * more than 25 characters, NaN / Inf / -Inf right-justified. */
static void f25(char* o, double x) {
  char b[512];
  if (isnan(x)) snprintf(b, sizeof b, "%25s", "NaN");
  else if (isinf(x)) snprintf(b, sizeof b, "%25s", x < 0 ? "-Inf" : "Inf");
  else snprintf(b, sizeof b, "%25.12f", x);
  if (strlen(b) > 25) memset(b, '*', 25);
  memcpy(o, b, 25);
  o[25] = 0;
}
/* Iw: right-justified, w asterisks when it does not fit. */
static void iw(char* o, int w, int v) {
  char b[32];
  snprintf(b, sizeof b, "%*d", w, v);
  if ((int)strlen(b) > w) memset(b, '*', w);
  b[w] = 0;
  strcpy(o, b);
}

/* WriteQ (bsaminputoutput.f90:60-82) with OutputQ (:85-129) and
 * WriteQ2D/3D (:132-181): time (F25.12), finestlevel (I3), then every patch
 * of levels 0..finest in level order: a list-directed blank line, level,
 * ndims, 2, nrvars (I3,3(1x,I3)), dx, xlower, xupper (F25.12 with 1x),
 * mx and mglobal (I8 with 1X), and q(0:mx+1, 0:my+1[, 0:mz+1]) i fastest,
 * one cell per record with its nrvars values edited 10(F25.12,2X) (no
 * trailing blanks).  geom: per patch dx[3], xlower[3], xupper[3]; NULL
 * takes dx and xlower from the descriptors and xupper = xlower + mx*dx. */
int32_t orc_write_frame(orc_ctx* c, const char* path, double time, const double* geom) {
  FILE* fh = fopen(path, "wb");
  if (!fh) return fail("cannot open %s", path);
  char t[64], u[64];
  f25(t, time);
  fprintf(fh, "%s\n", t);
  iw(t, 3, c->finestlevel);
  fprintf(fh, "%s\n", t);
  const int nd = c->ndims, nv = c->nv;
  for (int level = 0; level <= c->finestlevel; level++) {
    FOR_LEVEL(c, level, P) {
      const int pid = (int)(P - c->pt);
      double g[9];
      for (int a = 0; a < 3; a++) {
        g[a] = geom ? geom[9 * pid + a] : P->d.dx;
        g[3 + a] = geom ? geom[9 * pid + 3 + a] : P->d.xlower[a];
        g[6 + a] = geom ? geom[9 * pid + 6 + a] : P->d.xlower[a] + P->mx[a] * P->d.dx;
      }
      fprintf(fh, "  \n");
      int hv[4] = {level, nd, 2, nv};
      for (int n = 0; n < 4; n++) {
        iw(t, 3, hv[n]);
        fprintf(fh, n ? " %s" : "%s", t);
      }
      fprintf(fh, "\n");
      for (int row = 0; row < 3; row++) {
        for (int a = 0; a < nd; a++) {
          f25(u, g[3 * row + a]);
          fprintf(fh, a ? " %s" : "%s", u);
        }
        fprintf(fh, "\n");
      }
      for (int a = 0; a < nd; a++) {
        iw(t, 8, P->mx[a]);
        fprintf(fh, a ? " %s" : "%s", t);
      }
      fprintf(fh, "\n");
      for (int a = 0; a < nd; a++)
        for (int e = 0; e < 2; e++) {
          iw(t, 8, P->d.mglobal[a][e]);
          fprintf(fh, (a || e) ? " %s" : "%s", t);
        }
      fprintf(fh, "\n");
      const int k1 = nd == 3 ? P->mx[2] + 1 : 1, k0 = nd == 3 ? 0 : 1;
      for (int k = k0; k <= k1; k++)
        for (int j = 0; j <= P->mx[1] + 1; j++)
          for (int i = 0; i <= P->mx[0] + 1; i++) {
            for (int v = 0; v < nv; v++) {
              f25(u, Q(P, i, j, k, v));
              fprintf(fh, v ? "  %s" : "%s", u);
            }
            fprintf(fh, "\n");
          }
    }
  }
  if (fclose(fh)) return fail("cannot write %s", path);
  return BSAM_OK;
}

/* F25.12 of n values into out (25 bytes each): the oracle's edit, for
 * checking it against the Fortran runtime's. */
void orc_format_f25(const double* x, int64_t n, char* out) {
  char t[32];
  for (int64_t e = 0; e < n; e++) {
    f25(t, x[e]);
    memcpy(out + 25 * e, t, 25);
  }
}
