/*
 * bsam_gpu.h -- C ABI of the MI355X FAS multigrid V-cycle library.
 *
 * The library replaces the per-level work of BSAM 1.2's Fortran V-cycle
 * (LIBRARY/afasroutines.f90, LIBRARY/boundary.f90, LIBRARY/gridutilities.f90
 * and the PROBLEM-module hooks) with HIP kernels for gfx950.  The host keeps
 * its own patch tree; it describes the mesh hierarchy with a flat table of
 * bsam_patch_desc records (one per patch, including the below-root
 * multigrid levels), moves field arrays in and out with bsam_upload /
 * bsam_download in the reference's own array extents, and calls the
 * per-level entry points below, each of which replaces one
 * `CALL ApplyOnLevel(level, <callback>, dummy)` of the reference, or the
 * coarse-grained bsam_vcycle / bsam_multigrid_iterations.
 *
 * Conventions (as in the reference):
 *   - indices are 1-based with a ghost layer of width mbc=1 (lower bound 0);
 *   - arrays are column-major, i fastest, variable index slowest;
 *   - field arrays are FP64 on the host side of the ABI;
 *   - every call returns int32 status: 0 = ok, <0 = error (see BSAM_E*),
 *     >0 = informational (BSAM_NOT_CONVERGED).  No call ever exits the
 *     process; bsam_last_error() returns the message of the last failure
 *     on the calling thread.
 *   - one host thread drives one context; calls are stream-ordered on the
 *     context's HIP stream.  Only bsam_error_afas, bsam_multigrid_iterations,
 *     bsam_download, bsam_estimate_level_errors, bsam_write_frame and
 *     bsam_synchronize wait for the device.
 *   - the library keeps no host pointer past the call that received it.
 */
#ifndef BSAM_GPU_H
#define BSAM_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define BSAM_OK              0
#define BSAM_NOT_CONVERGED   1   /* maxvcycles exhausted (the reference exits its loop silently, afasroutines.f90:79-108) */
#define BSAM_EINVAL         -1   /* invalid argument or unsupported configuration */
#define BSAM_EHIP           -2   /* HIP runtime error */
#define BSAM_ENOMEM         -3   /* device allocation failed */
#define BSAM_ESTATE         -4   /* call out of order (e.g. no hierarchy) */
#define BSAM_EDIVERGED      -5   /* residual became non-finite */
#define BSAM_ENODEV         -6   /* no usable GPU */

/* ---- models (the PROBLEM-module hook sets the library implements) ---- */
#define BSAM_MODEL_CH        1   /* PROBLEM/CH/problem.f90 hooks */
#define BSAM_MODEL_BG        2   /* PROBLEM/BioGrowth/problem.f90 2D hooks */
#define BSAM_MODEL_TUMOUR    3   /* tumour growth (phi, mu, nutrient, pressure; nrvars = 4):
                                    a model the reference does not ship, written with
                                    the hook signatures of PROBLEM/CH (the oracle's probe
                                    PROBLEM module runs it inside the reference LIBRARY) */

/* ---- field identifiers for upload/download (nodeinfodef.f90:196-214) - */
#define BSAM_FIELD_Q         0   /* q   (1-mbc:mx+mbc)^d x nrvars */
#define BSAM_FIELD_QOLD      1   /* qold, same extents as q */
#define BSAM_FIELD_QC        2   /* qc  (1-mbc:mx/2+mbc)^d x nrvars */
#define BSAM_FIELD_F         3   /* f   (1:mx)^d x nrvars */
#define BSAM_FIELD_RF        4   /* rf  (1:mx)^d x nrvars */
#define BSAM_FIELD_FTMP      5   /* ftmp (1:mx)^d x nrvars */

/* ---- precision modes -------------------------------------------------- */
#define BSAM_PREC_FP64       0   /* FP64 everywhere (the reference's r8) */
#define BSAM_PREC_MIXED      1   /* FP64 outer defect correction, FP32 V-cycles
                                    (CH and tumour models; BSAM_EINVAL for BG) */

/* ---- per-field selector for bsam_fill_down (nodeinfodef.f90:69-70) --- */
#define BSAM_SOURCEFIELD     1
#define BSAM_SOLUTIONFIELD   2

/*
 * One patch of the hierarchy.  Mirrors the mesh part of TYPE(nodeinfo)
 * (nodeinfodef.f90:118-220): level, mx, mbounds, mglobal, mthbc, dx,
 * xlower.  `parent` replaces the tree's Parent pointer (treeops.f90:86-101,
 * GetParentInfo :889-898) with an index into the same descriptor table;
 * it is -1 for the single patch on the coarsest level (minlevel).  Below
 * the root the reference keeps one patch per level (bsamroutines.f90:
 * 443-510); above it any number of patches per level.  In 2D the third
 * entries are ignored (mx[2]=1).  The order of the descriptors within a
 * level is the order the library visits patches (the reference visits
 * youngest to eldest, treeops.f90:539-550); only the floating-point order
 * of the residual-norm sum depends on it.
 */
typedef struct bsam_patch_desc {
  int32_t level;
  int32_t parent;
  int32_t mx[3];
  int32_t mglobal[3][2];
  int32_t mbounds[3][2];
  int32_t mthbc[6];
  double  dx;
  double  xlower[3];
} bsam_patch_desc;

/*
 * Run parameters.  The V-cycle knobs are the reference's rundata NAMELIST
 * (bsamroutines.f90:77-96); the model constants replace the PROBLEM module
 * globals (CH/problemdef.f90:5, set in CH/problem.f90:32) and `dt`
 * (nodeinfodef.f90:81).
 */
typedef struct bsam_params {
  int32_t ndims;             /* 2 or 3 */
  int32_t nrvars;            /* 2 for CH and BG, 4 for the tumour model */
  int32_t mbc;               /* must be 1 */
  int32_t maux;              /* must be 0 (CH and BG use no aux fields) */
  int32_t model;             /* BSAM_MODEL_* */
  int32_t nsmoothingpasses;  /* V(nu,nu); reference default 2 */
  int32_t maxvcycles;        /* reference default 20 */
  int32_t errortype;         /* 1 (RMS) or 2 (max component RMS) */
  int32_t updateauxfreq;     /* reference default 1 */
  int32_t precision;         /* BSAM_PREC_* */
  int32_t batching;          /* 1: one launch per level; 0: one launch per patch */
  int32_t use_graph;         /* 1: capture the V-cycle loop body as a hipGraph */
  double  dt;
  double  qerrortol;         /* reference default 1e-6 */
  double  model_params[8];   /* CH: [0]=eps^2.  BG: [0]=eps^2,[1]=gamma,[2]=lambda,[3]=czero
                                (BioGrowth/problemdef.f90:5).  Tumour: [0]=eps^2,
                                [1]=proliferation, [2]=apoptosis, [3]=nutrient consumption */
} bsam_params;

typedef struct bsam_ctx bsam_ctx;

/* Create a context on the current HIP device.  Fails with BSAM_ENODEV when
 * no GPU is visible.  Replaces the NAMELIST read (bsamroutines.f90:99-105)
 * and SetProb (CH/problem.f90:12-34) as the source of parameters. */
int32_t bsam_create(const bsam_params* params, bsam_ctx** out);
int32_t bsam_destroy(bsam_ctx* ctx);
/* Change dt between time steps (nodeinfodef.f90:81 is a global the host may
 * change, e.g. for syncelliptic, bsamroutines.f90:641-643). */
int32_t bsam_set_dt(bsam_ctx* ctx, double dt);

/* Install a mesh hierarchy (after every regrid, bsamroutines.f90:682-689).
 * Allocates the device pools in the reference's extents (bsamstorage.f90:
 * 54-144), zero-initialised as AllocFields does, and builds the per-level
 * work lists and ghost-fill gather maps.  `mxmax_root` is the root grid's
 * mx, from which the periodic offsets per level follow (bsamroutines.f90:
 * 387-390, boundary.f90:875-954).  Periodicity is taken from the root's
 * mthbc (boundary.f90:805-872). */
int32_t bsam_set_hierarchy(bsam_ctx* ctx, int32_t npatch,
                           const bsam_patch_desc* desc,
                           int32_t minlevel, int32_t finestlevel);

/* Flat copies of one field of one patch, in the reference's extents and
 * column-major order (see BSAM_FIELD_*).  `count` is the number of doubles
 * the host buffer holds; it must equal the field's size. */
int32_t bsam_upload(bsam_ctx* ctx, int32_t patch, int32_t field,
                    const double* host, int64_t count);
int32_t bsam_download(bsam_ctx* ctx, int32_t patch, int32_t field,
                      double* host, int64_t count);
/* Number of doubles of one field of one patch. */
int64_t bsam_field_size(const bsam_ctx* ctx, int32_t patch, int32_t field);

/* ---- per-level entry points: each replaces one ApplyOnLevel sweep ----- */
/* RelaxPatch{iswitch=ipass} over the level (afasroutines.f90:560-607). */
int32_t bsam_relax_level(bsam_ctx* ctx, int32_t level, int32_t ipass);
/* SetGhost(level, ipass) (boundary.f90:67-97). */
int32_t bsam_set_ghost(bsam_ctx* ctx, int32_t level, int32_t ipass);
/* LevelRelax(level) (afasroutines.f90:532-557). */
int32_t bsam_level_relax(bsam_ctx* ctx, int32_t level);
/* RestrictSolution over the level (afasroutines.f90:610-652). */
int32_t bsam_restrict_solution(bsam_ctx* ctx, int32_t level);
/* GetCoarseGhostPoints over the level (boundary.f90:957-1015). */
int32_t bsam_coarse_ghost_points(bsam_ctx* ctx, int32_t level);
/* GetCoarseLevelLoading(level) (afasroutines.f90:655-736). */
int32_t bsam_coarse_loading(bsam_ctx* ctx, int32_t level);
/* CorrectFine over the level (afasroutines.f90:860-980). */
int32_t bsam_correct_fine(bsam_ctx* ctx, int32_t level);
/* FillDown(field) (afasroutines.f90:236-327). */
int32_t bsam_fill_down(bsam_ctx* ctx, int32_t field);
/* GetSourceFunction / UpdateSourceFunction / UpdateAuxInVcycle over the
 * level (afasroutines.f90:330-492). */
int32_t bsam_get_source(bsam_ctx* ctx, int32_t level);
int32_t bsam_update_source(bsam_ctx* ctx, int32_t level);
int32_t bsam_update_aux(bsam_ctx* ctx, int32_t level);
/* Residual(info) over the level: rf = f - L(q) (afasroutines.f90:817-857). */
int32_t bsam_residual(bsam_ctx* ctx, int32_t level);
/* ErrorAFAS(level) (afasroutines.f90:113-233); waits for the device. */
int32_t bsam_error_afas(bsam_ctx* ctx, int32_t level, double* norm);
/* CopyQToQold over the level (bsamroutines.f90:592-605). */
int32_t bsam_copy_q_to_qold(bsam_ctx* ctx, int32_t level);
/* Regrid tagging of one level: EstimateLevelErrors (bsamroutines.f90:
 * 994-1040) and ErrFlag (:1059-1131) up to BufferAndList, which, with
 * InflateEdgeTags and the new grids, stays with the caller.
 *   errflagopt 1 (errflagdefault): RestrictSolution(level),
 *     SetGhost(level-1, 0), GetCoarseGhostPoints(level), CopyQToQold on
 *     level and level-1 (changing q, qc and qold as the reference does),
 *     then RelativeTruncationError (afasroutines.f90:739-814) and
 *     SetErrFlags2D/3D (bsamroutines.f90:1134-1203) with tolerance
 *     qtolerance/h^ndims;
 *   errflagopt 10 (errflaguser): the second-difference test of the CH and
 *     BioGrowth SetErrFlagsUser2D/3D (CH/problem.f90:248-307,
 *     BioGrowth/problem.f90:185-240) on variable 1 of q as it is, against
 *     ndims*qtolerance (BSAM_EINVAL for the tumour model, whose user
 *     tagging is a different test);
 *   errflagopt 11: the gradient test of the probe PROBLEM module that
 *     built the AMR workloads and is the tumour model's user tagging
 *     (oracle/ref_problem/problem.f90 SetErrFlagsUser2D/3D), on variable 1
 *     of q as it is: sqrt(gx^2 + gy^2 (+ gz^2)) > qtolerance with
 *     g = (q(+1) - q(-1))/2.
 * flags (host, int32) receives info%errorflags(1:mx) of every patch of the
 * level, patches in descriptor order, each column-major; *ntagged the
 * number of flagged cells.  level must be above minlevel.  Synchronous. */
int32_t bsam_estimate_level_errors(bsam_ctx* ctx, int32_t level, int32_t errflagopt, double qtolerance,
                                   int32_t* flags, int64_t* ntagged);
/* The same tagging without any allocation (no counterpart in the
 * reference): for callers that cannot allocate device memory during the
 * time loop (HIP graph capture around it, tight serving loops, harnesses
 * that forbid hipMalloc).  bsam_estimate_level_errors grows a device buffer
 * for the level's flags on first use; this call stages them in a fixed
 * device array of 16 MiB instead, so a level may hold at most 4194304 cells
 * (BSAM_EINVAL beyond that, with nothing run).  Same arguments, results and
 * side effects.  The array is one per GPU, shared by every context on it
 * and kept until the process ends: calls that overlap on one GPU (two host
 * threads, two contexts) race and must be serialised by the caller. */
int32_t bsam_estimate_level_errors_noalloc(bsam_ctx* ctx, int32_t level, int32_t errflagopt,
                                           double qtolerance, int32_t* flags, int64_t* ntagged);
/* Output frame: WriteQ (bsaminputoutput.f90:60-82) with OutputQ (:85-129)
 * and WriteQ2D/3D (:132-181), written to `path` (the reference's
 * './OUT/m%05d.dat' is the caller's choice) byte for byte as the reference
 * writes it: time, finestlevel, then every patch of levels 0..finest in
 * descriptor order with its header and q(0:mx+1, ...), values edited F25.12
 * as the Fortran runtime does.  geom holds per descriptor dx[3], xlower[3],
 * xupper[3] as nodeinfo has them (nodeinfodef.f90:118-220); NULL takes dx
 * and xlower from the descriptors and xupper = xlower + mx*dx.  Synchronous. */
int32_t bsam_write_frame(bsam_ctx* ctx, const char* path, double time, const double* geom);
/* Host self-test of the frame writer's F25.12 edit (no counterpart in the
 * reference): out receives 25 bytes per value, as bsam_write_frame writes
 * them; no GPU is used. */
int32_t bsam_selftest_f25(const double* x, int64_t n, char* out);

/* ---- coarse-grained entry points --------------------------------------- */
/* AFASVcycle(level) (afasroutines.f90:495-529). */
int32_t bsam_vcycle(bsam_ctx* ctx, int32_t level);
/* MultigridIterations (afasroutines.f90:48-110).  On return *ncycles holds
 * the number of V-cycles done and *residual the last ErrorAFAS value.
 * Returns BSAM_NOT_CONVERGED when maxvcycles were used up, BSAM_EDIVERGED
 * when the residual is not finite.  `history` (may be NULL) receives up to
 * `history_len` residuals, one per V-cycle. */
int32_t bsam_multigrid_iterations(bsam_ctx* ctx, int32_t* ncycles,
                                  double* residual, double* history,
                                  int32_t history_len);
/* One time step of the time loop on a frozen mesh (bsamroutines.f90:
 * 673-758): the ghost fills the regrid AMR(rootlevel) (:966-991) does on an
 * unchanged mesh, SetGhost(level, 0) from the root to the finest level;
 * FillDown on the first step; CopyQToQold on every level;
 * MultigridIterations.  With maxlevel = 0 this is the reference's uniform
 * time step exactly. */
int32_t bsam_time_step(bsam_ctx* ctx, int32_t first_step, int32_t* ncycles,
                       double* residual);

/* The field set-up of a regrid onto a given new hierarchy (GridAdapt,
 * bsamroutines.f90:1663-1951, called from AMR at :966-991): replaces the
 * installed hierarchy like bsam_set_hierarchy, but carries the fields.  The
 * levels at and below the root must keep their one patch and extent (their
 * q, qold are kept); then level by level from 1 up, the new patches get
 * zero f and rf (MakeNewGrid), InitFields' mass-corrected P1 prolongation
 * of their parents with qold = q (:1787-1856), TransferValues' copy of q
 * from the old patches of the same level where their global boxes overlap
 * (:1859-1951), and SetGhost(level, 0).  Launch shapes set for a level are
 * kept while the level exists.  A descriptor table refused by the checks
 * above leaves the context as it was; a later failure leaves it with no
 * hierarchy. */
int32_t bsam_regrid(bsam_ctx* ctx, int32_t npatch, const bsam_patch_desc* desc,
                    int32_t minlevel, int32_t finestlevel);

/* The new patches of level+1 from the tags of `level` (host only, no
 * GPU): GridAdapt's RefineGrid over the level (bsamroutines.f90:1366-1405).
 * `flags` are the level's tags as bsam_estimate_level_errors returns them
 * (the level's patches in descriptor order, each column-major over mx); a
 * copy is buffered by `ibuffer` cells (BufferAndList :1206-1247 inside a
 * patch, InflateEdgeTags :1267-1349 across patches and periodic images),
 * every patch of the level is split by NewSubGrids (:1408-1605, Berger-
 * Rigoutsos signatures with the reference's trimming, hole and inflection
 * rules, maxsubgrids 1024, 15 passes) with `minimumgridpoints` (even, >= 4)
 * and `desiredfillratio`, and each subgrid becomes MakeNewGrid's child
 * descriptor (:1663-1735) with `parent` = its parent's index in `desc`.
 * The descriptors go to out[0..*nout) in the order the reference's tree
 * visits them (the youngest first).  BSAM_EINVAL, with *nout the count,
 * when they do not fit in `maxout`. */
int32_t bsam_new_level(int32_t ndims, int32_t npatch, const bsam_patch_desc* desc, int32_t level,
                       const int32_t* flags, int32_t ibuffer, int32_t minimumgridpoints,
                       double desiredfillratio, int32_t maxout, bsam_patch_desc* out,
                       int32_t* nout);

/* Grid parameters of AMR, per level 0..maxlevel-1: the rundata NAMELIST
 * griddata (bsamroutines.f90:243-258) as the regrid reads them. */
#define BSAM_MAX_AMR_LEVELS 16
typedef struct bsam_grid_params {
  int32_t maxlevel;                                /* finest level AMR may make */
  int32_t errflagopt[BSAM_MAX_AMR_LEVELS];         /* 1, 10 or 11 (bsam_estimate_level_errors) */
  int32_t ibuffer[BSAM_MAX_AMR_LEVELS];            /* tag buffer width, >= 0 */
  int32_t minimumgridpoints[BSAM_MAX_AMR_LEVELS];  /* even, >= 4 */
  double  qtolerance[BSAM_MAX_AMR_LEVELS];
  double  desiredfillratios[BSAM_MAX_AMR_LEVELS];
} bsam_grid_params;
/* AMR(rootlevel) (bsamroutines.f90:966-991), the regrid that opens every
 * time step of the reference's loop (:672-688), on the device: level by
 * level from the root, SetGhost(level, 0); below maxlevel the tagging of
 * bsam_estimate_level_errors (with its effect on q and qold for errflagopt
 * 1), the new patches of level+1 (bsam_new_level), and their fields
 * (InitFields from the parents, TransferValues from the mesh this call
 * replaces); the pass ends at the first level that makes no new patches.
 * The levels at and below the root are kept.  The new mesh replaces the
 * installed one (bsam_get_hierarchy reads it); levels that a pass leaves
 * as they were keep their device tables and launch shapes.  On failure the
 * context may hold no hierarchy. */
int32_t bsam_amr(bsam_ctx* ctx, const bsam_grid_params* gp);
/* The installed hierarchy: *npatch (and *minlevel, *finestlevel when not
 * NULL); with `out` non-NULL, the descriptors in installed order
 * (BSAM_EINVAL when they do not fit in maxout). */
int32_t bsam_get_hierarchy(bsam_ctx* ctx, int32_t maxout, bsam_patch_desc* out, int32_t* npatch,
                           int32_t* minlevel, int32_t* finestlevel);
/* One pass of the reference's time loop (bsamroutines.f90:672-758):
 * bsam_amr, FillDown on the first step (firstamr), CopyQToQold on every
 * level, MultigridIterations; returns as bsam_multigrid_iterations. */
int32_t bsam_time_step_amr(bsam_ctx* ctx, const bsam_grid_params* gp, int32_t first_step,
                           int32_t* ncycles, double* residual);

/* ReadQ's parse of a restart frame (bsaminputoutput.f90:184-443), host
 * only: the frame WriteQ (bsam_write_frame) writes, for a run of `ndims`
 * and `nrvars`.  *time and *finestlevel are the frame's; the patches of
 * levels 0..finest go to desc[0..*npatch) in the order the reference's tree
 * visits them after ReadQ (the root, then each level's patches from the
 * last in the file), with `parent` the index in this table of the patch one
 * level down that contains the patch's coarse box and mbounds from it (the
 * reference leaves both broken), mthbc 999 and the root's parent -1; q of
 * every patch (q(0:mx+1, ...) column-major, variable slowest) goes to q,
 * patch after patch in that order.  With desc and q NULL only *npatch and
 * *nq are returned; BSAM_EINVAL when they do not fit maxpatch / maxq, when
 * the file is not such a frame (the message names the line), or when a
 * refined patch lies in no patch of the level below. */
int32_t bsam_parse_frame(const char* path, int32_t ndims, int32_t nrvars, double* time, int32_t* finestlevel,
                         int32_t maxpatch, bsam_patch_desc* desc, int32_t* npatch, double* q, int64_t maxq,
                         int64_t* nq);
/* ReadQ onto the installed run: the frame's root must match the installed
 * root as the reference checks it (nrvars, dx, xlower within 1e-8, mx,
 * mglobal); the levels below the root stay; the frame's refined patches
 * (bsam_parse_frame) replace the levels above it, with MakeNewGrid's
 * inheritance of physical and periodic faces; q of the root and of every
 * refined patch is the frame's, every other field zero.  The reference
 * regrids before its next solve (AMR opens every step), which bsam_amr /
 * bsam_time_step_amr do with this mesh as the source of TransferValues. */
int32_t bsam_read_frame(bsam_ctx* ctx, const char* path, double* time, int32_t* finestlevel);

/* Wait for all work queued on the context's stream. */
int32_t bsam_synchronize(bsam_ctx* ctx);
/* Message of the last failed call on this thread ("" if none). */
const char* bsam_last_error(void);

/* ---- instrumentation (no counterpart in the reference, which only times
 * the whole run with CPU_TIME, bsamdriver.f90:56-60) -------------------- */
#define BSAM_KSTAT_RELAX_FINEST   0   /* relax sweeps on the finest level */
#define BSAM_KSTAT_RESID_FINEST   1   /* fused residual-norm on the finest level */
#define BSAM_KSTAT_FASRHS_FINEST  2   /* fused FAS right-hand side, finest -> next */
#define BSAM_KSTAT_COUNT          3
/* Enable (1) or disable (0) HIP-event timing of selected kernels; enabling
 * resets the counters. */
int32_t bsam_kstat_enable(bsam_ctx* ctx, int32_t on);
/* Launch count, summed kernel milliseconds and algorithmic bytes per launch
 * of one instrumented kernel class. */
int32_t bsam_kstat_read(bsam_ctx* ctx, int32_t which, int64_t* launches,
                        double* total_ms, double* bytes_per_launch);
/* Digest of a level's flattened ghost tables (a test hook): for each table
 * m = 0..6, out[5m .. 5m+4] = copy rules, interpolation rules, the copy and
 * interpolation rules of the first launch, and a 64-bit FNV-1a hash of the
 * tables' bytes (copies, post-operations, interpolation rules).
 * BSAM_ESTATE when the level has no flattened tables. */
int32_t bsam_ghost_table_digest(bsam_ctx* ctx, int32_t level, int64_t* out);
/* Wall-clock phases of a time step with regrid.  While enabled, the
 * context's stream is synchronised at every phase boundary, so the GPU work
 * queued in a phase is counted in it; graph capture is unaffected. */
#define BSAM_PHASE_AMR_TAG      0   /* SetGhost and EstimateLevelErrors per level */
#define BSAM_PHASE_AMR_CLUSTER  1   /* bsam_new_level: buffering, clustering */
#define BSAM_PHASE_AMR_INSTALL  2   /* installs of the new levels (tables, pools) */
#define BSAM_PHASE_AMR_FIELDS   3   /* InitFields, TransferValues, pool carries */
#define BSAM_PHASE_SOLVE        4   /* CopyQToQold and MultigridIterations */
#define BSAM_PHASE_GRAPH        5   /* capture and instantiation of the iteration graph (inside SOLVE) */
#define BSAM_PHASE_COUNT        6
/* Enable (1) or disable (0) the phase clocks; enabling resets them. */
int32_t bsam_phase_enable(bsam_ctx* ctx, int32_t on);
/* Times entered and summed milliseconds of one phase. */
int32_t bsam_phase_read(bsam_ctx* ctx, int32_t which, int64_t* calls, double* total_ms);

/* ---- launch-shape autotuning (no counterpart in the reference, whose
 * per-patch loops have no launch shape) --------------------------------
 * Times the candidate launch shapes of every level with at least 32768
 * cells on scratch copies of the fields and keeps the fastest; every shape
 * gives bit-identical fields, so tuning changes speed only.  Knobs: the
 * block order of a level's kernels and of its ghost gather in FP64 and
 * FP32 (one contiguous range of blocks per XCD, or runs of 1 to 256
 * consecutive blocks per XCD in turn; 1 is the hardware's round-robin
 * order), and on one large
 * patch the planes per workgroup of the marching sweeps (FP64, FP32) and
 * residuals.  Call after bsam_set_hierarchy (and bsam_upload),
 * outside graph capture; a captured iteration is re-captured at the next
 * call.  Synchronous.  The fields are not changed. */
int32_t bsam_autotune(bsam_ctx* ctx);
/* JSON report of the last bsam_autotune: per tuned level the patch count,
 * cells, a histogram of the largest patch extent, each knob's candidates
 * with their mean launch time in microseconds, and the choice ("{}" before
 * any tuning).  BSAM_EINVAL if `len` bytes do not hold it. */
int32_t bsam_tuning_report(const bsam_ctx* ctx, char* buf, int32_t len);

/* The same knobs set or read one at a time (no counterpart in the
 * reference).  A knob only moves work between workgroups, so every value
 * gives bit-identical fields; tests use these calls to check a value the
 * tuner may pick against the oracle.  Per-level knobs take the level; a
 * captured graph is re-captured at the next call.  Outside graph capture. */
#define BSAM_KNOB_ORDER              0  /* block order of the level kernels: 0 one range of
                                           blocks per XCD, c >= 1 runs of c blocks per XCD */
#define BSAM_KNOB_GATHER_ORDER_FP64  1  /* block order of the level's FP64 ghost gather */
#define BSAM_KNOB_GATHER_ORDER_FP32  2  /* block order of the level's FP32 ghost gather */
#define BSAM_KNOB_SWEEP_PLANES_FP64  3  /* one large 3D patch: planes per workgroup of the
                                           marching FP64 sweep (0: the default rule) */
#define BSAM_KNOB_SWEEP_PLANES_FP32  4  /* the same for the FP32 correction sweep */
#define BSAM_KNOB_RESID_PLANES       5  /* the same for the marching residual */
#define BSAM_KNOB_CHAIN_CELLS        6  /* whole hierarchy (level ignored): the largest level,
                                           in cells, at which the one-workgroup coarse V-cycle
                                           chain starts (0: no chain); defaults 4096 (CH) and
                                           512 (tumour model) */
#define BSAM_KNOB_REGRID_TABLES      7  /* whole hierarchy (level ignored): 1 (default), the
                                           levels bsam_regrid and bsam_amr build get the
                                           flattened ghost tables; 0, they keep the staged
                                           fill (interpolation rules, then box stages), which
                                           is planned in a fraction of the time, for meshes
                                           regridded every step */
#define BSAM_KNOB_GATHER_BLOCK_FP64  8  /* workgroup size of the level's FP64 ghost gather:
                                           0 (256), 64, 128, 256, 512 or 1024 */
#define BSAM_KNOB_GATHER_BLOCK_FP32  9  /* the same for the FP32 gather */
#define BSAM_KNOB_TABLE_PLANNER     10  /* whole hierarchy (level ignored): who plans the
                                           flattened ghost tables of the levels above the
                                           root that an install builds (batched launches):
                                           1 (default) the device, 0 the host; the tables
                                           are the same rule for rule */
#define BSAM_KNOB_SWEEP_IMAGES      11  /* whole hierarchy (level ignored): 1 (default), a
                                           colour sweep of a level with flattened tables that
                                           is followed by a reduced fill writes the fill's
                                           copies of the cells it updates (their ghost images
                                           in sibling and periodic partners) itself, and the
                                           gather does the rest; 0, the gather does it all.
                                           0 takes effect at once; 1 applies to levels
                                           installed after it is set.  Levels that bsam_amr
                                           builds get no images (they are rebuilt every
                                           step, and building the tables costs more than
                                           one step's sweeps save). */
int32_t bsam_set_knob(bsam_ctx* ctx, int32_t knob, int32_t level, int32_t value);
int32_t bsam_get_knob(bsam_ctx* ctx, int32_t knob, int32_t level, int32_t* value);

/* ---- failure handling of the mixed path (no counterpart in the reference,
 * whose loop is FP64 only and exits silently at maxvcycles,
 * afasroutines.f90:79-108) ---------------------------------------------
 * With BSAM_PREC_MIXED, from the second outer iteration on, an iteration
 * whose residual exceeds stall_ratio times the previous one (the FP32
 * correction cycle has stopped converging) hands the rest of that
 * MultigridIterations call to the FP64 FAS V-cycle, from the current
 * solution.  Default 0.9; 0 forces the hand-over after the second
 * iteration (a test hook).  bsam_mixed_fallbacks counts the calls that
 * handed over since the context was created. */
int32_t bsam_set_mixed_fallback(bsam_ctx* ctx, double stall_ratio);
int32_t bsam_mixed_fallbacks(bsam_ctx* ctx, int64_t* count);

/* ---- diagnostics (no counterpart in the reference) ----------------------
 * Host-only check of the flattened ghost-fill tables that per-level
 * launches use in place of the staged SetGhost (boundary.f90:67-97); needs
 * no GPU.  For every level > 0 of the hierarchy, on seeded random host
 * data and for the solution and correction forms of the boundary
 * conditions: (1) each table SetGhost(level, ipass), ipass = 0, 1, 2,
 * gives the staged fill's values bit for bit; (2) a red-black sweep
 * sequence using the reduced tables, whose last SetGhost is full, leaves
 * every face ghost the next sweep reads, and at its end every cell, equal
 * to the staged fill's.  A level whose fill cannot be flattened (a stage
 * reading an edge or corner ghost's value from before the call, as 3D
 * physical corners do) keeps the staged fill and is skipped.  *checked
 * receives the number of values compared, *levels_flat the number of
 * levels checked.  Returns BSAM_EINVAL with a message naming the first
 * difference. */
int32_t bsam_selftest_ghost_maps(const bsam_params* params, int32_t npatch,
                                 const bsam_patch_desc* desc, int32_t minlevel,
                                 int32_t finestlevel, uint64_t seed,
                                 int64_t* checked, int32_t* levels_flat);

/* Host self-test of the library's thread pool (no counterpart in the
 * reference; host planning and clustering run on it): `rounds` rounds of
 * jobs of 1..n_max tasks, the task counts chosen so that the pool grows
 * between jobs, and jobs that throw, each checking that every task ran
 * exactly once and that no task was still running when its job returned.
 * *jobs receives the number of jobs run.  No GPU is used.  Returns
 * BSAM_EINVAL with a message on the first violation. */
int32_t bsam_selftest_thread_pool(int32_t rounds, int32_t n_max, int64_t* jobs);

#ifdef __cplusplus
}
#endif
#endif /* BSAM_GPU_H */
