/*
 * bsam_oracle.h -- CPU restatement of BSAM 1.2's FAS V-cycle (TEST
 * INFRASTRUCTURE ONLY).
 *
 * This library is the checker for the HIP implementation: the parity tests,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py call it.
 * Nothing in the product links or calls it.  Every routine cites the
 * reference file:line it restates.  Arrays are held exactly as the
 * reference holds them (bsamstorage.f90:54-144): per patch, column-major,
 * ghost layer of width 1, variable index slowest.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   - against the unmodified reference built from /root/reference with
 *     amdflang (oracle/Makefile, oracle/_ref): residual histories and
 *     per-sweep / per-V-cycle snapshots committed under tests/golden/;
 *   - closed forms: linear fields reproduced by every ghost interpolation
 *     and by prolongation, Fourier eigenvalues of the undivided Laplacian;
 *   - invariants: a relaxation sweep is stationary at f = L(q); mass
 *     conservation of a converged CH step on a periodic mesh;
 *   - regrid tagging (orc_estimate_level_errors): closed forms of the
 *     second-difference test and of the truncation error on linear fields;
 *   - output frame (orc_write_frame): the reference's own frames byte for
 *     byte and the Fortran runtime's F25.12 edit of edge values.
 */
#ifndef BSAM_ORACLE_H
#define BSAM_ORACLE_H

#include <stdint.h>
#include "../include/bsam_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

orc_ctx* orc_create(const bsam_params* p);
void     orc_destroy(orc_ctx* c);
void     orc_set_dt(orc_ctx* c, double dt);
int32_t  orc_set_hierarchy(orc_ctx* c, int32_t npatch, const bsam_patch_desc* d,
                           int32_t minlevel, int32_t finestlevel);
int64_t  orc_field_size(orc_ctx* c, int32_t patch, int32_t field);
double*  orc_field_ptr(orc_ctx* c, int32_t patch, int32_t field);

int32_t  orc_relax_level(orc_ctx* c, int32_t level, int32_t ipass);
int32_t  orc_set_ghost(orc_ctx* c, int32_t level, int32_t ipass);
int32_t  orc_level_relax(orc_ctx* c, int32_t level);
int32_t  orc_restrict_solution(orc_ctx* c, int32_t level);
int32_t  orc_coarse_ghost_points(orc_ctx* c, int32_t level);
int32_t  orc_coarse_loading(orc_ctx* c, int32_t level);
int32_t  orc_correct_fine(orc_ctx* c, int32_t level);
int32_t  orc_fill_down(orc_ctx* c, int32_t field);
int32_t  orc_get_source(orc_ctx* c, int32_t level);
int32_t  orc_update_source(orc_ctx* c, int32_t level);
int32_t  orc_residual(orc_ctx* c, int32_t level);
int32_t  orc_error_afas(orc_ctx* c, int32_t level, double* norm);
int32_t  orc_copy_q_to_qold(orc_ctx* c, int32_t level);
int32_t  orc_write_frame(orc_ctx* c, const char* path, double time, const double* geom);
void     orc_format_f25(const double* x, int64_t n, char* out);
int32_t  orc_estimate_level_errors(orc_ctx* c, int32_t level, int32_t errflagopt, double qtolerance,
                                   int32_t* flags, int64_t* ntagged);
int32_t  orc_vcycle(orc_ctx* c, int32_t level);
int32_t  orc_multigrid_iterations(orc_ctx* c, int32_t* ncycles, double* residual,
                                  double* history, int32_t history_len);
int32_t  orc_regrid(orc_ctx* c, int32_t npatch, const bsam_patch_desc* d,
                    int32_t minlevel, int32_t finestlevel);
int32_t  orc_time_step(orc_ctx* c, int32_t first_step, int32_t* ncycles,
                       double* residual);
const char* orc_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
