// Mixed-precision (FP32 V-cycles inside an FP64 defect-correction loop)
// entry points, implemented in bsam_gpu.hip next to the FP64 path.
#pragma once
#include <stdint.h>

struct bsam_ctx;

namespace bsam {
// FP32 pools of the correction V-cycle (same element offsets as the FP64
// pools of the q and qc geometries).
struct MixedState {
  float* e = nullptr;     // correction
  float* rhs = nullptr;   // right-hand side of the correction equation
  float* coef = nullptr;  // 3 phi^2 of the FP64 iterate
  float* ec = nullptr;    // coarse-under-fine correction
};
int32_t mixed_init(bsam_ctx* c);
void mixed_free(MixedState* m);
int32_t mixed_multigrid_iterations(bsam_ctx* c, int32_t* ncycles, double* residual, double* history,
                                   int32_t history_len);
}  // namespace bsam
