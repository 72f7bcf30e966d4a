// Device-side data layout and helpers shared by the HIP kernels.
//
// Every cell-centred field of a patch (q, qold, f, rf, ftmp) is stored with
// the ghost-inclusive extents of q, (mx+2) x (my+2) [x (mz+2)], and split
// by cell colour: for variable v, colour c = (i+j+k) mod 2 (k omitted in
// 2D), plane k and row j, the cells of colour c sit contiguously at half
// index m = i/2.  A red-black sweep therefore reads and writes whole cache
// lines of each colour array instead of every other element, and the
// x-neighbours of a colour-c cell are the elements m-1/m or m/m+1 of the
// other colour's row.  The coarse-under-fine array qc uses the same scheme
// with its own extents.  Interior-only reference arrays (f, rf, ftmp,
// bsamstorage.f90:81-91) simply leave their ghost slots unused.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bsam {

constexpr int kMaxPatchDims = 3;

// One patch as the kernels see it.
struct DPatch {
  int mx[3];    // interior cells (1 in unused dims)
  int nq[3];    // q storage extents (mx+2; 1 for the unused 3rd dim in 2D)
  int hq;       // half-row length of the q geometry, (nq[0]+1)/2
  int cmx[3];   // qc interior (AllocFields: mx/2, or 1 for odd mx)
  int nc[3];    // qc storage extents
  int hc;       // half-row length of qc
  int parent;   // descriptor index of the parent, -1 on the coarsest level
  int level;
  int mb[3][2]; // mbounds (parent-local box covered by this patch)
  int mthbc[6];
  int qoff;     // element offset of variable 0 in the q-geometry pools
  int coff;     // element offset of variable 0 in the qc pool
  int qvs;      // variable stride in the q-geometry pools
  int cvs;      // variable stride in the qc pool
  double dx;
};

// Offset of cell (i,j,k) (reference indices, k = 0 in 2D) of variable 0.
__host__ __device__ __forceinline__ int qidx(const DPatch& P, int i, int j, int k) {
  return P.qoff + (((((i + j + k) & 1) * P.nq[2] + k) * P.nq[1] + j) * P.hq + (i >> 1));
}
__host__ __device__ __forceinline__ int cidx(const DPatch& P, int i, int j, int k) {
  return P.coff + (((((i + j + k) & 1) * P.nc[2] + k) * P.nc[1] + j) * P.hc + (i >> 1));
}

// Ghost-fill box operation: dst(patch dp, box dlo+ext) <- op(src(patch sp,
// box slo+ext)), all variables.  Indices are storage indices (k = 0 in 2D).
// kTumourBC: variables 0, 1 copied, 2 <- 2 - x, 3 <- -x.
enum BoxOp : int { kCopy = 0, kCopyPlus0 = 1, kDirichlet2 = 2, kBgMixed = 3, kTumourBC = 4 };
struct BoxDesc {
  int dp, sp, op, pad;
  int dlo[3], slo[3], ext[3];
};

// Model constants of one launch (the reference's module globals dt, eps2,
// gamma, lambda, czero: nodeinfodef.f90:81, CH/problemdef.f90:5,
// BioGrowth/problemdef.f90:5).
struct ModelK {
  double dt, eps2, gam, lam, cz;
};

}  // namespace bsam
