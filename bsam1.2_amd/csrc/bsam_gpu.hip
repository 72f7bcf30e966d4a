// bsam_gpu.hip -- MI355X (gfx950) implementation of BSAM 1.2's FAS V-cycle
// behind the C ABI of include/bsam_gpu.h.
//
// One launch covers all patches of a level: threads are numbered over the
// level's cells (flat, patch after patch, prefix sums per level) and find
// their patch by a binary search of the prefix array, so 8^3 and 256^3
// patches share the same kernels.  Block indices are remapped so that each
// of the 8 XCDs works on a contiguous range of cells (its own L2 then holds
// the neighbouring rows and planes of the stencil).  Ghost filling is a
// per-patch interpolation kernel plus a list of box copies per stage, in
// the reference's order (boundary.f90:67-97).
//
// Floating point: the file is compiled with -ffp-contract=off and every
// expression keeps the reference's evaluation order, so results are the
// reference's to the last bit except for the residual-norm sum.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <execinfo.h>
#include <signal.h>
#include <unistd.h>

#include <algorithm>
#include <array>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "../../include/bsam_gpu.h"
#include "bsam_grid.h"
#include "bsam_kernels.h"
#include "bsam_mixed.h"

using namespace bsam;

// ============================================================================
// error handling
// ============================================================================
static thread_local std::string g_err;

static int32_t set_err(int32_t code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}
int32_t bsam::fail(int32_t code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return set_err(BSAM_EHIP, "%s failed: %s", #expr, hipGetErrorString(e_));           \
  } while (0)

#define RETCHK(expr)              \
  do {                            \
    int32_t r_ = (expr);          \
    if (r_ < 0) return r_;        \
  } while (0)

extern "C" const char* bsam_last_error(void) { return g_err.c_str(); }

// ============================================================================
// device helpers
// ============================================================================
namespace {

constexpr int kBlock = 256;
constexpr int kNvMax = 4;  // fields of the largest model (the tumour model)
constexpr int kSeqBoxElems = 8192;  // box values up to which a level's stages run in one workgroup
// Largest level (cells) at which the tumour model's one-workgroup coarse
// chain starts.  On the 3D tumour AMR mesh (64^3 root) starting at 8^3 ran
// 65.6 (mixed) / 54.3 (FP64) time-steps/s, at 16^3 64.8 / 53.1 and at 4^3
// 64.6 / 53.6: the 16^3 level's sweeps are faster as level launches.
constexpr long kTmChainCells = 512;
// The same for CH: 16^3 (2D: 64^2).  Starting at 8^3 ran the 3D AMR bench
// at 74.5 instead of 75.0 (mixed) time-steps/s.
constexpr long kChChainCells = 4096;

// Fixed device scratch of the allocation-free tagging call
// (bsam_estimate_level_errors_noalloc): per-patch flag offsets and the
// level's flags, 16 MiB (a level of about 4.19 M cells).  Each device that
// loads the library holds one copy until the process ends; host code uses
// the address hipGetSymbolAddress returns for the current device, never
// the name.
constexpr size_t kScratchBytes = size_t(16) << 20;
__device__ __align__(256) char scratch[kScratchBytes];

// Interpolation constants, boundary.f90:54-63.
__constant__ double cA[3] = {-1.0 / 5.0, 2.0 / 3.0, 8.0 / 15.0};
__constant__ double cB[3] = {5.0 / 32.0, 15.0 / 16.0, -3.0 / 32.0};
__constant__ double cC[4] = {9.0 / 16.0, 3.0 / 16.0, 3.0 / 16.0, 1.0 / 16.0};

// Logical block index with XCD-contiguous ranges: hardware places block b on
// XCD b % 8; blocks of one XCD get consecutive logical indices.
__device__ __forceinline__ int xcd_block() {
  const int nb = gridDim.x, b = blockIdx.x;
  const int q = nb >> 3, r = nb & 7, x = b & 7, s = b >> 3;
  return (x < r) ? x * (q + 1) + s : r * (q + 1) + (x - r) * q + s;
}

// Logical block index in one of the block orders the autotuner picks from:
// order 0 is xcd_block (one contiguous range per XCD); order c >= 1 gives
// each XCD runs of c consecutive logical blocks in turn (1 is the hardware's
// round-robin order), over whole groups of 8c blocks, the rest in order.
__device__ __forceinline__ int ordered_block(int order) {
  if (order == 0) return xcd_block();
  const int b = blockIdx.x, g8 = 8 * order;
  const int full = ((int)gridDim.x / g8) * g8;
  if (b >= full) return b;
  const int grp = b / g8, r = b - grp * g8;
  return grp * g8 + (r & 7) * order + (r >> 3);
}

__device__ __forceinline__ int find_patch(const int* __restrict__ prefix, int n, int g) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (prefix[mid] <= g) lo = mid; else hi = mid;
  }
  return lo;
}

// A kernel's domain over one level (half-rows, cells, coarse cells, ...):
// patch p owns the flat range [prefix[p], prefix[p] + cnt[p]), and every
// range starts on a block boundary, so a block of a whole-level launch lies
// in one patch and reads that patch's record with scalar loads.
struct LevelArgs {
  const DPatch* __restrict__ pt;
  const int* __restrict__ patches;  // descriptor indices of the level
  const int* __restrict__ prefix;   // block-aligned start of each patch's range
  const int* __restrict__ cnt;      // elements of each patch
  const int* __restrict__ bp;       // patch of each block (whole-level launches)
  int npatch;
  int total;  // end of the launch's range of the domain
  int gbase;  // start of the range: 0 for a whole level, prefix[p] for patch p alone
  int fixed;  // the patch of a one-patch launch, else -1
  int order;  // block order (ordered_block)
};

// Logical block index of a level launch (the level's tuned block order).
__device__ __forceinline__ int level_block(const LevelArgs& L) { return ordered_block(L.order); }

// floor(a / b) for a >= 0 and b >= 1: below 2^22 the hardware
// reciprocal's estimate, corrected by one step (its error is below one), in
// place of the integer division sequence, since the flat level kernels are
// bound by their address arithmetic; from 2^22 on (patches of more than 4 M
// cells of a colour), where the float estimate may be off by more, the
// integer division.
__device__ __forceinline__ int fdiv(int a, int b) {
  if (a >= (1 << 22)) return a / b;
  int q = (int)((float)a * __builtin_amdgcn_rcpf((float)b));
  const int r = a - q * b;
  if (r < 0) q--;
  else if (r >= b) q++;
  return q;
}

// Cell (i, j, k) of element `local` of a patch's kHalf range for colour c
// (half-row position t, then j, then k), with t; false for the padding
// element of an odd row length.
template <int DIM>
__device__ __forceinline__ bool half_cell(const DPatch& P, int c, int local, int& i, int& j, int& k, int& t) {
  const int cnt = (P.mx[0] + 1) >> 1;
  const int row = fdiv(local, cnt);
  t = local - row * cnt;
  if (DIM == 3) {
    const int kk = fdiv(row, P.mx[1]);
    j = row - kk * P.mx[1] + 1;
    k = kk + 1;
  } else {
    j = row + 1;
    k = 0;
  }
  i = 2 * t + 2 - ((c + j + k) & 1);
  return i <= P.mx[0];
}

// Offsets of a colour-c cell's stencil in the colour-split layout: the
// other colour's row of (j, k) starts at ro; the x-neighbours are ro + t
// and ro + t + 1, the y- and z-neighbours the own half index m = i >> 1 in
// the rows +-1 (+-hq) and planes +-1 (+-ny hq); the cell itself is in the
// colour-c block, one colour block (nz ny hq) away.
struct Stencil {
  int o, xm, xp, ym, yp, zm, zp;
};

// Post-operations of a flattened ghost rule, three bits each from the low
// end, applied in that order: 1 adds +0.0 (the qpo term of an x wrap), 2
// takes the Dirichlet image 2 - x (-x for a correction, `homog`), 3 does
// that on variables >= 1 only (the BioGrowth boundary condition), 4 is the
// tumour model's condition (2 - x on variable 2, -x on variable 3).
template <typename T>
__host__ __device__ __forceinline__ T ghost_post(T x, int post, int v, int homog) {
  for (; post; post >>= 3) {
    const int op = post & 7;
    if (op == 1) x = x + (T)0;
    else if (op == 2 || (op == 3 && v >= 1) || (op == 4 && v == 2)) x = homog ? -x : (T)2 - x;
    else if (op == 4 && v == 3) x = -x;
  }
  return x;
}

// Ghost images that a colour sweep of a level writes for the reduced fill
// that follows it: the fill's copy rules whose source is a cell the sweep
// updates, grouped by the sweep's logical workgroup.  Entry (dst, dst
// variable stride, thread of the workgroup whose cell is the source, the
// rule's post-operation chain); off[b] .. off[b + 1] are workgroup b's.
struct ImgTab {
  const int* __restrict__ off;
  const int4* __restrict__ ent;
  int homog;
};
// Writes a workgroup's images from the new values its threads left in
// `nv` (one row of kBlock per variable): every thread of the workgroup
// must call it (it waits at a barrier).
template <typename T, int NV>
__device__ __forceinline__ void write_images(const ImgTab& I, int lb, T (*nv)[kBlock], T* __restrict__ q) {
  __syncthreads();
  const int e1 = I.off[lb + 1];
  for (int e = I.off[lb] + (int)threadIdx.x; e < e1; e += kBlock) {
    const int4 r = I.ent[e];
#pragma unroll
    for (int v = 0; v < NV; v++) q[r.x + v * r.y] = ghost_post<T>(nv[v][r.z], r.w, v, I.homog);
  }
}
template <int DIM>
__device__ __forceinline__ Stencil stencil_of(const DPatch& P, int c, int i, int j, int k, int t) {
  const int hq = P.hq, ny = P.nq[1], blk = P.nq[2] * ny * hq;
  const int cc = c ^ 1, m = i >> 1;
  const int ro = P.qoff + ((cc * P.nq[2] + k) * ny + j) * hq;
  Stencil S;
  S.o = ro + (c - cc) * blk + m;
  S.xm = ro + t;
  S.xp = ro + t + 1;
  S.ym = ro - hq + m;
  S.yp = ro + hq + m;
  S.zm = (DIM == 3) ? ro - ny * hq + m : 0;
  S.zp = (DIM == 3) ? ro + ny * hq + m : 0;
  return S;
}

// Cell (i, j, k) of element `local` of a patch's kCell range, which runs
// colour by colour, then row by row (j fastest, then k), then along the
// colour's half-row, so that consecutive elements are consecutive cells of
// one colour array; false for the padding element of an odd row length.
template <int DIM>
__device__ __forceinline__ bool cell_of(const DPatch& P, int local, int& i, int& j, int& k) {
  const int hx = (P.mx[0] + 1) >> 1;
  const int per = hx * P.mx[1] * (DIM == 3 ? P.mx[2] : 1);
  const int c = local >= per;
  int t;
  return half_cell<DIM>(P, c, local - c * per, i, j, k, t);
}

// Patch (position in the level's list) and local index of element
// gbase + lb*kBlock + threadIdx.x; false for the padding after a patch's
// range and past the launch's end.
__device__ __forceinline__ bool level_elem(const LevelArgs& L, int lb, int& pi, int& local) {
  const int g = L.gbase + lb * kBlock + threadIdx.x;
  if (g >= L.total) return false;
  pi = (L.fixed >= 0) ? L.fixed : L.bp[lb];
  local = g - L.prefix[pi];
  return local < L.cnt[pi];
}

// ---------------------------------------------------------------------------
// CH relaxation: RelaxGrid2D/3D, CH/problem.f90:318-369, 554-614.
// Domain: cells of colour c, (mx+1)/2 per row.  Colour c relaxes cells with
// (i+j+k) mod 2 == c; the reference's ipass = c+1 (problem.f90:350, 591).
// ---------------------------------------------------------------------------
// Ghost images of an updated interior cell of a patch that is periodic in
// every dimension (the net effect of SelfPeriodicBC, boundary.f90:480-577,
// on the colour just relaxed): every ghost whose periodic image is (i,j,k)
// takes its value, plus 0.0 on the x ghost planes (the qpo term).  The
// images have the cell's colour because mx is even.
template <int DIM, typename T>
__device__ __forceinline__ void wrap_images(T* __restrict__ q, const DPatch& P, int i, int j, int k, T v0,
                                            T v1) {
  const int mx = P.mx[0], my = P.mx[1], mz = P.mx[2];
  const bool ex = (i == 1 || i == mx), ey = (j == 1 || j == my), ez = (DIM == 3) && (k == 1 || k == mz);
  if (!(ex || ey || ez)) return;
  const int gi = (i == 1) ? mx + 1 : 0, gj = (j == 1) ? my + 1 : 0, gk = (k == 1) ? mz + 1 : 0;
  for (int a = 0; a <= (ex ? 1 : 0); a++)
    for (int b = 0; b <= (ey ? 1 : 0); b++)
      for (int d = 0; d <= (ez ? 1 : 0); d++) {
        if (a + b + d == 0) continue;
        const int ii = a ? gi : i, jj = b ? gj : j, kk = d ? gk : k;
        const int o = qidx(P, ii, jj, kk);
        q[o] = a ? v0 + (T)0 : v0;
        q[o + P.qvs] = a ? v1 + (T)0 : v1;
      }
}

// One cell of the CH correction smoother: RelaxGrid2D/3D's pointwise Newton
// step (CH/problem.f90:318-369, 554-614) written for the correction e of
// the base state q0 (phi0 = p0, the FP64 iterate the cycle started from).
// With phi = p0 + e_phi the Newton coefficient is -3 phi^2 and the cubic
// part of its right-hand side, -2 phi^3 minus what the residual of q0
// already holds, is -e_phi^2 (3 p0 + 2 e_phi); so the FP32 sweep takes the
// same step as the FP64 sweep on q, up to rounding of the correction.
// ec0 is the cell's own e_phi, rhs the residual f - L(q0).
__device__ __forceinline__ void ch_corr_update(float p0, float ec0, float r0, float r1, float smu, float sphi,
                                               float tmp1, float tmp2, float tmp3, float a12, float& n0,
                                               float& n1) {
  const float ph = p0 + ec0;
  const float a21 = -3.0f * (ph * ph) - tmp3;
  const float b1 = r0 + tmp1 * smu;
  const float b2 = r1 - tmp2 * sphi - (ec0 * ec0) * (3.0f * p0 + 2.0f * ec0);
  const float det = 1.0f - a12 * a21;
  n0 = (b1 - a12 * b2) / det;
  n1 = (b2 - a21 * b1) / det;
}

// The CH operator's action on the correction, N(e) = L(q0 + e) - L(q0):
// N1 = e_phi - t1 ULap(e_mu), N2 = e_mu - e_phi (3 p0^2 + 3 p0 e_phi +
// e_phi^2) + t2 ULap(e_phi), the cubic difference expanded so that no
// cancellation of O(1) terms happens in FP32.
__device__ __forceinline__ void ch_corr_op(float p0, float ep, float em, float lphi, float lmu, float t1,
                                           float t2, float& a1, float& a2) {
  a1 = ep - t1 * lmu;
  a2 = em - ep * (3.0f * (p0 * p0) + ep * (3.0f * p0 + ep)) + t2 * lphi;
}

template <int DIM, bool WRAP, bool IMG = false>
__global__ __launch_bounds__(kBlock) void k_relax_ch(LevelArgs L, int c, double dt, double eps2,
                                                     double h, double* __restrict__ q,
                                                     const double* __restrict__ f, ImgTab I = ImgTab{}) {
  // IMG: the workgroup then writes the ghost images of its updated cells
  // (write_images), so no thread leaves before that
  const int lb = level_block(L);
  int pi, local;
  double n0 = 0.0, n1 = 0.0;
  bool ok = level_elem(L, lb, pi, local);
  if (!IMG && !ok) return;
  if (ok) {
    const DPatch P = L.pt[L.patches[pi]];
    int i, j, k, t;
    ok = half_cell<DIM>(P, c, local, i, j, k, t);
    if (!IMG && !ok) return;
    if (ok) {
      const double h2 = h * h;
      const double tmp1 = dt / h2, tmp2 = eps2 / h2;
      const double tmp3 = (DIM == 3 ? 6.0 : 4.0) * eps2 / h2;
      const double a12 = (DIM == 3 ? 6.0 : 4.0) * tmp1;
      const int vs = P.qvs;
      const Stencil S = stencil_of<DIM>(P, c, i, j, k, t);
      const int o = S.o;
      const double phi = q[o];
      const double f0 = f[o], f1 = f[o + vs];
      const int xm = S.xm, xp = S.xp, ym = S.ym, yp = S.yp;
      double smu, sphi;
      if (DIM == 3) {
        const int zm = S.zm, zp = S.zp;
        smu = q[xm + vs] + q[xp + vs] + q[ym + vs] + q[yp + vs] + q[zm + vs] + q[zp + vs];
        sphi = q[xp] + q[xm] + q[yp] + q[ym] + q[zp] + q[zm];
      } else {
        smu = q[xm + vs] + q[xp + vs] + q[ym + vs] + q[yp + vs];
        sphi = q[xp] + q[xm] + q[yp] + q[ym];
      }
      const double a21 = -3.0 * (phi * phi) - tmp3;
      const double b1 = f0 + tmp1 * smu;
      const double b2 = f1 - tmp2 * sphi - 2.0 * (phi * phi * phi);
      const double det = 1.0 - a12 * a21;
      n0 = (b1 - a12 * b2) / det;
      n1 = (b2 - a21 * b1) / det;
      q[o] = n0;
      q[o + vs] = n1;
      if (WRAP) wrap_images<DIM, double>(q, P, i, j, k, n0, n1);
    }
  }
  if (IMG) {
    __shared__ double nv[2][kBlock];
    nv[0][threadIdx.x] = n0;
    nv[1][threadIdx.x] = n1;
    write_images<double, 2>(I, lb, nv, q);
  }
}

// 3D colour sweep of one large patch, marching in k.  A block owns a tile
// of 64 half-row positions x 16 rows and KC planes; the other colour's phi
// and mu of planes k-1, k, k+1 (tile plus one halo row each side and one
// halo position) sit in a 4-slot LDS ring, so every value of the other
// colour is read from HBM about once, and the own colour's phi and f once.
// LIN=false: CH RelaxGrid3D (CH/problem.f90:554-614) in FP64 with f.
// LIN=true:  the linearised correction smoother (FP32) with rhs and coef.
// WRAP: also write the periodic ghost images of the updated cells.
// FINEST only separates the finest level's launches in kernel traces.
// EMODE (LIN only) 1: the correction starts from zero, so the other
// colour's values are taken as zero and not read; 2: the last sweep of the
// cycle adds the correction into the FP64 solution q64 for both cells of
// each thread's pair (q64 += e, ghost images included) instead of storing e.
constexpr int kSweepTY = 16, kSweepThreads = 64 * kSweepTY;
template <typename T, bool LIN, bool WRAP, int FINEST, int EMODE = 0>
__global__ __launch_bounds__(kSweepThreads) void k_relax3d_march(const DPatch* __restrict__ pt, int pid, int c,
                                                                int KC, T tmp1, T tmp2, T tmp3, T a12,
                                                                T* __restrict__ q, const T* __restrict__ f,
                                                                const T* __restrict__ coef,
                                                                double* __restrict__ q64 = nullptr) {
  constexpr int TX = 64, TY = kSweepTY, NR = TY + 2, NP = TX + 1, NT = kSweepThreads;
  __shared__ T sh[4][2][NR][NP];
  const DPatch P = pt[pid];
  const int mx = P.mx[0], my = P.mx[1], mz = P.mx[2];
  const int nx2 = mx >> 1;
  const int ntx = (nx2 + TX - 1) / TX, nty = (my + TY - 1) / TY;
  int lb = xcd_block();
  const int bx = lb % ntx;
  lb /= ntx;
  const int by = lb % nty;
  const int bk = lb / nty;
  const int t0 = bx * TX, j0 = by * TY;
  const int k0 = bk * KC + 1, k1 = min(mz, k0 + KC - 1);
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int cc = c ^ 1;
  const int hq = P.hq, ny = P.nq[1], nz = P.nq[2], vs = P.qvs, base = P.qoff;
  auto rowoff = [&](int col, int k, int j) { return base + ((col * nz + k) * ny + j) * hq; };
  // One plane of the other colour's tile is NE entries; each thread owns
  // the entries threadIdx.x + NT*n.  Loads for plane k+3 are issued at the
  // end of iteration k into registers and stored to LDS during iteration
  // k+1, so their latency overlaps a plane of arithmetic.
  constexpr int NE = 2 * NR * NP;
  constexpr int NPER = (NE + NT - 1) / NT;
  auto fetch = [&](int k, T (&reg)[NPER]) {
#pragma unroll
    for (int n = 0; n < NPER; n++) {
      const int e = threadIdx.x + NT * n;
      T val = (T)0;
      if (EMODE != 1 && e < NE) {
        const int v = e / (NR * NP);
        const int r = (e / NP) % NR;
        const int p = e % NP;
        const int jj = j0 + r, pp = t0 + p;
        if (jj <= my + 1 && pp <= nx2) val = q[rowoff(cc, k, jj) + pp + v * vs];
      }
      reg[n] = val;
    }
  };
  auto store = [&](int k, const T (&reg)[NPER]) {
    T* sl = &sh[k & 3][0][0][0];
#pragma unroll
    for (int n = 0; n < NPER; n++) {
      const int e = threadIdx.x + NT * n;
      if (e < NE) sl[e] = reg[n];
    }
  };
  T pre[NPER];
  fetch(k0 - 1, pre);
  store(k0 - 1, pre);
  fetch(k0, pre);
  store(k0, pre);
  fetch(k0 + 1, pre);
  store(k0 + 1, pre);
  if (k0 + 2 <= k1 + 1) fetch(k0 + 2, pre);
  const int t = t0 + tx, j = j0 + 1 + ty;
  const bool active = (t < nx2) && (j <= my);
  // own-colour values of the current plane, prefetched one plane ahead:
  // f (or the correction's right-hand side), phi (or the base phi0) and,
  // for the correction, the cell's own e_phi
  T own0 = (T)0, own1 = (T)0, own2 = (T)0, own3 = (T)0;
  auto own_off = [&](int k) { return rowoff(c, k, j) + t + 1 - ((c + j + k) & 1); };
  if (active) {
    const int o = own_off(k0);
    own0 = f[o];
    own1 = f[o + vs];
    if (LIN) {
      own2 = coef[o];
      if (EMODE != 1) own3 = q[o];
    } else {
      own2 = q[o];
    }
  }
  __syncthreads();
  for (int k = k0; k <= k1; k++) {
    T nx0 = (T)0, nx1 = (T)0, nx2v = (T)0, nx3 = (T)0;
    if (active && k + 1 <= k1) {
      const int o = own_off(k + 1);
      nx0 = f[o];
      nx1 = f[o + vs];
      if (LIN) {
        nx2v = coef[o];
        if (EMODE != 1) nx3 = q[o];
      } else {
        nx2v = q[o];
      }
    }
    if (active) {
      const int ib = (c + j + k) & 1;
      const int m = t + 1 - ib;
      const int pr = tx + 1 - ib;
      const int sk = k & 3, skm = (k - 1) & 3, skp = (k + 1) & 3;
      const int r = ty + 1;
      const T smu = sh[sk][1][r][tx] + sh[sk][1][r][tx + 1] + sh[sk][1][r - 1][pr] + sh[sk][1][r + 1][pr] +
                    sh[skm][1][r][pr] + sh[skp][1][r][pr];
      const T sphi = sh[sk][0][r][tx + 1] + sh[sk][0][r][tx] + sh[sk][0][r + 1][pr] + sh[sk][0][r - 1][pr] +
                     sh[skp][0][r][pr] + sh[skm][0][r][pr];
      const int o = rowoff(c, k, j) + m;
      T n0, n1;
      if constexpr (LIN) {
        ch_corr_update(own2, own3, own0, own1, smu, sphi, tmp1, tmp2, tmp3, a12, n0, n1);
      } else {
        const T phi = own2;
        const T a21 = (T)-3 * (phi * phi) - tmp3;
        const T b1 = own0 + tmp1 * smu;
        const T b2 = own1 - tmp2 * sphi - (T)2 * (phi * phi * phi);
        const T det = (T)1 - a12 * a21;
        n0 = (b1 - a12 * b2) / det;
        n1 = (b2 - a21 * b1) / det;
      }
      if (EMODE == 2) {
        // own cell: q64 += the new correction; the pair's other cell:
        // q64 += its correction, final since the other colour's sweep
        const double a0 = q64[o] + (double)n0, a1 = q64[o + vs] + (double)n1;
        q64[o] = a0;
        q64[o + vs] = a1;
        if (WRAP) wrap_images<3, double>(q64, P, 2 * t + 2 - ib, j, k, a0, a1);
        const int io = 2 * t + 1 + ib;
        const int oo = rowoff(cc, k, j) + (io >> 1);
        const double b0 = q64[oo] + (double)sh[sk][0][r][tx + ib];
        const double b1 = q64[oo + vs] + (double)sh[sk][1][r][tx + ib];
        q64[oo] = b0;
        q64[oo + vs] = b1;
        if (WRAP) wrap_images<3, double>(q64, P, io, j, k, b0, b1);
      } else {
        q[o] = n0;
        q[o + vs] = n1;
        if (WRAP) wrap_images<3, T>(q, P, 2 * t + 2 - ib, j, k, n0, n1);
      }
    }
    if (k + 2 <= k1 + 1) {
      store(k + 2, pre);
      if (k + 3 <= k1 + 1) fetch(k + 3, pre);
    }
    own0 = nx0;
    own1 = nx1;
    own2 = nx2v;
    own3 = nx3;
    __syncthreads();
  }
}

// One cell of the tumour model's relaxation (the probe PROBLEM module's
// TumourRelax2D/3D): nutrient, then the phi/mu 2x2 Newton solve, then
// pressure, each with the updated nutrient.
template <int DIM>
__device__ __forceinline__ void relax_tm_cell(const DPatch& P, int i, int j, int k, const ModelK& m,
                                              double* __restrict__ q, const double* __restrict__ qold,
                                              const double* __restrict__ f) {
  const double h = P.dx, h2 = h * h, dd = (DIM == 3) ? 6.0 : 4.0;
  const double cdt = m.dt / h2, ceps = m.eps2 / h2, cdiag = dd * m.eps2 / h2, c12 = dd * cdt;
  const double chh = 1.0 / h2, cd = dd / h2;
  const int vs = P.qvs;
  const int o = qidx(P, i, j, k);
  const int xm = qidx(P, i - 1, j, k), xp = qidx(P, i + 1, j, k);
  const int ym = qidx(P, i, j - 1, k), yp = qidx(P, i, j + 1, k);
  const int zm = (DIM == 3) ? qidx(P, i, j, k - 1) : 0, zp = (DIM == 3) ? qidx(P, i, j, k + 1) : 0;
  // neighbour sums in the order i-1, i+1, j-1, j+1, k-1, k+1 (and i+1,
  // i-1, ... for phi), as TumourRelax writes them
  auto nb = [&](int off) {
    double s0 = q[xm + off] + q[xp + off] + q[ym + off] + q[yp + off];
    if (DIM == 3) s0 = s0 + q[zm + off] + q[zp + off];
    return s0;
  };
  double sphi = q[xp] + q[xm] + q[yp] + q[ym];
  if (DIM == 3) sphi = sphi + q[zp] + q[zm];
  const double t0 = 0.5 * (qold[o] + 1.0);
  const double n = (f[o + 2 * vs] + chh * nb(2 * vs)) / (cd + m.cz * t0);
  q[o + 2 * vs] = n;
  const double p = q[o];
  const double c21 = -3.0 * (p * p) - cdiag;
  const double r1 = f[o] + cdt * nb(vs) + m.dt * m.gam * t0 * n;
  const double r2 = f[o + vs] - ceps * sphi - 2.0 * (p * p * p);
  const double d = 1.0 - c12 * c21;
  q[o] = (r1 - c12 * r2) / d;
  q[o + vs] = (r2 - c21 * r1) / d;
  q[o + 3 * vs] = (f[o + 3 * vs] + chh * nb(3 * vs) + m.gam * t0 * n) / cd;
}

// The tumour model's relaxation (the probe PROBLEM module's TumourRelax2D/
// 3D): per cell of colour c, the nutrient n from its equation, the CH pair
// by RelaxGrid's Newton step with dt*lamp*t0*n in its first right-hand
// side, then the pressure; t0 = (1 + phi_old)/2.  ModelK: lamp in `gam`,
// kappa in `cz`.  Domain: cells of colour c, (mx+1)/2 per row.
template <typename T>
__host__ __device__ __forceinline__ T box_value(int op, int v, int homog, T x);

// The ghost images of interior cell (i, j, k) of a single patch with
// condition 10 on every face: every ghost cell whose mirrored interior
// cell (0 -> 1, mx+1 -> mx per dimension) is this one, set to the face
// operation applied once per dimension in which it lies outside, x first --
// the value SetPhysicalBC's staged faces give it.  A sweep that calls this
// for every cell it updates leaves every ghost as a full SetGhost would,
// provided the ghosts were consistent before: a ghost whose mirror cell the
// sweep did not touch still holds the image of its unchanged value.
template <int DIM, typename T>
__device__ __forceinline__ void phys_cell_images(const DPatch& P, int op, int i, int j, int k, int nv, int homog,
                                                 T* __restrict__ q) {
  int opt[3][3], no[3];
  const int c3[3] = {i, j, k};
  for (int a = 0; a < 3; a++) {
    no[a] = 1;
    opt[a][0] = c3[a];
    if (a < DIM) {
      if (c3[a] == 1) opt[a][no[a]++] = 0;
      if (c3[a] == P.mx[a]) opt[a][no[a]++] = P.mx[a] + 1;
    }
  }
  if (no[0] * no[1] * no[2] == 1) return;
  const int src = qidx(P, i, j, k);
  for (int z = 0; z < no[2]; z++)
    for (int y = 0; y < no[1]; y++)
      for (int x = 0; x < no[0]; x++) {
        if (x == 0 && y == 0 && z == 0) continue;
        const int d = qidx(P, opt[0][x], opt[1][y], opt[2][z]);
        for (int v = 0; v < nv; v++) {
          T w = q[src + v * P.qvs];
          if (x) w = box_value<T>(op, v, homog, w);
          if (y) w = box_value<T>(op, v, homog, w);
          if (z) w = box_value<T>(op, v, homog, w);
          q[d + v * P.qvs] = w;
        }
      }
}

// IMG: the patch is a single one with condition 10 on every face, and the
// sweep also writes the ghost images of the cells it updates
// (phys_cell_images), so no SetGhost need follow.
template <int DIM, bool IMG = false>
__global__ __launch_bounds__(kBlock) void k_relax_tm(LevelArgs L, int c, ModelK m, double* __restrict__ q,
                                                     const double* __restrict__ qold,
                                                     const double* __restrict__ f) {
  int pi, local;
  if (!level_elem(L, level_block(L), pi, local)) return;
  const DPatch P = L.pt[L.patches[pi]];
  const int cnt = (P.mx[0] + 1) >> 1;
  const int t = local % cnt;
  const int row = local / cnt;
  const int j = row % P.mx[1] + 1;
  const int k = (DIM == 3) ? row / P.mx[1] + 1 : 0;
  const int ib = (c + j + k) & 1;
  const int i = 2 * t + 2 - ib;
  if (i > P.mx[0]) return;
  relax_tm_cell<DIM>(P, i, j, k, m, q, qold, f);
  if (IMG) phys_cell_images<DIM, double>(P, kTumourBC, i, j, k, 4, 0, q);
}


// BioGrowth RelaxGrid2D, BioGrowth/problem.f90:255-296 (decoupled GS).
__global__ __launch_bounds__(kBlock) void k_relax_bg2d(LevelArgs L, int c, ModelK m, double h,
                                                       double* __restrict__ q,
                                                       const double* __restrict__ qold,
                                                       const double* __restrict__ f) {
  int pi, local;
  if (!level_elem(L, level_block(L), pi, local)) return;
  const DPatch P = L.pt[L.patches[pi]];
  const int cnt = (P.mx[0] + 1) >> 1;
  const int t = local % cnt;
  const int j = local / cnt + 1;
  const int ib = (c + j) & 1;
  const int i = 2 * t + 2 - ib;
  if (i > P.mx[0]) return;
  const double h2 = h * h;
  const double tmp1 = m.dt * m.eps2 / h2, tmp3 = 1.0 / h2;
  const int vs = P.qvs;
  const int o = qidx(P, i, j, 0);
  const int xm = qidx(P, i - 1, j, 0), xp = qidx(P, i + 1, j, 0);
  const int ym = qidx(P, i, j - 1, 0), yp = qidx(P, i, j + 1, 0);
  const double p1 = q[o];
  const double tmp2 = 4.0 * tmp1 + 1.0 + m.dt * (p1 * p1);
  const double po = qold[o];
  const double grow = m.gam * 0.5 * (po + 1.0);
  const double upt = m.lam * 0.5 * (po + 1.0);
  q[o] = (f[o] + m.dt * grow * (q[o + vs] - m.cz) + tmp1 * (q[xm] + q[xp] + q[ym] + q[yp])) / tmp2;
  const double tmp4 = 4.0 * tmp3 + upt;
  q[o + vs] = (f[o + vs] + tmp3 * (q[xp + vs] + q[xm + vs] + q[yp + vs] + q[ym + vs])) / tmp4;
}

// ---------------------------------------------------------------------------
// Operators (Operator2D/3D with ULap2D/3D): CH/problem.f90:393-422, 638-669,
// gridutilities.f90:50-72, 372-399; BioGrowth/problem.f90:320-353.  `po` is
// the qold value of the cell (BG only).
// ---------------------------------------------------------------------------
template <int DIM>
__device__ __forceinline__ double ulap(const double* __restrict__ a, const DPatch& P, int i, int j,
                                       int k, int off) {
  double s = a[qidx(P, i + 1, j, k) + off] + a[qidx(P, i - 1, j, k) + off] +
             a[qidx(P, i, j + 1, k) + off] + a[qidx(P, i, j - 1, k) + off];
  if (DIM == 3) {
    s = s + a[qidx(P, i, j, k + 1) + off];
    s = s + a[qidx(P, i, j, k - 1) + off];
    return s - 6.0 * a[qidx(P, i, j, k) + off];
  }
  return s - 4.0 * a[qidx(P, i, j, k) + off];
}
// Same on the qc geometry.
template <int DIM>
__device__ __forceinline__ double ulap_c(const double* __restrict__ a, const DPatch& P, int i, int j,
                                         int k, int off) {
  double s = a[cidx(P, i + 1, j, k) + off] + a[cidx(P, i - 1, j, k) + off] +
             a[cidx(P, i, j + 1, k) + off] + a[cidx(P, i, j - 1, k) + off];
  if (DIM == 3) {
    s = s + a[cidx(P, i, j, k + 1) + off];
    s = s + a[cidx(P, i, j, k - 1) + off];
    return s - 6.0 * a[cidx(P, i, j, k) + off];
  }
  return s - 4.0 * a[cidx(P, i, j, k) + off];
}

template <int DIM, int MODEL>
__device__ __forceinline__ void op_from(double p, double mu, double lphi, double lmu, double po,
                                        const ModelK& m, double h, double& o1, double& o2) {
  const double h2 = h * h;
  if (MODEL == BSAM_MODEL_CH) {
    const double tmp1 = m.dt / h2, tmp2 = m.eps2 / h2;
    o1 = p - tmp1 * lmu;
    o2 = mu - (p * p * p) + tmp2 * lphi;
  } else {
    const double tmp1 = m.dt * m.eps2 / h2, tmp3 = 1.0 / h2;
    const double grow = m.gam * 0.5 * (po + 1.0), upt = m.lam * 0.5 * (po + 1.0);
    o1 = p - tmp1 * lphi + m.dt * (p * p * p) - m.dt * grow * (mu - m.cz);
    o2 = -tmp3 * lmu + upt * mu;
  }
}

template <int DIM, int MODEL>
__device__ __forceinline__ void op_q(const double* __restrict__ q, const double* __restrict__ qold,
                                     const DPatch& P, int i, int j, int k, const ModelK& m, double h,
                                     double& o1, double& o2) {
  const int o = qidx(P, i, j, k);
  const double lmu = ulap<DIM>(q, P, i, j, k, P.qvs);
  const double lphi = ulap<DIM>(q, P, i, j, k, 0);
  const double po = (MODEL == BSAM_MODEL_BG) ? qold[o] : 0.0;
  op_from<DIM, MODEL>(q[o], q[o + P.qvs], lphi, lmu, po, m, h, o1, o2);
}

// The tumour model's operator (the probe PROBLEM module's Operator2D/3D
// with model = 3): with t0 = (1 + phi_old)/2,
//   o1 = phi - (dt/h^2) ULap(mu) - dt lamp t0 n
//   o2 = mu - phi^3 + (eps^2/h^2) ULap(phi)
//   o3 = -ULap(n)/h^2 + kappa t0 n
//   o4 = -ULap(p)/h^2 - lamp t0 n
// ModelK carries lamp in `gam`, lama in `lam` and kappa in `cz`.
template <int DIM>
__device__ __forceinline__ void op_tm_from(double p, double mu, double n, double l1, double l2, double l3,
                                           double l4, double po, const ModelK& m, double h, double* o) {
  const double h2 = h * h;
  const double cdt = m.dt / h2, ceps = m.eps2 / h2, chh = 1.0 / h2;
  const double t0 = 0.5 * (po + 1.0);
  o[0] = p - cdt * l2 - m.dt * m.gam * t0 * n;
  o[1] = mu - (p * p * p) + ceps * l1;
  o[2] = -chh * l3 + m.cz * t0 * n;
  o[3] = -chh * l4 - m.gam * t0 * n;
}

// Operator of any model at cell (i,j,k): nrvars outputs.
template <int DIM, int MODEL>
__device__ __forceinline__ void op_cell(const double* __restrict__ q, const double* __restrict__ qold,
                                        const DPatch& P, int i, int j, int k, const ModelK& m, double h,
                                        double* o) {
  if (MODEL == BSAM_MODEL_TUMOUR) {
    const int c0 = qidx(P, i, j, k), vs = P.qvs;
    const double l2 = ulap<DIM>(q, P, i, j, k, vs), l1 = ulap<DIM>(q, P, i, j, k, 0);
    const double l3 = ulap<DIM>(q, P, i, j, k, 2 * vs), l4 = ulap<DIM>(q, P, i, j, k, 3 * vs);
    op_tm_from<DIM>(q[c0], q[c0 + vs], q[c0 + 2 * vs], l1, l2, l3, l4, qold[c0], m, h, o);
  } else {
    op_q<DIM, MODEL>(q, qold, P, i, j, k, m, h, o[0], o[1]);
  }
}

// Residual (afasroutines.f90:817-857): rf = f - L(q) over the interior.
// Also accumulates sum(rf^2) per variable into per-block partials when
// `partial` is non-null (ErrorAFAS / L2Error / SquareSum, :113-233), and
// writes rf when `rf` is non-null.  Domain: kCell.
template <int DIM, int MODEL, typename TO>
__global__ __launch_bounds__(kBlock) void k_residual(LevelArgs L, ModelK m,
                                                     const double* __restrict__ q,
                                                     const double* __restrict__ qold,
                                                     const double* __restrict__ f,
                                                     TO* __restrict__ rf,
                                                     double* __restrict__ partial,
                                                     float* __restrict__ base = nullptr,
                                                     const float* __restrict__ eg = nullptr) {
  constexpr int NV = (MODEL == BSAM_MODEL_TUMOUR) ? 4 : 2;
  const int lb = level_block(L);
  double sv[kNvMax] = {};
  int pi, local;
  if (level_elem(L, lb, pi, local)) {
    const DPatch P = L.pt[L.patches[pi]];
    int i, j, k;
    if (cell_of<DIM>(P, local, i, j, k)) {
      double o[kNvMax];
      op_cell<DIM, MODEL>(q, qold, P, i, j, k, m, P.dx, o);
      const int c0 = qidx(P, i, j, k);
      // eg (CH, mixed path): the norm is that of the residual with the
      // ghosts the FAS cycle holds, q + eg on the ghost neighbours
      double dg[2] = {0.0, 0.0};
      if (MODEL == BSAM_MODEL_CH && eg) {
        const bool edge = i == 1 || i == P.mx[0] || j == 1 || j == P.mx[1] ||
                          (DIM == 3 && (k == 1 || k == P.mx[2]));
        if (edge) {
          const double h2 = P.dx * P.dx;
          const int c = (i + j + k) & 1;
          const Stencil S = stencil_of<DIM>(P, c, i, j, k, (i - 1) >> 1);
          double smu = 0.0, sphi = 0.0;
          auto add = [&](int x) {
            smu += (double)eg[x + P.qvs];
            sphi += (double)eg[x];
          };
          if (i == 1) add(S.xm);
          if (i == P.mx[0]) add(S.xp);
          if (j == 1) add(S.ym);
          if (j == P.mx[1]) add(S.yp);
          if (DIM == 3 && k == 1) add(S.zm);
          if (DIM == 3 && k == P.mx[2]) add(S.zp);
          dg[0] = (m.dt / h2) * smu;
          dg[1] = -(m.eps2 / h2) * sphi;
        }
      }
#pragma unroll
      for (int v = 0; v < NV; v++) {
        const double r = f[c0 + v * P.qvs] - o[v];
        if (rf) rf[c0 + v * P.qvs] = (TO)r;
        const double rn = (MODEL == BSAM_MODEL_CH && eg && v < 2) ? r + dg[v] : r;
        sv[v] += rn * rn;
      }
      // the base state phi0 of the next correction cycle (mixed path)
      if (base) base[c0] = (float)q[c0];
    }
  }
  if (!partial) return;
  // wavefront then block reduction in a fixed order (deterministic)
  for (int d = 32; d > 0; d >>= 1)
#pragma unroll
    for (int v = 0; v < NV; v++) sv[v] += __shfl_down(sv[v], d, 64);
  __shared__ double red[kNvMax][kBlock / 64];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0)
#pragma unroll
    for (int v = 0; v < kNvMax; v++) red[v][w] = sv[v];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int v = 0; v < kNvMax; v++) {
      double a = 0.0;
      for (int x = 0; x < kBlock / 64; x++) a += red[v][x];
      partial[kNvMax * lb + v] = a;
    }
  }
}

// Residual of one large 3D CH patch, marching in k.  A block owns 64 cell
// pairs (iA, iB) = (2t+1, 2t+2) x 8 rows x KC planes; a thread keeps its
// pair's phi and mu on planes k-1, k, k+1 in registers (the z neighbours),
// and plane k of the block's rows plus a halo row and cell on each side sits
// in LDS for the x and y neighbours.  The loads of plane k+2 (own pair), of
// f on plane k+1 and of the halo of plane k+1 are issued before the barrier
// of plane k, so their latency overlaps a plane of arithmetic.  The
// expressions are those of op_q / k_residual, so the residual is bit-identical.
//   MODE 0: r = f - L(q) into `out` (rf, or the FP32 right-hand side of the
//           mixed path; may be null) and squared sums into `partial`.
//   MODE 1: k_fas_rhs (CoarseLoadingFunction): the odd row hands its values
//           of planes 2kc-1 and 2kc to the even row through LDS, which sums
//           the eight children of coarse cell (t+1, j/2, kc) in k_fas_rhs's
//           order and adds the coarse operator on qc into the parent's f.
//           KC is even, so a block holds whole coarse planes.
template <int MODE, typename TO>
__global__ __launch_bounds__(512) void k_ch3d_resid_march(const DPatch* __restrict__ pt, int pid, int KC,
                                                         double tmp1, double tmp2, ModelK m,
                                                         const double* __restrict__ q,
                                                         const double* __restrict__ f, TO* __restrict__ out,
                                                         double* __restrict__ partial,
                                                         const double* __restrict__ qc,
                                                         double* __restrict__ fpar,
                                                         float* __restrict__ phi0 = nullptr) {
  constexpr int TX = 64, TY = 8, NR = TY + 2, NC = 2 * TX + 2;
  // halo entries of one plane: rows 0 and NR-1 (NC cells each), cells 0 and
  // NC-1 of rows 1..TY; two fields
  constexpr int NH = 2 * (2 * NC + 2 * TY);
  constexpr int HPER = (NH + 511) / 512;
  __shared__ double sh[2][2][NR][NC];
  __shared__ double xb[MODE == 1 ? TY / 2 : 1][8][TX];
  __shared__ double red[2][8];
  const DPatch P = pt[pid];
  const int mx = P.mx[0], my = P.mx[1], mz = P.mx[2];
  const int nx2 = mx >> 1;
  const int ntx = (nx2 + TX - 1) / TX, nty = (my + TY - 1) / TY;
  const int lb = xcd_block();
  int rem = lb;
  const int bx = rem % ntx;
  rem /= ntx;
  const int by = rem % nty;
  const int bk = rem / nty;
  const int t0 = bx * TX, j0 = by * TY;
  const int k0 = bk * KC + 1, k1 = min(mz, k0 + KC - 1);
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int hq = P.hq, ny = P.nq[1], nz = P.nq[2], vs = P.qvs, base = P.qoff;
  auto cell = [&](int i, int jj, int k) { return base + (((((i + jj + k) & 1) * nz + k) * ny + jj) * hq + (i >> 1)); };
  const int t = t0 + tx, j = j0 + 1 + ty;
  const bool active = (t < nx2) && (j <= my);
  const int iA = 2 * t + 1, iB = 2 * t + 2;
  // the pair is stored (interior or ghost) for these threads; the others of
  // a partial tile hold zeros, which no active thread reads
  const bool inA = iA <= mx + 1 && j <= my + 1, inB = iB <= mx + 1 && j <= my + 1;
  auto load_pair = [&](int k, double (&a)[2], double (&b)[2]) {
    const int oa = cell(iA, j, k), ob = cell(iB, j, k);
    a[0] = inA ? q[oa] : 0.0;
    a[1] = inA ? q[oa + vs] : 0.0;
    b[0] = inB ? q[ob] : 0.0;
    b[1] = inB ? q[ob + vs] : 0.0;
  };
  auto load_f = [&](int k, double (&fv)[4]) {
    if (active) {
      const int fa = cell(iA, j, k), fb = cell(iB, j, k);
      fv[0] = f[fa]; fv[1] = f[fa + vs];
      fv[2] = f[fb]; fv[3] = f[fb + vs];
    } else {
      fv[0] = fv[1] = fv[2] = fv[3] = 0.0;
    }
  };
  auto halo_pos = [&](int e, int& v, int& r, int& cx) {
    v = e / (NH / 2);
    const int h = e % (NH / 2);
    if (h < 2 * NC) {
      r = (h < NC) ? 0 : NR - 1;
      cx = h % NC;
    } else {
      const int h2 = h - 2 * NC;
      r = 1 + (h2 >> 1);
      cx = (h2 & 1) ? NC - 1 : 0;
    }
  };
  auto load_halo = [&](int k, double (&hv)[HPER]) {
#pragma unroll
    for (int n = 0; n < HPER; n++) {
      const int e = threadIdx.x + 512 * n;
      double val = 0.0;
      if (e < NH) {
        int v, r, cx;
        halo_pos(e, v, r, cx);
        const int ii = 2 * t0 + cx, jj = j0 + r;
        if (ii <= mx + 1 && jj <= my + 1) val = q[cell(ii, jj, k) + v * vs];
      }
      hv[n] = val;
    }
  };
  double A[3][2], B[3][2], An[2], Bn[2], F[4], Fn[4], H[HPER], Hn[HPER];
  load_pair(k0 - 1, A[1], B[1]);
  load_pair(k0, A[2], B[2]);
  load_pair(k0 + 1, An, Bn);
  load_f(k0, Fn);
  load_halo(k0, Hn);
  double s0 = 0.0, s1 = 0.0;
  // MODE 1 state: residuals of the odd plane, and the coarse cell of the
  // previous even plane, completed after the next barrier
  double ro[2][2] = {{0.0, 0.0}, {0.0, 0.0}}, re[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
  bool pending = false;
  int pkc = 0;
  const bool evenrow = (j & 1) == 0;
  auto finish = [&]() {
    const int rr = ty >> 1;
    double sv[2];
    for (int v = 0; v < 2; v++) {
      double a = re[v][1];            // (i2, j2, k2)
      a = a + re[v][0];               // (i1, j2, k2)
      a = a + xb[rr][4 * v + 1][tx];  // (i2, j1, k2)
      a = a + xb[rr][4 * v + 0][tx];  // (i1, j1, k2)
      a = a + ro[v][1];               // (i2, j2, k1)
      a = a + ro[v][0];               // (i1, j2, k1)
      a = a + xb[rr][4 * v + 3][tx];  // (i2, j1, k1)
      a = a + xb[rr][4 * v + 2][tx];  // (i1, j1, k1)
      sv[v] = 0.125 * a;
    }
    const DPatch& Q = pt[P.parent];
    const int ic = t + 1, jc = j >> 1, kc = pkc;
    const double lmu = ulap_c<3>(qc, P, ic, jc, kc, P.cvs);
    const double lphi = ulap_c<3>(qc, P, ic, jc, kc, 0);
    const int oc = cidx(P, ic, jc, kc);
    const int ip = P.mb[0][0] + ic - 1, jp = P.mb[1][0] + jc - 1, kp = P.mb[2][0] + kc - 1;
    double o1, o2;
    op_from<3, BSAM_MODEL_CH>(qc[oc], qc[oc + P.cvs], lphi, lmu, 0.0, m, Q.dx, o1, o2);
    const int opar = qidx(Q, ip, jp, kp);
    fpar[opar] = sv[0] + o1;
    fpar[opar + Q.qvs] = sv[1] + o2;
  };
  for (int k = k0; k <= k1; k++) {
    const int sl = k & 1;
    // window: plane k-1 -> [0], k -> [1], k+1 -> [2]
    for (int v = 0; v < 2; v++) {
      A[0][v] = A[1][v]; A[1][v] = A[2][v]; A[2][v] = An[v];
      B[0][v] = B[1][v]; B[1][v] = B[2][v]; B[2][v] = Bn[v];
    }
    for (int n = 0; n < 4; n++) F[n] = Fn[n];
    for (int n = 0; n < HPER; n++) H[n] = Hn[n];
    for (int v = 0; v < 2; v++) {
      sh[sl][v][ty + 1][2 * tx + 1] = A[1][v];
      sh[sl][v][ty + 1][2 * tx + 2] = B[1][v];
    }
#pragma unroll
    for (int n = 0; n < HPER; n++) {
      const int e = threadIdx.x + 512 * n;
      if (e < NH) {
        int v, r, cx;
        halo_pos(e, v, r, cx);
        sh[sl][v][r][cx] = H[n];
      }
    }
    if (k < k1) {
      load_pair(k + 2, An, Bn);
      load_f(k + 1, Fn);
      load_halo(k + 1, Hn);
    }
    __syncthreads();
    if (MODE == 1 && pending) {
      finish();
      pending = false;
    }
    if (active) {
      const int r = ty + 1, cA = 2 * tx + 1, cB = 2 * tx + 2;
      double lapA[2], lapB[2];
      for (int v = 0; v < 2; v++) {
        double sA = B[1][v] + sh[sl][v][r][cA - 1] + sh[sl][v][r + 1][cA] + sh[sl][v][r - 1][cA];
        sA = sA + A[2][v];
        sA = sA + A[0][v];
        lapA[v] = sA - 6.0 * A[1][v];
        double sB = sh[sl][v][r][cB + 1] + A[1][v] + sh[sl][v][r + 1][cB] + sh[sl][v][r - 1][cB];
        sB = sB + B[2][v];
        sB = sB + B[0][v];
        lapB[v] = sB - 6.0 * B[1][v];
      }
      // CH Operator3D: o1 = phi - tmp1*lap(mu), o2 = mu - phi^3 + tmp2*lap(phi)
      const double pa = A[1][0], pb = B[1][0];
      double ra[2], rb[2];
      ra[0] = F[0] - (pa - tmp1 * lapA[1]);
      ra[1] = F[1] - (A[1][1] - (pa * pa * pa) + tmp2 * lapA[0]);
      rb[0] = F[2] - (pb - tmp1 * lapB[1]);
      rb[1] = F[3] - (B[1][1] - (pb * pb * pb) + tmp2 * lapB[0]);
      if (MODE == 0) {
        s0 += ra[0] * ra[0] + rb[0] * rb[0];
        s1 += ra[1] * ra[1] + rb[1] * rb[1];
        if (out) {
          const int oa = cell(iA, j, k), ob = cell(iB, j, k);
          out[oa] = (TO)ra[0]; out[oa + vs] = (TO)ra[1];
          out[ob] = (TO)rb[0]; out[ob + vs] = (TO)rb[1];
          if (phi0) {
            phi0[oa] = (float)pa;
            phi0[ob] = (float)pb;
          }
        }
      } else if (k & 1) {
        for (int v = 0; v < 2; v++) {
          ro[v][0] = ra[v];
          ro[v][1] = rb[v];
        }
      } else if (!evenrow) {
        // odd row: hand this pair's values to the even row after it
        const int rr = ty >> 1;
        for (int v = 0; v < 2; v++) {
          xb[rr][4 * v + 0][tx] = ra[v];
          xb[rr][4 * v + 1][tx] = rb[v];
          xb[rr][4 * v + 2][tx] = ro[v][0];
          xb[rr][4 * v + 3][tx] = ro[v][1];
        }
      } else {
        for (int v = 0; v < 2; v++) {
          re[v][0] = ra[v];
          re[v][1] = rb[v];
        }
        pending = true;
        pkc = k >> 1;
      }
    }
  }
  if (MODE == 1) {
    __syncthreads();
    if (pending) finish();
    return;
  }
  if (!partial) return;
  for (int d = 32; d > 0; d >>= 1) {
    s0 += __shfl_down(s0, d, 64);
    s1 += __shfl_down(s1, d, 64);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][w] = s0;
    red[1][w] = s1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int x = 0; x < 8; x++) {
      a += red[0][x];
      b += red[1][x];
    }
    partial[kNvMax * lb] = a;
    partial[kNvMax * lb + 1] = b;
    partial[kNvMax * lb + 2] = 0.0;
    partial[kNvMax * lb + 3] = 0.0;
  }
}

// Sum the per-block partials in a fixed order: block b of this launch sums
// the chunk [b*chunk, (b+1)*chunk) of `partial` into out[2b..2b+1]
// (launched with many blocks, then with one).
__global__ __launch_bounds__(kBlock) void k_reduce_partials(const double* __restrict__ partial,
                                                            int nb, int chunk, double* __restrict__ out) {
  __shared__ double s[kNvMax][kBlock];
  double a[kNvMax] = {};
  const int lo = blockIdx.x * chunk, hi = min(nb, lo + chunk);
  for (int x = lo + threadIdx.x; x < hi; x += kBlock)
    for (int v = 0; v < kNvMax; v++) a[v] += partial[kNvMax * x + v];
  for (int v = 0; v < kNvMax; v++) s[v][threadIdx.x] = a[v];
  __syncthreads();
  for (int d = kBlock / 2; d > 0; d >>= 1) {
    if ((int)threadIdx.x < d)
      for (int v = 0; v < kNvMax; v++) s[v][threadIdx.x] += s[v][threadIdx.x + d];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int v = 0; v < kNvMax; v++) out[kNvMax * blockIdx.x + v] = s[v][0];
}

// ---------------------------------------------------------------------------
// Restriction (Restriction2D/3D, gridutilities.f90:103-126, 437-466) in the
// reference's summation order: (e,e,e),(o,e,e),(e,o,e),(o,o,e),(e,e,o),...
// ---------------------------------------------------------------------------
template <int DIM, typename T>
__device__ __forceinline__ T restrict_cell(const T* __restrict__ a, const DPatch& P, int ic, int jc, int kc,
                                           int off) {
  const int i2 = 2 * ic, i1 = i2 - 1, j2 = 2 * jc, j1 = j2 - 1;
  if (DIM == 2) {
    return (T)0.25 * (a[qidx(P, i2, j2, 0) + off] + a[qidx(P, i1, j2, 0) + off] +
                   a[qidx(P, i2, j1, 0) + off] + a[qidx(P, i1, j1, 0) + off]);
  }
  const int k2 = 2 * kc, k1 = k2 - 1;
  return (T)0.125 * (a[qidx(P, i2, j2, k2) + off] + a[qidx(P, i1, j2, k2) + off] +
                  a[qidx(P, i2, j1, k2) + off] + a[qidx(P, i1, j1, k2) + off] +
                  a[qidx(P, i2, j2, k1) + off] + a[qidx(P, i1, j2, k1) + off] +
                  a[qidx(P, i2, j1, k1) + off] + a[qidx(P, i1, j1, k1) + off]);
}

// Decode a flat coarse-cell index of the level (domain: cmx^d per patch).
struct CoarseCell {
  int pid, ic, jc, kc;
};
template <int DIM>
__device__ __forceinline__ bool decode_coarse(const LevelArgs& L, int lb, CoarseCell& cc) {
  int pi, local;
  if (!level_elem(L, lb, pi, local)) return false;
  cc.pid = L.patches[pi];
  const DPatch P = L.pt[cc.pid];
  cc.ic = local % P.cmx[0] + 1;
  const int row = local / P.cmx[0];
  cc.jc = row % P.cmx[1] + 1;
  cc.kc = (DIM == 3) ? row / P.cmx[1] + 1 : 0;
  return true;
}

// RestrictSolution (afasroutines.f90:610-652) when `qc` is non-null: qc =
// R q and parent%q(mb) = qc.  FillDownLevel (afasroutines.f90:257-327) when
// `qc` is null: parent field(mb) = R field (field = q or f).
// WRAP: the parent is one patch periodic in every dimension that this
// level's single patch covers completely; the restriction then also writes
// the parent's ghost images, which is the whole effect of the following
// SetGhost(L-1, 0).
template <int DIM, typename T, bool WRAP = false>
__global__ __launch_bounds__(kBlock) void k_restrict(LevelArgs L, int nv, const T* __restrict__ src,
                                                     T* __restrict__ dst_parent, T* __restrict__ qc) {
  CoarseCell c;
  if (!decode_coarse<DIM>(L, level_block(L), c)) return;
  const DPatch P = L.pt[c.pid];
  const DPatch Q = L.pt[P.parent];
  const int ip = P.mb[0][0] + c.ic - 1, jp = P.mb[1][0] + c.jc - 1;
  const int kp = (DIM == 3) ? P.mb[2][0] + c.kc - 1 : 0;
  T rv[2] = {(T)0, (T)0};
  for (int v = 0; v < nv; v++) {
    const T r = restrict_cell<DIM, T>(src, P, c.ic, c.jc, c.kc, v * P.qvs);
    if (qc) qc[cidx(P, c.ic, c.jc, c.kc) + v * P.cvs] = r;
    dst_parent[qidx(Q, ip, jp, kp) + v * Q.qvs] = r;
    if (v < 2) rv[v] = r;
  }
  if (WRAP) wrap_images<DIM, T>(dst_parent, Q, ip, jp, kp, rv[0], rv[1]);
}

// GetCoarseGhostPoints (boundary.f90:957-1015): the ring of qc <- parent q.
// Domain: the (cmx+2)^d qc box; interior cells are skipped.  The reference's
// face loops cover every ring cell exactly once with the value
// parent%q(mb-1+i, mb-1+j, mb-1+k).
template <int DIM, typename T>
__global__ __launch_bounds__(kBlock) void k_coarse_ring(LevelArgs L, int nv, const T* __restrict__ q,
                                                        T* __restrict__ qc) {
  int pi, local;
  if (!level_elem(L, level_block(L), pi, local)) return;
  const DPatch P = L.pt[L.patches[pi]];
  const DPatch Q = L.pt[P.parent];
  const int ic = local % P.nc[0];
  const int row = local / P.nc[0];
  const int jc = row % P.nc[1];
  const int kc = (DIM == 3) ? row / P.nc[1] : 0;
  const bool interior = ic >= 1 && ic <= P.cmx[0] && jc >= 1 && jc <= P.cmx[1] &&
                        (DIM == 2 || (kc >= 1 && kc <= P.cmx[2]));
  if (interior) return;
  const int ip = P.mb[0][0] - 1 + ic, jp = P.mb[1][0] - 1 + jc;
  const int kp = (DIM == 3) ? P.mb[2][0] - 1 + kc : 0;
  for (int v = 0; v < nv; v++) qc[cidx(P, ic, jc, kc) + v * P.cvs] = q[qidx(Q, ip, jp, kp) + v * Q.qvs];
}

// CoarseLoadingFunction (afasroutines.f90:673-736), fused with Residual:
// parent%f(mb) = R(f - L_h q) + L_2h(qc).  The fine residual is formed on
// the fly for the 2^d children of each coarse cell, in the order the
// restriction sums them.
template <int DIM, int MODEL>
__global__ __launch_bounds__(kBlock) void k_fas_rhs(LevelArgs L, ModelK m, const double* __restrict__ q,
                                                    const double* __restrict__ qold,
                                                    const double* __restrict__ f,
                                                    const double* __restrict__ qc,
                                                    double* __restrict__ fpar) {
  constexpr int NV = (MODEL == BSAM_MODEL_TUMOUR) ? 4 : 2;
  CoarseCell c;
  if (!decode_coarse<DIM>(L, level_block(L), c)) return;
  const DPatch P = L.pt[c.pid];
  const DPatch Q = L.pt[P.parent];
  const double h = P.dx;
  const int i2 = 2 * c.ic, j2 = 2 * c.jc, k2 = 2 * c.kc;
  const int nch = (DIM == 3) ? 8 : 4;
  double s[NV];
#pragma unroll
  for (int n = 0; n < nch; n++) {
    // order: (e,e,e),(o,e,e),(e,o,e),(o,o,e),(e,e,o),(o,e,o),(e,o,o),(o,o,o)
    const int i = i2 - (n & 1), j = j2 - ((n >> 1) & 1);
    const int k = (DIM == 3) ? k2 - ((n >> 2) & 1) : 0;
    double o[kNvMax];
    op_cell<DIM, MODEL>(q, qold, P, i, j, k, m, h, o);
    const int ofs = qidx(P, i, j, k);
#pragma unroll
    for (int v = 0; v < NV; v++) {
      const double a = f[ofs + v * P.qvs] - o[v];
      s[v] = (n == 0) ? a : s[v] + a;
    }
  }
  // coarse operator on qc with ch = parent%dx(1); BG and the tumour model
  // read parent%qold.
  const int ip = P.mb[0][0] + c.ic - 1, jp = P.mb[1][0] + c.jc - 1;
  const int kp = (DIM == 3) ? P.mb[2][0] + c.kc - 1 : 0;
  const int oc = cidx(P, c.ic, c.jc, c.kc);
  const int opar = qidx(Q, ip, jp, kp);
  double o[kNvMax];
  if (MODEL == BSAM_MODEL_TUMOUR) {
    const int vs = P.cvs;
    const double l2 = ulap_c<DIM>(qc, P, c.ic, c.jc, c.kc, vs), l1 = ulap_c<DIM>(qc, P, c.ic, c.jc, c.kc, 0);
    const double l3 = ulap_c<DIM>(qc, P, c.ic, c.jc, c.kc, 2 * vs);
    const double l4 = ulap_c<DIM>(qc, P, c.ic, c.jc, c.kc, 3 * vs);
    op_tm_from<DIM>(qc[oc], qc[oc + vs], qc[oc + 2 * vs], l1, l2, l3, l4, qold[opar], m, Q.dx, o);
  } else {
    const double lmu = ulap_c<DIM>(qc, P, c.ic, c.jc, c.kc, P.cvs);
    const double lphi = ulap_c<DIM>(qc, P, c.ic, c.jc, c.kc, 0);
    const double po = (MODEL == BSAM_MODEL_BG) ? qold[opar] : 0.0;
    op_from<DIM, MODEL>(qc[oc], qc[oc + P.cvs], lphi, lmu, po, m, Q.dx, o[0], o[1]);
  }
  const double w = (DIM == 3) ? 0.125 : 0.25;
#pragma unroll
  for (int v = 0; v < NV; v++) fpar[opar + v * Q.qvs] = w * s[v] + o[v];
}

// Regrid tagging, errflagopt 1: RelativeTruncationError
// (afasroutines.f90:739-814), qrte = R(L_h q) - L_2h(qc) with the coarse
// operator on parent%qold and ch, and SetErrFlags2D/3D
// (bsamroutines.f90:1134-1203): the 2^d children of a coarse cell whose
// largest |qrte| over the variables exceeds `tol` are flagged.  The fine
// operator values are summed in the restriction's child order, as in
// k_fas_rhs.  Domain: coarse cells.  flags: per patch (foff[position in
// the level]) column-major over 1..mx, cleared beforehand.
template <int DIM, int MODEL>
__global__ __launch_bounds__(kBlock) void k_rte_flags(LevelArgs L, ModelK m, const double* __restrict__ q,
                                                      const double* __restrict__ qold,
                                                      const double* __restrict__ qc, double tol,
                                                      const long* __restrict__ foff, int32_t* __restrict__ flags) {
  constexpr int NV = (MODEL == BSAM_MODEL_TUMOUR) ? 4 : 2;
  int pi, local;
  if (!level_elem(L, level_block(L), pi, local)) return;
  const DPatch P = L.pt[L.patches[pi]];
  const DPatch Q = L.pt[P.parent];
  const int ic = local % P.cmx[0] + 1, row = local / P.cmx[0];
  const int jc = row % P.cmx[1] + 1, kc = (DIM == 3) ? row / P.cmx[1] + 1 : 0;
  const int i2 = 2 * ic, j2 = 2 * jc, k2 = 2 * kc;
  const int nch = (DIM == 3) ? 8 : 4;
  double s[NV];
#pragma unroll
  for (int n = 0; n < nch; n++) {
    const int i = i2 - (n & 1), j = j2 - ((n >> 1) & 1);
    const int k = (DIM == 3) ? k2 - ((n >> 2) & 1) : 0;
    double o[kNvMax];
    op_cell<DIM, MODEL>(q, qold, P, i, j, k, m, P.dx, o);
#pragma unroll
    for (int v = 0; v < NV; v++) s[v] = (n == 0) ? o[v] : s[v] + o[v];
  }
  const int ip = P.mb[0][0] + ic - 1, jp = P.mb[1][0] + jc - 1;
  const int kp = (DIM == 3) ? P.mb[2][0] + kc - 1 : 0;
  const int oc = cidx(P, ic, jc, kc);
  const int opar = qidx(Q, ip, jp, kp);
  double o[kNvMax];
  if (MODEL == BSAM_MODEL_TUMOUR) {
    const int vs = P.cvs;
    const double l2 = ulap_c<DIM>(qc, P, ic, jc, kc, vs), l1 = ulap_c<DIM>(qc, P, ic, jc, kc, 0);
    const double l3 = ulap_c<DIM>(qc, P, ic, jc, kc, 2 * vs);
    const double l4 = ulap_c<DIM>(qc, P, ic, jc, kc, 3 * vs);
    op_tm_from<DIM>(qc[oc], qc[oc + vs], qc[oc + 2 * vs], l1, l2, l3, l4, qold[opar], m, Q.dx, o);
  } else {
    const double lmu = ulap_c<DIM>(qc, P, ic, jc, kc, P.cvs);
    const double lphi = ulap_c<DIM>(qc, P, ic, jc, kc, 0);
    const double po = (MODEL == BSAM_MODEL_BG) ? qold[opar] : 0.0;
    op_from<DIM, MODEL>(qc[oc], qc[oc + P.cvs], lphi, lmu, po, m, Q.dx, o[0], o[1]);
  }
  const double w = (DIM == 3) ? 0.125 : 0.25;
  double mv = -DBL_MAX;
#pragma unroll
  for (int v = 0; v < NV; v++) {
    const double a = fabs(w * s[v] - o[v]);
    if (a > mv) mv = a;
  }
  if (!(mv > tol)) return;
  int32_t* fl = flags + foff[pi];
  const int mxx = P.mx[0], myy = P.mx[1];
#pragma unroll
  for (int n = 0; n < nch; n++) {
    const int i = i2 - (n & 1), j = j2 - ((n >> 1) & 1);
    const int k = (DIM == 3) ? k2 - ((n >> 2) & 1) : 1;
    fl[((long)(k - 1) * myy + (j - 1)) * mxx + (i - 1)] = 1;
  }
}

// Regrid tagging on variable 1 of q as it is.  errflagopt 10: the
// second-difference test of the CH and BioGrowth SetErrFlagsUser2D/3D
// (CH/problem.f90:248-307, BioGrowth/problem.f90:185-240):
//   |q(i+1)-2q(i)+q(i-1)| + |..j..| (+ |..k..|) > thr (= ndims*qtolerance).
// GRAD (errflagopt 11): the gradient test of the probe PROBLEM module that
// built the AMR workloads and the tumour model's meshes
// (oracle/ref_problem/problem.f90 SetErrFlagsUser2D/3D):
//   sqrt(gx^2 + gy^2 (+ gz^2)) > thr (= qtolerance), g = (q(+1)-q(-1))/2.
// Domain: cells; every flag of the level is written.
template <int DIM, bool GRAD = false>
__global__ __launch_bounds__(kBlock) void k_user_flags(LevelArgs L, const double* __restrict__ q, double thr,
                                                       const long* __restrict__ foff, int32_t* __restrict__ flags) {
  int pi, local;
  if (!level_elem(L, level_block(L), pi, local)) return;
  const DPatch P = L.pt[L.patches[pi]];
  int i, j, k;
  if (!cell_of<DIM>(P, local, i, j, k)) return;
  bool tag;
  if (GRAD) {
    const double gx = 0.5 * (q[qidx(P, i + 1, j, k)] - q[qidx(P, i - 1, j, k)]);
    const double gy = 0.5 * (q[qidx(P, i, j + 1, k)] - q[qidx(P, i, j - 1, k)]);
    double g2 = gx * gx + gy * gy;
    if (DIM == 3) {
      const double gz = 0.5 * (q[qidx(P, i, j, k + 1)] - q[qidx(P, i, j, k - 1)]);
      g2 = g2 + gz * gz;
    }
    tag = sqrt(g2) > thr;
  } else {
    const double c0 = q[qidx(P, i, j, k)];
    double d = fabs(q[qidx(P, i + 1, j, k)] - 2.0 * c0 + q[qidx(P, i - 1, j, k)]) +
               fabs(q[qidx(P, i, j + 1, k)] - 2.0 * c0 + q[qidx(P, i, j - 1, k)]);
    if (DIM == 3) d = d + fabs(q[qidx(P, i, j, k + 1)] - 2.0 * c0 + q[qidx(P, i, j, k - 1)]);
    tag = d > thr;
  }
  const int kk = (DIM == 3) ? k - 1 : 0;
  flags[foff[pi] + ((long)kk * P.mx[1] + (j - 1)) * P.mx[0] + (i - 1)] = tag ? 1 : 0;
}

__global__ __launch_bounds__(kBlock) void k_zero_i32(int32_t* __restrict__ a, long n) {
  const long g = (long)blockIdx.x * kBlock + threadIdx.x;
  if (g < n) a[g] = 0;
}

// Tag buffering (BufferAndList and InflateEdgeTags, bsamroutines.f90:
// 1206-1349) as a dilation over the level's whole index space: every tag
// (flag 1) sets its bit in a bit map of the level's global cells
// (k_tag_bits), then a cell of a level patch is tagged when a bit lies
// within ibuffer cells in every dimension, with the root's periodic
// dimensions wrapped (k_tag_dilate).  The reference buffers a tag inside
// its patch when its buffer box fits there and otherwise applies the box,
// and its periodic images, to every patch of the level: the same union of
// boxes, cut to each patch.  glo: each level patch's global lower corner
// (0-based); ext: the level's extents in cells.
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_tag_bits(LevelArgs L, const long* __restrict__ foff,
                                                      const int* __restrict__ glo, const int32_t* __restrict__ flags,
                                                      uint32_t* __restrict__ bits, int ex, int ey) {
  int pi, local;
  if (!level_elem(L, level_block(L), pi, local)) return;
  const DPatch P = L.pt[L.patches[pi]];
  int i, j, k;
  if (!cell_of<DIM>(P, local, i, j, k)) return;
  const int kk = (DIM == 3) ? k - 1 : 0;
  if (flags[foff[pi] + ((long)kk * P.mx[1] + (j - 1)) * P.mx[0] + (i - 1)] != 1) return;
  const long b = ((long)(glo[3 * pi + 2] + kk) * ey + (glo[3 * pi + 1] + j - 1)) * ex + (glo[3 * pi] + i - 1);
  atomicOr(&bits[b >> 5], 1u << (b & 31));
}

template <int DIM>
__global__ __launch_bounds__(kBlock) void k_tag_dilate(LevelArgs L, const long* __restrict__ foff,
                                                        const int* __restrict__ glo, const uint32_t* __restrict__ bits,
                                                        int ex, int ey, int ez, int wrapx, int wrapy, int wrapz, int ib,
                                                        int32_t* __restrict__ out) {
  int pi, local;
  if (!level_elem(L, level_block(L), pi, local)) return;
  const DPatch P = L.pt[L.patches[pi]];
  int i, j, k;
  if (!cell_of<DIM>(P, local, i, j, k)) return;
  const int kk = (DIM == 3) ? k - 1 : 0;
  const int gi = glo[3 * pi] + i - 1, gj = glo[3 * pi + 1] + j - 1, gk = glo[3 * pi + 2] + kk;
  const int zb = (DIM == 3) ? ib : 0;
  bool hit = false;
  for (int dz = -zb; dz <= zb && !hit; dz++) {
    int z = gk + dz;
    if (z < 0 || z >= ez) {
      if (!wrapz) continue;
      z += z < 0 ? ez : -ez;
    }
    for (int dy = -ib; dy <= ib && !hit; dy++) {
      int y = gj + dy;
      if (y < 0 || y >= ey) {
        if (!wrapy) continue;
        y += y < 0 ? ey : -ey;
      }
      for (int dx = -ib; dx <= ib; dx++) {
        int x = gi + dx;
        if (x < 0 || x >= ex) {
          if (!wrapx) continue;
          x += x < 0 ? ex : -ex;
        }
        const long b = ((long)z * ey + y) * ex + x;
        if ((bits[b >> 5] >> (b & 31)) & 1u) {
          hit = true;
          break;
        }
      }
    }
  }
  out[foff[pi] + ((long)kk * P.mx[1] + (j - 1)) * P.mx[0] + (i - 1)] = hit ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Output frame (WriteQ / OutputQ / WriteQ2D/3D, bsaminputoutput.f90:60-181)
// ---------------------------------------------------------------------------
// Fortran F25.12 of x as the amdflang runtime edits it, into o[0..24]: the
// value correctly rounded to twelve decimals (halfway cases to even), the
// sign kept when a negative value rounds to zero, 25 asterisks when the
// edited value needs more than 25 characters, NaN / Inf / -Inf
// right-justified.  Exact by integer arithmetic on the binary value
// x = m 2^e: the fraction f / 2^s (s = -e) times 10^12 is formed in 128
// bits (f < 2^53, so f 10^12 < 2^93) and shifted, the bits shifted out
// deciding the rounding; no floating-point operation is involved.
__host__ __device__ inline void f25_edit(double x, char* o) {
  const uint64_t b = __builtin_bit_cast(uint64_t, x);
  const bool neg = b >> 63;
  const int E = (int)((b >> 52) & 0x7ff);
  uint64_t m = b & ((1ull << 52) - 1);
  for (int t = 0; t < 25; t++) o[t] = ' ';
  if (E == 0x7ff) {
    const char* w = m ? "NaN" : (neg ? "-Inf" : "Inf");
    int n = 0;
    while (w[n]) n++;
    for (int t = 0; t < n; t++) o[25 - n + t] = w[t];
    return;
  }
  int e;
  if (E == 0) {
    e = -1074;
  } else {
    m |= 1ull << 52;
    e = E - 1075;
  }
  if (e >= 0) {  // |x| >= 2^52: more than 25 characters
    for (int t = 0; t < 25; t++) o[t] = '*';
    return;
  }
  const int s = -e;
  uint64_t ip = s >= 64 ? 0 : (m >> s);
  const uint64_t f = s >= 64 ? m : (m & ((1ull << s) - 1));
  const unsigned __int128 P = (unsigned __int128)f * 1000000000000ull;
  uint64_t q = 0;
  if (s < 128) {
    q = (uint64_t)(P >> s);
    const unsigned __int128 r = P & ((((unsigned __int128)1) << s) - 1);
    const unsigned __int128 half = ((unsigned __int128)1) << (s - 1);
    if (r > half || (r == half && (q & 1))) {
      if (++q == 1000000000000ull) {
        q = 0;
        ip++;
      }
    }
  }
  char d[32];
  int n = 0;
  for (int t = 0; t < 12; t++) {
    d[n++] = (char)('0' + q % 10);
    q /= 10;
  }
  d[n++] = '.';
  do {
    d[n++] = (char)('0' + ip % 10);
    ip /= 10;
  } while (ip && n < 30);
  if (neg) d[n++] = '-';
  if (n > 25 || ip) {
    for (int t = 0; t < 25; t++) o[t] = '*';
    return;
  }
  for (int t = 0; t < n; t++) o[24 - t] = d[t];
}

// One patch's share of a frame chunk: its first cell in the chunk's flat
// cell range and the byte where its body starts.
struct FramePatch {
  int pid;
  long cell0, byte0;
};

// The body of WriteQ2D/3D (bsaminputoutput.f90:132-181) for the patches of
// a chunk: cell (i, j, k) of q(0:mx+1, ...), i fastest, is one record of
// its nv values edited 10(F25.12,2X) -- fields of 25 characters two blanks
// apart, no trailing blanks -- at a fixed length of 27 nv - 1 bytes with
// the newline.  Domain: the cells of the chunk.
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_frame_body(const DPatch* __restrict__ pt,
                                                       const FramePatch* __restrict__ fp, int np, long ncell,
                                                       int nv, const double* __restrict__ q,
                                                       char* __restrict__ out) {
  const long g = (long)blockIdx.x * kBlock + threadIdx.x;
  if (g >= ncell) return;
  int lo = 0, hi = np;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (fp[mid].cell0 <= g) lo = mid; else hi = mid;
  }
  const FramePatch F = fp[lo];
  const DPatch P = pt[F.pid];
  const long c = g - F.cell0;
  const int i = (int)(c % P.nq[0]);
  const long r = c / P.nq[0];
  const int j = (int)(r % P.nq[1]);
  const int k = (DIM == 3) ? (int)(r / P.nq[1]) : 0;
  const int len = 27 * nv - 1;
  char* o = out + F.byte0 + c * len;
  const int base = qidx(P, i, j, k);
  for (int v = 0; v < nv; v++) {
    f25_edit(q[base + v * P.qvs], o + 27 * v);
    if (v < nv - 1) {
      o[27 * v + 25] = ' ';
      o[27 * v + 26] = ' ';
    }
  }
  o[len - 1] = '\n';
}

// CorrectFine, first half (afasroutines.f90:892-895, 934-938):
// qc(0:cmx+1) = parent%q(mb-1:mb+1) - qc.  ASSIGN: qc(0:cmx+1) =
// parent%q(mb-1:mb+1), InitFields' coarse box (bsamroutines.f90:1819-1826).
// Domain: the qc box.
template <int DIM, typename T, bool ASSIGN = false>
__global__ __launch_bounds__(kBlock) void k_correct_qc(LevelArgs L, int nv, const T* __restrict__ q,
                                                       T* __restrict__ qc) {
  int pi, local;
  if (!level_elem(L, level_block(L), pi, local)) return;
  const DPatch P = L.pt[L.patches[pi]];
  const DPatch Q = L.pt[P.parent];
  const int ic = local % P.nc[0];
  const int row = local / P.nc[0];
  const int jc = row % P.nc[1];
  const int kc = (DIM == 3) ? row / P.nc[1] : 0;
  const int ip = P.mb[0][0] - 1 + ic, jp = P.mb[1][0] - 1 + jc;
  const int kp = (DIM == 3) ? P.mb[2][0] - 1 + kc : 0;
  for (int v = 0; v < nv; v++) {
    const int o = cidx(P, ic, jc, kc) + v * P.cvs;
    const T pq = q[qidx(Q, ip, jp, kp) + v * Q.qvs];
    qc[o] = ASSIGN ? pq : pq - qc[o];
  }
}

// Linear weights of Bi/TriLinProlongationP1 (gridutilities.f90:156-195,
// 510-566): even fine index 3*a(lo)+a(lo+1), odd a(lo)+3*a(lo+1).
template <typename T>
__device__ __forceinline__ T linw(T lo, T hi, int fine) {
  return (fine & 1) ? lo + (T)3 * hi : (T)3 * lo + hi;
}

// CorrectFine, second half: q(0:mx+1) += P1(qc).  One thread per dual
// cell (ic,jc[,kc]) in 0..cmx: it loads the 2^d coarse values once and
// updates the 2^d fine cells 2ic..2ic+1 (x) etc., all of which interpolate
// from exactly those coarse values.  Same expressions as below, so the
// result is bit-identical.  Domain: (cmx+1)^d.
template <int DIM, typename T, bool WRAP = false>
__global__ __launch_bounds__(kBlock) void k_prolong_dual(LevelArgs L, int nv, T* __restrict__ q,
                                                         const T* __restrict__ qc) {
  int pi, local;
  if (!level_elem(L, level_block(L), pi, local)) return;
  const DPatch P = L.pt[L.patches[pi]];
  const int n0 = P.cmx[0] + 1, n1 = P.cmx[1] + 1;
  const int ic = local % n0;
  const int row = local / n0;
  const int jc = row % n1;
  const int kc = (DIM == 3) ? row / n1 : 0;
  constexpr int ncell = (DIM == 3) ? 8 : 4;
  auto fine = [&](int n, int& i, int& j, int& k) {
    i = 2 * ic + (n & 1);
    j = 2 * jc + ((n >> 1) & 1);
    k = (DIM == 3) ? 2 * kc + (n >> 2) : 0;
  };
  auto is_ghost = [&](int i, int j, int k) {
    return i == 0 || i == P.mx[0] + 1 || j == 0 || j == P.mx[1] + 1 || (DIM == 3 && (k == 0 || k == P.mx[2] + 1));
  };
  // variables in pairs (CH and BG have one pair, the tumour model two)
  for (int vp = 0; vp < nv; vp += 2) {
    // coarse values of the pair around the dual cell
    T a[2][2][2][2];
#pragma unroll
    for (int v = 0; v < 2; v++)
#pragma unroll
      for (int d = 0; d < 2; d++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
          for (int e = 0; e < 2; e++)
            a[v][d][b][e] = (DIM == 3 || d == 0) ? qc[cidx(P, ic + e, jc + b, kc + d) + (vp + v) * P.cvs] : (T)0;
    // all loads of the 2^d fine cells first (ghost cells are storage too), so
    // they are in flight together
    T old[ncell][2];
#pragma unroll
    for (int n = 0; n < ncell; n++) {
      int i, j, k;
      fine(n, i, j, k);
      const int o = qidx(P, i, j, k) + vp * P.qvs;
      old[n][0] = q[o];
      old[n][1] = q[o + P.qvs];
    }
#pragma unroll
    for (int n = 0; n < ncell; n++) {
      int i, j, k;
      fine(n, i, j, k);
      if (WRAP && is_ghost(i, j, k)) continue;
#pragma unroll
      for (int v = 0; v < 2; v++) {
        T pr;
        if (DIM == 2) {
          const T b0 = linw<T>(a[v][0][0][0], a[v][0][0][1], i), b1 = linw<T>(a[v][0][1][0], a[v][0][1][1], i);
          pr = linw<T>(b0, b1, j) / (T)16;
        } else {
          T cc2[2];
#pragma unroll
          for (int d = 0; d < 2; d++) {
            const T b0 = linw<T>(a[v][d][0][0], a[v][d][0][1], i);
            const T b1 = linw<T>(a[v][d][1][0], a[v][d][1][1], i);
            cc2[d] = linw<T>(b0, b1, j);
          }
          pr = linw<T>(cc2[0], cc2[1], k) / (T)64;
        }
        q[qidx(P, i, j, k) + (vp + v) * P.qvs] = old[n][v] + pr;
      }
    }
  }
  if (WRAP) {
    // dual cells on the patch boundary also write the periodic images of
    // their interior fine cells, read back from this thread's own stores
    const bool edge = ic == 0 || ic == P.cmx[0] || jc == 0 || jc == P.cmx[1] ||
                      (DIM == 3 && (kc == 0 || kc == P.cmx[2]));
    if (edge) {
#pragma unroll 1
      for (int n = 0; n < ncell; n++) {
        int i, j, k;
        fine(n, i, j, k);
        if (is_ghost(i, j, k)) continue;
        const int o = qidx(P, i, j, k);
        wrap_images<DIM, T>(q, P, i, j, k, q[o], q[o + P.qvs]);
      }
    }
  }
}

// CorrectFine, second half: q(0:mx+1) += P1(qc).  Domain: (mx+2)^d.
template <int DIM, typename T>
__global__ __launch_bounds__(kBlock) void k_prolong_add(LevelArgs L, int nv, T* __restrict__ q,
                                                        const T* __restrict__ qc) {
  int pi, local;
  if (!level_elem(L, level_block(L), pi, local)) return;
  const DPatch P = L.pt[L.patches[pi]];
  const int i = local % P.nq[0];
  const int row = local / P.nq[0];
  const int j = row % P.nq[1];
  const int k = (DIM == 3) ? row / P.nq[1] : 0;
  const int il = i >> 1, jl = j >> 1, kl = k >> 1;
  for (int v = 0; v < nv; v++) {
    const int cv = v * P.cvs;
    T pr;
    if (DIM == 2) {
      const T b0 = linw<T>(qc[cidx(P, il, jl, 0) + cv], qc[cidx(P, il + 1, jl, 0) + cv], i);
      const T b1 = linw<T>(qc[cidx(P, il, jl + 1, 0) + cv], qc[cidx(P, il + 1, jl + 1, 0) + cv], i);
      pr = linw<T>(b0, b1, j) / (T)16;
    } else {
      T cc[2];
#pragma unroll
      for (int dk = 0; dk < 2; dk++) {
        const T b0 = linw<T>(qc[cidx(P, il, jl, kl + dk) + cv], qc[cidx(P, il + 1, jl, kl + dk) + cv], i);
        const T b1 =
            linw<T>(qc[cidx(P, il, jl + 1, kl + dk) + cv], qc[cidx(P, il + 1, jl + 1, kl + dk) + cv], i);
        cc[dk] = linw<T>(b0, b1, j);
      }
      pr = linw<T>(cc[0], cc[1], k) / (T)64;
    }
    const int o = qidx(P, i, j, k) + v * P.qvs;
    q[o] = q[o] + pr;
  }
}

// Source#D (CH/problem.f90:425-450, 672-695; BioGrowth/problem.f90:356-379)
// into ftmp, then f = ftmp + SourceUpdate#D (= 0 for both models)
// (afasroutines.f90:330-381, 438-492).  mode 0: source only, 1: update only.
template <int DIM, int MODEL>
__global__ __launch_bounds__(kBlock) void k_source(LevelArgs L, int mode, ModelK m,
                                                   const double* __restrict__ qold,
                                                   double* __restrict__ ftmp, double* __restrict__ f) {
  constexpr int NV = (MODEL == BSAM_MODEL_TUMOUR) ? 4 : 2;
  const double dt = m.dt;
  int pi, local;
  if (!level_elem(L, level_block(L), pi, local)) return;
  const DPatch P = L.pt[L.patches[pi]];
  int i, j, k;
  if (!cell_of<DIM>(P, local, i, j, k)) return;
  const int o = qidx(P, i, j, k);
  if (mode != 1) {
    const double po = qold[o];
    if (MODEL == BSAM_MODEL_CH) {
      ftmp[o] = po;
      ftmp[o + P.qvs] = -po;
    } else if (MODEL == BSAM_MODEL_TUMOUR) {
      // the probe PROBLEM module's Source2D/3D with model = 3
      ftmp[o] = po - dt * m.lam * (0.5 * (po + 1.0));
      ftmp[o + P.qvs] = -po;
      ftmp[o + 2 * P.qvs] = 0.0;
      ftmp[o + 3 * P.qvs] = -m.lam * (0.5 * (po + 1.0));
    } else {
      ftmp[o] = (1.0 + dt) * po;
      ftmp[o + P.qvs] = 0.0;
    }
  }
  if (mode != 0) {
#pragma unroll
    for (int v = 0; v < NV; v++) f[o + v * P.qvs] = ftmp[o + v * P.qvs] + 0.0;
  }
}

// InitFields on the patches of a new level (bsamroutines.f90:1787-1856),
// after their qc boxes hold the parent's values (k_correct_qc<ASSIGN>):
// q(0:mx+1) = P1(qc) (Bi/TriLinProlongationP1, the expressions of
// k_prolong_dual), then the 2^d children of every interior coarse cell are
// shifted by one common correction, qc - 2^-d (their sum), summed in the
// order of the P1MC loop (gridutilities.f90:628-705), and qold = q.  One
// thread per interior coarse cell (ic, jc[, kc]): its children 2ic-1..2ic
// per dimension, and the ghost cells 0 and mx+1 next to them when the cell
// is on the patch boundary (P1 only), so every fine cell has one writer.
// Domain: kCoarse.
template <int DIM, typename T>
__global__ __launch_bounds__(kBlock) void k_init_fields(LevelArgs L, int nv, T* __restrict__ q,
                                                        T* __restrict__ qold, const T* __restrict__ qc) {
  CoarseCell cc;
  if (!decode_coarse<DIM>(L, level_block(L), cc)) return;
  const DPatch P = L.pt[cc.pid];
  const int cx[3] = {cc.ic, cc.jc, cc.kc};
  int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  for (int a = 0; a < DIM; a++) {
    lo[a] = cx[a] == 1 ? 0 : 2 * cx[a] - 1;
    hi[a] = cx[a] == P.cmx[a] ? P.mx[a] + 1 : 2 * cx[a];
  }
  auto p1 = [&](const T* a, int i, int j, int k) -> T {
    const int il = i >> 1, jl = j >> 1, kl = k >> 1;
    if (DIM == 2) {
      const T b0 = linw<T>(a[cidx(P, il, jl, 0)], a[cidx(P, il + 1, jl, 0)], i);
      const T b1 = linw<T>(a[cidx(P, il, jl + 1, 0)], a[cidx(P, il + 1, jl + 1, 0)], i);
      return linw<T>(b0, b1, j) / (T)16;
    }
    T c2[2];
#pragma unroll
    for (int d = 0; d < 2; d++) {
      const T b0 = linw<T>(a[cidx(P, il, jl, kl + d)], a[cidx(P, il + 1, jl, kl + d)], i);
      const T b1 = linw<T>(a[cidx(P, il, jl + 1, kl + d)], a[cidx(P, il + 1, jl + 1, kl + d)], i);
      c2[d] = linw<T>(b0, b1, j);
    }
    return linw<T>(c2[0], c2[1], k) / (T)64;
  };
  constexpr int nch = (DIM == 3) ? 8 : 4;
  constexpr int nz = (DIM == 3) ? 3 : 1;
  for (int v = 0; v < nv; v++) {
    const T* a = qc + v * P.cvs;
    // the coarse cell's 3^d neighbourhood, loaded once: a child's p1 reads
    // 2^d of these (il = i >> 1 is ic - 1 for the odd child, ic for the
    // even one, and il + 1)
    T nb[nz][3][3];
#pragma unroll
    for (int dz = 0; dz < nz; dz++)
#pragma unroll
      for (int dy = 0; dy < 3; dy++)
#pragma unroll
        for (int dx = 0; dx < 3; dx++)
          nb[dz][dy][dx] = a[cidx(P, cc.ic - 1 + dx, cc.jc - 1 + dy, DIM == 3 ? cc.kc - 1 + dz : 0)];
    // p1 of child n from the neighbourhood, in p1's operation order; the
    // neighbourhood offsets (i >> 1) - (ic - 1) etc. are 1 for an even
    // child index and 0 for an odd one, written from n so that the array
    // stays in registers
    auto p1n = [&](int n, int i, int j, int k) -> T {
      const int ox = 1 - (n & 1), oy = 1 - ((n >> 1) & 1);
      if (DIM == 2) {
        const T b0 = linw<T>(nb[0][oy][ox], nb[0][oy][ox + 1], i);
        const T b1 = linw<T>(nb[0][oy + 1][ox], nb[0][oy + 1][ox + 1], i);
        return linw<T>(b0, b1, j) / (T)16;
      }
      const int oz = 1 - (n >> 2);
      T c2[2];
#pragma unroll
      for (int d = 0; d < 2; d++) {
        const T b0 = linw<T>(nb[oz + d][oy][ox], nb[oz + d][oy][ox + 1], i);
        const T b1 = linw<T>(nb[oz + d][oy + 1][ox], nb[oz + d][oy + 1][ox + 1], i);
        c2[d] = linw<T>(b0, b1, j);
      }
      return linw<T>(c2[0], c2[1], k) / (T)64;
    };
    // children (2ic - n&1, 2jc - (n>>1)&1, 2kc - n>>2) in the loop's order
    T ch[nch];
    T sum = (T)0;
#pragma unroll
    for (int n = 0; n < nch; n++) {
      ch[n] = p1n(n, 2 * cc.ic - (n & 1), 2 * cc.jc - ((n >> 1) & 1), DIM == 3 ? 2 * cc.kc - (n >> 2) : 0);
      sum = n == 0 ? ch[0] : sum + ch[n];
    }
    const T cor = nb[nz / 2][1][1] - (DIM == 3 ? (T)0.125 : (T)0.25) * sum;
#pragma unroll
    for (int n = 0; n < nch; n++) {
      const int o = qidx(P, 2 * cc.ic - (n & 1), 2 * cc.jc - ((n >> 1) & 1), DIM == 3 ? 2 * cc.kc - (n >> 2) : 0) +
                    v * P.qvs;
      const T x = ch[n] + cor;
      q[o] = x;
      qold[o] = x;
    }
    // the ghost cells this coarse cell owns
    for (int k = lo[2]; k <= hi[2]; k++)
      for (int j = lo[1]; j <= hi[1]; j++)
        for (int i = lo[0]; i <= hi[0]; i++) {
          const bool child = i >= 2 * cc.ic - 1 && i <= 2 * cc.ic && j >= 2 * cc.jc - 1 && j <= 2 * cc.jc &&
                             (DIM == 2 || (k >= 2 * cc.kc - 1 && k <= 2 * cc.kc));
          if (child) continue;
          const int o = qidx(P, i, j, k) + v * P.qvs;
          const T x = p1(a, i, j, k);
          q[o] = x;
          qold[o] = x;
        }
  }
}

// TransferValues (bsamroutines.f90:1859-1951): a new patch takes q over the
// interior overlap of its global index box with an old patch of its level.
// One box per (new, old) pair that overlaps; the old patch's descriptor
// travels in the box, since the old hierarchy's table is gone.
struct XferBox {
  DPatch S;       // old patch, offsets into the old q pool
  int dp;         // new patch (descriptor index)
  int slo[3];     // first overlap cell in the old patch's local indices
  int dlo[3];     // ... and in the new patch's
  int ext[3];     // overlap extents (1 in unused dimensions)
  long cell0;     // first element of the box in the launch's range
};

// Domain: the cells of the launch's boxes, box after box.
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_transfer(const DPatch* __restrict__ pt,
                                                     const XferBox* __restrict__ bx, int nbox, long ncell,
                                                     int nv, const double* __restrict__ oq,
                                                     double* __restrict__ q) {
  const long g = (long)blockIdx.x * kBlock + threadIdx.x;
  if (g >= ncell) return;
  int lo = 0, hi = nbox;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (bx[mid].cell0 <= g) lo = mid; else hi = mid;
  }
  const XferBox& B = bx[lo];
  const DPatch D = pt[B.dp];
  const long c = g - B.cell0;
  const int x = (int)(c % B.ext[0]);
  const long r = c / B.ext[0];
  const int y = (int)(r % B.ext[1]);
  const int z = (DIM == 3) ? (int)(r / B.ext[1]) : 0;
  const int so = qidx(B.S, B.slo[0] + x, B.slo[1] + y, B.slo[2] + z);
  const int d0 = qidx(D, B.dlo[0] + x, B.dlo[1] + y, B.dlo[2] + z);
  for (int v = 0; v < nv; v++) q[d0 + v * D.qvs] = oq[so + v * B.S.qvs];
}

// CopyQToQold (bsamroutines.f90:592-605): whole arrays including ghosts.
__global__ __launch_bounds__(kBlock) void k_copy_span(const double* __restrict__ src,
                                                      double* __restrict__ dst, long n) {
  long g = (long)blockIdx.x * kBlock + threadIdx.x;
  if (g < n) dst[g] = src[g];
}

// ---------------------------------------------------------------------------
// Ghost filling.
// ---------------------------------------------------------------------------
// GetFaceIndex (boundary.f90:1018-1045): 0 ghost, 1 second interior,
// 2 first interior, 3 coarse outside, 4 coarse inside.
__host__ __device__ inline void face_index(const DPatch& P, int dim, int parity, int fi[5]) {
  if (parity == 1) {
    fi[0] = 0; fi[1] = 2; fi[2] = 1; fi[3] = P.mb[dim][0] - 1; fi[4] = P.mb[dim][0];
  } else {
    const int mm = P.mx[dim];
    fi[0] = mm + 1; fi[1] = mm - 1; fi[2] = mm; fi[3] = P.mb[dim][1] + 1; fi[4] = P.mb[dim][1];
  }
}

// Transverse position kind of Interpolate2D/3D's statements: 1 -> L,
// mx -> H, even -> E (sources at +2/+1), odd -> O (sources at -2/-1).
enum { KE = 0, KO = 1, KL = 2, KH = 3 };
__host__ __device__ inline int pos_kind(int pos, int m) {
  return pos == 1 ? KL : (pos == m ? KH : ((pos & 1) ? KO : KE));
}
__host__ __device__ inline int kind_s1(int k) { return k == KE ? 2 : (k == KO ? -2 : 0); }
__host__ __device__ inline int kind_s2(int k) { return k == KE ? 1 : (k == KO ? -1 : 0); }
// coarse index of an E/O position, or the three quadratic nodes of L/H
__host__ __device__ inline int kind_c(const DPatch& P, int dim, int k, int pos, int r) {
  const int mb1 = P.mb[dim][0], mb2 = P.mb[dim][1];
  if (k == KE) return mb1 + pos / 2 - 1;
  if (k == KO) return mb1 + (pos - 1) / 2;
  if (k == KL) return mb1 - 1 + r;   // mb1-1, mb1, mb1+1
  return mb2 + 1 - r;                // mb2+1, mb2, mb2-1
}
__host__ __device__ inline bool bc_interp(int code) { return code == 2 || code == 999; }

// One coarse-fine ghost value of InterpolateCoarseFine (boundary.f90:
// 100-140, Interpolate2D :143-333, Interpolate3D :1048-1550), reduced to a
// gather: destination and sources are element offsets of variable 0 in the
// q-geometry pools (pvs / cvs: variable strides of the patch and of its
// parent).  kind 1: one coarse value; 3: the transverse quadratic
// (5/32, 15/16, -3/32); 4: the bilinear (9, 3, 3, 1)/16 of face corners.
struct GRule {
  int dst, s1, s2;
  int c[4];
  int kind;
  int pvs, cvs;
};

// Apply the rules [lo, hi) in the reference's expression order.  Every
// destination is a ghost cell written by one rule and every source an
// interior or parent cell, so the rules are independent.
template <typename T>
__device__ __forceinline__ T interp_value(const GRule& R, int kind, int v, const T* __restrict__ q) {
  const int fv = v * R.pvs, pv = v * R.cvs;
  T coarse;
  if (kind == 1) {
    coarse = q[R.c[0] + pv];
  } else if (kind == 3) {
    coarse = (T)cB[0] * q[R.c[0] + pv] + (T)cB[1] * q[R.c[1] + pv] + (T)cB[2] * q[R.c[2] + pv];
  } else {
    coarse = (T)cC[0] * q[R.c[0] + pv] + (T)cC[1] * q[R.c[1] + pv] + (T)cC[2] * q[R.c[2] + pv] +
             (T)cC[3] * q[R.c[3] + pv];
  }
  return (T)cA[0] * q[R.s1 + fv] + (T)cA[1] * q[R.s2 + fv] + (T)cA[2] * coarse;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_interp_map(const GRule* __restrict__ rules, int lo, int hi, int nv,
                                                       T* __restrict__ q) {
  const int g = lo + xcd_block() * kBlock + threadIdx.x;
  if (g >= hi) return;
  const GRule R = rules[g];
  for (int v = 0; v < nv; v++) q[R.dst + v * R.pvs] = interp_value<T>(R, R.kind, v, q);
}


// One whole SetGhost of a level from its flattened rules: every rule
// writes one ghost cell (all variables) from interior or parent cells, so
// the rules are independent and one launch applies them all.  Copy rules
// are (dst, src, dst variable stride, src variable stride); interpolation
// rules carry their post-operations in bits 8-15 of `kind`.
//
// REBASE (mixed path, FP64 solution): the ghost values before the call plus
// the correction's ghosts `eg` are the ghosts the FAS cycle would hold (see
// rebase_ghosts); each rule writes the fresh value into q and the
// difference (old + eg - new) into eg.
//
// The workgroup size is the level's tuned gather block (64 to 1024 threads).
template <typename T, bool REBASE = false>
__global__ __launch_bounds__(1024) void k_ghost_map(const int4* __restrict__ cr, const uint16_t* __restrict__ cpost,
                                                    int nc, const GRule* __restrict__ ir, int ni, int nv,
                                                    int homog, T* __restrict__ q, int order,
                                                    float* __restrict__ eg = nullptr) {
  const int g = ordered_block(order) * (int)blockDim.x + threadIdx.x;
  auto put = [&](int d, T x) {
    if (REBASE) eg[d] = (float)((q[d] + (T)eg[d]) - x);
    q[d] = x;
  };
  if (g < nc) {
    const int4 r = cr[g];
    const int post = cpost ? cpost[g] : 0;
    // a plain copy (or a copy + 0.0) of an interior cell: the ghost the
    // FAS cycle holds is the same copy of the same value, the source having
    // been updated by the same FP64 addition, so the offset is exactly zero
    // and neither old value needs reading
    const bool plain = REBASE && (post == 0 || post == 1);
    for (int v = 0; v < nv; v++) {
      const int d = r.x + v * r.z;
      const T x = ghost_post<T>(q[r.y + v * r.w], post, v, homog);
      if (plain) {
        eg[d] = 0.0f;
        q[d] = x;
      } else {
        put(d, x);
      }
    }
    return;
  }
  if (g >= nc + ni) return;
  const GRule R = ir[g - nc];
  const int kind = R.kind & 0xff, post = R.kind >> 8;
  for (int v = 0; v < nv; v++) put(R.dst + v * R.pvs, ghost_post<T>(interp_value<T>(R, kind, v, q), post, v, homog));
}

// The ghost shell of every patch of a level (domain kFineG: the
// (mx+2)^d box, interior skipped).  STASH 1: save = q + eg (FP64);
// STASH 0: eg = save - q.  The staged-fill form of REBASE.
template <int DIM, int STASH>
__global__ __launch_bounds__(kBlock) void k_ghost_stash(LevelArgs L, int nv, double* __restrict__ q,
                                                        float* __restrict__ eg, double* __restrict__ save) {
  int pi, local;
  if (!level_elem(L, level_block(L), pi, local)) return;
  const DPatch P = L.pt[L.patches[pi]];
  const int i = local % P.nq[0], r = local / P.nq[0];
  const int j = r % P.nq[1], k = (DIM == 3) ? r / P.nq[1] : 0;
  const bool ghost = i == 0 || i == P.mx[0] + 1 || j == 0 || j == P.mx[1] + 1 ||
                     (DIM == 3 && (k == 0 || k == P.mx[2] + 1));
  if (!ghost) return;
  const int o = qidx(P, i, j, k);
  for (int v = 0; v < nv; v++) {
    const int d = o + v * P.qvs;
    if (STASH) save[d] = q[d] + (double)eg[d];
    else eg[d] = (float)(save[d] - q[d]);
  }
}

// Box copies of one ghost-fill stage: sibling transfers (GhostOverlap,
// boundary.f90:359-431), self-periodic wraps (SelfPeriodicBC :480-577),
// periodic transfers (TransferPeriodicBC :580-611) and physical boundary
// conditions (SetPhysicalBC + UserBC).  Cell c of a box (x-fastest) is
// done for every variable, with the index arithmetic once per cell; the
// cells of one box are independent (its sources are interior or
// earlier-written cells, its destinations distinct).
// `homog` selects the boundary condition of a correction (mixed precision):
// q_ghost = 2 - q_in becomes e_ghost = -e_in.
template <typename T>
__host__ __device__ __forceinline__ T box_value(int op, int v, int homog, T x) {
  const T dir = homog ? -x : (T)2 - x;
  switch (op) {
    case kCopyPlus0: return x + (T)0;
    case kDirichlet2: return dir;
    case kBgMixed: return (v == 0) ? x : dir;
    case kTumourBC: return (v == 2) ? dir : ((v == 3) ? -x : x);
    default: return x;
  }
}
template <typename T>
__device__ __forceinline__ void box_cell(const DPatch& D, const DPatch& S, const BoxDesc& B, int c, int nv,
                                         int homog, T* __restrict__ q) {
  const int a = c % B.ext[0], r = c / B.ext[0];
  const int b = r % B.ext[1], z = r / B.ext[1];
  const int so = qidx(S, B.slo[0] + a, B.slo[1] + b, B.slo[2] + z);
  const int dof = qidx(D, B.dlo[0] + a, B.dlo[1] + b, B.dlo[2] + z);
  for (int v = 0; v < nv; v++) q[dof + v * D.qvs] = box_value<T>(B.op, v, homog, q[so + v * S.qvs]);
}

// One stage: box blockIdx.x, split over gridDim.y workgroups.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_boxops(const DPatch* __restrict__ pt,
                                                   const BoxDesc* __restrict__ boxes, int nv, int homog,
                                                   T* __restrict__ q) {
  const BoxDesc B = boxes[blockIdx.x];
  const DPatch D = pt[B.dp];
  const DPatch S = pt[B.sp];
  const int n = B.ext[0] * B.ext[1] * B.ext[2];
  for (int e = blockIdx.y * blockDim.x + threadIdx.x; e < n; e += blockDim.x * gridDim.y)
    box_cell<T>(D, S, B, e, nv, homog, q);
}

// Every stage of a small level's fill in one workgroup, with a barrier
// after each stage: the reference's sequential order in one launch (the
// physical conditions of the small levels at and below the root).  The
// boxes of one stage are independent (as the one-launch-per-stage path
// also relies on), so a stage is one flat range of cells; each thread loads
// up to kSeqCells cells of it, every variable, before it stores any, so a
// stage costs about one memory round trip and one barrier.
constexpr int kMaxStages = 8;
struct BoxSeq {
  const BoxDesc* d[kMaxStages];
  int n[kMaxStages];
  int ns;
};
constexpr int kSeqMaxBoxes = 48;  // boxes of a sequence staged in LDS
constexpr int kSeqCells = 4;
constexpr int kSeqMaxVars = 4;
template <typename T>
__device__ void run_boxseq(const DPatch* __restrict__ pt, const BoxSeq& S, int nv, int homog, T* __restrict__ q) {
  __shared__ BoxDesc sbox[kSeqMaxBoxes];
  __shared__ int scell[kSeqMaxBoxes];  // first cell of each box within its stage
  __shared__ int stot[kMaxStages];     // cells of each stage
  __shared__ DPatch spat[2];
  int nbox = 0;
  for (int st = 0; st < S.ns; st++) nbox += S.n[st];
  if (nbox > kSeqMaxBoxes) {
    for (int st = 0; st < S.ns; st++)
      for (int bx = 0; bx < S.n[st]; bx++) {
        const BoxDesc B = S.d[st][bx];
        const DPatch D = pt[B.dp];
        const DPatch Sp = pt[B.sp];
        const int n = B.ext[0] * B.ext[1] * B.ext[2];
        for (int e = threadIdx.x; e < n; e += blockDim.x) box_cell<T>(D, Sp, B, e, nv, homog, q);
        __syncthreads();
      }
    return;
  }
  if ((int)threadIdx.x < nbox) {
    int t = threadIdx.x, st = 0;
    while (t >= S.n[st]) {
      t -= S.n[st];
      st++;
    }
    sbox[threadIdx.x] = S.d[st][t];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int b = 0;
    for (int st = 0; st < S.ns; st++) {
      int acc = 0;
      for (int i = 0; i < S.n[st]; i++, b++) {
        scell[b] = acc;
        acc += sbox[b].ext[0] * sbox[b].ext[1] * sbox[b].ext[2];
      }
      stot[st] = acc;
    }
  }
  if (nbox > 0 && threadIdx.x >= 64 && threadIdx.x < 66)
    spat[threadIdx.x - 64] = pt[threadIdx.x == 65 ? sbox[0].sp : sbox[0].dp];
  __syncthreads();
  const int dp0 = nbox > 0 ? sbox[0].dp : -1, sp0 = nbox > 0 ? sbox[0].sp : -1;
  int b0 = 0;
  for (int st = 0; st < S.ns; st++) {
    const int nb = S.n[st], tot = stot[st];
    for (int base = threadIdx.x; base < tot; base += kSeqCells * blockDim.x) {
      T y[kSeqCells][kSeqMaxVars];
      int dof[kSeqCells], dvs[kSeqCells];
#pragma unroll
      for (int u = 0; u < kSeqCells; u++) {
        const int e = base + u * (int)blockDim.x;
        dof[u] = -1;
        dvs[u] = 0;
        if (e < tot) {
          int b = b0;
          while (b + 1 < b0 + nb && scell[b + 1] <= e) b++;
          const BoxDesc B = sbox[b];
          const DPatch D = B.dp == dp0 ? spat[0] : (B.dp == sp0 ? spat[1] : pt[B.dp]);
          const DPatch Sp = B.sp == sp0 ? spat[1] : (B.sp == dp0 ? spat[0] : pt[B.sp]);
          const int c = e - scell[b];
          const int a = c % B.ext[0], r = c / B.ext[0];
          const int bb = r % B.ext[1], z = r / B.ext[1];
          const int so = qidx(Sp, B.slo[0] + a, B.slo[1] + bb, B.slo[2] + z);
          dof[u] = qidx(D, B.dlo[0] + a, B.dlo[1] + bb, B.dlo[2] + z);
          dvs[u] = D.qvs;
#pragma unroll
          for (int v = 0; v < kSeqMaxVars; v++)
            if (v < nv) y[u][v] = box_value<T>(B.op, v, homog, q[so + v * Sp.qvs]);
        }
      }
#pragma unroll
      for (int u = 0; u < kSeqCells; u++)
        if (dof[u] >= 0) {
#pragma unroll
          for (int v = 0; v < kSeqMaxVars; v++)
            if (v < nv) q[dof[u] + v * dvs[u]] = y[u][v];
        }
    }
    __syncthreads();
    b0 += nb;
  }
}
// SetPhysicalBC of one patch with condition 10 on every face, in one pass
// of the workgroup over the ghost shell.  The staged fill runs the x faces,
// then y, then z, each face box spanning the ghost rows of the other
// dimensions and reading the cells one step inside; so a ghost cell ends
// as the face operation applied once per dimension in which it lies
// outside, in x, y, z order, to the interior cell mirrored into the patch
// (0 -> 1, mx+1 -> mx) -- the same operations on the same value.  Shell
// cells: the two full z planes (3D), then per interior plane the two full
// y rows and the x ends of the interior rows.  Sources are interior cells
// only, so the cells are independent; each thread loads up to kSeqCells
// cells before it stores.
template <int DIM>
__host__ __device__ __forceinline__ int shell_cells(const DPatch& P) {
  const int caps = DIM == 3 ? 2 * P.nq[0] * P.nq[1] : 0;
  return caps + (DIM == 3 ? P.mx[2] : 1) * (2 * P.nq[0] + 2 * P.mx[1]);
}
// Shell cell e: its storage indices and its mirrored interior source; the
// value is the face operation applied once per outside dimension, x first.
template <int DIM>
__host__ __device__ __forceinline__ void shell_cell(const DPatch& P, int e, int& i, int& j, int& k, int& so,
                                                    int out[3]) {
  const int n0 = P.nq[0], n1 = P.nq[1], n2 = DIM == 3 ? P.nq[2] : 1;
  const int plane = n0 * n1, caps = DIM == 3 ? 2 * plane : 0;
  const int ring = 2 * n0 + 2 * P.mx[1];  // shell cells of one interior plane
  if (e < caps) {
    const int r = e % plane;
    k = e < plane ? 0 : n2 - 1;
    i = r % n0;
    j = r / n0;
  } else {
    const int e2 = e - caps, r = e2 % ring;
    k = DIM == 3 ? e2 / ring + 1 : 0;
    if (r < 2 * n0) {
      j = r < n0 ? 0 : n1 - 1;
      i = r % n0;
    } else {
      j = (r - 2 * n0) / 2 + 1;
      i = (r & 1) ? n0 - 1 : 0;
    }
  }
  out[0] = i == 0 || i == n0 - 1;
  out[1] = j == 0 || j == n1 - 1;
  out[2] = DIM == 3 && (k == 0 || k == n2 - 1);
  const int si = i == 0 ? 1 : (i == n0 - 1 ? P.mx[0] : i);
  const int sj = j == 0 ? 1 : (j == n1 - 1 ? P.mx[1] : j);
  const int sk = DIM == 3 ? (k == 0 ? 1 : (k == n2 - 1 ? P.mx[2] : k)) : 0;
  so = qidx(P, si, sj, sk);
}
template <int DIM, typename T>
__device__ void phys_shell(const DPatch& P, int op, int nv, int homog, T* __restrict__ q) {
  const int nshell = shell_cells<DIM>(P);
  for (int base = threadIdx.x; base < nshell; base += kSeqCells * blockDim.x) {
    T y[kSeqCells][kSeqMaxVars];
    int dof[kSeqCells];
#pragma unroll
    for (int u = 0; u < kSeqCells; u++) {
      const int e = base + u * (int)blockDim.x;
      dof[u] = -1;
      if (e < nshell) {
        int i, j, k, so, out[3];
        shell_cell<DIM>(P, e, i, j, k, so, out);
        dof[u] = qidx(P, i, j, k);
#pragma unroll
        for (int v = 0; v < kSeqMaxVars; v++)
          if (v < nv) {
            T x = q[so + v * P.qvs];
            if (out[0]) x = box_value<T>(op, v, homog, x);
            if (out[1]) x = box_value<T>(op, v, homog, x);
            if (out[2]) x = box_value<T>(op, v, homog, x);
            y[u][v] = x;
          }
      }
    }
#pragma unroll
    for (int u = 0; u < kSeqCells; u++)
      if (dof[u] >= 0) {
#pragma unroll
        for (int v = 0; v < kSeqMaxVars; v++)
          if (v < nv) q[dof[u] + v * P.qvs] = y[u][v];
      }
  }
}

template <typename T>
__global__ __launch_bounds__(1024) void k_boxops_seq(const DPatch* __restrict__ pt, BoxSeq S, int nv, int homog,
                                                     T* __restrict__ q) {
  run_boxseq<T>(pt, S, nv, homog, q);
}

// Self-periodic wrap of a patch periodic in every dimension (SelfPeriodicBC
// applied x, then y, then z): the net effect is that every ghost cell takes
// the value of its periodic image, plus 0.0 (the qpo offset of the x wrap)
// when the cell lies on an x ghost plane.  Domain: the ghost shell.
template <int DIM, typename T>
__global__ __launch_bounds__(kBlock) void k_wrap_full(const DPatch* __restrict__ pt, int pid, int nv,
                                                      T* __restrict__ q) {
  const DPatch P = pt[pid];
  const int nx = P.nq[0], ny = P.nq[1], nz = P.nq[2];
  const int g = xcd_block() * kBlock + threadIdx.x;
  // shell enumeration: x planes (2*ny*nz), then y planes (2*(nx-2)*nz), then
  // z planes (2*(nx-2)*(ny-2)) in 3D
  const int nxp = 2 * ny * nz, nyp = 2 * (nx - 2) * nz, nzp = (DIM == 3) ? 2 * (nx - 2) * (ny - 2) : 0;
  if (g >= nxp + nyp + nzp) return;
  int i, j, k;
  if (g < nxp) {
    i = (g & 1) ? nx - 1 : 0;
    const int r = g >> 1;
    j = r % ny;
    k = r / ny;
  } else if (g < nxp + nyp) {
    const int r0 = g - nxp;
    j = (r0 & 1) ? ny - 1 : 0;
    const int r = r0 >> 1;
    i = r % (nx - 2) + 1;
    k = r / (nx - 2);
  } else {
    const int r0 = g - nxp - nyp;
    k = (r0 & 1) ? nz - 1 : 0;
    const int r = r0 >> 1;
    i = r % (nx - 2) + 1;
    j = r / (nx - 2) + 1;
  }
  const int mx = P.mx[0], my = P.mx[1], mz = P.mx[2];
  const int si = (i == 0) ? mx : (i == mx + 1 ? 1 : i);
  const int sj = (j == 0) ? my : (j == my + 1 ? 1 : j);
  const int sk = (DIM == 3) ? ((k == 0) ? mz : (k == mz + 1 ? 1 : k)) : 0;
  const bool xg = (i == 0 || i == mx + 1);
  for (int v = 0; v < nv; v++) {
    const T x = q[qidx(P, si, sj, sk) + v * P.qvs];
    q[qidx(P, i, j, k) + v * P.qvs] = xg ? x + (T)0 : x;
  }
}

// ---------------------------------------------------------------------------
// Mixed precision: FP32 correction V-cycles for the CH step linearised about
// the FP64 iterate q.  With e = (e_phi, e_mu) and c = 3*phi^2 the
// correction operator is
//   A1(e) = e_phi - (dt/h^2) ULap(e_mu)
//   A2(e) = e_mu - c e_phi + (eps^2/h^2) ULap(e_phi),
// the Jacobian of Operator2D/3D (CH/problem.f90:393-422, 638-669); its
// pointwise smoother is the 2x2 solve of RelaxGrid2D/3D with the cubic
// term dropped.  The outer FP64 loop forms r = f - L(q), solves A e = r
// approximately with one FP32 V-cycle, and updates q += e.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_zero_f(float* __restrict__ p, long n) {
  long g = (long)blockIdx.x * kBlock + threadIdx.x;
  if (g < n) p[g] = 0.0f;
}

// q(FP64) += e(FP32) over a contiguous span (ghost cells included), and
// e = 0 there for the next correction cycle.
__global__ __launch_bounds__(kBlock) void k_add_corr(double* __restrict__ q, float* __restrict__ e, long n) {
  long g = (long)blockIdx.x * kBlock + threadIdx.x;
  if (g < n) {
    q[g] = q[g] + (double)e[g];
    e[g] = 0.0f;
  }
}

// c = 3 phi^2 over the interior of every patch of a level; with TM (the
// tumour model) also t0 = (1 + phi_old)/2 in variable 1.
template <int DIM, bool TM = false>
__global__ __launch_bounds__(kBlock) void k_coef(LevelArgs L, const double* __restrict__ q,
                                                 const double* __restrict__ qold, float* __restrict__ coef) {
  int pi, local;
  if (!level_elem(L, level_block(L), pi, local)) return;
  const DPatch P = L.pt[L.patches[pi]];
  int i, j, k;
  if (!cell_of<DIM>(P, local, i, j, k)) return;
  const int o = qidx(P, i, j, k);
  const double p = q[o];
  coef[o] = (float)(3.0 * (p * p));
  if (TM) coef[o + P.qvs] = (float)(0.5 * (qold[o] + 1.0));
}

template <int DIM>
__device__ __forceinline__ float ulap_f(const float* __restrict__ a, const DPatch& P, int i, int j, int k,
                                        int off) {
  float s = a[qidx(P, i + 1, j, k) + off] + a[qidx(P, i - 1, j, k) + off] + a[qidx(P, i, j + 1, k) + off] +
            a[qidx(P, i, j - 1, k) + off];
  if (DIM == 3) {
    s = s + a[qidx(P, i, j, k + 1) + off] + a[qidx(P, i, j, k - 1) + off];
    return s - 6.0f * a[qidx(P, i, j, k) + off];
  }
  return s - 4.0f * a[qidx(P, i, j, k) + off];
}
template <int DIM>
__device__ __forceinline__ float ulap_cf(const float* __restrict__ a, const DPatch& P, int i, int j, int k,
                                         int off) {
  float s = a[cidx(P, i + 1, j, k) + off] + a[cidx(P, i - 1, j, k) + off] + a[cidx(P, i, j + 1, k) + off] +
            a[cidx(P, i, j - 1, k) + off];
  if (DIM == 3) {
    s = s + a[cidx(P, i, j, k + 1) + off] + a[cidx(P, i, j, k - 1) + off];
    return s - 6.0f * a[cidx(P, i, j, k) + off];
  }
  return s - 4.0f * a[cidx(P, i, j, k) + off];
}

// EMODE 1: the correction starts from zero, so the other colour's values
// are taken as zero and not read, and the other cell of the thread's pair
// (2t+1, 2t+2) is set to zero, since the coarse-fine interpolation after
// this sweep reads it; 2 (the finest level's last sweep): the
// new correction and that of the other cell of the thread's pair (2t+1,
// 2t+2), final since the previous sweep, are added into the FP64 solution
// q64 (interior cells; even mx).
template <int DIM, bool WRAP, int EMODE = 0, bool IMG = false>
__global__ __launch_bounds__(kBlock) void k_relax_lin(LevelArgs L, int c, float tmp1, float tmp2, float tmp3,
                                                      float a12, float* __restrict__ e,
                                                      const float* __restrict__ rhs,
                                                      const float* __restrict__ coef,
                                                      double* __restrict__ q64 = nullptr, ImgTab I = ImgTab{}) {
  // IMG: the workgroup then writes the ghost images of its updated cells
  // (write_images), so no thread leaves before that
  const int lb = level_block(L);
  int pi, local;
  float n0 = 0.0f, n1 = 0.0f;
  bool ok = level_elem(L, lb, pi, local);
  if (!IMG && !ok) return;
  if (ok) {
    const DPatch P = L.pt[L.patches[pi]];
    int i, j, k, t;
    ok = half_cell<DIM>(P, c, local, i, j, k, t);
    if (!IMG && !ok) return;
    if (ok) {
      const int ib = (c + j + k) & 1;
      const int vs = P.qvs;
      const Stencil S = stencil_of<DIM>(P, c, i, j, k, t);
      const int o = S.o;
      float smu = 0.0f, sphi = 0.0f;
      if (EMODE == 1) {
        // no interior neighbour holds a correction yet; a ghost neighbour
        // holds the offset of the base state's ghosts from a fresh fill
        // (the values the FAS cycle's first sweep reads, rebase_ghosts)
        auto add = [&](int x) {
          smu += e[x + vs];
          sphi += e[x];
        };
        if (i == 1) add(S.xm);
        if (i == P.mx[0]) add(S.xp);
        if (j == 1) add(S.ym);
        if (j == P.mx[1]) add(S.yp);
        if (DIM == 3 && k == 1) add(S.zm);
        if (DIM == 3 && k == P.mx[2]) add(S.zp);
      } else {
        smu = e[S.xm + vs] + e[S.xp + vs] + e[S.ym + vs] + e[S.yp + vs];
        sphi = e[S.xp] + e[S.xm] + e[S.yp] + e[S.ym];
        if (DIM == 3) {
          smu = smu + e[S.zm + vs] + e[S.zp + vs];
          sphi = sphi + e[S.zp] + e[S.zm];
        }
      }
      ch_corr_update(coef[o], EMODE == 1 ? 0.0f : e[o], rhs[o], rhs[o + vs], smu, sphi, tmp1, tmp2, tmp3, a12, n0,
                     n1);
      e[o] = n0;
      e[o + vs] = n1;
      if (WRAP) wrap_images<DIM, float>(e, P, i, j, k, n0, n1);
      // the other cell of the thread's pair (2t+1, 2t+2) is the other
      // colour's element t + ib of its row
      if (EMODE == 1) {
        const int po = S.xm + ib;
        e[po] = 0.0f;
        e[po + vs] = 0.0f;
      }
      if (EMODE == 2) {
        q64[o] = q64[o] + (double)n0;
        q64[o + vs] = q64[o + vs] + (double)n1;
        const int po = S.xm + ib;
        q64[po] = q64[po] + (double)e[po];
        q64[po + vs] = q64[po + vs] + (double)e[po + vs];
      }
    }
  }
  if (IMG) {
    __shared__ float nv[2][kBlock];
    nv[0][threadIdx.x] = n0;
    nv[1][threadIdx.x] = n1;
    write_images<float, 2>(I, lb, nv, e);
  }
}

// Coarse right-hand side of the correction V-cycle (the FAS form of
// CoarseLoadingFunction, afasroutines.f90:673-736, for the operator A):
// parent rhs(mb) = R(rhs - A_h e) + A_2h(ec).
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_fas_rhs_lin(LevelArgs L, double dt, double eps2,
                                                        const float* __restrict__ e,
                                                        const float* __restrict__ rhs,
                                                        const float* __restrict__ ec,
                                                        const float* __restrict__ coef,
                                                        float* __restrict__ rpar) {
  CoarseCell c;
  if (!decode_coarse<DIM>(L, level_block(L), c)) return;
  const DPatch P = L.pt[c.pid];
  const DPatch Q = L.pt[P.parent];
  const float t1 = (float)(dt / (P.dx * P.dx)), t2 = (float)(eps2 / (P.dx * P.dx));
  const int nch = (DIM == 3) ? 8 : 4;
  float s0 = 0.0f, s1 = 0.0f;
#pragma unroll
  for (int n = 0; n < nch; n++) {
    const int i = 2 * c.ic - (n & 1), j = 2 * c.jc - ((n >> 1) & 1);
    const int k = (DIM == 3) ? 2 * c.kc - ((n >> 2) & 1) : 0;
    const int o = qidx(P, i, j, k);
    float a1, a2;
    ch_corr_op(coef[o], e[o], e[o + P.qvs], ulap_f<DIM>(e, P, i, j, k, 0), ulap_f<DIM>(e, P, i, j, k, P.qvs), t1,
               t2, a1, a2);
    s0 += rhs[o] - a1;
    s1 += rhs[o + P.qvs] - a2;
  }
  const float w = (DIM == 3) ? 0.125f : 0.25f;
  const int ip = P.mb[0][0] + c.ic - 1, jp = P.mb[1][0] + c.jc - 1;
  const int kp = (DIM == 3) ? P.mb[2][0] + c.kc - 1 : 0;
  const int opar = qidx(Q, ip, jp, kp);
  const int oc = cidx(P, c.ic, c.jc, c.kc);
  const float T1 = (float)(dt / (Q.dx * Q.dx)), T2 = (float)(eps2 / (Q.dx * Q.dx));
  float a1, a2;
  ch_corr_op(coef[opar], ec[oc], ec[oc + P.cvs], ulap_cf<DIM>(ec, P, c.ic, c.jc, c.kc, 0),
             ulap_cf<DIM>(ec, P, c.ic, c.jc, c.kc, P.cvs), T1, T2, a1, a2);
  rpar[opar] = w * s0 + a1;
  rpar[opar + Q.qvs] = w * s1 + a2;
}

// The tumour model's correction operator: the Jacobian of its Operator,
//   A1(e) = e_phi - (dt/h^2) ULap(e_mu) - dt lamp t0 e_n,
//   A2(e) = e_mu - c e_phi + (eps^2/h^2) ULap(e_phi),
//   A3(e) = -ULap(e_n)/h^2 + kappa t0 e_n,
//   A4(e) = -ULap(e_p)/h^2 - lamp t0 e_n,
// with the frozen c = 3 phi^2 (coef variable 0) and the lagged tumour
// fraction t0 = (1 + phi_old)/2 (coef variable 1) of the cell.  CG selects
// the qc geometry.
template <int DIM, bool CG>
__device__ __forceinline__ void lin_tm(const float* __restrict__ a, const DPatch& P, int i, int j, int k,
                                       float cc, float t0, double h, const ModelK& m, float out[4]) {
  const int vs = CG ? P.cvs : P.qvs;
  auto lap = [&](int off) { return CG ? ulap_cf<DIM>(a, P, i, j, k, off) : ulap_f<DIM>(a, P, i, j, k, off); };
  const int o = CG ? cidx(P, i, j, k) : qidx(P, i, j, k);
  const double h2 = h * h;
  const float t1 = (float)(m.dt / h2), t2 = (float)(m.eps2 / h2), hh = (float)(1.0 / h2);
  const float e0 = a[o], e1 = a[o + vs], e2 = a[o + 2 * vs];
  out[0] = e0 - t1 * lap(vs) - (float)(m.dt * m.gam) * t0 * e2;
  out[1] = e1 - cc * e0 + t2 * lap(0);
  out[2] = -hh * lap(2 * vs) + (float)m.cz * t0 * e2;
  out[3] = -hh * lap(3 * vs) - (float)m.gam * t0 * e2;
}

// One cell (i, j, k) of the tumour correction smoother; `ip` is the other
// cell of the thread's pair (EMODE 1, 2).
template <int DIM, int EMODE>
__device__ __forceinline__ void relax_lin_tm_cell(const DPatch& P, int i, int j, int k, int ip, const ModelK& m,
                                                  float* __restrict__ e, const float* __restrict__ rhs,
                                                  const float* __restrict__ coef, double* __restrict__ q64) {
  const double h2 = P.dx * P.dx, dd = (DIM == 3) ? 6.0 : 4.0;
  const float cdt = (float)(m.dt / h2), ceps = (float)(m.eps2 / h2), cdiag = (float)(dd * m.eps2 / h2);
  const float c12 = (float)(dd * (m.dt / h2)), chh = (float)(1.0 / h2), cd = (float)(dd / h2);
  const int vs = P.qvs;
  const int o = qidx(P, i, j, k);
  float s[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  if (EMODE != 1) {
    const int xm = qidx(P, i - 1, j, k), xp = qidx(P, i + 1, j, k);
    const int ym = qidx(P, i, j - 1, k), yp = qidx(P, i, j + 1, k);
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int off = v * vs;
      s[v] = e[xm + off] + e[xp + off] + e[ym + off] + e[yp + off];
      if (DIM == 3) s[v] = s[v] + e[qidx(P, i, j, k - 1) + off] + e[qidx(P, i, j, k + 1) + off];
    }
  }
  const float t0 = coef[o + vs];
  const float en = (rhs[o + 2 * vs] + chh * s[2]) / (cd + (float)m.cz * t0);
  const float c21 = -coef[o] - cdiag;
  const float r1 = rhs[o] + cdt * s[1] + (float)(m.dt * m.gam) * t0 * en;
  const float r2 = rhs[o + vs] - ceps * s[0];
  const float d = 1.0f - c12 * c21;
  const float n[4] = {(r1 - c12 * r2) / d, (r2 - c21 * r1) / d, en,
                      (rhs[o + 3 * vs] + chh * s[3] + (float)m.gam * t0 * en) / cd};
  const int po = qidx(P, ip, j, k);
#pragma unroll
  for (int v = 0; v < 4; v++) {
    e[o + v * vs] = n[v];
    if (EMODE == 1) e[po + v * vs] = 0.0f;
    if (EMODE == 2) {
      q64[o + v * vs] = q64[o + v * vs] + (double)n[v];
      q64[po + v * vs] = q64[po + v * vs] + (double)e[po + v * vs];
    }
  }
}

// The tumour model's correction smoother: the pointwise update of
// k_relax_tm (nutrient first, then the 2x2 phi/mu solve, then pressure) for
// the operator A above.  EMODE as in k_relax_lin, over all four variables.
template <int DIM, int EMODE = 0, bool IMG = false>
__global__ __launch_bounds__(kBlock) void k_relax_lin_tm(LevelArgs L, int c, ModelK m, float* __restrict__ e,
                                                         const float* __restrict__ rhs,
                                                         const float* __restrict__ coef,
                                                         double* __restrict__ q64 = nullptr) {
  int pi, local;
  if (!level_elem(L, level_block(L), pi, local)) return;
  const DPatch P = L.pt[L.patches[pi]];
  const int cnt = (P.mx[0] + 1) >> 1;
  const int t = local % cnt;
  const int row = local / cnt;
  const int j = row % P.mx[1] + 1;
  const int k = (DIM == 3) ? row / P.mx[1] + 1 : 0;
  const int ib = (c + j + k) & 1;
  const int i = 2 * t + 2 - ib;
  if (i > P.mx[0]) return;
  relax_lin_tm_cell<DIM, EMODE>(P, i, j, k, 2 * t + 1 + ib, m, e, rhs, coef, q64);
  if (IMG) phys_cell_images<DIM, float>(P, kTumourBC, i, j, k, 4, 1, e);
}


// Coarse right-hand side of the tumour model's correction V-cycle:
// parent rhs(mb) = R(rhs - A_h e) + A_2h(ec), as k_fas_rhs_lin.
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_fas_rhs_lin_tm(LevelArgs L, ModelK m, const float* __restrict__ e,
                                                           const float* __restrict__ rhs,
                                                           const float* __restrict__ ec,
                                                           const float* __restrict__ coef,
                                                           float* __restrict__ rpar) {
  CoarseCell c;
  if (!decode_coarse<DIM>(L, level_block(L), c)) return;
  const DPatch P = L.pt[c.pid];
  const DPatch Q = L.pt[P.parent];
  const int nch = (DIM == 3) ? 8 : 4;
  float s[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int n = 0; n < nch; n++) {
    const int i = 2 * c.ic - (n & 1), j = 2 * c.jc - ((n >> 1) & 1);
    const int k = (DIM == 3) ? 2 * c.kc - ((n >> 2) & 1) : 0;
    const int o = qidx(P, i, j, k);
    float a[4];
    lin_tm<DIM, false>(e, P, i, j, k, coef[o], coef[o + P.qvs], P.dx, m, a);
#pragma unroll
    for (int v = 0; v < 4; v++) s[v] += rhs[o + v * P.qvs] - a[v];
  }
  const float w = (DIM == 3) ? 0.125f : 0.25f;
  const int ip = P.mb[0][0] + c.ic - 1, jp = P.mb[1][0] + c.jc - 1;
  const int kp = (DIM == 3) ? P.mb[2][0] + c.kc - 1 : 0;
  const int opar = qidx(Q, ip, jp, kp);
  float a[4];
  lin_tm<DIM, true>(ec, P, c.ic, c.jc, c.kc, coef[opar], coef[opar + Q.qvs], Q.dx, m, a);
#pragma unroll
  for (int v = 0; v < 4; v++) rpar[opar + v * Q.qvs] = w * s[v] + a[v];
}

// ---------------------------------------------------------------------------
// Coarse end of the V-cycle in one workgroup.  Below a level with at most
// 16^3 (3D) or 64^2 (2D) cells every level is one patch periodic in every
// dimension (the multigrid levels under the root, bsamroutines.f90:443-510),
// and AFASVcycle (afasroutines.f90:495-529) on those levels is a chain of
// small dependent steps.  This kernel runs that chain -- LevelRelax,
// RestrictSolution, SetGhost(L-1,0), GetCoarseGhostPoints, the coarse load,
// recursion, CorrectFine, SetGhost(L,0), LevelRelax -- with workgroup
// barriers instead of kernel launches.  Each step uses the same per-cell
// expressions as the level-wide kernels, so the results are bit-identical.
// LIN=false: the FP64 CH V-cycle; LIN=true: the FP32 correction V-cycle.
// ---------------------------------------------------------------------------
struct CoarseArgs {
  int pid[16];       // patch of each level, index 0 = the top level of the chain
  int nlev;          // levels in the chain (top .. minlevel)
  double dt, eps2;
};

// img: also write the periodic ghost images of the updated cells (only for
// even extents, where every image has the cell's colour; see wrap_images).
template <int DIM, typename T, bool LIN>
__device__ void cv_relax(const DPatch& P, int c, double dt, double eps2, T* __restrict__ q,
                         const T* __restrict__ f, const T* __restrict__ coef, bool img) {
  const double h2 = P.dx * P.dx;
  const double d2 = (DIM == 3) ? 6.0 : 4.0;
  const T tmp1 = (T)(dt / h2), tmp2 = (T)(eps2 / h2), tmp3 = (T)(d2 * eps2 / h2);
  const T a12 = (T)(d2 * (dt / h2));
  const int cnt = (P.mx[0] + 1) >> 1;
  const int total = cnt * P.mx[1] * (DIM == 3 ? P.mx[2] : 1);
  const int vs = P.qvs;
  for (int g = threadIdx.x; g < total; g += blockDim.x) {
    const int t = g % cnt, row = g / cnt;
    const int j = row % P.mx[1] + 1;
    const int k = (DIM == 3) ? row / P.mx[1] + 1 : 0;
    const int ib = (c + j + k) & 1;
    const int i = 2 * t + 2 - ib;
    if (i > P.mx[0]) continue;
    const int o = qidx(P, i, j, k);
    const int xm = qidx(P, i - 1, j, k), xp = qidx(P, i + 1, j, k);
    const int ym = qidx(P, i, j - 1, k), yp = qidx(P, i, j + 1, k);
    T smu, sphi;
    if (DIM == 3) {
      const int zm = qidx(P, i, j, k - 1), zp = qidx(P, i, j, k + 1);
      smu = q[xm + vs] + q[xp + vs] + q[ym + vs] + q[yp + vs] + q[zm + vs] + q[zp + vs];
      sphi = q[xp] + q[xm] + q[yp] + q[ym] + q[zp] + q[zm];
    } else {
      smu = q[xm + vs] + q[xp + vs] + q[ym + vs] + q[yp + vs];
      sphi = q[xp] + q[xm] + q[yp] + q[ym];
    }
    T n0, n1;
    if constexpr (LIN) {
      ch_corr_update(coef[o], q[o], f[o], f[o + vs], smu, sphi, tmp1, tmp2, tmp3, a12, n0, n1);
    } else {
      const T phi = q[o];
      const T a21 = (T)-3 * (phi * phi) - tmp3;
      const T b1 = f[o] + tmp1 * smu;
      const T b2 = f[o + vs] - tmp2 * sphi - (T)2 * (phi * phi * phi);
      const T det = (T)1 - a12 * a21;
      n0 = (b1 - a12 * b2) / det;
      n1 = (b2 - a21 * b1) / det;
    }
    q[o] = n0;
    q[o + vs] = n1;
    if (img) wrap_images<DIM, T>(q, P, i, j, k, n0, n1);
  }
}

template <int DIM, typename T>
__device__ void cv_wrap(const DPatch& P, T* __restrict__ q) {
  const int nx = P.nq[0], ny = P.nq[1], nz = P.nq[2];
  const int nxp = 2 * ny * nz, nyp = 2 * (nx - 2) * nz, nzp = (DIM == 3) ? 2 * (nx - 2) * (ny - 2) : 0;
  const int mx = P.mx[0], my = P.mx[1], mz = P.mx[2];
  for (int g = threadIdx.x; g < nxp + nyp + nzp; g += blockDim.x) {
    int i, j, k;
    if (g < nxp) {
      i = (g & 1) ? nx - 1 : 0;
      const int r = g >> 1;
      j = r % ny;
      k = r / ny;
    } else if (g < nxp + nyp) {
      const int r0 = g - nxp;
      j = (r0 & 1) ? ny - 1 : 0;
      const int r = r0 >> 1;
      i = r % (nx - 2) + 1;
      k = r / (nx - 2);
    } else {
      const int r0 = g - nxp - nyp;
      k = (r0 & 1) ? nz - 1 : 0;
      const int r = r0 >> 1;
      i = r % (nx - 2) + 1;
      j = r / (nx - 2) + 1;
    }
    const int si = (i == 0) ? mx : (i == mx + 1 ? 1 : i);
    const int sj = (j == 0) ? my : (j == my + 1 ? 1 : j);
    const int sk = (DIM == 3) ? ((k == 0) ? mz : (k == mz + 1 ? 1 : k)) : 0;
    const bool xg = (i == 0 || i == mx + 1);
    for (int v = 0; v < 2; v++) {
      const T x = q[qidx(P, si, sj, sk) + v * P.qvs];
      q[qidx(P, i, j, k) + v * P.qvs] = xg ? x + (T)0 : x;
    }
  }
}

template <int DIM, typename T, bool LIN>
__device__ void cv_down(const DPatch& P, const DPatch& Q, double dt, double eps2, T* __restrict__ q,
                        T* __restrict__ f, T* __restrict__ qc, T* __restrict__ coef) {
  const int ncc = P.cmx[0] * P.cmx[1] * (DIM == 3 ? P.cmx[2] : 1);
  // RestrictSolution: qc = R q, parent q(mb) = qc
  for (int g = threadIdx.x; g < ncc; g += blockDim.x) {
    const int ic = g % P.cmx[0] + 1, row = g / P.cmx[0];
    const int jc = row % P.cmx[1] + 1, kc = (DIM == 3) ? row / P.cmx[1] + 1 : 0;
    const int ip = P.mb[0][0] + ic - 1, jp = P.mb[1][0] + jc - 1, kp = (DIM == 3) ? P.mb[2][0] + kc - 1 : 0;
    for (int v = 0; v < 2; v++) {
      const T r = restrict_cell<DIM, T>(q, P, ic, jc, kc, v * P.qvs);
      qc[cidx(P, ic, jc, kc) + v * P.cvs] = r;
      q[qidx(Q, ip, jp, kp) + v * Q.qvs] = r;
    }
    // the correction's base state follows the restriction: phi0 of the
    // coarse level is R(phi0) of this one
    if (LIN) coef[qidx(Q, ip, jp, kp)] = restrict_cell<DIM, T>(coef, P, ic, jc, kc, 0);
  }
  __syncthreads();
  cv_wrap<DIM, T>(Q, q);  // SetGhost(L-1, 0)
  __syncthreads();
  // GetCoarseGhostPoints
  const int nbox = P.nc[0] * P.nc[1] * P.nc[2];
  for (int g = threadIdx.x; g < nbox; g += blockDim.x) {
    const int ic = g % P.nc[0], row = g / P.nc[0];
    const int jc = row % P.nc[1], kc = (DIM == 3) ? row / P.nc[1] : 0;
    const bool interior = ic >= 1 && ic <= P.cmx[0] && jc >= 1 && jc <= P.cmx[1] &&
                          (DIM == 2 || (kc >= 1 && kc <= P.cmx[2]));
    if (interior) continue;
    const int ip = P.mb[0][0] - 1 + ic, jp = P.mb[1][0] - 1 + jc, kp = (DIM == 3) ? P.mb[2][0] - 1 + kc : 0;
    for (int v = 0; v < 2; v++) qc[cidx(P, ic, jc, kc) + v * P.cvs] = q[qidx(Q, ip, jp, kp) + v * Q.qvs];
  }
  __syncthreads();
  // coarse load: parent f(mb) = R(f - L_h q) + L_2h(qc)
  const int nch = (DIM == 3) ? 8 : 4;
  for (int g = threadIdx.x; g < ncc; g += blockDim.x) {
    const int ic = g % P.cmx[0] + 1, row = g / P.cmx[0];
    const int jc = row % P.cmx[1] + 1, kc = (DIM == 3) ? row / P.cmx[1] + 1 : 0;
    const int ip = P.mb[0][0] + ic - 1, jp = P.mb[1][0] + jc - 1, kp = (DIM == 3) ? P.mb[2][0] + kc - 1 : 0;
    const int opar = qidx(Q, ip, jp, kp);
    const int oc = cidx(P, ic, jc, kc);
    T s0 = (T)0, s1 = (T)0;
    if constexpr (LIN) {
      const T t1 = (T)(dt / (P.dx * P.dx)), t2 = (T)(eps2 / (P.dx * P.dx));
      for (int n = 0; n < nch; n++) {
        const int i = 2 * ic - (n & 1), j = 2 * jc - ((n >> 1) & 1), k = (DIM == 3) ? 2 * kc - ((n >> 2) & 1) : 0;
        const int o = qidx(P, i, j, k);
        T a1, a2;
        ch_corr_op(coef[o], q[o], q[o + P.qvs], ulap_f<DIM>(q, P, i, j, k, 0), ulap_f<DIM>(q, P, i, j, k, P.qvs), t1,
                   t2, a1, a2);
        s0 += f[o] - a1;
        s1 += f[o + P.qvs] - a2;
      }
      const T w = (DIM == 3) ? (T)0.125 : (T)0.25;
      const T T1 = (T)(dt / (Q.dx * Q.dx)), T2 = (T)(eps2 / (Q.dx * Q.dx));
      T a1, a2;
      ch_corr_op(coef[opar], qc[oc], qc[oc + P.cvs], ulap_cf<DIM>(qc, P, ic, jc, kc, 0),
                 ulap_cf<DIM>(qc, P, ic, jc, kc, P.cvs), T1, T2, a1, a2);
      f[opar] = w * s0 + a1;
      f[opar + Q.qvs] = w * s1 + a2;
    } else {
      ModelK m;
      m.dt = dt; m.eps2 = eps2; m.gam = 0.0; m.lam = 0.0; m.cz = 0.0;
      for (int n = 0; n < nch; n++) {
        const int i = 2 * ic - (n & 1), j = 2 * jc - ((n >> 1) & 1), k = (DIM == 3) ? 2 * kc - ((n >> 2) & 1) : 0;
        double o1, o2;
        op_q<DIM, BSAM_MODEL_CH>(q, q, P, i, j, k, m, P.dx, o1, o2);
        const int o = qidx(P, i, j, k);
        const T a0 = f[o] - o1, a1 = f[o + P.qvs] - o2;
        if (n == 0) { s0 = a0; s1 = a1; } else { s0 = s0 + a0; s1 = s1 + a1; }
      }
      const T r0 = ((DIM == 3) ? (T)0.125 : (T)0.25) * s0;
      const T r1 = ((DIM == 3) ? (T)0.125 : (T)0.25) * s1;
      const double lmu = ulap_c<DIM>(qc, P, ic, jc, kc, P.cvs);
      const double lphi = ulap_c<DIM>(qc, P, ic, jc, kc, 0);
      double o1, o2;
      op_from<DIM, BSAM_MODEL_CH>(qc[oc], qc[oc + P.cvs], lphi, lmu, 0.0, m, Q.dx, o1, o2);
      f[opar] = r0 + o1;
      f[opar + Q.qvs] = r1 + o2;
    }
  }
}

template <int DIM, typename T>
__device__ void cv_up(const DPatch& P, const DPatch& Q, T* __restrict__ q, T* __restrict__ qc) {
  // CorrectFine: qc = parent q(mb-1:mb+1) - qc, then q += P1(qc)
  const int nbox = P.nc[0] * P.nc[1] * P.nc[2];
  for (int g = threadIdx.x; g < nbox; g += blockDim.x) {
    const int ic = g % P.nc[0], row = g / P.nc[0];
    const int jc = row % P.nc[1], kc = (DIM == 3) ? row / P.nc[1] : 0;
    const int ip = P.mb[0][0] - 1 + ic, jp = P.mb[1][0] - 1 + jc, kp = (DIM == 3) ? P.mb[2][0] - 1 + kc : 0;
    for (int v = 0; v < 2; v++) {
      const int o = cidx(P, ic, jc, kc) + v * P.cvs;
      qc[o] = q[qidx(Q, ip, jp, kp) + v * Q.qvs] - qc[o];
    }
  }
  __syncthreads();
  const int n0 = P.cmx[0] + 1, n1 = P.cmx[1] + 1, n2 = (DIM == 3) ? P.cmx[2] + 1 : 1;
  for (int g = threadIdx.x; g < n0 * n1 * n2; g += blockDim.x) {
    const int ic = g % n0, row = g / n0;
    const int jc = row % n1, kc = (DIM == 3) ? row / n1 : 0;
    for (int v = 0; v < 2; v++) {
      const int cv = v * P.cvs, fv = v * P.qvs;
      if (DIM == 2) {
        const T a00 = qc[cidx(P, ic, jc, 0) + cv], a01 = qc[cidx(P, ic + 1, jc, 0) + cv];
        const T a10 = qc[cidx(P, ic, jc + 1, 0) + cv], a11 = qc[cidx(P, ic + 1, jc + 1, 0) + cv];
        for (int b = 0; b < 2; b++)
          for (int a = 0; a < 2; a++) {
            const int i = 2 * ic + a, j = 2 * jc + b;
            const T b0 = linw<T>(a00, a01, i), b1 = linw<T>(a10, a11, i);
            const T pr = linw<T>(b0, b1, j) / (T)16;
            const int o = qidx(P, i, j, 0) + fv;
            q[o] = q[o] + pr;
          }
      } else {
        T a[2][2][2];
        for (int d = 0; d < 2; d++)
          for (int b = 0; b < 2; b++)
            for (int e = 0; e < 2; e++) a[d][b][e] = qc[cidx(P, ic + e, jc + b, kc + d) + cv];
        for (int dk = 0; dk < 2; dk++)
          for (int db = 0; db < 2; db++)
            for (int da = 0; da < 2; da++) {
              const int i = 2 * ic + da, j = 2 * jc + db, k = 2 * kc + dk;
              T cc2[2];
              for (int d = 0; d < 2; d++) {
                const T b0 = linw<T>(a[d][0][0], a[d][0][1], i);
                const T b1 = linw<T>(a[d][1][0], a[d][1][1], i);
                cc2[d] = linw<T>(b0, b1, j);
              }
              const T pr = linw<T>(cc2[0], cc2[1], k) / (T)64;
              const int o = qidx(P, i, j, k) + fv;
              q[o] = q[o] + pr;
            }
      }
    }
  }
  __syncthreads();
  cv_wrap<DIM, T>(P, q);  // SetGhost(L, 0)
}

// LevelRelax on a chain level: each colour sweep followed by SetGhost, which
// on these levels is the periodic wrap.  With even extents the sweep writes
// the images itself; with an odd extent (a coarse level such as 50x9x4 or
// 12x6x5) the wrap runs after the sweep, as SelfPeriodicBC does.
template <int DIM, typename T, bool LIN>
__device__ void cv_level_relax(const DPatch& P, int nsm, double dt, double eps2, T* q, const T* f,
                               const T* coef) {
  const bool img = !(P.mx[0] & 1) && !(P.mx[1] & 1) && (DIM == 2 || !(P.mx[2] & 1));
  for (int s = 0; s < nsm; s++)
    for (int c = 0; c < 2; c++) {
      cv_relax<DIM, T, LIN>(P, c, dt, eps2, q, f, coef, img);
      __syncthreads();
      if (!img) {
        cv_wrap<DIM, T>(P, q);
        __syncthreads();
      }
    }
}

template <int DIM, typename T, bool LIN>
__global__ __launch_bounds__(1024) void k_coarse_vcycle(const DPatch* __restrict__ pt, CoarseArgs A, int nsm,
                                                        T* __restrict__ q, T* __restrict__ f,
                                                        T* __restrict__ qc, T* __restrict__ coef) {
  // descend
  for (int l = 0; l < A.nlev - 1; l++) {
    const DPatch P = pt[A.pid[l]];
    cv_level_relax<DIM, T, LIN>(P, nsm, A.dt, A.eps2, q, f, coef);
    cv_down<DIM, T, LIN>(P, pt[A.pid[l + 1]], A.dt, A.eps2, q, f, qc, coef);
    __syncthreads();
  }
  cv_level_relax<DIM, T, LIN>(pt[A.pid[A.nlev - 1]], nsm, A.dt, A.eps2, q, f, coef);
  // ascend
  for (int l = A.nlev - 2; l >= 0; l--) {
    const DPatch P = pt[A.pid[l]];
    cv_up<DIM, T>(P, pt[A.pid[l + 1]], q, qc);
    __syncthreads();
    cv_level_relax<DIM, T, LIN>(P, nsm, A.dt, A.eps2, q, f, coef);
  }
}

// The coarse end of the V-cycle in one workgroup for levels whose ghost
// fill is a short sequence of box stages (physical boundaries), with the
// tumour model: the steps of k_coarse_vcycle -- LevelRelax, RestrictSolution,
// SetGhost(L-1,0), GetCoarseGhostPoints, the coarse load, recursion,
// CorrectFine, SetGhost(L,0), LevelRelax -- each with the per-cell
// expressions of the level-wise kernels (relax_tm_cell / relax_lin_tm_cell,
// restrict_cell, op_cell and op_tm_from or lin_tm, the prolongation of
// k_prolong_dual) and the box stages of k_boxops_seq, so the results are
// bit-identical.  LIN=false: FP64; LIN=true: the FP32 correction cycle.
constexpr int kMaxChain = 12;
// The chain's levels, fixed for a hierarchy: set where the chain is planned
// and read by the kernel from device memory (indexing a by-value kernel
// argument of this size by level would copy it to every lane's scratch).
struct TmChainArgs {
  int pid[kMaxChain];
  BoxSeq seq[kMaxChain];
  int physop[kMaxChain];  // face operation when every face has condition 10, else -1
  int nlev;
};

// SetGhost(level, ipass) of a chain level: the face stages in order, or,
// for a patch with condition 10 on every face, one pass over the shell.
template <int DIM, typename T>
__device__ __forceinline__ void tc_fill(const DPatch* __restrict__ pt, const TmChainArgs& A, int l, int homog,
                                        T* __restrict__ q) {
  if (A.physop[l] >= 0) {
    const DPatch P = pt[A.pid[l]];
    phys_shell<DIM, T>(P, A.physop[l], 4, homog, q);
    __syncthreads();
  } else {
    run_boxseq<T>(pt, A.seq[l], 4, homog, q);
  }
}

template <int DIM, bool LIN>
__device__ __forceinline__ void tc_level_relax(const DPatch* __restrict__ pt, const TmChainArgs& A, const ModelK& m,
                                               int l, int nsm,
                               typename std::conditional<LIN, float, double>::type* __restrict__ q,
                               const typename std::conditional<LIN, float, double>::type* __restrict__ f,
                               const double* __restrict__ qold, const float* __restrict__ coef) {
  using T = typename std::conditional<LIN, float, double>::type;
  const DPatch P = pt[A.pid[l]];
  const int cnt = (P.mx[0] + 1) >> 1;
  const int total = cnt * P.mx[1] * (DIM == 3 ? P.mx[2] : 1);
  for (int s = 0; s < nsm; s++)
    for (int c = 0; c < 2; c++) {
      for (int g = threadIdx.x; g < total; g += blockDim.x) {
        const int t = g % cnt, row = g / cnt;
        const int j = row % P.mx[1] + 1;
        const int k = (DIM == 3) ? row / P.mx[1] + 1 : 0;
        const int ib = (c + j + k) & 1;
        const int i = 2 * t + 2 - ib;
        if (i > P.mx[0]) continue;
        if constexpr (LIN) relax_lin_tm_cell<DIM, 0>(P, i, j, k, 0, m, q, f, coef, nullptr);
        else relax_tm_cell<DIM>(P, i, j, k, m, q, qold, f);
      }
      __syncthreads();
      tc_fill<DIM, T>(pt, A, l, LIN ? 1 : 0, q);
    }
}

template <int DIM, bool LIN>
__global__ __launch_bounds__(1024) void k_coarse_vcycle_tm(
    const DPatch* __restrict__ pt, const TmChainArgs* __restrict__ Ap, ModelK m, int nsm,
    typename std::conditional<LIN, float, double>::type* __restrict__ q,
    typename std::conditional<LIN, float, double>::type* __restrict__ f,
    typename std::conditional<LIN, float, double>::type* __restrict__ qc, const double* __restrict__ qold,
    const float* __restrict__ coef) {
  using T = typename std::conditional<LIN, float, double>::type;
  constexpr int NV = 4;
  const TmChainArgs& A = *Ap;
  // descend
  for (int l = 0; l < A.nlev - 1; l++) {
    tc_level_relax<DIM, LIN>(pt, A, m, l, nsm, q, f, qold, coef);
    const DPatch P = pt[A.pid[l]];
    const DPatch Q = pt[A.pid[l + 1]];
    const int ncc = P.cmx[0] * P.cmx[1] * (DIM == 3 ? P.cmx[2] : 1);
    // RestrictSolution: qc = R q, parent q(mb) = qc
    for (int g = threadIdx.x; g < ncc; g += blockDim.x) {
      const int ic = g % P.cmx[0] + 1, row = g / P.cmx[0];
      const int jc = row % P.cmx[1] + 1, kc = (DIM == 3) ? row / P.cmx[1] + 1 : 0;
      const int ip = P.mb[0][0] + ic - 1, jp = P.mb[1][0] + jc - 1, kp = (DIM == 3) ? P.mb[2][0] + kc - 1 : 0;
      for (int v = 0; v < NV; v++) {
        const T r = restrict_cell<DIM, T>(q, P, ic, jc, kc, v * P.qvs);
        qc[cidx(P, ic, jc, kc) + v * P.cvs] = r;
        q[qidx(Q, ip, jp, kp) + v * Q.qvs] = r;
      }
    }
    __syncthreads();
    tc_fill<DIM, T>(pt, A, l + 1, LIN ? 1 : 0, q);  // SetGhost(L-1, 0)
    // GetCoarseGhostPoints
    const int nbox = P.nc[0] * P.nc[1] * P.nc[2];
    for (int g = threadIdx.x; g < nbox; g += blockDim.x) {
      const int ic = g % P.nc[0], row = g / P.nc[0];
      const int jc = row % P.nc[1], kc = (DIM == 3) ? row / P.nc[1] : 0;
      const bool interior = ic >= 1 && ic <= P.cmx[0] && jc >= 1 && jc <= P.cmx[1] &&
                            (DIM == 2 || (kc >= 1 && kc <= P.cmx[2]));
      if (interior) continue;
      const int ip = P.mb[0][0] - 1 + ic, jp = P.mb[1][0] - 1 + jc, kp = (DIM == 3) ? P.mb[2][0] - 1 + kc : 0;
      for (int v = 0; v < NV; v++) qc[cidx(P, ic, jc, kc) + v * P.cvs] = q[qidx(Q, ip, jp, kp) + v * Q.qvs];
    }
    __syncthreads();
    // the coarse load: parent f(mb) = R(f - L_h q) + L_2h(qc)
    const int nch = (DIM == 3) ? 8 : 4;
    for (int g = threadIdx.x; g < ncc; g += blockDim.x) {
      const int ic = g % P.cmx[0] + 1, row = g / P.cmx[0];
      const int jc = row % P.cmx[1] + 1, kc = (DIM == 3) ? row / P.cmx[1] + 1 : 0;
      const int ip = P.mb[0][0] + ic - 1, jp = P.mb[1][0] + jc - 1, kp = (DIM == 3) ? P.mb[2][0] + kc - 1 : 0;
      const int opar = qidx(Q, ip, jp, kp);
      if constexpr (LIN) {
        float s[NV] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int n = 0; n < nch; n++) {
          const int i = 2 * ic - (n & 1), j = 2 * jc - ((n >> 1) & 1), k = (DIM == 3) ? 2 * kc - ((n >> 2) & 1) : 0;
          const int o = qidx(P, i, j, k);
          float a[NV];
          lin_tm<DIM, false>(q, P, i, j, k, coef[o], coef[o + P.qvs], P.dx, m, a);
          for (int v = 0; v < NV; v++) s[v] += f[o + v * P.qvs] - a[v];
        }
        const float w = (DIM == 3) ? 0.125f : 0.25f;
        float a[NV];
        lin_tm<DIM, true>(qc, P, ic, jc, kc, coef[opar], coef[opar + Q.qvs], Q.dx, m, a);
        for (int v = 0; v < NV; v++) f[opar + v * Q.qvs] = w * s[v] + a[v];
      } else {
        double s[NV];
        for (int n = 0; n < nch; n++) {
          const int i = 2 * ic - (n & 1), j = 2 * jc - ((n >> 1) & 1), k = (DIM == 3) ? 2 * kc - ((n >> 2) & 1) : 0;
          double o[kNvMax];
          op_cell<DIM, BSAM_MODEL_TUMOUR>(q, qold, P, i, j, k, m, P.dx, o);
          const int ofs = qidx(P, i, j, k);
          for (int v = 0; v < NV; v++) {
            const double a = f[ofs + v * P.qvs] - o[v];
            s[v] = (n == 0) ? a : s[v] + a;
          }
        }
        const int oc = cidx(P, ic, jc, kc), vs = P.cvs;
        const double l2 = ulap_c<DIM>(qc, P, ic, jc, kc, vs), l1 = ulap_c<DIM>(qc, P, ic, jc, kc, 0);
        const double l3 = ulap_c<DIM>(qc, P, ic, jc, kc, 2 * vs);
        const double l4 = ulap_c<DIM>(qc, P, ic, jc, kc, 3 * vs);
        double o[kNvMax];
        op_tm_from<DIM>(qc[oc], qc[oc + vs], qc[oc + 2 * vs], l1, l2, l3, l4, qold[opar], m, Q.dx, o);
        const double w = (DIM == 3) ? 0.125 : 0.25;
        for (int v = 0; v < NV; v++) f[opar + v * Q.qvs] = w * s[v] + o[v];
      }
    }
    __syncthreads();
  }
  tc_level_relax<DIM, LIN>(pt, A, m, A.nlev - 1, nsm, q, f, qold, coef);
  // ascend
  for (int l = A.nlev - 2; l >= 0; l--) {
    const DPatch P = pt[A.pid[l]];
    const DPatch Q = pt[A.pid[l + 1]];
    // CorrectFine: qc = parent q(mb-1:mb+1) - qc, then q += P1(qc)
    const int nbox = P.nc[0] * P.nc[1] * P.nc[2];
    for (int g = threadIdx.x; g < nbox; g += blockDim.x) {
      const int ic = g % P.nc[0], row = g / P.nc[0];
      const int jc = row % P.nc[1], kc = (DIM == 3) ? row / P.nc[1] : 0;
      const int ip = P.mb[0][0] - 1 + ic, jp = P.mb[1][0] - 1 + jc, kp = (DIM == 3) ? P.mb[2][0] - 1 + kc : 0;
      for (int v = 0; v < NV; v++) {
        const int o = cidx(P, ic, jc, kc) + v * P.cvs;
        qc[o] = q[qidx(Q, ip, jp, kp) + v * Q.qvs] - qc[o];
      }
    }
    __syncthreads();
    const int n0 = P.cmx[0] + 1, n1 = P.cmx[1] + 1, n2 = (DIM == 3) ? P.cmx[2] + 1 : 1;
    for (int g = threadIdx.x; g < n0 * n1 * n2; g += blockDim.x) {
      const int ic = g % n0, row = g / n0;
      const int jc = row % n1, kc = (DIM == 3) ? row / n1 : 0;
      for (int v = 0; v < NV; v++) {
        const int cv = v * P.cvs, fv = v * P.qvs;
        if (DIM == 2) {
          const T a00 = qc[cidx(P, ic, jc, 0) + cv], a01 = qc[cidx(P, ic + 1, jc, 0) + cv];
          const T a10 = qc[cidx(P, ic, jc + 1, 0) + cv], a11 = qc[cidx(P, ic + 1, jc + 1, 0) + cv];
          for (int b = 0; b < 2; b++)
            for (int a = 0; a < 2; a++) {
              const int i = 2 * ic + a, j = 2 * jc + b;
              const T b0 = linw<T>(a00, a01, i), b1 = linw<T>(a10, a11, i);
              const T pr = linw<T>(b0, b1, j) / (T)16;
              const int o = qidx(P, i, j, 0) + fv;
              q[o] = q[o] + pr;
            }
        } else {
          T a[2][2][2];
          for (int d = 0; d < 2; d++)
            for (int b = 0; b < 2; b++)
              for (int e = 0; e < 2; e++) a[d][b][e] = qc[cidx(P, ic + e, jc + b, kc + d) + cv];
          for (int dk = 0; dk < 2; dk++)
            for (int db = 0; db < 2; db++)
              for (int da = 0; da < 2; da++) {
                const int i = 2 * ic + da, j = 2 * jc + db, k = 2 * kc + dk;
                T cc2[2];
                for (int d = 0; d < 2; d++) {
                  const T b0 = linw<T>(a[d][0][0], a[d][0][1], i);
                  const T b1 = linw<T>(a[d][1][0], a[d][1][1], i);
                  cc2[d] = linw<T>(b0, b1, j);
                }
                const T pr = linw<T>(cc2[0], cc2[1], k) / (T)64;
                const int o = qidx(P, i, j, k) + fv;
                q[o] = q[o] + pr;
              }
        }
      }
    }
    __syncthreads();
    tc_fill<DIM, T>(pt, A, l, LIN ? 1 : 0, q);  // SetGhost(L, 0)
    tc_level_relax<DIM, LIN>(pt, A, m, l, nsm, q, f, qold, coef);
  }
}

// ---------------------------------------------------------------------------
// Layout conversion between the reference's column-major arrays (a staging
// buffer) and the colour-split pools.  kind 0: q-shaped (ghosts included),
// 1: interior-only (f, rf, ftmp), 2: qc-shaped.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_layout(const DPatch* __restrict__ pt, int pid, int kind,
                                                   int nv, int to_pool, double* __restrict__ stage,
                                                   double* __restrict__ pool) {
  const DPatch P = pt[pid];
  int e0, e1, e2;
  if (kind == 0) { e0 = P.nq[0]; e1 = P.nq[1]; e2 = P.nq[2]; }
  else if (kind == 1) { e0 = P.mx[0]; e1 = P.mx[1]; e2 = P.mx[2]; }
  else { e0 = P.nc[0]; e1 = P.nc[1]; e2 = P.nc[2]; }
  const long n = (long)e0 * e1 * e2;
  const long g = (long)blockIdx.x * kBlock + threadIdx.x;
  if (g >= n * nv) return;
  const int v = (int)(g / n);
  const long c = g % n;
  int i = (int)(c % e0);
  const long r = c / e0;
  int j = (int)(r % e1);
  int k = (int)(r / e1);
  int o;
  if (kind == 1) {
    i += 1; j += 1;
    if (P.nq[2] > 1) k += 1;
    o = qidx(P, i, j, k) + v * P.qvs;
  } else if (kind == 0) {
    o = qidx(P, i, j, k) + v * P.qvs;
  } else {
    o = cidx(P, i, j, k) + v * P.cvs;
  }
  if (to_pool) pool[o] = stage[g];
  else stage[g] = pool[o];
}

}  // namespace

// ============================================================================
// host side
// ============================================================================
namespace {

enum Dom { kHalf = 0, kCell, kCoarse, kFineG, kQcBox, kDual, kNumDom };

struct Stage {
  BoxDesc* d = nullptr;
  int n = 0;
  int maxel = 0;  // cells of the largest box
};

// A flattened ghost fill on the device: copy rules (dst, src, dst variable
// stride, src variable stride) with optional post-operation chains, then
// interpolation rules.
struct DevMap {
  int4* cr = nullptr;
  uint16_t* cpost = nullptr;
  GRule* ir = nullptr;
  int nc = 0, ni = 0;
  int nc1 = 0, ni1 = 0;  // first launch: rules reading ghost values from before the call
};

struct Level {
  int level = 0;
  std::vector<int> patches;
  int* d_patches = nullptr;
  int* d_prefix[kNumDom] = {};
  int* d_bp[kNumDom] = {};
  std::vector<int> hprefix[kNumDom];  // block-aligned starts of each patch's range
  std::vector<int> hcnt[kNumDom];     // elements of each patch
  int* d_cnt[kNumDom] = {};
  int total[kNumDom] = {};
  std::vector<Stage> stages;      // box stages after interpolation
  // coarse-fine interpolation rules (levels > 0), grouped by colour-filter
  // segment s (0: ipass 0 and 1; 1: every ipass; 2: ipass 0 and 2) and,
  // within a segment, by patch: rules of segment s, patch p are
  // [roff[s][p], roff[s][p+1])
  GRule* d_rules = nullptr;
  std::vector<int> roff[3];
  // the same ghost fill flattened to one rule per ghost cell (per-level
  // launches; not for a single patch periodic in every dimension at or
  // below the root, which keeps its wrap kernel, nor where the symbolic
  // run of the stages refuses): gmap[ipass] is SetGhost(level, ipass) in full,
  // gmap[2 + ipass] (ipass 1, 2) and gmap[5] (ipass 0) the ghosts a red-black sweep
  // sequence needs, gmap[6] those of gmap[5] a correction does not leave
  // exact (see plan_ghost_maps)
  DevMap gmap[7];
  // the reduced fills of ipass 1 and 2 (gmap[3], gmap[4]) split between
  // the colour sweep before them, which writes the copies of the cells it
  // updates (img_off / img_ent, ImgTab), and the gather of the rest (gimg)
  bool img_ok[2] = {false, false};
  int img_n[2] = {0, 0};
  const int* img_off[2] = {};
  const int4* img_ent[2] = {};
  DevMap gimg[2];
  bool maps_ok = false;
  // built without its ghost fill yet (build_level with defer_ghosts; the
  // caller finishes it with finish_level_ghosts)
  bool ghosts_deferred = false;
  std::string why;                // why a level keeps the staged fill
  int wrap_pid = -1;              // fully periodic single patch: fused wrap
  // wrap_pid's extents are all even: a colour sweep may write the ghost
  // images of its cells (wrap_images), since every image then has the
  // cell's colour and only the other colour reads it.  With an odd extent
  // the image of cell 1 is read by cell mx of the same colour in the same
  // sweep, so the sweep runs without images and the wrap follows it.
  bool wrap_img = false;
  int chain = 0;                 // >0: AFASVcycle from here down runs in one workgroup (chain length)
  // launch shapes chosen by bsam_autotune: block order (ordered_block) of
  // the level kernels and of the ghost gather, planes per workgroup of the
  // marching sweeps (FP64, FP32) and residuals (0: the default rule)
  int order = 0, gorder[2] = {0, 0};  // gather: FP64, FP32
  int gblock[2] = {0, 0};             // gather workgroup size (0: 256): FP64, FP32
  int kc_sweep[2] = {0, 0}, kc_resid = 0;
  double h = 0.0;
  long qspan_off = 0, qspan_len = 0;  // contiguous q-pool span of the level
  long ncell = 0;                     // interior cells of the level
  long* d_foff = nullptr;             // start of each patch's cells in a dense level array (tagging)
  std::vector<void*> blocks;          // pooled device blocks holding every array above
};

struct KStat {
  bool on = false;
  std::vector<hipEvent_t> ev;  // pairs
  int64_t launches = 0;
  double ms = 0.0;
  double bytes = 0.0;
};

}  // namespace

struct bsam_ctx {
  bsam_params p{};
  hipStream_t s = nullptr;
  int dev = 0;
  // hierarchy
  int npatch = 0, minlevel = 0, finest = 0;
  bool have_hier = false;
  std::vector<bsam_patch_desc> desc;
  std::vector<DPatch> hp;
  DPatch* d_pt = nullptr;
  double *q = nullptr, *qold = nullptr, *f = nullptr, *rf = nullptr, *ftmp = nullptr, *qc = nullptr;
  long nq = 0, nc = 0;
  std::vector<Level> lv;  // index level - minlevel
  // reductions
  double* d_partial = nullptr;
  double* d_partial2 = nullptr;
  int partial_cap = 0;
  double* d_norm = nullptr;
  double* h_norm = nullptr;
  // the tumour model's coarse chain (k_coarse_vcycle_tm), as plan_chain
  // sets it, and its device copy
  TmChainArgs h_tmchain{};
  TmChainArgs* d_tmchain = nullptr;
  // staging
  double* d_stage = nullptr;
  long stage_cap = 0;
  // regrid tagging: per-patch flag offsets, then the level's flags
  char* d_tag = nullptr;
  size_t tag_cap = 0;
  // AMR's tag buffering on the device: bit map of a level's index space
  // (capacity in words) and the level patches' global lower corners
  // workspace of the device planner of the flattened ghost tables
  char* dplan = nullptr;
  size_t dplan_cap = 0;
  int table_planner = 1;  // levels above the root: 1 device planner, 0 host planner
  uint32_t* d_tagbits = nullptr;
  size_t tagbits_cap = 0;
  int* d_tagglo = nullptr;
  size_t tagglo_cap = 0;
  std::vector<bsam_patch_desc> amr_out;  // the new patches of a level (bsam_amr)
  int32_t* h_flags = nullptr;             // pinned host copy of a level's tags (bsam_amr)
  size_t h_flags_cap = 0;
  // this device's copy of the file-scope `scratch` array (its device
  // address, looked up once in bsam_create with the device current)
  char* scratch = nullptr;
  // graph of the V-cycle loop body
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  bool capturing = false;
  // the installed mesh was built by bsam_amr: the loop body runs eagerly
  bool amr_mesh = false;
  std::string tune_report;  // JSON of the last bsam_autotune
  long chain_cells = 0;     // largest level (cells) where the coarse chain starts
  int regrid_tables = 1;  // flattened ghost tables for the levels a regrid builds
  int sweep_images = 1;   // BSAM_KNOB_SWEEP_IMAGES
  // field pools an install released, kept for the next install (a regrid
  // installs every step): buffer and capacity in bytes; pool_cap holds the
  // capacity of the pools in use
  std::vector<std::pair<void*, size_t>> spare, spare_tab;  // field pools, level-table blocks
  std::unordered_map<void*, std::pair<size_t, int>> pool_cap;  // capacity, reserve (0 fields, 1 tables)
  // mixed precision: an outer iteration whose residual exceeds mixed_stall
  // times the previous one ends the step in FP64 V-cycles
  double mixed_stall = 0.9;
  int64_t mixed_fallbacks = 0;  // MultigridIterations calls that fell back
  // instrumentation
  KStat ks[BSAM_KSTAT_COUNT];
  bool phase_on = false;
  int64_t phase_calls[BSAM_PHASE_COUNT] = {};
  double phase_ms[BSAM_PHASE_COUNT] = {};
  // mixed precision state
  bsam::MixedState* mixed = nullptr;

  Level& L(int level) { return lv[level - minlevel]; }
  int nv() const { return p.nrvars; }
  ModelK mk() const {
    ModelK m;
    m.dt = p.dt;
    m.eps2 = p.model_params[0];
    m.gam = p.model_params[1];
    m.lam = p.model_params[2];
    m.cz = p.model_params[3];
    return m;
  }
};

namespace {

inline int nblocks(long n) { return (int)((n + kBlock - 1) / kBlock); }

LevelArgs largs(bsam_ctx* c, Level& L, Dom d) {
  LevelArgs a;
  a.pt = c->d_pt;
  a.patches = L.d_patches;
  a.prefix = L.d_prefix[d];
  a.cnt = L.d_cnt[d];
  a.bp = L.d_bp[d];
  a.npatch = (int)L.patches.size();
  a.total = L.total[d];
  a.gbase = 0;
  a.fixed = -1;
  a.order = L.order;
  return a;
}

// The launches of one per-level operation: the whole level in one launch
// (batching = 1), or one launch per patch (batching = 0, the reference's
// per-patch ApplyOnLevel granularity).  `boff` is the number of blocks
// launched before the current one, for kernels that write per-block
// partial results.
struct LaunchIter {
  bsam_ctx* c;
  Level& L;
  Dom d;
  size_t i = 0;
  int boff = 0, nb_prev = 0;
  bool first = true;
  LaunchIter(bsam_ctx* c_, Level& L_, Dom d_) : c(c_), L(L_), d(d_) {}
  bool next(LevelArgs& a, int& nb) {
    boff += nb_prev;
    nb_prev = 0;
    const LevelArgs full = largs(c, L, d);
    if (c->p.batching || L.patches.size() <= 1) {
      if (!first) return false;
      first = false;
      a = full;
      nb = nblocks(full.total);
      nb_prev = nb;
      return nb > 0;
    }
    while (i < L.patches.size()) {
      a = full;
      a.gbase = L.hprefix[d][i];
      a.total = L.hprefix[d][i] + L.hcnt[d][i];
      a.fixed = (int)i;
      i++;
      nb = nblocks((long)a.total - a.gbase);
      if (nb > 0) {
        nb_prev = nb;
        return true;
      }
    }
    return false;
  }
  int blocks() const { return boff + nb_prev; }
};
#define FOR_LAUNCHES(c, L, dom, a, nb) \
  LevelArgs a;                          \
  int nb = 0;                           \
  for (LaunchIter it_(c, L, dom); it_.next(a, nb);)

int32_t check_level(bsam_ctx* c, int level) {
  if (!c->have_hier) return set_err(BSAM_ESTATE, "no hierarchy installed (bsam_set_hierarchy)");
  if (level < c->minlevel || level > c->finest)
    return set_err(BSAM_EINVAL, "level %d outside [%d, %d]", level, c->minlevel, c->finest);
  return BSAM_OK;
}

#define LAUNCH_CHECK()                                                                   \
  do {                                                                                   \
    hipError_t e_ = hipGetLastError();                                                   \
    if (e_ != hipSuccess) return set_err(BSAM_EHIP, "kernel launch: %s", hipGetErrorString(e_)); \
  } while (0)

void kstat_begin(bsam_ctx* c, int which) {
  KStat& k = c->ks[which];
  if (!k.on || c->capturing) return;
  hipEvent_t a, b;
  hipEventCreate(&a);
  hipEventCreate(&b);
  hipEventRecord(a, c->s);
  k.ev.push_back(a);
  k.ev.push_back(b);
}
void kstat_end(bsam_ctx* c, int which, double bytes) {
  KStat& k = c->ks[which];
  if (!k.on || c->capturing) return;
  hipEventRecord(k.ev.back(), c->s);
  k.launches++;
  k.bytes = bytes;
}

// Wall clock of one phase of a time step (bsam_phase_enable): the stream is
// drained on entry and exit, so the phase holds the GPU work it queued.
struct PhaseClock {
  bsam_ctx* c;
  int which;
  std::chrono::steady_clock::time_point t0;
  PhaseClock(bsam_ctx* c_, int w) : c(c_), which(w) { start(); }
  ~PhaseClock() { stop(); }
  void start() {
    if (which < 0 || !c->phase_on || c->capturing) return;
    hipStreamSynchronize(c->s);
    t0 = std::chrono::steady_clock::now();
  }
  void stop() {
    if (which < 0 || !c->phase_on || c->capturing) return;
    hipStreamSynchronize(c->s);
    c->phase_ms[which] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    c->phase_calls[which]++;
    which = -1;
  }
  // end this phase and start phase w
  void next(int w) {
    stop();
    which = w;
    start();
  }
};

}  // namespace

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
// With BSAM_BACKTRACE set in the environment, a segmentation fault or bus
// error of the process first writes the native call stack to stderr, then
// goes to the handler that was there before (the default one ends the
// process as before).
namespace {
struct sigaction g_prev_segv, g_prev_bus;
void crash_backtrace(int sig, siginfo_t* info, void*) {
  static const char head[] = "bsam: fatal signal, native stack:\n";
  ssize_t w = write(2, head, sizeof head - 1);
  (void)w;
  char buf[64];
  const int n = snprintf(buf, sizeof buf, "fault address %p\n", info ? info->si_addr : nullptr);
  if (n > 0) w = write(2, buf, (size_t)n);
  void* frames[64];
  backtrace_symbols_fd(frames, backtrace(frames, 64), 2);
  sigaction(sig, sig == SIGSEGV ? &g_prev_segv : &g_prev_bus, nullptr);
}
void install_backtrace() {
  static bool done = false;
  if (done || !getenv("BSAM_BACKTRACE")) return;
  done = true;
  struct sigaction sa {};
  sa.sa_sigaction = crash_backtrace;
  sa.sa_flags = SA_SIGINFO;
  sigemptyset(&sa.sa_mask);
  sigaction(SIGSEGV, &sa, &g_prev_segv);
  sigaction(SIGBUS, &sa, &g_prev_bus);
}
}  // namespace

extern "C" int32_t bsam_create(const bsam_params* params, bsam_ctx** out) {
  if (!params || !out) return set_err(BSAM_EINVAL, "null argument");
  install_backtrace();
  *out = nullptr;
  const bsam_params& p = *params;
  if (p.ndims != 2 && p.ndims != 3) return set_err(BSAM_EINVAL, "ndims must be 2 or 3");
  if (p.model != BSAM_MODEL_CH && p.model != BSAM_MODEL_BG && p.model != BSAM_MODEL_TUMOUR)
    return set_err(BSAM_EINVAL, "unknown model %d", p.model);
  if (p.nrvars != (p.model == BSAM_MODEL_TUMOUR ? 4 : 2))
    return set_err(BSAM_EINVAL, "nrvars must be 2 for CH and BG, 4 for the tumour model");
  if (p.mbc != 1) return set_err(BSAM_EINVAL, "only mbc=1 is supported");
  if (p.maux != 0) return set_err(BSAM_EINVAL, "maux must be 0");
  if (p.model == BSAM_MODEL_BG && p.ndims != 2)
    return set_err(BSAM_EINVAL, "the BioGrowth hooks exist in 2D only (their 3D forms are stubs)");
  if (p.errortype != 1 && p.errortype != 2)
    return set_err(BSAM_EINVAL, "ErrorAFAS: only errortype = 1,2 are supported");
  if (p.nsmoothingpasses < 1 || p.maxvcycles < 1) return set_err(BSAM_EINVAL, "bad V-cycle knobs");
  if (p.precision != BSAM_PREC_FP64 && p.precision != BSAM_PREC_MIXED)
    return set_err(BSAM_EINVAL, "unknown precision mode %d", p.precision);
  if (p.precision == BSAM_PREC_MIXED && p.model != BSAM_MODEL_CH && p.model != BSAM_MODEL_TUMOUR)
    return set_err(BSAM_EINVAL, "mixed precision is implemented for the CH and tumour models only");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return set_err(BSAM_ENODEV, "no HIP device visible");
  bsam_ctx* c = new bsam_ctx();
  c->p = p;
  c->chain_cells = (p.model == BSAM_MODEL_TUMOUR) ? kTmChainCells : kChChainCells;
  hipGetDevice(&c->dev);
  if (hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return set_err(BSAM_EHIP, "hipStreamCreate failed");
  }
  {
    void* sp = nullptr;
    const hipError_t e = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
      bsam_destroy(c);
      return set_err(BSAM_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    }
    c->scratch = static_cast<char*>(sp);
  }
  if (hipMalloc(&c->d_norm, 4 * sizeof(double)) != hipSuccess ||
      hipHostMalloc(&c->h_norm, 4 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
    bsam_destroy(c);
    return set_err(BSAM_ENOMEM, "allocation of reduction buffers failed");
  }
  *out = c;
  return BSAM_OK;
}

// Destroys the captured V-cycle graph; the next call captures it again
// (after a change of launch shapes, dt or hierarchy).
static void drop_graph(bsam_ctx* c) {
  if (c->gexec) {
    hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
  }
  if (c->graph) {
    hipGraphDestroy(c->graph);
    c->graph = nullptr;
  }
}

// Field pools and level-table blocks: a released buffer is kept in its
// reserve and handed to the next request of that reserve it can hold
// without wasting more than half of itself.  Every release and reuse
// follows a synchronisation of the context's stream (installs and
// destruction synchronise first), so no work in flight reads a reused
// buffer.
static hipError_t pool_alloc(bsam_ctx* c, void** p, size_t bytes, int kind = 0) {
  auto& spare = kind ? c->spare_tab : c->spare;
  // an eighth of headroom, in whole 2 MiB pages, so that a mesh that grows
  // a little from one regrid to the next finds its buffers in the reserve
  const size_t page = size_t(2) << 20;
  const size_t want = ((std::max<size_t>(bytes, 1) + bytes / 8) + page - 1) & ~(page - 1);
  int best = -1;
  for (size_t i = 0; i < spare.size(); i++) {
    const size_t cap = spare[i].second;
    if (cap >= bytes && cap <= 2 * want && (best < 0 || cap < spare[best].second)) best = (int)i;
  }
  if (best >= 0) {
    *p = spare[best].first;
    c->pool_cap[*p] = {spare[best].second, kind};
    spare.erase(spare.begin() + best);
    return hipSuccess;
  }
  const hipError_t e = hipMalloc(p, want);
  if (e == hipSuccess) c->pool_cap[*p] = {want, kind};
  return e;
}
template <typename T>
static hipError_t pool_alloc(bsam_ctx* c, T** p, size_t bytes, int kind = 0) {
  void* v = nullptr;
  const hipError_t e = pool_alloc(c, &v, bytes, kind);
  *p = static_cast<T*>(v);
  return e;
}
static void pool_release(bsam_ctx* c, void* p) {
  if (!p) return;
  auto it = c->pool_cap.find(p);
  if (it == c->pool_cap.end()) {
    hipFree(p);
    return;
  }
  const int kind = it->second.second;
  auto& spare = kind ? c->spare_tab : c->spare;
  spare.push_back({p, it->second.first});
  c->pool_cap.erase(it);
  // a bounded reserve (a regrid pass installs its levels one at a time,
  // each install with its own set of pools, and keeps the old mesh until
  // it ends): at most 64 buffers and the larger of 4 GiB and twice the
  // reserve's bytes in use; the oldest spare buffers go back to the device
  size_t inuse = 0, held = 0;
  for (const auto& kv : c->pool_cap)
    if (kv.second.second == kind) inuse += kv.second.first;
  for (const auto& sp : spare) held += sp.second;
  const size_t limit = std::max<size_t>(size_t(4) << 30, 2 * inuse);
  while (!spare.empty() && (spare.size() > 64 || held > limit)) {
    held -= spare.front().second;
    hipFree(spare.front().first);
    spare.erase(spare.begin());
  }
}
static void pool_drain(bsam_ctx* c) {
  for (auto& sp : c->spare) hipFree(sp.first);
  for (auto& sp : c->spare_tab) hipFree(sp.first);
  c->spare.clear();
  c->spare_tab.clear();
}

// The device arrays of one level (launch lists, ghost-fill stages, rules
// and tables) live in pooled blocks the level owns; they go back to the
// reserve and the level is left empty.
static void free_level(bsam_ctx* c, Level& L) {
  for (void* b : L.blocks) pool_release(c, b);
  L = Level();
}

static void free_hierarchy(bsam_ctx* c) {
  if (c->gexec) { hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
  if (c->graph) { hipGraphDestroy(c->graph); c->graph = nullptr; }
  for (Level& L : c->lv) free_level(c, L);
  c->lv.clear();
  pool_release(c, c->d_pt);
  for (double* pool : {c->q, c->qold, c->f, c->rf, c->ftmp, c->qc}) pool_release(c, pool);
  c->d_pt = nullptr; c->q = c->qold = c->f = c->rf = c->ftmp = c->qc = nullptr;
  if (c->mixed) {
    for (float* pool : {c->mixed->e, c->mixed->rhs, c->mixed->coef, c->mixed->ec}) pool_release(c, pool);
    c->mixed->e = c->mixed->rhs = c->mixed->coef = c->mixed->ec = nullptr;
    bsam::mixed_free(c->mixed);
    c->mixed = nullptr;
  }
  c->have_hier = false;
}

extern "C" int32_t bsam_destroy(bsam_ctx* c) {
  if (!c) return BSAM_OK;
  if (c->s) hipStreamSynchronize(c->s);
  free_hierarchy(c);
  pool_drain(c);
  hipFree(c->d_partial);
  hipFree(c->d_partial2);
  hipFree(c->d_norm);
  hipFree(c->d_stage);
  hipFree(c->d_tag);
  hipFree(c->d_tagbits);
  hipFree(c->d_tagglo);
  hipFree(c->dplan);
  hipFree(c->d_tmchain);
  if (c->h_norm) hipHostFree(c->h_norm);
  if (c->h_flags) hipHostFree(c->h_flags);
  for (auto& k : c->ks)
    for (hipEvent_t e : k.ev) hipEventDestroy(e);
  if (c->s) hipStreamDestroy(c->s);
  delete c;
  return BSAM_OK;
}

extern "C" int32_t bsam_set_dt(bsam_ctx* c, double dt) {
  if (!c) return set_err(BSAM_EINVAL, "null context");
  c->p.dt = dt;
  // the graph bakes dt into its kernel arguments
  if (c->gexec) { hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
  if (c->graph) { hipGraphDestroy(c->graph); c->graph = nullptr; }
  return BSAM_OK;
}

// ---------------------------------------------------------------------------
// hierarchy installation: pools, level tables, ghost-fill stages
// ---------------------------------------------------------------------------
namespace {

// GhostOverlap (boundary.f90:359-431) as two box copies; `off` shifts grid 2.
void ghost_overlap_boxes(const bsam_ctx* c, int g1, int g2, const int off[3], int op,
                         std::vector<BoxDesc>& out) {
  const int nd = c->p.ndims;
  const bsam_patch_desc& A = c->desc[g1];
  const bsam_patch_desc& B = c->desc[g2];
  int mg1[3][2], mg2[3][2], ovg[3][2];
  for (int n = 0; n < nd; n++)
    for (int s = 0; s < 2; s++) {
      mg1[n][s] = A.mglobal[n][s];
      mg2[n][s] = B.mglobal[n][s] + (off ? off[n] : 0);
    }
  for (int n = 0; n < nd; n++) {
    ovg[n][0] = std::max(mg1[n][0] - 1, mg2[n][0] - 1);
    ovg[n][1] = std::min(mg1[n][1] + 1, mg2[n][1] + 1);
    if (ovg[n][1] - ovg[n][0] < 0) return;
  }
  for (int dir = 0; dir < 2; dir++) {
    // dir 0: grid1 <- grid2 over ovg & mg2; dir 1: grid2 <- grid1 over ovg & mg1
    const int (*ms)[2] = dir == 0 ? mg2 : mg1;  // source interior box
    const int (*md)[2] = dir == 0 ? mg1 : mg2;
    BoxDesc b{};
    b.dp = dir == 0 ? g1 : g2;
    b.sp = dir == 0 ? g2 : g1;
    b.op = op;
    bool empty = false;
    for (int n = 0; n < 3; n++) {
      if (n < nd) {
        const int lo = std::max(ovg[n][0], ms[n][0]), hi = std::min(ovg[n][1], ms[n][1]);
        if (hi < lo) empty = true;
        b.dlo[n] = lo - md[n][0] + 1;
        b.slo[n] = lo - ms[n][0] + 1;
        b.ext[n] = hi - lo + 1;
      } else {
        b.dlo[n] = b.slo[n] = 0;
        b.ext[n] = 1;
      }
    }
    if (!empty) out.push_back(b);
  }
}

// PeriodicSetup + GetPeriodicOffsets of the installed root (bsam_grid.cpp).
void periodic_offsets(const bsam_ctx* c, int level, std::vector<std::array<int, 3>>& offs, bool& periodic) {
  const bsam_patch_desc* root = nullptr;
  for (auto& d : c->desc)
    if (d.level == 0) { root = &d; break; }
  bsam::periodic_offsets(c->p.ndims, *root, level, offs, periodic);
}

// Host plan of one level's SetGhost, built by plan_ghost_* and uploaded by
// upload_ghost: the staged form (interpolation rules, then box stages in
// the reference's order) and, where it is valid, the flattened maps.
struct HostMap {
  std::vector<int4> copies;    // dst, src, dst variable stride, src variable stride
  std::vector<uint16_t> cpost;  // post-operation chain per copy (empty: none anywhere)
  std::vector<GRule> interps;  // kind | post << 8
  int nc1 = 0, ni1 = 0;        // leading rules that read ghost values from before the call
};
struct GhostPlan {
  int wrap_pid = -1;
  std::vector<std::vector<BoxDesc>> stages;
  std::vector<GRule> rules;
  std::vector<int> roff[3];
  bool maps_ok = false;
  HostMap map[7];
  std::string why;  // why the flattened maps are not used
};

// The coarse-fine interpolation of one patch as gather rules, enumerating
// its ghost cells as InterpolateCoarseFine does: 2D faces then corners
// (boundary.f90:143-333), 3D faces (with the grey face-corner cells), edges
// and corners (:1048-1550).  seg receives the rules of the three
// colour-filter segments: a face cell refreshes only after the colour of
// its parity (segment 0 or 2); 3D edges and corners refresh after every
// colour except that ipass = 1 skips those touching a non-interpolated
// face (segment 2 then, else 1), including the comparison with 99 instead
// of 999 at boundary.f90:1104; grey cells refresh always (segment 1).
// Elements of a patch's coarse-fine interpolation (InterpolateCoarseFine,
// boundary.f90:100-140, Interpolate2D :143-333, Interpolate3D :1048-1550):
// the face cells, then (3D) the edge cells, then the corners of the ghost
// shell, in the order below.
__host__ __device__ inline int interp_elems(const DPatch& P, int ndims) {
  if (ndims == 2) return 2 * P.mx[1] + 2 * P.mx[0] + 4;
  const int mx = P.mx[0], my = P.mx[1], mz = P.mx[2];
  return 2 * (my * mz + mx * mz + mx * my) + 4 * (my + mz + mx) + 8;
}

// The rule of element e of patch P (parent Q) and its colour-filter
// segment; false when the element's face has no coarse-fine interpolation.
// Every coordinate array below is indexed by compile-time positions only
// (the face or edge direction is a template argument), so on the GPU the
// points stay in registers.
struct IrPts {
  int g[3], s1[3], s2[3], cq[4][3];
  int nco;
};
__host__ __device__ __forceinline__ bool ir_emit(const DPatch& P, const DPatch& Q, int ndims, const IrPts& t,
                                                 GRule& rule) {
  const bool d3 = ndims == 3;
  rule.dst = qidx(P, t.g[0], t.g[1], d3 ? t.g[2] : 0);
  rule.s1 = qidx(P, t.s1[0], t.s1[1], d3 ? t.s1[2] : 0);
  rule.s2 = qidx(P, t.s2[0], t.s2[1], d3 ? t.s2[2] : 0);
#pragma unroll
  for (int n = 0; n < 4; n++) rule.c[n] = (n < t.nco) ? qidx(Q, t.cq[n][0], t.cq[n][1], d3 ? t.cq[n][2] : 0) : 0;
  rule.kind = t.nco;
  rule.pvs = P.qvs;
  rule.cvs = Q.qvs;
  return true;
}
// the boundary code of face (dim, parity)
template <int D>
__host__ __device__ __forceinline__ int ir_mth(const DPatch& P, int parity) {
  return parity == 1 ? P.mthbc[2 * D] : P.mthbc[2 * D + 1];
}
// 2D face cell pos of face (D, p) (boundary.f90:143-333)
template <int D>
__host__ __device__ __forceinline__ bool ir_face2(const DPatch& P, const DPatch& Q, int p, int pos, GRule& rule,
                                                  int& sg) {
  constexpr int T = 1 - D;
  if (!bc_interp(ir_mth<D>(P, p))) return false;
  IrPts t;
  t.nco = 1;
  t.g[2] = t.s1[2] = t.s2[2] = 1;
#pragma unroll
  for (int n = 0; n < 4; n++) t.cq[n][2] = 1;
  int fi[5];
  face_index(P, D, p, fi);
  sg = ((fi[0] + pos) & 1) ? 2 : 0;
  const int k = pos_kind(pos, P.mx[T]);
  t.g[D] = fi[0]; t.s1[D] = fi[1]; t.s2[D] = fi[2];
  t.g[T] = pos; t.s1[T] = pos + kind_s1(k); t.s2[T] = pos + kind_s2(k);
  if (k == KE || k == KO) {
    t.cq[0][D] = fi[3]; t.cq[0][T] = kind_c(P, T, k, pos, 0);
  } else {
    t.nco = 3;
#pragma unroll
    for (int r = 0; r < 3; r++) { t.cq[r][D] = fi[3]; t.cq[r][T] = kind_c(P, T, k, pos, r); }
  }
  return ir_emit(P, Q, 2, t, rule);
}
// 2D corner cn (boundary.f90:155-181): applied whatever mthbc is
__host__ __device__ __forceinline__ bool ir_corner2(const DPatch& P, const DPatch& Q, int cn, GRule& rule, int& sg) {
  IrPts t;
  t.nco = 1;
  t.g[2] = t.s1[2] = t.s2[2] = 1;
#pragma unroll
  for (int n = 0; n < 4; n++) t.cq[n][2] = 1;
  int f0[5], f1[5];
  face_index(P, 0, 1 + (cn & 1), f0);
  face_index(P, 1, 1 + (cn >> 1), f1);
  sg = ((f0[0] + f1[0]) & 1) ? 2 : 0;
  t.g[0] = f0[0]; t.s1[0] = f0[1]; t.s2[0] = f0[2]; t.cq[0][0] = f0[3];
  t.g[1] = f1[0]; t.s1[1] = f1[1]; t.s2[1] = f1[2]; t.cq[0][1] = f1[3];
  return ir_emit(P, Q, 2, t, rule);
}
// 3D face cell `local` of face (D, p), with the grey face-corner cells
// (boundary.f90:1048-1550)
template <int D>
__host__ __device__ __forceinline__ bool ir_face3(const DPatch& P, const DPatch& Q, int p, int local, GRule& rule,
                                                  int& sg) {
  constexpr int T1 = (D == 0) ? 1 : 0, T2 = (D == 2) ? 1 : 2;
  if (!bc_interp(ir_mth<D>(P, p))) return false;
  IrPts t;
  t.nco = 1;
  const int a = local % P.mx[T1] + 1, b = local / P.mx[T1] + 1;
  const int k1 = pos_kind(a, P.mx[T1]), k2 = pos_kind(b, P.mx[T2]);
  const bool q1 = k1 >= KL, q2 = k2 >= KL;
  int fi[5];
  face_index(P, D, p, fi);
  t.g[D] = fi[0]; t.s1[D] = fi[1]; t.s2[D] = fi[2];
  if (q1 && q2) {
    // face-corner grey cells (boundary.f90:1320-1333 and the like)
    int ft1[5], ft2[5];
    face_index(P, T1, a == 1 ? 1 : 2, ft1);
    face_index(P, T2, b == 1 ? 1 : 2, ft2);
    t.g[T1] = t.s1[T1] = t.s2[T1] = ft1[2];
    t.g[T2] = t.s1[T2] = t.s2[T2] = ft2[2];
    t.nco = 4;
#pragma unroll
    for (int r = 0; r < 4; r++) t.cq[r][D] = fi[3];
    t.cq[0][T1] = ft1[4]; t.cq[0][T2] = ft2[4];
    t.cq[1][T1] = ft1[3]; t.cq[1][T2] = ft2[4];
    t.cq[2][T1] = ft1[4]; t.cq[2][T2] = ft2[3];
    t.cq[3][T1] = ft1[3]; t.cq[3][T2] = ft2[3];
    sg = 1;
  } else {
    sg = ((fi[0] + a + b) & 1) ? 2 : 0;
    t.g[T1] = a; t.s1[T1] = a + kind_s1(k1); t.s2[T1] = a + kind_s2(k1);
    t.g[T2] = b; t.s1[T2] = b + kind_s1(k2); t.s2[T2] = b + kind_s2(k2);
    if (!q1 && !q2) {
      t.cq[0][D] = fi[3]; t.cq[0][T1] = kind_c(P, T1, k1, a, 0); t.cq[0][T2] = kind_c(P, T2, k2, b, 0);
    } else {
      t.nco = 3;
#pragma unroll
      for (int r = 0; r < 3; r++) {
        t.cq[r][D] = fi[3];
        t.cq[r][T1] = kind_c(P, T1, k1, a, q1 ? r : 0);
        t.cq[r][T2] = kind_c(P, T2, k2, b, q2 ? r : 0);
      }
    }
  }
  return ir_emit(P, Q, 3, t, rule);
}
// 3D edge cell: edges along dim 1 (EE 0), dim 2 (EE 1), dim 0 (EE 2);
// `rem` counts from the first edge of that group
template <int EE>
__host__ __device__ __forceinline__ bool ir_edge3(const DPatch& P, const DPatch& Q, int rem, GRule& rule, int& sg) {
  constexpr int ED = (EE == 0) ? 1 : (EE == 1 ? 2 : 0);
  constexpr int FA = (ED == 0) ? 1 : 0, FB = (ED == 2) ? 1 : 2;
  const int ne = P.mx[ED];
  const int combo = rem / ne, pos = rem % ne + 1;
  const int pa = 1 + (combo >> 1), pb = 1 + (combo & 1);
  bool skip1 = !bc_interp(ir_mth<FA>(P, pa));
  const int cb = ir_mth<FB>(P, pb);
  if (EE == 0) skip1 = skip1 || (cb != 2 && cb != 99);  // boundary.f90:1104
  else skip1 = skip1 || !bc_interp(cb);
  sg = skip1 ? 2 : 1;
  IrPts t;
  t.nco = 1;
  int fa[5], fb[5];
  face_index(P, FA, pa, fa);
  face_index(P, FB, pb, fb);
  const int k = pos_kind(pos, P.mx[ED]);
  t.g[FA] = fa[0]; t.s1[FA] = fa[1]; t.s2[FA] = fa[2];
  t.g[FB] = fb[0]; t.s1[FB] = fb[1]; t.s2[FB] = fb[2];
  t.g[ED] = pos; t.s1[ED] = pos + kind_s1(k); t.s2[ED] = pos + kind_s2(k);
  if (k == KE || k == KO) {
    t.cq[0][FA] = fa[3]; t.cq[0][FB] = fb[3]; t.cq[0][ED] = kind_c(P, ED, k, pos, 0);
  } else {
    t.nco = 3;
#pragma unroll
    for (int r = 0; r < 3; r++) {
      t.cq[r][FA] = fa[3]; t.cq[r][FB] = fb[3]; t.cq[r][ED] = kind_c(P, ED, k, pos, r);
    }
  }
  return ir_emit(P, Q, 3, t, rule);
}
// 3D corner cn of the ghost shell
__host__ __device__ __forceinline__ bool ir_corner3(const DPatch& P, const DPatch& Q, int cn, GRule& rule, int& sg) {
  const int p1 = 1 + (cn & 1), p2 = 1 + ((cn >> 1) & 1), p3 = 1 + (cn >> 2);
  const bool skip1 = !bc_interp(ir_mth<0>(P, p1)) || !bc_interp(ir_mth<1>(P, p2)) || !bc_interp(ir_mth<2>(P, p3));
  sg = skip1 ? 2 : 1;
  IrPts t;
  t.nco = 1;
  int f0[5], f1[5], f2[5];
  face_index(P, 0, p1, f0);
  face_index(P, 1, p2, f1);
  face_index(P, 2, p3, f2);
  t.g[0] = f0[0]; t.s1[0] = f0[1]; t.s2[0] = f0[2]; t.cq[0][0] = f0[3];
  t.g[1] = f1[0]; t.s1[1] = f1[1]; t.s2[1] = f1[2]; t.cq[0][1] = f1[3];
  t.g[2] = f2[0]; t.s1[2] = f2[1]; t.s2[2] = f2[2]; t.cq[0][2] = f2[3];
  return ir_emit(P, Q, 3, t, rule);
}
__host__ __device__ inline bool interp_rule_at(const DPatch& P, const DPatch& Q, int ndims, int e, GRule& rule,
                                               int& sg) {
  if (ndims == 2) {
    const int nf0 = P.mx[1], nf1 = P.mx[0];
    if (e < 2 * nf0) return ir_face2<0>(P, Q, 1 + e / nf0, e % nf0 + 1, rule, sg);
    if (e < 2 * nf0 + 2 * nf1) {
      const int e2 = e - 2 * nf0;
      return ir_face2<1>(P, Q, 1 + e2 / nf1, e2 % nf1 + 1, rule, sg);
    }
    return ir_corner2(P, Q, e - (2 * nf0 + 2 * nf1), rule, sg);
  }
  const int mx = P.mx[0], my = P.mx[1], mz = P.mx[2];
  const int nf0 = my * mz, nf1 = mx * mz, nf2 = mx * my;
  int rem = e;
  if (rem < 2 * nf0) return ir_face3<0>(P, Q, 1 + rem / nf0, rem % nf0, rule, sg);
  rem -= 2 * nf0;
  if (rem < 2 * nf1) return ir_face3<1>(P, Q, 1 + rem / nf1, rem % nf1, rule, sg);
  rem -= 2 * nf1;
  if (rem < 2 * nf2) return ir_face3<2>(P, Q, 1 + rem / nf2, rem % nf2, rule, sg);
  rem -= 2 * nf2;
  // edges along dim 1 (4 my cells), dim 2 (4 mz), dim 0 (4 mx)
  if (rem < 4 * my) return ir_edge3<0>(P, Q, rem, rule, sg);
  rem -= 4 * my;
  if (rem < 4 * mz) return ir_edge3<1>(P, Q, rem, rule, sg);
  rem -= 4 * mz;
  if (rem < 4 * mx) return ir_edge3<2>(P, Q, rem, rule, sg);
  rem -= 4 * mx;
  return ir_corner3(P, Q, rem, rule, sg);
}

static void interp_rules(const DPatch& P, const DPatch& Q, int ndims, std::vector<GRule> (&seg)[3]) {
  const int total = interp_elems(P, ndims);
  for (int e = 0; e < total; e++) {
    GRule r;
    int sg = 0;
    if (interp_rule_at(P, Q, ndims, e, r, sg)) seg[sg].push_back(r);
  }
}

// The interpolation rules of a level built on the device, for levels that
// keep the staged fill: the elements of the level's patches (interp_elems
// each, patch after patch, from eoff) are classified by colour-filter
// segment, compacted in order segment by segment (hipCUB DeviceSelect), and
// written, so the table equals plan_interp_rules' (segment, then patch,
// then element order).
__device__ __forceinline__ int elem_patch(const int* __restrict__ eoff, int np, long g) {
  int lo = 0, hi = np;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (eoff[mid] <= g) lo = mid; else hi = mid;
  }
  return lo;
}
__global__ __launch_bounds__(kBlock) void k_interp_segment(const DPatch* __restrict__ pt, const int* __restrict__ patches,
                                                           const int* __restrict__ eoff, int np, int ndims, long total,
                                                           uint8_t* __restrict__ seg) {
  const long g = (long)blockIdx.x * kBlock + threadIdx.x;
  if (g >= total) return;
  const int a = elem_patch(eoff, np, g);
  const DPatch P = pt[patches[a]];
  GRule r;
  int sg = 0;
  seg[g] = interp_rule_at(P, pt[P.parent], ndims, (int)(g - eoff[a]), r, sg) ? (uint8_t)sg : (uint8_t)3;
}
__global__ __launch_bounds__(kBlock) void k_interp_write(const DPatch* __restrict__ pt, const int* __restrict__ patches,
                                                         const int* __restrict__ eoff, int np, int ndims,
                                                         const int* __restrict__ idx, int n, GRule* __restrict__ out) {
  const int j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= n) return;
  const int g = idx[j];
  const int a = elem_patch(eoff, np, g);
  const DPatch P = pt[patches[a]];
  GRule r;
  int sg = 0;
  interp_rule_at(P, pt[P.parent], ndims, g - eoff[a], r, sg);
  out[j] = r;
}
struct SegmentIs {
  const uint8_t* seg;
  uint8_t want;
  __host__ __device__ bool operator()(const int& e) const { return seg[e] == want; }
};

// Host threads for the hierarchy set-up (at most 16).
static size_t host_threads() {
  const unsigned n = std::thread::hardware_concurrency();
  return n ? std::min(16u, n) : 4u;
}

// fn(t) for t = 0 .. n-1 on host threads; the calls a thread could not be
// started for run on the calling thread.
template <typename F>
static void run_threads(int n, F&& fn) {
  const std::function<void(int)> f = [&fn](int t) { fn(t); };
  bsam::parallel_for(n, f);
}

static void plan_interp_rules(const bsam_ctx* c, const Level& L, GhostPlan& G) {
  const size_t np = L.patches.size();
  for (int sg = 0; sg < 3; sg++) G.roff[sg].assign(np + 1, 0);
  if (L.level <= 0) return;
  std::vector<GRule>& all = G.rules;
  std::vector<std::vector<GRule>> per[3];
  for (int sg = 0; sg < 3; sg++) per[sg].resize(np);
  // patches are independent: interleaved over the host's threads
  const int nth = (int)std::max<size_t>(1, std::min<size_t>(np, host_threads()));
  run_threads(nth, [&](int t) {
    for (size_t a = t; a < np; a += nth) {
      const DPatch& P = c->hp[L.patches[a]];
      std::vector<GRule> seg[3];
      interp_rules(P, c->hp[P.parent], c->p.ndims, seg);
      for (int sg = 0; sg < 3; sg++) per[sg][a].swap(seg[sg]);
    }
  });
  // one table, segment by segment and patch by patch, copied in parallel
  size_t total = all.size();
  for (int sg = 0; sg < 3; sg++) {
    for (size_t a = 0; a < np; a++) {
      G.roff[sg][a] = (int)total;
      total += per[sg][a].size();
    }
    G.roff[sg][np] = (int)total;
  }
  all.resize(total);
  run_threads(nth, [&](int t) {
    for (size_t a = t; a < np; a += nth)
      for (int sg = 0; sg < 3; sg++) std::copy(per[sg][a].begin(), per[sg][a].end(), all.begin() + G.roff[sg][a]);
  });
}

// For each patch a of a level, the patches b > a whose ghost-extended box
// can meet a's, directly or through one of the periodic offsets, in
// ascending order: a uniform grid of bins over the level's index space,
// each listing the patches whose box it touches, replaces the test of
// every pair (ApplyOnLevelPairs, treeops.f90:555-590), so the work grows
// with the number of neighbours rather than with the square of the patch
// count.  Pairs whose boxes do not meet produce no transfer boxes, so the
// boxes come out as from the all-pairs loop, in the same order.
// (Each candidate b > a of patch a comes with the shifts under which the
// two ghost-extended boxes may meet: bit 0 for none, bit 1 + q for shift q,
// q = 2 * offset + (1 for +offset, 0 for -offset).)
using Cand = std::pair<int, uint32_t>;
static std::vector<std::vector<Cand>> neighbour_candidates(const bsam_ctx* c, const std::vector<int>& ps,
                                                           const std::vector<std::array<int, 3>>& offs) {
  const int nd = c->p.ndims;
  const size_t np = ps.size();
  std::vector<std::vector<Cand>> cand(np);
  if (np < 2) return cand;
  // bins of the largest ghost-extended patch extent, over the bounding box
  int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, ext = 1;
  for (size_t a = 0; a < np; a++) {
    const bsam_patch_desc& d = c->desc[ps[a]];
    for (int n = 0; n < nd; n++) {
      lo[n] = a ? std::min(lo[n], d.mglobal[n][0] - 1) : d.mglobal[n][0] - 1;
      hi[n] = a ? std::max(hi[n], d.mglobal[n][1] + 1) : d.mglobal[n][1] + 1;
      ext = std::max(ext, d.mglobal[n][1] - d.mglobal[n][0] + 3);
    }
  }
  int nb[3] = {1, 1, 1};
  for (int n = 0; n < nd; n++) nb[n] = (hi[n] - lo[n]) / ext + 1;
  const long nbins = (long)nb[0] * nb[1] * nb[2];
  auto bin_range = [&](const int (*box)[2], int* b0, int* b1) {
    for (int n = 0; n < 3; n++) {
      if (n >= nd) { b0[n] = b1[n] = 0; continue; }
      b0[n] = std::max(0, (box[n][0] - lo[n]) / ext);
      b1[n] = std::min(nb[n] - 1, (box[n][1] - lo[n]) / ext);
      if (box[n][1] < lo[n] || box[n][0] > hi[n]) return false;
    }
    return true;
  };
  // the bins' patch lists in one array (counted, then filled in patch
  // order, so each list is ascending)
  std::vector<int> bstart(nbins + 1, 0), bins, fill;
  for (int pass = 0; pass < 2; pass++) {
    if (pass == 1) {
      for (long b = 0; b < nbins; b++) bstart[b + 1] += bstart[b];
      bins.resize(bstart[nbins]);
      fill.assign(bstart.begin(), bstart.end() - 1);
    }
    for (size_t a = 0; a < np; a++) {
      const bsam_patch_desc& d = c->desc[ps[a]];
      int box[3][2], b0[3], b1[3];
      for (int n = 0; n < 3; n++) {
        box[n][0] = n < nd ? d.mglobal[n][0] - 1 : 0;
        box[n][1] = n < nd ? d.mglobal[n][1] + 1 : 0;
      }
      bin_range(box, b0, b1);
      for (int z = b0[2]; z <= b1[2]; z++)
        for (int y = b0[1]; y <= b1[1]; y++)
          for (int x = b0[0]; x <= b1[0]; x++) {
            const long b = ((long)z * nb[1] + y) * nb[0] + x;
            if (pass == 0) bstart[b + 1]++;
            else bins[fill[b]++] = (int)a;
          }
    }
  }
  // patches on host threads, each thread with its own marks
  const int nth = (int)std::max<size_t>(1, std::min<size_t>(np / 64 + 1, host_threads()));
  run_threads(nth, [&](int t) {
    std::vector<int> seen(np, -1);
    std::vector<uint32_t> mask(np, 0);
    std::vector<int> found;
    for (size_t a = t; a < np; a += nth) {
      const bsam_patch_desc& d = c->desc[ps[a]];
      found.clear();
      // a's box shifted by -off meets b's box exactly when b's box shifted
      // by +off meets a's: both polarities of every offset, and no offset
      for (int q = -1; q < 2 * (int)offs.size(); q++) {
        int box[3][2], b0[3], b1[3];
        for (int n = 0; n < 3; n++) {
          const int sh = q < 0 ? 0 : (q & 1 ? 1 : -1) * offs[q >> 1][n];
          box[n][0] = n < nd ? d.mglobal[n][0] - 1 - sh : 0;
          box[n][1] = n < nd ? d.mglobal[n][1] + 1 - sh : 0;
        }
        if (!bin_range(box, b0, b1)) continue;
        for (int z = b0[2]; z <= b1[2]; z++)
          for (int y = b0[1]; y <= b1[1]; y++)
            for (int x = b0[0]; x <= b1[0]; x++)
              for (int e = bstart[((long)z * nb[1] + y) * nb[0] + x], e1 = bstart[((long)z * nb[1] + y) * nb[0] + x + 1];
                   e < e1; e++) {
                const int b = bins[e];
                if (b <= (int)a) continue;
                if (seen[b] != (int)a) {
                  seen[b] = (int)a;
                  mask[b] = 0;
                  found.push_back(b);
                }
                mask[b] |= 1u << (q + 1);
              }
      }
      std::sort(found.begin(), found.end());
      cand[a].reserve(found.size());
      for (int b : found) cand[a].push_back({b, mask[b]});
    }
  });
  return cand;
}

// Operation of a physical face (condition 10, SetPhysicalBC + UserBC):
// 3D CH 2 - q_in, 2D CH a copy; the tumour model copies variables 0 and 1,
// sets 2 - q_in for variable 2 and -q_in for variable 3; the BG model
// copies variable 0 and sets 2 - q_in for the others.
static int phys_face_op(const bsam_ctx* c) {
  if (c->p.model == BSAM_MODEL_CH) return c->p.ndims == 3 ? kDirichlet2 : kCopy;
  if (c->p.model == BSAM_MODEL_TUMOUR) return kTumourBC;
  return kBgMixed;
}

int32_t plan_ghost_stages(const bsam_ctx* c, const Level& L, GhostPlan& G) {
  const int nd = c->p.ndims;
  std::vector<std::array<int, 3>> offs;
  bool periodic = false;
  periodic_offsets(c, L.level, offs, periodic);
  std::vector<BoxDesc> sib, px, py, pz, pp, phys[3];
  const auto& ps = L.patches;
  const std::vector<std::vector<Cand>> cand =
      L.level > 0 ? neighbour_candidates(c, ps, periodic ? offs : std::vector<std::array<int, 3>>{})
                  : std::vector<std::vector<Cand>>();
  // the pairs of each patch on the host's threads, joined in patch order
  // (each thread appends to one buffer of its own and records every
  // patch's range in it, so the boxes are allocated per thread, not per
  // patch)
  auto per_patch = [&](auto&& boxes_of, std::vector<BoxDesc>& out) {
    const int nth = (int)std::max<size_t>(1, std::min<size_t>(ps.size() / 64 + 1, host_threads()));
    std::vector<std::vector<BoxDesc>> buf(nth);
    std::vector<std::pair<size_t, size_t>> span(ps.size());
    run_threads(nth, [&](int t) {
      std::vector<BoxDesc>& v = buf[t];
      for (size_t a = t; a < ps.size(); a += nth) {
        const size_t s0 = v.size();
        boxes_of(a, v);
        span[a] = {s0, v.size()};
      }
    });
    size_t total = out.size();
    for (const auto& sp : span) total += sp.second - sp.first;
    out.reserve(total);
    for (size_t a = 0; a < ps.size(); a++) {
      const std::vector<BoxDesc>& v = buf[a % nth];
      out.insert(out.end(), v.begin() + span[a].first, v.begin() + span[a].second);
    }
  };
  if (L.level > 0)
    per_patch([&](size_t a, std::vector<BoxDesc>& v) {
      for (const Cand& b : cand[a])
        if (b.second & 1u) ghost_overlap_boxes(c, ps[a], ps[b.first], nullptr, kCopy, v);
    }, sib);
  bool all_wrap = periodic && ps.size() == 1;
  if (periodic) {
    for (int pid : ps) {
      const bsam_patch_desc& d = c->desc[pid];
      const DPatch& P = c->hp[pid];
      for (int dim = 0; dim < nd; dim++) {
        if (!(d.mthbc[2 * dim] == 2 && d.mthbc[2 * dim + 1] == 2)) {
          all_wrap = false;
          continue;
        }
        std::vector<BoxDesc>& dst = dim == 0 ? px : (dim == 1 ? py : pz);
        for (int side = 0; side < 2; side++) {
          BoxDesc b{};
          b.dp = b.sp = pid;
          b.op = (dim == 0 && side == 0) ? kCopyPlus0 : kCopy;
          for (int n = 0; n < 3; n++) {
            b.dlo[n] = b.slo[n] = 0;
            b.ext[n] = P.nq[n];
          }
          b.ext[dim] = 1;
          b.dlo[dim] = side == 0 ? P.mx[dim] + 1 : 0;
          b.slo[dim] = side == 0 ? 1 : P.mx[dim];
          dst.push_back(b);
        }
      }
    }
    if (L.level > 0)
      per_patch([&](size_t a, std::vector<BoxDesc>& v) {
        for (const Cand& b : cand[a])
          for (int pol = -1; pol <= 1; pol += 2)
            for (size_t oi = 0; oi < offs.size(); oi++) {
              // only the shifts under which the boxes may meet
              if (!((b.second >> (1 + 2 * oi + (pol > 0 ? 1 : 0))) & 1u)) continue;
              const auto& o = offs[oi];
              int off[3] = {pol * o[0], pol * o[1], pol * o[2]};
              ghost_overlap_boxes(c, ps[a], ps[b.first], off, kCopy, v);
            }
      }, pp);
  }
  // physical faces (SetPhysicalBC2D/3D, boundary.f90:642-802)
  bool has_phys = false;
  for (int pid : ps) {
    const bsam_patch_desc& d = c->desc[pid];
    const DPatch& P = c->hp[pid];
    for (int face = 0; face < 2 * nd; face++) {
      const int code = d.mthbc[face];
      if (code == 2 || code == 999) continue;
      if (code != 10)
        return set_err(BSAM_EINVAL, "SetPhysicalBC: user boundary condition %d on boundary %d not coded", code,
                       face + 1);
      has_phys = true;
      const int dim = face / 2, hi = face % 2;
      BoxDesc b{};
      b.dp = b.sp = pid;
      b.op = phys_face_op(c);
      for (int n = 0; n < 3; n++) {
        b.dlo[n] = b.slo[n] = 0;
        b.ext[n] = P.nq[n];
      }
      b.ext[dim] = 1;
      b.dlo[dim] = hi ? P.mx[dim] + 1 : 0;
      b.slo[dim] = hi ? P.mx[dim] : 1;
      phys[dim].push_back(b);
    }
  }
  if (has_phys) all_wrap = false;
  G.stages.clear();
  G.wrap_pid = -1;
  if (all_wrap && L.level <= 0) {
    G.wrap_pid = ps[0];
    return BSAM_OK;
  }
  auto add = [&](std::vector<BoxDesc>& v) {
    if (!v.empty()) G.stages.push_back(std::move(v));
  };
  const bool no_self = px.empty() && py.empty() && pz.empty();
  if (no_self) {
    sib.insert(sib.end(), pp.begin(), pp.end());
    add(sib);
  } else {
    add(sib);
    add(px);
    add(py);
    add(pz);
    add(pp);
  }
  for (int dim = 0; dim < 3; dim++) add(phys[dim]);
  return BSAM_OK;
}

// ---------------------------------------------------------------------------
// Flattened ghost fill.  SetGhost(level, ipass) runs its stages in order
// (interpolation, sibling transfers, self-periodic, periodic pairs,
// physical conditions), later stages overwriting earlier ones.  Executing
// the stages symbolically on the host gives, for every ghost cell the call
// writes, the expression its final value is: a copy of one cell or one
// interpolation rule, followed by a chain of +0.0 / Dirichlet images when
// a stage read a ghost cell that an earlier stage of the same call wrote.
// When every expression reads only interior cells of the level and parent
// cells (which no stage writes), the expressions are independent and one
// gather launch evaluates them all with the bits of the staged order.
// ---------------------------------------------------------------------------
struct FExpr {
  int kind;   // 0: copy of cell `src`; 1: interpolation rule `src` of the plan;
              // 2: a value read from an edge or corner ghost the call has not
              // written (allowed only where a later stage overwrites it)
  int src;
  int svs;    // copy: source variable stride
  int dvs;    // destination variable stride
  int spid;   // copy: source patch
  int dpid;   // destination patch
  int post;   // post-operation chain (ghost_post)
  int npost;
  int early;  // reads a face ghost as it was before the call
};

__host__ __device__ inline void qdecode(const DPatch& P, int o, int& i, int& j, int& k) {
  int r = o - P.qoff;
  const int m = r % P.hq;
  r /= P.hq;
  j = r % P.nq[1];
  r /= P.nq[1];
  k = r % P.nq[2];
  const int col = r / P.nq[2];
  i = 2 * m + ((col + j + k) & 1);
}
__host__ __device__ inline int n_outside(const DPatch& P, int nd, int i, int j, int k) {
  return (i < 1 || i > P.mx[0]) + (j < 1 || j > P.mx[1]) + (nd == 3 && (k < 1 || k > P.mx[2]));
}

// Expressions of one flattened call, indexed by the variable-0 cell of a
// level patch: dense(pid, off) = base[pos[pid]] + off - qoff(pid).
struct LevelCells {
  std::vector<int> pos;     // patch -> position in the level, or -1
  std::vector<long> base;
  long n = 0;
  LevelCells(const bsam_ctx* c, const Level& L) : pos(c->npatch, -1), base(L.patches.size()) {
    for (size_t a = 0; a < L.patches.size(); a++) {
      pos[L.patches[a]] = (int)a;
      base[a] = n;
      n += c->hp[L.patches[a]].qvs;
    }
  }
  long at(const bsam_ctx* c, int pid, int off) const { return base[pos[pid]] + (off - c->hp[pid].qoff); }
};
struct FlatCall {
  std::vector<int> at;     // dense cell -> index into ex, or -1
  std::vector<FExpr> ex;
  std::vector<int> dst;    // pool offset of the cell of ex[n]
  const FExpr* find(long cell) const { return at[cell] < 0 ? nullptr : &ex[at[cell]]; }
};

// Contiguous ranges of a level's patches with about equal dense cell
// counts: chunk t is patches [pb[t], pb[t+1]) and dense cells
// [LC.base[pb[t]], LC.base[pb[t+1]]).
static std::vector<int> patch_chunks(const LevelCells& LC, int np, int nch) {
  std::vector<int> pb(nch + 1, np);
  pb[0] = 0;
  int a = 0;
  for (int t = 1; t < nch; t++) {
    const long goal = LC.n * t / nch;
    while (a < np && LC.base[a] < goal) a++;
    pb[t] = a;
  }
  return pb;
}

// The symbolic run of SetGhost(level, ipass), on `nch` host threads that own
// contiguous ranges of the level's patches.  A stage's boxes read only
// cells that earlier stages (or no stage) wrote, and each cell has at most
// one writer per stage, so the stages run one after another with the
// threads writing the expressions of their own patches' cells in parallel.
// A chunk's expressions go to its own array, reserved to its ghost-cell
// count, so that other chunks read finished entries while it appends.
static bool flatten_setghost(const bsam_ctx* c, const Level& L, const GhostPlan& G, const LevelCells& LC,
                             int ipass, FlatCall& E, std::string& why, int nch) {
  const int nd = c->p.ndims;
  const int np = (int)L.patches.size();
  nch = std::max(1, std::min(nch, np));
  const std::vector<int> pb = patch_chunks(LC, np, nch);
  std::vector<int> owner(np);
  for (int t = 0; t < nch; t++)
    for (int a = pb[t]; a < pb[t + 1]; a++) owner[a] = t;
  constexpr int kShift = 27;
  constexpr int kMask = (1 << kShift) - 1;
  E.at.assign(LC.n, -1);
  std::vector<std::vector<FExpr>> ex(nch);
  std::vector<std::vector<int>> dst(nch);
  for (int t = 0; t < nch; t++) {
    long ghosts = 0;
    for (int a = pb[t]; a < pb[t + 1]; a++) {
      const DPatch& P = c->hp[L.patches[a]];
      ghosts += (long)P.nq[0] * P.nq[1] * P.nq[2] - (long)P.mx[0] * P.mx[1] * P.mx[2];
    }
    if (ghosts > kMask || nch > 15) {
      why = "a level too large for the chunked symbolic fill";
      return false;
    }
    ex[t].reserve(ghosts);
    dst[t].reserve(ghosts);
  }
  auto find = [&](long cell) -> const FExpr* {
    const int v = E.at[cell];
    return v < 0 ? nullptr : &ex[v >> kShift][v & kMask];
  };
  std::vector<std::string> err(nch);
  // a chunk appends only for cells of its own patches
  auto set = [&](int t, long cell, int d, const FExpr& e) -> bool {
    const int v = E.at[cell];
    if (v >= 0) {
      ex[t][v & kMask] = e;
      return true;
    }
    if (ex[t].size() >= ex[t].capacity()) {
      err[t] = "more ghost expressions than ghost cells";
      return false;
    }
    E.at[cell] = (t << kShift) | (int)ex[t].size();
    ex[t].push_back(e);
    dst[t].push_back(d);
    return true;
  };
  const int s0 = (ipass == 2) ? 1 : 0, s1 = (ipass == 1) ? 1 : 2;
  run_threads(nch, [&](int t) {
    for (int sg = s0; sg <= s1; sg++)
      for (int a = pb[t]; a < pb[t + 1]; a++)
        for (int r = G.roff[sg][a]; r < G.roff[sg][a + 1]; r++) {
          const FExpr e{1, r, 0, G.rules[r].pvs, -1, L.patches[a], 0, 0, 0};
          if (!set(t, LC.at(c, L.patches[a], G.rules[r].dst), G.rules[r].dst, e)) return;
        }
  });
  for (const std::string& m : err)
    if (!m.empty()) {
      why = m;
      return false;
    }
  for (const auto& st : G.stages) {
    std::vector<std::vector<int>> mine(nch);
    for (size_t n = 0; n < st.size(); n++) mine[owner[LC.pos[st[n].dp]]].push_back((int)n);
    run_threads(nch, [&](int t) {
      for (int n : mine[t]) {
        const BoxDesc& B = st[n];
        const DPatch& D = c->hp[B.dp];
        const DPatch& S = c->hp[B.sp];
        const int code = (B.op == kCopy) ? 0
                         : (B.op == kCopyPlus0 ? 1 : (B.op == kDirichlet2 ? 2 : (B.op == kBgMixed ? 3 : 4)));
        for (int z = 0; z < B.ext[2]; z++)
          for (int y = 0; y < B.ext[1]; y++)
            for (int x = 0; x < B.ext[0]; x++) {
              const int si = B.slo[0] + x, sj = B.slo[1] + y, sk = B.slo[2] + z;
              const int s = qidx(S, si, sj, sk);
              const int d = qidx(D, B.dlo[0] + x, B.dlo[1] + y, B.dlo[2] + z);
              if (!n_outside(D, nd, B.dlo[0] + x, B.dlo[1] + y, B.dlo[2] + z)) {
                err[t] = "a ghost-fill stage writes an interior cell (overlapping patches)";
                return;
              }
              FExpr e;
              const FExpr* prev = find(LC.at(c, B.sp, s));
              if (prev) {
                e = *prev;
                if (e.kind == 1 && e.dpid != B.dp) {
                  err[t] = "an interpolated ghost value is copied into another patch";
                  return;
                }
              } else {
                // a ghost not written so far in this call is read as it was
                // before the call (physical conditions over ranges that
                // include the ghost rows, boundary.f90:642-802); an edge or
                // corner ghost read so is only tolerated in values that a
                // later stage overwrites (the x faces of a 3D box read the
                // y-z edges, whose images the z faces rewrite)
                const int out = n_outside(S, nd, si, sj, sk);
                e = FExpr{out > 1 ? 2 : 0, s, S.qvs, 0, B.sp, 0, 0, 0, out};
              }
              e.dvs = D.qvs;
              e.dpid = B.dp;
              if (code) {
                if (e.npost >= 5) {
                  err[t] = "post-operation chain longer than five";
                  return;
                }
                e.post |= code << (3 * e.npost);
                e.npost++;
              }
              if (!set(t, LC.at(c, B.dp, d), d, e)) return;
            }
      }
    });
    for (const std::string& m : err)
      if (!m.empty()) {
        why = m;
        return false;
      }
  }
  // one expression array, chunk after chunk; each chunk renumbers its cells
  std::vector<int> base(nch + 1, 0);
  for (int t = 0; t < nch; t++) base[t + 1] = base[t] + (int)ex[t].size();
  E.ex.resize(base[nch]);
  E.dst.resize(base[nch]);
  std::vector<char> bad(nch, 0);
  run_threads(nch, [&](int t) {
    std::copy(ex[t].begin(), ex[t].end(), E.ex.begin() + base[t]);
    std::copy(dst[t].begin(), dst[t].end(), E.dst.begin() + base[t]);
    for (const FExpr& e : ex[t])
      if (e.kind == 2) bad[t] = 1;
    const long c0 = LC.base[pb[t]], c1 = pb[t + 1] < np ? LC.base[pb[t + 1]] : LC.n;
    for (long cell = c0; cell < c1; cell++)
      if (E.at[cell] >= 0) E.at[cell] = base[t] + (E.at[cell] & kMask);
  });
  for (int t = 0; t < nch; t++)
    if (bad[t]) {
      why = "a ghost value is read from an edge or corner ghost the call has not written";
      return false;
    }
  return true;
}

static bool same_expr(const GhostPlan& G, const FExpr& a, const FExpr& b) {
  if (a.kind != b.kind || a.post != b.post || a.npost != b.npost) return false;
  if (a.kind == 0) return a.src == b.src && a.svs == b.svs;
  const GRule &r = G.rules[a.src], &s = G.rules[b.src];
  if (r.kind != s.kind || r.s1 != s.s1 || r.s2 != s.s2 || r.cvs != s.cvs) return false;
  for (int n = 0; n < r.kind; n++)
    if (r.c[n] != s.c[n]) return false;
  return true;
}

// The flattened maps of a level (at and below the root only where the fill
// is box stages: a single patch periodic in every dimension keeps its wrap
// kernel).
//
// map[p], p = 0, 1, 2: SetGhost(level, p) in full.
// map[2 + p], p = 1, 2: the face ghosts X that SetGhost(level, p) writes
// and for which (a) X has the colour p-1 just relaxed, or (b) X's
// expression reads a cell of colour p-1, or (c) X's expression differs
// from its expression in SetGhost(level, 3-p); and the edge and corner
// ghosts SetGhost(level, p) writes but SetGhost(level, 3-p) does not (2D
// corners are colour-filtered, boundary.f90:155-181).
//
// Use of map[2 + p] inside a red-black sweep sequence (LevelRelax: sweep
// colour 0, SetGhost(1), sweep colour 1, SetGhost(2), ..., the last
// SetGhost in full) leaves every value a 7-point sweep, residual or FAS
// right-hand side reads equal to the staged fill's: a colour-c sweep reads
// only face ghosts of colour 1-c.  A face ghost X the full call would
// write but map[2 + p] skips has colour 2-p, its expression E is the same
// at ipass 1 and 2 and reads nothing of colour p-1: after the first
// SetGhost(2) of the sequence (which writes every colour-1 face ghost of
// map[2]) X holds E evaluated after the previous call, whose sources the
// sweep in between did not change.  Before that point only colour-1 face
// ghosts can differ, and the colour-1 sweep does not read them.  Edge and
// corner ghosts are not read by 7-point stencils; those both ipasses write
// are restored by the last SetGhost of the sequence, which is full, the
// others are rewritten at every call of their ipass.
//
// map[5], the colour-1 face ghosts of SetGhost(level, 0), may stand for it
// when a sweep sequence follows (the post-smoothing after CorrectFine):
// the first sweep (colour 0) reads exactly those ghosts.  Every colour-0
// face ghost is rewritten by the SetGhost(1) after it; a colour-1 face
// ghost that SetGhost(1) writes has the same expression at ipass 0, 1 and
// 2 (a copy, since copies follow interpolation), so after map[5] it holds
// what the full calls would give until SetGhost(2) rewrites it; and
// SetGhost(level, 0) writes no cell that neither ipass 1 nor 2 writes, so
// the sequence ends with every cell as the staged fills leave it.
// Cells along dimension a of the whole domain at `level` (the root's extent
// refined level times).
static int level_period(const bsam_ctx* c, int level, int a) {
  for (const bsam_patch_desc& d : c->desc)
    if (d.level == 0) return level >= 0 ? d.mx[a] << level : d.mx[a] >> -level;
  return 1;
}

// Cells of the whole domain at `level` (x fastest) that lie inside one of
// the level's patches.
static std::vector<uint8_t> parent_interior(const bsam_ctx* c, int level) {
  const int nd = c->p.ndims;
  long n = 1;
  int per[3] = {1, 1, 1};
  for (int a = 0; a < nd; a++) {
    per[a] = level_period(c, level, a);
    n *= per[a];
  }
  std::vector<uint8_t> occ(n, 0);
  for (const bsam_patch_desc& d : c->desc) {
    if (d.level != level) continue;
    const int k0 = nd == 3 ? d.mglobal[2][0] : 1, k1 = nd == 3 ? d.mglobal[2][1] : 1;
    for (int k = k0; k <= k1; k++)
      for (int j = d.mglobal[1][0]; j <= d.mglobal[1][1]; j++)
        for (int i = d.mglobal[0][0]; i <= d.mglobal[0][1]; i++)
          occ[((long)(k - 1) * per[1] + (j - 1)) * per[0] + (i - 1)] = 1;
  }
  return occ;
}

int32_t plan_ghost_maps(const bsam_ctx* c, const Level& L, GhostPlan& G) {
  G.maps_ok = false;
  if (G.wrap_pid >= 0 || (L.level <= 0 && G.stages.empty())) return BSAM_OK;
  const int nd = c->p.ndims;
  const LevelCells LC(c, L);
  FlatCall E[3];
  bool ok[3] = {false, false, false};
  std::string why[3];
  // the three calls are independent, and each runs on its share of the
  // host threads (levels of a few patches on one thread each)
  const int share = LC.n < (1L << 20) ? 1 : std::max<int>(1, (int)host_threads() / 3);
  run_threads(3, [&](int p) { ok[p] = flatten_setghost(c, L, G, LC, p, E[p], why[p], share); });
  for (int p = 0; p < 3; p++) {
    if (!ok[p]) {
      G.why = why[p];
      return BSAM_OK;
    }
    // the rules that read ghost values from before the call run in a first
    // launch; none of them may write a cell another of them reads
    for (size_t n = 0; n < E[p].ex.size(); n++) {
      const FExpr& e = E[p].ex[n];
      if (!e.early) continue;
      const FExpr* w = E[p].find(LC.at(c, e.spid, e.src));
      if (w && w->early) {
        G.why = "a ghost read from before the call is rewritten by another such rule";
        return BSAM_OK;
      }
    }
  }
  // Each table holds its rules in destination order (a level's patches are
  // in ascending pool offset, so the dense cell index ascends with the
  // destination offset), those that read ghosts from before the call
  // (phase 1) ahead of the others.  The dense range is cut into chunks of
  // whole patches that host threads emit in parallel; the parts are joined
  // in chunk order, phase 1 first.
  struct Part {
    std::vector<int4> cp[2];
    std::vector<uint16_t> post[2];
    std::vector<GRule> in[2];
    bool anypost = false;
  };
  auto emit = [&](const FlatCall& M, int n, Part& P) {
    const FExpr& e = M.ex[n];
    const int d = M.dst[n];
    const int ph = e.early ? 1 : 0;
    if (e.kind == 0) {
      P.cp[ph].push_back(make_int4(d, e.src, e.dvs, e.svs));
      P.post[ph].push_back((uint16_t)e.post);
      P.anypost = P.anypost || e.post;
    } else {
      GRule r = G.rules[e.src];
      r.dst = d;
      r.kind = (r.kind & 0xff) | (e.post << 8);
      P.in[ph].push_back(r);
    }
  };
  auto join = [&](std::vector<Part>& parts, HostMap& H) {
    const int K = (int)parts.size();
    std::vector<size_t> oc(2 * K + 1, 0), oi(2 * K + 1, 0);
    bool anypost = false;
    int slot = 0;
    for (int ph = 1; ph >= 0; ph--)
      for (int k = 0; k < K; k++, slot++) {
        oc[slot + 1] = oc[slot] + parts[k].cp[ph].size();
        oi[slot + 1] = oi[slot] + parts[k].in[ph].size();
        anypost = anypost || parts[k].anypost;
      }
    H.copies.resize(oc[2 * K]);
    H.cpost.resize(anypost ? oc[2 * K] : 0);
    H.interps.resize(oi[2 * K]);
    H.nc1 = (int)oc[K];
    H.ni1 = (int)oi[K];
    run_threads(K, [&](int k) {
      for (int ph = 1; ph >= 0; ph--) {
        const int sl = (1 - ph) * K + k;
        const Part& P = parts[k];
        std::copy(P.cp[ph].begin(), P.cp[ph].end(), H.copies.begin() + oc[sl]);
        if (anypost) std::copy(P.post[ph].begin(), P.post[ph].end(), H.cpost.begin() + oc[sl]);
        std::copy(P.in[ph].begin(), P.in[ph].end(), H.interps.begin() + oi[sl]);
      }
    });
  };
  // map[6]'s test of the coarse cells a prolongation reads (see below)
  const std::vector<uint8_t> occ = parent_interior(c, L.level - 1);
  int cper[3] = {1, 1, 1}, cwrap[3] = {0, 0, 0};
  {
    const bsam_patch_desc* root = nullptr;
    for (const bsam_patch_desc& d : c->desc)
      if (d.level == 0) root = &d;
    for (int a = 0; a < nd; a++) {
      cper[a] = level_period(c, L.level - 1, a);
      cwrap[a] = root->mthbc[2 * a] == 2;
    }
  }
  auto covered = [&](const int* g) {
    long idx = 0;
    for (int a = nd - 1; a >= 0; a--) {
      int x = g[a] - 1;
      if (x < 0 || x >= cper[a]) {
        if (!cwrap[a]) return false;
        x = ((x % cper[a]) + cper[a]) % cper[a];
      }
      idx = idx * cper[a] + x;
    }
    return occ[idx] != 0;
  };
  // map[5]: the colour-1 face ghosts of SetGhost(level, 0).  map[6]: map[5]
  // less the copies that CorrectFine leaves exact.  A ghost whose value is
  // a plain copy (no post-operation) of an interior cell at the same global
  // position (modulo the period) held that cell's value before the
  // correction (the last SetGhost of the pre-smoothing wrote it), and both
  // receive the prolongation of the same coarse correction, by the same
  // expression, provided every coarse cell of the prolongation stencil
  // holds one value for both patches: a cell inside a patch of the parent
  // level (its ghost images in other parents are copies of it, and the
  // coarse-under-fine values are restrictions of the same fine cells).  A
  // coarse cell outside every parent-level patch is a coarse-fine ghost
  // that each parent interpolates from its own cells; such ghosts, and
  // reflecting copies (mirrored position), stay in the table.
  // Returns 0 (not in map[5]), 1 (map[5] and map[6]), 2 (map[5] only), -1
  // (refused).
  auto sel56 = [&](int n, std::string& why) -> int {
    const FExpr& e = E[0].ex[n];
    const DPatch& D = c->hp[e.dpid];
    int i, j, k;
    qdecode(D, E[0].dst[n], i, j, k);
    if (!(n_outside(D, nd, i, j, k) == 1 && ((i + j + k) & 1) == 1)) return 0;
    if (e.early) {
      why = "a colour-1 face ghost of SetGhost(level, 0) reads a ghost from before the call";
      return -1;
    }
    bool exact = e.kind == 0 && e.post == 0;
    if (exact) {
      const DPatch& S = c->hp[e.spid];
      int s3[3], d3[3];
      qdecode(S, e.src, s3[0], s3[1], s3[2]);
      qdecode(D, E[0].dst[n], d3[0], d3[1], d3[2]);
      exact = n_outside(S, nd, s3[0], s3[1], s3[2]) == 0;
      for (int a = 0; a < nd && exact; a++) {
        const int per = level_period(c, L.level, a);
        const int ds = (c->desc[e.spid].mglobal[a][0] + s3[a]) - (c->desc[e.dpid].mglobal[a][0] + d3[a]);
        exact = ((ds % per) + per) % per == 0;
      }
      // the coarse cells of the prolongation of the destination: along
      // each dimension an odd fine index g takes coarse (g+1)/2 - 1 and
      // (g+1)/2, an even one g/2 and g/2 + 1 (BiLinProlongationP1,
      // TriLinProlongationP1, gridutilities.f90:156-195, 510-566)
      for (int corner = 0; corner < (1 << nd) && exact; corner++) {
        int g[3] = {1, 1, 1};
        for (int a = 0; a < nd; a++) {
          const int gf = c->desc[e.dpid].mglobal[a][0] + d3[a] - 1;
          const int lo = (gf & 1) ? (gf + 1) / 2 - 1 : gf / 2;
          g[a] = lo + ((corner >> a) & 1);
        }
        exact = covered(g);
      }
    }
    return exact ? 2 : 1;
  };
  // map[2 + p], p = 1, 2 (described above): 1 in the table, 0 not, -1 refused
  auto lean = [&](int p, int n, std::string& why) -> int {
    const int col = p - 1;
    const FlatCall& other = E[3 - p];
    const FExpr& e = E[p].ex[n];
    const int d = E[p].dst[n];
    const DPatch& D = c->hp[e.dpid];
    int i, j, k;
    qdecode(D, d, i, j, k);
    const FExpr* o = other.find(LC.at(c, e.dpid, d));
    if (n_outside(D, nd, i, j, k) != 1) {
      // an edge or corner ghost: only if the other ipass does not write it
      if (!o && e.early) {
        why = "a colour-filtered edge or corner ghost reads a ghost from before the call";
        return -1;
      }
      return o ? 0 : 1;
    }
    bool need = ((i + j + k) & 1) == col;
    if (!need) {
      int si, sj, sk;
      if (e.kind == 0) {
        qdecode(c->hp[e.spid], e.src, si, sj, sk);
        need = ((si + sj + sk) & 1) == col;
      } else {
        const GRule& r = G.rules[e.src];
        qdecode(D, r.s1, si, sj, sk);
        need = ((si + sj + sk) & 1) == col;
        qdecode(D, r.s2, si, sj, sk);
        need = need || ((si + sj + sk) & 1) == col;
      }
    }
    if (!need) need = !o || !same_expr(G, e, *o);
    if (need && e.early) {
      why = "a face ghost of a reduced table reads a ghost from before the call";
      return -1;
    }
    return need ? 1 : 0;
  };
  // the three ipasses' tables are independent; each runs on its share of
  // the host threads
  const int np = (int)L.patches.size();
  const int K = std::max(1, std::min(share, np));
  const std::vector<int> pb = patch_chunks(LC, np, K);
  auto plan_p = [&](int p, std::string& why) -> bool {
    std::vector<Part> full(K), sub(K), sub6(K);
    std::vector<std::string> err(K);
    run_threads(K, [&](int k) {
      const long c0 = LC.base[pb[k]], c1 = pb[k + 1] < np ? LC.base[pb[k + 1]] : LC.n;
      for (long cell = c0; cell < c1; cell++) {
        const int n = E[p].at[cell];
        if (n < 0) continue;
        emit(E[p], n, full[k]);
        if (p == 0) {
          const int r = sel56(n, err[k]);
          if (r < 0) return;
          if (r >= 1) emit(E[0], n, sub[k]);
          if (r == 1) emit(E[0], n, sub6[k]);
        } else {
          const int r = lean(p, n, err[k]);
          if (r < 0) return;
          if (r) emit(E[p], n, sub[k]);
        }
      }
    });
    for (const std::string& m : err)
      if (!m.empty()) {
        why = m;
        return false;
      }
    join(full, G.map[p]);
    if (p == 0) {
      join(sub, G.map[5]);
      join(sub6, G.map[6]);
    } else {
      join(sub, G.map[2 + p]);
    }
    return true;
  };
  {
    bool ok[3] = {false, false, false};
    std::string why[3];
    run_threads(3, [&](int p) { ok[p] = plan_p(p, why[p]); });
    for (int p = 0; p < 3; p++)
      if (!ok[p]) {
        G.why = why[p];
        return BSAM_OK;
      }
  }
  G.maps_ok = true;
  return BSAM_OK;
}

// Small device arrays of one level staged in one host buffer (256-byte
// aligned), then copied with one copy into one pooled block the level owns;
// each array's device address is stored where add() was told.  The host
// buffer must outlive the copy: upload() waits for it.
struct BlockBuilder {
  std::vector<char> h;
  std::vector<std::pair<size_t, void**>> fix;
  template <typename T>
  void add(const T* src, size_t count, T** dst) {
    const size_t off = (h.size() + 255) & ~size_t(255);
    h.resize(off + std::max<size_t>(count * sizeof(T), 1));
    if (count) memcpy(h.data() + off, src, count * sizeof(T));
    fix.push_back({off, reinterpret_cast<void**>(dst)});
  }
  int32_t upload(bsam_ctx* c, Level& L) {
    if (h.empty()) return BSAM_OK;
    void* b = nullptr;
    if (pool_alloc(c, &b, h.size(), 1) != hipSuccess) return set_err(BSAM_ENOMEM, "level tables of %zu bytes", h.size());
    L.blocks.push_back(b);
    HIPCHK(hipMemcpyAsync(b, h.data(), h.size(), hipMemcpyHostToDevice, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    for (auto& f : fix) *f.second = static_cast<char*>(b) + f.first;
    h.clear();
    fix.clear();
    return BSAM_OK;
  }
};

// The ghost fill of a level planned on the host, staged for the level's
// block: box stages, host-built interpolation rules, flattened tables.
void upload_ghost(bsam_ctx* c, Level& L, const GhostPlan& G, BlockBuilder& Bk) {
  L.wrap_pid = G.wrap_pid;
  L.wrap_img = false;
  if (L.wrap_pid >= 0) {
    const DPatch& P = c->hp[L.wrap_pid];
    L.wrap_img = !(P.mx[0] & 1) && !(P.mx[1] & 1) && (c->p.ndims == 2 || !(P.mx[2] & 1));
  }
  L.stages.clear();
  for (const auto& boxes : G.stages) {
    Stage s;
    s.n = (int)boxes.size();
    for (const BoxDesc& b : boxes) s.maxel = std::max(s.maxel, b.ext[0] * b.ext[1] * b.ext[2]);
    L.stages.push_back(s);
  }
  for (size_t st = 0; st < G.stages.size(); st++) Bk.add(G.stages[st].data(), G.stages[st].size(), &L.stages[st].d);
  for (int sg = 0; sg < 3; sg++) L.roff[sg] = G.roff[sg];
  if (!G.rules.empty()) Bk.add(G.rules.data(), G.rules.size(), &L.d_rules);
  L.maps_ok = G.maps_ok;
  if (!G.maps_ok) return;
  for (int m = 0; m < 7; m++) {
    const HostMap& H = G.map[m];
    DevMap& D = L.gmap[m];
    D.nc = (int)H.copies.size();
    D.ni = (int)H.interps.size();
    D.nc1 = H.nc1;
    D.ni1 = H.ni1;
    if (D.nc) Bk.add(H.copies.data(), H.copies.size(), &D.cr);
    if (!H.cpost.empty()) Bk.add(H.cpost.data(), H.cpost.size(), &D.cpost);
    if (D.ni) Bk.add(H.interps.data(), H.interps.size(), &D.ir);
  }
}

// The device planners' workspace (kept by the context, grown as needed).
static char* plan_workspace(bsam_ctx* c, size_t bytes) {
  if (bytes > c->dplan_cap) {
    if (hipStreamSynchronize(c->s) != hipSuccess) return nullptr;
    hipFree(c->dplan);
    c->dplan = nullptr;
    c->dplan_cap = 0;
    if (hipMalloc(&c->dplan, bytes) != hipSuccess) return nullptr;
    c->dplan_cap = bytes;
  }
  return c->dplan;
}

// ---------------------------------------------------------------------------
// The flattened ghost fill planned on the device, for levels above the root
// (batched launches): flatten_setghost's symbolic run of SetGhost(level,
// ipass) and plan_ghost_maps' selection of the seven tables, cell by cell.
// The expressions of one call live in a dense array over the level's q
// cells (variable 0, LevelCells' numbering).  Within a stage the host runs
// the boxes in order, so the last box that writes a cell wins; here a
// claim pass (atomicMax of the box index) picks that box and a write pass
// lets it write.  The interpolation stage likewise keeps the last rule in
// (segment, patch, element) order.  Every table is compacted in the host's
// order (dense cell order, the rules that read ghosts from before the call
// first), so the tables equal plan_ghost_maps' rule for rule.  What the
// host refuses is refused here too, and so is a box reading a cell that
// its own stage writes (an order the claim cannot reproduce): the level
// then keeps the staged fill.
// ---------------------------------------------------------------------------
struct DExpr {
  int src, spid, dpid, meta;
};
// meta: valid bit, kind (2 bits: 0 copy, 1 interpolation rule, 2 a value
// read from an unwritten edge or corner ghost), early (2 bits: outside
// dimensions of a ghost read from before the call), post-operation count
// (3 bits) and chain (15 bits)
constexpr int kDxValid = 1 << 30;
__host__ __device__ inline int dx_kind(int m) { return m & 3; }
__host__ __device__ inline int dx_early(int m) { return (m >> 2) & 3; }
__host__ __device__ inline int dx_npost(int m) { return (m >> 4) & 7; }
__host__ __device__ inline int dx_post(int m) { return (m >> 7) & 0x7fff; }
__host__ __device__ inline int dx_make(int kind, int early, int npost, int post) {
  return kDxValid | kind | (early << 2) | (npost << 4) | (post << 7);
}
// refusal bits of the device planner
enum : int {
  kDpInterior = 1,    // a stage writes an interior cell
  kDpOrder = 2,       // a box reads a cell its own stage writes
  kDpInterpCopy = 4,  // an interpolated ghost copied into another patch
  kDpChain = 8,       // post-operation chain longer than five
  kDpUnknown = 16,    // a value read from an unwritten edge or corner ghost survives
  kDpEarly = 32,      // a ghost read from before the call is rewritten by another such rule
  kDpSel5 = 64,       // a colour-1 face ghost of SetGhost(level, 0) reads a ghost from before the call
  kDpLean = 128,      // a reduced table's ghost reads a ghost from before the call
};

// Ghost cells of a patch numbered in pool order (colour, plane, row, half
// index), for an even row length nx = mx + 2 (every refined patch): per
// colour c, the plane k = 0 (ny rows of nx/2 cells of colour c), the inner
// planes (row 0 and row ny - 1 with nx/2 each, and between them the one
// colour-c end cell of each inner row: i = 0 and i = nx - 1 differ in
// colour), then the plane nz - 1.  2D: one inner plane per colour.
__host__ __device__ inline long shell_count(const DPatch& P, int nd) {
  const long nz = nd == 3 ? P.nq[2] : 1, mz = nd == 3 ? P.mx[2] : 1;
  return (long)P.nq[0] * P.nq[1] * nz - (long)P.mx[0] * P.mx[1] * mz;
}
// the index of ghost cell (i, j, k) (storage indices; not for an interior
// cell)
__device__ __forceinline__ int shell_index(const DPatch& P, int nd, int i, int j, int k) {
  const int nx = P.nq[0], ny = P.nq[1], h = nx >> 1, inner = nx + ny - 2;
  const int c = (i + j + k) & 1, m = i >> 1;
  int base;
  if (nd == 3) {
    const int nz = P.nq[2];
    base = c * (ny * nx + (nz - 2) * inner);
    if (k == 0) return base + j * h + m;
    base += ny * h;
    if (k == nz - 1) return base + (nz - 2) * inner + j * h + m;
    base += (k - 1) * inner;
  } else {
    base = c * inner;
  }
  if (j == 0) return base + m;
  if (j == ny - 1) return base + h + (ny - 2) + m;
  return base + h + (j - 1);
}
// the cell of ghost index s
__device__ __forceinline__ void shell_cell(const DPatch& P, int nd, int s, int& i, int& j, int& k) {
  const int nx = P.nq[0], ny = P.nq[1], h = nx >> 1, inner = nx + ny - 2;
  int c;
  k = 0;
  if (nd == 3) {
    const int nz = P.nq[2], per = ny * nx + (nz - 2) * inner;
    c = s >= per;
    s -= c * per;
    bool full = s < ny * h;
    if (!full) {
      s -= ny * h;
      if (s >= (nz - 2) * inner) {
        s -= (nz - 2) * inner;
        full = true;
        k = nz - 1;
      } else {
        k = 1 + s / inner;
        s -= (k - 1) * inner;
      }
    }
    if (full) {
      j = s / h;
      i = 2 * (s - j * h) + ((c + j + k) & 1);
      return;
    }
  } else {
    c = s >= inner;
    s -= c * inner;
  }
  if (s < h) {
    j = 0;
    i = 2 * s + ((c + j + k) & 1);
  } else if (s < h + ny - 2) {
    j = 1 + (s - h);
    i = ((j + k) & 1) == c ? 0 : nx - 1;
  } else {
    j = ny - 1;
    i = 2 * (s - h - (ny - 2)) + ((c + j + k) & 1);
  }
}
struct DPlanArgs {
  const DPatch* __restrict__ pt;
  const int* __restrict__ patches;  // the level's descriptor indices
  const int* __restrict__ pos;      // descriptor index -> position in the level, or -1
  const long* __restrict__ gbase;   // first ghost index of each level patch (np + 1 entries)
  const int* __restrict__ bp;       // workgroup -> level position, over the ghost cells
  const int* __restrict__ bfirst;   // first workgroup of each level patch there
  const int* __restrict__ eoff;     // first interpolation element of each level patch (np + 1)
  int np, nd;
};
__device__ __forceinline__ long dp_ghost(const DPlanArgs& A, int pid, int i, int j, int k) {
  return A.gbase[A.pos[pid]] + shell_index(A.pt[pid], A.nd, i, j, k);
}
// The level's ghost cells in pool order, each patch's from a workgroup
// boundary: thread t of workgroup b is the patch's ghost (b - bfirst) *
// kBlock + t.  Returns its ghost index (-1 past the patch's ghosts), its
// patch (level position) and storage indices.  The tables list their rules
// in this order, as the host planner does.
__device__ __forceinline__ long dp_cell(const DPlanArgs& A, int& a, int& i, int& j, int& k) {
  a = A.bp[blockIdx.x];
  const int local = (blockIdx.x - A.bfirst[a]) * kBlock + threadIdx.x;
  if (local >= A.gbase[a + 1] - A.gbase[a]) return -1;
  shell_cell(A.pt[A.patches[a]], A.nd, local, i, j, k);
  return A.gbase[a] + local;
}
// Interpolation element e of level patch pid (interp_rule_at) as a rule.
__device__ __forceinline__ GRule dp_rule(const DPlanArgs& A, int pid, int e) {
  const DPatch P = A.pt[pid];
  GRule r;
  int sg = 0;
  interp_rule_at(P, A.pt[P.parent], A.nd, e, r, sg);
  return r;
}

__global__ __launch_bounds__(kBlock) void k_dp_fill(int* __restrict__ a, long n, int v) {
  const long g = (long)blockIdx.x * kBlock + threadIdx.x;
  if (g < n) a[g] = v;
}

// The interpolation stage of the three calls at once, over the level's
// interpolation elements (patch a's are [eoff[a], eoff[a + 1])): element e
// of colour-filter segment sg belongs to the calls whose filter passes it
// (segment 0: ipass 0 and 1; 1: all; 2: ipass 0 and 2), and in each of them
// the last rule in (segment, patch, element) order writing a cell wins: the
// claim is that order as a 64-bit key.  The expression keeps the element
// (src) and its patch (dpid); the rule is made again where it is needed.
// claim and ex hold one array of n per call; ecell keeps each element's
// ghost index and segment (sg << 29 | cell; -1 for no rule) from the claim
// pass for the write pass.  write 0: claim; 1: the last claimant writes.
__global__ __launch_bounds__(kBlock) void k_dp_interp(DPlanArgs A, long total, long n,
                                                      unsigned long long* __restrict__ claim, DExpr* __restrict__ ex,
                                                      int* __restrict__ ecell, int write) {
  const long g = (long)blockIdx.x * kBlock + threadIdx.x;
  if (g >= total) return;
  const int a = elem_patch(A.eoff, A.np, g);
  const int pid = A.patches[a];
  const int e = (int)(g - A.eoff[a]);
  int sg = 0;
  long cell;
  if (!write) {
    const DPatch P = A.pt[pid];
    GRule r;
    if (!interp_rule_at(P, A.pt[P.parent], A.nd, e, r, sg)) {
      ecell[g] = -1;
      return;
    }
    int i, j, k;
    qdecode(P, r.dst, i, j, k);
    cell = A.gbase[a] + shell_index(P, A.nd, i, j, k);
    ecell[g] = (sg << 29) | (int)cell;
  } else {
    const int v = ecell[g];
    if (v < 0) return;
    sg = v >> 29;
    cell = v & ((1 << 29) - 1);
  }
  const unsigned long long key = (((unsigned long long)sg << 56) | ((unsigned long long)a << 28) | (unsigned)e) + 1ull;
  for (int p = 0; p < 3; p++) {
    if ((p == 1 && sg == 2) || (p == 2 && sg == 0)) continue;
    if (!write) atomicMax(&claim[p * n + cell], key);
    else if (claim[p * n + cell] == key) ex[p * n + cell] = DExpr{e, -1, pid, dx_make(1, 0, 0, 0)};
  }
}

// One box stage over the flat cells of its boxes (bpre: each box's first
// flat cell; x fastest within a box), for the three calls at once (the box
// stages are the same at every ipass; ex holds one array of n per call).
// `tag` = (stage number + 1) << 24 in the claims.
__global__ __launch_bounds__(kBlock) void k_dp_box(DPlanArgs A, const BoxDesc* __restrict__ boxes,
                                                   const int* __restrict__ bpre, int nbox, int ncell, int tag, long n,
                                                   int* __restrict__ claim, DExpr* __restrict__ ex,
                                                   int* __restrict__ err, int write) {
  const int g = blockIdx.x * kBlock + threadIdx.x;
  if (g >= ncell) return;
  int lo = 0, hi = nbox;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (bpre[mid] <= g) lo = mid;
    else hi = mid;
  }
  const BoxDesc B = boxes[lo];
  const int loc = g - bpre[lo];
  const int x = loc % B.ext[0], r = loc / B.ext[0];
  const int y = r % B.ext[1], z = r / B.ext[1];
  const DPatch D = A.pt[B.dp];
  const int di = B.dlo[0] + x, dj = B.dlo[1] + y, dk = B.dlo[2] + z;
  if (!n_outside(D, A.nd, di, dj, dk)) {
    atomicOr(err, kDpInterior);
    return;
  }
  const long dcell = dp_ghost(A, B.dp, di, dj, dk);
  const int me = tag | lo;
  if (!write) {
    atomicMax(&claim[dcell], me);
    return;
  }
  if (claim[dcell] != me) return;
  const DPatch S = A.pt[B.sp];
  const int si = B.slo[0] + x, sj = B.slo[1] + y, sk = B.slo[2] + z;
  const int s = qidx(S, si, sj, sk);
  const int out = n_outside(S, A.nd, si, sj, sk);
  long scell = -1;
  if (out) {
    // a ghost source: the expression an earlier stage left there, if any
    scell = dp_ghost(A, B.sp, si, sj, sk);
    if ((claim[scell] & 0x7f000000) == tag) {
      atomicOr(err, kDpOrder);
      return;
    }
  }
  // the post-operation the box's operation adds (ghost_post's codes)
  const int code = (B.op == kCopy) ? 0
                   : (B.op == kCopyPlus0 ? 1 : (B.op == kDirichlet2 ? 2 : (B.op == kBgMixed ? 3 : 4)));
  for (int p = 0; p < 3; p++) {
    DExpr e{s, B.sp, 0, dx_make(out > 1 ? 2 : 0, out, 0, 0)};
    if (scell >= 0) {
      const DExpr prev = ex[p * n + scell];
      if (prev.meta & kDxValid) {
        if (dx_kind(prev.meta) == 1 && prev.dpid != B.dp) {
          atomicOr(err, kDpInterpCopy);
          return;
        }
        e = prev;
      }
    }
    e.dpid = B.dp;
    if (code) {
      const int k = dx_npost(e.meta);
      if (k >= 5) {
        atomicOr(err, kDpChain);
        return;
      }
      e.meta = dx_make(dx_kind(e.meta), dx_early(e.meta), k + 1, dx_post(e.meta) | (code << (3 * k)));
    }
    ex[p * n + dcell] = e;
  }
}

// The call's end checks: no unknown value survives, and no rule that reads
// a ghost from before the call reads one that another such rule writes.
__global__ __launch_bounds__(kBlock) void k_dp_check(DPlanArgs A, const DExpr* __restrict__ ex, long n,
                                                     int* __restrict__ err) {
  const long g = (long)blockIdx.x * kBlock + threadIdx.x;
  if (g >= 3 * n) return;
  const DExpr e = ex[g];
  if (!(e.meta & kDxValid)) return;
  if (dx_kind(e.meta) == 2) atomicOr(err, kDpUnknown);
  if (dx_early(e.meta)) {
    const DPatch S = A.pt[e.spid];
    int i, j, k;
    qdecode(S, e.src, i, j, k);
    if (n_outside(S, A.nd, i, j, k)) {
      const DExpr w = ex[(g / n) * n + dp_ghost(A, e.spid, i, j, k)];
      if ((w.meta & kDxValid) && dx_early(w.meta)) atomicOr(err, kDpEarly);
    }
  }
}

// What plan_ghost_maps' sel56 and lean need besides the expressions.
struct DPlanMaps {
  const DExpr* __restrict__ ex[3];
  const int* __restrict__ mglo;      // descriptor index -> mglobal lower corner (1-based, 3 ints)
  const uint32_t* __restrict__ occ;  // parent level's cells inside its patches (bit map)
  int per[3], cper[3], cwrap[3];
};

// same_expr of the host planner; ra, rb the rules of a and b when they are
// interpolation expressions
__device__ __forceinline__ bool dp_same_expr(const DPlanArgs& A, const DExpr& a, const DExpr& b, const GRule& ra,
                                             const GRule& rb) {
  if (dx_kind(a.meta) != dx_kind(b.meta) || dx_post(a.meta) != dx_post(b.meta) || dx_npost(a.meta) != dx_npost(b.meta))
    return false;
  if (dx_kind(a.meta) == 0) return a.src == b.src && A.pt[a.spid].qvs == A.pt[b.spid].qvs;
  const GRule& r = ra;
  const GRule& s = rb;
  if (r.kind != s.kind || r.s1 != s.s1 || r.s2 != s.s2 || r.cvs != s.cvs) return false;
#pragma unroll
  for (int n = 0; n < 4; n++)
    if (n < r.kind && r.c[n] != s.c[n]) return false;
  return true;
}

// Table category of a cell: 0 none, 1 copy reading a ghost from before the
// call, 2 other copy, 3 and 4 the same for interpolation rules.
__device__ __forceinline__ uint8_t dp_cat(const DExpr& e) {
  const int late = dx_early(e.meta) ? 0 : 1;
  return (uint8_t)((dx_kind(e.meta) == 0 ? 1 : 3) + late);
}
// the call whose expression table m holds
__host__ __device__ inline int dp_call(int m) { return m < 3 ? m : (m == 3 ? 1 : (m == 4 ? 2 : 0)); }

constexpr int kDpWaves = kBlock / 64;

// Per ghost cell g (over the level's nf cells in pool order), the category
// in each of the seven tables (cat[m * n + g]), the refusals of sel56 and
// lean, which tables have a copy with a post-operation, and per workgroup
// the cells of every (table, category) (bcnt[(5 m + k) * nblk + block],
// counted by wave ballots).
__global__ __launch_bounds__(kBlock) void k_dp_select(DPlanArgs A, DPlanMaps M, long n, int nblk,
                                                      uint8_t* __restrict__ cat, int* __restrict__ err,
                                                      int* __restrict__ anypost, int* __restrict__ bcnt) {
  __shared__ int wc[kDpWaves][35];
  uint8_t out[7] = {0, 0, 0, 0, 0, 0, 0};
  int a = 0, d3[3] = {0, 0, 0}, post = 0;
  const long g = dp_cell(A, a, d3[0], d3[1], d3[2]);
  if (g >= 0) {
    const int pid = A.patches[a];
    const DPatch D = A.pt[pid];
    DExpr e[3];
    bool ok[3];
#pragma unroll
    for (int p = 0; p < 3; p++) {
      e[p] = M.ex[p][g];
      ok[p] = (e[p].meta & kDxValid) != 0;
      if (ok[p]) out[p] = dp_cat(e[p]);
    }
    const int outside = n_outside(D, A.nd, d3[0], d3[1], d3[2]);
    // map[5] and map[6]: colour-1 face ghosts of SetGhost(level, 0), map[6]
    // without the copies a correction leaves exact
    if (ok[0] && outside == 1 && ((d3[0] + d3[1] + d3[2]) & 1) == 1) {
      const DExpr& x = e[0];
      if (dx_early(x.meta)) {
        atomicOr(err, kDpSel5);
      } else {
        bool exact = dx_kind(x.meta) == 0 && dx_post(x.meta) == 0;
        if (exact) {
          const DPatch S = A.pt[x.spid];
          int s3[3];
          qdecode(S, x.src, s3[0], s3[1], s3[2]);
          exact = n_outside(S, A.nd, s3[0], s3[1], s3[2]) == 0;
          // (constant bounds with a guard on nd keep s3 and d3 in
          // registers)
#pragma unroll
          for (int b = 0; b < 3; b++) {
            if (b >= A.nd || !exact) continue;
            const int per = M.per[b];
            const int ds = (M.mglo[3 * x.spid + b] + s3[b]) - (M.mglo[3 * pid + b] + d3[b]);
            exact = ((ds % per) + per) % per == 0;
          }
          for (int corner = 0; corner < (1 << A.nd) && exact; corner++) {
            long idx = 0;
#pragma unroll
            for (int b = 2; b >= 0; b--) {
              if (b >= A.nd || !exact) continue;
              const int gf = M.mglo[3 * pid + b] + d3[b] - 1;
              const int lo = (gf & 1) ? (gf + 1) / 2 - 1 : gf / 2;
              int xx = lo + ((corner >> b) & 1) - 1;
              if (xx < 0 || xx >= M.cper[b]) {
                if (!M.cwrap[b]) exact = false;
                xx = ((xx % M.cper[b]) + M.cper[b]) % M.cper[b];
              }
              idx = idx * M.cper[b] + xx;
            }
            exact = exact && ((M.occ[idx >> 5] >> (idx & 31)) & 1u);
          }
        }
        out[5] = out[0];
        out[6] = exact ? 0 : out[0];
      }
    }
    // map[2 + p], p = 1, 2: what a red-black sweep sequence needs (the
    // interpolation rules of the two calls' expressions made once)
    GRule rr[3];
#pragma unroll
    for (int p = 1; p <= 2; p++)
      if (ok[p] && dx_kind(e[p].meta) == 1) {
        if (p == 2 && ok[1] && dx_kind(e[1].meta) == 1 && e[1].dpid == e[2].dpid && e[1].src == e[2].src) rr[2] = rr[1];
        else rr[p] = dp_rule(A, e[p].dpid, e[p].src);
      }
#pragma unroll
    for (int p = 1; p <= 2; p++) {
      if (!ok[p]) continue;
      const DExpr& x = e[p];
      const int col = p - 1;
      const bool have_o = ok[3 - p];
      bool need;
      if (outside != 1) {
        if (!have_o && dx_early(x.meta)) {
          atomicOr(err, kDpLean);
          continue;
        }
        need = !have_o;
      } else {
        need = ((d3[0] + d3[1] + d3[2]) & 1) == col;
        if (!need) {
          int si, sj, sk;
          if (dx_kind(x.meta) == 0) {
            qdecode(A.pt[x.spid], x.src, si, sj, sk);
            need = ((si + sj + sk) & 1) == col;
          } else {
            const GRule& R = rr[p];
            qdecode(D, R.s1, si, sj, sk);
            need = ((si + sj + sk) & 1) == col;
            qdecode(D, R.s2, si, sj, sk);
            need = need || ((si + sj + sk) & 1) == col;
          }
        }
        if (!need) need = !have_o || !dp_same_expr(A, x, e[3 - p], rr[p], rr[3 - p]);
        if (need && dx_early(x.meta)) {
          atomicOr(err, kDpLean);
          continue;
        }
      }
      if (need) out[2 + p] = out[p];
    }
#pragma unroll
    for (int m = 0; m < 7; m++) {
      cat[m * n + g] = out[m];
      if (out[m] && out[m] <= 2 && dx_post(e[dp_call(m)].meta)) post |= 1 << m;
    }
  }
  // the tables with a post-operation copy: one atomic per wave
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int m = 0; m < 7; m++)
    if (__ballot((post >> m) & 1) && lane == 0) atomicOr(anypost, 1 << m);
  // counts per (table, category): one ballot each, summed over the waves
  for (int m = 0; m < 7; m++)
    for (int k = 1; k <= 4; k++) {
      const unsigned long long b = __ballot(out[m] == k);
      if (lane == 0) wc[wave][5 * m + k] = __popcll(b);
    }
  __syncthreads();
  if (threadIdx.x < 35) {
    int s = 0;
    if (threadIdx.x % 5)
      for (int w = 0; w < kDpWaves; w++) s += wc[w][threadIdx.x];
    bcnt[(long)threadIdx.x * nblk + blockIdx.x] = s;
  }
}

// Start of every (table, category) bin in the scanned counts and the total.
__global__ void k_dp_bins(const int* __restrict__ scan, const int* __restrict__ bcnt, int nblk,
                          int* __restrict__ start) {
  const int t = threadIdx.x;
  if (t < 35) start[t] = scan[(long)t * nblk];
  if (t == 35) start[35] = scan[35L * nblk - 1] + bcnt[35L * nblk - 1];
}

struct DpOut {
  int4* cr[7];
  uint16_t* cpost[7];
  GRule* ir[7];
  int nc1[7], ni1[7];
};

// The rules of the seven tables at their compacted positions: the
// workgroup's start in its bin, the waves before it, the lanes before it
// (the same workgroups over the pool-order cells as k_dp_select).
__global__ __launch_bounds__(kBlock) void k_dp_emit(DPlanArgs A, DPlanMaps M, long n, int nblk,
                                                    const uint8_t* __restrict__ cat, const int* __restrict__ scan,
                                                    const int* __restrict__ start, DpOut O) {
  __shared__ int wc[kDpWaves][35];
  int a = 0, d3[3] = {0, 0, 0};
  const long g = dp_cell(A, a, d3[0], d3[1], d3[2]);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned long long below = (1ull << lane) - 1ull;
  uint8_t c[7];
  int rank[7];
#pragma unroll
  for (int m = 0; m < 7; m++) {
    c[m] = g >= 0 ? cat[m * n + g] : 0;
    rank[m] = 0;
    for (int k = 1; k <= 4; k++) {
      const unsigned long long b = __ballot(c[m] == k);
      if (c[m] == k) rank[m] = __popcll(b & below);
      if (lane == 0) wc[wave][5 * m + k] = __popcll(b);
    }
  }
  __syncthreads();
  if (g < 0) return;
  const int pid = A.patches[a];
  const int dst = qidx(A.pt[pid], d3[0], d3[1], d3[2]);
  GRule rule{};
  int rpid = -1, relem = -1;
#pragma unroll
  for (int m = 0; m < 7; m++) {
    const int k = c[m];
    if (!k) continue;
    const int bin = 5 * m + k;
    int at = scan[(long)bin * nblk + blockIdx.x] - start[bin] + rank[m];
    for (int w = 0; w < wave; w++) at += wc[w][bin];
    const DExpr e = M.ex[dp_call(m)][g];
    const int post = dx_post(e.meta);
    if (k <= 2) {
      if (k == 2) at += O.nc1[m];
      O.cr[m][at] = make_int4(dst, e.src, A.pt[pid].qvs, A.pt[e.spid].qvs);
      if (O.cpost[m]) O.cpost[m][at] = (uint16_t)post;
    } else {
      if (k == 4) at += O.ni1[m];
      // the rule of the last element made, if it is this one
      if (e.dpid != rpid || e.src != relem) {
        rule = dp_rule(A, e.dpid, e.src);
        rpid = e.dpid;
        relem = e.src;
      }
      GRule r = rule;
      r.dst = dst;
      r.kind = (r.kind & 0xff) | (post << 8);
      O.ir[m][at] = r;
    }
  }
}

// Cells of a level's patches as bits of the level's index space (the
// occupancy test of map[6]).  Over the kHalf domain, whose elements with
// t = 0 are the patch's rows (j, k) once each: such a thread sets its row's
// bits a word at a time.
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_dp_occupy(LevelArgs L, const int* __restrict__ mglo,
                                                      uint32_t* __restrict__ bits, int ex, int ey) {
  int pi, local;
  if (!level_elem(L, level_block(L), pi, local)) return;
  const int pid = L.patches[pi];
  const DPatch P = L.pt[pid];
  const int cnt = (P.mx[0] + 1) >> 1;
  const int row = local / cnt;
  if (local - row * cnt != 0) return;
  const int j = row % P.mx[1] + 1;
  const int kk = (DIM == 3) ? row / P.mx[1] : 0;
  const long base = ((long)(mglo[3 * pid + 2] - 1 + kk) * ey + (mglo[3 * pid + 1] - 1 + j - 1)) * ex;
  const long b0 = base + (mglo[3 * pid] - 1), b1 = b0 + P.mx[0] - 1;
  for (long w = b0 >> 5; w <= (b1 >> 5); w++) {
    const int lo = w == (b0 >> 5) ? (int)(b0 & 31) : 0, hi = w == (b1 >> 5) ? (int)(b1 & 31) : 31;
    const uint32_t m = (hi == 31 ? 0xffffffffu : ((1u << (hi + 1)) - 1u)) & ~((1u << lo) - 1u);
    atomicOr(&bits[w], m);
  }
}

int32_t plan_maps_device(bsam_ctx* c, Level& L, const GhostPlan& G, std::string& why) {
  L.maps_ok = false;
  const int np = (int)L.patches.size(), nd = c->p.ndims;
  if (L.level <= 0 || np == 0) {
    why = "not a refined level";
    return BSAM_OK;
  }
  if (G.stages.size() > 120) {
    why = "too many ghost-fill stages";
    return BSAM_OK;
  }
  // ghost cells of the level, patch after patch, each patch's from a
  // workgroup boundary; the interpolation elements likewise
  std::vector<int> pos(c->npatch, -1), eoff(np + 1, 0);
  std::vector<long> gbase(np + 1);
  std::vector<int> bfirst(np), bp;
  long n = 0, nel = 0;
  for (int a = 0; a < np; a++) {
    const DPatch& P = c->hp[L.patches[a]];
    if (P.nq[0] & 1) {
      why = "an odd row length (the ghost numbering needs even ones)";
      return BSAM_OK;
    }
    pos[L.patches[a]] = a;
    gbase[a] = n;
    const long cnt = shell_count(P, nd);
    bfirst[a] = (int)bp.size();
    bp.insert(bp.end(), nblocks(cnt), a);
    n += cnt;
    eoff[a] = (int)nel;
    nel += interp_elems(P, nd);
  }
  gbase[np] = n;
  eoff[np] = (int)nel;
  if (n >= (1L << 31) / 16 || nel >= (1L << 28) || np >= (1 << 27)) {
    why = "level too large for the device planner";
    return BSAM_OK;
  }
  // box stages: each box's first flat cell
  std::vector<std::vector<int>> bpre(G.stages.size());
  for (size_t st = 0; st < G.stages.size(); st++) {
    long acc = 0;
    for (const BoxDesc& b : G.stages[st]) {
      bpre[st].push_back((int)acc);
      acc += (long)b.ext[0] * b.ext[1] * b.ext[2];
    }
    bpre[st].push_back((int)acc);
    if (acc >= (1L << 31) || G.stages[st].size() >= (1u << 24)) {
      why = "a ghost-fill stage too large for the device planner";
      return BSAM_OK;
    }
  }
  // the root's periods and the parent level's extents
  const bsam_patch_desc* root = nullptr;
  for (const bsam_patch_desc& d : c->desc)
    if (d.level == 0) { root = &d; break; }
  int per[3] = {1, 1, 1}, cper[3] = {1, 1, 1}, cwrap[3] = {0, 0, 0};
  for (int a = 0; a < nd; a++) {
    per[a] = level_period(c, L.level, a);
    cper[a] = level_period(c, L.level - 1, a);
    cwrap[a] = root->mthbc[2 * a] == 2;
  }
  const long ocells = (long)cper[0] * cper[1] * cper[2];
  if (ocells > (1L << 33)) {
    why = "parent level too large for the occupancy bit map";
    return BSAM_OK;
  }
  const size_t owords = (size_t)((ocells + 31) / 32);
  const int nblk = std::max<int>(1, (int)bp.size());
  // workspace: three expression arrays, two claim arrays, the categories,
  // the per-workgroup counts and their scan, then the small tables
  size_t scan_bytes = 0;
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (int*)nullptr, (int*)nullptr, 35 * nblk, c->s));
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  size_t boxints = 0;
  for (auto& v : bpre) boxints += v.size();
  const size_t o_ex = 0, o_ci = o_ex + al(3 * n * sizeof(DExpr)), o_cb = o_ci + al(3 * n * sizeof(unsigned long long)),
               o_ec = o_cb + al(n * sizeof(int)), o_cat = o_ec + al(nel * sizeof(int)), o_bc = o_cat + al(7 * n),
               o_bs = o_bc + al(35 * (size_t)nblk * sizeof(int)), o_pos = o_bs + al(35 * (size_t)nblk * sizeof(int)),
               o_gb = o_pos + al(c->npatch * sizeof(int)), o_bp = o_gb + al((np + 1) * sizeof(long)),
               o_bf = o_bp + al(bp.size() * sizeof(int)), o_eo = o_bf + al(np * sizeof(int)),
               o_bpre = o_eo + al((np + 1) * sizeof(int)), o_mglo = o_bpre + al(boxints * sizeof(int)),
               o_occ = o_mglo + al(3 * c->npatch * sizeof(int)), o_small = o_occ + al(owords * sizeof(uint32_t)),
               o_scan = o_small + 512, total = o_scan + al(scan_bytes);
  char* W = plan_workspace(c, total);
  if (!W) return set_err(BSAM_ENOMEM, "ghost-table planner workspace of %zu bytes", total);
  DExpr* ex = reinterpret_cast<DExpr*>(W + o_ex);
  auto* claim_i = reinterpret_cast<unsigned long long*>(W + o_ci);
  int* claim_b = reinterpret_cast<int*>(W + o_cb);
  uint8_t* cat = reinterpret_cast<uint8_t*>(W + o_cat);
  int* d_ecell = reinterpret_cast<int*>(W + o_ec);
  int* bcnt = reinterpret_cast<int*>(W + o_bc);
  int* bscan = reinterpret_cast<int*>(W + o_bs);
  int* d_pos = reinterpret_cast<int*>(W + o_pos);
  long* d_gb = reinterpret_cast<long*>(W + o_gb);
  int* d_bp = reinterpret_cast<int*>(W + o_bp);
  int* d_bf = reinterpret_cast<int*>(W + o_bf);
  int* d_eo = reinterpret_cast<int*>(W + o_eo);
  int* d_bpre = reinterpret_cast<int*>(W + o_bpre);
  int* d_mglo = reinterpret_cast<int*>(W + o_mglo);
  uint32_t* d_occ = reinterpret_cast<uint32_t*>(W + o_occ);
  int* d_err = reinterpret_cast<int*>(W + o_small);
  int* d_post = d_err + 1;
  int* d_start = reinterpret_cast<int*>(W + o_small + 64);
  // the small tables in one staging copy; the bit map and the flags start
  // at zero
  {
    std::vector<char> h(o_occ - o_pos, 0);
    auto put = [&](size_t off, const void* src, size_t bytes) { memcpy(h.data() + (off - o_pos), src, bytes); };
    put(o_pos, pos.data(), pos.size() * sizeof(int));
    put(o_gb, gbase.data(), gbase.size() * sizeof(long));
    put(o_bp, bp.data(), bp.size() * sizeof(int));
    put(o_bf, bfirst.data(), bfirst.size() * sizeof(int));
    put(o_eo, eoff.data(), eoff.size() * sizeof(int));
    size_t k = 0;
    for (auto& v : bpre) {
      put(o_bpre + k * sizeof(int), v.data(), v.size() * sizeof(int));
      k += v.size();
    }
    std::vector<int> mg(3 * c->npatch, 1);
    for (int i = 0; i < c->npatch; i++)
      for (int a = 0; a < nd; a++) mg[3 * i + a] = c->desc[i].mglobal[a][0];
    put(o_mglo, mg.data(), mg.size() * sizeof(int));
    HIPCHK(hipMemcpyAsync(W + o_pos, h.data(), h.size(), hipMemcpyHostToDevice, c->s));
    const long zw = (long)((o_scan - o_occ) / sizeof(int));
    k_dp_fill<<<nblocks(zw), kBlock, 0, c->s>>>(reinterpret_cast<int*>(W + o_occ), zw, 0);
    LAUNCH_CHECK();
    HIPCHK(hipStreamSynchronize(c->s));
  }
  DPlanArgs A{c->d_pt, L.d_patches, d_pos, d_gb, d_bp, d_bf, d_eo, np, nd};
  // the parent level's cells
  {
    Level& Pl = c->L(L.level - 1);
    FOR_LAUNCHES(c, Pl, kHalf, a, nb) {
      if (nd == 2) k_dp_occupy<2><<<nb, kBlock, 0, c->s>>>(a, d_mglo, d_occ, cper[0], cper[1]);
      else k_dp_occupy<3><<<nb, kBlock, 0, c->s>>>(a, d_mglo, d_occ, cper[0], cper[1]);
      LAUNCH_CHECK();
    }
  }
  // the three calls' symbolic runs, stage by stage together
  k_dp_fill<<<nblocks(12 * n), kBlock, 0, c->s>>>(reinterpret_cast<int*>(ex), 12 * n, 0);
  k_dp_fill<<<nblocks(6 * n), kBlock, 0, c->s>>>(reinterpret_cast<int*>(claim_i), 6 * n, 0);
  k_dp_fill<<<nblocks(n), kBlock, 0, c->s>>>(claim_b, n, -1);
  LAUNCH_CHECK();
  if (nel > 0) {
    for (int w = 0; w < 2; w++) k_dp_interp<<<nblocks(nel), kBlock, 0, c->s>>>(A, nel, n, claim_i, ex, d_ecell, w);
    LAUNCH_CHECK();
  }
  {
    size_t k = 0;
    for (size_t st = 0; st < G.stages.size(); st++) {
      const int nbox = (int)G.stages[st].size(), ncell = bpre[st].back();
      if (ncell > 0)
        for (int w = 0; w < 2; w++)
          k_dp_box<<<nblocks(ncell), kBlock, 0, c->s>>>(A, L.stages[st].d, d_bpre + k, nbox, ncell,
                                                          (int)(st + 1) << 24, n, claim_b, ex, d_err, w);
      LAUNCH_CHECK();
      k += bpre[st].size();
    }
  }
  k_dp_check<<<nblocks(3 * n), kBlock, 0, c->s>>>(A, ex, n, d_err);
  LAUNCH_CHECK();
  DPlanMaps M{};
  for (int p = 0; p < 3; p++) M.ex[p] = ex + (size_t)p * n;
  M.mglo = d_mglo;
  M.occ = d_occ;
  for (int a = 0; a < 3; a++) M.per[a] = per[a], M.cper[a] = cper[a], M.cwrap[a] = cwrap[a];
  k_dp_select<<<nblk, kBlock, 0, c->s>>>(A, M, n, nblk, cat, d_err, d_post, bcnt);
  LAUNCH_CHECK();
  size_t sb = scan_bytes;
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(W + o_scan, sb, bcnt, bscan, 35 * nblk, c->s));
  k_dp_bins<<<1, 64, 0, c->s>>>(bscan, bcnt, nblk, d_start);
  LAUNCH_CHECK();
  struct {
    int err, post;
    int pad[14];
    int start[36];
  } h{};
  HIPCHK(hipMemcpyAsync(&h, W + o_small, 64 + 36 * sizeof(int), hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  if (h.err) {
    static const char* what[8] = {"a ghost-fill stage writes an interior cell",
                                  "a box reads a cell its own stage writes",
                                  "an interpolated ghost value is copied into another patch",
                                  "post-operation chain longer than five",
                                  "a ghost value is read from an edge or corner ghost the call has not written",
                                  "a ghost read from before the call is rewritten by another such rule",
                                  "a colour-1 face ghost of SetGhost(level, 0) reads a ghost from before the call",
                                  "a face ghost of a reduced table reads a ghost from before the call"};
    for (int b = 0; b < 8; b++)
      if (h.err & (1 << b)) {
        why = what[b];
        break;
      }
    return BSAM_OK;
  }
  // the seven tables in one pooled block
  size_t tbytes = 0;
  size_t at[7][3];
  DpOut O{};
  for (int m = 0; m < 7; m++) {
    DevMap& D = L.gmap[m];
    auto cnt = [&](int k) { return h.start[5 * m + k + 1] - h.start[5 * m + k]; };
    D.nc1 = cnt(1);
    D.nc = cnt(1) + cnt(2);
    D.ni1 = cnt(3);
    D.ni = cnt(3) + cnt(4);
    at[m][0] = tbytes;
    tbytes = al(tbytes + (size_t)D.nc * sizeof(int4));
    at[m][1] = tbytes;
    if (h.post & (1 << m)) tbytes = al(tbytes + (size_t)D.nc * sizeof(uint16_t));
    at[m][2] = tbytes;
    tbytes = al(tbytes + (size_t)D.ni * sizeof(GRule));
  }
  char* T = nullptr;
  if (pool_alloc(c, &T, std::max<size_t>(tbytes, 1), 1) != hipSuccess)
    return set_err(BSAM_ENOMEM, "ghost tables of %zu bytes", tbytes);
  L.blocks.push_back(T);
  for (int m = 0; m < 7; m++) {
    DevMap& D = L.gmap[m];
    D.cr = D.nc ? reinterpret_cast<int4*>(T + at[m][0]) : nullptr;
    D.cpost = (D.nc && (h.post & (1 << m))) ? reinterpret_cast<uint16_t*>(T + at[m][1]) : nullptr;
    D.ir = D.ni ? reinterpret_cast<GRule*>(T + at[m][2]) : nullptr;
    O.cr[m] = D.cr;
    O.cpost[m] = D.cpost;
    O.ir[m] = D.ir;
    O.nc1[m] = D.nc1;
    O.ni1[m] = D.ni1;
  }
  k_dp_emit<<<nblk, kBlock, 0, c->s>>>(A, M, n, nblk, cat, bscan, d_start, O);
  LAUNCH_CHECK();
  HIPCHK(hipStreamSynchronize(c->s));
  L.maps_ok = true;
  return BSAM_OK;
}

}  // namespace

namespace {
// Coarse chains: from the highest level with at most c->chain_cells cells
// down to minlevel, AFASVcycle runs in one workgroup.  CH: every level of
// the chain a single patch periodic in every dimension (k_coarse_vcycle);
// the tumour model: levels of one patch each whose fill is a short
// sequence of box stages that one workgroup runs in order
// (k_coarse_vcycle_tm).  A chain has at least two levels.
int32_t plan_chain(bsam_ctx* c) {
  for (Level& L : c->lv) L.chain = 0;
  int top = c->minlevel - 1;
  if (c->p.model == BSAM_MODEL_CH) {
    for (int l = c->minlevel; l <= c->finest; l++) {
      const Level& L = c->L(l);
      if (L.wrap_pid < 0) break;
      const DPatch& P = c->hp[L.wrap_pid];
      if ((long)P.mx[0] * P.mx[1] * P.mx[2] > c->chain_cells) break;
      top = l;
    }
    if (top >= c->minlevel && top - c->minlevel + 1 >= 2 && top - c->minlevel + 1 <= 16)
      c->L(top).chain = top - c->minlevel + 1;
  } else if (c->p.model == BSAM_MODEL_TUMOUR) {
    for (int l = c->minlevel; l <= c->finest && l <= 0; l++) {
      const Level& L = c->L(l);
      if (L.patches.size() != 1 || L.wrap_pid >= 0 || L.d_rules || L.stages.empty() ||
          (int)L.stages.size() > kMaxStages)
        break;
      int maxel = 0;
      for (const Stage& st : L.stages) maxel = std::max(maxel, st.maxel);
      const DPatch& P = c->hp[L.patches[0]];
      if (maxel * c->p.nrvars > kSeqBoxElems || (long)P.mx[0] * P.mx[1] * P.mx[2] > c->chain_cells) break;
      top = l;
    }
    if (top >= c->minlevel && top - c->minlevel + 1 >= 2 && top - c->minlevel + 1 <= kMaxChain) {
      Level& T = c->L(top);
      T.chain = top - c->minlevel + 1;
      // the chain's levels for the kernel, top first
      TmChainArgs& A = c->h_tmchain;
      A = TmChainArgs{};
      A.nlev = T.chain;
      for (int n = 0; n < T.chain; n++) {
        const Level& Ln = c->L(top - n);
        A.pid[n] = Ln.patches[0];
        A.seq[n].ns = (int)Ln.stages.size();
        for (int st = 0; st < A.seq[n].ns; st++) {
          A.seq[n].d[st] = Ln.stages[st].d;
          A.seq[n].n[st] = Ln.stages[st].n;
        }
        // one patch, condition 10 on every face: its stages are the faces
        // x, y, z, which phys_shell evaluates in one pass
        bool allphys = Ln.wrap_pid < 0 && (int)Ln.stages.size() == c->p.ndims;
        for (int f = 0; f < 2 * c->p.ndims; f++) allphys = allphys && c->desc[A.pid[n]].mthbc[f] == 10;
        A.physop[n] = allphys ? phys_face_op(c) : -1;
      }
      if (!c->d_tmchain) HIPCHK(hipMalloc(&c->d_tmchain, sizeof(TmChainArgs)));
      // after every earlier chain launch, and complete on return
      HIPCHK(hipMemcpyAsync(c->d_tmchain, &c->h_tmchain, sizeof(TmChainArgs), hipMemcpyHostToDevice, c->s));
      HIPCHK(hipStreamSynchronize(c->s));
    }
  }
  return BSAM_OK;
}

// The host half of bsam_set_hierarchy: descriptor validation, the device
// patch records with their pool offsets, and each level's patch list.
int32_t plan_patches(bsam_ctx* c, int32_t npatch, const bsam_patch_desc* desc, int32_t minlevel,
                     int32_t finestlevel) {
  if (!desc || npatch < 1) return set_err(BSAM_EINVAL, "bad arguments");
  if (minlevel > 0 || finestlevel < 0 || minlevel < -30 || finestlevel > 30)
    return set_err(BSAM_EINVAL, "level range [%d, %d] invalid", minlevel, finestlevel);
  const int nd = c->p.ndims, nv = c->p.nrvars;
  c->desc.assign(desc, desc + npatch);
  c->npatch = npatch;
  c->minlevel = minlevel;
  c->finest = finestlevel;
  c->hp.resize(npatch);
  long qo = 0, co = 0;
  int nroot = 0;
  for (int n = 0; n < npatch; n++) {
    const bsam_patch_desc& d = desc[n];
    DPatch& P = c->hp[n];
    memset(&P, 0, sizeof P);
    if (d.level < minlevel || d.level > finestlevel)
      return set_err(BSAM_EINVAL, "patch %d: level %d outside [%d, %d]", n, d.level, minlevel, finestlevel);
    if (d.level == 0) nroot++;
    if (d.level > minlevel) {
      if (d.parent < 0 || d.parent >= npatch || desc[d.parent].level != d.level - 1)
        return set_err(BSAM_EINVAL, "patch %d: parent %d is not on level %d", n, d.parent, d.level - 1);
    }
    for (int a = 0; a < 3; a++) {
      const bool used = a < nd;
      P.mx[a] = used ? d.mx[a] : 1;
      if (used && P.mx[a] < 2) return set_err(BSAM_EINVAL, "patch %d: mx[%d]=%d < 2", n, a, P.mx[a]);
      P.cmx[a] = used ? ((P.mx[a] % 2) ? 1 : P.mx[a] / 2) : 1;
      P.nq[a] = used ? P.mx[a] + 2 : 1;
      P.nc[a] = used ? P.cmx[a] + 2 : 1;
      P.mb[a][0] = used ? d.mbounds[a][0] : 0;
      P.mb[a][1] = used ? d.mbounds[a][1] : 0;
    }
    if (d.level > minlevel) {
      for (int a = 0; a < nd; a++) {
        if (P.mx[a] % 2) return set_err(BSAM_EINVAL, "patch %d: odd mx on a refined level", n);
        if (d.mbounds[a][1] - d.mbounds[a][0] + 1 != P.mx[a] / 2)
          return set_err(BSAM_EINVAL, "patch %d: mbounds do not span mx/2 in dim %d", n, a);
        const int pmx = desc[d.parent].mx[a];
        if (d.mbounds[a][0] < 1 || d.mbounds[a][1] > pmx)
          return set_err(BSAM_EINVAL, "patch %d: mbounds outside the parent in dim %d", n, a);
      }
    }
    for (int f = 0; f < 6; f++) P.mthbc[f] = d.mthbc[f];
    P.hq = (P.nq[0] + 1) / 2;
    P.hc = (P.nc[0] + 1) / 2;
    P.parent = d.level > minlevel ? d.parent : -1;
    P.level = d.level;
    P.dx = d.dx;
    P.qvs = 2 * P.nq[2] * P.nq[1] * P.hq;
    P.cvs = 2 * P.nc[2] * P.nc[1] * P.hc;
    P.qoff = (int)qo;
    P.coff = (int)co;
    qo += (long)P.qvs * nv;
    co += (long)P.cvs * nv;
    if (qo > 0x7fffffffL || co > 0x7fffffffL)
      return set_err(BSAM_EINVAL, "hierarchy too large for 32-bit offsets (%ld elements)", qo);
  }
  if (nroot != 1) return set_err(BSAM_EINVAL, "exactly one root-level patch required (found %d)", nroot);
  c->nq = qo;
  c->nc = co;
  c->lv.assign(finestlevel - minlevel + 1, Level());
  for (int l = minlevel; l <= finestlevel; l++) c->L(l).level = l;
  for (int n = 0; n < npatch; n++) c->L(desc[n].level).patches.push_back(n);
  for (Level& L : c->lv) {
    if (L.patches.empty()) return set_err(BSAM_EINVAL, "level %d has no patches", L.level);
    L.h = c->hp[L.patches[0]].dx;
  }
  return BSAM_OK;
}
}  // namespace

namespace {
// The level's interpolation rules on the device (k_interp_segment,
// DeviceSelect per segment, k_interp_write).  Only the segment bounds of
// the whole level are kept (roff[s][0], roff[s][np]): the batched staged
// fill needs no others.
int32_t build_rules_device(bsam_ctx* c, Level& L) {
  const int np = (int)L.patches.size(), nd = c->p.ndims;
  std::vector<int> eoff(np + 1, 0);
  long total = 0;
  for (int a = 0; a < np; a++) {
    eoff[a] = (int)total;
    total += interp_elems(c->hp[L.patches[a]], nd);
  }
  if (total > 0x7fffffffL / 3) return set_err(BSAM_EINVAL, "level %d: too many interpolation elements", L.level);
  eoff[np] = (int)total;
  for (int sg = 0; sg < 3; sg++) L.roff[sg].assign(np + 1, 0);
  L.d_rules = nullptr;
  if (total == 0) return BSAM_OK;
  // workspace: element offsets, segments, three index lists, counts, scan
  size_t bytes = 0;
  hipcub::CountingInputIterator<int> elems(0);
  HIPCHK(hipcub::DeviceSelect::If(nullptr, bytes, elems, (int*)nullptr, (int*)nullptr, (int)total,
                                  SegmentIs{nullptr, 0}, c->s));
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t o_eoff = 0, o_seg = al((np + 1) * sizeof(int)), o_idx = o_seg + al(total),
               o_num = o_idx + al(3 * total * sizeof(int)), o_scr = o_num + 256, need = o_scr + al(bytes);
  char* W = plan_workspace(c, need);
  if (!W) return set_err(BSAM_ENOMEM, "interpolation-rule workspace of %zu bytes", need);
  int* t_eoff = reinterpret_cast<int*>(W + o_eoff);
  uint8_t* t_seg = reinterpret_cast<uint8_t*>(W + o_seg);
  int* t_idx = reinterpret_cast<int*>(W + o_idx);
  int* t_num = reinterpret_cast<int*>(W + o_num);
  HIPCHK(hipMemcpyAsync(t_eoff, eoff.data(), (np + 1) * sizeof(int), hipMemcpyHostToDevice, c->s));
  k_interp_segment<<<nblocks(total), kBlock, 0, c->s>>>(c->d_pt, L.d_patches, t_eoff, np, nd, total, t_seg);
  LAUNCH_CHECK();
  // each segment's elements in order, into its own list; one wait for the
  // three counts
  for (int sg = 0; sg < 3; sg++) {
    size_t b2 = bytes;
    HIPCHK(hipcub::DeviceSelect::If(W + o_scr, b2, elems, t_idx + (size_t)sg * total, t_num + sg, (int)total,
                                    SegmentIs{t_seg, (uint8_t)sg}, c->s));
  }
  int n[3] = {0, 0, 0};
  HIPCHK(hipMemcpyAsync(n, t_num, sizeof(n), hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  int off = 0;
  for (int sg = 0; sg < 3; sg++) {
    L.roff[sg][0] = off;
    off += n[sg];
    L.roff[sg][np] = off;
  }
  if (off > 0) {
    if (pool_alloc(c, &L.d_rules, (size_t)off * sizeof(GRule), 1) != hipSuccess)
      return set_err(BSAM_ENOMEM, "interpolation rules of level %d", L.level);
    L.blocks.push_back(L.d_rules);
    for (int sg = 0; sg < 3; sg++)
      if (n[sg] > 0) {
        k_interp_write<<<nblocks(n[sg]), kBlock, 0, c->s>>>(c->d_pt, L.d_patches, t_eoff, np, nd,
                                                            t_idx + (size_t)sg * total, n[sg],
                                                            L.d_rules + L.roff[sg][0]);
        LAUNCH_CHECK();
      }
  }
  HIPCHK(hipStreamSynchronize(c->s));
  return BSAM_OK;
}
// The device side of one level of a new hierarchy: its patch list, the
// block-aligned launch ranges of every domain, the tagging offsets and the
// ghost fill (stages, interpolation rules, flattened tables).
// The ghost fill of a level whose box stages are planned (G.stages): the
// interpolation rules, on the host or on the device, the flattened tables
// (host planner, or device planner for batched levels above the root), and
// their upload in a pooled block of the level.
int32_t build_images(bsam_ctx* c, Level& L);
int32_t finish_level_ghosts(bsam_ctx* c, Level& L, GhostPlan& G, bool tables, bool images = true) {
  // batched levels above the root build their interpolation rules, and
  // (with the device planner) their flattened tables, on the device
  const bool dev_tables = tables && c->p.batching && L.level > 0 && c->table_planner == 1;
  const bool dev_rules = (!tables || dev_tables) && c->p.batching && L.level > 0;
  // (with the device planner the interpolation rules are needed only by a
  // level it refuses, which keeps the staged fill)
  if (dev_rules) {
    for (int sg = 0; sg < 3; sg++) G.roff[sg].assign(L.patches.size() + 1, 0);
  } else {
    plan_interp_rules(c, L, G);
  }
  if (c->p.batching && tables && !dev_tables) RETCHK(plan_ghost_maps(c, L, G));
  BlockBuilder Bk;
  upload_ghost(c, L, G, Bk);
  RETCHK(Bk.upload(c, L));
  L.why = G.why;
  if (dev_tables) RETCHK(plan_maps_device(c, L, G, L.why));
  if (dev_rules && !L.maps_ok) RETCHK(build_rules_device(c, L));
  if (images) RETCHK(build_images(c, L));
  L.ghosts_deferred = false;
  return BSAM_OK;
}

// ---------------------------------------------------------------------------
// Ghost images written by the colour sweeps.  The reduced fill after the
// sweep of colour col (gmap[3 + col]) copies, among others, cells the sweep
// has just updated into face ghosts of the same colour in sibling and
// periodic partners.  The sweep's workgroup has those values in hand, so it
// writes them itself (write_images), and the gather keeps the interpolation
// rules and the other copies.  A rule moves to the sweep when its source is
// an interior cell of colour col and its destination a ghost of colour col:
// the sweep of colour col reads no cell of its own colour but its own, so
// no workgroup reads a ghost another one writes, and the value written is
// the one the gather would read.  Maps with rules that read ghost values
// from before the fill are left whole.
// ---------------------------------------------------------------------------
// position of the level patch holding pool offset x (patches in pool order)
__device__ __forceinline__ int img_patch(const DPatch* __restrict__ pt, const int* __restrict__ patches, int np,
                                         int x) {
  int lo = 0, hi = np;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (pt[patches[mid]].qoff <= x) lo = mid;
    else hi = mid;
  }
  return lo;
}
// colour and storage indices of pool offset x of patch P (qidx inverted)
__device__ __forceinline__ bool img_decode(const DPatch& P, int x, int& col, int& i, int& j, int& k) {
  const int rel = x - P.qoff;
  if (rel < 0 || rel >= P.qvs) return false;
  const int row = P.nq[1] * P.hq, blk = P.nq[2] * row;
  col = rel / blk;
  const int r1 = rel - col * blk;
  k = r1 / row;
  const int r2 = r1 - k * row;
  j = r2 / P.hq;
  const int m = r2 - j * P.hq;
  i = 2 * m + ((col + j + k) & 1);
  return true;
}
// per copy rule r: key[r] = the logical workgroup of the sweep (domain
// kHalf) whose cell is its source and tid[r] the thread, or key -1
__global__ __launch_bounds__(kBlock) void k_img_class(const DPatch* __restrict__ pt, const int* __restrict__ patches,
                                                      int np, const int* __restrict__ hprefix,
                                                      const int4* __restrict__ cr, int nc, int col, int nd,
                                                      int* __restrict__ key, int* __restrict__ tid,
                                                      int* __restrict__ bcnt) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= nc) return;
  const int4 x = cr[r];
  int b = -1;
  const int a = img_patch(pt, patches, np, x.y);
  const DPatch P = pt[patches[a]];
  const DPatch D = pt[patches[img_patch(pt, patches, np, x.x)]];
  int cs, i, j, k, cd, di, dj, dk;
  if (img_decode(P, x.y, cs, i, j, k) && img_decode(D, x.x, cd, di, dj, dk) && cs == col && cd == col &&
      i >= 1 && i <= P.mx[0] && j >= 1 && j <= P.mx[1] && (nd == 2 || (k >= 1 && k <= P.mx[2]))) {
    const int cnt = (P.mx[0] + 1) >> 1;
    const int row = (nd == 3) ? (k - 1) * P.mx[1] + (j - 1) : (j - 1);
    const int t = (i + ((col + j + k) & 1)) / 2 - 1;
    const long g = (long)hprefix[a] + (long)row * cnt + t;
    b = (int)(g / kBlock);
    tid[r] = (int)(g - (long)b * kBlock);
    atomicAdd(&bcnt[b], 1);
  }
  key[r] = b;
}
__global__ __launch_bounds__(kBlock) void k_img_write(const int4* __restrict__ cr, const uint16_t* __restrict__ cpost,
                                                      int nc, const int* __restrict__ key, const int* __restrict__ tid,
                                                      const int* __restrict__ boff, int* __restrict__ fill,
                                                      const int* __restrict__ kpos, int4* __restrict__ ent,
                                                      int4* __restrict__ gcr, uint16_t* __restrict__ gpost) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= nc) return;
  const int4 x = cr[r];
  const int post = cpost ? cpost[r] : 0;
  const int b = key[r];
  if (b >= 0) {
    ent[boff[b] + atomicAdd(&fill[b], 1)] = make_int4(x.x, x.z, tid[r], post);
  } else {
    const int at = kpos[r];
    gcr[at] = x;
    if (gpost) gpost[at] = (uint16_t)post;
  }
}
__global__ void k_img_keep(const int* __restrict__ key, int n, int* __restrict__ keep) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r < n) keep[r] = key[r] < 0 ? 1 : 0;
  else if (r == n) keep[r] = 0;
}

int32_t build_images(bsam_ctx* c, Level& L) {
  L.img_ok[0] = L.img_ok[1] = false;
  if (!c->sweep_images || !L.maps_ok || !c->p.batching || c->p.model != BSAM_MODEL_CH || L.wrap_pid >= 0 ||
      L.patches.empty())
    return BSAM_OK;
  const int np = (int)L.patches.size(), nd = c->p.ndims;
  const int nblk = nblocks(L.total[kHalf]);
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  for (int col = 0; col < 2; col++) {
    const DevMap& M = L.gmap[3 + col];
    if (M.nc1 || M.ni1 || M.nc == 0) continue;
    const int nc = M.nc;
    size_t scan_a = 0, scan_b = 0;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_a, (int*)nullptr, (int*)nullptr, nblk + 1, c->s));
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_b, (int*)nullptr, (int*)nullptr, nc + 1, c->s));
    const size_t o_key = 0, o_tid = o_key + al((size_t)nc * 4), o_keep = o_tid + al((size_t)nc * 4),
                 o_kpos = o_keep + al((size_t)(nc + 1) * 4), o_bcnt = o_kpos + al((size_t)(nc + 1) * 4),
                 o_boff = o_bcnt + al((size_t)(nblk + 1) * 4), o_fill = o_boff + al((size_t)(nblk + 1) * 4),
                 o_scan = o_fill + al((size_t)nblk * 4), total = o_scan + al(std::max(scan_a, scan_b));
    char* W = plan_workspace(c, total);
    if (!W) return set_err(BSAM_ENOMEM, "sweep-image workspace of %zu bytes", total);
    int* key = reinterpret_cast<int*>(W + o_key);
    int* tid = reinterpret_cast<int*>(W + o_tid);
    int* keep = reinterpret_cast<int*>(W + o_keep);
    int* kpos = reinterpret_cast<int*>(W + o_kpos);
    int* bcnt = reinterpret_cast<int*>(W + o_bcnt);
    int* boff = reinterpret_cast<int*>(W + o_boff);
    int* fill = reinterpret_cast<int*>(W + o_fill);
    k_dp_fill<<<nblocks(nblk + 1), kBlock, 0, c->s>>>(bcnt, nblk + 1, 0);
    k_dp_fill<<<nblocks(nblk), kBlock, 0, c->s>>>(fill, nblk, 0);
    k_img_class<<<nblocks(nc), kBlock, 0, c->s>>>(c->d_pt, L.d_patches, np, L.d_prefix[kHalf], M.cr, nc, col, nd,
                                                  key, tid, bcnt);
    k_img_keep<<<nblocks(nc + 1), kBlock, 0, c->s>>>(key, nc, keep);
    LAUNCH_CHECK();
    size_t sa = scan_a, sb = scan_b;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(W + o_scan, sa, bcnt, boff, nblk + 1, c->s));
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(W + o_scan, sb, keep, kpos, nc + 1, c->s));
    int n[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(&n[0], boff + nblk, sizeof(int), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipMemcpyAsync(&n[1], kpos + nc, sizeof(int), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    const int nimg = n[0], nkeep = n[1];
    if (nimg == 0 || nimg + nkeep != nc) continue;
    // one pooled block: workgroup offsets, images, the gather's copies
    const size_t b_off = 0, b_ent = al((size_t)(nblk + 1) * 4), b_cr = b_ent + al((size_t)nimg * sizeof(int4)),
                 b_post = b_cr + al((size_t)std::max(nkeep, 1) * sizeof(int4)),
                 bytes = b_post + (M.cpost ? al((size_t)std::max(nkeep, 1) * 2) : 0);
    char* B = nullptr;
    if (pool_alloc(c, &B, bytes, 1) != hipSuccess) return set_err(BSAM_ENOMEM, "sweep images of %zu bytes", bytes);
    L.blocks.push_back(B);
    int* off = reinterpret_cast<int*>(B + b_off);
    int4* ent = reinterpret_cast<int4*>(B + b_ent);
    int4* gcr = reinterpret_cast<int4*>(B + b_cr);
    uint16_t* gpost = M.cpost ? reinterpret_cast<uint16_t*>(B + b_post) : nullptr;
    HIPCHK(hipMemcpyAsync(off, boff, (size_t)(nblk + 1) * sizeof(int), hipMemcpyDeviceToDevice, c->s));
    k_img_write<<<nblocks(nc), kBlock, 0, c->s>>>(M.cr, M.cpost, nc, key, tid, boff, fill, kpos, ent, gcr, gpost);
    LAUNCH_CHECK();
    HIPCHK(hipStreamSynchronize(c->s));
    DevMap G;
    G.cr = nkeep ? gcr : nullptr;
    G.cpost = nkeep ? gpost : nullptr;
    G.ir = M.ir;
    G.nc = nkeep;
    G.ni = M.ni;
    L.gimg[col] = G;
    L.img_off[col] = off;
    L.img_ent[col] = ent;
    L.img_n[col] = nimg;
    L.img_ok[col] = true;
  }
  return BSAM_OK;
}

// A level's device side: the launch domains, the tagging offsets and, unless
// `defer_ghosts`, its ghost fill (plan_ghost_stages, finish_level_ghosts).
int32_t build_level(bsam_ctx* c, Level& L, bool tables = true, bool defer_ghosts = false) {
  const int nv = c->p.nrvars;
  BlockBuilder Bk;
  Bk.add(L.patches.data(), L.patches.size(), &L.d_patches);
  for (int d = 0; d < kNumDom; d++) {
    std::vector<int> pre(L.patches.size() + 1, 0), cnts(L.patches.size(), 0);
    long acc = 0;
    for (size_t a = 0; a < L.patches.size(); a++) {
      const DPatch& P = c->hp[L.patches[a]];
      long cnt;
      switch (d) {
        case kHalf: cnt = (long)((P.mx[0] + 1) / 2) * P.mx[1] * P.mx[2]; break;
        case kCell: cnt = 2L * ((P.mx[0] + 1) / 2) * P.mx[1] * P.mx[2]; break;
        case kCoarse: cnt = (long)P.cmx[0] * P.cmx[1] * P.cmx[2]; break;
        case kFineG: cnt = (long)P.nq[0] * P.nq[1] * P.nq[2]; break;
        case kDual:
          cnt = (long)(P.cmx[0] + 1) * (P.cmx[1] + 1) * (c->p.ndims == 3 ? P.cmx[2] + 1 : 1);
          break;
        default: cnt = (long)P.nc[0] * P.nc[1] * P.nc[2]; break;
      }
      // every patch's range starts on a block boundary
      pre[a] = (int)acc;
      cnts[a] = (int)cnt;
      acc += (cnt + kBlock - 1) / kBlock * kBlock;
    }
    pre[L.patches.size()] = (int)acc;
    if (acc > 0x7fffffffL) return set_err(BSAM_EINVAL, "level %d too large", L.level);
    L.total[d] = (int)acc;
    L.hprefix[d] = pre;
    L.hcnt[d] = cnts;
    Bk.add(pre.data(), pre.size(), &L.d_prefix[d]);
    Bk.add(cnts.data(), cnts.size(), &L.d_cnt[d]);
    // block -> patch, for whole-level launches
    std::vector<int> bp(std::max(1, nblocks(acc)), 0);
    size_t pi = 0;
    for (size_t b = 0; b < bp.size(); b++) {
      const long g0 = (long)b * kBlock;
      while (pi + 1 < L.patches.size() && pre[pi + 1] <= g0) pi++;
      bp[b] = (int)pi;
    }
    Bk.add(bp.data(), bp.size(), &L.d_bp[d]);
  }
  // contiguous span of the level in the q pools (patches are laid out in
  // descriptor order; levels are spans only if their patches are
  // consecutive -- otherwise copies go patch by patch)
  L.qspan_off = c->hp[L.patches[0]].qoff;
  long end = L.qspan_off;
  bool contiguous = true;
  for (int pid : L.patches) {
    if (c->hp[pid].qoff != end) contiguous = false;
    end = c->hp[pid].qoff + (long)c->hp[pid].qvs * nv;
  }
  L.qspan_len = contiguous ? end - L.qspan_off : -1;
  // dense cell offsets of the level's patches (the tagging flags)
  {
    std::vector<long> foff(L.patches.size());
    long n = 0;
    for (size_t a = 0; a < L.patches.size(); a++) {
      foff[a] = n;
      const DPatch& P = c->hp[L.patches[a]];
      n += (long)P.mx[0] * P.mx[1] * P.mx[2];
    }
    L.ncell = n;
    Bk.add(foff.data(), foff.size(), &L.d_foff);
  }
  RETCHK(Bk.upload(c, L));
  if (defer_ghosts) {
    L.ghosts_deferred = true;
    return BSAM_OK;
  }
  GhostPlan G;
  RETCHK(plan_ghost_stages(c, L, G));
  return finish_level_ghosts(c, L, G, tables);
}

bool same_desc(const bsam_patch_desc& a, const bsam_patch_desc& b) {
  if (a.level != b.level || a.parent != b.parent || a.dx != b.dx) return false;
  for (int i = 0; i < 3; i++) {
    if (a.mx[i] != b.mx[i] || a.xlower[i] != b.xlower[i]) return false;
    for (int e = 0; e < 2; e++)
      if (a.mglobal[i][e] != b.mglobal[i][e] || a.mbounds[i][e] != b.mbounds[i][e]) return false;
  }
  for (int f = 0; f < 6; f++)
    if (a.mthbc[f] != b.mthbc[f]) return false;
  return true;
}

// The highest level up to which a new descriptor table equals an older
// table `tab`, whose levels minlevel.. hold the patches plist[l - minlevel]
// (minlevel - 1 if none): the older table's patches of the levels up to it
// are its first n rows, the new table starts with the same n rows, and no
// later row of the new table is on those levels.  Such a level's device
// side depends on no higher level (its ghost fill reads its parent level,
// and the transfers between levels belong to the finer one), and its pool
// offsets are the same in both tables.
int kept_levels(const std::vector<std::vector<int>>& plist, int tmin, const std::vector<bsam_patch_desc>& tab,
                int32_t npatch, const bsam_patch_desc* desc, int32_t minlevel, int32_t finestlevel) {
  int keep = minlevel - 1;
  if (tmin != minlevel || !desc) return keep;
  const int tfin = tmin + (int)plist.size() - 1;
  int n = 0;
  for (int l = minlevel; l <= std::min(tfin, finestlevel); l++) {
    const std::vector<int>& ps = plist[l - tmin];
    for (size_t a = 0; a < ps.size(); a++) {
      const int i = n + (int)a;
      if (ps[a] != i || i >= npatch || !same_desc(desc[i], tab[i])) return keep;
    }
    n += (int)ps.size();
    for (int i = n; i < npatch; i++)
      if (desc[i].level <= l) return keep;
    keep = l;
  }
  return keep;
}

// The mesh a regrid replaces: its q pools (the source of TransferValues),
// descriptors, each level's patches and launch shapes, and each level's
// device side, which an install of an equal level takes over.  Taking it
// releases the rest of the old hierarchy (the context then has none).
struct OldMesh {
  bsam_ctx* c;
  double *q = nullptr, *qold = nullptr;
  std::vector<DPatch> hp;
  std::vector<bsam_patch_desc> desc;
  std::vector<Level> lv;                 // device side, empty once taken
  std::vector<std::vector<int>> plist;   // patches per level
  std::vector<std::array<int, 8>> shape;  // launch shapes per level
  int minlevel = 0, finest = -1;
  XferBox* d_box = nullptr;
  size_t box_cap = 0;
  explicit OldMesh(bsam_ctx* c_) : c(c_) {
    q = c->q;
    qold = c->qold;
    c->q = c->qold = nullptr;
    hp = c->hp;
    desc = c->desc;
    minlevel = c->minlevel;
    finest = c->finest;
    for (int l = minlevel; l <= finest; l++) {
      Level& L = c->L(l);
      plist.push_back(L.patches);
      shape.push_back({L.order, L.gorder[0], L.gorder[1], L.kc_sweep[0], L.kc_sweep[1], L.kc_resid, L.gblock[0],
                       L.gblock[1]});
      lv.push_back(std::move(L));
      L = Level();
    }
    free_hierarchy(c);
  }
  ~OldMesh() {
    hipStreamSynchronize(c->s);
    for (Level& L : lv) free_level(c, L);
    pool_release(c, q);
    pool_release(c, qold);
    hipFree(d_box);
  }
  const std::vector<int>* patches(int l) const {
    return (l >= minlevel && l <= finest) ? &plist[l - minlevel] : nullptr;
  }
  // the launch shapes set for a level carry over while the level exists
  void restore_shapes(int l) const {
    if (l < minlevel || l > finest || l < c->minlevel || l > c->finest) return;
    const std::array<int, 8>& k = shape[l - minlevel];
    Level& L = c->L(l);
    L.order = k[0];
    L.gorder[0] = k[1], L.gorder[1] = k[2];
    L.kc_sweep[0] = k[3], L.kc_sweep[1] = k[4];
    L.kc_resid = k[5];
    L.gblock[0] = k[6], L.gblock[1] = k[7];
  }
};

// Install a hierarchy with zero fields.  The levels kept_levels() finds
// unchanged keep their device side (and launch shapes); the others are
// built.  Returns the highest kept level through *kept.
int32_t install(bsam_ctx* c, int32_t npatch, const bsam_patch_desc* desc, int32_t minlevel, int32_t finestlevel,
                int* kept, OldMesh* donor = nullptr, bool tables = true, int defer_level = -1000) {
  HIPCHK(hipStreamSynchronize(c->s));
  int keep = minlevel - 1;
  if (c->have_hier && desc && npatch > 0) {
    std::vector<std::vector<int>> cur;
    for (const Level& L : c->lv) cur.push_back(L.patches);
    keep = kept_levels(cur, c->minlevel, c->desc, npatch, desc, minlevel, finestlevel);
  }
  // past those, levels equal to the replaced mesh's take its device side
  const int dkeep = donor && desc && npatch > 0
                        ? kept_levels(donor->plist, donor->minlevel, donor->desc, npatch, desc, minlevel, finestlevel)
                        : minlevel - 1;
  if (kept) *kept = keep;
  struct Held {
    bsam_ctx* c;
    std::vector<Level> v;
    ~Held() {
      for (Level& L : v) free_level(c, L);
    }
  } held{c, {}};
  for (int l = minlevel; l <= keep; l++) {
    held.v.push_back(std::move(c->L(l)));
    c->L(l) = Level();
  }
  free_hierarchy(c);
  RETCHK(plan_patches(c, npatch, desc, minlevel, finestlevel));
  const long qo = c->nq, co = c->nc;
  if (pool_alloc(c, &c->d_pt, npatch * sizeof(DPatch), 1) != hipSuccess) return set_err(BSAM_ENOMEM, "patch table");
  HIPCHK(hipMemcpyAsync(c->d_pt, c->hp.data(), npatch * sizeof(DPatch), hipMemcpyHostToDevice, c->s));
  double** pools[5] = {&c->q, &c->qold, &c->f, &c->rf, &c->ftmp};
  for (double** pp : pools) {
    if (pool_alloc(c, pp, qo * sizeof(double)) != hipSuccess)
      return set_err(BSAM_ENOMEM, "device pool of %ld doubles", qo);
    HIPCHK(hipMemsetAsync(*pp, 0, qo * sizeof(double), c->s));
  }
  if (pool_alloc(c, &c->qc, co * sizeof(double)) != hipSuccess) return set_err(BSAM_ENOMEM, "qc pool");
  HIPCHK(hipMemsetAsync(c->qc, 0, co * sizeof(double), c->s));
  int maxblocks = 1;
  for (int l = minlevel; l <= finestlevel; l++) {
    Level& L = c->L(l);
    if (l <= keep) {
      L = std::move(held.v[l - minlevel]);
      held.v[l - minlevel] = Level();
    } else if (l <= dkeep && !donor->lv[l - donor->minlevel].patches.empty()) {
      L = std::move(donor->lv[l - donor->minlevel]);
      donor->lv[l - donor->minlevel] = Level();
    } else {
      RETCHK(build_level(c, L, tables, l == defer_level));
    }
    // per-patch launches round every patch up to whole blocks
    maxblocks = std::max(maxblocks, nblocks(L.total[kCell]) + (int)L.patches.size());
  }
  RETCHK(plan_chain(c));
  if (maxblocks > c->partial_cap) {
    hipFree(c->d_partial);
    hipFree(c->d_partial2);
    HIPCHK(hipMalloc(&c->d_partial, kNvMax * (long)maxblocks * sizeof(double)));
    HIPCHK(hipMalloc(&c->d_partial2, kNvMax * (long)(maxblocks / 256 + 2) * sizeof(double)));
    c->partial_cap = maxblocks;
  }
  c->have_hier = true;
  if (c->p.precision == BSAM_PREC_MIXED) {
    RETCHK(bsam::mixed_init(c));
  }
  HIPCHK(hipStreamSynchronize(c->s));
  return BSAM_OK;
}
}  // namespace

extern "C" int32_t bsam_set_hierarchy(bsam_ctx* c, int32_t npatch, const bsam_patch_desc* desc,
                                      int32_t minlevel, int32_t finestlevel) {
  if (!c) return set_err(BSAM_EINVAL, "bad arguments");
  int keep = minlevel - 1;
  RETCHK(install(c, npatch, desc, minlevel, finestlevel, &keep));
  c->amr_mesh = false;
  // a fresh install: the kept levels' launch shapes go back to the defaults
  for (int l = minlevel; l <= keep; l++) {
    Level& L = c->L(l);
    L.order = 0;
    L.gorder[0] = L.gorder[1] = 0;
    L.kc_sweep[0] = L.kc_sweep[1] = 0;
    L.kc_resid = 0;
    L.gblock[0] = L.gblock[1] = 0;
  }
  return BSAM_OK;
}

// ---------------------------------------------------------------------------
// host self-test of the flattened ghost-fill tables
// ---------------------------------------------------------------------------
namespace {

// Host evaluation of both forms of SetGhost on one pool, in FP64, with the
// interpolation constants of boundary.f90:54-63.
const double hA[3] = {-1.0 / 5.0, 2.0 / 3.0, 8.0 / 15.0};
const double hB[3] = {5.0 / 32.0, 15.0 / 16.0, -3.0 / 32.0};
const double hC[4] = {9.0 / 16.0, 3.0 / 16.0, 3.0 / 16.0, 1.0 / 16.0};

double host_interp(const GRule& R, int kind, int v, const std::vector<double>& q) {
  const int fv = v * R.pvs, pv = v * R.cvs;
  double coarse;
  if (kind == 1) coarse = q[R.c[0] + pv];
  else if (kind == 3) coarse = hB[0] * q[R.c[0] + pv] + hB[1] * q[R.c[1] + pv] + hB[2] * q[R.c[2] + pv];
  else
    coarse = hC[0] * q[R.c[0] + pv] + hC[1] * q[R.c[1] + pv] + hC[2] * q[R.c[2] + pv] + hC[3] * q[R.c[3] + pv];
  return hA[0] * q[R.s1 + fv] + hA[1] * q[R.s2 + fv] + hA[2] * coarse;
}
double host_post(double x, int post, int v, int homog) { return ghost_post<double>(x, post, v, homog); }

// The staged order: interpolation rules of the ipass's segments, then the
// box stages one box after another (the reference's sequential order).
void host_staged(const bsam_ctx* c, const Level& L, const GhostPlan& G, int ipass, int homog,
                 std::vector<double>& q) {
  const int nv = c->p.nrvars;
  const int s0 = (ipass == 2) ? 1 : 0, s1 = (ipass == 1) ? 1 : 2;
  const int np = (int)L.patches.size();
  for (int sg = s0; sg <= s1; sg++)
    for (int r = G.roff[sg][0]; r < G.roff[sg][np]; r++)
      for (int v = 0; v < nv; v++) q[G.rules[r].dst + v * G.rules[r].pvs] = host_interp(G.rules[r], G.rules[r].kind, v, q);
  for (const auto& st : G.stages)
    for (const BoxDesc& B : st) {
      const DPatch& D = c->hp[B.dp];
      const DPatch& S = c->hp[B.sp];
      for (int v = 0; v < nv; v++)
        for (int z = 0; z < B.ext[2]; z++)
          for (int y = 0; y < B.ext[1]; y++)
            for (int x = 0; x < B.ext[0]; x++) {
              const double a = q[qidx(S, B.slo[0] + x, B.slo[1] + y, B.slo[2] + z) + v * S.qvs];
              const double dir = homog ? -a : 2.0 - a;
              double y0;
              switch (B.op) {
                case kCopyPlus0: y0 = a + 0.0; break;
                case kDirichlet2: y0 = dir; break;
                case kBgMixed: y0 = (v == 0) ? a : dir; break;
                case kTumourBC: y0 = (v == 2) ? dir : ((v == 3) ? -a : a); break;
                default: y0 = a;
              }
              q[qidx(D, B.dlo[0] + x, B.dlo[1] + y, B.dlo[2] + z) + v * D.qvs] = y0;
            }
    }
}

// The flattened form: all right-hand sides first, then the stores (the
// device evaluates the rules concurrently).
void host_map(const bsam_ctx* c, const HostMap& H, int homog, std::vector<double>& q) {
  const int nv = c->p.nrvars;
  const int cs[3] = {0, H.nc1, (int)H.copies.size()}, is[3] = {0, H.ni1, (int)H.interps.size()};
  for (int ph = 0; ph < 2; ph++) {
    std::vector<double> val;
    for (int n = cs[ph]; n < cs[ph + 1]; n++)
      for (int v = 0; v < nv; v++)
        val.push_back(host_post(q[H.copies[n].y + v * H.copies[n].w], H.cpost.empty() ? 0 : H.cpost[n], v, homog));
    for (int n = is[ph]; n < is[ph + 1]; n++) {
      const GRule& R = H.interps[n];
      for (int v = 0; v < nv; v++) val.push_back(host_post(host_interp(R, R.kind & 0xff, v, q), R.kind >> 8, v, homog));
    }
    size_t t = 0;
    for (int n = cs[ph]; n < cs[ph + 1]; n++)
      for (int v = 0; v < nv; v++) q[H.copies[n].x + v * H.copies[n].z] = val[t++];
    for (int n = is[ph]; n < is[ph + 1]; n++)
      for (int v = 0; v < nv; v++) q[H.interps[n].dst + v * H.interps[n].pvs] = val[t++];
  }
}

struct Rng {
  uint64_t s;
  double next() {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return (double)(s >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
  }
};

}  // namespace

extern "C" int32_t bsam_selftest_ghost_maps(const bsam_params* params, int32_t npatch, const bsam_patch_desc* desc,
                                            int32_t minlevel, int32_t finestlevel, uint64_t seed, int64_t* checked,
                                            int32_t* levels_flat) {
  if (!params || !checked || !levels_flat) return set_err(BSAM_EINVAL, "null argument");
  *checked = 0;
  *levels_flat = 0;
  bsam_ctx tmp;
  bsam_ctx* c = &tmp;
  c->p = *params;
  c->p.batching = 1;
  RETCHK(plan_patches(c, npatch, desc, minlevel, finestlevel));
  const int nd = c->p.ndims, nv = c->p.nrvars;
  Rng rng{seed * 2654435761u + 12345u};
  for (Level& L : c->lv) {
    GhostPlan G;
    RETCHK(plan_ghost_stages(c, L, G));
    plan_interp_rules(c, L, G);
    RETCHK(plan_ghost_maps(c, L, G));
    if (!G.maps_ok) continue;  // the level keeps the staged fill
    ++*levels_flat;
    // cells of the level by kind
    std::vector<int> face[2], inner[2];
    for (int pid : L.patches) {
      const DPatch& P = c->hp[pid];
      for (int k = (nd == 3 ? 0 : 0); k < P.nq[2]; k++)
        for (int j = 0; j < P.nq[1]; j++)
          for (int i = 0; i < P.nq[0]; i++) {
            const int n = n_outside(P, nd, i, j, k);
            const int o = qidx(P, i, j, k), col = (i + j + k) & 1;
            for (int v = 0; v < nv; v++) {
              if (n == 0) inner[col].push_back(o + v * P.qvs);
              if (n == 1) face[col].push_back(o + v * P.qvs);
            }
          }
    }
    for (int homog = 0; homog <= 1; homog++) {
      std::vector<double> a(c->nq), b;
      for (double& x : a) x = rng.next();
      // (1) every full table against the staged call
      for (int p = 0; p < 3; p++) {
        b = a;
        std::vector<double> s = a;
        host_staged(c, L, G, p, homog, s);
        host_map(c, G.map[p], homog, b);
        for (long n = 0; n < c->nq; n++)
          if (memcmp(&s[n], &b[n], sizeof(double)))
            return set_err(BSAM_EINVAL, "level %d ipass %d homog %d: table differs from the staged fill at %ld",
                           L.level, p, homog, n);
        *checked += c->nq;
      }
      // (1b) a single patch with condition 10 on every face: the one-pass
      // shell fill (phys_shell) against the staged call
      bool allphys = L.patches.size() == 1 && G.wrap_pid < 0 && (int)G.stages.size() == nd;
      for (int f = 0; f < 2 * nd && allphys; f++) allphys = c->desc[L.patches[0]].mthbc[f] == 10;
      if (allphys) {
        const DPatch& P = c->hp[L.patches[0]];
        std::vector<double> s = a;
        b = a;
        host_staged(c, L, G, 0, homog, s);
        const int ns = nd == 3 ? shell_cells<3>(P) : shell_cells<2>(P);
        std::vector<uint8_t> seen((size_t)P.qvs, 0);
        for (int e = 0; e < ns; e++) {
          int i, j, k, so, out[3];
          if (nd == 3) shell_cell<3>(P, e, i, j, k, so, out);
          else shell_cell<2>(P, e, i, j, k, so, out);
          const int d = qidx(P, i, j, k);
          if (!n_outside(P, nd, i, j, k) || n_outside(P, nd, i, j, k) != out[0] + out[1] + out[2] ||
              seen[d - P.qoff]++)
            return set_err(BSAM_EINVAL, "level %d: shell cell %d (%d,%d,%d) is not a new ghost cell", L.level, e,
                           i, j, k);
          for (int v = 0; v < nv; v++) {
            double x = a[so + v * P.qvs];
            for (int ax = 0; ax < 3; ax++)
              if (out[ax]) x = box_value<double>(phys_face_op(c), v, homog, x);
            b[d + v * P.qvs] = x;
          }
        }
        long nghost = 1, ninner = 1;
        for (int ax = 0; ax < nd; ax++) {
          nghost *= P.nq[ax];
          ninner *= P.mx[ax];
        }
        if (ns != nghost - ninner)
          return set_err(BSAM_EINVAL, "level %d: %d shell cells, %ld ghost cells", L.level, ns, nghost - ninner);
        for (long n = 0; n < c->nq; n++)
          if (memcmp(&s[n], &b[n], sizeof(double)))
            return set_err(BSAM_EINVAL, "level %d homog %d: shell fill differs from the staged fill at %ld",
                           L.level, homog, n);
        *checked += c->nq;
      }
      // (2) a red-black sequence (three passes, the last SetGhost full)
      // from an arbitrary state, each sweep replaced by new values of its
      // colour's interior cells; (3) the same after SetGhost(level, 0)
      // against its reduced table map[5]; (4) the same after a correction:
      // from a state SetGhost(level, 0) has made consistent, every cell
      // (ghosts included) shifted by one value per variable and global
      // position modulo the period, as the prolongation of a coarse
      // correction does, then SetGhost(level, 0) against map[6]
      for (int variant = 0; variant < 3; variant++) {
      for (long n = 0; n < c->nq; n++) a[n] = rng.next();
      if (variant == 2) {
        host_staged(c, L, G, 0, homog, a);
        for (int pid : L.patches) {
          const DPatch& P = c->hp[pid];
          const bsam_patch_desc& d = c->desc[pid];
          for (int k = 0; k < P.nq[2]; k++)
            for (int j = 0; j < P.nq[1]; j++)
              for (int i = 0; i < P.nq[0]; i++) {
                const int loc[3] = {i, j, k};
                uint64_t h = 1469598103934665603ull;
                for (int ax = 0; ax < nd; ax++) {
                  const int per = level_period(c, L.level, ax);
                  const int g = (((d.mglobal[ax][0] + loc[ax]) % per) + per) % per;
                  h = (h ^ (uint64_t)g) * 1099511628211ull;
                }
                for (int v = 0; v < nv; v++) {
                  const uint64_t hv = (h ^ (uint64_t)(v + 1)) * 0x9E3779B97F4A7C15ull;
                  a[qidx(P, i, j, k) + v * P.qvs] += (double)(hv >> 11) * (1.0 / 9007199254740992.0) - 0.5;
                }
              }
        }
      }
      b = a;
      if (variant >= 1) {
        host_staged(c, L, G, 0, homog, a);
        host_map(c, G.map[variant == 1 ? 5 : 6], homog, b);
        for (int o : face[1])
          if (memcmp(&a[o], &b[o], sizeof(double)))
            return set_err(BSAM_EINVAL, "level %d homog %d: colour-1 face ghost %d differs after map[%d]", L.level,
                           homog, o, variant == 1 ? 5 : 6);
        *checked += face[1].size();
      }
      const int ns = 3;
      for (int sweep = 0; sweep < 2 * ns; sweep++) {
        const int col = sweep & 1, ipass = col + 1;
        const bool last = sweep == 2 * ns - 1;
        for (int o : inner[col]) a[o] = b[o] = rng.next();
        host_staged(c, L, G, ipass, homog, a);
        host_map(c, G.map[last ? ipass : 2 + ipass], homog, b);
        const std::vector<int>& f0 = face[col];
        for (int o : f0)
          if (memcmp(&a[o], &b[o], sizeof(double)))
            return set_err(BSAM_EINVAL, "level %d homog %d sweep %d: colour-%d face ghost %d differs", L.level,
                           homog, sweep, col, o);
        *checked += f0.size();
        if (ipass == 2) {
          for (int o : face[0])
            if (memcmp(&a[o], &b[o], sizeof(double)))
              return set_err(BSAM_EINVAL, "level %d homog %d sweep %d: face ghost %d differs", L.level, homog,
                             sweep, o);
          *checked += face[0].size();
        }
      }
      for (long n = 0; n < c->nq; n++)
        if (memcmp(&a[n], &b[n], sizeof(double)))
          return set_err(BSAM_EINVAL, "level %d homog %d variant %d: state after the sequence differs at %ld",
                         L.level, homog, variant, n);
      *checked += c->nq;
      }
    }
  }
  return BSAM_OK;
}

// ---------------------------------------------------------------------------
// upload / download
// ---------------------------------------------------------------------------
namespace {
int field_kind(int field) {
  switch (field) {
    case BSAM_FIELD_Q: case BSAM_FIELD_QOLD: return 0;
    case BSAM_FIELD_F: case BSAM_FIELD_RF: case BSAM_FIELD_FTMP: return 1;
    case BSAM_FIELD_QC: return 2;
  }
  return -1;
}
double* field_pool(bsam_ctx* c, int field) {
  switch (field) {
    case BSAM_FIELD_Q: return c->q;
    case BSAM_FIELD_QOLD: return c->qold;
    case BSAM_FIELD_F: return c->f;
    case BSAM_FIELD_RF: return c->rf;
    case BSAM_FIELD_FTMP: return c->ftmp;
    case BSAM_FIELD_QC: return c->qc;
  }
  return nullptr;
}
}  // namespace

extern "C" int64_t bsam_field_size(const bsam_ctx* c, int32_t patch, int32_t field) {
  if (!c || !c->have_hier || patch < 0 || patch >= c->npatch) return -1;
  const DPatch& P = c->hp[patch];
  const int kind = field_kind(field);
  if (kind == 0) return (int64_t)P.nq[0] * P.nq[1] * P.nq[2] * c->p.nrvars;
  if (kind == 1) return (int64_t)P.mx[0] * P.mx[1] * P.mx[2] * c->p.nrvars;
  if (kind == 2) return (int64_t)P.nc[0] * P.nc[1] * P.nc[2] * c->p.nrvars;
  return -1;
}

static int32_t xfer(bsam_ctx* c, int32_t patch, int32_t field, double* host, int64_t count, int to_dev) {
  if (!c || !host) return set_err(BSAM_EINVAL, "null argument");
  if (!c->have_hier) return set_err(BSAM_ESTATE, "no hierarchy installed");
  if (patch < 0 || patch >= c->npatch) return set_err(BSAM_EINVAL, "patch %d out of range", patch);
  const int kind = field_kind(field);
  if (kind < 0) return set_err(BSAM_EINVAL, "unknown field %d", field);
  const int64_t n = bsam_field_size(c, patch, field);
  if (count != n) return set_err(BSAM_EINVAL, "field holds %lld doubles, buffer %lld", (long long)n, (long long)count);
  if (n > c->stage_cap) {
    HIPCHK(hipStreamSynchronize(c->s));
    hipFree(c->d_stage);
    c->d_stage = nullptr;
    if (hipMalloc(&c->d_stage, n * sizeof(double)) != hipSuccess) return set_err(BSAM_ENOMEM, "staging buffer");
    c->stage_cap = n;
  }
  double* pool = field_pool(c, field);
  if (to_dev) {
    HIPCHK(hipMemcpyAsync(c->d_stage, host, n * sizeof(double), hipMemcpyHostToDevice, c->s));
    k_layout<<<nblocks(n), kBlock, 0, c->s>>>(c->d_pt, patch, kind, c->p.nrvars, 1, c->d_stage, pool);
    LAUNCH_CHECK();
    HIPCHK(hipStreamSynchronize(c->s));
  } else {
    k_layout<<<nblocks(n), kBlock, 0, c->s>>>(c->d_pt, patch, kind, c->p.nrvars, 0, c->d_stage, pool);
    LAUNCH_CHECK();
    HIPCHK(hipMemcpyAsync(host, c->d_stage, n * sizeof(double), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
  }
  return BSAM_OK;
}

extern "C" int32_t bsam_upload(bsam_ctx* c, int32_t patch, int32_t field, const double* host, int64_t count) {
  return xfer(c, patch, field, const_cast<double*>(host), count, 1);
}
extern "C" int32_t bsam_download(bsam_ctx* c, int32_t patch, int32_t field, double* host, int64_t count) {
  return xfer(c, patch, field, host, count, 0);
}

// ---------------------------------------------------------------------------
// per-level operations
// ---------------------------------------------------------------------------
#define DISPATCH_DIM(nd, KCALL2, KCALL3) \
  do {                                   \
    if ((nd) == 2) { KCALL2; } else { KCALL3; } \
  } while (0)

// A tumour level that is one patch with condition 10 on every face: its
// sweeps write their ghost images (phys_cell_images) instead of a SetGhost
// after each colour.
static bool phys_images_level(const bsam_ctx* c, const Level& L) {
  if (c->p.model != BSAM_MODEL_TUMOUR || L.patches.size() != 1 || L.wrap_pid >= 0 || L.level > 0 ||
      (int)L.stages.size() != c->p.ndims)
    return false;
  for (int f = 0; f < 2 * c->p.ndims; f++)
    if (c->desc[L.patches[0]].mthbc[f] != 10) return false;
  return true;
}

static int32_t relax_level_impl(bsam_ctx* c, int32_t level, int32_t ipass, bool img);
extern "C" int32_t bsam_relax_level(bsam_ctx* c, int32_t level, int32_t ipass) {
  return relax_level_impl(c, level, ipass, false);
}
static int32_t relax_level_impl(bsam_ctx* c, int32_t level, int32_t ipass, bool img) {
  RETCHK(check_level(c, level));
  if (ipass != 1 && ipass != 2) return set_err(BSAM_EINVAL, "ipass must be 1 or 2");
  Level& L = c->L(level);
  const int col = ipass - 1;
  const bool fin = level == c->finest;
  if (fin) kstat_begin(c, BSAM_KSTAT_RELAX_FINEST);
  FOR_LAUNCHES(c, L, kHalf, a, nb) {
    if (c->p.model == BSAM_MODEL_CH) {
      const double eps2 = c->p.model_params[0];
      if (c->p.ndims == 2)
        k_relax_ch<2, false><<<nb, kBlock, 0, c->s>>>(a, col, c->p.dt, eps2, L.h, c->q, c->f);
      else
        k_relax_ch<3, false><<<nb, kBlock, 0, c->s>>>(a, col, c->p.dt, eps2, L.h, c->q, c->f);
    } else if (c->p.model == BSAM_MODEL_TUMOUR) {
      if (img)
        DISPATCH_DIM(c->p.ndims, (k_relax_tm<2, true><<<nb, kBlock, 0, c->s>>>(a, col, c->mk(), c->q, c->qold, c->f)),
                     (k_relax_tm<3, true><<<nb, kBlock, 0, c->s>>>(a, col, c->mk(), c->q, c->qold, c->f)));
      else
        DISPATCH_DIM(c->p.ndims, (k_relax_tm<2><<<nb, kBlock, 0, c->s>>>(a, col, c->mk(), c->q, c->qold, c->f)),
                     (k_relax_tm<3><<<nb, kBlock, 0, c->s>>>(a, col, c->mk(), c->q, c->qold, c->f)));
    } else {
      k_relax_bg2d<<<nb, kBlock, 0, c->s>>>(a, col, c->mk(), L.h, c->q, c->qold, c->f);
    }
    LAUNCH_CHECK();
  }
  if (fin) {
    // algorithmic bytes of one colour sweep (CH): phi over N+G_f, mu of the
    // other colour over (N+G_f)/2, f of the colour (2 fields) N/2, and the
    // colour's phi and mu written, N/2 each.
    double N = 0, G = 0;
    for (int pid : L.patches) {
      const DPatch& P = c->hp[pid];
      const double n = (double)P.mx[0] * P.mx[1] * P.mx[2];
      N += n;
      G += (c->p.ndims == 3) ? 2.0 * ((double)P.mx[0] * P.mx[1] + (double)P.mx[1] * P.mx[2] + (double)P.mx[0] * P.mx[2])
                             : 2.0 * ((double)P.mx[0] + P.mx[1]);
    }
    // tumour: the other colour's four fields over (N+G)/2 each, the cell's
    // own phi and qold N/2 each, f (four fields) N/2 each, four fields
    // written N/2 each
    const double bytes = (c->p.model == BSAM_MODEL_CH)
                             ? 8.0 * ((N + G) + 0.5 * (N + G) + N + N)
                             : (c->p.model == BSAM_MODEL_TUMOUR) ? 8.0 * (2.0 * (N + G) + 5.0 * N)
                                                                 : 8.0 * (2.0 * (N + G) + 0.5 * N + N + N);
    kstat_end(c, BSAM_KSTAT_RELAX_FINEST, bytes);
  }
  return BSAM_OK;
}

// SetGhost on any field of the q geometry (FP64 solution or FP32 correction).
// form kLean (ipass 1, 2 inside a red-black sweep sequence whose last
// SetGhost is not lean) refreshes only the face ghosts such a sequence
// reads; kAfterCorrection (ipass 0, right after CorrectFine, before such a
// sequence) also leaves out the ghosts the correction left exact.  Levels
// without flattened tables do the full SetGhost.
// kLeanImg: kLean after a sweep that wrote the fill's images (ImgTab):
// the rest of the reduced table.
enum GhostForm : int { kFull = 0, kLean = 1, kAfterCorrection = 2, kLeanImg = 3 };
template <typename T>
static int32_t set_ghost_t(bsam_ctx* c, int level, int ipass, T* q, int homog, int form = kFull) {
  Level& L = c->L(level);
  const int nv = c->p.nrvars;
  if (L.maps_ok && c->p.batching) {
    const int mi = form == kAfterCorrection ? 6 : (form == kLean ? (ipass > 0 ? 2 + ipass : 5) : ipass);
    const DevMap& M = form == kLeanImg ? L.gimg[ipass - 1] : L.gmap[mi];
    // rules reading ghosts from before the call first, then the rest
    const int pr = sizeof(T) == 8 ? 0 : 1;
    const int bs = L.gblock[pr] > 0 ? L.gblock[pr] : kBlock;
    if (M.nc1 + M.ni1 > 0) {
      k_ghost_map<T><<<(M.nc1 + M.ni1 + bs - 1) / bs, bs, 0, c->s>>>(M.cr, M.cpost, M.nc1, M.ir, M.ni1, nv, homog, q,
                                                                     L.gorder[pr]);
      LAUNCH_CHECK();
    }
    const int nc = M.nc - M.nc1, ni = M.ni - M.ni1;
    if (nc + ni > 0) {
      k_ghost_map<T><<<(nc + ni + bs - 1) / bs, bs, 0, c->s>>>(M.cr + M.nc1, M.cpost ? M.cpost + M.nc1 : nullptr, nc,
                                                                M.ir + M.ni1, ni, nv, homog, q, L.gorder[pr]);
      LAUNCH_CHECK();
    }
    return BSAM_OK;
  }
  if (level > 0 && L.d_rules) {
    // colour-filter segments this ipass refreshes: 0 -> all, 1 -> 0..1,
    // 2 -> 1..2
    const int s0 = (ipass == 2) ? 1 : 0, s1 = (ipass == 1) ? 1 : 2;
    const int np = (int)L.patches.size();
    if (c->p.batching) {
      const int lo = L.roff[s0][0], hi = L.roff[s1][np];
      if (hi > lo) {
        k_interp_map<T><<<nblocks(hi - lo), kBlock, 0, c->s>>>(L.d_rules, lo, hi, nv, q);
        LAUNCH_CHECK();
      }
    } else {
      for (int a = 0; a < np; a++)
        for (int sg = s0; sg <= s1; sg++) {
          const int lo = L.roff[sg][a], hi = L.roff[sg][a + 1];
          if (hi <= lo) continue;
          k_interp_map<T><<<nblocks(hi - lo), kBlock, 0, c->s>>>(L.d_rules, lo, hi, nv, q);
          LAUNCH_CHECK();
        }
    }
  }
  if (L.wrap_pid >= 0) {
    const DPatch& P = c->hp[L.wrap_pid];
    const long shell = (long)P.nq[0] * P.nq[1] * P.nq[2] - (long)P.mx[0] * P.mx[1] * P.mx[2];
    if (c->p.ndims == 2) k_wrap_full<2, T><<<nblocks(shell), kBlock, 0, c->s>>>(c->d_pt, L.wrap_pid, nv, q);
    else k_wrap_full<3, T><<<nblocks(shell), kBlock, 0, c->s>>>(c->d_pt, L.wrap_pid, nv, q);
    LAUNCH_CHECK();
  }
  // a small level whose fill is box stages only: one workgroup, in order
  if (c->p.batching && !L.d_rules && !L.stages.empty() && (int)L.stages.size() <= kMaxStages) {
    int maxel = 0;
    for (const Stage& s : L.stages) maxel = std::max(maxel, s.maxel);
    if (maxel * nv <= kSeqBoxElems) {
      BoxSeq S{};
      S.ns = (int)L.stages.size();
      for (int st = 0; st < S.ns; st++) {
        S.d[st] = L.stages[st].d;
        S.n[st] = L.stages[st].n;
      }
      k_boxops_seq<T><<<1, 1024, 0, c->s>>>(c->d_pt, S, nv, homog, q);
      LAUNCH_CHECK();
      return BSAM_OK;
    }
  }
  for (const Stage& s : L.stages) {
    // each box over as many workgroups as its size needs; per-patch mode
    // launches one box (one patch pair or face) at a time, as the
    // reference's pairwise callbacks go
    const int per = c->p.batching ? s.n : 1;
    const int chunks = std::max(1, std::min(64, (s.maxel * nv + kBlock - 1) / kBlock));
    for (int b0 = 0; b0 < s.n; b0 += per) {
      k_boxops<T><<<dim3(per, chunks), kBlock, 0, c->s>>>(c->d_pt, s.d + b0, nv, homog, q);
      LAUNCH_CHECK();
    }
  }
  return BSAM_OK;
}

extern "C" int32_t bsam_set_ghost(bsam_ctx* c, int32_t level, int32_t ipass) {
  RETCHK(check_level(c, level));
  if (ipass < 0 || ipass > 2) return set_err(BSAM_EINVAL, "ipass must be 0, 1 or 2");
  return set_ghost_t<double>(c, level, ipass, c->q, 0);
}

// Launch shape of the marching 3D colour sweep (tiles of 64 half-rows x
// kSweepTY rows) for a single large patch, or 0 when the level uses the
// flat kernel.  Two 1024-thread workgroups fit on a CU; the plane count per
// workgroup is chosen for about 1024 workgroups in FP64 and 512 in FP32,
// the fastest of the shapes measured on the 256^3 patch.
static int sweep_kc(bsam_ctx* c, const Level& L, int elem_bytes, int* nblk) {
  if (c->p.ndims != 3 || L.patches.size() != 1) return 0;
  const DPatch& P = c->hp[L.patches[0]];
  const int nx2 = P.mx[0] / 2;
  if (nx2 < 64 || P.mx[1] < kSweepTY || (P.mx[0] & 1)) return 0;
  const int ntile = ((nx2 + 63) / 64) * ((P.mx[1] + kSweepTY - 1) / kSweepTY);
  // the tuned planes per workgroup, else about 1024 (FP64) / 512 (FP32)
  // workgroups of at least 4 planes
  int kc = L.kc_sweep[elem_bytes == 8 ? 0 : 1];
  if (kc <= 0) {
    const int target = (elem_bytes == 8) ? 1024 : 512;
    const int nkc = std::max(1, (target + ntile - 1) / ntile);
    kc = std::max(4, (P.mx[2] + nkc - 1) / nkc);
  }
  *nblk = ntile * ((P.mx[2] + kc - 1) / kc);
  return kc;
}

// Launch shape of the marching residual kernels (tiles of 64 half-rows x 8
// rows, 512 threads, about 768 workgroups), or 0.
static int march_kc(bsam_ctx* c, const Level& L, int* nblk) {
  if (c->p.ndims != 3 || L.patches.size() != 1) return 0;
  const DPatch& P = c->hp[L.patches[0]];
  const int nx2 = P.mx[0] / 2;
  if (nx2 < 64 || P.mx[1] < 8 || (P.mx[0] & 1)) return 0;
  const int ntile = ((nx2 + 63) / 64) * ((P.mx[1] + 7) / 8);
  int kc = L.kc_resid;
  if (kc <= 0) {
    const int nkc = std::max(1, (768 + ntile - 1) / ntile);
    kc = std::max(4, (P.mx[2] + nkc - 1) / nkc);
  }
  *nblk = ntile * ((P.mx[2] + kc - 1) / kc);
  return kc;
}

// One colour sweep followed by SetGhost(level, ipass) (the body of
// LevelRelax, afasroutines.f90:546-555).  For a single patch periodic in
// every dimension the sweep writes the ghost images itself, which is the
// whole effect of that SetGhost on the values that changed.
// The colour sweep of ipass writes the images of the reduced fill after it
// (build_images) when the level has them and the knob allows.
static bool use_images(const bsam_ctx* c, const Level& L, int ipass) {
  return c->sweep_images && c->p.batching && L.img_ok[ipass - 1];
}
// bytes an image moves: its entry, and each variable written
static double image_bytes(const Level& L, int ipass, double value_bytes, int nv) {
  return (double)L.img_n[ipass - 1] * (16.0 + nv * value_bytes);
}

// The CH FP64 colour sweep of a level with its fill's images (batched).
static int32_t relax_ch_images(bsam_ctx* c, int level, int ipass) {
  Level& L = c->L(level);
  const bool fin = level == c->finest;
  if (fin) kstat_begin(c, BSAM_KSTAT_RELAX_FINEST);
  const ImgTab I{L.img_off[ipass - 1], L.img_ent[ipass - 1], 0};
  const double eps2 = c->p.model_params[0];
  FOR_LAUNCHES(c, L, kHalf, a, nb) {
    DISPATCH_DIM(c->p.ndims,
                 (k_relax_ch<2, false, true><<<nb, kBlock, 0, c->s>>>(a, ipass - 1, c->p.dt, eps2, L.h, c->q, c->f, I)),
                 (k_relax_ch<3, false, true><<<nb, kBlock, 0, c->s>>>(a, ipass - 1, c->p.dt, eps2, L.h, c->q, c->f, I)));
    LAUNCH_CHECK();
  }
  if (fin) {
    // relax_level_impl's algorithmic bytes of the sweep, plus the images
    double N = 0, G = 0;
    for (int pid : L.patches) {
      const DPatch& P = c->hp[pid];
      N += (double)P.mx[0] * P.mx[1] * P.mx[2];
      G += (c->p.ndims == 3) ? 2.0 * ((double)P.mx[0] * P.mx[1] + (double)P.mx[1] * P.mx[2] + (double)P.mx[0] * P.mx[2])
                             : 2.0 * ((double)P.mx[0] + P.mx[1]);
    }
    kstat_end(c, BSAM_KSTAT_RELAX_FINEST, 8.0 * ((N + G) + 0.5 * (N + G) + N + N) + image_bytes(L, ipass, 8.0, 2));
  }
  return BSAM_OK;
}

static int32_t relax_and_ghost(bsam_ctx* c, int level, int ipass, bool lean) {
  Level& L = c->L(level);
  const bool wrap = L.wrap_img;
  int nblk = 0;
  const int kc = (c->p.model == BSAM_MODEL_CH) ? sweep_kc(c, L, 8, &nblk) : 0;
  if (phys_images_level(c, L)) return relax_level_impl(c, level, ipass, true);  // no SetGhost
  if (!kc && !wrap && lean && use_images(c, L, ipass)) {
    RETCHK(relax_ch_images(c, level, ipass));
    return set_ghost_t<double>(c, level, ipass, c->q, 0, kLeanImg);
  }
  if (!kc && !(wrap && c->p.model == BSAM_MODEL_CH)) {
    RETCHK(bsam_relax_level(c, level, ipass));
    return set_ghost_t<double>(c, level, ipass, c->q, 0, lean ? kLean : kFull);
  }
  const int col = ipass - 1;
  const bool fin = level == c->finest;
  const double h2 = L.h * L.h, eps2 = c->p.model_params[0];
  if (fin) kstat_begin(c, BSAM_KSTAT_RELAX_FINEST);
  if (kc) {
    const double tmp1 = c->p.dt / h2, tmp2 = eps2 / h2, tmp3 = 6.0 * eps2 / h2, a12 = 6.0 * tmp1;
    const int pid = L.patches[0];
    if (wrap) {
      if (fin) k_relax3d_march<double, false, true, 1><<<nblk, kSweepThreads, 0, c->s>>>(c->d_pt, pid, col, kc, tmp1, tmp2, tmp3, a12, c->q, c->f, nullptr);
      else k_relax3d_march<double, false, true, 0><<<nblk, kSweepThreads, 0, c->s>>>(c->d_pt, pid, col, kc, tmp1, tmp2, tmp3, a12, c->q, c->f, nullptr);
    } else {
      if (fin) k_relax3d_march<double, false, false, 1><<<nblk, kSweepThreads, 0, c->s>>>(c->d_pt, pid, col, kc, tmp1, tmp2, tmp3, a12, c->q, c->f, nullptr);
      else k_relax3d_march<double, false, false, 0><<<nblk, kSweepThreads, 0, c->s>>>(c->d_pt, pid, col, kc, tmp1, tmp2, tmp3, a12, c->q, c->f, nullptr);
    }
  } else {
    LevelArgs a = largs(c, L, kHalf);
    DISPATCH_DIM(c->p.ndims, (k_relax_ch<2, true><<<nblocks(a.total), kBlock, 0, c->s>>>(a, col, c->p.dt, eps2, L.h, c->q, c->f)),
                 (k_relax_ch<3, true><<<nblocks(a.total), kBlock, 0, c->s>>>(a, col, c->p.dt, eps2, L.h, c->q, c->f)));  // one patch
  }
  LAUNCH_CHECK();
  if (fin) {
    double N = 0, G = 0;
    for (int pid : L.patches) {
      const DPatch& P = c->hp[pid];
      N += (double)P.mx[0] * P.mx[1] * P.mx[2];
      G += (c->p.ndims == 3) ? 2.0 * ((double)P.mx[0] * P.mx[1] + (double)P.mx[1] * P.mx[2] + (double)P.mx[0] * P.mx[2])
                             : 2.0 * ((double)P.mx[0] + P.mx[1]);
    }
    kstat_end(c, BSAM_KSTAT_RELAX_FINEST, 8.0 * ((N + G) + 0.5 * (N + G) + N + N));
  }
  if (!wrap) RETCHK(set_ghost_t<double>(c, level, ipass, c->q, 0, lean ? kLean : kFull));
  return BSAM_OK;
}

// LevelRelax; with `descend` (the pre-smoothing of a level the V-cycle
// then restricts) the last SetGhost is the reduced table as well: after
// the two reduced fills of a pass every face ghost is current, and until
// the fill after CorrectFine only face ghosts are read (the coarse
// right-hand side's stencils).  An edge or corner ghost left stale there is
// changed by CorrectFine and then rewritten by the post-smoothing's last
// fill, which is full and whose every expression reads interior and parent
// cells only; the copies the fill after CorrectFine leaves out are
// colour-1 face ghosts, which the reduced fill of ipass 2 writes.
static int32_t level_relax_t(bsam_ctx* c, int32_t level, bool descend) {
  RETCHK(check_level(c, level));
  for (int s = 0; s < c->p.nsmoothingpasses; s++)
    for (int rb = 1; rb <= 2; rb++)
      RETCHK(relax_and_ghost(c, level, rb, descend || !(s == c->p.nsmoothingpasses - 1 && rb == 2)));
  return BSAM_OK;
}

extern "C" int32_t bsam_level_relax(bsam_ctx* c, int32_t level) { return level_relax_t(c, level, false); }

static int32_t need_parent(bsam_ctx* c, int level) {
  RETCHK(check_level(c, level));
  if (level <= c->minlevel) return set_err(BSAM_EINVAL, "level %d has no parent level", level);
  return BSAM_OK;
}

template <typename T>
static int32_t restrict_t(bsam_ctx* c, int level, const T* src, T* dst, T* qc, bool wrap = false) {
  Level& L = c->L(level);
  FOR_LAUNCHES(c, L, kCoarse, a, nb) {
    if (wrap)
      DISPATCH_DIM(c->p.ndims, (k_restrict<2, T, true><<<nb, kBlock, 0, c->s>>>(a, c->nv(), src, dst, qc)),
                   (k_restrict<3, T, true><<<nb, kBlock, 0, c->s>>>(a, c->nv(), src, dst, qc)));
    else
      DISPATCH_DIM(c->p.ndims, (k_restrict<2, T><<<nb, kBlock, 0, c->s>>>(a, c->nv(), src, dst, qc)),
                   (k_restrict<3, T><<<nb, kBlock, 0, c->s>>>(a, c->nv(), src, dst, qc)));
    LAUNCH_CHECK();
  }
  return BSAM_OK;
}

// True when SetGhost(level-1, 0) after a restriction from `level` reduces to
// the periodic images of the parent values just written: the parent level
// is one patch periodic in every dimension, completely covered by this
// level's single patch.
static bool restrict_wraps_parent(bsam_ctx* c, int level) {
  if (level <= c->minlevel) return false;
  const Level& L = c->L(level);
  const Level& Lp = c->L(level - 1);
  if (L.patches.size() != 1 || Lp.wrap_pid < 0 || level - 1 > 0) return false;
  const DPatch& P = c->hp[L.patches[0]];
  const DPatch& Q = c->hp[Lp.wrap_pid];
  for (int a = 0; a < c->p.ndims; a++)
    if (P.mb[a][0] != 1 || P.mb[a][1] != Q.mx[a]) return false;
  return true;
}

template <typename T>
static int32_t ring_t(bsam_ctx* c, int level, const T* q, T* qc) {
  Level& L = c->L(level);
  FOR_LAUNCHES(c, L, kQcBox, a, nb) {
    DISPATCH_DIM(c->p.ndims, (k_coarse_ring<2, T><<<nb, kBlock, 0, c->s>>>(a, c->nv(), q, qc)),
                 (k_coarse_ring<3, T><<<nb, kBlock, 0, c->s>>>(a, c->nv(), q, qc)));
    LAUNCH_CHECK();
  }
  return BSAM_OK;
}

template <typename T>
static int32_t correct_t(bsam_ctx* c, int level, T* q, T* qc, bool wrap = false) {
  Level& L = c->L(level);
  {
    FOR_LAUNCHES(c, L, kQcBox, a, nb) {
      DISPATCH_DIM(c->p.ndims, (k_correct_qc<2, T><<<nb, kBlock, 0, c->s>>>(a, c->nv(), q, qc)),
                   (k_correct_qc<3, T><<<nb, kBlock, 0, c->s>>>(a, c->nv(), q, qc)));
      LAUNCH_CHECK();
    }
  }
  FOR_LAUNCHES(c, L, kDual, b, nb) {
    if (wrap)
      DISPATCH_DIM(c->p.ndims, (k_prolong_dual<2, T, true><<<nb, kBlock, 0, c->s>>>(b, c->nv(), q, qc)),
                   (k_prolong_dual<3, T, true><<<nb, kBlock, 0, c->s>>>(b, c->nv(), q, qc)));
    else
      DISPATCH_DIM(c->p.ndims, (k_prolong_dual<2, T><<<nb, kBlock, 0, c->s>>>(b, c->nv(), q, qc)),
                   (k_prolong_dual<3, T><<<nb, kBlock, 0, c->s>>>(b, c->nv(), q, qc)));
    LAUNCH_CHECK();
  }
  return BSAM_OK;
}

extern "C" int32_t bsam_restrict_solution(bsam_ctx* c, int32_t level) {
  RETCHK(need_parent(c, level));
  return restrict_t<double>(c, level, c->q, c->q, c->qc);
}

extern "C" int32_t bsam_coarse_ghost_points(bsam_ctx* c, int32_t level) {
  RETCHK(need_parent(c, level));
  return ring_t<double>(c, level, c->q, c->qc);
}

extern "C" int32_t bsam_coarse_loading(bsam_ctx* c, int32_t level) {
  RETCHK(need_parent(c, level));
  Level& L = c->L(level);
  const ModelK m = c->mk();
  const bool fin = level == c->finest;
  if (fin) kstat_begin(c, BSAM_KSTAT_FASRHS_FINEST);
  int nblk = 0;
  int kc = (c->p.model == BSAM_MODEL_CH) ? march_kc(c, L, &nblk) : 0;
  if (kc) {
    // whole coarse planes per block
    const DPatch& P = c->hp[L.patches[0]];
    kc += kc & 1;
    const int ntile = ((P.mx[0] / 2 + 63) / 64) * ((P.mx[1] + 7) / 8);
    nblk = ntile * ((P.mx[2] + kc - 1) / kc);
    const double h2 = L.h * L.h;
    k_ch3d_resid_march<1, double><<<nblk, 512, 0, c->s>>>(c->d_pt, L.patches[0], kc, m.dt / h2, m.eps2 / h2, m,
                                                          c->q, c->f, nullptr, nullptr, c->qc, c->f);
  } else {
    FOR_LAUNCHES(c, L, kCoarse, a, nb) {
      if (c->p.model == BSAM_MODEL_CH) {
        DISPATCH_DIM(c->p.ndims,
                     (k_fas_rhs<2, BSAM_MODEL_CH><<<nb, kBlock, 0, c->s>>>(a, m, c->q, c->qold, c->f, c->qc, c->f)),
                     (k_fas_rhs<3, BSAM_MODEL_CH><<<nb, kBlock, 0, c->s>>>(a, m, c->q, c->qold, c->f, c->qc, c->f)));
      } else if (c->p.model == BSAM_MODEL_TUMOUR) {
        DISPATCH_DIM(c->p.ndims,
                     (k_fas_rhs<2, BSAM_MODEL_TUMOUR><<<nb, kBlock, 0, c->s>>>(a, m, c->q, c->qold, c->f, c->qc, c->f)),
                     (k_fas_rhs<3, BSAM_MODEL_TUMOUR><<<nb, kBlock, 0, c->s>>>(a, m, c->q, c->qold, c->f, c->qc, c->f)));
      } else {
        k_fas_rhs<2, BSAM_MODEL_BG><<<nb, kBlock, 0, c->s>>>(a, m, c->q, c->qold, c->f, c->qc, c->f);
      }
      LAUNCH_CHECK();
    }
  }
  LAUNCH_CHECK();
  if (fin) {
    double N = 0, G = 0;
    for (int pid : L.patches) {
      const DPatch& P = c->hp[pid];
      N += (double)P.mx[0] * P.mx[1] * P.mx[2];
      G += (c->p.ndims == 3) ? 2.0 * ((double)P.mx[0] * P.mx[1] + (double)P.mx[1] * P.mx[2] + (double)P.mx[0] * P.mx[2])
                             : 2.0 * ((double)P.mx[0] + P.mx[1]);
    }
    const double r = (c->p.ndims == 3) ? 8.0 : 4.0;
    // q over N+G_f, f over N, qc over N/2^d plus its ring (~G_f/2^(d-1)), parent f write N/2^d
    const double bytes = 8.0 * 2.0 * ((N + G) + N + (N / r + G / (r / 2.0)) + N / r);
    kstat_end(c, BSAM_KSTAT_FASRHS_FINEST, bytes);
  }
  return BSAM_OK;
}

extern "C" int32_t bsam_correct_fine(bsam_ctx* c, int32_t level) {
  RETCHK(need_parent(c, level));
  return correct_t<double>(c, level, c->q, c->qc);
}

// FillDown restricted to the levels min(finest, highest) .. max(minlevel+1,
// lowest), in the reference's order (finest first).
static int32_t fill_down_range(bsam_ctx* c, int32_t field, int highest, int lowest) {
  double* fld = (field == BSAM_SOURCEFIELD) ? c->f : c->q;
  for (int level = std::min(c->finest, highest); level >= std::max(c->minlevel + 1, lowest); level--) {
    RETCHK(restrict_t<double>(c, level, fld, fld, nullptr));
    if (field == BSAM_SOLUTIONFIELD) RETCHK(set_ghost_t<double>(c, level - 1, 0, c->q, 0));
  }
  return BSAM_OK;
}

extern "C" int32_t bsam_fill_down(bsam_ctx* c, int32_t field) {
  if (!c || !c->have_hier) return set_err(BSAM_ESTATE, "no hierarchy installed");
  if (field != BSAM_SOURCEFIELD && field != BSAM_SOLUTIONFIELD) return set_err(BSAM_EINVAL, "bad field %d", field);
  double* fld = (field == BSAM_SOURCEFIELD) ? c->f : c->q;
  for (int level = c->finest; level >= c->minlevel + 1; level--) {
    RETCHK(restrict_t<double>(c, level, fld, fld, nullptr));
    if (field == BSAM_SOLUTIONFIELD) RETCHK(set_ghost_t<double>(c, level - 1, 0, c->q, 0));
  }
  return BSAM_OK;
}

static int32_t source_launch(bsam_ctx* c, int level, int mode) {
  RETCHK(check_level(c, level));
  Level& L = c->L(level);
  FOR_LAUNCHES(c, L, kCell, a, nb) {
    const ModelK m = c->mk();
    if (c->p.model == BSAM_MODEL_CH) {
      DISPATCH_DIM(c->p.ndims,
                   (k_source<2, BSAM_MODEL_CH><<<nb, kBlock, 0, c->s>>>(a, mode, m, c->qold, c->ftmp, c->f)),
                   (k_source<3, BSAM_MODEL_CH><<<nb, kBlock, 0, c->s>>>(a, mode, m, c->qold, c->ftmp, c->f)));
    } else if (c->p.model == BSAM_MODEL_TUMOUR) {
      DISPATCH_DIM(c->p.ndims,
                   (k_source<2, BSAM_MODEL_TUMOUR><<<nb, kBlock, 0, c->s>>>(a, mode, m, c->qold, c->ftmp, c->f)),
                   (k_source<3, BSAM_MODEL_TUMOUR><<<nb, kBlock, 0, c->s>>>(a, mode, m, c->qold, c->ftmp, c->f)));
    } else {
      k_source<2, BSAM_MODEL_BG><<<nb, kBlock, 0, c->s>>>(a, mode, m, c->qold, c->ftmp, c->f);
    }
    LAUNCH_CHECK();
  }
  return BSAM_OK;
}

extern "C" int32_t bsam_get_source(bsam_ctx* c, int32_t level) { return source_launch(c, level, 0); }
extern "C" int32_t bsam_update_source(bsam_ctx* c, int32_t level) { return source_launch(c, level, 1); }
// GetSourceFunction then UpdateSourceFunction on one level in one pass.
static int32_t source_and_update(bsam_ctx* c, int32_t level) { return source_launch(c, level, 2); }

extern "C" int32_t bsam_update_aux(bsam_ctx* c, int32_t level) {
  // maux = 0: UpdateAuxInVcycle returns before calling the hook
  // (afasroutines.f90:417).
  return check_level(c, level);
}

// Two-stage deterministic reduction of nb block partials into d_norm[0..1].
static int32_t reduce_norm(bsam_ctx* c, int nb) {
  const int chunk = 256;
  const int nb2 = (nb + chunk - 1) / chunk;
  if (nb2 > 1) {
    k_reduce_partials<<<nb2, kBlock, 0, c->s>>>(c->d_partial, nb, chunk, c->d_partial2);
    LAUNCH_CHECK();
    k_reduce_partials<<<1, kBlock, 0, c->s>>>(c->d_partial2, nb2, nb2, c->d_norm);
  } else {
    k_reduce_partials<<<1, kBlock, 0, c->s>>>(c->d_partial, nb, nb, c->d_norm);
  }
  LAUNCH_CHECK();
  return BSAM_OK;
}

static int32_t residual_launch(bsam_ctx* c, int level, double* rf, double* partial) {
  Level& L = c->L(level);
  const ModelK m = c->mk();
  int total = 0;
  LevelArgs a;
  int nb = 0;
  for (LaunchIter it(c, L, kCell); it.next(a, nb);) {
    double* part = partial ? partial + (long)kNvMax * it.boff : nullptr;
    if (c->p.model == BSAM_MODEL_CH) {
      DISPATCH_DIM(c->p.ndims,
                   (k_residual<2, BSAM_MODEL_CH, double><<<nb, kBlock, 0, c->s>>>(a, m, c->q, c->qold, c->f, rf, part)),
                   (k_residual<3, BSAM_MODEL_CH, double><<<nb, kBlock, 0, c->s>>>(a, m, c->q, c->qold, c->f, rf, part)));
    } else if (c->p.model == BSAM_MODEL_TUMOUR) {
      DISPATCH_DIM(c->p.ndims,
                   (k_residual<2, BSAM_MODEL_TUMOUR, double><<<nb, kBlock, 0, c->s>>>(a, m, c->q, c->qold, c->f, rf, part)),
                   (k_residual<3, BSAM_MODEL_TUMOUR, double><<<nb, kBlock, 0, c->s>>>(a, m, c->q, c->qold, c->f, rf, part)));
    } else {
      k_residual<2, BSAM_MODEL_BG, double><<<nb, kBlock, 0, c->s>>>(a, m, c->q, c->qold, c->f, rf, part);
    }
    LAUNCH_CHECK();
    total = it.blocks();
  }
  return total;
}

extern "C" int32_t bsam_residual(bsam_ctx* c, int32_t level) {
  RETCHK(check_level(c, level));
  RETCHK(residual_launch(c, level, c->rf, nullptr));
  return BSAM_OK;
}

// Enqueue the residual norm of `level` into d_norm[0..1] (no wait).
// Marching residual for a single large 3D CH patch; returns the number of
// partial sums written, or 0 when the level does not qualify.
template <typename TO>
static int resid_march(bsam_ctx* c, int level, TO* out, double* partial, float* base = nullptr) {
  Level& L = c->L(level);
  int nblk = 0;
  if (c->p.model != BSAM_MODEL_CH) return 0;
  const int kc = march_kc(c, L, &nblk);
  if (!kc) return 0;
  const double h2 = L.h * L.h;
  k_ch3d_resid_march<0, TO><<<nblk, 512, 0, c->s>>>(c->d_pt, L.patches[0], kc, c->p.dt / h2,
                                                     c->p.model_params[0] / h2, c->mk(), c->q, c->f, out, partial,
                                                     nullptr, nullptr, base);
  if (hipGetLastError() != hipSuccess) return -1;
  return nblk;
}

static int32_t enqueue_norm(bsam_ctx* c, int level, bool write_rf) {
  const bool fin = level == c->finest;
  if (fin) kstat_begin(c, BSAM_KSTAT_RESID_FINEST);
  int nb = resid_march<double>(c, level, write_rf ? c->rf : nullptr, c->d_partial);
  if (nb < 0) return set_err(BSAM_EHIP, "residual kernel launch failed");
  if (nb == 0) nb = residual_launch(c, level, write_rf ? c->rf : nullptr, c->d_partial);
  if (nb < 0) return nb;
  if (fin) {
    double N = 0, G = 0;
    for (int pid : c->L(level).patches) {
      const DPatch& P = c->hp[pid];
      N += (double)P.mx[0] * P.mx[1] * P.mx[2];
      G += (c->p.ndims == 3) ? 2.0 * ((double)P.mx[0] * P.mx[1] + (double)P.mx[1] * P.mx[2] + (double)P.mx[0] * P.mx[2])
                             : 2.0 * ((double)P.mx[0] + P.mx[1]);
    }
    kstat_end(c, BSAM_KSTAT_RESID_FINEST, 8.0 * 2.0 * ((N + G) + N + (write_rf ? N : 0.0)));
  }
  return reduce_norm(c, nb);
}

static double norm_from(bsam_ctx* c, const double* s, int level) {
  double cells = 0;
  for (int pid : c->L(level).patches) {
    const DPatch& P = c->hp[pid];
    cells += (double)P.mx[0] * P.mx[1] * P.mx[2];
  }
  const int nv = c->p.nrvars;
  if (c->p.errortype == 1) {
    double t = 0.0;
    for (int v = 0; v < nv; v++) t += s[v];
    return std::sqrt(t / (nv * cells));
  }
  double e = 0.0;
  for (int v = 0; v < nv; v++) e = std::max(e, std::sqrt(s[v] / cells));
  return e;
}

extern "C" int32_t bsam_error_afas(bsam_ctx* c, int32_t level, double* norm) {
  RETCHK(check_level(c, level));
  if (!norm) return set_err(BSAM_EINVAL, "null norm");
  RETCHK(enqueue_norm(c, level, true));
  HIPCHK(hipMemcpyAsync(c->h_norm, c->d_norm, kNvMax * sizeof(double), hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  *norm = norm_from(c, c->h_norm, level);
  return BSAM_OK;
}

extern "C" int32_t bsam_copy_q_to_qold(bsam_ctx* c, int32_t level) {
  RETCHK(check_level(c, level));
  Level& L = c->L(level);
  if (L.qspan_len > 0) {
    // one copy of the level's span covers every variable of every patch
    k_copy_span<<<nblocks(L.qspan_len), kBlock, 0, c->s>>>(c->q + L.qspan_off, c->qold + L.qspan_off, L.qspan_len);
    LAUNCH_CHECK();
  } else {
    for (int pid : L.patches) {
      const DPatch& P = c->hp[pid];
      const long n = (long)P.qvs * c->p.nrvars;
      k_copy_span<<<nblocks(n), kBlock, 0, c->s>>>(c->q + P.qoff, c->qold + P.qoff, n);
      LAUNCH_CHECK();
    }
  }
  return BSAM_OK;
}

// The frame writer's F25.12 edit of n values on the host (25 bytes each):
// the same function the device runs, for checking without a GPU.
extern "C" int32_t bsam_selftest_f25(const double* x, int64_t n, char* out) {
  if (!x || !out || n < 0) return set_err(BSAM_EINVAL, "null argument");
  for (int64_t e = 0; e < n; e++) f25_edit(x[e], out + 25 * e);
  return BSAM_OK;
}

// Iw edit: right-justified in w characters, w asterisks when it does not fit.
static std::string frame_int(int v, int w) {
  char b[32];
  snprintf(b, sizeof b, "%*d", w, v);
  std::string t(b);
  return (int)t.size() > w ? std::string(w, '*') : t;
}
static std::string frame_f25(double x) {
  char b[25];
  f25_edit(x, b);
  return std::string(b, 25);
}

// WriteQ (bsaminputoutput.f90:60-82) with OutputQ (:85-129) and
// WriteQ2D/3D (:132-181): the header is edited on the host, the patch
// bodies on the device in chunks of whole patches (at most 256 MB, or one
// patch), copied back through pinned memory and written in level order.
extern "C" int32_t bsam_write_frame(bsam_ctx* c, const char* path, double time, const double* geom) {
  if (!c || !path) return set_err(BSAM_EINVAL, "null argument");
  if (!c->have_hier) return set_err(BSAM_ESTATE, "no hierarchy installed (bsam_set_hierarchy)");
  const int nd = c->p.ndims, nv = c->p.nrvars;
  const long len = 27L * nv - 1;
  std::vector<int> order;  // patches of levels 0..finest in level order
  for (int level = 0; level <= c->finest; level++)
    for (int pid : c->L(level).patches) order.push_back(pid);
  auto cells = [&](int pid) {
    const DPatch& P = c->hp[pid];
    return (long)P.nq[0] * P.nq[1] * (nd == 3 ? P.nq[2] : 1);
  };
  long cap = 256L << 20;
  for (int pid : order) cap = std::max(cap, cells(pid) * len);
  FILE* fh = fopen(path, "wb");
  if (!fh) return set_err(BSAM_EINVAL, "cannot open %s for writing", path);
  char* d_out = nullptr;
  char* h_out = nullptr;
  FramePatch* d_fp = nullptr;
  int32_t rc = BSAM_OK;
  auto fail = [&](int32_t r) {
    rc = r;
    return r;
  };
  do {
    if (hipMalloc(&d_out, cap) != hipSuccess || hipHostMalloc(&h_out, cap) != hipSuccess ||
        hipMalloc(&d_fp, order.size() * sizeof(FramePatch)) != hipSuccess) {
      fail(set_err(BSAM_EHIP, "frame buffers of %ld bytes could not be allocated", cap));
      break;
    }
    std::string head = frame_f25(time) + "\n" + frame_int(c->finest, 3) + "\n";
    if (fwrite(head.data(), 1, head.size(), fh) != head.size()) {
      fail(set_err(BSAM_EINVAL, "cannot write %s", path));
      break;
    }
    for (size_t a0 = 0; a0 < order.size() && rc == BSAM_OK;) {
      // one chunk of whole patches
      std::vector<FramePatch> fp;
      long ncell = 0, nbyte = 0;
      size_t a1 = a0;
      while (a1 < order.size() && (fp.empty() || nbyte + cells(order[a1]) * len <= cap)) {
        fp.push_back(FramePatch{order[a1], ncell, nbyte});
        ncell += cells(order[a1]);
        nbyte += cells(order[a1]) * len;
        a1++;
      }
      if (hipMemcpyAsync(d_fp, fp.data(), fp.size() * sizeof(FramePatch), hipMemcpyHostToDevice, c->s) !=
          hipSuccess) {
        fail(set_err(BSAM_EHIP, "frame: patch table upload failed"));
        break;
      }
      if (nd == 2) k_frame_body<2><<<nblocks(ncell), kBlock, 0, c->s>>>(c->d_pt, d_fp, (int)fp.size(), ncell, nv, c->q, d_out);
      else k_frame_body<3><<<nblocks(ncell), kBlock, 0, c->s>>>(c->d_pt, d_fp, (int)fp.size(), ncell, nv, c->q, d_out);
      if (hipGetLastError() != hipSuccess ||
          hipMemcpyAsync(h_out, d_out, nbyte, hipMemcpyDeviceToHost, c->s) != hipSuccess ||
          hipStreamSynchronize(c->s) != hipSuccess) {
        fail(set_err(BSAM_EHIP, "frame: body kernel or copy failed"));
        break;
      }
      for (const FramePatch& F : fp) {
        const int pid = F.pid;
        const DPatch& P = c->hp[pid];
        const bsam_patch_desc& D = c->desc[pid];
        double g[9];
        for (int x = 0; x < 3; x++) {
          g[x] = geom ? geom[9L * pid + x] : D.dx;
          g[3 + x] = geom ? geom[9L * pid + 3 + x] : D.xlower[x];
          g[6 + x] = geom ? geom[9L * pid + 6 + x] : D.xlower[x] + P.mx[x] * D.dx;
        }
        std::string h = "  \n" + frame_int(D.level, 3) + " " + frame_int(nd, 3) + " " + frame_int(2, 3) + " " +
                        frame_int(nv, 3) + "\n";
        for (int row = 0; row < 3; row++) {
          for (int x = 0; x < nd; x++) h += (x ? " " : "") + frame_f25(g[3 * row + x]);
          h += "\n";
        }
        for (int x = 0; x < nd; x++) h += (x ? " " : "") + frame_int(P.mx[x], 8);
        h += "\n";
        for (int x = 0; x < nd; x++)
          for (int e = 0; e < 2; e++) h += ((x || e) ? " " : "") + frame_int(D.mglobal[x][e], 8);
        h += "\n";
        const size_t nb = (size_t)(cells(pid) * len);
        if (fwrite(h.data(), 1, h.size(), fh) != h.size() || fwrite(h_out + F.byte0, 1, nb, fh) != nb) {
          fail(set_err(BSAM_EINVAL, "cannot write %s", path));
          break;
        }
      }
      a0 = a1;
    }
  } while (false);
  hipFree(d_out);
  hipFree(d_fp);
  if (h_out) hipHostFree(h_out);
  if (fclose(fh) != 0 && rc == BSAM_OK) rc = set_err(BSAM_EINVAL, "cannot write %s", path);
  return rc;
}

// EstimateLevelErrors (bsamroutines.f90:994-1040) and the tagging of
// ErrFlag (:1059-1131); BufferAndList, InflateEdgeTags and the new grids
// stay with the caller.  errflagopt 1 runs RestrictSolution, SetGhost(L-1,
// 0), GetCoarseGhostPoints and CopyQToQold on L and L-1 (the reference's
// order, with their effect on q, qc and qold), then the truncation-error
// flags; errflagopt 10 applies the second-difference test to q as it is.
static int32_t check_tag_args(bsam_ctx* c, int32_t level, int32_t errflagopt, const int32_t* flags,
                              const int64_t* ntagged) {
  if (!c || !flags || !ntagged) return set_err(BSAM_EINVAL, "null argument");
  RETCHK(check_level(c, level));
  if (level <= c->minlevel) return set_err(BSAM_EINVAL, "level %d has no coarser level to estimate against", level);
  if (errflagopt != 1 && errflagopt != 10 && errflagopt != 11)
    return set_err(BSAM_EINVAL,
                   "errflagopt must be 1 (errflagdefault), 10 (errflaguser) or 11 (gradient), not %d", errflagopt);
  // errflaguser is the second-difference test of the CH and BioGrowth
  // modules; the tumour model's user tagging is the gradient test (11)
  if (errflagopt == 10 && c->p.model == BSAM_MODEL_TUMOUR)
    return set_err(BSAM_EINVAL,
                   "errflagopt 10 (errflaguser) is not implemented for the tumour model: its user tagging is "
                   "the gradient test, errflagopt 11");
  return BSAM_OK;
}

// The tagging of one level with the level's flags staged in `d_flags`
// (device memory of at least L.ncell int32); shared by the entry point that
// grows a library buffer and the one that uses the fixed scratch array.
// The per-patch flag offsets were uploaded with the hierarchy.
static int32_t estimate_level_errors_in(bsam_ctx* c, int32_t level, int32_t errflagopt, double qtolerance,
                                        int32_t* flags, int64_t* ntagged, int32_t* d_flags) {
  Level& L = c->L(level);
  const int nd = c->p.ndims;
  const long n = L.ncell;
  const long* d_foff = L.d_foff;
  if (errflagopt == 1) {
    RETCHK(bsam_restrict_solution(c, level));
    RETCHK(bsam_set_ghost(c, level - 1, 0));
    RETCHK(bsam_coarse_ghost_points(c, level));
    RETCHK(bsam_copy_q_to_qold(c, level));
    RETCHK(bsam_copy_q_to_qold(c, level - 1));
    if (n > 0) {
      k_zero_i32<<<nblocks(n), kBlock, 0, c->s>>>(d_flags, n);
      LAUNCH_CHECK();
    }
    const double h = c->hp[L.patches[0]].dx;
    const double tol = nd == 3 ? qtolerance / h / h / h : qtolerance / h / h;
    const ModelK m = c->mk();
    FOR_LAUNCHES(c, L, kCoarse, a, nb) {
      if (c->p.model == BSAM_MODEL_CH)
        DISPATCH_DIM(nd, (k_rte_flags<2, BSAM_MODEL_CH><<<nb, kBlock, 0, c->s>>>(a, m, c->q, c->qold, c->qc, tol, d_foff, d_flags)),
                     (k_rte_flags<3, BSAM_MODEL_CH><<<nb, kBlock, 0, c->s>>>(a, m, c->q, c->qold, c->qc, tol, d_foff, d_flags)));
      else if (c->p.model == BSAM_MODEL_TUMOUR)
        DISPATCH_DIM(nd, (k_rte_flags<2, BSAM_MODEL_TUMOUR><<<nb, kBlock, 0, c->s>>>(a, m, c->q, c->qold, c->qc, tol, d_foff, d_flags)),
                     (k_rte_flags<3, BSAM_MODEL_TUMOUR><<<nb, kBlock, 0, c->s>>>(a, m, c->q, c->qold, c->qc, tol, d_foff, d_flags)));
      else
        k_rte_flags<2, BSAM_MODEL_BG><<<nb, kBlock, 0, c->s>>>(a, m, c->q, c->qold, c->qc, tol, d_foff, d_flags);
      LAUNCH_CHECK();
    }
  } else if (errflagopt == 10) {
    const double thr = (nd == 3 ? 3.0 : 2.0) * qtolerance;
    FOR_LAUNCHES(c, L, kCell, a, nb) {
      DISPATCH_DIM(nd, (k_user_flags<2><<<nb, kBlock, 0, c->s>>>(a, c->q, thr, d_foff, d_flags)),
                   (k_user_flags<3><<<nb, kBlock, 0, c->s>>>(a, c->q, thr, d_foff, d_flags)));
      LAUNCH_CHECK();
    }
  } else {
    FOR_LAUNCHES(c, L, kCell, a, nb) {
      DISPATCH_DIM(nd, (k_user_flags<2, true><<<nb, kBlock, 0, c->s>>>(a, c->q, qtolerance, d_foff, d_flags)),
                   (k_user_flags<3, true><<<nb, kBlock, 0, c->s>>>(a, c->q, qtolerance, d_foff, d_flags)));
      LAUNCH_CHECK();
    }
  }
  // without `flags` the level's flags stay on the device
  if (!flags) return BSAM_OK;
  if (n > 0) HIPCHK(hipMemcpyAsync(flags, d_flags, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  int64_t nt = 0;
  for (long e = 0; e < n; e++) nt += flags[e] != 0;
  *ntagged = nt;
  return BSAM_OK;
}

// The tags AMR clusters on one level: EstimateLevelErrors into device
// flags, buffered by ibuffer on the device where the level's bit map fits
// (k_tag_bits, k_tag_dilate: at most 2^31 cells, ibuffer below every
// extent), then copied to `flags` (L.ncell entries).  *buffered tells the
// caller whether the buffering is done (else bsam_new_level buffers).
static int32_t amr_level_tags(bsam_ctx* c, int level, int errflagopt, double qtolerance, int ibuffer,
                              int32_t*& flags, bool& buffered) {
  Level& L = c->L(level);
  const int nd = c->p.ndims;
  const long n = L.ncell;
  // the flags come back through a pinned host buffer of the context
  if ((size_t)std::max<long>(1, n) > c->h_flags_cap) {
    HIPCHK(hipStreamSynchronize(c->s));
    if (c->h_flags) hipHostFree(c->h_flags);
    c->h_flags = nullptr;
    c->h_flags_cap = 0;
    HIPCHK(hipHostMalloc(&c->h_flags, std::max<long>(1, n) * sizeof(int32_t), hipHostMallocDefault));
    c->h_flags_cap = std::max<long>(1, n);
  }
  flags = c->h_flags;
  buffered = false;
  const bsam_patch_desc* root = nullptr;
  for (const bsam_patch_desc& d : c->desc)
    if (d.level == 0) { root = &d; break; }
  long ext[3] = {1, 1, 1};
  bool fits = ibuffer > 0 && root && level >= 0;
  for (int a = 0; a < nd && fits; a++) {
    ext[a] = (long)root->mx[a] << level;
    fits = ibuffer < ext[a];
  }
  const long nbits = ext[0] * ext[1] * ext[2];
  fits = fits && nbits <= (1L << 31);
  for (size_t a = 0; a < L.patches.size() && fits; a++)
    for (int d = 0; d < nd; d++) {
      const bsam_patch_desc& D = c->desc[L.patches[a]];
      fits = fits && D.mglobal[d][0] >= 1 && D.mglobal[d][1] <= ext[d] &&
             D.mglobal[d][1] - D.mglobal[d][0] + 1 == D.mx[d];
    }
  const size_t words = fits ? (size_t)((nbits + 31) / 32) : 0;
  const size_t need = (size_t)n * sizeof(int32_t) * (fits ? 2 : 1);
  if (need > c->tag_cap) {
    HIPCHK(hipStreamSynchronize(c->s));
    hipFree(c->d_tag);
    c->d_tag = nullptr;
    c->tag_cap = 0;
    HIPCHK(hipMalloc(&c->d_tag, need));
    c->tag_cap = need;
  }
  int32_t* d_flags = reinterpret_cast<int32_t*>(c->d_tag);
  RETCHK(estimate_level_errors_in(c, level, errflagopt, qtolerance, nullptr, nullptr, d_flags));
  if (n == 0) return BSAM_OK;
  if (fits) {
    if (words > c->tagbits_cap) {
      HIPCHK(hipStreamSynchronize(c->s));
      hipFree(c->d_tagbits);
      c->d_tagbits = nullptr;
      c->tagbits_cap = 0;
      HIPCHK(hipMalloc(&c->d_tagbits, words * sizeof(uint32_t)));
      c->tagbits_cap = words;
    }
    if (L.patches.size() * 3 > c->tagglo_cap) {
      HIPCHK(hipStreamSynchronize(c->s));
      hipFree(c->d_tagglo);
      c->d_tagglo = nullptr;
      c->tagglo_cap = 0;
      HIPCHK(hipMalloc(&c->d_tagglo, L.patches.size() * 3 * sizeof(int)));
      c->tagglo_cap = L.patches.size() * 3;
    }
    std::vector<int> glo(L.patches.size() * 3, 0);
    for (size_t a = 0; a < L.patches.size(); a++)
      for (int d = 0; d < nd; d++) glo[3 * a + d] = c->desc[L.patches[a]].mglobal[d][0] - 1;
    HIPCHK(hipMemcpyAsync(c->d_tagglo, glo.data(), glo.size() * sizeof(int), hipMemcpyHostToDevice, c->s));
    k_zero_i32<<<nblocks((long)words), kBlock, 0, c->s>>>(reinterpret_cast<int32_t*>(c->d_tagbits), (long)words);
    LAUNCH_CHECK();
    int32_t* d_out = d_flags + n;
    const int wrap[3] = {root->mthbc[0] == 2, nd > 1 && root->mthbc[2] == 2, nd > 2 && root->mthbc[4] == 2};
    for (int pass = 0; pass < 2; pass++) {
      FOR_LAUNCHES(c, L, kCell, a, nb) {
        if (pass == 0) {
          DISPATCH_DIM(nd,
                       (k_tag_bits<2><<<nb, kBlock, 0, c->s>>>(a, L.d_foff, c->d_tagglo, d_flags, c->d_tagbits,
                                                               (int)ext[0], (int)ext[1])),
                       (k_tag_bits<3><<<nb, kBlock, 0, c->s>>>(a, L.d_foff, c->d_tagglo, d_flags, c->d_tagbits,
                                                               (int)ext[0], (int)ext[1])));
        } else {
          DISPATCH_DIM(nd,
                       (k_tag_dilate<2><<<nb, kBlock, 0, c->s>>>(a, L.d_foff, c->d_tagglo, c->d_tagbits, (int)ext[0],
                                                                 (int)ext[1], (int)ext[2], wrap[0], wrap[1], wrap[2],
                                                                 ibuffer, d_out)),
                       (k_tag_dilate<3><<<nb, kBlock, 0, c->s>>>(a, L.d_foff, c->d_tagglo, c->d_tagbits, (int)ext[0],
                                                                 (int)ext[1], (int)ext[2], wrap[0], wrap[1], wrap[2],
                                                                 ibuffer, d_out)));
        }
        LAUNCH_CHECK();
      }
    }
    d_flags = d_out;
    buffered = true;
  }
  HIPCHK(hipMemcpyAsync(flags, d_flags, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  return BSAM_OK;
}

extern "C" int32_t bsam_estimate_level_errors(bsam_ctx* c, int32_t level, int32_t errflagopt, double qtolerance,
                                              int32_t* flags, int64_t* ntagged) {
  RETCHK(check_tag_args(c, level, errflagopt, flags, ntagged));
  const size_t need = (size_t)c->L(level).ncell * sizeof(int32_t);
  if (need > c->tag_cap) {
    HIPCHK(hipStreamSynchronize(c->s));
    hipFree(c->d_tag);
    c->d_tag = nullptr;
    c->tag_cap = 0;
    HIPCHK(hipMalloc(&c->d_tag, need));
    c->tag_cap = need;
  }
  return estimate_level_errors_in(c, level, errflagopt, qtolerance, flags, ntagged,
                                  reinterpret_cast<int32_t*>(c->d_tag));
}

extern "C" int32_t bsam_estimate_level_errors_noalloc(bsam_ctx* c, int32_t level, int32_t errflagopt,
                                                      double qtolerance, int32_t* flags, int64_t* ntagged) {
  RETCHK(check_tag_args(c, level, errflagopt, flags, ntagged));
  if (!c->scratch) return set_err(BSAM_ESTATE, "the scratch array of device %d has no address", c->dev);
  const long n = c->L(level).ncell;
  if ((size_t)n * sizeof(int32_t) > kScratchBytes)
    return set_err(BSAM_EINVAL,
                   "level %d has %ld cells; the allocation-free tagging call stages at most %zu flags "
                   "(%zu bytes of scratch): use bsam_estimate_level_errors",
                   level, n, kScratchBytes / sizeof(int32_t), kScratchBytes);
  return estimate_level_errors_in(c, level, errflagopt, qtolerance, flags, ntagged,
                                  reinterpret_cast<int32_t*>(c->scratch));
}

// ---------------------------------------------------------------------------
// V-cycle and MultigridIterations
// ---------------------------------------------------------------------------
static int32_t launch_coarse_chain(bsam_ctx* c, int level, bool lin) {
  Level& L = c->L(level);
  if (c->p.model == BSAM_MODEL_TUMOUR) {
    // the chain's levels were set by plan_chain for this level
    if (!c->d_tmchain || c->h_tmchain.nlev != L.chain || c->h_tmchain.pid[0] != L.patches[0])
      return set_err(BSAM_ESTATE, "level %d: coarse chain not planned", level);
    const ModelK mk = c->mk();
    const int nsm = c->p.nsmoothingpasses;
    const TmChainArgs* A = c->d_tmchain;
    if (!lin) {
      DISPATCH_DIM(c->p.ndims,
                   (k_coarse_vcycle_tm<2, false><<<1, 1024, 0, c->s>>>(c->d_pt, A, mk, nsm, c->q, c->f, c->qc, c->qold, nullptr)),
                   (k_coarse_vcycle_tm<3, false><<<1, 1024, 0, c->s>>>(c->d_pt, A, mk, nsm, c->q, c->f, c->qc, c->qold, nullptr)));
    } else {
      bsam::MixedState* m = c->mixed;
      DISPATCH_DIM(c->p.ndims,
                   (k_coarse_vcycle_tm<2, true><<<1, 1024, 0, c->s>>>(c->d_pt, A, mk, nsm, m->e, m->rhs, m->ec, c->qold, m->coef)),
                   (k_coarse_vcycle_tm<3, true><<<1, 1024, 0, c->s>>>(c->d_pt, A, mk, nsm, m->e, m->rhs, m->ec, c->qold, m->coef)));
    }
    LAUNCH_CHECK();
    return BSAM_OK;
  }
  CoarseArgs A;
  A.nlev = L.chain;
  for (int n = 0; n < L.chain; n++) A.pid[n] = c->L(level - n).wrap_pid;
  A.dt = c->p.dt;
  A.eps2 = c->p.model_params[0];
  const int nsm = c->p.nsmoothingpasses;
  if (!lin) {
    DISPATCH_DIM(c->p.ndims,
                 (k_coarse_vcycle<2, double, false><<<1, 1024, 0, c->s>>>(c->d_pt, A, nsm, c->q, c->f, c->qc, nullptr)),
                 (k_coarse_vcycle<3, double, false><<<1, 1024, 0, c->s>>>(c->d_pt, A, nsm, c->q, c->f, c->qc, nullptr)));
  } else {
    bsam::MixedState* m = c->mixed;
    DISPATCH_DIM(c->p.ndims,
                 (k_coarse_vcycle<2, float, true><<<1, 1024, 0, c->s>>>(c->d_pt, A, nsm, m->e, m->rhs, m->ec, m->coef)),
                 (k_coarse_vcycle<3, float, true><<<1, 1024, 0, c->s>>>(c->d_pt, A, nsm, m->e, m->rhs, m->ec, m->coef)));
  }
  LAUNCH_CHECK();
  return BSAM_OK;
}

extern "C" int32_t bsam_vcycle(bsam_ctx* c, int32_t level) {
  RETCHK(check_level(c, level));
  if (c->L(level).chain > 0) return launch_coarse_chain(c, level, false);
  RETCHK(level_relax_t(c, level, level > c->minlevel));
  if (level > c->minlevel) {
    // RestrictSolution + SetGhost(level-1, 0), fused when the parent is a
    // fully covered periodic patch
    const bool rw = restrict_wraps_parent(c, level);
    RETCHK(restrict_t<double>(c, level, c->q, c->q, c->qc, rw));
    if (!rw) RETCHK(bsam_set_ghost(c, level - 1, 0));
    RETCHK(bsam_coarse_ghost_points(c, level));
    RETCHK(bsam_coarse_loading(c, level));
    RETCHK(bsam_vcycle(c, level - 1));
    // CorrectFine + SetGhost(level, 0), fused for a periodic single patch
    const bool cw = c->L(level).wrap_pid >= 0;
    RETCHK(correct_t<double>(c, level, c->q, c->qc, cw));
    // SetGhost(level, 0); the post-smoothing sweeps that follow read only
    // the reduced table's ghosts before rewriting the rest
    // (a level whose sweeps write their ghost images needs every ghost
    // consistent before them: the full call)
    if (!cw)
      RETCHK(set_ghost_t<double>(c, level, 0, c->q, 0,
                                 phys_images_level(c, c->L(level)) ? kFull : kAfterCorrection));
    RETCHK(bsam_level_relax(c, level));
  }
  return BSAM_OK;
}

// Loop body of MultigridIterations after the V-cycle (afasroutines.f90:
// 83-104), ending with the residual norm enqueued in d_norm.
//
// The reference's loop body also calls UpdateSourceFunction on every level
// >= root and FillDown(sourcefield) (afasroutines.f90:96-100).  With
// SourceUpdate#D = 0 (both the CH and the BG hooks, CH/problem.f90:453-476,
// BioGrowth/problem.f90:382-405) those stores are dead: the only f values
// the V-cycle changes are parent cells under a child, written by
// CoarseLoadingFunction, which rewrites exactly those cells before any read
// in the next V-cycle; every other f value already equals ftmp + 0.  They
// are therefore done once after the loop (so f leaves the call as the
// reference leaves it) instead of once per V-cycle.
//
// FillDown(solutionfield) (afasroutines.f90:83) is needed inside the loop
// only for levels >= 1: it refreshes the parent cells that the next
// V-cycle's coarse-fine ghost interpolation of the child reads.  Levels at
// and below the root have no interpolation, and every parent value the
// restriction would write there is overwritten by RestrictSolution and
// SetGhost of the next V-cycle before it is read; those levels are filled
// down once after the loop.
static int32_t enqueue_iteration(bsam_ctx* c) {
  RETCHK(bsam_vcycle(c, c->finest));
  RETCHK(fill_down_range(c, BSAM_SOLUTIONFIELD, c->finest, 1));
  RETCHK(enqueue_norm(c, c->finest, false));
  return BSAM_OK;
}

static bool kstat_any(bsam_ctx* c) {
  for (auto& k : c->ks)
    if (k.on) return true;
  return false;
}

// The iteration body is captured as a graph only with per-level launches:
// one launch per patch (batching = 0) is the plain per-patch mode, whose
// launch count grows with the patch count (tens of thousands per V-cycle on
// large AMR hierarchies) and is not captured.
// One loop body on the context's stream, then the norm's copy to the host.
// With use_graph the body is captured once and replayed; `eager` runs it
// without the graph (the context keeps one graph, that of its own path).
// On a mesh that bsam_amr built the body always runs eagerly: a host-side
// segmentation fault inside hipGraphLaunch was seen on such meshes (FP64,
// the graph captured again after every regrid) whose cause is not known,
// so that combination is not used.
static int32_t run_iteration(bsam_ctx* c, int32_t (*body)(bsam_ctx*), bool eager = false) {
  if (c->p.use_graph && c->p.batching && !kstat_any(c) && !eager && !c->amr_mesh) {
    if (!c->gexec) {
      PhaseClock pc(c, BSAM_PHASE_GRAPH);
      HIPCHK(hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
      c->capturing = true;
      int32_t r = body(c);
      c->capturing = false;
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(c->s, &g);
      if (r < 0) {
        if (g) hipGraphDestroy(g);
        return r;
      }
      if (e != hipSuccess) return set_err(BSAM_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
      c->graph = g;
      HIPCHK(hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
    }
    HIPCHK(hipGraphLaunch(c->gexec, c->s));
  } else {
    RETCHK(body(c));
  }
  HIPCHK(hipMemcpyAsync(c->h_norm, c->d_norm, kNvMax * sizeof(double), hipMemcpyDeviceToHost, c->s));
  return BSAM_OK;
}

extern "C" int32_t bsam_multigrid_iterations(bsam_ctx* c, int32_t* ncycles, double* residual, double* history,
                                             int32_t history_len) {
  if (!c || !ncycles || !residual) return set_err(BSAM_EINVAL, "null argument");
  if (!c->have_hier) return set_err(BSAM_ESTATE, "no hierarchy installed");
  if (c->p.precision == BSAM_PREC_MIXED) return bsam::mixed_multigrid_iterations(c, ncycles, residual, history, history_len);
  RETCHK(bsam_fill_down(c, BSAM_SOLUTIONFIELD));
  for (int level = c->finest; level >= c->minlevel; level--) RETCHK(bsam_update_aux(c, level));
  for (int level = c->finest; level >= 0; level--) RETCHK(source_and_update(c, level));
  RETCHK(bsam_fill_down(c, BSAM_SOURCEFIELD));
  double res = 0.0;
  int it;
  for (it = 1; it <= c->p.maxvcycles; it++) {
    RETCHK(run_iteration(c, enqueue_iteration));
    HIPCHK(hipStreamSynchronize(c->s));
    res = norm_from(c, c->h_norm, c->finest);
    if (history && it - 1 < history_len) history[it - 1] = res;
    if (!std::isfinite(res)) {
      *ncycles = it;
      *residual = res;
      return set_err(BSAM_EDIVERGED, "residual not finite after V-cycle %d", it);
    }
    if (res < c->p.qerrortol) break;
  }
  // the finest level's f is never written by the V-cycle
  for (int level = c->finest - 1; level >= 0; level--) RETCHK(bsam_update_source(c, level));
  RETCHK(bsam_fill_down(c, BSAM_SOURCEFIELD));
  // the part of the last FillDown(solutionfield) skipped in the loop
  RETCHK(fill_down_range(c, BSAM_SOLUTIONFIELD, 0, c->minlevel + 1));
  *ncycles = it > c->p.maxvcycles ? c->p.maxvcycles : it;
  *residual = res;
  return it > c->p.maxvcycles ? BSAM_NOT_CONVERGED : BSAM_OK;
}

extern "C" int32_t bsam_time_step(bsam_ctx* c, int32_t first_step, int32_t* ncycles, double* residual) {
  if (!c || !c->have_hier) return set_err(BSAM_ESTATE, "no hierarchy installed");
  // the ghost fills of AMR(rootlevel) on an unchanged mesh: SetGhost(level,
  // 0) from the root up (bsamroutines.f90:978-989); its tagging changes
  // nothing the step reads (RestrictSolution rewrites the values the last
  // FillDown left, and qold, qc are rewritten below and in the V-cycle)
  for (int level = 0; level <= c->finest; level++) RETCHK(bsam_set_ghost(c, level, 0));
  PhaseClock pc(c, BSAM_PHASE_SOLVE);
  if (first_step) RETCHK(bsam_fill_down(c, BSAM_SOLUTIONFIELD));
  for (int level = c->minlevel; level <= c->finest; level++) RETCHK(bsam_copy_q_to_qold(c, level));
  return bsam_multigrid_iterations(c, ncycles, residual, nullptr, 0);
}

// ---------------------------------------------------------------------------
// regrid: GridAdapt's field set-up (bsamroutines.f90:1663-1951) and AMR
// ---------------------------------------------------------------------------
namespace {
// The levels at and below the root keep their one patch each, with the
// same extents: q and qold come over from the old pools.
int32_t carry_root_levels(bsam_ctx* c, const OldMesh& o) {
  const int nv = c->p.nrvars;
  for (int lvl = c->minlevel; lvl <= 0; lvl++) {
    const DPatch& P = c->hp[c->L(lvl).patches[0]];
    const DPatch& O = o.hp[(*o.patches(lvl))[0]];
    const size_t n = (size_t)P.qvs * nv * sizeof(double);
    HIPCHK(hipMemcpyAsync(c->q + P.qoff, o.q + O.qoff, n, hipMemcpyDeviceToDevice, c->s));
    HIPCHK(hipMemcpyAsync(c->qold + P.qoff, o.qold + O.qoff, n, hipMemcpyDeviceToDevice, c->s));
  }
  return BSAM_OK;
}

// A level of new patches above the root: InitFields' prolongation of the
// parents with qold = q (bsamroutines.f90:1787-1856), then TransferValues'
// copy of q from the old mesh's patches of the level over the interior
// overlap of their global boxes (:1859-1951).  f and rf stay zero
// (MakeNewGrid).
int32_t init_level(bsam_ctx* c, int lvl, OldMesh& o) {
  const int nv = c->p.nrvars, nd = c->p.ndims;
  Level& L = c->L(lvl);
  {
    FOR_LAUNCHES(c, L, kQcBox, a, nb) {
      DISPATCH_DIM(nd, (k_correct_qc<2, double, true><<<nb, kBlock, 0, c->s>>>(a, nv, c->q, c->qc)),
                   (k_correct_qc<3, double, true><<<nb, kBlock, 0, c->s>>>(a, nv, c->q, c->qc)));
      LAUNCH_CHECK();
    }
  }
  {
    FOR_LAUNCHES(c, L, kCoarse, a, nb) {
      DISPATCH_DIM(nd, (k_init_fields<2, double><<<nb, kBlock, 0, c->s>>>(a, nv, c->q, c->qold, c->qc)),
                   (k_init_fields<3, double><<<nb, kBlock, 0, c->s>>>(a, nv, c->q, c->qold, c->qc)));
      LAUNCH_CHECK();
    }
  }
  const std::vector<int>* old = o.patches(lvl);
  if (!old) return BSAM_OK;
  // the old patches each new patch overlaps, through a grid of bins over
  // the level's global index space (the patches of a level are disjoint)
  constexpr int kBin = 16;
  int nb[3] = {1, 1, 1}, lo0[3] = {1, 1, 1};
  {
    int hi0[3] = {1, 1, 1};
    bool first = true;
    for (int op : *old)
      for (int a = 0; a < nd; a++) {
        lo0[a] = first ? o.desc[op].mglobal[a][0] : std::min(lo0[a], o.desc[op].mglobal[a][0]);
        hi0[a] = first ? o.desc[op].mglobal[a][1] : std::max(hi0[a], o.desc[op].mglobal[a][1]);
        if (a == nd - 1) first = false;
      }
    for (int a = 0; a < nd; a++) nb[a] = (hi0[a] - lo0[a]) / kBin + 1;
  }
  std::vector<std::vector<int>> bins((size_t)nb[0] * nb[1] * nb[2]);
  auto bin_of = [&](int a, int g) { return std::min(nb[a] - 1, std::max(0, (g - lo0[a]) / kBin)); };
  for (int op : *old) {
    int b0[3] = {0, 0, 0}, b1[3] = {0, 0, 0};
    for (int a = 0; a < nd; a++) b0[a] = bin_of(a, o.desc[op].mglobal[a][0]), b1[a] = bin_of(a, o.desc[op].mglobal[a][1]);
    for (int z = b0[2]; z <= b1[2]; z++)
      for (int y = b0[1]; y <= b1[1]; y++)
        for (int x = b0[0]; x <= b1[0]; x++) bins[((size_t)z * nb[1] + y) * nb[0] + x].push_back(op);
  }
  std::vector<int> cand;
  std::vector<XferBox> boxes;
  long ncell = 0;
  for (int pid : L.patches) {
    cand.clear();
    int b0[3] = {0, 0, 0}, b1[3] = {0, 0, 0};
    for (int a = 0; a < nd; a++) b0[a] = bin_of(a, c->desc[pid].mglobal[a][0]), b1[a] = bin_of(a, c->desc[pid].mglobal[a][1]);
    for (int z = b0[2]; z <= b1[2]; z++)
      for (int y = b0[1]; y <= b1[1]; y++)
        for (int x = b0[0]; x <= b1[0]; x++)
          for (int op : bins[((size_t)z * nb[1] + y) * nb[0] + x]) cand.push_back(op);
    // the old patches in their table order, once each (TransferValues'
    // pairs; a cell lies in one old patch, so the order changes no value)
    std::sort(cand.begin(), cand.end());
    cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    for (int op : cand) {
      const bsam_patch_desc& D = c->desc[pid];
      const bsam_patch_desc& S = o.desc[op];
      XferBox B{};
      B.S = o.hp[op];
      B.dp = pid;
      bool hit = true;
      for (int a = 0; a < 3; a++) {
        if (a >= nd) {
          B.slo[a] = B.dlo[a] = 0;
          B.ext[a] = 1;
          continue;
        }
        const int lo = std::max(D.mglobal[a][0], S.mglobal[a][0]);
        const int hi = std::min(D.mglobal[a][1], S.mglobal[a][1]);
        if (hi < lo) { hit = false; break; }
        B.slo[a] = lo - S.mglobal[a][0] + 1;
        B.dlo[a] = lo - D.mglobal[a][0] + 1;
        B.ext[a] = hi - lo + 1;
      }
      if (!hit) continue;
      B.cell0 = ncell;
      ncell += (long)B.ext[0] * B.ext[1] * B.ext[2];
      boxes.push_back(B);
    }
  }
  if (boxes.empty()) return BSAM_OK;
  // the box table of the previous level may still be read: wait for it
  HIPCHK(hipStreamSynchronize(c->s));
  if (boxes.size() > o.box_cap) {
    hipFree(o.d_box);
    o.d_box = nullptr;
    o.box_cap = 0;
    HIPCHK(hipMalloc(&o.d_box, boxes.size() * sizeof(XferBox)));
    o.box_cap = boxes.size();
  }
  HIPCHK(hipMemcpy(o.d_box, boxes.data(), boxes.size() * sizeof(XferBox), hipMemcpyHostToDevice));
  const int nbox = (int)boxes.size();
  DISPATCH_DIM(nd, (k_transfer<2><<<nblocks(ncell), kBlock, 0, c->s>>>(c->d_pt, o.d_box, nbox, ncell, nv, o.q, c->q)),
               (k_transfer<3><<<nblocks(ncell), kBlock, 0, c->s>>>(c->d_pt, o.d_box, nbox, ncell, nv, o.q, c->q)));
  LAUNCH_CHECK();
  return BSAM_OK;
}
}  // namespace

// GridAdapt's field set-up for a given new hierarchy: the levels at and
// below the root keep their one patch and its q, qold; then level by level
// from 1 up, the new patches are made (init_level) and SetGhost(level, 0)
// runs (AMR(level) at bsamroutines.f90:978-989 starts with it, before the
// next level is made).  The tagging and clustering of AMR change no field
// (with errflagopt 10 or 11), so for the same new hierarchy the fields equal
// the reference's after its regrid.
extern "C" int32_t bsam_regrid(bsam_ctx* c, int32_t npatch, const bsam_patch_desc* desc, int32_t minlevel,
                               int32_t finestlevel) {
  if (!c || !desc || npatch <= 0) return set_err(BSAM_EINVAL, "bad arguments");
  if (!c->have_hier) return set_err(BSAM_ESTATE, "no hierarchy installed (bsam_set_hierarchy)");
  if (minlevel != c->minlevel)
    return set_err(BSAM_EINVAL, "regrid keeps the coarsest level %d (got %d)", c->minlevel, minlevel);
  if (finestlevel < 0) return set_err(BSAM_EINVAL, "finest level %d below the root", finestlevel);
  for (int lvl = minlevel; lvl <= 0; lvl++) {
    int n = 0, id = -1;
    for (int i = 0; i < npatch; i++)
      if (desc[i].level == lvl) n++, id = i;
    if (n != 1) return set_err(BSAM_EINVAL, "regrid: level %d must hold one patch (has %d)", lvl, n);
    const DPatch& O = c->hp[c->L(lvl).patches[0]];
    for (int a = 0; a < 3; a++)
      if (desc[id].mx[a] != O.mx[a]) return set_err(BSAM_EINVAL, "regrid: level %d changed its extent", lvl);
  }
  HIPCHK(hipStreamSynchronize(c->s));
  OldMesh old(c);
  RETCHK(install(c, npatch, desc, minlevel, finestlevel, nullptr, &old, c->regrid_tables != 0));
  for (int lvl = minlevel; lvl <= finestlevel; lvl++) old.restore_shapes(lvl);
  RETCHK(carry_root_levels(c, old));
  RETCHK(bsam_set_ghost(c, 0, 0));
  for (int lvl = 1; lvl <= finestlevel; lvl++) {
    RETCHK(init_level(c, lvl, old));
    RETCHK(bsam_set_ghost(c, lvl, 0));
  }
  return BSAM_OK;
}

// AMR(rootlevel) (bsamroutines.f90:966-991), the regrid that opens every
// time step (:672-688): level by level from the root, SetGhost(level, 0),
// EstimateLevelErrors (bsam_estimate_level_errors, with its effect on q and
// qold for errflagopt 1) below maxlevel, the new patches of level+1 from the
// tags (bsam_new_level), and their fields (init_level), until a level makes
// no new patches.  Each new level is installed on top of the levels below
// it, which keep their device side and fields; the old mesh stays the
// source of TransferValues until the pass ends.
namespace {
// plan_ghost_stages of a level on a thread of its own (when `run`), so that
// the caller can go on with device work; join() returns its status with its
// message, and the destructor joins on every other way out.
struct StagePlanner {
  std::thread th;
  int32_t rc = BSAM_OK;
  std::string msg;
  StagePlanner(const bsam_ctx* c, const Level& L, GhostPlan& G, bool run) {
    if (!run) return;
    auto work = [this, c, &L, &G]() {
      try {
        rc = plan_ghost_stages(c, L, G);
        if (rc != BSAM_OK) msg = bsam_last_error();
      } catch (...) {
        rc = BSAM_ENOMEM;
        msg = "host stage planning failed";
      }
    };
    try {
      th = std::thread(work);
    } catch (...) {
      work();  // no thread: plan here
    }
  }
  int32_t join() {
    if (th.joinable()) th.join();
    return rc == BSAM_OK ? BSAM_OK : set_err(rc, "%s", msg.c_str());
  }
  ~StagePlanner() {
    if (th.joinable()) th.join();
  }
};
}  // namespace

extern "C" int32_t bsam_amr(bsam_ctx* c, const bsam_grid_params* gp) {
  if (!c || !gp) return set_err(BSAM_EINVAL, "bad arguments");
  if (!c->have_hier) return set_err(BSAM_ESTATE, "no hierarchy installed (bsam_set_hierarchy)");
  if (gp->maxlevel < 0 || gp->maxlevel >= BSAM_MAX_AMR_LEVELS)
    return set_err(BSAM_EINVAL, "maxlevel %d outside [0, %d)", gp->maxlevel, BSAM_MAX_AMR_LEVELS);
  for (int l = 0; l < gp->maxlevel; l++) {
    const int o = gp->errflagopt[l];
    if (o != 1 && o != 10 && o != 11) return set_err(BSAM_EINVAL, "level %d: errflagopt %d", l, o);
    if (gp->ibuffer[l] < 0) return set_err(BSAM_EINVAL, "level %d: ibuffer %d", l, gp->ibuffer[l]);
    if (gp->minimumgridpoints[l] < 4 || gp->minimumgridpoints[l] % 2)
      return set_err(BSAM_EINVAL, "level %d: minimumgridpoints must be even and at least 4", l);
  }
  const int minlevel = c->minlevel, nd = c->p.ndims;
  HIPCHK(hipStreamSynchronize(c->s));
  PhaseClock pc(c, BSAM_PHASE_AMR_INSTALL);
  OldMesh old(c);
  c->amr_mesh = true;
  // the new table starts with the old levels at and below the root
  std::vector<bsam_patch_desc> H;
  std::vector<int> at(old.desc.size(), -1);
  for (int lvl = minlevel; lvl <= 0; lvl++)
    for (int pid : *old.patches(lvl)) {
      bsam_patch_desc d = old.desc[pid];
      if (d.parent >= 0) d.parent = at[d.parent];
      at[pid] = (int)H.size();
      H.push_back(d);
    }
  RETCHK(install(c, (int)H.size(), H.data(), minlevel, 0, nullptr, &old, c->regrid_tables != 0));
  for (int lvl = minlevel; lvl <= 0; lvl++) old.restore_shapes(lvl);
  RETCHK(carry_root_levels(c, old));
  int32_t* flags = nullptr;
  // room for the new patches of a level: what the last pass made, with
  // margin (a second clustering call only when a level grows past it)
  std::vector<bsam_patch_desc>& nw = c->amr_out;
  for (int lvl = 0;; lvl++) {
    pc.next(BSAM_PHASE_AMR_TAG);
    RETCHK(bsam_set_ghost(c, lvl, 0));
    if (lvl >= gp->maxlevel) break;
    {
      const int32_t f0 = 0;
      const int64_t n0 = 0;
      RETCHK(check_tag_args(c, lvl, gp->errflagopt[lvl], &f0, &n0));
    }
    bool buffered = false;
    RETCHK(amr_level_tags(c, lvl, gp->errflagopt[lvl], gp->qtolerance[lvl], gp->ibuffer[lvl], flags, buffered));
    pc.next(BSAM_PHASE_AMR_CLUSTER);
    const int ib = buffered ? 0 : gp->ibuffer[lvl];
    {
      const std::vector<int>* was = old.patches(lvl + 1);
      const size_t want = 64 + 2 * (was ? was->size() : 0);
      if (nw.size() < want) nw.resize(want);
    }
    int32_t nn = 0;
    int32_t r = bsam_new_level(nd, (int)H.size(), H.data(), lvl, flags, ib, gp->minimumgridpoints[lvl],
                               gp->desiredfillratios[lvl], (int)nw.size(), nw.data(), &nn);
    if (r == BSAM_EINVAL && nn > (int)nw.size()) {
      nw.resize(nn);
      r = bsam_new_level(nd, (int)H.size(), H.data(), lvl, flags, ib, gp->minimumgridpoints[lvl],
                         gp->desiredfillratios[lvl], (int)nw.size(), nw.data(), &nn);
    }
    RETCHK(r);
    if (nn == 0) break;
    pc.next(BSAM_PHASE_AMR_INSTALL);
    H.insert(H.end(), nw.begin(), nw.begin() + nn);
    // the levels so far are the table's first nkeep rows: their fields are
    // the whole of the current pools, the start of the next ones
    const long span = c->nq;
    struct Prev {
      bsam_ctx* c;
      double *q, *qold;
      ~Prev() {
        hipStreamSynchronize(c->s);
        pool_release(c, q);
        pool_release(c, qold);
      }
    } prev{c, c->q, c->qold};
    c->q = c->qold = nullptr;
    int kept = minlevel - 1;
    RETCHK(install(c, (int)H.size(), H.data(), minlevel, lvl + 1, &kept, &old, c->regrid_tables != 0, lvl + 1));
    if (kept != lvl) return set_err(BSAM_EINVAL, "AMR: levels up to %d were not kept (%d)", lvl, kept);
    // a rebuilt level's box stages are planned on a host thread while its
    // fields are made (InitFields and TransferValues do not read the ghost
    // fill); its ghost tables follow
    Level& Ln = c->L(lvl + 1);
    GhostPlan G;
    StagePlanner stages(c, Ln, G, Ln.ghosts_deferred);
    pc.next(BSAM_PHASE_AMR_FIELDS);
    HIPCHK(hipMemcpyAsync(c->q, prev.q, span * sizeof(double), hipMemcpyDeviceToDevice, c->s));
    HIPCHK(hipMemcpyAsync(c->qold, prev.qold, span * sizeof(double), hipMemcpyDeviceToDevice, c->s));
    old.restore_shapes(lvl + 1);
    RETCHK(init_level(c, lvl + 1, old));
    RETCHK(stages.join());
    if (Ln.ghosts_deferred) {
      pc.next(BSAM_PHASE_AMR_INSTALL);
      // no sweep images: building them costs more than they save in the one
      // time step a level built here usually lives
      RETCHK(finish_level_ghosts(c, Ln, G, c->regrid_tables != 0, false));
    }
  }
  return BSAM_OK;
}

// ReadQ (bsaminputoutput.f90:184-443) onto the installed run: the frame's
// root must match the installed root as the reference checks it (nrvars,
// dx and xlower/xupper within 1e-8, mx, mglobal = 1..mx); the levels below
// the root stay as installed, the frame's refined patches replace the
// levels above it (bsam_parse_frame: visiting order, parents and mbounds
// from the boxes), their faces inherit the parent's physical or periodic
// codes where they reach the parent's faces (MakeNewGrid's rule), and q of
// the root and every refined patch is the frame's.  The other fields are
// zero, as after AllocFields.
extern "C" int32_t bsam_read_frame(bsam_ctx* c, const char* path, double* time, int32_t* finestlevel) {
  if (!c || !path || !time || !finestlevel) return set_err(BSAM_EINVAL, "bad arguments");
  if (!c->have_hier) return set_err(BSAM_ESTATE, "no hierarchy installed (bsam_set_hierarchy)");
  const int nd = c->p.ndims, nv = c->p.nrvars;
  int32_t np = 0, fin = 0;
  int64_t nqv = 0;
  double t = 0.0;
  RETCHK(bsam_parse_frame(path, nd, nv, &t, &fin, 0, nullptr, &np, nullptr, 0, &nqv));
  std::vector<bsam_patch_desc> fd(np);
  std::vector<double> fq((size_t)nqv);
  RETCHK(bsam_parse_frame(path, nd, nv, &t, &fin, np, fd.data(), &np, fq.data(), nqv, &nqv));
  const bsam_patch_desc& root = c->desc[c->L(0).patches[0]];
  const double small = 1.0e-8;
  const bsam_patch_desc& fr = fd[0];
  for (int a = 0; a < nd; a++) {
    // the reference's test of dx is one-sided (bsaminputoutput.f90:306)
    if (fr.dx - root.dx > small) return set_err(BSAM_EINVAL, "ReadQ: root-level grid spacing (dx) inconsistent");
    if (fabs(fr.xlower[a] - root.xlower[a]) > small)
      return set_err(BSAM_EINVAL, "ReadQ: root-level grid location (xlower) inconsistent");
    if (fr.mx[a] != root.mx[a]) return set_err(BSAM_EINVAL, "ReadQ: root-level grid size (mx) inconsistent");
    if (fr.mglobal[a][0] != 1 || fr.mglobal[a][1] != root.mx[a])
      return set_err(BSAM_EINVAL, "ReadQ: root-level grid size (mglobal) inconsistent");
  }
  // the table: the installed levels below the root, the root, the frame's
  // refined patches
  std::vector<bsam_patch_desc> T;
  std::vector<int> at(c->npatch, -1);
  for (int l = c->minlevel; l <= 0; l++)
    for (int pid : c->L(l).patches) {
      bsam_patch_desc d = c->desc[pid];
      if (d.parent >= 0) d.parent = at[d.parent];
      at[pid] = (int)T.size();
      T.push_back(d);
    }
  const int base = (int)T.size() - 1;  // the root's row
  for (int n = 1; n < np; n++) {
    bsam_patch_desc d = fd[n];
    d.parent = base + d.parent;
    const bsam_patch_desc& P = T[d.parent];
    for (int f = 0; f < 6; f++) d.mthbc[f] = 999;
    for (int a = 0; a < nd; a++) {
      if (P.mthbc[2 * a] < 999 && d.mbounds[a][0] == 1) d.mthbc[2 * a] = P.mthbc[2 * a];
      if (P.mthbc[2 * a + 1] < 999 && d.mbounds[a][1] == P.mx[a]) d.mthbc[2 * a + 1] = P.mthbc[2 * a + 1];
    }
    T.push_back(d);
  }
  RETCHK(bsam_set_hierarchy(c, (int)T.size(), T.data(), c->minlevel, fin));
  size_t off = 0;
  for (int n = 0; n < np; n++) {
    long cells = 1;
    for (int a = 0; a < nd; a++) cells *= fd[n].mx[a] + 2;
    const int64_t cnt = (int64_t)cells * nv;
    RETCHK(bsam_upload(c, base + n, BSAM_FIELD_Q, fq.data() + off, cnt));
    off += (size_t)cnt;
  }
  HIPCHK(hipStreamSynchronize(c->s));
  *time = t;
  *finestlevel = fin;
  return BSAM_OK;
}

extern "C" int32_t bsam_get_hierarchy(bsam_ctx* c, int32_t maxout, bsam_patch_desc* out, int32_t* npatch,
                                      int32_t* minlevel, int32_t* finestlevel) {
  if (!c || !npatch) return set_err(BSAM_EINVAL, "bad arguments");
  if (!c->have_hier) return set_err(BSAM_ESTATE, "no hierarchy installed (bsam_set_hierarchy)");
  *npatch = c->npatch;
  if (minlevel) *minlevel = c->minlevel;
  if (finestlevel) *finestlevel = c->finest;
  if (!out) return BSAM_OK;
  if (maxout < c->npatch) return set_err(BSAM_EINVAL, "%d descriptors do not fit in %d", c->npatch, maxout);
  std::copy(c->desc.begin(), c->desc.end(), out);
  return BSAM_OK;
}

// One pass of the reference's time loop (bsamroutines.f90:672-758): AMR,
// FillDown on the first step (firstamr), CopyQToQold on every level,
// MultigridIterations.
extern "C" int32_t bsam_time_step_amr(bsam_ctx* c, const bsam_grid_params* gp, int32_t first_step,
                                      int32_t* ncycles, double* residual) {
  RETCHK(bsam_amr(c, gp));
  PhaseClock pc(c, BSAM_PHASE_SOLVE);
  if (first_step) RETCHK(bsam_fill_down(c, BSAM_SOLUTIONFIELD));
  for (int level = c->minlevel; level <= c->finest; level++) RETCHK(bsam_copy_q_to_qold(c, level));
  return bsam_multigrid_iterations(c, ncycles, residual, nullptr, 0);
}

extern "C" int32_t bsam_synchronize(bsam_ctx* c) {
  if (!c) return set_err(BSAM_EINVAL, "null context");
  HIPCHK(hipStreamSynchronize(c->s));
  return BSAM_OK;
}

// ---------------------------------------------------------------------------
// instrumentation
// ---------------------------------------------------------------------------
extern "C" int32_t bsam_kstat_enable(bsam_ctx* c, int32_t on) {
  if (!c) return set_err(BSAM_EINVAL, "null context");
  HIPCHK(hipStreamSynchronize(c->s));
  for (auto& k : c->ks) {
    for (hipEvent_t e : k.ev) hipEventDestroy(e);
    k.ev.clear();
    k.launches = 0;
    k.ms = 0.0;
    k.on = on != 0;
  }
  return BSAM_OK;
}

extern "C" int32_t bsam_ghost_table_digest(bsam_ctx* c, int32_t level, int64_t* out) {
  if (!c || !out) return set_err(BSAM_EINVAL, "null argument");
  RETCHK(check_level(c, level));
  const Level& L = c->L(level);
  if (!L.maps_ok) return set_err(BSAM_ESTATE, "level %d has no flattened ghost tables (%s)", level, L.why.c_str());
  HIPCHK(hipStreamSynchronize(c->s));
  for (int m = 0; m < 7; m++) {
    const DevMap& D = L.gmap[m];
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const void* dev, size_t bytes) -> hipError_t {
      if (!dev || !bytes) return hipSuccess;
      std::vector<unsigned char> b(bytes);
      const hipError_t e = hipMemcpy(b.data(), dev, bytes, hipMemcpyDeviceToHost);
      for (unsigned char x : b) h = (h ^ x) * 1099511628211ull;
      return e;
    };
    HIPCHK(mix(D.cr, (size_t)D.nc * sizeof(int4)));
    HIPCHK(mix(D.cpost, (size_t)D.nc * sizeof(uint16_t)));
    HIPCHK(mix(D.ir, (size_t)D.ni * sizeof(GRule)));
    out[5 * m] = D.nc;
    out[5 * m + 1] = D.ni;
    out[5 * m + 2] = D.nc1;
    out[5 * m + 3] = D.ni1;
    out[5 * m + 4] = (int64_t)h;
  }
  return BSAM_OK;
}

extern "C" int32_t bsam_phase_enable(bsam_ctx* c, int32_t on) {
  if (!c) return set_err(BSAM_EINVAL, "null context");
  HIPCHK(hipStreamSynchronize(c->s));
  c->phase_on = on != 0;
  for (int i = 0; i < BSAM_PHASE_COUNT; i++) c->phase_calls[i] = 0, c->phase_ms[i] = 0.0;
  return BSAM_OK;
}

extern "C" int32_t bsam_phase_read(bsam_ctx* c, int32_t which, int64_t* calls, double* total_ms) {
  if (!c || which < 0 || which >= BSAM_PHASE_COUNT) return set_err(BSAM_EINVAL, "bad phase index");
  if (calls) *calls = c->phase_calls[which];
  if (total_ms) *total_ms = c->phase_ms[which];
  return BSAM_OK;
}

extern "C" int32_t bsam_kstat_read(bsam_ctx* c, int32_t which, int64_t* launches, double* total_ms,
                                   double* bytes_per_launch) {
  if (!c || which < 0 || which >= BSAM_KSTAT_COUNT) return set_err(BSAM_EINVAL, "bad kstat index");
  HIPCHK(hipStreamSynchronize(c->s));
  KStat& k = c->ks[which];
  double ms = 0.0;
  for (size_t i = 0; i + 1 < k.ev.size(); i += 2) {
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, k.ev[i], k.ev[i + 1]));
    ms += t;
  }
  if (launches) *launches = k.launches;
  if (total_ms) *total_ms = ms;
  if (bytes_per_launch) *bytes_per_launch = k.bytes;
  return BSAM_OK;
}

// ---------------------------------------------------------------------------
// mixed precision: FP64 defect correction around FP32 correction V-cycles
// ---------------------------------------------------------------------------
namespace bsam {
int32_t mixed_init(bsam_ctx* c) {
  MixedState* m = new MixedState();
  c->mixed = m;
  float** qs[3] = {&m->e, &m->rhs, &m->coef};
  for (float** pp : qs) {
    if (pool_alloc(c, pp, c->nq * sizeof(float)) != hipSuccess) return set_err(BSAM_ENOMEM, "FP32 pools");
    HIPCHK(hipMemsetAsync(*pp, 0, c->nq * sizeof(float), c->s));
  }
  if (pool_alloc(c, &m->ec, c->nc * sizeof(float)) != hipSuccess) return set_err(BSAM_ENOMEM, "FP32 qc pool");
  HIPCHK(hipMemsetAsync(m->ec, 0, c->nc * sizeof(float), c->s));
  return BSAM_OK;
}

void mixed_free(MixedState* m) {
  if (!m) return;
  (void)hipFree(m->e);
  (void)hipFree(m->rhs);
  (void)hipFree(m->coef);
  (void)hipFree(m->ec);
  delete m;
}
}  // namespace bsam

namespace {

// True when the correction cycle's first and last finest-level sweeps take
// over the zeroing of e and the update q += e: a single level (finest = 0)
// that is one large patch periodic in every dimension, swept by the
// marching kernel.
static bool mixed_fused(bsam_ctx* c) {
  if (c->finest != 0 || c->p.nsmoothingpasses < 1 || c->p.model != BSAM_MODEL_CH) return false;
  Level& L = c->L(0);
  int nblk = 0;
  return L.wrap_img && sweep_kc(c, L, 4, &nblk) > 0;
}

// A refined mesh whose finest level runs the flat correction sweeps: its
// first sweep starts from a zero correction and its last adds the
// correction into q (k_relax_lin EMODE 1, 2), so the finest level's q += e
// and the clearing of e need no pass of their own.
static bool mixed_fused_amr(bsam_ctx* c) {
  if (c->finest == 0 || c->p.nsmoothingpasses < 1) return false;
  Level& L = c->L(c->finest);
  int nblk = 0;
  if (L.wrap_pid >= 0) return false;
  if (c->p.model == BSAM_MODEL_CH && sweep_kc(c, L, 4, &nblk) > 0) return false;
  for (int pid : L.patches)
    if (c->hp[pid].mx[0] % 2) return false;
  return true;
}

// One colour sweep of the correction smoother; returns 1 when the sweep
// also wrote the periodic ghost images (the following SetGhost is then
// redundant), 0 otherwise.  emode 1: the first sweep from e = 0; emode 2:
// the last sweep, adding the correction into q (the finest level of
// mixed_fused or mixed_fused_amr meshes).
int32_t mixed_relax(bsam_ctx* c, int level, int ipass, int emode = 0, bool lean_fill = false) {
  bsam::MixedState* m = c->mixed;
  Level& L = c->L(level);
  const double h2 = L.h * L.h, eps2 = c->p.model_params[0];
  const double d2 = 2.0 * c->p.ndims;
  const float tmp1 = (float)(c->p.dt / h2), tmp2 = (float)(eps2 / h2), tmp3 = (float)(d2 * eps2 / h2);
  const float a12 = (float)(d2 * (c->p.dt / h2));
  const bool fin = level == c->finest;
  const bool tm = c->p.model == BSAM_MODEL_TUMOUR;
  // the tumour sweep writes no periodic images: its SetGhost follows
  const bool wrap = L.wrap_img && !tm;
  int nblk = 0;
  const int kc = tm ? 0 : sweep_kc(c, L, 4, &nblk);
  const bool timed = fin && emode == 0;
  if (timed) kstat_begin(c, BSAM_KSTAT_RELAX_FINEST);
  const bool img = tm && emode == 0 && phys_images_level(c, L);
  // the flat CH sweep writes the images of the reduced fill after it
  const bool sweep_img = !tm && !kc && !wrap && lean_fill && use_images(c, L, ipass);
  if (tm) {
    const ModelK mk = c->mk();
    FOR_LAUNCHES(c, L, kHalf, a, nb) {
      if (img)
        DISPATCH_DIM(c->p.ndims,
                     (k_relax_lin_tm<2, 0, true><<<nb, kBlock, 0, c->s>>>(a, ipass - 1, mk, m->e, m->rhs, m->coef)),
                     (k_relax_lin_tm<3, 0, true><<<nb, kBlock, 0, c->s>>>(a, ipass - 1, mk, m->e, m->rhs, m->coef)));
      else if (emode == 1)
        DISPATCH_DIM(c->p.ndims,
                     (k_relax_lin_tm<2, 1><<<nb, kBlock, 0, c->s>>>(a, ipass - 1, mk, m->e, m->rhs, m->coef)),
                     (k_relax_lin_tm<3, 1><<<nb, kBlock, 0, c->s>>>(a, ipass - 1, mk, m->e, m->rhs, m->coef)));
      else if (emode == 2)
        DISPATCH_DIM(c->p.ndims,
                     (k_relax_lin_tm<2, 2><<<nb, kBlock, 0, c->s>>>(a, ipass - 1, mk, m->e, m->rhs, m->coef, c->q)),
                     (k_relax_lin_tm<3, 2><<<nb, kBlock, 0, c->s>>>(a, ipass - 1, mk, m->e, m->rhs, m->coef, c->q)));
      else
        DISPATCH_DIM(c->p.ndims,
                     (k_relax_lin_tm<2><<<nb, kBlock, 0, c->s>>>(a, ipass - 1, mk, m->e, m->rhs, m->coef)),
                     (k_relax_lin_tm<3><<<nb, kBlock, 0, c->s>>>(a, ipass - 1, mk, m->e, m->rhs, m->coef)));
      LAUNCH_CHECK();
    }
  } else if (kc) {
    const int pid = L.patches[0];
    if (wrap) {
      if (emode == 1)
        k_relax3d_march<float, true, true, 1, 1><<<nblk, kSweepThreads, 0, c->s>>>(c->d_pt, pid, ipass - 1, kc, tmp1, tmp2, tmp3, a12, m->e, m->rhs, m->coef);
      else if (emode == 2)
        k_relax3d_march<float, true, true, 1, 2><<<nblk, kSweepThreads, 0, c->s>>>(c->d_pt, pid, ipass - 1, kc, tmp1, tmp2, tmp3, a12, m->e, m->rhs, m->coef, c->q);
      else if (fin) k_relax3d_march<float, true, true, 1><<<nblk, kSweepThreads, 0, c->s>>>(c->d_pt, pid, ipass - 1, kc, tmp1, tmp2, tmp3, a12, m->e, m->rhs, m->coef);
      else k_relax3d_march<float, true, true, 0><<<nblk, kSweepThreads, 0, c->s>>>(c->d_pt, pid, ipass - 1, kc, tmp1, tmp2, tmp3, a12, m->e, m->rhs, m->coef);
    } else {
      k_relax3d_march<float, true, false, 0><<<nblk, kSweepThreads, 0, c->s>>>(c->d_pt, pid, ipass - 1, kc, tmp1, tmp2, tmp3, a12, m->e, m->rhs, m->coef);
    }
  } else if (sweep_img) {
    const ImgTab I{L.img_off[ipass - 1], L.img_ent[ipass - 1], 1};
    FOR_LAUNCHES(c, L, kHalf, a, nb) {
      if (emode == 1)
        DISPATCH_DIM(c->p.ndims,
                     (k_relax_lin<2, false, 1, true><<<nb, kBlock, 0, c->s>>>(a, ipass - 1, tmp1, tmp2, tmp3, a12, m->e, m->rhs, m->coef, nullptr, I)),
                     (k_relax_lin<3, false, 1, true><<<nb, kBlock, 0, c->s>>>(a, ipass - 1, tmp1, tmp2, tmp3, a12, m->e, m->rhs, m->coef, nullptr, I)));
      else if (emode == 2)
        DISPATCH_DIM(c->p.ndims,
                     (k_relax_lin<2, false, 2, true><<<nb, kBlock, 0, c->s>>>(a, ipass - 1, tmp1, tmp2, tmp3, a12, m->e, m->rhs, m->coef, c->q, I)),
                     (k_relax_lin<3, false, 2, true><<<nb, kBlock, 0, c->s>>>(a, ipass - 1, tmp1, tmp2, tmp3, a12, m->e, m->rhs, m->coef, c->q, I)));
      else
        DISPATCH_DIM(c->p.ndims,
                     (k_relax_lin<2, false, 0, true><<<nb, kBlock, 0, c->s>>>(a, ipass - 1, tmp1, tmp2, tmp3, a12, m->e, m->rhs, m->coef, nullptr, I)),
                     (k_relax_lin<3, false, 0, true><<<nb, kBlock, 0, c->s>>>(a, ipass - 1, tmp1, tmp2, tmp3, a12, m->e, m->rhs, m->coef, nullptr, I)));
      LAUNCH_CHECK();
    }
  } else {
    FOR_LAUNCHES(c, L, kHalf, a, nb) {
      if (wrap)
        DISPATCH_DIM(c->p.ndims,
                     (k_relax_lin<2, true><<<nb, kBlock, 0, c->s>>>(a, ipass - 1, tmp1, tmp2, tmp3, a12, m->e, m->rhs, m->coef)),
                     (k_relax_lin<3, true><<<nb, kBlock, 0, c->s>>>(a, ipass - 1, tmp1, tmp2, tmp3, a12, m->e, m->rhs, m->coef)));
      else if (emode == 1)
        DISPATCH_DIM(c->p.ndims,
                     (k_relax_lin<2, false, 1><<<nb, kBlock, 0, c->s>>>(a, ipass - 1, tmp1, tmp2, tmp3, a12, m->e, m->rhs, m->coef)),
                     (k_relax_lin<3, false, 1><<<nb, kBlock, 0, c->s>>>(a, ipass - 1, tmp1, tmp2, tmp3, a12, m->e, m->rhs, m->coef)));
      else if (emode == 2)
        DISPATCH_DIM(c->p.ndims,
                     (k_relax_lin<2, false, 2><<<nb, kBlock, 0, c->s>>>(a, ipass - 1, tmp1, tmp2, tmp3, a12, m->e, m->rhs, m->coef, c->q)),
                     (k_relax_lin<3, false, 2><<<nb, kBlock, 0, c->s>>>(a, ipass - 1, tmp1, tmp2, tmp3, a12, m->e, m->rhs, m->coef, c->q)));
      else
        DISPATCH_DIM(c->p.ndims,
                     (k_relax_lin<2, false><<<nb, kBlock, 0, c->s>>>(a, ipass - 1, tmp1, tmp2, tmp3, a12, m->e, m->rhs, m->coef)),
                     (k_relax_lin<3, false><<<nb, kBlock, 0, c->s>>>(a, ipass - 1, tmp1, tmp2, tmp3, a12, m->e, m->rhs, m->coef)));
      LAUNCH_CHECK();
    }
  }
  LAUNCH_CHECK();
  if (timed) {
    double N = 0, G = 0;
    for (int pid : L.patches) {
      const DPatch& P = c->hp[pid];
      N += (double)P.mx[0] * P.mx[1] * P.mx[2];
      G += (c->p.ndims == 3) ? 2.0 * ((double)P.mx[0] * P.mx[1] + (double)P.mx[1] * P.mx[2] + (double)P.mx[0] * P.mx[2])
                             : 2.0 * ((double)P.mx[0] + P.mx[1]);
    }
    // e_phi and e_mu of the other colour over (N+G)/2 each, rhs (2 fields),
    // coef and the cell's own e_phi (the Newton step's phi0 + e) of the
    // colour N/2 each, the colour's e written N/2 x 2 fields; the tumour
    // model: e (4 fields) of the other colour, rhs (4) and coef (2) of the
    // colour, e (4) written
    if (tm) kstat_end(c, BSAM_KSTAT_RELAX_FINEST, 4.0 * (2.0 * (N + G) + 2.0 * N + N + 2.0 * N));
    else
      kstat_end(c, BSAM_KSTAT_RELAX_FINEST,
                4.0 * ((N + G) + N + 0.5 * N + 0.5 * N + N) + (sweep_img ? image_bytes(L, ipass, 4.0, 2) : 0.0));
  }
  if (sweep_img) return 2;
  return (wrap || img) ? 1 : 0;
}

// nsmoothingpasses red-black passes; phase 0 = pre-smoothing, 1 = post.
int32_t mixed_level_relax(bsam_ctx* c, int level, int phase) {
  const bool fused = level == c->finest && (mixed_fused(c) || mixed_fused_amr(c));
  const int ns = c->p.nsmoothingpasses;
  for (int s = 0; s < ns; s++)
    for (int rb = 1; rb <= 2; rb++) {
      int emode = 0;
      if (fused && phase == 0 && s == 0 && rb == 1) emode = 1;
      if (fused && phase == 1 && s == ns - 1 && rb == 2) emode = 2;
      const bool last = s == ns - 1 && rb == 2;
      // After the finest level's last sweep (emode 2) nothing reads its
      // correction's ghosts: q has been updated and the next cycle starts
      // from zero.  (CH on a refined mesh: the reduced fill that follows
      // leaves the ghosts that are part of the state the next cycle and the
      // norm read, those of the FAS cycle after its last SetGhost.)  The
      // pre-smoothing's last fill is reduced too when the cycle descends:
      // after two reduced fills every face ghost is current, and until the
      // post-correction fill only face ghosts are read (the coarse right-
      // hand side); edge and corner ghosts are rewritten by the post-
      // smoothing's last, full, fill.
      const bool lean_fill =
          emode == 2 ? (c->p.model == BSAM_MODEL_CH && c->finest > 0) : (!last || (phase == 0 && level > c->minlevel));
      const int32_t done = mixed_relax(c, level, rb, emode, lean_fill);
      RETCHK(done);
      if (done == 2) {
        // the sweep wrote the fill's images; the gather does the rest
        RETCHK(set_ghost_t<float>(c, level, rb, c->mixed->e, 1, kLeanImg));
        continue;
      }
      if (emode == 2) {
        if (c->p.model == BSAM_MODEL_CH && c->finest > 0)
          RETCHK(set_ghost_t<float>(c, level, rb, c->mixed->e, 1, kLean));
        continue;
      }
      if (!done) RETCHK(set_ghost_t<float>(c, level, rb, c->mixed->e, 1, lean_fill ? kLean : kFull));
    }
  return BSAM_OK;
}

int32_t mixed_vcycle(bsam_ctx* c, int level) {
  bsam::MixedState* m = c->mixed;
  if (c->L(level).chain > 0) return launch_coarse_chain(c, level, true);
  RETCHK(mixed_level_relax(c, level, 0));
  if (level > c->minlevel) {
    const bool rw = restrict_wraps_parent(c, level);
    if (level <= 0 && c->p.model == BSAM_MODEL_CH) {
      // a level below the root is wholly covered: its base state phi0 is
      // the restriction of this level's (the FP64 state there is stale
      // inside the loop)
      Level& L = c->L(level);
      FOR_LAUNCHES(c, L, kCoarse, a, nb) {
        DISPATCH_DIM(c->p.ndims, (k_restrict<2, float><<<nb, kBlock, 0, c->s>>>(a, 1, m->coef, m->coef, nullptr)),
                     (k_restrict<3, float><<<nb, kBlock, 0, c->s>>>(a, 1, m->coef, m->coef, nullptr)));
        LAUNCH_CHECK();
      }
    }
    RETCHK(restrict_t<float>(c, level, m->e, m->e, m->ec, rw));
    if (!rw) RETCHK(set_ghost_t<float>(c, level - 1, 0, m->e, 1));
    RETCHK(ring_t<float>(c, level, m->e, m->ec));
    Level& L = c->L(level);
    const ModelK mk = c->mk();
    FOR_LAUNCHES(c, L, kCoarse, a, nb) {
      if (c->p.model == BSAM_MODEL_TUMOUR)
        DISPATCH_DIM(c->p.ndims,
                     (k_fas_rhs_lin_tm<2><<<nb, kBlock, 0, c->s>>>(a, mk, m->e, m->rhs, m->ec, m->coef, m->rhs)),
                     (k_fas_rhs_lin_tm<3><<<nb, kBlock, 0, c->s>>>(a, mk, m->e, m->rhs, m->ec, m->coef, m->rhs)));
      else
        DISPATCH_DIM(c->p.ndims,
                     (k_fas_rhs_lin<2><<<nb, kBlock, 0, c->s>>>(a, c->p.dt, c->p.model_params[0], m->e, m->rhs, m->ec, m->coef, m->rhs)),
                     (k_fas_rhs_lin<3><<<nb, kBlock, 0, c->s>>>(a, c->p.dt, c->p.model_params[0], m->e, m->rhs, m->ec, m->coef, m->rhs)));
      LAUNCH_CHECK();
    }
    RETCHK(mixed_vcycle(c, level - 1));
    const bool cw = c->L(level).wrap_pid >= 0;
    RETCHK(correct_t<float>(c, level, m->e, m->ec, cw));
    if (!cw)
      RETCHK(set_ghost_t<float>(c, level, 0, m->e, 1, phys_images_level(c, c->L(level)) ? kFull : kAfterCorrection));
    RETCHK(mixed_level_relax(c, level, 1));
  }
  return BSAM_OK;
}

// A level's ghosts at a new base state of the correction cycle (levels
// 0..finest, root first).  The FAS cycle does not refill ghosts after
// FillDown: until a level's next SetGhost, whatever reads its ghosts -- the
// level's own first sweep and the norm (afasroutines.f90:104) on the
// finest level, the coarse-fine interpolation of the level above during
// that level's pre-smoothing -- sees the ghosts of the level's last
// SetGhost in the previous cycle, filled before FillDown changed coarse
// cells under finer patches (stale ghosts G).  In correction form those
// are q's ghosts plus the correction's, q_g + e_g.  Every fill of e is
// linear and reads the parent's e, so for the fills of e to give the FAS
// cycle's values, q's ghosts must be a fresh fill F of the base state, each
// reading the parent's fresh fill: this writes F into q and G - F into e_g
// (a read before the level's next fill of e sees G; that fill replaces the
// offset, and on the level above it carries the parent's G).  One gather
// (REBASE) on flattened levels; else the shell is saved in rf, filled, and
// differenced.
int32_t rebase_ghosts(bsam_ctx* c, int level) {
  Level& L = c->L(level);
  float* e = c->mixed->e;
  const int nv = c->p.nrvars;
  if (L.maps_ok && c->p.batching) {
    const DevMap& M = L.gmap[0];
    if (M.nc1 + M.ni1 > 0) {
      k_ghost_map<double, true><<<nblocks(M.nc1 + M.ni1), kBlock, 0, c->s>>>(M.cr, M.cpost, M.nc1, M.ir, M.ni1, nv,
                                                                            0, c->q, L.gorder[0], e);
      LAUNCH_CHECK();
    }
    const int nc = M.nc - M.nc1, ni = M.ni - M.ni1;
    if (nc + ni > 0) {
      k_ghost_map<double, true><<<nblocks(nc + ni), kBlock, 0, c->s>>>(
          M.cr + M.nc1, M.cpost ? M.cpost + M.nc1 : nullptr, nc, M.ir + M.ni1, ni, nv, 0, c->q, L.gorder[0], e);
      LAUNCH_CHECK();
    }
    return BSAM_OK;
  }
  {
    FOR_LAUNCHES(c, L, kFineG, a, nb) {
      DISPATCH_DIM(c->p.ndims, (k_ghost_stash<2, 1><<<nb, kBlock, 0, c->s>>>(a, nv, c->q, e, c->rf)),
                   (k_ghost_stash<3, 1><<<nb, kBlock, 0, c->s>>>(a, nv, c->q, e, c->rf)));
      LAUNCH_CHECK();
    }
  }
  RETCHK(set_ghost_t<double>(c, level, 0, c->q, 0));
  {
    FOR_LAUNCHES(c, L, kFineG, a, nb) {
      DISPATCH_DIM(c->p.ndims, (k_ghost_stash<2, 0><<<nb, kBlock, 0, c->s>>>(a, nv, c->q, e, c->rf)),
                   (k_ghost_stash<3, 0><<<nb, kBlock, 0, c->s>>>(a, nv, c->q, e, c->rf)));
      LAUNCH_CHECK();
    }
  }
  return BSAM_OK;
}

// FP64 residual of levels >= 0 into the FP32 right-hand side; the finest
// level's squared sums go to d_norm.
int32_t mixed_residuals(bsam_ctx* c) {
  bsam::MixedState* m = c->mixed;
  const ModelK mk = c->mk();
  int nbf = 0;
  for (int level = c->finest; level >= 0; level--) {
    Level& L = c->L(level);
    double* part = (level == c->finest) ? c->d_partial : nullptr;
    const bool fin = level == c->finest;
    if (fin) kstat_begin(c, BSAM_KSTAT_RESID_FINEST);
    // CH: the base state phi0 of the next correction cycle goes to coef;
    // on a refined mesh the finest norm uses the FAS cycle's ghosts
    // (rebase_ghosts)
    float* base = (c->p.model == BSAM_MODEL_CH) ? m->coef : nullptr;
    const float* eg = (c->p.model == BSAM_MODEL_CH && fin && c->finest > 0) ? m->e : nullptr;
    int nbm = resid_march<float>(c, level, m->rhs, part, base);
    if (nbm < 0) return set_err(BSAM_EHIP, "residual kernel launch failed");
    int nb = nbm;
    if (nbm == 0) {
      LevelArgs a;
      int nbl = 0;
      for (LaunchIter it(c, L, kCell); it.next(a, nbl);) {
        double* pp = part ? part + (long)kNvMax * it.boff : nullptr;
        if (c->p.model == BSAM_MODEL_TUMOUR)
          DISPATCH_DIM(c->p.ndims,
                       (k_residual<2, BSAM_MODEL_TUMOUR, float><<<nbl, kBlock, 0, c->s>>>(a, mk, c->q, c->qold, c->f, m->rhs, pp)),
                       (k_residual<3, BSAM_MODEL_TUMOUR, float><<<nbl, kBlock, 0, c->s>>>(a, mk, c->q, c->qold, c->f, m->rhs, pp)));
        else
          DISPATCH_DIM(c->p.ndims,
                       (k_residual<2, BSAM_MODEL_CH, float><<<nbl, kBlock, 0, c->s>>>(a, mk, c->q, c->qold, c->f, m->rhs, pp, base, eg)),
                       (k_residual<3, BSAM_MODEL_CH, float><<<nbl, kBlock, 0, c->s>>>(a, mk, c->q, c->qold, c->f, m->rhs, pp, base, eg)));
        LAUNCH_CHECK();
        nb = it.blocks();
      }
    }
    if (fin) {
      double N = 0, G = 0;
      for (int pid : L.patches) {
        const DPatch& P = c->hp[pid];
        N += (double)P.mx[0] * P.mx[1] * P.mx[2];
        G += (c->p.ndims == 3) ? 2.0 * ((double)P.mx[0] * P.mx[1] + (double)P.mx[1] * P.mx[2] + (double)P.mx[0] * P.mx[2])
                               : 2.0 * ((double)P.mx[0] + P.mx[1]);
      }
      // q over the interior and face ghosts, f, the FP32 right-hand side
      // (all fields) and, for the tumour model, qold of phi
      const double nv = c->p.nrvars;
      kstat_end(c, BSAM_KSTAT_RESID_FINEST,
                8.0 * nv * (N + G) + 8.0 * nv * N + 4.0 * nv * N + (c->p.model == BSAM_MODEL_TUMOUR ? 8.0 * N : 0.0));
      nbf = nb;
    }
  }
  return reduce_norm(c, nbf);
}

// One outer iteration: correction V-cycle, q += e, FillDown, ghost refresh
// of the finest level, new residual and norm.
// The correction e starts every cycle at zero on the levels >= 0: they are
// cleared by the q += e update of the previous cycle (or the set-up).  A
// level below the root lies wholly under the next finer level, so the
// restriction writes every interior cell of it and the SetGhost after it
// every ghost before anything reads them: its old values need no clearing.
int32_t mixed_iteration(bsam_ctx* c) {
  bsam::MixedState* m = c->mixed;
  RETCHK(mixed_vcycle(c, c->finest));
  // with the fused finest sweeps, q += e has been done by the last sweep
  // (on a refined mesh for the finest level only)
  const int top = mixed_fused(c) ? -1 : (mixed_fused_amr(c) ? c->finest - 1 : c->finest);
  for (int level = 0; level <= top; level++) {
    Level& L = c->L(level);
    if (L.qspan_len > 0) {
      k_add_corr<<<nblocks(L.qspan_len), kBlock, 0, c->s>>>(c->q + L.qspan_off, m->e + L.qspan_off, L.qspan_len);
      LAUNCH_CHECK();
    } else {
      for (int pid : L.patches) {
        const DPatch& P = c->hp[pid];
        const long n = (long)P.qvs * c->p.nrvars;
        k_add_corr<<<nblocks(n), kBlock, 0, c->s>>>(c->q + P.qoff, m->e + P.qoff, n);
        LAUNCH_CHECK();
      }
    }
  }
  // With refined levels the finest residual sees the coarse level through
  // the coarse-fine ghost interpolation, so the coarse q must be the
  // restriction of the updated fine q first (levels >= 1 into their
  // parents; nothing in the loop reads q below the root).  A single-level
  // mesh has no such coupling (its ghosts were updated with e).  The rest
  // of FillDown is done once after the loop.
  if (c->finest > 0) {
    RETCHK(fill_down_range(c, BSAM_SOLUTIONFIELD, c->finest, 1));
    if (c->p.model == BSAM_MODEL_CH) {
      for (int level = 0; level <= c->finest; level++) RETCHK(rebase_ghosts(c, level));
    } else {
      RETCHK(set_ghost_t<double>(c, c->finest, 0, c->q, 0));
    }
  }
  RETCHK(mixed_residuals(c));
  return BSAM_OK;
}

}  // namespace

namespace bsam {
int32_t mixed_multigrid_iterations(bsam_ctx* c, int32_t* ncycles, double* residual, double* history,
                                   int32_t history_len) {
  RETCHK(bsam_fill_down(c, BSAM_SOLUTIONFIELD));
  for (int level = c->finest; level >= 0; level--) RETCHK(source_and_update(c, level));
  RETCHK(bsam_fill_down(c, BSAM_SOURCEFIELD));
  // the tumour model's correction operator is linearised about the step's
  // start (3 phi^2 and t0 frozen); the CH correction works on the base
  // state phi0 that every FP64 residual evaluation writes
  if (c->p.model == BSAM_MODEL_TUMOUR)
    for (int level = c->minlevel; level <= c->finest; level++) {
      Level& L = c->L(level);
      FOR_LAUNCHES(c, L, kCell, a, nb) {
        DISPATCH_DIM(c->p.ndims, (k_coef<2, true><<<nb, kBlock, 0, c->s>>>(a, c->q, c->qold, c->mixed->coef)),
                     (k_coef<3, true><<<nb, kBlock, 0, c->s>>>(a, c->q, c->qold, c->mixed->coef)));
        LAUNCH_CHECK();
      }
    }
  k_zero_f<<<nblocks(c->nq), kBlock, 0, c->s>>>(c->mixed->e, c->nq);
  LAUNCH_CHECK();
  if (c->p.model == BSAM_MODEL_CH && c->finest > 0)
    for (int level = 0; level <= c->finest; level++) RETCHK(rebase_ghosts(c, level));
  RETCHK(mixed_residuals(c));
  double res = 0.0, prev = 0.0;
  bool fp64 = false;  // fell back to the FP64 FAS V-cycle
  int it;
  for (it = 1; it <= c->p.maxvcycles; it++) {
    if (fp64) RETCHK(run_iteration(c, enqueue_iteration, true));
    else RETCHK(run_iteration(c, mixed_iteration));
    HIPCHK(hipStreamSynchronize(c->s));
    res = norm_from(c, c->h_norm, c->finest);
    if (history && it - 1 < history_len) history[it - 1] = res;
    if (!std::isfinite(res)) {
      *ncycles = it;
      *residual = res;
      return set_err(BSAM_EDIVERGED, "residual not finite after outer iteration %d", it);
    }
    if (res < c->p.qerrortol) break;
    // an outer iteration that does not reduce the residual by the stall
    // ratio hands the rest of the step to the FP64 V-cycle, which starts
    // from the current q with f as the source (the state every outer
    // iteration leaves: q updated, filled down to level 1, the finest
    // level's ghosts set)
    if (!fp64 && it >= 2 && res > c->mixed_stall * prev) {
      fp64 = true;
      c->mixed_fallbacks++;
    }
    prev = res;
  }
  if (fp64) {
    // as the FP64 path leaves f: the source, not the tau terms
    for (int level = c->finest - 1; level >= 0; level--) RETCHK(bsam_update_source(c, level));
    RETCHK(bsam_fill_down(c, BSAM_SOURCEFIELD));
  }
  RETCHK(fill_down_range(c, BSAM_SOLUTIONFIELD, 0, c->minlevel + 1));
  *ncycles = it > c->p.maxvcycles ? c->p.maxvcycles : it;
  *residual = res;
  return it > c->p.maxvcycles ? BSAM_NOT_CONVERGED : BSAM_OK;
}
}  // namespace bsam

// ---------------------------------------------------------------------------
// Autotuning of launch shapes
// ---------------------------------------------------------------------------
namespace {
// Mean duration (us) of `reps` calls of fn on the context's stream, after
// one untimed call.
template <typename F>
static float time_launches(bsam_ctx* c, int reps, F&& fn) {
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess) return -1.0f;
  if (hipEventCreate(&b) != hipSuccess) {
    hipEventDestroy(a);
    return -1.0f;
  }
  float ms = -1.0f;
  if (fn() >= 0 && hipEventRecord(a, c->s) == hipSuccess) {
    bool ok = true;
    for (int r = 0; r < reps && ok; r++) ok = fn() >= 0;
    if (ok && hipEventRecord(b, c->s) == hipSuccess && hipEventSynchronize(b) == hipSuccess)
      hipEventElapsedTime(&ms, a, b);
  }
  hipEventDestroy(a);
  hipEventDestroy(b);
  return ms < 0.0f ? -1.0f : 1e3f * ms / reps;
}

// Best of the candidate values: sets *knob to each, times fn, keeps the
// fastest; appends "name":[[value, us], ...] to the report.
template <typename F>
static int32_t tune_knob(bsam_ctx* c, const char* name, int* knob, const std::vector<int>& cand, F&& fn,
                         std::string& rep) {
  int best = *knob;
  float tbest = 0.0f;
  char buf[64];
  rep += std::string(",\"") + name + "\":[";
  for (size_t n = 0; n < cand.size(); n++) {
    *knob = cand[n];
    const float t = time_launches(c, 3, fn);
    if (t < 0.0f) {
      *knob = best;
      return set_err(BSAM_EHIP, "autotune: %s candidate %d failed", name, cand[n]);
    }
    snprintf(buf, sizeof buf, "%s[%d,%.1f]", n ? "," : "", cand[n], t);
    rep += buf;
    if (n == 0 || t < tbest) {
      tbest = t;
      best = cand[n];
    }
  }
  *knob = best;
  snprintf(buf, sizeof buf, "],\"%s_choice\":%d", name, best);
  rep += buf;
  return BSAM_OK;
}
}  // namespace

// Times the candidate launch shapes of every level with at least 32768
// cells on scratch copies of the fields and keeps the fastest (all shapes
// give bit-identical fields): the block order of the level kernels and of
// the ghost gather (ordered_block: one contiguous range per XCD, or runs of
// 1 to 256 blocks per XCD in turn), and the planes per workgroup of the
// marching sweeps and residuals on one large patch.  The report lists each level's
// patch-size histogram with the timings.
extern "C" int32_t bsam_autotune(bsam_ctx* c) {
  if (!c || !c->have_hier) return set_err(BSAM_ESTATE, "no hierarchy installed");
  if (c->capturing) return set_err(BSAM_ESTATE, "bsam_autotune inside a graph capture");
  std::string rep = "{\"levels\":[";
  bool firstlev = true;
  if (c->p.batching) {
    double* sq = nullptr;
    float* se = nullptr;
    if (hipMalloc(&sq, c->nq * sizeof(double)) != hipSuccess) return set_err(BSAM_ENOMEM, "autotune scratch");
    HIPCHK(hipMemcpyAsync(sq, c->q, c->nq * sizeof(double), hipMemcpyDeviceToDevice, c->s));
    if (c->mixed) {
      if (hipMalloc(&se, c->nq * sizeof(float)) != hipSuccess) {
        hipFree(sq);
        return set_err(BSAM_ENOMEM, "autotune scratch");
      }
      HIPCHK(hipMemcpyAsync(se, c->mixed->e, c->nq * sizeof(float), hipMemcpyDeviceToDevice, c->s));
    }
    double* q0 = c->q;
    float* e0 = c->mixed ? c->mixed->e : nullptr;
    int32_t rc = BSAM_OK;
    for (int level = c->minlevel; level <= c->finest && rc == BSAM_OK; level++) {
      Level& L = c->L(level);
      long cells = 0;
      std::vector<int> hist(8, 0);
      for (int pid : L.patches) {
        const DPatch& P = c->hp[pid];
        cells += (long)P.mx[0] * P.mx[1] * P.mx[2];
        const int ext = std::max(P.mx[0], std::max(P.mx[1], c->p.ndims == 3 ? P.mx[2] : 1));
        int b = 0;
        while (b < 7 && (8 << b) < ext) b++;
        hist[b]++;
      }
      if (cells < 32768) continue;
      char buf[160];
      snprintf(buf, sizeof buf, "%s{\"level\":%d,\"patches\":%d,\"cells\":%ld,\"extent_hist\":{", firstlev ? "" : ",",
               level, (int)L.patches.size(), cells);
      rep += buf;
      firstlev = false;
      bool fh = true;
      for (int b = 0; b < 8; b++)
        if (hist[b]) {
          snprintf(buf, sizeof buf, "%s\"<=%d\":%d", fh ? "" : ",", 8 << b, hist[b]);
          rep += buf;
          fh = false;
        }
      rep += "}";
      c->q = sq;
      int nb = 0;
      const bool march = c->p.model == BSAM_MODEL_CH && sweep_kc(c, L, 8, &nb) > 0;
      const bool wrap = L.wrap_img;
      // the level's colour sweep (with the periodic images it writes on a
      // fully periodic patch): in mixed mode the FP32 correction sweep,
      // which is the sweep that path runs
      if (!march) {
        if (c->mixed) c->mixed->e = se;
        rc = tune_knob(c, "order", &L.order, {0, 1, 4, 16, 64}, [&]() {
          if (c->mixed) return mixed_relax(c, level, 1);
          return (wrap && c->p.model == BSAM_MODEL_CH) ? relax_and_ghost(c, level, 1, false)
                                                         : bsam_relax_level(c, level, 1);
        }, rep);
        if (c->mixed) c->mixed->e = e0;
      }
      if (rc == BSAM_OK && L.maps_ok)
        rc = tune_knob(c, "gather_order_fp64", &L.gorder[0], {0, 1, 16, 64, 256},
                       [&]() { return set_ghost_t<double>(c, level, 0, sq, 0); }, rep);
      if (rc == BSAM_OK && L.maps_ok && se)
        rc = tune_knob(c, "gather_order_fp32", &L.gorder[1], {0, 1, 16, 64, 256},
                       [&]() { return set_ghost_t<float>(c, level, 0, se, 1); }, rep);
      // the gather's workgroup size, with the block order just chosen
      if (rc == BSAM_OK && L.maps_ok)
        rc = tune_knob(c, "gather_block_fp64", &L.gblock[0], {128, 256, 512, 1024},
                       [&]() { return set_ghost_t<double>(c, level, 0, sq, 0); }, rep);
      if (rc == BSAM_OK && L.maps_ok && se)
        rc = tune_knob(c, "gather_block_fp32", &L.gblock[1], {128, 256, 512, 1024},
                       [&]() { return set_ghost_t<float>(c, level, 0, se, 1); }, rep);
      if (rc == BSAM_OK && march) {
        std::vector<int> cand;
        for (int k : {4, 8, 16, 32})
          if (k <= c->hp[L.patches[0]].mx[2]) cand.push_back(k);
        rc = tune_knob(c, "sweep_planes_fp64", &L.kc_sweep[0], cand,
                       [&]() { return relax_and_ghost(c, level, 1, false); }, rep);
        if (rc == BSAM_OK && c->mixed) {
          c->mixed->e = se;
          rc = tune_knob(c, "sweep_planes_fp32", &L.kc_sweep[1], cand,
                         [&]() { return mixed_relax(c, level, 1); }, rep);
          c->mixed->e = e0;
        }
      }
      if (rc == BSAM_OK && c->p.model == BSAM_MODEL_CH && march_kc(c, L, &nb) > 0) {
        std::vector<int> cand;
        for (int k : {4, 8, 16, 32})
          if (k <= c->hp[L.patches[0]].mx[2]) cand.push_back(k);
        rc = tune_knob(c, "resid_planes", &L.kc_resid, cand,
                       [&]() { return resid_march<double>(c, level, c->rf, nullptr); }, rep);
      }
      c->q = q0;
      rep += "}";
    }
    c->q = q0;
    if (c->mixed) c->mixed->e = e0;
    hipStreamSynchronize(c->s);
    hipFree(sq);
    hipFree(se);
    RETCHK(rc);
  }
  rep += "]}";
  c->tune_report = rep;
  // launches captured with the old shapes are rebuilt at the next call
  drop_graph(c);
  return BSAM_OK;
}

// Per-level knob storage, or nullptr.
static int* level_knob(bsam_ctx* c, int32_t knob, Level& L) {
  switch (knob) {
    case BSAM_KNOB_ORDER: return &L.order;
    case BSAM_KNOB_GATHER_ORDER_FP64: return &L.gorder[0];
    case BSAM_KNOB_GATHER_ORDER_FP32: return &L.gorder[1];
    case BSAM_KNOB_SWEEP_PLANES_FP64: return &L.kc_sweep[0];
    case BSAM_KNOB_SWEEP_PLANES_FP32: return &L.kc_sweep[1];
    case BSAM_KNOB_RESID_PLANES: return &L.kc_resid;
    case BSAM_KNOB_GATHER_BLOCK_FP64: return &L.gblock[0];
    case BSAM_KNOB_GATHER_BLOCK_FP32: return &L.gblock[1];
    default: return nullptr;
  }
}

extern "C" int32_t bsam_set_knob(bsam_ctx* c, int32_t knob, int32_t level, int32_t value) {
  if (!c) return set_err(BSAM_EINVAL, "null context");
  if (c->capturing) return set_err(BSAM_ESTATE, "bsam_set_knob inside a graph capture");
  if (knob == BSAM_KNOB_CHAIN_CELLS) {
    if (value < 0) return set_err(BSAM_EINVAL, "chain cells must be >= 0");
    c->chain_cells = value;
    if (c->have_hier) RETCHK(plan_chain(c));
    drop_graph(c);
    return BSAM_OK;
  }
  if (knob == BSAM_KNOB_REGRID_TABLES) {
    if (value != 0 && value != 1) return set_err(BSAM_EINVAL, "regrid tables must be 0 or 1");
    c->regrid_tables = value;
    return BSAM_OK;
  }
  if (knob == BSAM_KNOB_TABLE_PLANNER) {
    if (value != 0 && value != 1) return set_err(BSAM_EINVAL, "table planner must be 0 (host) or 1 (device)");
    c->table_planner = value;
    return BSAM_OK;
  }
  if (knob == BSAM_KNOB_SWEEP_IMAGES) {
    if (value != 0 && value != 1) return set_err(BSAM_EINVAL, "sweep images must be 0 or 1");
    c->sweep_images = value;
    drop_graph(c);
    return BSAM_OK;
  }
  if (!(knob >= 0 && knob <= BSAM_KNOB_RESID_PLANES) && knob != BSAM_KNOB_GATHER_BLOCK_FP64 &&
      knob != BSAM_KNOB_GATHER_BLOCK_FP32)
    return set_err(BSAM_EINVAL, "unknown knob %d", knob);
  RETCHK(check_level(c, level));
  Level& L = c->L(level);
  if (value < 0 || value > 4096) return set_err(BSAM_EINVAL, "knob value %d out of range", value);
  if ((knob == BSAM_KNOB_GATHER_BLOCK_FP64 || knob == BSAM_KNOB_GATHER_BLOCK_FP32) && value != 0 &&
      value != 64 && value != 128 && value != 256 && value != 512 && value != 1024)
    return set_err(BSAM_EINVAL, "gather block %d is not 0, 64, 128, 256, 512 or 1024", value);
  if (knob >= BSAM_KNOB_SWEEP_PLANES_FP64 && knob <= BSAM_KNOB_RESID_PLANES && value > 0) {
    // planes per workgroup: one large 3D patch, at most its z extent
    if (c->p.ndims != 3 || L.patches.size() != 1 || value > c->hp[L.patches[0]].mx[2])
      return set_err(BSAM_EINVAL, "plane count %d does not fit level %d", value, level);
  }
  *level_knob(c, knob, L) = value;
  drop_graph(c);
  return BSAM_OK;
}

extern "C" int32_t bsam_set_mixed_fallback(bsam_ctx* c, double stall_ratio) {
  if (!c) return set_err(BSAM_EINVAL, "null context");
  if (!(stall_ratio >= 0.0)) return set_err(BSAM_EINVAL, "stall ratio must be >= 0");
  c->mixed_stall = stall_ratio;
  return BSAM_OK;
}

extern "C" int32_t bsam_mixed_fallbacks(bsam_ctx* c, int64_t* count) {
  if (!c || !count) return set_err(BSAM_EINVAL, "null argument");
  *count = c->mixed_fallbacks;
  return BSAM_OK;
}

extern "C" int32_t bsam_get_knob(bsam_ctx* c, int32_t knob, int32_t level, int32_t* value) {
  if (!c || !value) return set_err(BSAM_EINVAL, "null argument");
  if (knob == BSAM_KNOB_CHAIN_CELLS) {
    *value = (int32_t)c->chain_cells;
    return BSAM_OK;
  }
  if (knob == BSAM_KNOB_REGRID_TABLES) {
    *value = c->regrid_tables;
    return BSAM_OK;
  }
  if (knob == BSAM_KNOB_TABLE_PLANNER) {
    *value = c->table_planner;
    return BSAM_OK;
  }
  if (knob == BSAM_KNOB_SWEEP_IMAGES) {
    *value = c->sweep_images;
    return BSAM_OK;
  }
  if (!(knob >= 0 && knob <= BSAM_KNOB_RESID_PLANES) && knob != BSAM_KNOB_GATHER_BLOCK_FP64 &&
      knob != BSAM_KNOB_GATHER_BLOCK_FP32)
    return set_err(BSAM_EINVAL, "unknown knob %d", knob);
  RETCHK(check_level(c, level));
  *value = *level_knob(c, knob, c->L(level));
  return BSAM_OK;
}

extern "C" int32_t bsam_tuning_report(const bsam_ctx* c, char* buf, int32_t len) {
  if (!c || !buf || len <= 0) return set_err(BSAM_EINVAL, "null argument");
  const std::string& r = c->tune_report.empty() ? std::string("{}") : c->tune_report;
  snprintf(buf, (size_t)len, "%s", r.c_str());
  return (int32_t)r.size() < len ? BSAM_OK : set_err(BSAM_EINVAL, "report needs %d bytes", (int)r.size() + 1);
}
