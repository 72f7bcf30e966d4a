// Regrid clustering on the host: the patches of a new level from the tagged
// cells of the level below it, as GridAdapt's RefineGrid makes them
// (LIBRARY/bsamroutines.f90:1366-1405): the tags are buffered (BufferAndList
// :1206-1247, InflateEdgeTags :1267-1349), every patch of the level is split
// into subgrids by the signature method of Berger and Rigoutsos as the
// reference modifies it (NewSubGrids :1408-1605, GetSignatures :1608-1632,
// GridFlagRatio :1635-1660), and each subgrid becomes a child descriptor
// (MakeNewGrid :1663-1735).  No device code: clustering reads one int per
// cell of the level once per split pass and is a small part of a regrid.
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <chrono>
#include <stdexcept>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "bsam_grid.h"

namespace {

// The host threads of bsam::parallel_for: workers 1..size() wait for a job
// (a generation count), run their share and count down.  A worker starts
// at the generation current when it is created, so one added for a job
// takes that job and no earlier one (whose function no longer exists).
struct HostPool {
  std::mutex m, busy;
  std::condition_variable cv, done;
  std::vector<std::thread> th;
  const std::function<void(int)>* job = nullptr;
  int njob = 0, left = 0;
  long gen = 0;
  bool stop = false;
  const pid_t pid = getpid();
  ~HostPool() {
    {
      std::lock_guard<std::mutex> g(m);
      stop = true;
    }
    cv.notify_all();
    for (auto& t : th)
      if (t.joinable()) t.join();
  }
  void loop(int w, long seen) {
    for (;;) {
      const std::function<void(int)>* f;
      int n;
      {
        std::unique_lock<std::mutex> lk(m);
        cv.wait(lk, [&] { return stop || gen != seen; });
        if (stop) return;
        seen = gen;
        f = job;
        n = njob;
      }
      if (w < n) {
        try {
          (*f)(w);
        } catch (...) {
        }
      }
      std::lock_guard<std::mutex> g(m);
      if (--left == 0) done.notify_all();
    }
  }
};
HostPool& host_pool() {
  static HostPool p;
  return p;
}

// fn(t), t = 0 .. n-1, on threads started for the call
void spawn_for(int n, const std::function<void(int)>& fn) {
  std::vector<std::thread> th;
  int started = 1;
  try {
    for (; started < n; started++) th.emplace_back([&fn, started]() {
        try {
          fn(started);
        } catch (...) {
        }
      });
  } catch (...) {
  }
  std::exception_ptr err;
  try {
    for (int t = started; t < n; t++) fn(t);
    fn(0);
  } catch (...) {
    err = std::current_exception();
  }
  for (auto& x : th) x.join();
  if (err) std::rethrow_exception(err);
}

}  // namespace

namespace bsam {

void parallel_for(int n, const std::function<void(int)>& fn) {
  if (n <= 1) {
    if (n == 1) fn(0);
    return;
  }
  HostPool& P = host_pool();
  std::unique_lock<std::mutex> busy(P.busy, std::try_to_lock);
  if (!busy.owns_lock() || getpid() != P.pid) {
    spawn_for(n, fn);
    return;
  }
  try {
    while ((int)P.th.size() < n - 1) {
      const int w = (int)P.th.size() + 1;
      // gen changes only under `busy`, which this thread holds
      const long g0 = P.gen;
      P.th.emplace_back([&P, w, g0]() { P.loop(w, g0); });
    }
  } catch (...) {
  }
  const int nw = (int)P.th.size();
  {
    std::lock_guard<std::mutex> g(P.m);
    P.job = &fn;
    P.njob = n;
    P.left = nw;
    P.gen++;
  }
  P.cv.notify_all();
  // the workers use fn until they have counted down: wait for them before
  // an exception of this thread's share leaves
  std::exception_ptr err;
  try {
    fn(0);
    for (int t = nw + 1; t < n; t++) fn(t);
  } catch (...) {
    err = std::current_exception();
  }
  {
    std::unique_lock<std::mutex> lk(P.m);
    P.done.wait(lk, [&] { return P.left == 0; });
  }
  if (err) std::rethrow_exception(err);
}

void periodic_offsets(int ndims, const bsam_patch_desc& root, int level,
                      std::vector<std::array<int, 3>>& offs, bool& periodic) {
  // the offset table of GetPeriodicOffsets, cases 1-13
  static const int tab[13][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {-1, 1, 0},
                                 {0, 1, 1}, {0, 1, -1}, {1, 0, 1}, {-1, 0, 1}, {1, 1, 1},
                                 {1, 1, -1}, {-1, 1, 1}, {1, -1, 1}};
  int m[6];
  for (int i = 0; i < 6; i++) m[i] = 10;
  for (int i = 0; i < 2 * ndims; i++) m[i] = root.mthbc[i];
  // PeriodicSetup's choice of cases from the periodic dimensions
  std::vector<int> idx;
  periodic = true;
  if (m[0] == 2) {
    if (m[2] == 2) idx = (m[4] == 2) ? std::vector<int>{1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13}
                                     : std::vector<int>{1, 2, 4, 5};
    else idx = (m[4] == 2) ? std::vector<int>{1, 3, 8, 9} : std::vector<int>{1};
  } else {
    if (m[2] == 2) idx = (m[4] == 2) ? std::vector<int>{2, 3, 6, 7} : std::vector<int>{2};
    else if (m[4] == 2) idx = {3};
    else periodic = false;
  }
  offs.clear();
  if (level < 0) return;
  const int s = 1 << level;
  for (int x : idx) {
    std::array<int, 3> o;
    for (int a = 0; a < 3; a++) o[a] = tab[x - 1][a] * root.mx[a] * s;
    offs.push_back(o);
  }
}

}  // namespace bsam

namespace {

constexpr int kMaxSubgrids = 1024;   // maxsubgrids, nodeinfodef.f90:62
constexpr int kMaxSplitPasses = 15;  // NewSubGrids' maxsplitpasses
constexpr int kInternalBC = 999;     // internalbc, nodeinfodef.f90:62

struct Box {
  int lo[3], hi[3];
};

// One patch's tags, column-major over its interior (1-based access).
struct Tags {
  int mx[3];
  int32_t* f;
  int32_t& at(int i, int j, int k) const { return f[(i - 1) + (long)mx[0] * ((j - 1) + (long)mx[1] * (k - 1))]; }
};

// Summed tags of one patch: at(i, j, k) = the tags of [1, i] x [1, j] x
// [1, k] (0 where an index is 0), so the tags of any box are eight lookups.
struct TagSums {
  int sx = 0, sy = 0;
  std::vector<int> s;
  void build(const Tags& T) {
    sx = T.mx[0] + 1;
    sy = T.mx[1] + 1;
    s.assign((size_t)sx * sy * (T.mx[2] + 1), 0);
    for (int k = 1; k <= T.mx[2]; k++)
      for (int j = 1; j <= T.mx[1]; j++) {
        int row = 0;
        for (int i = 1; i <= T.mx[0]; i++) {
          row += T.at(i, j, k);
          at(i, j, k) = row + at(i, j - 1, k) + at(i, j, k - 1) - at(i, j - 1, k - 1);
        }
      }
  }
  int& at(int i, int j, int k) { return s[((size_t)k * sy + j) * sx + i]; }
  int at(int i, int j, int k) const { return s[((size_t)k * sy + j) * sx + i]; }
  // tags of the box [i0, i1] x [j0, j1] x [k0, k1]
  int box(int i0, int i1, int j0, int j1, int k0, int k1) const {
    i0--, j0--, k0--;
    return at(i1, j1, k1) - at(i0, j1, k1) - at(i1, j0, k1) - at(i1, j1, k0) + at(i0, j0, k1) + at(i0, j1, k0) +
           at(i1, j0, k0) - at(i0, j0, k0);
  }
};

// GetSignatures: sig[n][i] = the sum of the tags on the slab i of box b
// across dimension n (entries outside b are left alone), each slab from the
// summed tags.
void signatures(const TagSums& S, const Box& b, int nd, std::vector<int> sig[3]) {
  for (int i = b.lo[0]; i <= b.hi[0]; i++) sig[0][i] = S.box(i, i, b.lo[1], b.hi[1], b.lo[2], b.hi[2]);
  for (int j = b.lo[1]; j <= b.hi[1]; j++) sig[1][j] = S.box(b.lo[0], b.hi[0], j, j, b.lo[2], b.hi[2]);
  if (nd == 3)
    for (int k = b.lo[2]; k <= b.hi[2]; k++) sig[2][k] = S.box(b.lo[0], b.hi[0], b.lo[1], b.hi[1], k, k);
}

// NewSubGrids of one patch: the subgrid boxes (patch-local indices) in the
// order MakeNewGrid creates them; on failure BSAM_EINVAL with `msg` set.
int32_t new_subgrids(const Tags& T, int nd, int mgp, double desfill, int level, std::vector<Box>& sub,
                     std::string& msg, TagSums& S) {
  sub.clear();
  char buf[160];
  auto fail = [&](const char* what) {
    snprintf(buf, sizeof buf, "level %d: %s", level, what);
    msg = buf;
    return BSAM_EINVAL;
  };
  if (mgp < 4) return fail("minimumgridpoints cannot be less than 4");
  if (mgp % 2) return fail("minimumgridpoints must be even");
  Box whole;
  for (int n = 0; n < 3; n++) {
    whole.lo[n] = 1;
    whole.hi[n] = n < nd ? T.mx[n] : 1;
  }
  S.build(T);
  // GridFlagRatio of the whole patch
  const double cells0 = (double)T.mx[0] * T.mx[1] * T.mx[2];
  if ((double)S.box(1, T.mx[0], 1, T.mx[1], 1, T.mx[2]) / cells0 < 1.0e-10) return BSAM_OK;
  std::vector<int> sig[3], dd[3];
  for (int n = 0; n < nd; n++) {
    sig[n].assign(T.mx[n] + 2, 0);
    dd[n].assign(T.mx[n] + 2, 0);
  }
  sub.push_back(whole);
  std::vector<char> can(1, 1);
  for (size_t g = 0; sub.size() < (size_t)kMaxSubgrids && g < sub.size(); g++) {
    for (int pass = 0; can[g] && pass < kMaxSplitPasses; pass++) {
      // trim untagged edge slabs (signatures of the box before trimming),
      // then widen to whole coarse cell pairs: lo odd, hi even
      signatures(S, sub[g], nd, sig);
      for (int n = 0; n < nd; n++) {
        int i1 = sub[g].lo[n], i2 = sub[g].hi[n];
        while (sig[n][i1] == 0 && i1 < sub[g].hi[n] && i2 - i1 + 1 > mgp) i1++;
        while (sig[n][i2] == 0 && i2 > sub[g].lo[n] && i2 - i1 + 1 > mgp) i2--;
        sub[g].lo[n] = i1 - (i1 + 1) % 2;
        sub[g].hi[n] = i2 + i2 % 2;
      }
      // GridFlagRatio of the trimmed box, from its signatures (their sum
      // over one dimension counts the box's tags exactly)
      signatures(S, sub[g], nd, sig);
      long tagged = 0, cells = 1;
      for (int i = sub[g].lo[0]; i <= sub[g].hi[0]; i++) tagged += sig[0][i];
      for (int n = 0; n < nd; n++) cells *= (long)(sub[g].hi[n] - sub[g].lo[n] + 1);
      if (!((double)tagged / (double)cells < desfill)) {
        can[g] = 0;
        continue;
      }
      // a hole: the first untagged slab (last index of the first half) that
      // leaves both halves at least mgp wide
      int isplit[3] = {0, 0, 0};
      bool have = false;
      for (int n = 0; n < nd && !have; n++) {
        const int i1 = sub[g].lo[n], i2 = sub[g].hi[n];
        for (int i = i1 + mgp - 1; i <= i2 - mgp; i += 2)
          if (sig[n][i] == 0 && std::min(i - i1 + 1, i2 - i) >= mgp) {
            isplit[n] = i;
            have = true;
            break;
          }
      }
      if (!have) {
        // else the strongest inflection of the signatures' second
        // difference, the most central one among equals
        for (int n = 0; n < nd; n++)
          for (int i = sub[g].lo[n] + 1; i <= sub[g].hi[n] - 1; i++)
            dd[n][i] = sig[n][i - 1] - 2 * sig[n][i] + sig[n][i + 1];
        int inflect = 0, dist0 = 0;
        for (int n = 0; n < nd; n++) {
          const int i1 = sub[g].lo[n], i2 = sub[g].hi[n];
          for (int i = i1 + mgp; i <= i2 - mgp + 1; i += 2) {
            // i: first index of the second half
            const int del = std::abs(dd[n][i] - dd[n][i - 1]);
            if (del > inflect) {
              inflect = del;
              isplit[0] = isplit[1] = isplit[2] = 0;
              isplit[n] = i - 1;
              have = true;
              dist0 = std::min(i - i1, i2 - i + 1);
            } else if (del == inflect && inflect > 0) {
              const int dist1 = std::min(i - i1, i2 - i + 1);
              if (dist1 > dist0) {
                isplit[0] = isplit[1] = isplit[2] = 0;
                isplit[n] = i - 1;
                have = true;
                dist0 = dist1;
              }
            }
          }
        }
      }
      if (!have) {
        can[g] = 0;
        continue;
      }
      for (int n = 0; n < nd; n++) {
        if (isplit[n] > 0 && std::min(sub[g].hi[n] - isplit[n], isplit[n] - sub[g].lo[n] + 1) >= mgp) {
          if (sub.size() >= (size_t)kMaxSubgrids) return fail("more than 1024 subgrids in one patch");
          Box b = sub[g];
          b.lo[n] = isplit[n] + 1;
          sub[g].hi[n] = isplit[n];
          sub.push_back(b);
          can.push_back(1);
          break;
        }
      }
    }
  }
  for (const Box& b : sub)
    for (int n = 0; n < nd; n++) {
      const int len = b.hi[n] - b.lo[n] + 1;
      if (len < mgp) return fail("subgrid smaller than minimumgridpoints");
      if (len % 2 || b.lo[n] % 2 == 0 || b.hi[n] % 2) return fail("subgrid not aligned to coarse cell pairs");
    }
  return BSAM_OK;
}

// fn(t) for t = 0 .. n-1 on the host pool.
template <typename F>
void for_threads(int n, F&& fn) {
  const std::function<void(int)> f = [&fn](int t) { fn(t); };
  bsam::parallel_for(n, f);
}

int host_threads() {
  const unsigned hw = std::thread::hardware_concurrency();
  return hw ? (int)std::min(16u, hw) : 4;
}

// MakeNewGrid's descriptor of the child of patch `pidx` over its local box b.
bsam_patch_desc make_child(const bsam_patch_desc& P, int pidx, const Box& b, int nd) {
  bsam_patch_desc d{};
  d.level = P.level + 1;
  d.parent = pidx;
  for (int a = 0; a < 3; a++) {
    if (a < nd) {
      d.mx[a] = (b.hi[a] - b.lo[a] + 1) * 2;
      d.mbounds[a][0] = b.lo[a];
      d.mbounds[a][1] = b.hi[a];
      const int g0 = P.mglobal[a][0] + b.lo[a] - 1;
      const int g1 = g0 + b.hi[a] - b.lo[a];
      d.mglobal[a][0] = (g0 - 1) * 2 + 1;
      d.mglobal[a][1] = g1 * 2;
      d.xlower[a] = (double)(b.lo[a] - 1) * P.dx + P.xlower[a];
    } else {
      d.mx[a] = 1;
      d.mbounds[a][0] = d.mbounds[a][1] = 1;
      d.mglobal[a][0] = d.mglobal[a][1] = 1;
      d.xlower[a] = 0.0;
    }
  }
  // internal faces, except where the child reaches a physical (or
  // periodic) face of its parent, whose condition it inherits
  for (int f = 0; f < 6; f++) d.mthbc[f] = kInternalBC;
  for (int n = 0; n < nd; n++) {
    if (P.mthbc[2 * n] < kInternalBC && b.lo[n] == 1) d.mthbc[2 * n] = P.mthbc[2 * n];
    if (P.mthbc[2 * n + 1] < kInternalBC && b.hi[n] == P.mx[n]) d.mthbc[2 * n + 1] = P.mthbc[2 * n + 1];
  }
  d.dx = P.dx / 2.0;
  return d;
}

}  // namespace

// Host self-test of bsam::parallel_for (include/bsam_gpu.h).  Each task
// marks its index and, after a short wait, a "done" flag; a job returning
// before every flag is set, an index run twice or not at all, or an
// exception lost or invented fails the test.
extern "C" int32_t bsam_selftest_thread_pool(int32_t rounds, int32_t n_max, int64_t* jobs) {
  if (rounds < 1 || n_max < 2 || n_max > 256 || !jobs) return bsam::fail(BSAM_EINVAL, "bad arguments");
  *jobs = 0;
  for (int r = 0; r < rounds; r++) {
    // counts that grow within a round (new workers join mid-sequence), then
    // shrink; every third job throws from one task
    const int counts[5] = {1 + r % 3, 2 + (r * 7) % (n_max - 1), n_max, 2, 1 + (r * 5) % n_max};
    for (int n : counts) {
      std::vector<std::atomic<int>> hit(n), done(n);
      for (int t = 0; t < n; t++) hit[t] = 0, done[t] = 0;
      const bool thrower = (*jobs % 3) == 2;
      const int bad = n / 2;
      bool caught = false;
      const std::function<void(int)> f = [&](int t) {
        hit[t]++;
        std::this_thread::sleep_for(std::chrono::microseconds(t % 4 == 0 ? 200 : 20));
        done[t] = 1;
        if (thrower && t == bad) throw std::runtime_error("task failure");
      };
      try {
        bsam::parallel_for(n, f);
      } catch (const std::runtime_error&) {
        caught = true;
      }
      for (int t = 0; t < n; t++) {
        if (hit[t] != 1)
          return bsam::fail(BSAM_EINVAL, "job %lld (%d tasks): task %d ran %d times", (long long)*jobs, n, t,
                            (int)hit[t]);
        if (!done[t])
          return bsam::fail(BSAM_EINVAL, "job %lld (%d tasks): returned while task %d still ran", (long long)*jobs, n,
                            t);
      }
      // a task's exception reaches the caller when the calling thread ran
      // that task; the workers' are contained (a worker cannot rethrow)
      if (caught && !thrower) return bsam::fail(BSAM_EINVAL, "job %lld: an exception that no task threw", (long long)*jobs);
      (*jobs)++;
    }
  }
  return BSAM_OK;
}

extern "C" int32_t bsam_new_level(int32_t ndims, int32_t npatch, const bsam_patch_desc* desc, int32_t level,
                                  const int32_t* flags, int32_t ibuffer, int32_t minimumgridpoints,
                                  double desiredfillratio, int32_t maxout, bsam_patch_desc* out,
                                  int32_t* nout) {
  if ((ndims != 2 && ndims != 3) || npatch <= 0 || !desc || !flags || !nout || (maxout > 0 && !out))
    return bsam::fail(BSAM_EINVAL, "bad arguments");
  if (level < 0) return bsam::fail(BSAM_EINVAL, "level %d: new grids are made above the root only", level);
  if (ibuffer < 0) return bsam::fail(BSAM_EINVAL, "ibuffer must be >= 0");
  *nout = 0;
  const bsam_patch_desc* root = nullptr;
  std::vector<int> lv;  // the level's patches, visiting order
  for (int i = 0; i < npatch; i++) {
    if (desc[i].level == 0 && !root) root = &desc[i];
    if (desc[i].level == level) lv.push_back(i);
  }
  if (!root) return bsam::fail(BSAM_EINVAL, "no root patch in the table");
  if (lv.empty()) return BSAM_OK;
  // the level's tags, buffered in a copy
  std::vector<long> off(lv.size() + 1, 0);
  for (size_t a = 0; a < lv.size(); a++)
    off[a + 1] = off[a] + (long)desc[lv[a]].mx[0] * desc[lv[a]].mx[1] * desc[lv[a]].mx[2];
  // buffering writes into a copy of the tags; without it they are only read
  std::vector<int32_t> tg;
  if (ibuffer > 0) tg.assign(flags, flags + off.back());
  int32_t* tags = ibuffer > 0 ? tg.data() : const_cast<int32_t*>(flags);
  std::vector<Tags> T(lv.size());
  for (size_t a = 0; a < lv.size(); a++) {
    for (int n = 0; n < 3; n++) T[a].mx[n] = desc[lv[a]].mx[n];
    T[a].f = tags + off[a];
  }
  const int nth = std::max(1, std::min<int>((int)lv.size(), host_threads()));
  if (ibuffer > 0) {
    // BufferAndList: a tag whose buffer box stays inside its patch is
    // buffered there; the others are listed in global indices.  Buffering
    // only ever sets flags to 1 and reads the unbuffered tags, so the
    // patches, and then the listed tags, are spread over host threads
    // (flags set by two threads at once are written with relaxed atomic
    // stores).
    std::vector<std::vector<std::array<int, 3>>> edges(lv.size());
    for_threads(nth, [&](int t) {
      for (size_t a = t; a < lv.size(); a += nth) {
        const bsam_patch_desc& P = desc[lv[a]];
        const Tags O{{P.mx[0], P.mx[1], P.mx[2]}, const_cast<int32_t*>(flags) + off[a]};
        for (int k = 1; k <= P.mx[2]; k++)
          for (int j = 1; j <= P.mx[1]; j++)
            for (int i = 1; i <= P.mx[0]; i++) {
              if (O.at(i, j, k) != 1) continue;
              const int idx[3] = {i, j, k};
              int lo[3] = {1, 1, 1}, hi[3] = {1, 1, 1};
              bool inside = true;
              for (int n = 0; n < ndims; n++) {
                lo[n] = idx[n] - ibuffer;
                hi[n] = idx[n] + ibuffer;
                inside = inside && lo[n] >= 1 && hi[n] <= P.mx[n];
              }
              if (inside) {
                for (int z = lo[2]; z <= hi[2]; z++)
                  for (int y = lo[1]; y <= hi[1]; y++)
                    for (int x = lo[0]; x <= hi[0]; x++) T[a].at(x, y, z) = 1;
              } else {
                std::array<int, 3> g = {1, 1, 1};
                for (int n = 0; n < ndims; n++) g[n] = idx[n] + P.mglobal[n][0] - 1;
                edges[a].push_back(g);
              }
            }
      }
    });
    std::vector<std::array<int, 3>> edge;
    for (auto& e : edges) edge.insert(edge.end(), e.begin(), e.end());
    // InflateEdgeTags: the buffer box of every listed tag, and of its
    // periodic images, on every patch of the level it meets (a bin index
    // over the level's global index space finds the patches)
    if (!edge.empty()) {
      std::vector<std::array<int, 3>> offs;
      bool periodic = false;
      bsam::periodic_offsets(ndims, *root, level, offs, periodic);
      constexpr int kBin = 16;
      int ext[3] = {1, 1, 1}, nb[3] = {1, 1, 1};
      for (int n = 0; n < ndims; n++) {
        ext[n] = root->mx[n] << level;
        nb[n] = (ext[n] + kBin - 1) / kBin;
      }
      std::vector<std::vector<int>> bins((size_t)nb[0] * nb[1] * nb[2]);
      auto bin_range = [&](const int* lo, const int* hi, int* b0, int* b1) {
        for (int n = 0; n < 3; n++) {
          if (n >= ndims) { b0[n] = b1[n] = 0; continue; }
          b0[n] = (std::max(lo[n], 1) - 1) / kBin;
          b1[n] = (std::min(hi[n], ext[n]) - 1) / kBin;
        }
      };
      for (size_t a = 0; a < lv.size(); a++) {
        const bsam_patch_desc& P = desc[lv[a]];
        int lo[3], hi[3], b0[3], b1[3];
        for (int n = 0; n < 3; n++) lo[n] = P.mglobal[n][0], hi[n] = P.mglobal[n][1];
        bin_range(lo, hi, b0, b1);
        for (int z = b0[2]; z <= b1[2]; z++)
          for (int y = b0[1]; y <= b1[1]; y++)
            for (int x = b0[0]; x <= b1[0]; x++) bins[((size_t)z * nb[1] + y) * nb[0] + x].push_back((int)a);
      }
      const int eth = std::max(1, std::min<int>((int)(edge.size() / 256) + 1, host_threads()));
      for_threads(eth, [&](int t) {
        std::vector<long> stamp(lv.size(), -1);
        long tick = 0;
        auto apply = [&](const std::array<int, 3>& cc) {
          int lo[3] = {1, 1, 1}, hi[3] = {1, 1, 1}, b0[3], b1[3];
          for (int n = 0; n < ndims; n++) lo[n] = cc[n] - ibuffer, hi[n] = cc[n] + ibuffer;
          for (int n = 0; n < ndims; n++)
            if (hi[n] < 1 || lo[n] > ext[n]) return;
          bin_range(lo, hi, b0, b1);
          tick++;
          for (int z = b0[2]; z <= b1[2]; z++)
            for (int y = b0[1]; y <= b1[1]; y++)
              for (int x = b0[0]; x <= b1[0]; x++)
                for (int a : bins[((size_t)z * nb[1] + y) * nb[0] + x]) {
                  if (stamp[a] == tick) continue;
                  stamp[a] = tick;
                  const bsam_patch_desc& P = desc[lv[a]];
                  int ml[3] = {1, 1, 1}, mh[3] = {1, 1, 1};
                  bool hit = true;
                  for (int n = 0; n < ndims && hit; n++) {
                    const int o0 = std::max(lo[n], P.mglobal[n][0]), o1 = std::min(hi[n], P.mglobal[n][1]);
                    hit = o1 >= o0;
                    ml[n] = o0 - P.mglobal[n][0] + 1;
                    mh[n] = o1 - P.mglobal[n][0] + 1;
                  }
                  if (!hit) continue;
                  for (int z2 = ml[2]; z2 <= mh[2]; z2++)
                    for (int y2 = ml[1]; y2 <= mh[1]; y2++)
                      for (int x2 = ml[0]; x2 <= mh[0]; x2++) __atomic_store_n(&T[a].at(x2, y2, z2), 1, __ATOMIC_RELAXED);
                }
        };
        for (size_t e = t; e < edge.size(); e += eth) {
          const std::array<int, 3>& cc = edge[e];
          apply(cc);
          if (!periodic) continue;
          for (int pol = -1; pol <= 1; pol += 2)
            for (const auto& o : offs) {
              std::array<int, 3> sh = cc;
              for (int n = 0; n < ndims; n++) sh[n] = cc[n] + pol * o[n];
              apply(sh);
            }
        }
      });
    }
  }
  // RefineGrid over the level in visiting order: the patches' subgrids are
  // independent, found on host threads (at most 16), and MakeNewGrid runs
  // in the reference's order
  std::vector<std::vector<Box>> sub(lv.size());
  std::vector<int32_t> rc(lv.size(), BSAM_OK);
  std::vector<std::string> msg(lv.size());
  // largest patches first, taken from a shared counter: a few large
  // parents would otherwise hold up the threads that drew them
  std::vector<int> big(lv.size());
  for (size_t a = 0; a < lv.size(); a++) big[a] = (int)a;
  std::stable_sort(big.begin(), big.end(), [&](int x, int y) { return off[x + 1] - off[x] > off[y + 1] - off[y]; });
  std::atomic<size_t> next(0);
  for_threads(nth, [&](int) {
    TagSums S;
    for (size_t n; (n = next.fetch_add(1)) < big.size();) {
      const int a = big[n];
      rc[a] = new_subgrids(T[a], ndims, minimumgridpoints, desiredfillratio, level, sub[a], msg[a], S);
    }
  });
  std::vector<bsam_patch_desc> made;
  for (size_t a = 0; a < lv.size(); a++) {
    if (rc[a] != BSAM_OK) return bsam::fail(rc[a], "%s", msg[a].c_str());
    for (const Box& b : sub[a]) made.push_back(make_child(desc[lv[a]], lv[a], b, ndims));
  }
  *nout = (int32_t)made.size();
  if ((int32_t)made.size() > maxout)
    return bsam::fail(BSAM_EINVAL, "%d new patches do not fit in %d descriptors", (int)made.size(), maxout);
  // the tree's level list runs from the youngest node: visiting order is
  // the reverse of creation
  for (size_t n = 0; n < made.size(); n++) out[n] = made[made.size() - 1 - n];
  return BSAM_OK;
}

// ---------------------------------------------------------------------------
// ReadQ (LIBRARY/bsaminputoutput.f90:184-443): a frame as WriteQ writes it
// (time F25.12; finestlevel I3; per patch of levels 0..finest in the
// writer's order a blank record, level/ndims/r/nrvars (I3,3(1x,I3)), dx,
// xlower and xupper (F25.12 with 1x), mx and mglobal (I8 with 1X), then one
// record per cell of q(0:mx+1, ...), i fastest, nrvars values each).  The
// reference reads it into a tree whose level lists run from the youngest
// node, so a level's patches are visited in the reverse of the file's
// order; it leaves a refined patch's parent (the youngest node one level
// down), mbounds and mthbc broken, since only TransferValues reads those
// patches before the next regrid replaces them.  Here the parent is the
// patch one level down whose box contains the patch's coarse box, mbounds
// follow from it and mthbc is MakeNewGrid's inheritance, so the table is a
// valid hierarchy for any step.
// ---------------------------------------------------------------------------
namespace {

struct FrameReader {
  FILE* f = nullptr;
  std::string line;
  long lineno = 0;
  ~FrameReader() {
    if (f) fclose(f);
  }
  bool next() {
    line.clear();
    int ch;
    while ((ch = fgetc(f)) != EOF && ch != '\n') line.push_back((char)ch);
    if (ch == EOF && line.empty()) return false;
    lineno++;
    return true;
  }
};

// whitespace-separated numbers of the current line
bool read_numbers(const std::string& s, int n, double* out) {
  const char* p = s.c_str();
  for (int k = 0; k < n; k++) {
    char* end = nullptr;
    while (*p == ' ') p++;
    if (!*p || *p == '*') return false;
    out[k] = strtod(p, &end);
    if (end == p) return false;
    p = end;
  }
  return true;
}
bool read_ints(const std::string& s, int n, int* out) {
  double v[16];
  if (n > 16 || !read_numbers(s, n, v)) return false;
  for (int k = 0; k < n; k++) {
    out[k] = (int)v[k];
    if ((double)out[k] != v[k]) return false;
  }
  return true;
}

}  // namespace

extern "C" int32_t bsam_parse_frame(const char* path, int32_t ndims, int32_t nrvars, double* time,
                                    int32_t* finestlevel, int32_t maxpatch, bsam_patch_desc* desc, int32_t* npatch,
                                    double* q, int64_t maxq, int64_t* nq) {
  if (!path || (ndims != 2 && ndims != 3) || nrvars < 1 || nrvars > 10 || !time || !finestlevel || !npatch || !nq)
    return bsam::fail(BSAM_EINVAL, "bad arguments");
  FrameReader R;
  R.f = fopen(path, "rb");
  if (!R.f) return bsam::fail(BSAM_EINVAL, "ReadQ: cannot open %s", path);
  auto bad = [&](const char* what) {
    return bsam::fail(BSAM_EINVAL, "ReadQ: %s at line %ld of %s", what, R.lineno, path);
  };
  if (!R.next() || !read_numbers(R.line, 1, time)) return bad("no time");
  int finest = 0;
  if (!R.next() || !read_ints(R.line, 1, &finest)) return bad("no finest level");
  if (finest < 0 || finest >= BSAM_MAX_AMR_LEVELS) return bad("finest level out of range");
  struct P {
    bsam_patch_desc d;
    std::vector<double> q;
  };
  std::vector<std::vector<P>> lv(finest + 1);
  int prev = 0;
  while (R.next()) {
    if (R.line.find_first_not_of(' ') != std::string::npos) return bad("a patch does not start with a blank record");
    int h[4];
    if (!R.next() || !read_ints(R.line, 4, h)) return bad("bad patch header");
    const int level = h[0];
    if (h[1] != ndims) return bad("spatial dimension (ndims) inconsistent");
    if (h[2] != 2) return bad("refinement ratio other than 2");
    if (h[3] != nrvars) return bad("number of variables (nrvars) inconsistent");
    if (level < prev || level > finest) return bad("patch levels out of order");
    if (level == 0 && !lv[0].empty()) return bad("more than one root patch");
    prev = level;
    P p{};
    bsam_patch_desc& d = p.d;
    double dx[3], xl[3], xu[3];
    int mx[3], mg[6];
    if (!R.next() || !read_numbers(R.line, ndims, dx)) return bad("bad dx");
    if (!R.next() || !read_numbers(R.line, ndims, xl)) return bad("bad xlower");
    if (!R.next() || !read_numbers(R.line, ndims, xu)) return bad("bad xupper");
    if (!R.next() || !read_ints(R.line, ndims, mx)) return bad("bad mx");
    if (!R.next() || !read_ints(R.line, 2 * ndims, mg)) return bad("bad mglobal");
    d.level = level;
    d.parent = -1;
    d.dx = dx[0];
    for (int a = 0; a < 3; a++) {
      const bool used = a < ndims;
      d.mx[a] = used ? mx[a] : 1;
      if (used && (mx[a] < 2 || mg[2 * a + 1] - mg[2 * a] + 1 != mx[a])) return bad("mx and mglobal disagree");
      d.mglobal[a][0] = used ? mg[2 * a] : 1;
      d.mglobal[a][1] = used ? mg[2 * a + 1] : 1;
      d.mbounds[a][0] = 1;
      d.mbounds[a][1] = used ? (level == 0 ? mx[a] : mx[a] / 2) : 1;
      d.xlower[a] = used ? xl[a] : 0.0;
    }
    for (int f2 = 0; f2 < 6; f2++) d.mthbc[f2] = 999;
    long cells = 1;
    for (int a = 0; a < ndims; a++) cells *= mx[a] + 2;
    p.q.resize((size_t)cells * nrvars);
    double v[10];
    // ReadQ2D/3D: q(i,j[,k],1:nrvars) record by record, i fastest
    for (long cidx = 0; cidx < cells; cidx++) {
      if (!R.next() || !read_numbers(R.line, nrvars, v)) return bad("bad cell record");
      for (int k = 0; k < nrvars; k++) p.q[(size_t)k * cells + cidx] = v[k];
    }
    lv[level].push_back(std::move(p));
  }
  if (lv[0].empty()) return bad("no root patch");
  // the reference's visiting order: the root, then each level's patches
  // from the youngest (the last in the file)
  std::vector<P*> order;
  std::vector<int> first(finest + 2, 0);
  for (int l = 0; l <= finest; l++) {
    first[l] = (int)order.size();
    for (int n = (int)lv[l].size() - 1; n >= 0; n--) order.push_back(&lv[l][n]);
  }
  first[finest + 1] = (int)order.size();
  *finestlevel = finest;
  *npatch = (int32_t)order.size();
  long total = 0;
  for (P* p : order) total += (long)p->q.size();
  *nq = total;
  if ((desc && (int32_t)order.size() > maxpatch) || (q && total > maxq))
    return bsam::fail(BSAM_EINVAL, "ReadQ: %d patches and %ld values do not fit in %d and %ld",
                      (int)order.size(), total, maxpatch, (long)maxq);
  // the root keeps its own faces: the caller's root descriptor supplies
  // them; a child's parent is the patch one level down that holds its
  // coarse box, and MakeNewGrid's rules give mbounds and mthbc
  for (int l = 1; l <= finest; l++)
    for (int n = first[l]; n < first[l + 1]; n++) {
      bsam_patch_desc& d = order[n]->d;
      int pick = -1;
      for (int m = first[l - 1]; m < first[l] && pick < 0; m++) {
        const bsam_patch_desc& pd = order[m]->d;
        bool in = true;
        for (int a = 0; a < ndims && in; a++) {
          const int c0 = (d.mglobal[a][0] + 1) / 2, c1 = d.mglobal[a][1] / 2;
          in = (d.mglobal[a][0] % 2 == 1) && (d.mglobal[a][1] % 2 == 0) && c0 >= pd.mglobal[a][0] &&
               c1 <= pd.mglobal[a][1];
        }
        if (in) pick = m;
      }
      if (pick < 0) return bsam::fail(BSAM_EINVAL, "ReadQ: a level-%d patch lies in no patch of level %d", l, l - 1);
      d.parent = pick;
      const bsam_patch_desc& pd = order[pick]->d;
      for (int a = 0; a < ndims; a++) {
        d.mbounds[a][0] = (d.mglobal[a][0] + 1) / 2 - pd.mglobal[a][0] + 1;
        d.mbounds[a][1] = d.mglobal[a][1] / 2 - pd.mglobal[a][0] + 1;
      }
    }
  for (int n = 0; n < (int)order.size(); n++) {
    if (desc) desc[n] = order[n]->d;
  }
  if (q) {
    long off = 0;
    for (P* p : order) {
      std::copy(p->q.begin(), p->q.end(), q + off);
      off += (long)p->q.size();
    }
  }
  return BSAM_OK;
}
