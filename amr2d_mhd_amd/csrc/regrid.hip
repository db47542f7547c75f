// Regrid decision and quadtree surgery on the device: admesh.cpp:17-102 with
// markref.cpp:6-18, the thresholds of criteria.cpp:8-43, critneighup.cpp:6-35,
// critneighdown.cpp:6-33, updatelpup.cpp:4-118 and updatelpdown.cpp:5-54,
// over a device copy of each instance's tree, one workgroup per instance.
//
// The reference walks blocks one at a time.  Here every pass runs over all
// blocks of one level at once, with these equivalences:
//  * critneighup at level nl only writes iref of coarser blocks, which are read
//    at later levels: one pass per level gives the serial result.
//  * The refinements of one level write the same final pointers in any order
//    (two refined neighbours link their sons to each other whichever goes
//    first); new son groups are numbered in block order by a prefix count, so
//    every block gets the ids the serial loop gives it.
//  * critneighdown reads leaf flags that its loop does not change, and the
//    choice of the groups that coarsen reads only fields coarsening does not
//    change, so the groups of every level are chosen in one pass.
//  * The coarsenings of one level differ with order only in the stale sons of
//    two neighbouring dads: the serial loop updates a neighbour's sons when the
//    neighbour is still refined when the loop reaches this dad, i.e. unless its
//    own sons were coarsened earlier in the same level (a smaller first-son id).
// Passes that must follow block order give each lane a contiguous block range
// and order a whole level with one prefix sum over the lanes; levels with
// nothing to refine or coarsen are skipped.  The CFL minimum of
// main.cpp:208-230 (min over the instance's leaves, from 1e6) is taken by the
// same workgroup; min over non-NaN values does not depend on the order.
//
// The tree is staged in LDS (16-bit rows and byte flags) when the host's bound
// on the rows a regrid can reach fits, and written back only if it changed;
// otherwise the workgroup works on the device copy in place.
//
// Results per instance: the CFL minimum, the refinements (parent ids) and
// coarsenings (first-son ids) in the reference's order, the new lastActive,
// and an overflow status (admesh.cpp:42-45) at the op where the serial loop
// throws, with the tree as the serial loop leaves it there (kErrConfig: the
// op log would outgrow the instance's leaf range, which a consistent tree
// never does; the host then reports an internal error).
#include <hip/hip_runtime.h>

#include <climits>

#include "amr_types.hpp"
#include "kernels.hpp"
#include "physics.hpp"

namespace amr {

namespace {

constexpr int kRgThreads = 1024;
constexpr int kRgWaves = kRgThreads / 64;
constexpr int kStageUnroll = 8;  // 16-byte loads in flight per lane while staging a tree
static_assert(kRgLdsRows * 20 + 512 <= 160 * 1024 && kRgLdsRows % 4 == 0, "staged tree must fit the 160 KB of LDS");

// one instance's tree: rows [block][8] (level, father, first son, sibling
// id, left, right, bottom, top) of type R, and per-block flags; `mark` holds
// the first sons of the groups that coarsen in this regrid (cleared level by
// level as they are applied)
template <typename R>
struct TreeT {
  R* lp;
  uint8_t* leafs;
  uint8_t* iref;
  uint8_t* icoarse;
  uint8_t* mark;
};

// updatelpup.cpp:4-118: the four sons of nb (group s0 .. s0 + 3) linked into
// the neighbour table; a refined neighbour's adjacent son points back
template <typename R>
__device__ __forceinline__ void link_sons(const TreeT<R>& T, int nb) {
  R* lp = T.lp;
  const int s0 = lp[8 * nb + 2];
  const int s[4] = {s0, s0 + 1, s0 + 2, s0 + 3};
  const int lev = lp[8 * nb] + 1;
  auto outer = [&](int son, int side, int off, int back) {
    const int n = lp[8 * nb + side];
    if (n == -1 || T.leafs[n]) {
      lp[8 * son + side] = (R)n;
    } else {
      const int nsr = lp[8 * n + 2] + off;
      lp[8 * son + side] = (R)nsr;
      lp[8 * nsr + back] = (R)son;
    }
  };
  for (int q = 0; q < 4; q++) {
    lp[8 * s[q] + 0] = (R)lev;
    lp[8 * s[q] + 1] = (R)nb;
    lp[8 * s[q] + 3] = (R)q;
  }
  lp[8 * s[0] + 5] = (R)s[1];
  lp[8 * s[0] + 7] = (R)s[2];
  outer(s[0], 4, 1, 5);
  outer(s[0], 6, 2, 7);
  lp[8 * s[1] + 4] = (R)s[0];
  lp[8 * s[1] + 7] = (R)s[3];
  outer(s[1], 5, 0, 4);
  outer(s[1], 6, 3, 7);
  lp[8 * s[2] + 5] = (R)s[3];
  lp[8 * s[2] + 6] = (R)s[0];
  outer(s[2], 4, 3, 5);
  outer(s[2], 7, 0, 6);
  lp[8 * s[3] + 4] = (R)s[2];
  lp[8 * s[3] + 6] = (R)s[1];
  outer(s[3], 5, 2, 4);
  outer(s[3], 7, 1, 6);
  for (int q = 0; q < 4; q++) {
    T.leafs[s[q]] = 1;
    T.icoarse[s[q]] = 0;
  }
  T.iref[nb] = 0;
  T.icoarse[nb] = 0;
}

// updatelpdown.cpp:5-54, the pointer part, for the group whose first son is
// nb: the sons of each neighbour of the dad that is refined when the serial
// loop reaches this group point back to the dad.  Leaf flags are updated
// afterwards, by the caller, for all groups of the level at once.
template <typename R>
__device__ __forceinline__ void unlink_pointers(const TreeT<R>& T, int nb) {
  R* lp = T.lp;
  const int dad = lp[8 * nb + 1];
  auto refined = [&](int n) {
    if (n == -1 || T.leafs[n]) return false;
    const int g = lp[8 * n + 2];
    return !(T.mark[g] && g < nb);
  };
  int n = lp[8 * dad + 4];
  if (refined(n)) {
    const int a = lp[8 * lp[8 * n + 2] + 5];
    lp[8 * a + 5] = (R)dad;
    lp[8 * lp[8 * a + 7] + 5] = (R)dad;
  }
  n = lp[8 * dad + 5];
  if (refined(n)) {
    const int a = lp[8 * n + 2];
    lp[8 * a + 4] = (R)dad;
    lp[8 * lp[8 * a + 7] + 4] = (R)dad;
  }
  n = lp[8 * dad + 6];
  if (refined(n)) {
    const int a = lp[8 * lp[8 * n + 2] + 7];
    lp[8 * a + 7] = (R)dad;
    lp[8 * lp[8 * a + 5] + 7] = (R)dad;
  }
  n = lp[8 * dad + 7];
  if (refined(n)) {
    const int a = lp[8 * n + 2];
    lp[8 * a + 6] = (R)dad;
    lp[8 * lp[8 * a + 5] + 6] = (R)dad;
  }
}

// exclusive prefix sum of `v` over the workgroup's lanes in lane order;
// *total receives the sum (uniform)
__device__ int wg_scan_sum(int v, int* sw, int* total) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) sw[wv] = x;
  __syncthreads();
  int base = 0, tot = 0;
  for (int w = 0; w < kRgWaves; w++) {
    const int y = sw[w];
    base += w < wv ? y : 0;
    tot += y;
  }
  __syncthreads();
  *total = tot;
  return base + x - v;
}

// workgroup scratch of the admesh passes
struct Scratch {
  int* sw;  // [kRgWaves]
  int* fail_seq;
  int* fail_nb;
  int* bad;
  volatile unsigned* levels;  // bit per level: levels with blocks to refine, later levels that coarsen
};

struct AdmeshResult {
  int last_active, nref, ncoars, status, fail_nb;
};

// min over a lane's partial CFL minima (order-free: partials are never NaN)
__device__ double wg_min(double lam, double* sl) {
  for (int o = 32; o > 0; o >>= 1) lam = std_min(lam, __shfl_xor(lam, o));
  if ((threadIdx.x & 63) == 0) sl[threadIdx.x >> 6] = lam;
  __syncthreads();
  double m = sl[0];
  for (int w = 1; w < kRgWaves; w++) m = std_min(m, sl[w]);
  return m;
}

// admesh.cpp:17-102 for one instance (all lanes of the workgroup).  Passes
// that must follow block order give each lane a contiguous range of blocks,
// so that one prefix sum over the lanes orders a whole level.
template <typename R>
__device__ __forceinline__ AdmeshResult admesh(const TreeT<R>& T, const RegridInst& RI,
                                               const LeafItem* __restrict__ leaves,
                                               const double* __restrict__ lam_grad, int32_t* log,
                                               const Scratch& S, double* sl, double* lambda) {
  const int tid = threadIdx.x;
  const R* lp = T.lp;
  // every op consumes distinct leaves, so the log never outgrows the
  // instance's leaf range; a write past it (an inconsistent tree) is refused
  // and reported instead
  const int nlog = RI.leaf_end - RI.leaf_begin;
  const int ML = RI.maxlevs;
  int LA = RI.last_active;

  // markref.cpp:6-18 with criteria.cpp's thresholds on the leaf's indicator
  for (int nb = tid; nb < LA; nb += kRgThreads) {
    T.iref[nb] = 0;
    T.icoarse[nb] = 0;
  }
  if (tid == 0) *S.levels = 0u;
  __syncthreads();
  // the leaf pass also takes the CFL minimum (main.cpp:208-230)
  unsigned lv = 0;
  double lam = 1.E6;
  for (int l = RI.leaf_begin + tid; l < RI.leaf_end; l += kRgThreads) {
    const int nb = leaves[l].nb;
    lam = std_min(lam, lam_grad[2 * l]);
    if (lam_grad[2 * l + 1] >= 0.05) {
      const int lev = lp[8 * nb];
      if (lev < ML - 1) {
        T.iref[nb] = 1;
        lv |= 1u << lev;
      }
    } else {
      T.icoarse[nb] = 1;
    }
  }
  if (lv) atomicOr((unsigned*)S.levels, lv);
  *lambda = wg_min(lam, sl);  // its barrier publishes the marks

  int nref = 0, ncoars = 0, status = 0, fail_nb = -1;
  if (*S.levels) {
    // critneighup.cpp:6-35, levels that hold marked blocks.  A level's bit is
    // final once the passes of every finer level are done (each ends in a
    // barrier), so every lane takes the same branch.
    for (int nl = ML - 1; nl >= 1; nl--) {
      if (!((*S.levels >> nl) & 1u)) continue;
      unsigned add = 0;
      for (int nb = tid; nb < LA; nb += kRgThreads)
        if (T.iref[nb] && lp[8 * nb] == nl)
          for (int s = 4; s < 8; s++) {
            const int n1 = lp[8 * nb + s];
            if (n1 != -1) {
              const int l1 = lp[8 * n1];
              if (l1 < nl) {
                T.iref[n1] = 1;
                add |= 1u << l1;
              }
            }
          }
      if (add) atomicOr((unsigned*)S.levels, add);
      __syncthreads();
    }
    const unsigned refl = *S.levels;
    // admesh.cpp:30-58, levels that hold marked blocks, coarsest first
    for (int nl = 0; nl < ML - 1; nl++) {
      if (!((refl >> nl) & 1u)) continue;
      const int la0 = LA;
      // the first new group that takes lastActive - 1 to maxblocks or beyond:
      // the smallest k with la0 + 4 k + 3 >= maxblocks
      const int d = RI.maxblocks - la0 - 3;
      const int kfail = d <= 0 ? 0 : (d + 3) / 4;
      const int per = (la0 + kRgThreads - 1) / kRgThreads;
      const int b0 = tid * per, b1 = b0 + per < la0 ? b0 + per : la0;
      // marked blocks of this level (low 16 bits) and those among them that
      // get a new son group (high 16 bits) in this lane's range
      int cnt = 0;
      for (int nb = b0; nb < b1; nb++)
        if (T.iref[nb] && lp[8 * nb] == nl) cnt += 1 + (lp[8 * nb + 2] == 0 ? 0x10000 : 0);
      int tot;
      const int ex = wg_scan_sum(cnt, S.sw, &tot);
      const int nsel = tot & 0xffff, nnew = tot >> 16;
      // the op that overflows maxblocks, if any: the new group of rank kfail
      int fs = INT_MAX;
      if (nnew > kfail) {
        int is = ex & 0xffff, in = ex >> 16;
        if (kfail >= in && kfail < in + (cnt >> 16))
          for (int nb = b0; nb < b1; nb++)
            if (T.iref[nb] && lp[8 * nb] == nl) {
              if (lp[8 * nb + 2] == 0 && in++ == kfail) {
                *S.fail_seq = is;
                *S.fail_nb = nb;
              }
              is++;
            }
        __syncthreads();
        fs = *S.fail_seq;
      }
      {
        int is = ex & 0xffff, in = ex >> 16;
        for (int nb = b0; nb < b1; nb++) {
          if (!(T.iref[nb] && lp[8 * nb] == nl)) continue;
          const int seq = is++;
          const bool fresh = lp[8 * nb + 2] == 0;
          const int rk = fresh ? in++ : -1;
          if (seq > fs) continue;
          if (nref + seq >= nlog) {
            *S.bad = 1;
            continue;
          }
          if (fresh) T.lp[8 * nb + 2] = (R)(la0 + 4 * rk);
          log[nref + seq] = nb;
          // the op that overflows stops before updatelpup
          if (seq < fs) T.leafs[nb] = 0;
        }
      }
      __syncthreads();
      if (*S.bad) {
        status = kErrConfig;
        break;
      }
      const int napplied = fs == INT_MAX ? nsel : fs;
      for (int k = tid; k < napplied; k += kRgThreads) link_sons(T, log[nref + k]);
      __syncthreads();
      if (fs != INT_MAX) {
        status = kErrOverflow;
        fail_nb = *S.fail_nb;
        nref += fs + 1;
        LA = la0 + 4 * (kfail + 1);
        break;
      }
      nref += nsel;
      LA = la0 + 4 * nnew;
    }
  }

  if (!status) {
    // critneighdown.cpp:6-33, then the groups of four siblings that coarsen
    // (admesh.cpp:69-81) at every level at once: the choice reads the sons'
    // flags and the groups' inner pointers, which coarsening does not change.
    // mark holds their first sons; `cl`: levels that coarsen.
    for (int nb = tid; nb < LA; nb += kRgThreads) {
      if (!T.icoarse[nb]) continue;
      for (int s = 4; s < 8; s++) {
        const int n1 = lp[8 * nb + s];
        if (n1 != -1 && !T.leafs[n1]) {
          T.icoarse[nb] = 0;
          break;
        }
      }
    }
    if (tid == 0) *S.levels = 0u;
    __syncthreads();
    unsigned cl = 0;
    for (int nb = tid; nb < LA; nb += kRgThreads) {
      const int lev = lp[8 * nb];
      if (lev >= 1 && T.icoarse[nb] && lp[8 * nb + 3] == 0) {
        const int s1 = lp[8 * nb + 5], s2 = lp[8 * nb + 7], s3 = s1 >= 0 ? lp[8 * s1 + 7] : -1;
        if (s2 >= 0 && s3 >= 0 && T.icoarse[s1] && T.icoarse[s2] && T.icoarse[s3]) {
          T.mark[nb] = 1;
          cl |= 1u << lev;
        }
      }
    }
    if (cl) atomicOr((unsigned*)S.levels, cl);
    __syncthreads();
    const unsigned coarl = *S.levels;
    const int per = (LA + kRgThreads - 1) / kRgThreads;
    const int b0 = tid * per, b1 = b0 + per < LA ? b0 + per : LA;
    // admesh.cpp:66-95, finest level first
    for (int nl = ML - 1; nl >= 1; nl--) {
      if (!((coarl >> nl) & 1u)) continue;
      int cnt = 0;
      for (int nb = b0; nb < b1; nb++) cnt += T.mark[nb] && lp[8 * nb] == nl;
      int nsel;
      int pos = nref + ncoars + wg_scan_sum(cnt, S.sw, &nsel);
      for (int nb = b0; nb < b1; nb++) {
        if (!(T.mark[nb] && lp[8 * nb] == nl)) continue;
        if (pos < nlog) log[pos] = nb;
        else *S.bad = 1;
        pos++;
      }
      __syncthreads();
      if (*S.bad) {
        status = kErrConfig;
        break;
      }
      for (int k = tid; k < nsel; k += kRgThreads) unlink_pointers(T, log[nref + ncoars + k]);
      __syncthreads();
      for (int k = tid; k < nsel; k += kRgThreads) {
        const int nb = log[nref + ncoars + k];
        T.leafs[lp[8 * nb + 1]] = 1;
        for (int j = 0; j < 4; j++) T.leafs[nb + j] = 0;
        T.mark[nb] = 0;
      }
      __syncthreads();
      ncoars += nsel;
    }
    if (status) {
      // marks of the levels not reached
      for (int nb = tid; nb < LA; nb += kRgThreads) T.mark[nb] = 0;
    }
  }
  return AdmeshResult{LA, nref, ncoars, status, fail_nb};
}

__global__ __launch_bounds__(kRgThreads) void regrid_kernel(const RegridInst* __restrict__ ri,
                                                            const LeafItem* __restrict__ leaves,
                                                            const double* __restrict__ lam_grad, DeviceTree D,
                                                            RegridOut* out, int32_t* oplog) {
  __shared__ int sw[kRgWaves];
  __shared__ double sl[kRgWaves];
  __shared__ int s_fail_seq, s_fail_nb, s_bad;
  __shared__ unsigned s_levels;
  // the instance's tree staged in LDS (rows as 16-bit ids: block ids stay
  // below 2^15 for maxlevs <= 7) when its rows fit
  __shared__ int16_t s_lp[8 * kRgLdsRows];
  __shared__ __align__(16) uint8_t s_flags[4 * kRgLdsRows];
  const int q = blockIdx.x, tid = threadIdx.x;
  if (tid == 0) s_bad = 0;
  const RegridInst RI = ri[q];

  if (!RI.regrid) {
    // CFL minimum over the instance's leaves only (main.cpp:208-230)
    double lam = 1.E6;
    for (int l = RI.leaf_begin + tid; l < RI.leaf_end; l += kRgThreads) lam = std_min(lam, lam_grad[2 * l]);
    lam = wg_min(lam, sl);
    if (tid == 0) out[q] = RegridOut{lam, 0, 0, RI.last_active, 0, -1, 0};
    return;
  }
  const size_t off = (size_t)RI.off;
  const TreeT<int32_t> G{D.lp + 8 * off, D.leafs + off, D.iref + off, D.icoarse + off, D.mark + off};
  const Scratch S{sw, &s_fail_seq, &s_fail_nb, &s_bad, &s_levels};
  int32_t* log = oplog + RI.leaf_begin;
  AdmeshResult res;
  double lam;
  const int rows = RI.lds_rows;  // uniform: rows staged in LDS, 0 = work in global memory
  if (rows > 0) {
    // staged with 16-byte loads (a row is two), a batch of kStageUnroll per
    // lane issued before any of them is stored
    const int4* src = reinterpret_cast<const int4*>(G.lp);
    const int nv = 2 * rows;
    // leaf flags four at a time into flag planes of ps bytes each (instance
    // offsets and maxblocks are multiples of 4, so the last word stays inside
    // the instance's blocks, and ps <= kRgLdsRows), loaded with the first batch
    const int ps = (rows + 3) & ~3;
    const uint32_t* lsrc = reinterpret_cast<const uint32_t*>(G.leafs);
    uint32_t* ldst = reinterpret_cast<uint32_t*>(s_flags);
    uint32_t lw[2];
#pragma unroll
    for (int u = 0; u < 2; u++)
      if (tid + u * kRgThreads < ps / 4) lw[u] = lsrc[tid + u * kRgThreads];
    for (int i = tid + 2 * kRgThreads; i < ps / 4; i += kRgThreads) ldst[i] = lsrc[i];
    for (int i0 = tid; i0 < nv; i0 += kStageUnroll * kRgThreads) {
      int4 v[kStageUnroll];
#pragma unroll
      for (int u = 0; u < kStageUnroll; u++) {
        const int i = i0 + u * kRgThreads;
        if (i < nv) v[u] = src[i];
      }
#pragma unroll
      for (int u = 0; u < kStageUnroll; u++) {
        const int i = i0 + u * kRgThreads;
        if (i < nv) {
          s_lp[4 * i + 0] = (int16_t)v[u].x;
          s_lp[4 * i + 1] = (int16_t)v[u].y;
          s_lp[4 * i + 2] = (int16_t)v[u].z;
          s_lp[4 * i + 3] = (int16_t)v[u].w;
        }
      }
    }
#pragma unroll
    for (int u = 0; u < 2; u++)
      if (tid + u * kRgThreads < ps / 4) ldst[tid + u * kRgThreads] = lw[u];
    for (int i = tid; i < 3 * ps / 4; i += kRgThreads) ldst[ps / 4 + i] = 0;
    __syncthreads();
    const TreeT<int16_t> L{s_lp, s_flags, s_flags + ps, s_flags + 2 * ps, s_flags + 3 * ps};
    res = admesh(L, RI, leaves, lam_grad, log, S, sl, &lam);
    __syncthreads();
    // back to device memory when the tree changed (rows past lastActive that
    // the regrid did not touch stay as they were)
    const int wb = (res.nref + res.ncoars == 0) ? 0 : res.last_active < rows ? res.last_active : rows;
    int4* dst = reinterpret_cast<int4*>(G.lp);
    for (int i = tid; i < 2 * wb; i += kRgThreads)
      dst[i] = make_int4(s_lp[4 * i], s_lp[4 * i + 1], s_lp[4 * i + 2], s_lp[4 * i + 3]);
    for (int nb = tid; nb < wb; nb += kRgThreads) G.leafs[nb] = s_flags[nb];
  } else {
    res = admesh(G, RI, leaves, lam_grad, log, S, sl, &lam);
  }
  if (tid == 0) out[q] = RegridOut{lam, res.nref, res.ncoars, res.last_active, res.status, res.fail_nb, 0};
}

}  // namespace

void launch_regrid(const RegridInst* ri, int ninst, const LeafItem* leaves, const double* lam_grad,
                   const DeviceTree& tree, RegridOut* out, int32_t* oplog, hipStream_t s) {
  if (ninst <= 0) return;
  regrid_kernel<<<ninst, kRgThreads, 0, s>>>(ri, leaves, lam_grad, tree, out, oplog);
}

}  // namespace amr
