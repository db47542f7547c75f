// Host engine of the AMR2D-MHD hot path: quadtree bookkeeping mirrored into
// device work lists, the ghost-fill program compiler, the step driver and the
// C ABI declared in include/amr2d_mhd.h.
//
// The solution state of every slot of every instance stays in HBM for the
// whole run; per step only the per-leaf CFL bounds and refinement indicators
// come back to the host, and work lists / fill programs go down when a tree
// changes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <chrono>
#include <cstdlib>
#include <new>
#include <cstring>
#include <functional>
#include <thread>
#include <atomic>
#include <exception>
#include <condition_variable>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/amr2d_mhd.h"
#include "abi_guard.hpp"
#include "amr_types.hpp"
#include "flop_count.hpp"
#include "kernels.hpp"
#include "physics.hpp"

using namespace amr;

namespace {

constexpr int kMaxLev = 16;
constexpr int kNbRoots = 4;  // defs.hpp:29

// block capacity of a simulation with `maxlevs` levels (defs.hpp:30)
int maxblocks_of(int maxlevs) {
  const double p = std::pow(kNbRoots, maxlevs - 1);
  return kNbRoots > p ? kNbRoots : (int)(std::pow(4, maxlevs) * 2);
}
constexpr double kPiRef = 3.14159265359;  // defs.hpp:23

int lim_code(const char* s) {
  static const char* names[] = {"1st", "noSL", "minmod", "VanLeer", "Superbee", "MC", "Koren", "Cada3rd", "WENO3", "WENO5"};
  if (!std::strcmp(s, "WENO2D")) return kLimNoSL;
  for (int q = 0; q < kLimCount; q++)
    if (!std::strcmp(s, names[q])) return q;
  return -1;
}
int flux_code(const char* s) {
  static const char* names[] = {"RUSA", "HLLE", "HLLC", "HLLD"};
  for (int q = 0; q < kFluxCount; q++)
    if (!std::strcmp(s, names[q])) return q;
  return -1;
}
enum { IC_SST, IC_SST1, IC_SSTY, IC_BWX, IC_BWX1, IC_BWY, IC_BWY1, IC_ROTOR, IC_ROTOR_P, IC_OT, IC_ROTOR_BS, IC_COUNT };
int ic_code(const char* s) {
  static const char* names[] = {"SST",   "SST1",    "SSTy", "BWx",     "BWx1", "BWy", "BWy1",
                                "rotor", "rotor_p", "OT",   "rotor_bs"};
  for (int q = 0; q < IC_COUNT; q++)
    if (!std::strcmp(s, names[q])) return q;
  return -1;
}

// restriction node: `simple` when its four operand cells still hold their
// pre-fill values (the common case), else explicit operand references.
// During compilation references use the instance's own block ids (nb).
struct RNode {
  int32_t nb;
  uint16_t cell;
  bool simple;
  int64_t ref[4];
};

// page-locked growable array, so uploads are true async DMA (falls back to
// ordinary memory where no device is present, e.g. for the host test hooks)
template <class T>
struct PinnedVec {
  T* p = nullptr;
  size_t n = 0, cap = 0;
  bool pinned = false;
  PinnedVec() = default;
  PinnedVec(const PinnedVec&) = delete;
  PinnedVec& operator=(const PinnedVec&) = delete;
  PinnedVec(PinnedVec&& o) noexcept : p(o.p), n(o.n), cap(o.cap), pinned(o.pinned) { o.p = nullptr; o.n = o.cap = 0; }
  PinnedVec& operator=(PinnedVec&& o) noexcept {
    if (this != &o) {
      release(p, pinned);
      p = o.p; n = o.n; cap = o.cap; pinned = o.pinned;
      o.p = nullptr; o.n = o.cap = 0;
    }
    return *this;
  }
  ~PinnedVec() { release(p, pinned); }
  static void release(T* q, bool pin) {
    if (!q) return;
    if (pin) hipHostFree(q);
    else std::free(q);
  }
  void clear() { n = 0; }
  void reserve(size_t m) {
    if (m <= cap) return;
    size_t nc = std::max(m, cap * 2 + 1024);
    T* q = nullptr;
    bool pin = hipHostMalloc((void**)&q, nc * sizeof(T)) == hipSuccess;
    if (!pin) {
      q = (T*)std::malloc(nc * sizeof(T));
      if (!q) throw std::bad_alloc();
    }
    if (n) std::memcpy(q, p, n * sizeof(T));
    release(p, pinned);
    p = q;
    pinned = pin;
    cap = nc;
  }
  void push_back(const T& v) {
    if (n == cap) reserve(n + 1);
    p[n++] = v;
  }
  // size m, new elements left unset
  void resize_uninit(size_t m) {
    reserve(m);
    n = m;
  }
  void swap(PinnedVec& o) noexcept {
    std::swap(p, o.p);
    std::swap(n, o.n);
    std::swap(cap, o.cap);
    std::swap(pinned, o.pinned);
  }
  size_t size() const { return n; }
  T* data() { return p; }
};

// an op with an operand in a ghost ring (a possible stale read)
struct RingSrc {
  int32_t op;
  int32_t nref;
  int64_t ref[4];  // block-local operands
};

// Persistent helper threads for the ranges of one compile: run(R, f) calls
// f(r) for every r in [0, R) (R < 256) on the helpers and the caller and
// returns when all have finished.  The work counter carries its generation,
// so a helper that wakes late never takes an index of a later call.  Helpers
// and caller spin briefly before they block: the phases of one compile
// follow each other within microseconds.
class RangePool {
 public:
  explicit RangePool(int helpers) {
    for (int i = 0; i < helpers; i++) th_.emplace_back([this] { loop(); });
  }
  RangePool(const RangePool&) = delete;
  RangePool& operator=(const RangePool&) = delete;
  ~RangePool() {
    {
      std::lock_guard<std::mutex> lk(m_);
      stop_.store(true, std::memory_order_release);
    }
    cv_.notify_all();
    for (auto& t : th_) t.join();
  }
  int helpers() const { return (int)th_.size(); }
  template <class F>
  void run(int R, const F& f) {
    const std::function<void(int)> job = [&f](int r) { f(r); };
    uint64_t g;
    {
      std::lock_guard<std::mutex> lk(m_);
      g = gen_.load(std::memory_order_relaxed) + 1;
      job_ = &job;
      pending_.store(R, std::memory_order_relaxed);
      state_.store((g << 16) | ((uint64_t)R << 8), std::memory_order_release);
      gen_.store(g, std::memory_order_release);
    }
    cv_.notify_all();
    work(g);
    if (!spin([this] { return pending_.load(std::memory_order_acquire) == 0; })) {
      std::unique_lock<std::mutex> lk(m_);
      done_.wait(lk, [this] { return pending_.load(std::memory_order_acquire) == 0; });
    }
    // an exception of any range (host allocation) surfaces in the caller
    std::exception_ptr e;
    {
      std::lock_guard<std::mutex> lk(m_);
      e = error_;
      error_ = nullptr;
    }
    if (e) std::rethrow_exception(e);
  }

 private:
  void work(uint64_t g) {
    uint64_t v = state_.load(std::memory_order_acquire);
    for (;;) {
      if ((v >> 16) != g) return;
      const int R = (int)((v >> 8) & 0xff), r = (int)(v & 0xff);
      if (r >= R) return;
      if (!state_.compare_exchange_weak(v, v + 1, std::memory_order_acq_rel, std::memory_order_acquire)) continue;
      try {
        (*job_)(r);
      } catch (...) {
        std::lock_guard<std::mutex> lk(m_);
        if (!error_) error_ = std::current_exception();
      }
      if (pending_.fetch_sub(1, std::memory_order_acq_rel) == 1) {
        std::lock_guard<std::mutex> lk(m_);
        done_.notify_all();
      }
      v = state_.load(std::memory_order_acquire);
    }
  }
  // true once ready() holds, false after ~3 ms without.  A regridding run
  // compiles every few steps (every 2-3 ms on the rotor), so helpers that
  // keep spinning across that gap start the next compile without a blocking
  // wake-up; they sleep once compiles stop.
  template <class P>
  static bool spin(const P& ready) {
    const auto t0 = std::chrono::steady_clock::now();
    for (int i = 1;; i++) {
      if (ready()) return true;
      __builtin_ia32_pause();
      if ((i & 255) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(3000)) return false;
    }
  }
  void loop() {
    uint64_t seen = 0;
    for (;;) {
      auto fresh = [&] { return stop_.load(std::memory_order_acquire) || gen_.load(std::memory_order_acquire) != seen; };
      if (!spin(fresh)) {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, fresh);
      }
      if (stop_.load(std::memory_order_acquire)) return;
      seen = gen_.load(std::memory_order_acquire);
      work(seen);
    }
  }
  std::vector<std::thread> th_;
  std::mutex m_;
  std::condition_variable cv_, done_;
  std::atomic<uint64_t> state_{0};  // generation << 16 | R << 8 | next index
  std::atomic<int> pending_{0};
  std::atomic<uint64_t> gen_{0};
  std::atomic<bool> stop_{false};
  const std::function<void(int)>* job_ = nullptr;
  std::exception_ptr error_;  // first exception of the current call
};

// Symbolic fill state of one contiguous range of the schedule.  A compile
// splits the schedule into ranges executed independently (in parallel on
// large trees); each range sees the pre-range state as "pre-fill" values,
// and the merge substitutes what earlier ranges wrote.
struct FillWorker {
  // symbolic value per ghost-ring element (the only ones fills write),
  // [block][ring index]: -1 = not written in this range, >= 0 = the value
  // element (nb << 16) | cell held when the range began, <= -2 = restriction
  // node -(v + 2) of this range.  Reset only where touched.
  std::vector<int32_t> cur;
  std::vector<int32_t> touched;  // elements written in this range, in first-write order
  std::vector<RNode> nodes;
  bool interior_write = false;
  // elements first written (over all ranges) in this range, with their final
  // values in pre-fill terms; nkeep of them change
  std::vector<int32_t> first, fval;
  int nkeep = 0;
  // emission of this range's ops
  std::vector<FillGen> gen;
  std::vector<int32_t> gen_ops;  // ops that index gen
  std::vector<RingSrc> ringsrc;  // ops with an operand in a ghost ring
  bool nested = false;           // a restriction of a restriction (rejected)
  int64_t nkind[3] = {0, 0, 0};  // ops emitted per kind (copy, restrict, general)
  std::vector<std::array<int32_t, 2>> edges;  // (reader op, writer op) of stale reads
  bool self_read = false;        // an op reads (pre-fill) the element it writes
  std::vector<int32_t> linked, gid;           // linked ops of this range's op segment, their groups
};

// host scratch reused across compiles (tables reset only where touched)
struct FillScratch {
  std::vector<FillWorker> w;
  std::vector<uint16_t> wmask;   // per ring element: bit r set when range r writes it
  std::vector<RNode> rnodes;     // restriction nodes of all ranges, in pre-fill terms
  std::vector<int32_t> nodeoff;  // first rnodes index of each range
  std::vector<int32_t> seq;      // schedule entries in execution order
  std::vector<int32_t> uf;       // union-find over ops linked by stale reads
  std::vector<FillOp> moved;
  std::vector<uint8_t> ring;      // per cell of a tile: 1 in the ghost ring
  std::vector<int32_t> ring_idx;  // per cell: index among the ring cells, -1 inside
  int nring = 0;
  int ring_nt = 0;
  bool force_general = false;  // test hook: emit every restriction as kFillGeneral
  int force_ranges = 0;        // test hook: number of schedule ranges (0: by tree size)
  std::vector<int32_t> group_count, group_pos;
  std::unique_ptr<RangePool> pool;  // helpers of a compile split over threads
};

// A compiled ghost-fill program kept for reuse, with its device copy: AMR
// runs refine and coarsen the same blocks back and forth, so trees recur
// (the 5-level rotor changes its tree 1912 times but has 525 distinct trees;
// 1315 changes return to the tree of two changes earlier).  The program is a
// function of the key alone: lastActive, then per block its neighbour table
// row, leaf flag and arena slot.
struct FillProg {
  std::vector<int32_t> key;
  uint64_t hash = 0;
  FillOp* d_ops = nullptr;  // device copy only: the host arrays are compile scratch
  int cap_ops = 0, nops = 0;
  FillGen* d_gen = nullptr;
  int cap_gen = 0, ngen = 0;
  std::array<int64_t, 4> kinds{};  // ops per kind (copy, restrict, general, padding)
  int tail_chunk0 = 0;             // Instance::tail_chunk0 of this program
  uint64_t used = 0;  // LRU stamp
};

// ------------------------------------------------------------- instance
struct Instance {
  amr_sim_config cfg;
  int lim = 0, solver = 0, ic = 0, maxlevs = 1, maxblocks = 4;
  double gamma = 2.0, tEnd = 0.2, cfl = 0.4;
  double dx[kMaxLev] = {0};
  // quadtree (defs.hpp:52-58): lp = level, father, first son, sibling id,
  // left, right, bottom, top neighbour
  std::vector<std::array<int32_t, 8>> lp;
  std::vector<std::array<int32_t, 2>> icoord;
  std::vector<uint8_t> leafs, iref, icoarse;
  int lastActive = 0;
  std::vector<int32_t> gs;  // nb -> global arena slot
  std::vector<std::array<int32_t, 3>> sched;
  double t = 0, dt_main = 0, dt_dom = 0;
  int count = 0, status = 0;
  bool periodic = false;  // build-defined: wrapped root neighbours, domain [0, lenx]^2
  int ct_type = 1;        // corner EMF: CTtype 1 or 2 (CT2D.cpp:15-75)
  amr_diag err{};
  bool done = false;
  // the device copy of this tree (device-side regrid) is stale: refreshed
  // from this one at the next topology rebuild
  bool dev_tree_dirty = true;
  // ghost-fill program of this instance (global slot ids); recompiled only
  // when its tree changed, kept in its own device segment (nops ops, ngen
  // explicit-operand entries); ops/gen are the host arrays of the last compile
  PinnedVec<FillOp> ops;
  std::vector<FillGen> gen;
  bool prog_dirty = true;
  FillOp* d_ops = nullptr;
  int cap_ops = 0, nops = 0;
  FillGen* d_gen = nullptr;
  int cap_gen = 0, ngen = 0;
  std::array<int64_t, 4> op_kinds{};  // ops of the current program per kind (copy, restrict, general, padding)
  // first chunk of the current program's linked (stale-read) groups: applied
  // again to its own output, every op before it writes the value already
  // there, so a repeated fill needs only the chunks from here on (0: every
  // chunk, when an op reads its own destination)
  int tail_chunk0 = 0;
  // the tree the current program was compiled for, and earlier programs
  std::vector<int32_t> prog_key;
  uint64_t prog_hash = 0;
  std::vector<FillProg> prog_cache;
};

// --------------------------------------------------------------- context
struct Profile {
  double ms = 0;
  int64_t launches = 0;
};
enum { PF_STAGE = 0, PF_FILL, PF_CT, PF_CFL, PF_REGRID, PF_HOST_TOPO, PF_HOST_REGRID, PF_HOST_WAIT, PF_HOST_COMPILE,
       PF_GRAPH, PF_STEP, PF_CAPTURE, PF_COUNT };


}  // namespace

// stage-kernel tuner: trials per variant and leaf-count bucket
constexpr int kTuneTrials = 3;

struct amr_ctx {
  int NT = 20, ng = 2, nxi = 16;
  int rk = 2;  // RK stages of every instance: 2 (timeIntegral.cpp) or 3 (SSP-RK3, build-defined)
  std::vector<Instance> inst;
  FieldPtrs F{};
  int64_t cap_slots = 0, used_slots = 0;
  hipStream_t stream = nullptr;      // the stream every operation of the context runs on
  hipStream_t own_stream = nullptr;  // the context's own stream (created with it)
  // device work lists and buffers (capacities in elements)
  // topology block, rebuilt at each tree change: the fill-program segment
  // tables and the leaf work list in one device allocation, filled by one
  // upload from a pinned staging copy with the same layout
  char* d_topo = nullptr;
  char* h_topo = nullptr;
  size_t cap_topo = 0;
  LeafItem* d_leaves = nullptr;  // view into d_topo
  int nleaf = 0;
  std::vector<LeafItem> h_leaves;
  // stage launches split by specialisation (running instances of different
  // solver/limiter classes): the leaves ordered by class, split_off[k] ..
  // split_off[k + 1] those of class split_spec[k]
  bool stage_split = false;
  std::vector<LeafItem> h_split_leaves;
  int split_n = 0, split_spec[kStageSpecs] = {0}, split_off[kStageSpecs + 1] = {0};
  LeafItem* d_split_leaves = nullptr;  // view into d_topo
  std::vector<int> leaf_inst_begin;  // per active instance, first leaf index
  // host copies kept alive until the stream has consumed them
  std::vector<RegridOp> h_refs, h_coars;
  std::vector<FillScratch> scratch;  // one per compile thread
  // fill program tables: per-instance segment pointers and chunk-count prefix
  const FillOp** d_inst_ops = nullptr;  // views into d_topo
  const FillGen** d_inst_gen = nullptr;
  int32_t* d_prefix = nullptr;
  int nchunks = 0;
  // the same tables over each program's linked tail (Instance::tail_chunk0 on)
  const FillOp** d_tail_ops = nullptr;
  int32_t* d_tail_prefix = nullptr;
  int ntail_chunks = 0;
  // Set by a step that leaves every tree as it was (no regrid op, no instance
  // retired) and that ended, like every step, with the fills of Bi (before
  // the last cell sync) and of U (after it); cleared by every other entry
  // point that takes the context.  While set, U and Bi hold those fills'
  // results and nothing has written either since, so the next step's first
  // fill of U and Bi -- the same programs over the same values -- changes
  // only what its linked tails write, and runs only those.
  bool fill_fresh = false;
  bool tail_first_fill = false;  // the step being enqueued runs its first fill on the tails
  bool step_direct = false;      // the step in flight was enqueued by enqueue_step, not captured
  InstParams* d_ip = nullptr;
  PinnedVec<InstParams> h_ip;  // pinned: the step graph copies from it
  // the device work of one step replayed as a HIP graph; re-captured when
  // the topology (work lists, program tables, buffer addresses) changes
  bool graph_on = false, graph_dirty = true, regrid_launched = false;
  // launch geometry of the step's leaf and chunk kernels and of its download:
  // the work itself (direct launches, launch_counts = null) or capacities
  // covering it (a captured graph: its kernels read the counts from d_counts
  // and surplus workgroups exit, so one graph serves every tree whose counts
  // fit, and tree changes cost no capture)
  int grid_leaves = 0, grid_chunks = 0;
  // a host-to-device copy (IC, amr_upload, a cleared record) since the last
  // direct step: the next step launches directly, so a replayed graph only
  // ever reads data that kernels wrote
  bool dma_since_direct = true;

  const int32_t* d_counts = nullptr;       // {nleaf, nchunks}, view into d_topo
  const int32_t* launch_counts = nullptr;
  uint64_t topo_epoch = 0;                 // bumps when d_topo moves
  struct GraphKey {
    uint64_t epoch = ~0ull;
    int leaves = 0, chunks = 0, variant = -1, spec = -1;
    bool operator==(const GraphKey& o) const {
      return epoch == o.epoch && leaves == o.leaves && chunks == o.chunks && variant == o.variant && spec == o.spec;
    }
  } gkey;
  // steps launched since the last topology rebuild: a graph is captured only
  // for a topology that has already lasted a step (most trees of a regridding
  // run last one or two steps, and a capture used once costs more than it saves)
  int steps_same_topo = 0;
  // compiled fill programs kept per instance besides the current one (LRU)
  int prog_cache_max = 4;
  // a step launched by step_begin and not yet completed by step_end
  bool in_flight = false, step_trial = false;
  uint64_t prog_clock = 0;
  // stage-kernel variant: fixed (0..kStageVariants-1) or chosen per leaf-count
  // bucket by the tuner (-1): the first launches in a bucket try every variant
  // kTuneTrials times, timed by the kernel's own device-clock span, then the
  // one with the fastest trial per leaf is kept for that bucket
  int stage_variant = -1;
  struct TuneBucket {
    double us_per_leaf[kStageVariants] = {0};          // sum over the trials (reported as a mean)
    double trial_us[kStageVariants][kTuneTrials] = {};  // each timed trial, us per leaf (decides: median)
    int nval[kStageVariants] = {0};                     // timed trials
    int n[kStageVariants] = {0};                        // trials launched
    int chosen = -1;
  };
  TuneBucket tune_by_spec[kStageSpecs][24];  // per stage-kernel specialisation
  int stage_spec = kStageGeneric;   // of the running instances (set at each topology rebuild)
  TuneBucket* tune = tune_by_spec[0];
  unsigned long long* d_span = nullptr;  // [stage 1 start, end, stage 2 start, end]
  unsigned long long* h_span = nullptr;
  int span_variant[2] = {-1, -1};
  double wall_khz = 0;
  hipGraphExec_t gexec = nullptr;
  // persistent step kernel (kernels.hpp): 0 = per-phase launches (the
  // default: on MI355X a grid-wide barrier across the XCDs costs more than the
  // kernel boundary it replaces), 1 / 2 = one launch per step with
  // kStepLanes[0] / [1] lanes per workgroup
  int step_mode = 0;
  int step_grid[kStageSpecs][2] = {{-1, -1}, {-1, -1}, {-1, -1}, {-1, -1}};  // co-resident grid; -1: not queried yet
  bool step_broken = false;  // a barrier wait ran out: per-phase launches from then on
  // inside amr_run_group: steps of several contexts may run at once on their
  // own streams, and two cooperative grids sized for the whole device could
  // each hold part of it; grouped contexts launch per phase
  bool in_group = false;
  std::vector<StepPhase> phases;
  GridSync* d_sync = nullptr;
  int32_t *d_fail = nullptr, *h_fail = nullptr;  // views into d_rec / h_rec, after the error records
  unsigned long long *d_stamps = nullptr, *h_stamps = nullptr;
  int stamps_pending = 0;  // phases whose device-clock stamps await collection
  // per-leaf CFL bounds and refinement indicators as pairs (lam, grad) per
  // leaf, in one device and one pinned host buffer (one download per step)
  double *d_red = nullptr, *h_red = nullptr;
  // d_err/h_err records first, then the reduction arrays, in one device and
  // one pinned host block sized for every slot: one download per step brings
  // the errors and the per-leaf values
  char *d_rec = nullptr, *h_rec = nullptr;
  size_t rec_err_bytes = 0;
  DevError* d_err = nullptr;  // one record per instance
  DevError* h_err = nullptr;
  int call_fail = 0;          // code of the first instance that failed in the current call
  int fail_nb = -1;           // block of a host-detected regrid failure
  RegridOp* d_rops = nullptr;
  int cap_rops = 0;
  // device-side regrid (regrid.hip): the device copy of every instance's
  // tree (instance q's blocks from tree_off[q]), its per-instance parameters
  // (view into d_topo) and results (views into d_rec / h_rec: per-instance
  // records, then the op log indexed like the leaf list); off: the host
  // decides from the downloaded per-leaf indicators
  bool regrid_device = true;
  bool regrid_lds = true;  // stage trees that fit in LDS (else every tree is worked on in place)
  DeviceTree dtree{};
  std::vector<int32_t> tree_off;
  RegridInst* d_rgi = nullptr;
  RegridOut *d_rgo = nullptr, *h_rgo = nullptr;
  int32_t *d_oplog = nullptr, *h_oplog = nullptr;
  size_t rec_off_log = 0, rec_off_red = 0;
  std::vector<uint8_t> rg_on;  // instances whose regrid the device runs (set at each topology rebuild)
  int32_t *d_cmapU = nullptr, *d_cmapB = nullptr;
  double* d_W = nullptr;  // initial-condition primitives for the first regrid
  bool topo_dirty = true;
  amr_diag diag{};
  // kernel timing: families selected by prof_mask are bracketed by events on
  // the context stream; elapsed times are collected after the next sync
  unsigned prof_mask = 0;
  Profile prof[PF_COUNT];
  // kernel families are timed on one launch in prof_stride (the others run
  // without events); prof_seen counts each family's launches
  int prof_stride = 1;
  int64_t prof_seen[PF_COUNT] = {};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<std::array<hipEvent_t, 2>> ev_pool;
  std::vector<std::pair<int, int>> ev_pending;  // (family, pool index)
};

namespace {

#define HIPCHK(x)                                                                    \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess) {                                                          \
      std::fprintf(stderr, "amr2d_mhd: HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
      return AMR_ERR_HIP;                                                            \
    }                                                                                \
  } while (0)

int set_error(amr_ctx* c, int code, int inst, int nb, int i, int j) {
  if (c->diag.code == 0) c->diag = amr_diag{code, inst, nb, i, j, 0};
  return code;
}

// An instance that hits one of the reference's throw sites stops where the
// reference would abort; the other instances of the context keep running.
void fail_instance(amr_ctx* c, int q, int code, int nb, int i, int j) {
  Instance& I = c->inst[q];
  if (I.status) return;
  I.status = code;
  I.err = amr_diag{code, q, nb, i, j, 0};
  I.done = true;
  c->topo_dirty = true;
  set_error(c, code, q, nb, i, j);
  if (!c->call_fail) c->call_fail = code;
}

int running(const amr_ctx* c) {
  int n = 0;
  for (auto& I : c->inst) n += I.done ? 0 : 1;
  return n;
}

template <class T>
int ensure(T*& p, int& cap, int n) {
  if (n <= cap) return 0;
  if (p) hipFree(p);
  int nc = std::max(n, cap * 2);
  if (nc < 64) nc = 64;
  if (hipMalloc((void**)&p, sizeof(T) * (size_t)nc) != hipSuccess) { p = nullptr; cap = 0; return AMR_ERR_HIP; }
  cap = nc;
  return 0;
}


// ----------------------------------------------------- timing families
// With `attached`, the caller hands the two events to the launch itself
// (hipExtLaunchKernelGGL) so they time the kernel; otherwise they are
// recorded on the stream around the launch.
struct Timed {
  amr_ctx* c;
  int fam, slot = -1;
  bool attached;
  Timed(amr_ctx* c_, int f, bool attached_ = false) : c(c_), fam(f), attached(attached_) {
    if (!(c->prof_mask & (1u << fam))) return;
    if (c->prof_seen[fam]++ % c->prof_stride) return;
    slot = (int)c->ev_pending.size();
    if (slot >= (int)c->ev_pool.size()) {
      std::array<hipEvent_t, 2> e{};
      hipEventCreate(&e[0]);
      hipEventCreate(&e[1]);
      c->ev_pool.push_back(e);
    }
    if (!attached) hipEventRecord(c->ev_pool[slot][0], c->stream);
  }
  hipEvent_t start() const { return slot >= 0 ? c->ev_pool[slot][0] : nullptr; }
  hipEvent_t stop() const { return slot >= 0 ? c->ev_pool[slot][1] : nullptr; }
  ~Timed() {
    if (slot < 0) return;
    if (!attached) hipEventRecord(c->ev_pool[slot][1], c->stream);
    c->ev_pending.push_back({fam, slot});
  }
};

// host wall time of a scope, accumulated into a profile family
struct HostTimed {
  amr_ctx* c;
  int fam;
  std::chrono::steady_clock::time_point t0;
  HostTimed(amr_ctx* c_, int f) : c(c_), fam(f), t0(std::chrono::steady_clock::now()) {}
  ~HostTimed() {
    if (!(c->prof_mask & (1u << fam))) return;
    c->prof[fam].ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    c->prof[fam].launches += 1;
  }
};

// collect the elapsed times of the recorded launches (stream already synced)
void flush_profile(amr_ctx* c) {
  // phases of a persistent step kernel: device-clock stamps at the start and
  // after each phase's closing barrier
  for (int p = 0; p < c->stamps_pending && c->wall_khz > 0; p++) {
    static const int fam_of[] = {PF_FILL, PF_STAGE, PF_CT, PF_CT, PF_CT, PF_CFL};
    const int fam = fam_of[c->phases[p].kind];
    if (!(c->prof_mask & (1u << fam))) continue;
    c->prof[fam].ms += (double)(c->h_stamps[p + 1] - c->h_stamps[p]) / c->wall_khz;
    c->prof[fam].launches += 1;
  }
  c->stamps_pending = 0;
  for (auto& p : c->ev_pending) {
    float ms = 0;
    hipEventElapsedTime(&ms, c->ev_pool[p.second][0], c->ev_pool[p.second][1]);
    c->prof[p.first].ms += ms;
    c->prof[p.first].launches += 1;
  }
  c->ev_pending.clear();
}

// -------------------------------------------------------- slot arena
int slot_of(amr_ctx* c, Instance& I, int nb) {
  while ((int)I.gs.size() <= nb) {
    if (c->used_slots >= c->cap_slots) return -1;
    I.gs.push_back((int32_t)c->used_slots++);
  }
  return I.gs[nb];
}

// ---------------------------------------------------- initial condition
// IC2Dtype.cpp:6-190 for one root block: primitives of the whole tile, then
// face B from cell averages (index 0 copies index 1).  rotor_p generalises
// the reference rotor with (rho_in, v0, p, Bx, r0, r1).
void initial_tile(const Instance& I, int nb, int NT, int ng, std::vector<double>& W, std::vector<double>& B) {
  const int N2 = NT * NT;
  W.assign(8 * N2, 0.0);
  B.assign(2 * N2, 0.0);
  double r4[4] = {0}, p4[4] = {0}, bx4[4] = {0}, by4[4] = {0};
  auto set4 = [](double* a, double x0, double x1, double x2, double x3) { a[0] = x0; a[1] = x1; a[2] = x2; a[3] = x3; };
  switch (I.ic) {
    case IC_SST: set4(p4, 0.1, 1.0, 1.0, 0.1); set4(r4, 0.125, 1.0, 1.0, 0.125); break;
    case IC_SST1: set4(p4, 1.0, 0.1, 0.1, 1.0); set4(r4, 1.0, 0.125, 0.125, 1.0); break;
    case IC_SSTY: set4(p4, 0.1, 0.1, 1.0, 1.0); set4(r4, 0.125, 0.125, 1.0, 1.0); break;
    case IC_BWX: set4(p4, 0.1, 1.0, 1.0, 0.1); set4(r4, 0.125, 1.0, 1.0, 0.125);
      set4(bx4, 0.75, 0.75, 0.75, 0.75); set4(by4, -1.0, 1.0, 1.0, -1.0); break;
    case IC_BWX1: set4(p4, 1.0, 0.1, 0.1, 1.0); set4(r4, 1.0, 0.125, 0.125, 1.0);
      set4(bx4, 0.75, 0.75, 0.75, 0.75); set4(by4, 1.0, -1.0, -1.0, 1.0); break;
    case IC_BWY: set4(p4, 0.1, 0.1, 1.0, 1.0); set4(r4, 0.125, 0.125, 1.0, 1.0);
      set4(by4, 0.75, 0.75, 0.75, 0.75); set4(bx4, -1.0, -1.0, 1.0, 1.0); break;
    case IC_BWY1: set4(p4, 1.0, 1.0, 0.1, 0.1); set4(r4, 1.0, 1.0, 0.125, 0.125);
      set4(by4, 0.75, 0.75, 0.75, 0.75); set4(bx4, 1.0, 1.0, -1.0, -1.0); break;
    default: break;
  }
  const int lev = I.lp[nb][0];
  // cell centres (IC2Dtype.cpp:92-94); the periodic geometry starts at 0
  const double dxl = I.dx[lev], dx0 = I.periodic ? 0. : I.dx[0];
  const double* rp = I.cfg.ic_params;
  for (int i = 0; i < NT; i++) {
    const double x = ((i + I.icoord[nb][0] - ng) + 0.5) * dxl - dx0;
    for (int j = 0; j < NT; j++) {
      const double y = ((j + I.icoord[nb][1] - ng) + 0.5) * dxl - dx0;
      double w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      if (I.ic == IC_OT) {
        // build-defined Orszag-Tang vortex on [0,1]^2
        const double pi = 3.141592653589793, B0 = 1. / sqrt(4. * pi);
        w[0] = 25. / (36. * pi);
        w[1] = -sin(2. * pi * y);
        w[2] = sin(2. * pi * x);
        w[4] = 5. / (12. * pi);
        w[5] = -B0 * sin(2. * pi * y);
        w[6] = B0 * sin(4. * pi * x);
      } else if (I.ic == IC_ROTOR || I.ic == IC_ROTOR_BS) {
        // IC_ROTOR_BS: the Balsara-Spicer rotor of the rotor workload's text
        // (p = 1, v0 = 2, Bx = 5/sqrt(4 pi)) in the reference rotor's form;
        // scaling by v0 = 1 or 2 is exact
        const bool bs = I.ic == IC_ROTOR_BS;
        const double v0 = bs ? 2. : 1.;
        const double rad = sqrt(((x - 0.5) * (x - 0.5)) + ((y - 0.5) * (y - 0.5)));
        const double fr = (23. - 200. * rad) / 3.;
        w[4] = bs ? 1.0 : 0.5;
        w[5] = bs ? 5.0 / (sqrt(4 * kPiRef)) : 2.5 / (sqrt(4 * kPiRef));
        if (rad <= 0.1) {
          w[0] = 10.;
          w[1] = -v0 * (y - 0.5) / 0.1;
          w[2] = v0 * (x - 0.5) / 0.1;
        } else if (rad > 0.1 && rad < 0.115) {
          w[0] = (1 + 9. * fr);
          w[1] = (-v0 * (y - 0.5) * fr / 0.1);
          w[2] = v0 * (x - 0.5) * fr / 0.1;
        } else {
          w[0] = 1.;
        }
      } else if (I.ic == IC_ROTOR_P) {
        const double rho_in = rp[0], v0 = rp[1], r0 = rp[4], r1 = rp[5];
        const double rad = sqrt(((x - 0.5) * (x - 0.5)) + ((y - 0.5) * (y - 0.5)));
        const double fr = (r1 - rad) / (r1 - r0);
        w[4] = rp[2];
        w[5] = rp[3];
        if (rad <= r0) {
          w[0] = rho_in;
          w[1] = -v0 * (y - 0.5) / r0;
          w[2] = v0 * (x - 0.5) / r0;
        } else if (rad > r0 && rad < r1) {
          w[0] = 1. + (rho_in - 1.) * fr;
          w[1] = -v0 * (y - 0.5) * fr / r0;
          w[2] = v0 * (x - 0.5) * fr / r0;
        } else {
          w[0] = 1.;
        }
      } else {
        int q;
        if (x >= 0.5 && y >= 0.5) q = 0;
        else if (x < 0.5 && y >= 0.5) q = 1;
        else if (x < 0.5 && y < 0.5) q = 2;
        else q = 3;
        w[0] = r4[q];
        w[4] = p4[q];
        w[5] = bx4[q];
        w[6] = by4[q];
      }
      for (int k = 0; k < 8; k++) W[k * N2 + j * NT + i] = w[k];
    }
  }
  for (int i = 0; i < NT; i++)
    for (int j = 0; j < NT; j++) {
      if (i < NT - 1) B[j * NT + i + 1] = 0.5 * (W[5 * N2 + j * NT + i] + W[5 * N2 + j * NT + i + 1]);
      else B[j * NT + 0] = B[j * NT + 1];
      if (j < NT - 1) B[N2 + (j + 1) * NT + i] = 0.5 * (W[6 * N2 + j * NT + i] + W[6 * N2 + (j + 1) * NT + i]);
      else B[N2 + 0 * NT + i] = B[N2 + 1 * NT + i];
    }
}

// meshblock.cpp:181-199
void prim_to_cons_tile(const std::vector<double>& W, std::vector<double>& U, int N2, double g) {
  U.assign(8 * N2, 0.0);
  for (int c = 0; c < N2; c++) {
    const double w0 = W[c], w1 = W[N2 + c], w2 = W[2 * N2 + c], w3 = W[3 * N2 + c];
    const double b5 = W[5 * N2 + c], b6 = W[6 * N2 + c], b7 = W[7 * N2 + c];
    U[c] = w0;
    U[N2 + c] = w1 * w0;
    U[2 * N2 + c] = w2 * w0;
    U[3 * N2 + c] = w3 * w0;
    const double e = (W[4 * N2 + c] / (g - 1)) + 0.5 * w0 * (w1 * w1 + w2 * w2 + w3 * w3);
    U[4 * N2 + c] = e + 0.5 * (b5 * b5 + b6 * b6 + b7 * b7);
    U[5 * N2 + c] = b5;
    U[6 * N2 + c] = b6;
    U[7 * N2 + c] = b7;
  }
}

// ------------------------------------------------------ fill schedule
// locate_bounds.cpp:35-208: ordered (block, neighbour, code) list; every
// shared side of two leaves appears once, from the first block in id order
// (or from the coarser block for a level jump).
void build_schedule(Instance& I) {
  I.sched.clear();
  std::vector<uint8_t> seen((size_t)I.lastActive * 4, 0);
  // per side: neighbour slot in lp, opposite side, codes, son offset of the
  // finer neighbour's first adjacent son and the stride to the second
  struct SideInfo { int opp, box, same, finer, son, stride; };
  static const SideInfo S[4] = {{1, -1, 1, 2, 1, 2}, {0, -2, 4, 5, 0, 2}, {3, -3, 7, 8, 2, 1}, {2, -4, 10, 11, 0, 1}};
  for (int nb = 0; nb < I.lastActive; nb++) {
    if (!I.leafs[nb]) continue;
    for (int s = 0; s < 4; s++) {
      if (seen[nb * 4 + s]) continue;
      const int n = I.lp[nb][4 + s];
      if (n == -1) {
        seen[nb * 4 + s] = 1;
        I.sched.push_back({nb, -1, S[s].box});
      } else if (I.lp[n][0] >= I.lp[nb][0]) {
        seen[nb * 4 + s] = 1;
        if (I.leafs[n]) {
          I.sched.push_back({nb, n, S[s].same});
          seen[n * 4 + S[s].opp] = 1;
        } else {
          const int f1 = I.lp[n][2] + S[s].son, f2 = f1 + S[s].stride;
          I.sched.push_back({nb, f1, S[s].finer});
          seen[f1 * 4 + S[s].opp] = 1;
          seen[f2 * 4 + S[s].opp] = 1;
        }
      }
    }
  }
}

// ---------------------------------------------------- fill compiler
// Symbolic execution of the reference's sequential ghost fill
// (boundary.cpp / boundaryfc.cpp / boundaryec.cpp, extboundary.cpp): each
// element holds a symbolic value -- its own pre-fill value, a pre-fill value
// of another element (copies collapse) or a restriction node.  The final
// symbolic value of every written element becomes one FillOp, evaluated on
// the device as gather (from the pre-fill state) and scatter, which
// reproduces the sequential last-writer semantics without races.  One
// program serves all fields: only the restriction arithmetic depends on the
// field kind.

struct FillCompiler {
  const Instance& I;
  int NT, ng, N2;
  const FillScratch& S;
  FillWorker& W;

  FillCompiler(const Instance& inst, int NT_, int ng_, const FillScratch& s, FillWorker& w)
      : I(inst), NT(NT_), ng(ng_), N2(NT_ * NT_), S(s), W(w) {
    W.touched.clear();
    W.nodes.clear();
    W.interior_write = false;
  }
  static int32_t lref(int nb, int cell) { return (nb << 16) | cell; }
  // table slot of ring cell `cell` of block nb
  size_t slot(int nb, int cell) const { return (size_t)nb * S.nring + S.ring_idx[cell]; }
  // fills write ghost-ring elements only: an interior element always holds
  // its pre-fill value
  int32_t read(int nb, int i, int j) const {
    const int cell = j * NT + i;
    if (!S.ring[cell]) return lref(nb, cell);
    const int32_t v = W.cur[slot(nb, cell)];
    return v != -1 ? v : lref(nb, cell);
  }
  void write(int nb, int i, int j, int32_t v) {
    const int cell = j * NT + i;
    if (!S.ring[cell]) {
      W.interior_write = true;
      return;
    }
    const size_t a = slot(nb, cell);
    if (W.cur[a] == -1) W.touched.push_back((nb << 16) | cell);
    W.cur[a] = v;
  }
  bool unwritten(int nb, int cell) const { return !S.ring[cell] || W.cur[slot(nb, cell)] == -1; }
  void copy(int dn, int di, int dj, int sn, int si, int sj) { write(dn, di, dj, read(sn, si, sj)); }
  void restrict_(int dn, int di, int dj, int fn, int i1, int j1) {
    RNode nd{};
    nd.nb = fn;
    nd.cell = (uint16_t)(j1 * NT + i1);
    const int c0 = j1 * NT + i1;
    nd.simple = !S.force_general && unwritten(fn, c0) && unwritten(fn, c0 + NT) && unwritten(fn, c0 + 1) &&
                unwritten(fn, c0 + NT + 1);
    if (!nd.simple) {
      nd.ref[0] = read(fn, i1, j1);
      nd.ref[1] = read(fn, i1, j1 + 1);
      nd.ref[2] = read(fn, i1 + 1, j1);
      nd.ref[3] = read(fn, i1 + 1, j1 + 1);
    }
    W.nodes.push_back(nd);
    write(dn, di, dj, -(int32_t)W.nodes.size() - 1);  // node index size-1 -> -(index + 2)
  }
  // corner of an entry: the diagonal block `cb` found through the neighbour
  // table is copied when it is a leaf (with same-level indices even if it is
  // coarser), else its designated son is restricted when that is a leaf.
  template <class Dst, class SameIdx, class FineIdx>
  void corner(int nbs, int cb, int son_off, int ilo, int ihi, int jlo, int jhi, Dst dst, SameIdx same, FineIdx fine) {
    if (cb == -1) return;
    if (I.leafs[cb]) {
      for (int i = ilo; i < ihi; i++)
        for (int j = jlo; j < jhi; j++) {
          int di, dj, si, sj;
          dst(i, j, di, dj);
          same(i, j, si, sj);
          copy(nbs, di, dj, cb, si, sj);
        }
      return;
    }
    const int fb = I.lp[cb][2] + son_off;
    if (!I.leafs[fb]) return;
    for (int i = ilo; i < ihi; i++)
      for (int j = jlo; j < jhi; j++) {
        int di, dj, i1, j1;
        dst(i, j, di, dj);
        fine(i, j, i1, j1);
        restrict_(nbs, di, dj, fb, i1, j1);
      }
  }

  // entries seq[k0, k1) of the schedule in execution order
  void run(int k0, int k1) {
    for (int k = k0; k < k1; k++) entry(I.sched[S.seq[k]]);
  }

  void entry(const std::array<int32_t, 3>& e) {
    const int lo = ng, hi = NT - 1 - ng, lob = lo - 1, hib = hi + 1, h2 = (NT - 1) / 2, top = hib;
    auto id = [](int i, int j, int& a, int& b) { a = i; b = j; };
    {
      const int nbs = e[0], n1 = e[1];
      switch (e[2]) {
        case 1:   // left neighbour, same level (boundary.cpp:17-35)
        case 4:   // right neighbour, same level (boundary.cpp:109-127)
        {
          const bool left = e[2] == 1;
          for (int j = 0; j < NT; j++)
            for (int g = 0; g < ng; g++) {
              if (left) {
                copy(nbs, g, j, n1, hi - (ng - 1) + g, j);
                copy(n1, top + g, j, nbs, lo + g, j);
              } else {
                copy(nbs, top + g, j, n1, lo + g, j);
                copy(n1, g, j, nbs, hi - (ng - 1) + g, j);
              }
            }
          break;
        }
        case 7:   // bottom neighbour, same level (boundary.cpp:199-217)
        case 10:  // top neighbour, same level (boundary.cpp:289-307)
        {
          const bool bottom = e[2] == 7;
          for (int i = 0; i < NT; i++)
            for (int g = 0; g < ng; g++) {
              if (bottom) {
                copy(nbs, i, g, n1, i, hi - (ng - 1) + g);
                copy(n1, i, top + g, nbs, i, lo + g);
              } else {
                copy(nbs, i, top + g, n1, i, lo + g);
                copy(n1, i, g, nbs, i, hi - (ng - 1) + g);
              }
            }
          break;
        }
        case 2: {  // finer neighbours on the left (boundary.cpp:36-107)
          const int n2 = n1 + 2;
          for (int j = lo; j <= h2; j++)
            for (int i = 0; i < ng; i++) {
              const int i1 = 2 * (i - ng / 2) - ng / 2 + hi, j1 = 2 * j - lo;
              restrict_(nbs, i, j, n1, i1, j1);
              restrict_(nbs, i, h2 + j - 1 - (ng - 2), n2, i1, j1);
            }
          corner(nbs, I.lp[I.lp[n1][1]][6], 3, 0, ng, 0, ng, id,
                 [&](int i, int j, int& a, int& b) { a = hi - ng + i + 1; b = hi - ng + j + 1; },
                 [&](int i, int j, int& a, int& b) { a = hi - 2 * ng + 2 * i + 1; b = hi - 2 * ng + 2 * j + 1; });
          corner(nbs, I.lp[I.lp[n2][1]][7], 1, 0, ng, 0, ng,
                 [&](int i, int j, int& a, int& b) { a = i; b = NT - 1 - j; },
                 [&](int i, int j, int& a, int& b) { a = hi - ng + i + 1; b = lo + ng - (j + 1); },
                 [&](int i, int j, int& a, int& b) { a = hi - 2 * ng + 2 * i + 1; b = lo + 2 * ng - 2 * (j + 1); });
          for (int i = 1; i <= ng; i++)
            for (int j = 0; j < NT; j++) {
              copy(n1, i + hi, j, nbs, (i + 2 * ng - 1) / 2, (j + ng) / 2);
              copy(n2, i + hi, j, nbs, (i + 2 * ng - 1) / 2, (j - ng + 2) / 2 + h2);
            }
          break;
        }
        case 5: {  // finer neighbours on the right (boundary.cpp:128-197)
          const int n2 = n1 + 2;
          for (int i = top; i < NT; i++)
            for (int j = lo; j <= h2; j++) {
              const int i1 = 2 * (i - hi + (ng / 2 - 1)), j1 = 2 * j - lo;
              restrict_(nbs, i, j, n1, i1, j1);
              restrict_(nbs, i, h2 + j - 1 - (ng - 2), n2, i1, j1);
            }
          corner(nbs, I.lp[I.lp[n1][1]][6], 2, hib, NT, 0, ng, id,
                 [&](int i, int j, int& a, int& b) { a = lob + (i - hi); b = hi - ng + j + 1; },
                 [&](int i, int j, int& a, int& b) { a = lob + 2 * (i - hi) - 1; b = hi - 2 * ng + 2 * j + 1; });
          corner(nbs, I.lp[I.lp[n2][1]][7], 0, hib, NT, 0, ng,
                 [&](int i, int j, int& a, int& b) { a = i; b = NT - 1 - j; },
                 [&](int i, int j, int& a, int& b) { a = lob + (i - hi); b = lo + ng - (j + 1); },
                 [&](int i, int j, int& a, int& b) { a = lob + 2 * (i - hi) - 1; b = lo + 2 * ng - 2 * (j + 1); });
          for (int i = 1; i <= ng; i++)
            for (int j = 0; j < NT; j++) {
              copy(n1, i - 1, j, nbs, (i + 1) / 2 + hi - ng / 2, (j + ng) / 2);
              copy(n2, i - 1, j, nbs, (i + 1) / 2 + hi - ng / 2, (j - ng + 2) / 2 + h2);
            }
          break;
        }
        case 8: {  // finer neighbours below (boundary.cpp:218-287)
          const int n2 = n1 + 1;
          for (int i = lo; i <= h2; i++)
            for (int j = 0; j < ng; j++) {
              const int i1 = 2 * i - lo, j1 = 2 * (j - ng / 2) - ng / 2 + hi;
              restrict_(nbs, i, j, n1, i1, j1);
              restrict_(nbs, h2 + i - 1 - (ng - 2), j, n2, i1, j1);
            }
          corner(nbs, I.lp[I.lp[n1][1]][4], 3, 0, ng, 0, ng, id,
                 [&](int i, int j, int& a, int& b) { a = hi - ng + i + 1; b = hi - ng + j + 1; },
                 [&](int i, int j, int& a, int& b) { a = hi - 2 * ng + 2 * i + 1; b = hi - 2 * ng + 2 * j + 1; });
          corner(nbs, I.lp[I.lp[n2][1]][5], 2, 0, ng, 0, ng,
                 [&](int i, int j, int& a, int& b) { a = NT - 1 - i; b = j; },
                 [&](int i, int j, int& a, int& b) { a = lo + ng - (i + 1); b = hi - ng + j + 1; },
                 [&](int i, int j, int& a, int& b) { a = lo + 2 * ng - 2 * (i + 1); b = hi - 2 * ng + 2 * j + 1; });
          for (int j = 1; j <= ng; j++)
            for (int i = 0; i < NT; i++) {
              copy(n1, i, j + hi, nbs, (i + ng) / 2, (j + 2 * ng - 1) / 2);
              copy(n2, i, j + hi, nbs, (i - ng + 2) / 2 + h2, (j + 2 * ng - 1) / 2);
            }
          break;
        }
        case 11: {  // finer neighbours above (boundary.cpp:308-377)
          const int n2 = n1 + 1;
          for (int i = lo; i <= h2; i++)
            for (int j = top; j < NT; j++) {
              const int i1 = 2 * i - lo, j1 = 2 * (j - hi + (ng / 2 - 1));
              restrict_(nbs, i, j, n1, i1, j1);
              restrict_(nbs, h2 + i - 1 - (ng - 2), j, n2, i1, j1);
            }
          corner(nbs, I.lp[I.lp[n1][1]][4], 1, 0, ng, hib, NT, id,
                 [&](int i, int j, int& a, int& b) { a = hi - ng + i + 1; b = lob + (j - hi); },
                 [&](int i, int j, int& a, int& b) { a = hi - 2 * ng + 2 * i + 1; b = lob + 2 * (j - hi) - 1; });
          {  // right corner: the same-level copy mirrors i, the finer restriction does not
            const int cb = I.lp[I.lp[n2][1]][5];
            if (cb != -1) {
              if (I.leafs[cb]) {
                for (int i = 0; i < ng; i++)
                  for (int j = hib; j < NT; j++) copy(nbs, NT - 1 - i, j, cb, lo + ng - (i + 1), lob + (j - hi));
              } else {
                const int fb = I.lp[cb][2];
                if (I.leafs[fb])
                  for (int i = hib; i < NT; i++)
                    for (int j = hib; j < NT; j++)
                      restrict_(nbs, i, j, fb, lob + 2 * (i - hi) - 1, lob + 2 * (j - hi) - 1);
              }
            }
          }
          for (int j = 1; j <= ng; j++)
            for (int i = 0; i < NT; i++) {
              copy(n1, i, j - 1, nbs, (i + ng) / 2, (j + 1) / 2 + hi - ng / 2);
              copy(n2, i, j - 1, nbs, (i - ng + 2) / 2 + h2, (j + 1) / 2 + hi - ng / 2);
            }
          break;
        }
        // extboundary.cpp:5-61: mirror of the first ng interior rows/columns,
        // applied to every field kind alike, after all internal entries
        case -1:
          for (int g = 0; g < ng; g++)
            for (int j = 0; j < NT; j++) copy(nbs, g, j, nbs, lo + ng - 1 - g, j);
          break;
        case -2:
          for (int g = 0; g < ng; g++)
            for (int j = 0; j < NT; j++) copy(nbs, top + g, j, nbs, hi - g, j);
          break;
        case -3:
          for (int g = 0; g < ng; g++)
            for (int i = 0; i < NT; i++) copy(nbs, i, g, nbs, i, lo + ng - 1 - g);
          break;
        case -4:
          for (int g = 0; g < ng; g++)
            for (int i = 0; i < NT; i++) copy(nbs, i, top + g, nbs, i, hi - g);
          break;
        default: break;
      }
    }
  }
};

// f(r) for r in [0, R) on the scratch's helper threads and the caller
template <class F>
void for_ranges(FillScratch& S, int R, int nthreads, const F& f) {
  if (R == 1 || nthreads <= 1) {
    for (int r = 0; r < R; r++) f(r);
    return;
  }
  if (!S.pool || S.pool->helpers() != nthreads - 1) S.pool.reset(new RangePool(nthreads - 1));
  S.pool->run(R, f);
}

// ring elements (blocks x ring cells) from which a compile is split over
// threads (about 115 blocks of a 20x20 tile): below, the extra merge pass
// costs more than it saves.  Whole rotor run on the 16 host threads of a GPU,
// same box, interleaved: 1.407-1.412 s splitting from 2^15 elements with
// helpers that spin 200 us, 1.377-1.387 s with a 3 ms spin, 1.378-1.379 s
// with the 3 ms spin from 2^14.
constexpr size_t kSplitSlots = size_t(1) << 14;

// Program of one instance: 16-byte destination ops plus the explicit-operand
// table of restrictions that read an element already written in the same
// fill.  Nested restrictions (a restriction reading another restriction's
// destination) cannot occur for nghosts 2 and 4 -- the fine elements a
// restriction reads are interior cells or, for nghosts 4, ghost cells that
// only the same entry's later injection writes -- and are rejected.
//
// Parallel symbolic execution: the schedule (internal entries, then the
// external ones) is cut into R contiguous ranges run independently, each on
// its own table, treating the state at its start as "pre-fill".  Range r's
// reference to the start value of a ring element that an earlier range wrote
// is then replaced by that range's value (recursively, through the last
// earlier writer); a restriction whose operands an earlier range wrote loses
// its `simple` form the same way.  The final value of an element is the one
// of the last range that writes it, and ops are laid out in first-write order
// over the whole schedule -- so the program is the one a single sequential
// pass produces, whatever R is.
//
// Chunking: an op may read (as a pre-fill value) an element that another op
// writes -- the reference's stale corner reads.  Such readers and writers
// are linked into groups that are laid out inside one kFillChunk chunk, so
// the single-pass kernel (read all, barrier, write all per chunk) keeps the
// sequential semantics.  The segment is padded with kFillNop to whole chunks.
int compile_instance(Instance& I, int NT, int ng, FillScratch& S, int nthreads = 1) {
  I.ops.clear();
  I.gen.clear();
  const int N2 = NT * NT;
  const int off[4] = {0, NT, 1, NT + 1};
  if (S.ring_nt != NT) {
    S.ring.assign(N2, 0);
    S.ring_idx.assign(N2, -1);
    S.nring = 0;
    for (int c = 0; c < N2; c++) {
      const int i = c % NT, j = c / NT;
      S.ring[c] = (i < ng || i >= NT - ng || j < ng || j >= NT - ng) ? 1 : 0;
      if (S.ring[c]) S.ring_idx[c] = S.nring++;
    }
    S.ring_nt = NT;
    for (auto& w : S.w) w.cur.clear();
  }
  if (I.lastActive > 32767) return AMR_ERR_CONFIG;  // block ids are 15-bit in the symbolic values
  const uint8_t* ring = S.ring.data();
  const int32_t* ridx = S.ring_idx.data();
  const int32_t* gs = I.gs.data();
  const int nring = S.nring;
  // table slot of element t = (nb << 16) | cell
  auto slot = [&](int32_t t) { return (size_t)(t >> 16) * nring + ridx[t & 0xffff]; };
  // schedule in execution order: internal entries, then extboundary's
  S.seq.clear();
  for (size_t e = 0; e < I.sched.size(); e++)
    if (I.sched[e][2] > 0) S.seq.push_back((int32_t)e);
  for (size_t e = 0; e < I.sched.size(); e++)
    if (I.sched[e][2] < 0) S.seq.push_back((int32_t)e);
  const int nseq = (int)S.seq.size();
  const size_t nslots = (size_t)I.lastActive * nring;
  int R = 1;
  if (S.force_ranges > 0) R = S.force_ranges;
  else if (nthreads > 1 && nslots >= kSplitSlots) R = nthreads;
  R = std::max(1, std::min({R, 16, std::max(1, nseq)}));
  if ((int)S.w.size() < R) S.w.resize(R);
  if (R > 1 && S.wmask.size() < nslots) S.wmask.resize(nslots, 0);
  for (int r = 0; r < R; r++)
    if (S.w[r].cur.size() < nslots) S.w[r].cur.resize(nslots, -1);

  // 1. every range on its own table; with several, wmask records which
  //    ranges write what
  for_ranges(S, R, nthreads, [&](int r) {
    FillWorker& W = S.w[r];
    FillCompiler fc(I, NT, ng, S, W);
    fc.run((int)((int64_t)nseq * r / R), (int)((int64_t)nseq * (r + 1) / R));
    if (R > 1) {
      const uint16_t bit = (uint16_t)(1u << r);
      for (int32_t t : W.touched) __atomic_fetch_or(&S.wmask[slot(t)], bit, __ATOMIC_RELAXED);
    }
  });
  S.nodeoff.assign(R + 1, 0);
  for (int r = 0; r < R; r++) S.nodeoff[r + 1] = S.nodeoff[r] + (int32_t)S.w[r].nodes.size();

  // value v of range r in pre-fill terms: a start value of a ring element an
  // earlier range wrote becomes that range's value; nodes get global indices
  auto subst = [&](int r, int32_t v) -> int32_t {
    for (;;) {
      if (v <= -2) return -(2 + S.nodeoff[r] + (-v - 2));
      const int cell = v & 0xffff;
      if (!ring[cell]) return v;
      const size_t a = (size_t)(v >> 16) * nring + ridx[cell];
      const uint32_t m = S.wmask[a] & ((1u << r) - 1u);
      if (!m) return v;
      r = 31 - __builtin_clz(m);
      v = S.w[r].cur[a];
    }
  };

  // 2. several ranges: per range, its nodes in pre-fill terms and the final
  //    values of the elements it writes first (one range: its own table)
  if (R > 1) {
    S.rnodes.resize(S.nodeoff[R]);
    for_ranges(S, R, nthreads, [&](int r) {
      FillWorker& W = S.w[r];
      const uint32_t before = (1u << r) - 1u;
      for (size_t k = 0; k < W.nodes.size(); k++) {
        const RNode& nd = W.nodes[k];
        RNode o = nd;
        if (nd.simple) {
          bool still = true;
          for (int q = 0; q < 4 && r > 0; q++) {
            const int c = nd.cell + off[q];
            if (ring[c] && (S.wmask[(size_t)nd.nb * nring + ridx[c]] & before)) still = false;
          }
          o.simple = still;
          if (!still)
            for (int q = 0; q < 4; q++) o.ref[q] = subst(r, FillCompiler::lref(nd.nb, nd.cell + off[q]));
        } else {
          for (int q = 0; q < 4; q++) o.ref[q] = subst(r, (int32_t)nd.ref[q]);
        }
        S.rnodes[S.nodeoff[r] + k] = o;
      }
      W.first.clear();
      W.fval.clear();
      W.nkeep = 0;
      for (int32_t t : W.touched) {
        const size_t a = slot(t);
        const uint32_t m = S.wmask[a];
        if ((int)__builtin_ctz(m) != r) continue;
        const int rl = 31 - __builtin_clz(m);
        const int32_t v = subst(rl, S.w[rl].cur[a]);
        W.first.push_back(t);
        W.fval.push_back(v);
        W.nkeep += v != t;  // v == t: the element keeps its own value, no op
      }
    });
  }
  const RNode* nodes = R > 1 ? S.rnodes.data() : S.w[0].nodes.data();
  bool interior_write = false;
  std::vector<int32_t> opoff(R + 1, 0);
  for (int r = 0; r < R; r++) {
    interior_write |= S.w[r].interior_write;
    opoff[r + 1] = opoff[r] + (R > 1 ? S.w[r].nkeep : (int)S.w[0].touched.size());
  }
  I.ops.reserve((size_t)opoff[R] + kFillChunk);
  I.ops.resize_uninit(opoff[R]);

  // 3. ops in first-write order; the first-writing range's table entry of a
  //    written element becomes its op index (-1: no op) for the stale-read
  //    search below
  for_ranges(S, R, nthreads, [&](int r) {
    FillWorker& W = S.w[r];
    W.gen.clear();
    W.gen_ops.clear();
    W.ringsrc.clear();
    W.nested = false;
    int64_t nk[3] = {0, 0, 0};
    // locals, so that the stores below cannot be taken to alias them
    const uint8_t* __restrict rg = ring;
    const int32_t* __restrict ri = ridx;
    const int32_t* __restrict gsl = gs;
    const RNode* __restrict nds = nodes;
    int32_t* __restrict cur = W.cur.data();
    const int32_t* __restrict tl = R > 1 ? W.first.data() : W.touched.data();
    const int32_t* __restrict vl = R > 1 ? W.fval.data() : nullptr;
    FillOp* __restrict ops = I.ops.data();
    const int nr = nring, nt = NT;
    const size_t m = R > 1 ? W.first.size() : W.touched.size();
    int32_t o = opoff[r];
    for (size_t i = 0; i < m; i++) {
      const int32_t t = tl[i];
      const size_t a = (size_t)(t >> 16) * nr + ri[t & 0xffff];
      const int32_t v = vl ? vl[i] : cur[a];
      if (v == t) {  // the element keeps its own value: no op
        cur[a] = -1;
        continue;
      }
      FillOp op{};
      op.dst_gs = gsl[t >> 16];
      op.dst_cell = (uint16_t)(t & 0xffff);
      RingSrc rs;
      rs.nref = 0;
      if (v >= 0) {
        op.op = kFillCopy;
        op.src = gsl[v >> 16];
        op.src_cell = (uint16_t)(v & 0xffff);
        if (rg[op.src_cell]) { rs.nref = 1; rs.ref[0] = v; }
      } else {
        const RNode& nd = nds[-v - 2];
        if (nd.simple) {
          op.op = kFillRestrict;
          op.src = gsl[nd.nb];
          op.src_cell = nd.cell;
          if (rg[nd.cell] | rg[nd.cell + 1] | rg[nd.cell + nt] | rg[nd.cell + nt + 1]) {
            rs.nref = 4;
            for (int q = 0; q < 4; q++) rs.ref[q] = (nd.nb << 16) | (nd.cell + off[q]);
          }
        } else {
          FillGen g{};
          rs.nref = 4;
          for (int q = 0; q < 4; q++) {
            const int64_t ref = nd.ref[q];
            rs.ref[q] = ref;
            if (ref < 0) {  // a restriction of a restriction
              W.nested = true;
              continue;
            }
            g.ref[q] = src_ref(gsl[ref >> 16], (int)(ref & 0xffff));
          }
          op.op = kFillGeneral;
          op.src = (int32_t)W.gen.size();
          W.gen_ops.push_back(o);
          W.gen.push_back(g);
        }
      }
      if (rs.nref) {
        rs.op = o;
        W.ringsrc.push_back(rs);
      }
      cur[a] = o;  // from here on: the op writing the element
      nk[op.op]++;
      ops[o++] = op;
    }
    for (int k = 0; k < 3; k++) W.nkind[k] = nk[k];
    if (R == 1) opoff[1] = o;
  });
  const int n = opoff[R];
  I.ops.resize_uninit(n);
  int status = interior_write ? AMR_ERR_CONFIG : 0;
  // ops per kind (the padding added below is counted by the caller)
  I.op_kinds = {0, 0, 0, 0};
  for (int r = 0; r < R; r++)
    for (int k = 0; k < 3; k++) I.op_kinds[k] += S.w[r].nkind[k];
  for (int r = 0; r < R; r++) {
    FillWorker& W = S.w[r];
    if (W.nested) status = AMR_ERR_CONFIG;
    const int32_t base = (int32_t)I.gen.size();
    for (int32_t q : W.gen_ops) I.ops.data()[q].src += base;
    I.gen.insert(I.gen.end(), W.gen.begin(), W.gen.end());
  }

  // 4. stale reads: an op reading (pre-fill) an element another op writes is
  //    linked with that writer; fills write ghost-ring elements only, so only
  //    operands in a ring can be written ones
  for_ranges(S, R, nthreads, [&](int r) {
    FillWorker& W = S.w[r];
    W.edges.clear();
    W.self_read = false;
    for (const auto& rs : W.ringsrc)
      for (int t = 0; t < rs.nref; t++) {
        const int64_t ref = rs.ref[t];
        if (ref < 0) continue;
        const int rc = (int)(ref & 0xffff);
        if (!ring[rc]) continue;
        const size_t a = (size_t)(ref >> 16) * nring + ridx[rc];
        int32_t w = -1;
        if (R == 1) {
          w = W.cur[a];
        } else if (const uint32_t m = S.wmask[a]) {
          w = S.w[__builtin_ctz(m)].cur[a];
        }
        if (w == rs.op) W.self_read = true;
        if (w < 0 || w == rs.op) continue;
        W.edges.push_back({rs.op, w});
      }
  });
  size_t nedges = 0;
  bool self_read = false;
  for (int r = 0; r < R; r++) {
    nedges += S.w[r].edges.size();
    self_read |= S.w[r].self_read;
  }
  // groups = connected components (lock-free union-find, -1: not linked; a
  // root is linked under a smaller one, so every component's root is its
  // smallest op); the tables go back to "unwritten" where this compile
  // touched them
  if ((int)S.uf.size() < n) S.uf.resize(n);
  int32_t* uf = S.uf.data();
  if (nedges)
    for_ranges(S, R, nthreads, [&](int r) { std::fill(uf + opoff[r], uf + opoff[r + 1], -1); });
  auto find = [uf](int32_t x) {
    for (;;) {
      int32_t p = __atomic_load_n(&uf[x], __ATOMIC_ACQUIRE);
      if (p == x) return x;
      const int32_t gp = __atomic_load_n(&uf[p], __ATOMIC_ACQUIRE);
      if (gp != p) __atomic_compare_exchange_n(&uf[x], &p, gp, false, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE);
      x = gp;
    }
  };
  if (R == 1) {  // one thread: plain union-find, the same components and roots
    FillWorker& W = S.w[0];
    auto find1 = [uf](int32_t x) {
      while (uf[x] != x) x = uf[x] = uf[uf[x]];
      return x;
    };
    for (const auto& e : W.edges) {
      if (uf[e[0]] == -1) uf[e[0]] = e[0];
      if (uf[e[1]] == -1) uf[e[1]] = e[1];
      const int32_t a = find1(e[0]), b = find1(e[1]);
      if (a != b) uf[std::max(a, b)] = std::min(a, b);
    }
    for (int32_t t : W.touched) W.cur[slot(t)] = -1;
  } else {
    for_ranges(S, R, nthreads, [&](int r) {
      FillWorker& W = S.w[r];
      for (const auto& e : W.edges) {
        for (int k = 0; k < 2; k++) {
          int32_t unset = -1;
          __atomic_compare_exchange_n(&uf[e[k]], &unset, e[k], false, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE);
        }
        int32_t a = e[0], b = e[1];
        for (;;) {
          a = find(a);
          b = find(b);
          if (a == b) break;
          if (a < b) std::swap(a, b);
          int32_t expect = a;
          if (__atomic_compare_exchange_n(&uf[a], &expect, b, false, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE)) break;
        }
      }
      for (int32_t t : W.touched) {
        const size_t a = slot(t);
        W.cur[a] = -1;
        __atomic_store_n(&S.wmask[a], (uint16_t)0, __ATOMIC_RELAXED);
      }
    });
  }

  // 5. linked ops move behind the others, grouped -- groups in the order of
  //    their smallest op, ops of a group in op order -- each group inside one
  //    chunk; their old positions become padding; whole chunks
  FillOp nop{};
  nop.op = kFillNop;
  auto pad_to_chunk = [&]() {
    while (I.ops.size() % kFillChunk) I.ops.push_back(nop);
  };
  if (!nedges) {
    pad_to_chunk();
    I.tail_chunk0 = self_read ? 0 : (int)(I.ops.size() / kFillChunk);
    return status;
  }
  // linked ops of every range's op segment with their roots (all unions are
  // done: plain finds); roots numbered in op order, their union-find entry
  // holding -(group + 2) from then on
  std::vector<int32_t> gbase(R + 1, 0);
  for_ranges(S, R, nthreads, [&](int r) {
    FillWorker& W = S.w[r];
    W.linked.clear();
    W.gid.clear();
    int32_t nroots = 0;
    for (int32_t q = opoff[r]; q < opoff[r + 1]; q++)
      if (uf[q] != -1) {
        int32_t x = q;
        while (uf[x] != x) x = uf[x];
        W.linked.push_back(q);
        W.gid.push_back(x);
        nroots += x == q;
      }
    gbase[r + 1] = nroots;
  });
  for (int r = 0; r < R; r++) gbase[r + 1] += gbase[r];
  const int32_t ngroups = gbase[R];
  for_ranges(S, R, nthreads, [&](int r) {
    FillWorker& W = S.w[r];
    int32_t g = gbase[r];
    for (size_t i = 0; i < W.linked.size(); i++)
      if (W.gid[i] == W.linked[i]) uf[W.linked[i]] = -(g++ + 2);
  });
  for_ranges(S, R, nthreads, [&](int r) {
    FillWorker& W = S.w[r];
    for (size_t i = 0; i < W.linked.size(); i++) W.gid[i] = -uf[W.gid[i]] - 2;
  });
  // group sizes, then each group's place behind the unlinked ops (a group
  // that does not fit the rest of a chunk starts the next one)
  S.group_count.assign(ngroups, 0);
  for (int r = 0; r < R; r++)
    for (int32_t g : S.w[r].gid) S.group_count[g] += 1;
  S.group_pos.resize(ngroups);
  const size_t tail = ((size_t)n + kFillChunk - 1) / kFillChunk * kFillChunk;
  I.tail_chunk0 = self_read ? 0 : (int)(tail / kFillChunk);
  size_t pos = tail;
  for (int32_t g = 0; g < ngroups; g++) {
    const size_t m = (size_t)S.group_count[g];
    if (m > (size_t)kFillChunk) return AMR_ERR_CONFIG;
    if (pos % kFillChunk + m > (size_t)kFillChunk) pos = (pos + kFillChunk - 1) / kFillChunk * kFillChunk;
    S.group_pos[g] = (int32_t)pos;
    pos += m;
  }
  const size_t total = (pos + kFillChunk - 1) / kFillChunk * kFillChunk;
  // destinations, ops of a group in op order
  for (int r = 0; r < R; r++)
    for (int32_t& g : S.w[r].gid) g = S.group_pos[g]++;
  I.ops.reserve(total);
  I.ops.resize_uninit(total);
  FillOp* p = I.ops.data();
  for_ranges(S, R, nthreads, [&](int r) {
    // padding of this range's share of the tail, then its linked ops
    std::fill(p + n + (total - n) * r / R, p + n + (total - n) * (r + 1) / R, nop);
  });
  for_ranges(S, R, nthreads, [&](int r) {
    const FillWorker& W = S.w[r];
    for (size_t i = 0; i < W.linked.size(); i++) {
      p[W.gid[i]] = p[W.linked[i]];
      p[W.linked[i]] = nop;
    }
  });
  return status;
}

// ------------------------------------------------- tree bookkeeping
// updatelpup.cpp:4-118: link four new sons of nb into the neighbour table
void link_sons(Instance& I, int nb, const int s[4]) {
  auto& lp = I.lp;
  const int lev = lp[nb][0] + 1;
  // adjacent son of a refined neighbour, or the neighbour itself
  auto outer = [&](int son, int side, int nsr_off, int back_side) {
    const int n = lp[nb][side];
    if (n == -1 || I.leafs[n]) {
      lp[son][side] = n;
    } else {
      const int nsr = lp[n][2] + nsr_off;
      lp[son][side] = nsr;
      lp[nsr][back_side] = son;
    }
  };
  for (int q = 0; q < 4; q++) {
    lp[s[q]][0] = lev;
    lp[s[q]][1] = nb;
    lp[s[q]][3] = q;
  }
  lp[s[0]][5] = s[1]; lp[s[0]][7] = s[2];
  outer(s[0], 4, 1, 5);
  outer(s[0], 6, 2, 7);
  lp[s[1]][4] = s[0]; lp[s[1]][7] = s[3];
  outer(s[1], 5, 0, 4);
  outer(s[1], 6, 3, 7);
  lp[s[2]][5] = s[3]; lp[s[2]][6] = s[0];
  outer(s[2], 4, 3, 5);
  outer(s[2], 7, 0, 6);
  lp[s[3]][4] = s[2]; lp[s[3]][6] = s[1];
  outer(s[3], 5, 2, 4);
  outer(s[3], 7, 1, 6);
  I.leafs[nb] = 0;
  for (int q = 0; q < 4; q++) {
    I.leafs[s[q]] = 1;
    I.icoarse[s[q]] = 0;
  }
  I.iref[nb] = 0;
  I.icoarse[nb] = 0;
}

// updatelpdown.cpp:5-54.  A refined neighbour always has a son group
// (validate_tree); a neighbour without one means a corrupt tree, reported as
// AMR_ERR_CONFIG before anything is changed.
int unlink_sons(Instance& I, int dad, const int s[4]) {
  auto& lp = I.lp;
  for (int side = 4; side < 8; side++) {
    const int n = lp[dad][side];
    if (n != -1 && !I.leafs[n] && lp[n][2] < kNbRoots) return AMR_ERR_CONFIG;
  }
  I.leafs[dad] = 1;
  for (int q = 0; q < 4; q++) I.leafs[s[q]] = 0;
  // side, first son id offset to reach the adjacent column/row, pointer to fix, step to the second son
  int n = lp[dad][4];
  if (n != -1 && !I.leafs[n]) { n = lp[lp[n][2]][5]; lp[n][5] = dad; lp[lp[n][7]][5] = dad; }
  n = lp[dad][5];
  if (n != -1 && !I.leafs[n]) { n = lp[n][2]; lp[n][4] = dad; lp[lp[n][7]][4] = dad; }
  n = lp[dad][6];
  if (n != -1 && !I.leafs[n]) { n = lp[lp[n][2]][7]; lp[n][7] = dad; lp[lp[n][5]][7] = dad; }
  n = lp[dad][7];
  if (n != -1 && !I.leafs[n]) { n = lp[n][2]; lp[n][6] = dad; lp[lp[n][5]][6] = dad; }
  return 0;
}

// Structural check of a quadtree handed in through the C ABI (amr_set_tree,
// checkpoint load), before anything is changed.  The trees the reference
// builds (basegrid.cpp, updatelpup.cpp, updatelpdown.cpp) satisfy:
//  * blocks 0..3 are the roots: level 0, father 0, sibling id 0;
//  * every other block is son lp[nb][3] of the son group of its father,
//    one level below it (son groups are 4 consecutive ids, allocated once per
//    father and reused, admesh.cpp:33-51), including stale blocks;
//  * inside a son group the sibling pointers are those updatelpup.cpp sets;
//  * neighbour pointers are -1 or a block id, and a refined neighbour of a
//    leaf has a son group;
//  * leaf flags are 0/1 and icoord lies inside the finest-level index range.
// Every index the schedule, the fill compiler and the regrid follow is then
// in range, and every level indexes the per-level cell sizes.
int validate_tree(int n, int maxblocks, int maxlevs, int side, const int32_t* lp, const int32_t* ic,
                  const int32_t* lf) {
  if (!lp || !ic || !lf || n < kNbRoots || n > maxblocks) return AMR_ERR_CONFIG;
  auto L = [&](int nb, int s) { return lp[8 * nb + s]; };
  const int64_t coord_max = (int64_t)side << maxlevs;
  for (int nb = 0; nb < n; nb++) {
    if (lf[nb] != 0 && lf[nb] != 1) return AMR_ERR_CONFIG;
    if (ic[2 * nb] < 0 || ic[2 * nb] >= coord_max || ic[2 * nb + 1] < 0 || ic[2 * nb + 1] >= coord_max)
      return AMR_ERR_CONFIG;
    for (int s = 4; s < 8; s++)
      if (L(nb, s) < -1 || L(nb, s) >= n) return AMR_ERR_CONFIG;
    const int s0 = L(nb, 2);
    if (s0 != 0 && (s0 < kNbRoots || s0 > n - 4)) return AMR_ERR_CONFIG;
    if (nb < kNbRoots) {
      if (L(nb, 0) != 0 || L(nb, 1) != 0 || L(nb, 3) != 0) return AMR_ERR_CONFIG;
      continue;
    }
    const int f = L(nb, 1), q = L(nb, 3);
    if (f < 0 || f >= n || q < 0 || q > 3) return AMR_ERR_CONFIG;
    if (L(f, 2) + q != nb || L(nb, 0) != L(f, 0) + 1 || L(nb, 0) >= maxlevs) return AMR_ERR_CONFIG;
  }
  for (int nb = 0; nb < n; nb++) {
    const int s0 = L(nb, 2);
    if (s0 != 0) {
      // the group's internal pointers (updatelpup.cpp:9-16)
      if (L(s0, 5) != s0 + 1 || L(s0, 7) != s0 + 2 || L(s0 + 1, 4) != s0 || L(s0 + 1, 7) != s0 + 3 ||
          L(s0 + 2, 5) != s0 + 3 || L(s0 + 2, 6) != s0 || L(s0 + 3, 4) != s0 + 2 || L(s0 + 3, 6) != s0 + 1)
        return AMR_ERR_CONFIG;
    }
    if (!lf[nb]) continue;
    for (int s = 4; s < 8; s++) {
      const int m = L(nb, s);
      if (m != -1 && !lf[m] && L(m, 2) == 0) return AMR_ERR_CONFIG;
    }
  }
  return 0;
}

void grow_tree(Instance& I, int n) {
  if ((int)I.lp.size() < n) {
    I.lp.resize(n, std::array<int32_t, 8>{0, 0, 0, 0, 0, 0, 0, 0});
    I.icoord.resize(n, std::array<int32_t, 2>{0, 0});
    I.leafs.resize(n, 0);
    I.iref.resize(n, 0);
    I.icoarse.resize(n, 0);
  }
}

// admesh.cpp:33-58 for one block marked for refinement: its son group (a
// new one at lastActive, or the one it had before), the data-transfer op, the
// sons' coordinates and updatelpup.  `first_ref`: the instance's first op in refs.
int refine_block(amr_ctx* c, int ii, int nb, std::vector<RegridOp>& refs, size_t first_ref) {
  Instance& I = c->inst[ii];
  auto& lp = I.lp;
  const int side = c->NT - 2 * c->ng;
  int s[4];
  if (lp[nb][2] == 0) {
    s[0] = I.lastActive;
    s[1] = s[0] + 1; s[2] = s[0] + 2; s[3] = s[0] + 3;
    lp[nb][2] = s[0];
    I.lastActive += 4;
    if (I.lastActive - 1 >= I.maxblocks) { c->fail_nb = nb; return AMR_ERR_OVERFLOW; }
    grow_tree(I, I.lastActive);
  } else {
    s[0] = lp[nb][2]; s[1] = lp[s[0]][5]; s[2] = lp[s[0]][7]; s[3] = lp[s[2]][5];
  }
  RegridOp op{};
  op.parent_gs = slot_of(c, I, nb);
  for (int q = 0; q < 4; q++) {
    op.son_gs[q] = slot_of(c, I, s[q]);
    if (op.son_gs[q] < 0) { c->fail_nb = s[q]; return AMR_ERR_CAPACITY; }
  }
  op.inst = ii;
  op.nb = nb;
  op.son_nb = s[0];
  op.seq = (int32_t)(refs.size() - first_ref);
  refs.push_back(op);
  // refine.cpp:61-84
  const int x0 = I.icoord[nb][0] * 2, y0 = I.icoord[nb][1] * 2;
  for (int q = 0; q < 4; q++) {
    I.icoord[s[q]][0] = x0 + ((q & 1) ? side : 0);
    I.icoord[s[q]][1] = y0 + ((q & 2) ? side : 0);
  }
  link_sons(I, nb, s);
  return 0;
}

// admesh.cpp:69-92 for one group of four siblings coarsened into their dad
// (first son nb): the data-transfer op and updatelpdown.  `nrefs`: the
// instance's refinements of this regrid (coarsenings are numbered after them).
int coarsen_group(amr_ctx* c, int ii, int nb, size_t nrefs, std::vector<RegridOp>& coars, size_t first_coars) {
  Instance& I = c->inst[ii];
  auto& lp = I.lp;
  const int dad = lp[nb][1];
  int s[4] = {nb, lp[nb][5], lp[nb][7], 0};
  s[3] = lp[s[1]][7];
  RegridOp op{};
  op.parent_gs = slot_of(c, I, dad);
  for (int q = 0; q < 4; q++) op.son_gs[q] = slot_of(c, I, s[q]);
  op.inst = ii;
  op.nb = dad;
  op.son_nb = s[0];
  op.seq = (int32_t)(nrefs + coars.size() - first_coars);
  if (int e = unlink_sons(I, dad, s)) { c->fail_nb = dad; return e; }
  coars.push_back(op);
  return 0;
}

// admesh.cpp:17-102 on the host given the per-leaf indicator
// (criteria.cpp:8-43), with critneighup.cpp / critneighdown.cpp.  Emits the
// data-transfer ops.
int regrid_decide(amr_ctx* c, int ii, const std::vector<double>& grad_of_nb, std::vector<RegridOp>& refs,
                  std::vector<RegridOp>& coars) {
  Instance& I = c->inst[ii];
  auto& lp = I.lp;
  // this instance's ops in the reference's order (error-record ordering)
  const size_t first_ref = refs.size(), first_coars = coars.size();
  for (int nb = 0; nb < I.lastActive; nb++) { I.iref[nb] = 0; I.icoarse[nb] = 0; }
  for (int nb = 0; nb < I.lastActive; nb++) {
    if (!(I.leafs[nb] && lp[nb][0] >= 0)) continue;
    if (grad_of_nb[nb] >= 0.05) {
      if (lp[nb][0] < I.maxlevs - 1) { I.iref[nb] = 1; I.icoarse[nb] = 0; }
    } else {
      I.icoarse[nb] = 1;
    }
  }
  for (int nl = I.maxlevs - 1; nl >= 0; nl--)
    for (int nb = 0; nb < I.lastActive; nb++)
      if (I.iref[nb] && lp[nb][0] == nl)
        for (int s = 4; s < 8; s++) {
          const int n1 = lp[nb][s];
          if (n1 != -1 && lp[n1][0] < lp[nb][0]) I.iref[n1] = 1;
        }
  for (int nl = 0; nl < I.maxlevs - 1; nl++)
    for (int nb = 0; nb < I.lastActive; nb++)
      if (lp[nb][0] == nl && I.iref[nb])
        if (int e = refine_block(c, ii, nb, refs, first_ref)) return e;
  for (int nb = 0; nb < I.lastActive; nb++) {
    if (!I.icoarse[nb]) continue;
    for (int s = 4; s < 8; s++) {
      const int n1 = lp[nb][s];
      if (n1 != -1 && !I.leafs[n1]) I.icoarse[nb] = 0;
    }
  }
  for (int nl = I.maxlevs - 1; nl >= 1; nl--)
    for (int nb = 0; nb < I.lastActive; nb++) {
      if (!(I.icoarse[nb] && lp[nb][3] == 0 && lp[nb][0] == nl)) continue;
      const int s1 = lp[nb][5], s2 = lp[nb][7], s3 = lp[s1][7];
      if (!(I.icoarse[s1] && I.icoarse[s2] && I.icoarse[s3])) continue;
      if (int e = coarsen_group(c, ii, nb, refs.size() - first_ref, coars, first_coars)) return e;
    }
  return 0;
}

// The regrid the device decided (regrid.hip) applied to the host copy of
// instance ii's tree: its refinements, then its coarsenings, in the order the
// device recorded them, with the same surgery as regrid_decide.  The ids come
// from the device: each is checked against the host tree before it is used.
int regrid_replay(amr_ctx* c, int ii, const RegridOut& o, const int32_t* log, std::vector<RegridOp>& refs,
                  std::vector<RegridOp>& coars) {
  Instance& I = c->inst[ii];
  const size_t first_ref = refs.size(), first_coars = coars.size();
  auto bad = [&](const char* what, int nb) {
    std::fprintf(stderr, "amr2d_mhd: device regrid of instance %d: %s (block %d)\n", ii, what, nb);
    return AMR_ERR_HIP;
  };
  for (int k = 0; k < o.nref; k++) {
    const int nb = log[k];
    if (nb < 0 || nb >= I.lastActive || !I.leafs[nb] || I.lp[nb][0] >= I.maxlevs - 1)
      return bad("refinement of a block that cannot refine", nb);
    if (int e = refine_block(c, ii, nb, refs, first_ref)) return e;
  }
  for (int k = 0; k < o.ncoars; k++) {
    const int nb = log[o.nref + k];
    if (nb < kNbRoots || nb + 3 >= I.lastActive || I.lp[nb][3] != 0 || I.lp[nb][5] != nb + 1 ||
        I.lp[nb][7] != nb + 2 || I.lp[nb + 1][7] != nb + 3)
      return bad("coarsening of a block that is not a first son", nb);
    if (int e = coarsen_group(c, ii, nb, refs.size() - first_ref, coars, first_coars)) return e;
  }
  return 0;
}

// coarsen.cpp:14-148 / coarsenBi_CT.cpp:14-306 replayed once per tile size:
// for each dad cell the last write's son and lower-left source cell.
void coarsen_maps(int NT, int ng, std::vector<int32_t>& mU, std::vector<int32_t>& mB) {
  const int N2 = NT * NT, lo = ng, lob = lo - 1, hib = NT - ng, h2 = (NT - 1) / 2;
  mU.assign(N2, -1);
  mB.assign(2 * N2, -1);
  auto pk = [](int s, int i1, int j1) { return s | (i1 << 8) | (j1 << 16); };
  auto wU = [&](int i, int j, int s, int i1, int j1) { mU[j * NT + i] = pk(s, i1, j1); };
  auto wB = [&](int k, int i, int j, int s, int i1, int j1) { mB[k * N2 + j * NT + i] = pk(s, i1, j1); };
  // interior (both fields)
  for (int i = lo; i <= h2; i++)
    for (int j = lo; j <= h2; j++) {
      const int i1 = 2 * i - lo, j1 = 2 * j - lo, i2 = h2 + i - 1 - (ng - 2), j2 = h2 + j - 1 - (ng - 2);
      wU(i, j, 0, i1, j1); wU(i2, j, 1, i1, j1); wU(i, j2, 2, i1, j1); wU(i2, j2, 3, i1, j1);
      for (int k = 0; k < 2; k++) {
        wB(k, i, j, 0, i1, j1); wB(k, i2, j, 1, i1, j1); wB(k, i, j2, 2, i1, j1); wB(k, i2, j2, 3, i1, j1);
      }
    }
  // cell field ghost ring (coarsen.cpp:33-148)
  for (int i = 0; i < ng / 2; i++)
    for (int j = 0; j < ng / 2; j++) {
      wU(lob - i, lob - j, 0, lob - 2 * i - 1, lob - 2 * j - 1);
      wU(lob - i, hib + j, 2, lob - 2 * i - 1, hib + 2 * j);
    }
  for (int i = 0; i <= lob; i++)
    for (int j = lo; j <= h2; j++) {
      int i1 = lob - 2 * i - 1;
      if (i1 < 0) i1 = 0;
      wU(lob - i, j, 0, i1, 2 * j - lo);
      wU(lob - i, j + h2 - 1 - (ng - 2), 2, i1, 2 * j - lo);
    }
  for (int i = 0; i < ng / 2; i++)
    for (int j = 0; j < ng / 2; j++) {
      wU(hib + i, lob - j, 1, hib + 2 * i, lob - 2 * j - 1);
      wU(hib + i, hib + j, 3, hib + 2 * i, hib + 2 * j);
    }
  for (int i = 0; i < ng; i++)
    for (int j = lo; j <= h2; j++) {
      int i1 = hib + 2 * i;
      if (i1 + 1 > NT - 1) i1 = NT - 2;
      wU(hib + i, j, 1, i1, 2 * j - lo);
      wU(hib + i, j + h2 - 1 - (ng - 2), 3, i1, 2 * j - lo);
    }
  for (int j = 0; j < ng / 2; j++)
    for (int i = 0; i < ng / 2; i++) {
      wU(lob - i, lob - j, 0, lob - 2 * i - 1, lob - 2 * j - 1);
      wU(hib + i, lob - j, 1, hib + 2 * i, lob - 2 * j - 1);
    }
  for (int i = lo; i <= h2; i++)
    for (int j = 0; j <= lob; j++) {
      int j1 = lob - 2 * j - 1;
      if (j1 < 0) j1 = 0;
      wU(i, lob - j, 0, 2 * i - lo, j1);
      wU(i + h2 - 1 - (ng - 2), lob - j, 1, 2 * i - lo, j1);
    }
  for (int j = 0; j < ng / 2; j++)
    for (int i = 0; i < ng / 2; i++) {
      wU(lob - i, hib + j, 2, lob - 2 * i - 1, hib + 2 * j);
      wU(hib + i, hib + j, 3, hib + 2 * i, hib + 2 * j);
    }
  for (int i = lo; i <= h2; i++)
    for (int j = 0; j < ng; j++) {
      int j1 = hib + 2 * j;
      if (j1 + 1 > NT - 1) j1 = NT - 2;
      wU(i, hib + j, 2, 2 * i - lo, j1);
      wU(i + h2 - 1 - (ng - 2), hib + j, 3, 2 * i - lo, j1);
    }
  // face field ghost ring (coarsenBi_CT.cpp:189-306): same cells and clamps;
  // the pairing direction is applied on the device by component
  for (int k = 0; k < 2; k++) {
    for (int i = 0; i < ng / 2; i++)
      for (int j = 0; j < ng / 2; j++) {
        wB(k, lob - i, lob - j, 0, lob - 2 * i - 1, lob - 2 * j - 1);
        wB(k, lob - i, hib + j, 2, lob - 2 * i - 1, hib + 2 * j);
      }
    for (int i = 0; i <= lob; i++)
      for (int j = lo; j <= h2; j++) {
        int i1 = lob - 2 * i - 1;
        if (i1 < 0) i1 = 0;
        wB(k, lob - i, j, 0, i1, 2 * j - lo);
        wB(k, lob - i, j + h2 - 1 - (ng - 2), 2, i1, 2 * j - lo);
      }
    for (int i = 0; i < ng / 2; i++)
      for (int j = 0; j < ng / 2; j++) {
        wB(k, hib + i, lob - j, 1, hib + 2 * i, lob - 2 * j - 1);
        wB(k, hib + i, hib + j, 3, hib + 2 * i, hib + 2 * j);
      }
    for (int i = 0; i < ng; i++)
      for (int j = lo; j <= h2; j++) {
        int i1 = hib + 2 * i;
        if (i1 + 1 > NT - 1) i1 = NT - 2;
        wB(k, hib + i, j, 1, i1, 2 * j - lo);
        wB(k, hib + i, j + h2 - 1 - (ng - 2), 3, i1, 2 * j - lo);
      }
    for (int j = 0; j < ng / 2; j++)
      for (int i = 0; i < ng / 2; i++) {
        wB(k, lob - i, lob - j, 0, lob - 2 * i - 1, lob - 2 * j - 1);
        wB(k, hib + i, lob - j, 1, hib + 2 * i, lob - 2 * j - 1);
      }
    for (int i = lo; i <= h2; i++)
      for (int j = 0; j <= lob; j++) {
        int j1 = lob - 2 * j - 1;
        if (j1 < 0) j1 = 0;
        wB(k, i, lob - j, 0, 2 * i - lo, j1);
        wB(k, i + h2 - 1 - (ng - 2), lob - j, 1, 2 * i - lo, j1);
      }
    for (int j = 0; j < ng / 2; j++)
      for (int i = 0; i < ng / 2; i++) {
        wB(k, lob - i, hib + j, 2, lob - 2 * i - 1, hib + 2 * j);
        wB(k, hib + i, hib + j, 3, hib + 2 * i, hib + 2 * j);
      }
    for (int i = lo; i <= h2; i++)
      for (int j = 0; j < ng; j++) {
        int j1 = hib + 2 * j;
        if (j1 + 1 > NT - 1) j1 = NT - 2;
        wB(k, i, hib + j, 2, 2 * i - lo, j1);
        wB(k, i + h2 - 1 - (ng - 2), hib + j, 3, 2 * i - lo, j1);
      }
  }
}

}  // namespace

// ====================================================================
// context operations
// ====================================================================
namespace {

void fill_params(amr_ctx* c) {
  c->h_ip.reserve(c->inst.size());
  c->h_ip.n = c->inst.size();
  for (size_t q = 0; q < c->inst.size(); q++) {
    const Instance& I = c->inst[q];
    InstParams& P = c->h_ip.data()[q];
    P.gamma = I.gamma;
    P.dt = I.dt_dom;
    for (int l = 0; l < kMaxLev; l++) P.dx[l] = I.dx[l];
    P.limiter = I.lim;
    P.solver = I.solver;
    P.maxlevs = I.maxlevs;
    P.rk = c->rk;
    P.ct_type = I.ct_type;
    P.pad = 0;
  }
}

int upload_params(amr_ctx* c) {
  fill_params(c);
  HIPCHK(hipMemcpyAsync(c->d_ip, c->h_ip.data(), sizeof(InstParams) * c->h_ip.size(), hipMemcpyHostToDevice,
                        c->stream));
  c->dma_since_direct = true;
  return 0;
}

// the tree a fill program depends on (FillProg::key)
void program_key(const Instance& I, std::vector<int32_t>& key, uint64_t& hash) {
  const int n = I.lastActive;
  key.resize(1 + (size_t)n * 10);
  int32_t* k = key.data();
  *k++ = n;
  for (int nb = 0; nb < n; nb++) {
    std::memcpy(k, I.lp[nb].data(), sizeof(int32_t) * 8);
    k[8] = I.leafs[nb];
    k[9] = nb < (int)I.gs.size() ? I.gs[nb] : -1;
    k += 10;
  }
  // four independent multiply-xorshift chains (a single chain is bound by the
  // multiply latency), then combined
  const size_t m = key.size();
  uint64_t h[4] = {0x9e3779b97f4a7c15ull ^ m, 0xc2b2ae3d27d4eb4full ^ m, 0x165667b19e3779f9ull ^ m,
                   0x27d4eb2f165667c5ull ^ m};
  size_t i = 0;
  for (; i + 4 <= m; i += 4)
    for (int l = 0; l < 4; l++) {
      h[l] ^= (uint32_t)key[i + l];
      h[l] *= 0xff51afd7ed558ccdull;
      h[l] ^= h[l] >> 32;
    }
  for (; i < m; i++) {
    h[0] ^= (uint32_t)key[i];
    h[0] *= 0xff51afd7ed558ccdull;
    h[0] ^= h[0] >> 32;
  }
  uint64_t r = h[0];
  for (int l = 1; l < 4; l++) {
    r ^= h[l] + 0x9e3779b97f4a7c15ull + (r << 6) + (r >> 2);
    r *= 0xff51afd7ed558ccdull;
    r ^= r >> 32;
  }
  hash = r;
}

void swap_program(Instance& I, FillProg& P) {
  std::swap(I.d_ops, P.d_ops);
  std::swap(I.cap_ops, P.cap_ops);
  std::swap(I.nops, P.nops);
  std::swap(I.d_gen, P.d_gen);
  std::swap(I.cap_gen, P.cap_gen);
  std::swap(I.ngen, P.ngen);
  std::swap(I.op_kinds, P.kinds);
  std::swap(I.tail_chunk0, P.tail_chunk0);
  std::swap(I.prog_key, P.key);
  std::swap(I.prog_hash, P.hash);
}

// Makes the program of I's current tree current without compiling when that
// tree is the one of the current program or of a cached one (the device copy
// travels with it; the stream is idle here, so no launch still reads a
// buffer that changes role).  Otherwise the current program is cached (the
// least recently used entry gives up its device buffers for the new
// program), the key is set and false is returned: compile into I.ops/I.gen
// and upload.
bool reuse_program(amr_ctx* c, Instance& I) {
  std::vector<int32_t> key;
  uint64_t hash = 0;
  program_key(I, key, hash);
  if (!I.prog_key.empty() && hash == I.prog_hash && key == I.prog_key) return true;
  if (c->prog_cache_max <= 0) {
    I.prog_key.swap(key);
    I.prog_hash = hash;
    return false;
  }
  for (auto& P : I.prog_cache)
    if (P.hash == hash && !P.key.empty() && P.key == key) {
      swap_program(I, P);  // P now holds the previous program
      P.used = ++c->prog_clock;
      return true;
    }
  FillProg* slot = nullptr;
  if ((int)I.prog_cache.size() < c->prog_cache_max) {
    I.prog_cache.emplace_back();
    slot = &I.prog_cache.back();
  } else {
    slot = &I.prog_cache[0];
    for (auto& P : I.prog_cache)
      if (P.used < slot->used) slot = &P;
  }
  swap_program(I, *slot);  // the previous program is cached, its buffers are reused
  slot->used = ++c->prog_clock;
  I.prog_key.swap(key);
  I.prog_hash = hash;
  return false;
}

// leaf work list and fill program of the running instances
// launch geometry of direct launches: the grids are the work
void direct_geometry(amr_ctx* c) {
  c->grid_leaves = c->nleaf;
  c->grid_chunks = c->nchunks;
  c->launch_counts = nullptr;
}

// A graph's capacity for a count n: the current one while n fits and fills
// more than a quarter of it, else the next power of two (at most hard_max).
int graph_capacity(int cur, int n, int hard_max) {
  n = std::max(n, 1);
  if (cur >= n && cur / 4 < n) return cur;
  int g = 1;
  while (g < n) g <<= 1;
  return std::max(n, std::min(g, hard_max));
}

// A specialised stage kernel serves launches whose running instances all use
// HLLD with a MUSCL-family limiter, or all HLLD with WENO5 (24x24 tiles).
int stage_spec_of(const amr_ctx* c) {
  bool hlld_muscl = true, hlld_weno5 = c->NT == 24 && c->ng == 4;
  for (const Instance& I : c->inst) {
    if (I.done) continue;
    if (!(I.solver == kFluxHLLD && I.lim < kLimWENO3)) hlld_muscl = false;
    if (!(I.solver == kFluxHLLD && I.lim == kLimWENO5)) hlld_weno5 = false;
  }
  return hlld_muscl ? kStageHLLDMuscl : hlld_weno5 ? kStageHLLDWeno5 : kStageGeneric;
}

// stage specialisation an instance's leaves can run with on their own
int stage_class_of(const amr_ctx* c, const Instance& I) {
  if (I.solver == kFluxHLLD && I.lim < kLimWENO3) return kStageHLLDMuscl;
  if (I.solver == kFluxHLLD && I.lim == kLimWENO5 && c->NT == 24 && c->ng == 4) return kStageHLLDWeno5;
  if (I.solver == kFluxHLLC && I.lim < kLimWENO3) return kStageHLLCMuscl;
  return kStageGeneric;
}

// when the running instances fall into more than one class, order the leaf
// list by class for per-class stage launches (each leaf's stage is the same
// whichever launch runs it)
void plan_stage_split(amr_ctx* c) {
  c->stage_split = false;
  c->split_n = 0;
  if (c->stage_spec != kStageGeneric) return;
  std::vector<int> cls(c->inst.size(), -1);
  int present = 0;
  for (size_t q = 0; q < c->inst.size(); q++) {
    if (c->inst[q].done) continue;
    cls[q] = stage_class_of(c, c->inst[q]);
    present |= 1 << cls[q];
  }
  if (__builtin_popcount(present) < 2) return;
  static const int order[kStageSpecs] = {kStageHLLDMuscl, kStageHLLDWeno5, kStageHLLCMuscl, kStageGeneric};
  c->h_split_leaves.clear();
  for (int k = 0; k < kStageSpecs; k++) {
    if (!(present & (1 << order[k]))) continue;
    c->split_spec[c->split_n] = order[k];
    c->split_off[c->split_n] = (int)c->h_split_leaves.size();
    for (const LeafItem& L : c->h_leaves)
      if (cls[L.inst] == order[k]) c->h_split_leaves.push_back(L);
    c->split_n++;
  }
  c->split_off[c->split_n] = (int)c->h_split_leaves.size();
  c->stage_split = true;
}

// The device copy of instance q's tree refreshed from the host tree: rows
// and leaf flags of its blocks, zero beyond lastActive (blocks a regrid may
// allocate), work flags cleared.
int upload_tree(amr_ctx* c, int q) {
  Instance& I = c->inst[q];
  const size_t off = (size_t)c->tree_off[q], n = (size_t)I.lastActive, mb = (size_t)I.maxblocks;
  HIPCHK(hipMemcpyAsync(c->dtree.lp + 8 * off, I.lp.data(), sizeof(int32_t) * 8 * n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemsetAsync(c->dtree.lp + 8 * (off + n), 0, sizeof(int32_t) * 8 * (mb - n), c->stream));
  HIPCHK(hipMemcpyAsync(c->dtree.leafs + off, I.leafs.data(), n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemsetAsync(c->dtree.leafs + off + n, 0, mb - n, c->stream));
  HIPCHK(hipMemsetAsync(c->dtree.iref + off, 0, mb, c->stream));
  HIPCHK(hipMemsetAsync(c->dtree.icoarse + off, 0, mb, c->stream));
  HIPCHK(hipMemsetAsync(c->dtree.mark + off, 0, mb, c->stream));
  I.dev_tree_dirty = false;
  return 0;
}

int rebuild_topology(amr_ctx* c) {
  HostTimed ht(c, PF_HOST_TOPO);
  HIPCHK(hipStreamSynchronize(c->stream));  // previous uploads consumed
  c->h_leaves.clear();
  c->leaf_inst_begin.assign(c->inst.size() + 1, 0);
  std::vector<int> dirty;
  for (size_t q = 0; q < c->inst.size(); q++) {
    Instance& I = c->inst[q];
    c->leaf_inst_begin[q] = (int)c->h_leaves.size();
    if (I.done) continue;
    for (int nb = 0; nb < I.lastActive; nb++)
      if (I.leafs[nb]) c->h_leaves.push_back(LeafItem{I.gs[nb], (int32_t)q, nb, I.lp[nb][0]});
    if (I.prog_dirty) dirty.push_back((int)q);
  }
  c->leaf_inst_begin[c->inst.size()] = (int)c->h_leaves.size();
  c->nleaf = (int)c->h_leaves.size();
  {
    c->stage_spec = stage_spec_of(c);
    c->tune = c->tune_by_spec[c->stage_spec];
    plan_stage_split(c);
  }
  // programs of instances whose tree changed: the current one if the tree is
  // the same, a cached one if the tree recurs, else compiled (in parallel)
  {
    std::vector<int> todo;
    for (int qi : dirty)
      if (!reuse_program(c, c->inst[qi])) todo.push_back(qi);
      else c->inst[qi].prog_dirty = false;
    dirty.swap(todo);
  }
  const auto t_compile = std::chrono::steady_clock::now();
  const int hw = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  int nthreads = (int)std::min<size_t>(dirty.size(), hw);
  if (nthreads < 1) nthreads = 1;
  if ((int)c->scratch.size() < nthreads) c->scratch.resize(nthreads);
  std::vector<int> status(dirty.size(), 0);
  // several instances: one thread each; a single one: its schedule split
  // over the threads (large trees only)
  const int inner = dirty.size() == 1 ? hw : 1;
  auto work = [&](int t) {
    for (size_t q = t; q < dirty.size(); q += nthreads) {
      Instance& I = c->inst[dirty[q]];
      build_schedule(I);
      status[q] = compile_instance(I, c->NT, c->ng, c->scratch[t], inner);
    }
  };
  if (nthreads == 1 || dirty.size() < 4) {
    work(0);
    if (nthreads > 1)
      for (int t = 1; t < nthreads; t++) work(t);
  } else {
    std::vector<std::thread> th;
    for (int t = 1; t < nthreads; t++) th.emplace_back(work, t);
    work(0);
    for (auto& x : th) x.join();
  }
  if (!dirty.empty() && (c->prof_mask & (1u << PF_HOST_COMPILE))) {
    c->prof[PF_HOST_COMPILE].ms +=
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_compile).count();
    c->prof[PF_HOST_COMPILE].launches += (int64_t)dirty.size();
  }
  bool failed = false;
  for (size_t q = 0; q < dirty.size(); q++)
    if (status[q]) {
      fail_instance(c, dirty[q], status[q], -1, 0, 0);
      failed = true;
    }
  if (failed) return rebuild_topology(c);  // drop the failed instances' leaves
  for (int qi : dirty) {
    Instance& I = c->inst[qi];
    const int n = (int)I.ops.size(), ng = (int)I.gen.size();
    I.nops = n;
    I.op_kinds[3] = n - I.op_kinds[0] - I.op_kinds[1] - I.op_kinds[2];  // padding
    I.ngen = ng;
    if (int e = ensure(I.d_ops, I.cap_ops, std::max(1, n))) return e;
    if (int e = ensure(I.d_gen, I.cap_gen, std::max(1, ng))) return e;
    I.prog_dirty = false;
  }
  for (int qi : dirty) {
    Instance& I = c->inst[qi];
    const size_t n = I.ops.size(), ng = I.gen.size();
    if (n) HIPCHK(hipMemcpyAsync(I.d_ops, I.ops.data(), sizeof(FillOp) * n, hipMemcpyHostToDevice, c->stream));
    if (ng) HIPCHK(hipMemcpyAsync(I.d_gen, I.gen.data(), sizeof(FillGen) * ng, hipMemcpyHostToDevice, c->stream));
  }
  // segment tables over the running instances and the leaf work list: one
  // upload of [program pointers | explicit-operand pointers | chunk prefix |
  // leaves] from the pinned staging copy
  const int ni = (int)c->inst.size();
  auto align = [](size_t x) { return (x + 255) & ~size_t(255); };
  // [counts | program pointers | explicit-operand pointers | chunk prefix | leaves]
  const size_t off_ops = 256;
  const size_t off_gen = off_ops + align(sizeof(FillOp*) * (ni + 1));
  const size_t off_pre = off_gen + align(sizeof(FillGen*) * (ni + 1));
  const size_t off_tops = off_pre + align(sizeof(int32_t) * (ni + 1));
  const size_t off_tpre = off_tops + align(sizeof(FillOp*) * (ni + 1));
  const size_t off_rgi = off_tpre + align(sizeof(int32_t) * (ni + 1));
  const size_t off_leaf = off_rgi + align(sizeof(RegridInst) * ni);
  const size_t off_split = off_leaf + align(sizeof(LeafItem) * std::max(1, c->nleaf));
  const size_t topo_bytes = off_split + sizeof(LeafItem) * std::max(1, c->nleaf);
  if (topo_bytes > c->cap_topo) {
    const size_t cap = topo_bytes + topo_bytes / 2;
    hipFree(c->d_topo);
    if (c->h_topo) hipHostFree(c->h_topo);
    c->d_topo = nullptr;
    c->h_topo = nullptr;
    c->cap_topo = 0;
    HIPCHK(hipMalloc((void**)&c->d_topo, cap));
    HIPCHK(hipHostMalloc((void**)&c->h_topo, cap));
    c->cap_topo = cap;
    c->topo_epoch++;
  }
  const FillOp** h_ops = (const FillOp**)(c->h_topo + off_ops);
  const FillGen** h_gen = (const FillGen**)(c->h_topo + off_gen);
  int32_t* h_pre = (int32_t*)(c->h_topo + off_pre);
  const FillOp** h_tops = (const FillOp**)(c->h_topo + off_tops);
  int32_t* h_tpre = (int32_t*)(c->h_topo + off_tpre);
  int total = 0, ttotal = 0;
  h_pre[0] = 0;
  h_tpre[0] = 0;
  for (int q = 0; q < ni; q++) {
    const Instance& I = c->inst[q];
    h_ops[q] = I.d_ops;
    h_gen[q] = I.d_gen;
    total += I.done ? 0 : I.nops / kFillChunk;
    h_pre[q + 1] = total;
    const int t0 = std::min(std::max(I.tail_chunk0, 0), I.nops / kFillChunk);
    h_tops[q] = I.d_ops + (size_t)t0 * kFillChunk;
    ttotal += I.done ? 0 : I.nops / kFillChunk - t0;
    h_tpre[q + 1] = ttotal;
  }
  c->nchunks = total;
  c->ntail_chunks = ttotal;
  // the device regrid's per-instance parameters; device trees that are stale
  // are refreshed from the host trees first
  RegridInst* h_rgi = (RegridInst*)(c->h_topo + off_rgi);
  c->rg_on.assign(ni, 0);
  bool tree_up = false;
  for (int q = 0; q < ni; q++) {
    Instance& I = c->inst[q];
    const bool on = c->regrid_device && !I.done && I.maxlevs > 1;
    // every block a regrid can reach: the current ones and the new son groups
    // of leaves that can refine and have none yet
    int fresh = 0;
    if (on)
      for (int nb = 0; nb < I.lastActive; nb++) fresh += I.leafs[nb] && I.lp[nb][0] < I.maxlevs - 1 && I.lp[nb][2] == 0;
    const int64_t reach = std::min<int64_t>(I.maxblocks, (int64_t)I.lastActive + 4 * (int64_t)fresh);
    h_rgi[q] = RegridInst{c->tree_off[q], I.maxblocks, I.maxlevs, I.lastActive, c->leaf_inst_begin[q],
                          c->leaf_inst_begin[q + 1], on ? 1 : 0,
                          (c->regrid_lds && reach <= kRgLdsRows) ? (int32_t)reach : 0};
    c->rg_on[q] = on;
    if (on && I.dev_tree_dirty) {
      if (int e = upload_tree(c, q)) return e;
      tree_up = true;
    }
  }
  if (tree_up) HIPCHK(hipStreamSynchronize(c->stream));  // the host trees may change after this call
  ((int32_t*)c->h_topo)[kCountLeaves] = c->nleaf;
  ((int32_t*)c->h_topo)[kCountChunks] = c->nchunks;
  if (c->nleaf) std::memcpy(c->h_topo + off_leaf, c->h_leaves.data(), sizeof(LeafItem) * c->nleaf);
  size_t up = off_leaf + sizeof(LeafItem) * c->nleaf;
  if (c->stage_split) {
    std::memcpy(c->h_topo + off_split, c->h_split_leaves.data(), sizeof(LeafItem) * c->h_split_leaves.size());
    up = off_split + sizeof(LeafItem) * c->h_split_leaves.size();
  }
  HIPCHK(hipMemcpyAsync(c->d_topo, c->h_topo, up, hipMemcpyHostToDevice, c->stream));
  c->d_split_leaves = (LeafItem*)(c->d_topo + off_split);
  c->d_counts = (const int32_t*)c->d_topo;
  c->d_inst_ops = (const FillOp**)(c->d_topo + off_ops);
  c->d_inst_gen = (const FillGen**)(c->d_topo + off_gen);
  c->d_prefix = (int32_t*)(c->d_topo + off_pre);
  c->d_tail_ops = (const FillOp**)(c->d_topo + off_tops);
  c->d_tail_prefix = (int32_t*)(c->d_topo + off_tpre);
  c->d_rgi = (RegridInst*)(c->d_topo + off_rgi);
  c->d_leaves = (LeafItem*)(c->d_topo + off_leaf);
  c->topo_dirty = false;
  c->steps_same_topo = 0;
  direct_geometry(c);
  return 0;
}

FillSet make_set(amr_ctx* c, std::initializer_list<std::pair<int, int>> fk) {
  FillSet s{};
  for (auto& p : fk) {
    s.base[s.n] = c->F.f[p.first];
    s.nv[s.n] = kFieldVars[p.first];
    s.kind[s.n] = p.second;
    s.n++;
  }
  return s;
}

void run_fill(amr_ctx* c, const FillSet& s) {
  Timed tm(c, PF_FILL);
  launch_fill(c->NT, c->d_inst_ops, c->d_inst_gen, c->d_prefix, (int)c->inst.size(), c->grid_chunks, s, c->stream,
              c->launch_counts);
}

int tune_bucket(int nleaf) {
  int b = 0;
  while ((2 << b) <= nleaf && b < 23) b++;
  return b;
}

// variant of the next stage launch; sets *timed when the launch is a trial
// (trials run only inside full steps, whose spans are collected)
// The strip variants split a leaf over several workgroups, so they cannot run
// a stage that reads and writes the same field (SSP-RK3's second stage reads
// Us and writes Us: one strip's update would race with another's loads of
// its halo rows).  The tuner leaves them out of RK3 contexts, and a fixed
// strip variant runs such a stage as the whole-leaf 768-lane variant.
bool strip_variant(int v) { return v == 6 || v == 7; }
bool strips_allowed(const amr_ctx* c) { return c->rk != 3; }
int stage_variant_for(const amr_ctx* c, int v, int step) {
  return strip_variant(v) && c->rk == 3 && step == 2 ? 2 : v;
}

int stage_choice(amr_ctx* c, bool* timed, bool allow_trial = true) {
  *timed = false;
  if (c->stage_variant >= 0) return c->stage_variant;
  const auto& B = c->tune[tune_bucket(c->nleaf)];
  if (B.chosen >= 0) return B.chosen;
  if (!allow_trial) return 0;
  // round robin over the variants (the one with the fewest trials), so that
  // every variant's trials fall on about the same leaf counts
  int next = -1;
  for (int v = 0; v < kStageVariants; v++)
    if ((strips_allowed(c) || !strip_variant(v)) && B.n[v] < kTuneTrials && (next < 0 || B.n[v] < B.n[next]))
      next = v;
  if (next >= 0) {
    *timed = true;
    return next;
  }
  return 0;
}

bool tuning(amr_ctx* c) {
  if (c->stage_split) return false;  // split launches make no trials
  bool t;
  stage_choice(c, &t);
  return t;
}

void run_stage(amr_ctx* c, int step, bool allow_trial = true) {
  if (c->stage_split) {
    // one launch per class of instances, each with its specialisation
    // (the HLLC one exists for kSplitVariant only); no tuner trials
    const int v = stage_variant_for(c, c->stage_variant >= 0 ? c->stage_variant : kSplitVariant, step);
    c->span_variant[step - 1] = -1;
    for (int k = 0; k < c->split_n; k++) {
      const int n = c->split_off[k + 1] - c->split_off[k];
      if (n <= 0) continue;
      Timed tm(c, PF_STAGE, true);
      launch_stage(c->NT, c->ng, v, c->split_spec[k], c->F, c->d_split_leaves + c->split_off[k], n, c->d_ip, step,
                   c->d_err, c->stream, tm.start(), tm.stop(), nullptr, nullptr);
    }
    return;
  }
  Timed tm(c, PF_STAGE, true);
  bool timed;
  const int v = stage_variant_for(c, stage_choice(c, &timed, allow_trial), step);
  unsigned long long* span = timed ? c->d_span + 2 * (step - 1) : nullptr;
  c->span_variant[step - 1] = timed ? v : -1;
  if (timed) {
    // the trial counts as pending until the step's results come back
    c->tune[tune_bucket(c->nleaf)].n[v] += 1;
  }
  launch_stage(c->NT, c->ng, v, c->stage_spec, c->F, c->d_leaves, c->grid_leaves, c->d_ip, step, c->d_err, c->stream,
               tm.start(), tm.stop(), span, c->launch_counts);
}

// fold the spans of this step's trial launches into the tuner
void tune_collect(amr_ctx* c) {
  auto& B = c->tune[tune_bucket(c->nleaf)];
  for (int q = 0; q < 2; q++) {
    const int v = c->span_variant[q];
    if (v < 0) continue;
    c->span_variant[q] = -1;
    const unsigned long long t0 = ~c->h_span[2 * q], t1 = c->h_span[2 * q + 1];
    if (t1 > t0 && c->wall_khz > 0) {
      const double us = (double)(t1 - t0) / (c->wall_khz * 1e-3) / c->nleaf;
      B.us_per_leaf[v] += us;
      if (B.nval[v] < kTuneTrials) B.trial_us[v][B.nval[v]++] = us;
    }
  }
  bool all = true;
  for (int v = 0; v < kStageVariants; v++) all &= B.n[v] >= kTuneTrials || (!strips_allowed(c) && strip_variant(v));
  if (all && B.chosen < 0) {
    // the variant with the smallest median trial (one slow or fast outlier
    // does not decide)
    int best = -1;
    double best_us = 0.;
    for (int v = 0; v < kStageVariants; v++) {
      if (!B.nval[v]) continue;
      double t[kTuneTrials];
      std::copy(B.trial_us[v], B.trial_us[v] + B.nval[v], t);
      std::sort(t, t + B.nval[v]);
      const double med = t[(B.nval[v] - 1) / 2];
      if (best < 0 || med < best_us) {
        best = v;
        best_us = med;
      }
    }
    B.chosen = best < 0 ? 0 : best;
  }
}

// CT2D.cpp:11-124
void run_ct(amr_ctx* c, int step) {
  {
    Timed tm(c, PF_CT);
    launch_emf(c->NT, c->F, c->d_leaves, c->grid_leaves, c->d_ip, step, c->rk, c->d_err, c->stream, c->launch_counts);
  }
  run_fill(c, make_set(c, {{kEMF, kFillEdge}}));
  {
    Timed tm(c, PF_CT);
    launch_bupdate(c->NT, c->F, c->d_leaves, c->grid_leaves, c->d_ip, step, c->rk, c->stream, c->launch_counts);
  }
  run_fill(c, make_set(c, {{step == c->rk ? kBi : kBis, kFillFace}}));
  {
    Timed tm(c, PF_CT);
    launch_bsync(c->NT, c->F, c->d_leaves, c->grid_leaves, step == c->rk, c->stream, c->launch_counts);
  }
}

// waits for the stream (the error records already queued for download) and
// retires the instances whose kernels reported an error; returns nonzero only
// for a HIP failure
int wait_errors(amr_ctx* c) {
  const int ni = (int)c->inst.size();
  {
    HostTimed ht(c, PF_HOST_WAIT);
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  flush_profile(c);
  if (*c->h_fail) {
    // a grid-barrier wait of the step kernel ran out (its workgroups were not
    // all resident): the step's results are void; later steps use per-phase
    // launches
    std::fprintf(stderr, "amr2d_mhd: step kernel barrier timed out; using per-phase launches from now on\n");
    c->step_broken = true;
    *c->h_fail = 0;
    HIPCHK(hipMemsetAsync(c->d_fail, 0, sizeof(int32_t), c->stream));
    HIPCHK(hipMemsetAsync(c->d_sync, 0, sizeof(GridSync), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return set_error(c, AMR_ERR_HIP, -1, -1, 0, 0);
  }
  bool any = false;
  for (int q = 0; q < ni; q++) {
    const DevError& e = c->h_err[q];
    if (!e.nkey) continue;
    any = true;
    // error_key fields: code, block id, j, i
    const unsigned long long key = ~e.nkey;
    fail_instance(c, q, (int)(key & 0xf), (int)((key >> 4) & 0xffff), (int)((key >> 32) & 0xff),
                  (int)((key >> 24) & 0xff));
  }
  if (any) {
    HIPCHK(hipMemsetAsync(c->d_err, 0, sizeof(DevError) * ni, c->stream));
    c->dma_since_direct = true;
  }
  return 0;
}

int sync_errors(amr_ctx* c) {
  // the error records and the barrier-failure flag after them
  HIPCHK(hipMemcpyAsync(c->h_rec, c->d_rec, c->rec_err_bytes, hipMemcpyDeviceToHost, c->stream));
  return wait_errors(c);
}

// regrid data transfer in reference order: all refinements, then all
// coarsenings; ops that touch a slot an earlier op wrote run in a later batch
int apply_regrid_ops(amr_ctx* c, const std::vector<RegridOp>& ops, bool refine) {
  size_t b = 0;
  while (b < ops.size()) {
    std::vector<int32_t> written, read;
    size_t e = b;
    for (; e < ops.size(); e++) {
      const RegridOp& o = ops[e];
      std::vector<int32_t> r, w;
      if (refine) { r = {o.parent_gs}; w = {o.son_gs[0], o.son_gs[1], o.son_gs[2], o.son_gs[3]}; }
      else { r = {o.son_gs[0], o.son_gs[1], o.son_gs[2], o.son_gs[3], o.parent_gs}; w = {o.parent_gs}; }
      bool clash = false;
      for (auto x : r) clash |= std::find(written.begin(), written.end(), x) != written.end();
      for (auto x : w) clash |= std::find(written.begin(), written.end(), x) != written.end() ||
                               std::find(read.begin(), read.end(), x) != read.end();
      if (clash && e > b) break;
      written.insert(written.end(), w.begin(), w.end());
      read.insert(read.end(), r.begin(), r.end());
    }
    const int n = (int)(e - b);
    if (int er = ensure(c->d_rops, c->cap_rops, n)) return er;
    HIPCHK(hipMemcpyAsync(c->d_rops, ops.data() + b, sizeof(RegridOp) * n, hipMemcpyHostToDevice, c->stream));
    {
      Timed tm(c, PF_REGRID);
      if (refine) launch_refine(c->NT, c->ng, c->F, c->d_rops, n, c->d_ip, c->d_err, c->stream);
      else launch_coarsen(c->NT, c->ng, c->F, c->d_rops, n, c->d_cmapU, c->d_cmapB, c->d_ip, c->d_err, c->stream);
    }
    b = e;
  }
  return 0;
}

// The end of a step's device work: per-leaf CFL bounds and indicators
// (primitives from dW for the initial regrid, else from U), then, with the
// device regrid, the regrid decision and the CFL minimum per instance.
// `grid`: leaf-kernel grid (the leaf count, or a captured graph's capacity).
void launch_cfl_regrid(amr_ctx* c, const double* dW, int grid) {
  {
    Timed tm(c, PF_CFL);
    launch_cfl_criteria(c->NT, c->ng, c->F, dW, c->d_leaves, grid, c->d_ip, c->d_red, c->d_err, c->stream,
                        c->launch_counts);
  }
  if (c->regrid_device) {
    Timed tm(c, PF_REGRID);
    launch_regrid(c->d_rgi, (int)c->inst.size(), c->d_leaves, c->d_red, c->dtree, c->d_rgo, c->d_oplog, c->stream);
  }
}

// The step's one download: the error records, then the regrid records and op
// log (device regrid) or the per-leaf CFL bounds and indicators (host regrid).
int download_results(amr_ctx* c, int nleaf, bool regrid_records) {
  if (regrid_records) {
    HIPCHK(hipMemcpyAsync(c->h_rec, c->d_rec, c->rec_off_log + sizeof(int32_t) * nleaf, hipMemcpyDeviceToHost,
                          c->stream));
  } else {
    HIPCHK(hipMemcpyAsync(c->h_rec, c->d_rec, c->rec_err_bytes, hipMemcpyDeviceToHost, c->stream));
    if (nleaf > 0)
      HIPCHK(hipMemcpyAsync(c->h_red, c->d_red, sizeof(double) * 2 * nleaf, hipMemcpyDeviceToHost, c->stream));
  }
  return 0;
}

// per-leaf CFL bounds and indicators of the running instances -> host
int cfl_and_indicators(amr_ctx* c, const double* dW) {
  {
    Timed tm(c, PF_CFL);
    launch_cfl_criteria(c->NT, c->ng, c->F, dW, c->d_leaves, c->nleaf, c->d_ip, c->d_red, c->d_err,
                        c->stream);
  }
  if (int e = download_results(c, c->nleaf, false)) return e;
  return wait_errors(c);
}

// the same with the device regrid: CFL bounds, indicators, the regrid
// decision and the per-instance CFL minimum, downloaded and waited for
int cfl_and_regrid(amr_ctx* c, const double* dW) {
  launch_cfl_regrid(c, dW, c->nleaf);
  if (int e = download_results(c, c->nleaf, true)) return e;
  return wait_errors(c);
}

// admesh for every running instance from the downloaded indicators
int regrid_all(amr_ctx* c) {
  std::vector<RegridOp>& refs = c->h_refs;
  std::vector<RegridOp>& coars = c->h_coars;
  HIPCHK(hipStreamSynchronize(c->stream));  // previous ops consumed
  refs.clear();
  coars.clear();
  HostTimed ht(c, PF_HOST_REGRID);
  for (size_t q = 0; q < c->inst.size(); q++) {
    Instance& I = c->inst[q];
    if (I.done || I.maxlevs <= 1) continue;
    std::vector<double> grad(I.lastActive, 0.0);
    for (int l = c->leaf_inst_begin[q]; l < c->leaf_inst_begin[q + 1]; l++)
      grad[c->h_leaves[l].nb] = c->h_red[2 * l + 1];
    const int before = I.lastActive;
    const size_t nr = refs.size(), nc = coars.size();
    if (int e = regrid_decide(c, (int)q, grad, refs, coars)) {
      if (e == AMR_ERR_HIP) return e;
      fail_instance(c, (int)q, e, c->fail_nb, 0, 0);
      continue;
    }
    if (refs.size() != nr || coars.size() != nc || I.lastActive != before) {
      c->topo_dirty = true;
      I.prog_dirty = true;
    }
  }
  c->regrid_launched = !refs.empty() || !coars.empty();
  if (int e = apply_regrid_ops(c, refs, true)) return e;
  if (int e = apply_regrid_ops(c, coars, false)) return e;
  return 0;
}

// admesh for every running instance as the device decided it (downloaded
// records and op log): the host trees follow, the data-transfer ops run.
// An instance that failed in this step, or whose replay stops at an error,
// keeps a host tree the device copy no longer matches: its copy is marked
// stale.
int regrid_replay_all(amr_ctx* c) {
  std::vector<RegridOp>& refs = c->h_refs;
  std::vector<RegridOp>& coars = c->h_coars;
  HIPCHK(hipStreamSynchronize(c->stream));  // previous ops consumed
  refs.clear();
  coars.clear();
  HostTimed ht(c, PF_HOST_REGRID);
  // a device result the host trees do not bear out: an internal error; every
  // device copy is refreshed from the host trees before it is used again
  auto internal = [&]() {
    for (auto& J : c->inst) J.dev_tree_dirty = true;
    c->topo_dirty = true;
    return AMR_ERR_HIP;
  };
  for (size_t q = 0; q < c->inst.size(); q++) {
    Instance& I = c->inst[q];
    if (q >= c->rg_on.size() || !c->rg_on[q]) continue;
    if (I.done) {
      I.dev_tree_dirty = true;
      continue;
    }
    const RegridOut& o = c->h_rgo[q];
    const int lb = c->leaf_inst_begin[q], nl = c->leaf_inst_begin[q + 1] - lb;
    if (o.nref < 0 || o.ncoars < 0 || o.nref + o.ncoars > nl || (o.status != 0 && o.status != kErrOverflow)) {
      std::fprintf(stderr, "amr2d_mhd: device regrid of instance %d: %d + %d ops for %d leaves, status %d\n", (int)q,
                   o.nref, o.ncoars, nl, o.status);
      return internal();
    }
    const int before = I.lastActive;
    const size_t nr = refs.size(), nc = coars.size();
    const int e = regrid_replay(c, (int)q, o, c->h_oplog + lb, refs, coars);
    if (e == AMR_ERR_HIP) return internal();
    if ((!e || e == AMR_ERR_OVERFLOW) &&
        (I.lastActive != o.last_active || (e == AMR_ERR_OVERFLOW) != (o.status == kErrOverflow))) {
      std::fprintf(stderr, "amr2d_mhd: device regrid of instance %d: lastActive %d (host %d), status %d (host %d)\n",
                   (int)q, o.last_active, I.lastActive, o.status, e);
      return internal();
    }
    if (e) {
      I.dev_tree_dirty = true;
      fail_instance(c, (int)q, e, c->fail_nb, 0, 0);
      continue;
    }
    if (refs.size() != nr || coars.size() != nc || I.lastActive != before) {
      c->topo_dirty = true;
      I.prog_dirty = true;
    }
  }
  c->regrid_launched = !refs.empty() || !coars.empty();
  if (int e = apply_regrid_ops(c, refs, true)) return e;
  if (int e = apply_regrid_ops(c, coars, false)) return e;
  return 0;
}

// The phases of one step (main.cpp:128-230) for the persistent step kernel,
// in the order of the per-phase launches of enqueue_step below.
void build_phases(amr_ctx* c) {
  c->phases.clear();
  auto fill = [&](std::initializer_list<std::pair<int, int>> fk) {
    StepPhase p{};
    p.kind = kPhaseFill;
    for (auto& x : fk) {
      p.field[p.nf] = (int8_t)x.first;
      p.fkind[p.nf] = (int8_t)x.second;
      p.nf++;
    }
    c->phases.push_back(p);
  };
  auto op = [&](int kind, int step) {
    StepPhase p{};
    p.kind = (int8_t)kind;
    p.step = (int8_t)step;
    c->phases.push_back(p);
  };
  fill({{kU, kFillCell}, {kBi, kFillFace}});
  for (int st = 1; st <= c->rk; st++) {
    if (st > 1) fill({{kUs, kFillCell}});
    op(kPhaseStage, st);
    if (st == 1) fill({{kBi, kFillFace}, {kFF, kFillFace}, {kGG, kFillFace}});
    else fill({{kFF, kFillFace}, {kGG, kFillFace}});
    // CT2D.cpp:11-124
    op(kPhaseEmf, st);
    fill({{kEMF, kFillEdge}});
    op(kPhaseBupdate, st);
    fill({{st == c->rk ? kBi : kBis, kFillFace}});
    op(kPhaseBsync, st == c->rk ? 1 : 0);
  }
  fill({{kU, kFillCell}});
  op(kPhaseCfl, 0);
}

// whether this step runs as one persistent kernel
bool persistent_step(amr_ctx* c) {
  if (c->step_mode == 0 || c->step_broken || c->in_group || c->nleaf <= 0) return false;
  const int lanes = c->step_mode == 2 ? 1 : 0;
  int& g = c->step_grid[c->stage_spec][lanes];
  if (g < 0) g = step_kernel_grid(c->NT, c->ng, c->stage_spec, lanes);
  return g > 0;
}

// The device work of one step as one cooperative launch of the persistent
// step kernel, then the download of the CFL bounds, indicators and error
// records (and, when kernel families are timed, the phase stamps).
int enqueue_step_persistent(amr_ctx* c) {
  if (c->phases.empty()) build_phases(c);
  const int lanes = c->step_mode == 2 ? 1 : 0;
  StepArgs A{};
  A.F = c->F;
  A.leaves = c->d_leaves;
  A.nleaf = c->nleaf;
  A.nchunks = c->nchunks;
  A.ops = c->d_inst_ops;
  A.gen = c->d_inst_gen;
  A.cprefix = c->d_prefix;
  A.ninst = (int)c->inst.size();
  static_assert(sizeof(InstParams) % sizeof(double) == 0, "parameters copied as 8-byte words");
  A.copy_words = (int32_t)(sizeof(InstParams) * c->inst.size() / sizeof(double));
  A.copy_src = (const double*)c->h_ip.data();
  A.copy_dst = (double*)c->d_ip;
  A.ip = c->d_ip;
  A.lam_grad = c->d_red;
  A.err = c->d_err;
  A.sync = c->d_sync;
  A.fail = c->d_fail;
  const bool stamps = (c->prof_mask & 0xfu) != 0 && c->wall_khz > 0;
  A.stamps = stamps ? c->d_stamps : nullptr;
  A.nphases = (int)c->phases.size();
  for (int p = 0; p < A.nphases; p++) A.phase[p] = c->phases[p];
  {
    Timed tm(c, PF_STEP, true);
    if (launch_step(c->NT, c->ng, c->stage_spec, lanes, c->step_grid[c->stage_spec][lanes], A, c->stream,
                    tm.start(), tm.stop())) {
      std::fprintf(stderr, "amr2d_mhd: cooperative launch of the step kernel failed: %s\n",
                   hipGetErrorString(hipGetLastError()));
      return AMR_ERR_HIP;
    }
  }
  if (c->regrid_device) {
    Timed tm(c, PF_REGRID);
    launch_regrid(c->d_rgi, (int)c->inst.size(), c->d_leaves, c->d_red, c->dtree, c->d_rgo, c->d_oplog, c->stream);
  }
  if (int e = download_results(c, c->nleaf, c->regrid_device)) return e;
  if (stamps) {
    HIPCHK(hipMemcpyAsync(c->h_stamps, c->d_stamps, sizeof(unsigned long long) * (A.nphases + 1),
                          hipMemcpyDeviceToHost, c->stream));
    c->stamps_pending = A.nphases;
  }
  return 0;
}

// main.cpp:128-247 for all running instances
// Device work of one step (main.cpp:128-230) up to the download of the CFL
// bounds, indicators and error records: stream-ordered launches only, so it
// can be captured into a graph.
int enqueue_step(amr_ctx* c) {
  const bool trial = tuning(c);
  if (trial) HIPCHK(hipMemsetAsync(c->d_span, 0, 4 * sizeof(unsigned long long), c->stream));
  {
    // the step's parameters (dt) ride on the first fill launch
    FillSet s0 = make_set(c, {{kU, kFillCell}, {kBi, kFillFace}});
    static_assert(sizeof(InstParams) % sizeof(double) == 0, "parameters copied as 8-byte words");
    // U and Bi as the previous step's last fills of the same programs left
    // them (fill_fresh): only the programs' linked tails change anything
    const bool tail = c->tail_first_fill;
    const int chunks = tail ? c->ntail_chunks : c->grid_chunks;
    if (chunks > 0) {
      s0.copy_words = (int32_t)(sizeof(InstParams) * c->inst.size() / sizeof(double));
      s0.copy_src = (const double*)c->h_ip.data();
      s0.copy_dst = (double*)c->d_ip;
    } else {
      HIPCHK(hipMemcpyAsync(c->d_ip, c->h_ip.data(), sizeof(InstParams) * c->inst.size(), hipMemcpyHostToDevice,
                            c->stream));
    }
    if (tail) {
      Timed tm(c, PF_FILL);
      launch_fill(c->NT, c->d_tail_ops, c->d_inst_gen, c->d_tail_prefix, (int)c->inst.size(), chunks, s0, c->stream,
                  nullptr);
    } else {
      run_fill(c, s0);
    }
  }
  run_stage(c, 1);
  run_fill(c, make_set(c, {{kBi, kFillFace}, {kFF, kFillFace}, {kGG, kFillFace}}));
  run_ct(c, 1);
  run_fill(c, make_set(c, {{kUs, kFillCell}}));
  run_stage(c, 2);
  run_fill(c, make_set(c, {{kFF, kFillFace}, {kGG, kFillFace}}));
  run_ct(c, 2);
  if (c->rk == 3) {
    run_fill(c, make_set(c, {{kUs, kFillCell}}));
    run_stage(c, 3);
    run_fill(c, make_set(c, {{kFF, kFillFace}, {kGG, kFillFace}}));
    run_ct(c, 3);
  }
  run_fill(c, make_set(c, {{kU, kFillCell}}));
  launch_cfl_regrid(c, nullptr, c->grid_leaves);
  if (int e = download_results(c, c->grid_leaves, c->regrid_device)) return e;
  if (trial)
    HIPCHK(hipMemcpyAsync(c->h_span, c->d_span, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
  return 0;
}

// (Re)capture the step into the context's graph: update the executable graph
// in place when the captured topology allows it, else instantiate anew.
int capture_step(amr_ctx* c) {
  HostTimed ht(c, PF_CAPTURE);
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  const int e = enqueue_step(c);
  const hipError_t ce = hipStreamEndCapture(c->stream, &g);
  if (e || ce != hipSuccess) {
    if (g) hipGraphDestroy(g);
    if (!e) std::fprintf(stderr, "amr2d_mhd: graph capture failed: %s\n", hipGetErrorString(ce));
    return e ? e : AMR_ERR_HIP;
  }
  // a re-capture changes grids, the stage kernel's variant or the buffers:
  // instantiate anew rather than update the executable graph in place
  if (c->gexec) {
    hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
  }
  {
    const hipError_t ie = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
      hipGraphDestroy(g);
      std::fprintf(stderr, "amr2d_mhd: graph instantiate failed: %s\n", hipGetErrorString(ie));
      return AMR_ERR_HIP;
    }
  }
  hipGraphDestroy(g);
  c->graph_dirty = false;
  return 0;
}

// Returns nonzero for a HIP failure, or when no instance is left running
// and one failed during this call (its code); otherwise instance failures
// are reported through amr_info().status.
// First half of a step: rebuild the work lists and fill programs if the
// trees changed, then launch the step's device work (graph or direct) on the
// context's stream and return without waiting.
int step_begin(amr_ctx* c) {
  bool fresh = c->fill_fresh;
  c->fill_fresh = false;
  c->step_direct = false;
  c->call_fail = 0;
  c->in_flight = false;
  if (!running(c)) return 0;
  if (c->topo_dirty) {
    fresh = false;
    if (int e = rebuild_topology(c)) return e;
  }
  fill_params(c);
  const bool device_timing = (c->prof_mask & 0x1fu) != 0;
  if (persistent_step(c)) {
    c->step_trial = false;
    if (int e = enqueue_step_persistent(c)) return e;
    c->steps_same_topo++;
    c->in_flight = true;
    return 0;
  }
  c->step_trial = tuning(c);
  if (c->graph_on && !device_timing && !c->step_trial && !c->dma_since_direct) {
    // the captured step serves this tree when its grids cover the counts
    bool trial;
    amr_ctx::GraphKey k;
    k.epoch = c->topo_epoch;
    k.leaves = graph_capacity(c->gkey.leaves, c->nleaf, (int)std::min<int64_t>(c->cap_slots, INT32_MAX));
    k.chunks = graph_capacity(c->gkey.chunks, c->nchunks, INT32_MAX);
    k.variant = stage_choice(c, &trial, false);
    k.spec = c->stage_spec;
    c->grid_leaves = k.leaves;
    c->grid_chunks = k.chunks;
    c->launch_counts = c->d_counts;
    if (c->graph_dirty || !c->gexec || !(k == c->gkey)) {
      const int e = capture_step(c);
      if (e) {
        direct_geometry(c);
        return e;
      }
      c->gkey = k;
    }
    const hipError_t le = hipGraphLaunch(c->gexec, c->stream);
    direct_geometry(c);
    HIPCHK(le);
    if (c->prof_mask & (1u << PF_GRAPH)) c->prof[PF_GRAPH].launches += 1;
  } else {
    c->tail_first_fill = fresh;
    const int e = enqueue_step(c);
    c->tail_first_fill = false;
    if (e) return e;
    c->step_direct = true;
    c->dma_since_direct = false;
  }
  c->steps_same_topo++;
  c->in_flight = true;
  return 0;
}

// Second half: wait for the step, advance time and dt (main.cpp:231-238),
// regrid (admesh) and retire finished instances.
int step_end(amr_ctx* c, int* n_active) {
  c->fill_fresh = false;
  if (!c->in_flight) {
    if (n_active) *n_active = running(c);
    return running(c) ? 0 : c->diag.code;
  }
  c->in_flight = false;
  if (int e = wait_errors(c)) return e;
  if (c->step_trial) tune_collect(c);
  int active = 0;
  for (size_t q = 0; q < c->inst.size(); q++) {
    Instance& I = c->inst[q];
    if (I.done) continue;
    double lambda = 1.E6;
    if (c->regrid_device)
      lambda = c->h_rgo[q].lambda;
    else
      for (int l = c->leaf_inst_begin[q]; l < c->leaf_inst_begin[q + 1]; l++) lambda = std_min(lambda, c->h_red[2 * l]);
    I.t = I.t + I.dt_main;
    I.dt_dom = I.dt_main;
    I.count = I.count + 1;
    I.dt_main = I.cfl * lambda;
    if (I.dt_main < 0 || std::isnan(I.dt_main)) fail_instance(c, (int)q, AMR_ERR_DT, -1, 0, 0);
  }
  if (int e = c->regrid_device ? regrid_replay_all(c) : regrid_all(c)) return e;
  if (c->regrid_launched)
    if (int e = sync_errors(c)) return e;  // refine/coarsen errors
  for (auto& I : c->inst) {
    if (!I.done && !(I.t <= I.tEnd)) {
      I.done = true;
      c->topo_dirty = true;
    }
    active += I.done ? 0 : 1;
  }
  c->fill_fresh = c->step_direct && !c->topo_dirty && !c->regrid_launched && !c->call_fail;
  if (n_active) *n_active = active;
  return (active == 0 && c->call_fail) ? c->call_fail : 0;
}

// every entry point other than the step calls and the read-only queries:
// the next step runs its first fill whole
inline void fill_touched(amr_ctx* c) {
  if (c) c->fill_fresh = false;
}

int step_all(amr_ctx* c, int* n_active) {
  if (!running(c)) {
    c->call_fail = 0;
    if (n_active) *n_active = 0;
    return c->diag.code;
  }
  if (int e = step_begin(c)) return e;
  return step_end(c, n_active);
}

// leaf-cell count of each running instance (zone-updates of its next step)
int64_t running_cells(amr_ctx* c, std::vector<int64_t>& leaves) {
  const int64_t cells = (int64_t)c->nxi * c->nxi;
  leaves.assign(c->inst.size(), 0);
  int64_t total = 0;
  for (size_t q = 0; q < c->inst.size(); q++) {
    const Instance& I = c->inst[q];
    if (I.done) continue;
    for (int nb = 0; nb < I.lastActive; nb++) leaves[q] += I.leafs[nb];
    leaves[q] *= cells;
    total += leaves[q];
  }
  return total;
}

// zone-updates of a completed step: the instances that did not fail in it
int64_t completed_cells(amr_ctx* c, const std::vector<int64_t>& leaves) {
  int64_t zu = 0;
  for (size_t q = 0; q < c->inst.size(); q++)
    if (!c->inst[q].status) zu += leaves[q];
  return zu;
}

// end of a phase-level call: the first instance failure of the call, if any
int finish_call(amr_ctx* c) {
  if (int e = sync_errors(c)) return e;
  return c->call_fail;
}

int check_inst(amr_ctx* c, int inst) { return (c && inst >= 0 && inst < (int)c->inst.size()) ? 0 : AMR_ERR_CONFIG; }

}  // namespace

// ====================================================================
// C ABI
// ====================================================================
extern "C" {

int amr_create(int n_inst, const amr_sim_config* cfgs, int64_t slot_capacity, amr_ctx** out) {
  AMR_TRY
  if (!out || n_inst <= 0 || !cfgs) return AMR_ERR_CONFIG;
  *out = nullptr;
  amr_ctx* c = new amr_ctx();
  c->nxi = cfgs[0].nx;
  c->ng = cfgs[0].nghosts;
  c->NT = c->nxi + 2 * c->ng;
  if (!tile_supported(c->NT, c->ng)) { delete c; return AMR_ERR_CONFIG; }
  int64_t total_mb = 0;
  c->inst.resize(n_inst);
  for (int q = 0; q < n_inst; q++) {
    const amr_sim_config& g = cfgs[q];
    Instance& I = c->inst[q];
    I.cfg = g;
    I.lim = lim_code(g.limiter);
    I.solver = flux_code(g.flux);
    I.ic = ic_code(g.ic);
    I.maxlevs = g.maxlevs;
    I.cfl = g.cfl;
    if (g.nx != c->nxi || g.nghosts != c->ng || I.lim < 0 || I.solver < 0 || I.ic < 0 || g.maxlevs < 1 ||
        g.maxlevs > 7 || ((I.lim == kLimWENO5) && c->ng < 3) || g.ct_type < 0 || g.ct_type > 2) {
      delete c;
      return AMR_ERR_CONFIG;
    }
    I.maxblocks = maxblocks_of(g.maxlevs);
    total_mb += I.maxblocks;
    // meshblock.cpp:90-96; a periodic domain needs its exact length, so it
    // uses lenx / (sqrt(nbroots) * nx) instead
    I.periodic = g.periodic != 0;
    I.ct_type = g.ct_type == 2 ? 2 : 1;
    I.dx[0] = g.lenx / (std::sqrt((double)kNbRoots) * (c->NT - 2 * c->ng - (I.periodic ? 0 : 1)));
    for (int l = 1; l < kMaxLev; l++) I.dx[l] = I.dx[l - 1] / 2.;
    // IC2Dtype.cpp:18-80 (gamma, tEnd per problem)
    if (I.ic == IC_ROTOR || I.ic == IC_ROTOR_BS) { I.gamma = 1.4; I.tEnd = 0.295; }
    else if (I.ic == IC_ROTOR_P) { I.gamma = g.ic_params[6]; I.tEnd = g.ic_params[7]; }
    else if (I.ic == IC_OT) { I.gamma = 5. / 3.; I.tEnd = 0.5; }
    else { I.gamma = 2.0; I.tEnd = 0.2; }
    // all instances of a context share the RK scheme (one step sequence)
    const int rk = g.rk_stages == 0 ? 2 : g.rk_stages;
    if ((rk != 2 && rk != 3) || (q > 0 && rk != c->rk)) {
      delete c;
      return AMR_ERR_CONFIG;
    }
    c->rk = rk;
  }
  c->cap_slots = slot_capacity > 0 ? slot_capacity : total_mb;
  const size_t N2 = (size_t)c->NT * c->NT;
  for (int f = 0; f < kNumFields; f++) {
    const size_t bytes = (size_t)c->cap_slots * kFieldVars[f] * N2 * sizeof(double);
    if (hipMalloc((void**)&c->F.f[f], bytes) != hipSuccess || hipMemset(c->F.f[f], 0, bytes) != hipSuccess) {
      amr_destroy(c);
      return AMR_ERR_HIP;
    }
  }
  // error records, then the step kernel's barrier-failure flag
  c->rec_err_bytes = (sizeof(DevError) * n_inst + sizeof(int32_t) + 255) & ~size_t(255);
  // [error records | regrid records | op log | per-leaf CFL bounds and indicators]
  const size_t cap_leaves = (size_t)std::max<int64_t>(1, c->cap_slots);
  c->rec_off_log = c->rec_err_bytes + ((sizeof(RegridOut) * n_inst + 255) & ~size_t(255));
  c->rec_off_red = c->rec_off_log + ((sizeof(int32_t) * cap_leaves + 255) & ~size_t(255));
  const size_t rec_bytes = c->rec_off_red + sizeof(double) * 2 * cap_leaves;
  c->tree_off.assign(n_inst + 1, 0);
  for (int q = 0; q < n_inst; q++) c->tree_off[q + 1] = c->tree_off[q] + c->inst[q].maxblocks;
  const size_t tree_blocks = (size_t)c->tree_off[n_inst];
  if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc((void**)&c->d_ip, sizeof(InstParams) * n_inst) != hipSuccess ||
      hipMalloc((void**)&c->d_rec, rec_bytes) != hipSuccess ||
      hipMemset(c->d_rec, 0, rec_bytes) != hipSuccess ||
      hipMalloc((void**)&c->dtree.lp, sizeof(int32_t) * 8 * tree_blocks) != hipSuccess ||
      hipMemset(c->dtree.lp, 0, sizeof(int32_t) * 8 * tree_blocks) != hipSuccess ||
      hipMalloc((void**)&c->dtree.leafs, 4 * tree_blocks) != hipSuccess ||
      hipMemset(c->dtree.leafs, 0, 4 * tree_blocks) != hipSuccess ||
      hipHostMalloc((void**)&c->h_rec, rec_bytes) != hipSuccess ||
      hipMalloc((void**)&c->d_span, 4 * sizeof(unsigned long long)) != hipSuccess ||
      hipHostMalloc((void**)&c->h_span, 4 * sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc((void**)&c->d_sync, sizeof(GridSync)) != hipSuccess ||
      hipMemset(c->d_sync, 0, sizeof(GridSync)) != hipSuccess ||
      hipMalloc((void**)&c->d_stamps, sizeof(unsigned long long) * (kMaxPhases + 1)) != hipSuccess ||
      hipHostMalloc((void**)&c->h_stamps, sizeof(unsigned long long) * (kMaxPhases + 1)) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    amr_destroy(c);
    return AMR_ERR_HIP;
  }
  c->stream = c->own_stream;
  std::memset(c->h_rec, 0, rec_bytes);
  c->d_err = (DevError*)c->d_rec;
  c->h_err = (DevError*)c->h_rec;
  c->d_red = (double*)(c->d_rec + c->rec_off_red);
  c->d_fail = (int32_t*)(c->d_rec + sizeof(DevError) * n_inst);
  c->h_fail = (int32_t*)(c->h_rec + sizeof(DevError) * n_inst);
  c->h_red = (double*)(c->h_rec + c->rec_off_red);
  c->d_rgo = (RegridOut*)(c->d_rec + c->rec_err_bytes);
  c->h_rgo = (RegridOut*)(c->h_rec + c->rec_err_bytes);
  c->d_oplog = (int32_t*)(c->d_rec + c->rec_off_log);
  c->h_oplog = (int32_t*)(c->h_rec + c->rec_off_log);
  c->dtree.iref = c->dtree.leafs + tree_blocks;
  c->dtree.icoarse = c->dtree.iref + tree_blocks;
  c->dtree.mark = c->dtree.icoarse + tree_blocks;
  {
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess)
      c->wall_khz = khz;
    if (c->wall_khz <= 0) c->stage_variant = 0;  // no device clock: no tuning
  }
  std::vector<int32_t> mU, mB;
  coarsen_maps(c->NT, c->ng, mU, mB);
  if (hipMalloc((void**)&c->d_cmapU, sizeof(int32_t) * mU.size()) != hipSuccess ||
      hipMalloc((void**)&c->d_cmapB, sizeof(int32_t) * mB.size()) != hipSuccess ||
      hipMemcpy(c->d_cmapU, mU.data(), sizeof(int32_t) * mU.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->d_cmapB, mB.data(), sizeof(int32_t) * mB.size(), hipMemcpyHostToDevice) != hipSuccess) {
    amr_destroy(c);
    return AMR_ERR_HIP;
  }
  *out = c;
  return AMR_OK;
  AMR_CATCH
}

int amr_destroy(amr_ctx* c) {
  AMR_TRY
  if (!c) return AMR_OK;
  if (c->stream) hipStreamSynchronize(c->stream);
  if (c->own_stream) hipStreamSynchronize(c->own_stream);
  for (int f = 0; f < kNumFields; f++)
    if (c->F.f[f]) hipFree(c->F.f[f]);
  hipFree(c->d_topo);
  if (c->h_topo) hipHostFree(c->h_topo);
  for (auto& I : c->inst) {
    hipFree(I.d_ops);
    hipFree(I.d_gen);
    for (auto& P : I.prog_cache) {
      hipFree(P.d_ops);
      hipFree(P.d_gen);
    }
  }
  hipFree(c->d_ip); hipFree(c->d_rec); hipFree(c->d_rops);
  hipFree(c->dtree.lp); hipFree(c->dtree.leafs);
  hipFree(c->d_cmapU); hipFree(c->d_cmapB); hipFree(c->d_W);
  if (c->h_rec) hipHostFree(c->h_rec);
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  for (auto& e : c->ev_pool) {
    hipEventDestroy(e[0]);
    hipEventDestroy(e[1]);
  }
  if (c->gexec) hipGraphExecDestroy(c->gexec);
  hipFree(c->d_span);
  hipFree(c->d_sync);
  hipFree(c->d_stamps);
  if (c->h_stamps) hipHostFree(c->h_stamps);
  if (c->h_span) hipHostFree(c->h_span);
  if (c->own_stream) hipStreamDestroy(c->own_stream);
  delete c;
  return AMR_OK;
  AMR_CATCH
}

// basegrid.cpp:4-89 + IC2Dtype.cpp + main.cpp:97-110
int amr_init(amr_ctx* c) {
  AMR_TRY
  fill_touched(c);
  if (!c) return AMR_ERR_CONFIG;
  c->call_fail = 0;
  static const int32_t root_lp[4][8] = {{0, 0, 0, 0, -1, 1, -1, 2}, {0, 0, 0, 0, 0, -1, -1, 3},
                                        {0, 0, 0, 0, -1, 3, 0, -1}, {0, 0, 0, 0, 2, -1, 1, -1}};
  // periodic roots: the outer neighbour pointers wrap instead of -1
  static const int32_t root_lpp[4][8] = {{0, 0, 0, 0, 1, 1, 2, 2}, {0, 0, 0, 0, 0, 0, 3, 3},
                                         {0, 0, 0, 0, 3, 3, 0, 0}, {0, 0, 0, 0, 2, 2, 1, 1}};
  const int NT = c->NT, N2 = NT * NT, side = NT - 2 * c->ng;
  std::vector<double> Wall;
  for (size_t q = 0; q < c->inst.size(); q++) {
    Instance& I = c->inst[q];
    grow_tree(I, kNbRoots);
    for (int r = 0; r < kNbRoots; r++) {
      for (int s = 0; s < 8; s++) I.lp[r][s] = I.periodic ? root_lpp[r][s] : root_lp[r][s];
      I.icoord[r] = {(r & 1) ? side : 0, (r & 2) ? side : 0};
      I.leafs[r] = 1;
      I.iref[r] = 1;
      if (slot_of(c, I, r) < 0) return set_error(c, AMR_ERR_CAPACITY, (int)q, r, 0, 0);
    }
    I.lastActive = kNbRoots;
    I.dev_tree_dirty = true;
    I.t = 0; I.dt_main = 0; I.dt_dom = 0; I.count = 0; I.done = false;
    for (int r = 0; r < kNbRoots; r++) {
      std::vector<double> W, B, U;
      initial_tile(I, r, NT, c->ng, W, B);
      prim_to_cons_tile(W, U, N2, I.gamma);
      HIPCHK(hipMemcpy(c->F.f[kU] + (size_t)I.gs[r] * 8 * N2, U.data(), sizeof(double) * 8 * N2, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(c->F.f[kBi] + (size_t)I.gs[r] * 2 * N2, B.data(), sizeof(double) * 2 * N2, hipMemcpyHostToDevice));
      Wall.insert(Wall.end(), W.begin(), W.end());  // leaf order = instance, then nb
    }
  }
  c->topo_dirty = true;
  if (int e = rebuild_topology(c)) return e;
  if (int e = upload_params(c)) return e;
  // the initial regrid reads the IC primitives directly (main.cpp:110)
  if (c->d_W) hipFree(c->d_W);
  HIPCHK(hipMalloc((void**)&c->d_W, sizeof(double) * Wall.size()));
  HIPCHK(hipMemcpy(c->d_W, Wall.data(), sizeof(double) * Wall.size(), hipMemcpyHostToDevice));
  c->dma_since_direct = true;
  if (int e = c->regrid_device ? cfl_and_regrid(c, c->d_W) : cfl_and_indicators(c, c->d_W)) return e;
  hipFree(c->d_W);
  c->d_W = nullptr;
  if (int e = c->regrid_device ? regrid_replay_all(c) : regrid_all(c)) return e;
  c->topo_dirty = true;
  if (int e = sync_errors(c)) return e;
  return running(c) ? AMR_OK : c->call_fail;
  AMR_CATCH
}

int amr_step(amr_ctx* c, int* n_active) {
  AMR_TRY
  if (!c) return AMR_ERR_CONFIG;
  return step_all(c, n_active);
  AMR_CATCH
}

int amr_run(amr_ctx* c, int max_steps, int64_t* steps, int64_t* zone_updates) {
  AMR_TRY
  if (!c) return AMR_ERR_CONFIG;
  int64_t n = 0, zu = 0;
  int active = running(c);
  std::vector<int64_t> leaves;
  while (active > 0 && (max_steps < 0 || n < max_steps)) {
    running_cells(c, leaves);
    const int e = step_all(c, &active);
    // a step counts for the instances that completed it
    zu += completed_cells(c, leaves);
    if (e) {
      if (steps) *steps = n;
      if (zone_updates) *zone_updates = zu;
      return e;
    }
    n++;
  }
  if (steps) *steps = n;
  if (zone_updates) *zone_updates = zu;
  return AMR_OK;
  AMR_CATCH
}

int amr_step_begin(amr_ctx* c) {
  AMR_TRY
  if (!c || c->in_flight) return AMR_ERR_CONFIG;
  return step_begin(c);
  AMR_CATCH
}

int amr_step_end(amr_ctx* c, int* n_active) {
  AMR_TRY
  if (!c) return AMR_ERR_CONFIG;
  return step_end(c, n_active);
  AMR_CATCH
}

int amr_run_group(int n_ctx, amr_ctx** ctxs, int max_steps, int64_t* steps, int64_t* zone_updates) {
  AMR_TRY
  if (n_ctx <= 0 || !ctxs) return AMR_ERR_CONFIG;
  for (int k = 0; k < n_ctx; k++)
    if (!ctxs[k] || ctxs[k]->in_flight) return AMR_ERR_CONFIG;
  // one context runs its steps as it would alone
  struct GroupScope {
    int n;
    amr_ctx** c;
    GroupScope(int n_, amr_ctx** c_) : n(n_), c(c_) {
      for (int k = 0; k < n && n > 1; k++) c[k]->in_group = true;
    }
    ~GroupScope() {
      for (int k = 0; k < n; k++) c[k]->in_group = false;
    }
  } group_scope(n_ctx, ctxs);
  std::vector<std::vector<int64_t>> leaves(n_ctx);
  std::vector<int64_t> done(n_ctx, 0);
  int64_t zu = 0;
  int err = 0;
  auto launch = [&](int k) {
    amr_ctx* c = ctxs[k];
    if (!running(c) || (max_steps >= 0 && done[k] >= max_steps)) return 0;
    running_cells(c, leaves[k]);
    return step_begin(c);
  };
  // a context stops at its own end or error (like amr_run); the others go on
  // unless the runtime itself failed
  bool hip_failed = false;
  auto note = [&](int e) {
    if (e && !err) err = e;
    if (e == AMR_ERR_HIP) hip_failed = true;
  };
  for (int k = 0; k < n_ctx && !hip_failed; k++) note(launch(k));
  // round robin: while one context's host work (wait, regrid, rebuild,
  // launch) runs, the steps already launched on the others keep the device busy
  bool any = true;
  while (any) {
    any = false;
    for (int k = 0; k < n_ctx; k++) {
      amr_ctx* c = ctxs[k];
      if (!c->in_flight) continue;
      int active = 0;
      const int e = step_end(c, &active);
      zu += completed_cells(c, leaves[k]);
      done[k]++;
      note(e);
      if (!e && !hip_failed) note(launch(k));
      any = any || c->in_flight;
    }
  }
  int64_t n = 0;
  for (int k = 0; k < n_ctx; k++) n = std::max(n, done[k]);
  if (steps) *steps = n;
  if (zone_updates) *zone_updates = zu;
  return err;
  AMR_CATCH
}

int amr_info(amr_ctx* c, int inst, amr_inst_info* o) {
  AMR_TRY
  if (check_inst(c, inst) || !o) return AMR_ERR_CONFIG;
  const Instance& I = c->inst[inst];
  int nl = 0;
  for (int nb = 0; nb < I.lastActive; nb++) nl += I.leafs[nb];
  *o = amr_inst_info{I.t, I.dt_dom, I.dt_main, I.gamma, I.tEnd, I.lastActive, nl, I.count, I.done ? 1 : 0,
                     I.status, I.maxblocks};
  return AMR_OK;
  AMR_CATCH
}

int amr_last_error(amr_ctx* c, amr_diag* d) {
  AMR_TRY
  if (!c || !d) return AMR_ERR_CONFIG;
  *d = c->diag;
  return AMR_OK;
  AMR_CATCH
}

int amr_download(amr_ctx* c, int inst, int field, int nslots, double* host) {
  AMR_TRY
  if (check_inst(c, inst) || field < 0 || field >= kNumFields || !host) return AMR_ERR_CONFIG;
  Instance& I = c->inst[inst];
  if (nslots < 0 || nslots > I.maxblocks) return AMR_ERR_CONFIG;
  const size_t per = (size_t)kFieldVars[field] * c->NT * c->NT;
  HIPCHK(hipStreamSynchronize(c->stream));
  // blocks the instance has never used hold no data: zeros, no slot is taken
  for (int nb = 0; nb < nslots; nb++) {
    if (nb >= (int)I.gs.size()) {
      std::memset(host + per * nb, 0, per * sizeof(double));
      continue;
    }
    HIPCHK(hipMemcpy(host + per * nb, c->F.f[field] + per * I.gs[nb], per * sizeof(double), hipMemcpyDeviceToHost));
  }
  return AMR_OK;
  AMR_CATCH
}

int amr_upload(amr_ctx* c, int inst, int field, int nslots, const double* host) {
  AMR_TRY
  fill_touched(c);
  if (check_inst(c, inst) || field < 0 || field >= kNumFields || !host) return AMR_ERR_CONFIG;
  Instance& I = c->inst[inst];
  // only blocks the instance uses (amr_init, amr_set_tree, regrids) have slots
  if (nslots < 0 || nslots > (int)I.gs.size()) return AMR_ERR_CONFIG;
  const size_t per = (size_t)kFieldVars[field] * c->NT * c->NT;
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int nb = 0; nb < nslots; nb++)
    HIPCHK(hipMemcpy(c->F.f[field] + per * I.gs[nb], host + per * nb, per * sizeof(double), hipMemcpyHostToDevice));
  c->dma_since_direct = true;
  return AMR_OK;
  AMR_CATCH
}

int amr_get_tree(amr_ctx* c, int inst, int n, int32_t* lp, int32_t* ic, int32_t* lf) {
  AMR_TRY
  if (check_inst(c, inst)) return AMR_ERR_CONFIG;
  const Instance& I = c->inst[inst];
  if (n > (int)I.lp.size()) return AMR_ERR_CONFIG;
  for (int nb = 0; nb < n; nb++) {
    for (int s = 0; s < 8; s++) lp[8 * nb + s] = I.lp[nb][s];
    ic[2 * nb] = I.icoord[nb][0];
    ic[2 * nb + 1] = I.icoord[nb][1];
    lf[nb] = I.leafs[nb];
  }
  return AMR_OK;
  AMR_CATCH
}

int amr_set_tree(amr_ctx* c, int inst, int n, const int32_t* lp, const int32_t* ic, const int32_t* lf) {
  AMR_TRY
  fill_touched(c);
  if (check_inst(c, inst) || c->in_flight) return AMR_ERR_CONFIG;
  Instance& I = c->inst[inst];
  if (int e = validate_tree(n, I.maxblocks, I.maxlevs, c->nxi, lp, ic, lf)) return e;
  // every block gets its slot first, so a full arena leaves the instance unchanged
  for (int nb = 0; nb < n; nb++)
    if (slot_of(c, I, nb) < 0) return set_error(c, AMR_ERR_CAPACITY, inst, nb, 0, 0);
  grow_tree(I, n);
  for (int nb = 0; nb < n; nb++) {
    for (int s = 0; s < 8; s++) I.lp[nb][s] = lp[8 * nb + s];
    I.icoord[nb] = {ic[2 * nb], ic[2 * nb + 1]};
    I.leafs[nb] = (uint8_t)lf[nb];
  }
  // blocks past the new tree are unused: a regrid allocates its new son
  // groups there and expects them blank (a longer tree may have left rows)
  for (size_t nb = n; nb < I.lp.size(); nb++) {
    I.lp[nb] = std::array<int32_t, 8>{0, 0, 0, 0, 0, 0, 0, 0};
    I.icoord[nb] = std::array<int32_t, 2>{0, 0};
    I.leafs[nb] = I.iref[nb] = I.icoarse[nb] = 0;
  }
  I.lastActive = n;
  I.done = I.status != 0;
  I.prog_dirty = true;
  I.dev_tree_dirty = true;
  c->topo_dirty = true;
  return AMR_OK;
  AMR_CATCH
}

int amr_validate_tree(int nx, int maxlevs, int n, const int32_t* lp, const int32_t* icoord, const int32_t* leafs) {
  AMR_TRY
  if (nx <= 0 || maxlevs < 1 || maxlevs > 7) return AMR_ERR_CONFIG;
  return validate_tree(n, maxblocks_of(maxlevs), maxlevs, nx, lp, icoord, leafs);
  AMR_CATCH
}

int amr_set_time(amr_ctx* c, int inst, double t, double dt, double dt_next, int32_t count) {
  AMR_TRY
  fill_touched(c);
  if (check_inst(c, inst)) return AMR_ERR_CONFIG;
  Instance& I = c->inst[inst];
  I.t = t;
  I.dt_dom = dt;
  I.dt_main = dt_next;
  I.count = count;
  return AMR_OK;
  AMR_CATCH
}

}  // extern "C"

// ----------------------------------------------------------- checkpoint
// Binary checkpoint of one instance (no counterpart in the reference, whose
// only output, save4py in savedata.cpp:7-46, keeps 6 significant digits of
// the primitives of the leaves and cannot restart a run).  It holds what the
// next steps read: the configuration it belongs to, t and the two-deep dt lag
// (main.cpp:120-122, 231-234: the dt applied in the current step and the one
// computed for the next), the step count, the status, the whole quadtree of
// every block ever allocated (stale blocks included: their slots are reused,
// admesh.cpp:33-51) and all seven slot fields of those blocks (stale per-slot
// ff/gg/EMF values feed later ghost fills).  Little-endian, FNV-1a-64 of all
// preceding bytes at the end:
//   CkptHeader | lp int32[n][8] | icoord int32[n][2] | leafs int32[n] |
//   U, Us, Bi, Bis, ff, gg, EMF float64[n][nv][NT][NT] | uint64 checksum
namespace {

struct CkptHeader {
  char magic[8];
  int32_t version, header_bytes;
  int32_t nx, nghosts, maxlevs, NT, rk, periodic, lim, solver, ic, last_active, count, status;
  int32_t nfields, ct_type;
  double t, dt, dt_next, cfl, lenx, gamma, t_end;
  double ic_params[8];
};
constexpr char kCkptMagic[8] = {'A', 'M', 'R', '2', 'D', 'C', 'K', 'P'};
constexpr int32_t kCkptVersion = 1;

uint64_t fnv1a(uint64_t h, const void* p, size_t n) {
  const unsigned char* b = (const unsigned char*)p;
  for (size_t k = 0; k < n; k++) {
    h ^= b[k];
    h *= 1099511628211ull;
  }
  return h;
}
constexpr uint64_t kFnvBasis = 14695981039346656037ull;

CkptHeader ckpt_header(const amr_ctx* c, const Instance& I) {
  CkptHeader h{};
  std::memcpy(h.magic, kCkptMagic, 8);
  h.version = kCkptVersion;
  h.header_bytes = (int32_t)sizeof(CkptHeader);
  h.nx = c->nxi;
  h.nghosts = c->ng;
  h.maxlevs = I.maxlevs;
  h.NT = c->NT;
  h.rk = c->rk;
  h.periodic = I.periodic ? 1 : 0;
  h.lim = I.lim;
  h.solver = I.solver;
  h.ic = I.ic;
  h.last_active = I.lastActive;
  h.count = I.count;
  h.status = I.status;
  h.nfields = kNumFields;
  h.ct_type = I.ct_type;
  h.t = I.t;
  h.dt = I.dt_dom;
  h.dt_next = I.dt_main;
  h.cfl = I.cfl;
  h.lenx = I.cfg.lenx;
  h.gamma = I.gamma;
  h.t_end = I.tEnd;
  for (int k = 0; k < 8; k++) h.ic_params[k] = I.ic == IC_ROTOR_P ? I.cfg.ic_params[k] : 0.;
  return h;
}

// the configuration fields a checkpoint must agree on with the instance
bool ckpt_matches(const CkptHeader& a, const CkptHeader& b) {
  return a.nx == b.nx && a.nghosts == b.nghosts && a.maxlevs == b.maxlevs && a.NT == b.NT && a.rk == b.rk &&
         a.periodic == b.periodic && a.lim == b.lim && a.solver == b.solver && a.ic == b.ic &&
         a.nfields == b.nfields && (a.ct_type == 2) == (b.ct_type == 2) && a.cfl == b.cfl && a.lenx == b.lenx &&
         a.gamma == b.gamma &&
         a.t_end == b.t_end && std::memcmp(a.ic_params, b.ic_params, sizeof(a.ic_params)) == 0;
}

struct FileCloser {
  void operator()(std::FILE* f) const {
    if (f) std::fclose(f);
  }
};

}  // namespace

extern "C" {

int amr_checkpoint_save(amr_ctx* c, int inst, const char* path) {
  AMR_TRY
  if (check_inst(c, inst) || !path || c->in_flight) return AMR_ERR_CONFIG;
  const Instance& I = c->inst[inst];
  const int n = I.lastActive;
  HIPCHK(hipStreamSynchronize(c->stream));
  const std::string tmp = std::string(path) + ".tmp";
  std::unique_ptr<std::FILE, FileCloser> fp(std::fopen(tmp.c_str(), "wb"));
  if (!fp) return AMR_ERR_IO;
  uint64_t h = kFnvBasis;
  bool ok = true;
  auto put = [&](const void* p, size_t bytes) {
    h = fnv1a(h, p, bytes);
    ok = ok && std::fwrite(p, 1, bytes, fp.get()) == bytes;
  };
  const CkptHeader hd = ckpt_header(c, I);
  put(&hd, sizeof hd);
  std::vector<int32_t> lp(8 * (size_t)n), ic(2 * (size_t)n), lf(n);
  for (int nb = 0; nb < n; nb++) {
    for (int s = 0; s < 8; s++) lp[8 * nb + s] = I.lp[nb][s];
    ic[2 * nb] = I.icoord[nb][0];
    ic[2 * nb + 1] = I.icoord[nb][1];
    lf[nb] = I.leafs[nb];
  }
  put(lp.data(), lp.size() * 4);
  put(ic.data(), ic.size() * 4);
  put(lf.data(), lf.size() * 4);
  std::vector<double> tile;
  for (int f = 0; f < kNumFields && ok; f++) {
    const size_t per = (size_t)kFieldVars[f] * c->NT * c->NT;
    tile.resize(per);
    for (int nb = 0; nb < n && ok; nb++) {
      HIPCHK(hipMemcpy(tile.data(), c->F.f[f] + per * I.gs[nb], per * sizeof(double), hipMemcpyDeviceToHost));
      put(tile.data(), per * sizeof(double));
    }
  }
  ok = ok && std::fwrite(&h, 1, sizeof h, fp.get()) == sizeof h;
  ok = std::fclose(fp.release()) == 0 && ok;
  // the file appears under its name only when complete
  if (!ok || std::rename(tmp.c_str(), path) != 0) {
    std::remove(tmp.c_str());
    return AMR_ERR_IO;
  }
  return AMR_OK;
  AMR_CATCH
}

int amr_checkpoint_load(amr_ctx* c, int inst, const char* path) {
  AMR_TRY
  fill_touched(c);
  if (check_inst(c, inst) || !path || c->in_flight) return AMR_ERR_CONFIG;
  Instance& I = c->inst[inst];
  std::unique_ptr<std::FILE, FileCloser> fp(std::fopen(path, "rb"));
  if (!fp) return AMR_ERR_IO;
  uint64_t h = kFnvBasis;
  auto get = [&](void* p, size_t bytes) {
    if (std::fread(p, 1, bytes, fp.get()) != bytes) return false;
    h = fnv1a(h, p, bytes);
    return true;
  };
  CkptHeader hd{};
  if (!get(&hd, sizeof hd) || std::memcmp(hd.magic, kCkptMagic, 8) != 0 || hd.version != kCkptVersion ||
      hd.header_bytes != (int32_t)sizeof(CkptHeader))
    return AMR_ERR_IO;
  // a checkpoint restarts only the run it was written from
  if (!ckpt_matches(hd, ckpt_header(c, I))) return AMR_ERR_CONFIG;
  const int n = hd.last_active;
  if (n < kNbRoots || n > I.maxblocks) return AMR_ERR_IO;
  std::vector<int32_t> lp(8 * (size_t)n), ic(2 * (size_t)n), lf(n);
  if (!get(lp.data(), lp.size() * 4) || !get(ic.data(), ic.size() * 4) || !get(lf.data(), lf.size() * 4))
    return AMR_ERR_IO;
  std::vector<std::vector<double>> fields(kNumFields);
  for (int f = 0; f < kNumFields; f++) {
    fields[f].resize((size_t)n * kFieldVars[f] * c->NT * c->NT);
    if (!get(fields[f].data(), fields[f].size() * sizeof(double))) return AMR_ERR_IO;
  }
  uint64_t stored = 0;
  if (std::fread(&stored, 1, sizeof stored, fp.get()) != sizeof stored || stored != h) return AMR_ERR_IO;
  if (validate_tree(n, I.maxblocks, I.maxlevs, c->nxi, lp.data(), ic.data(), lf.data())) return AMR_ERR_IO;
  // the file is whole and consistent: apply it
  if (int e = amr_set_tree(c, inst, n, lp.data(), ic.data(), lf.data())) return e;
  for (int f = 0; f < kNumFields; f++)
    if (int e = amr_upload(c, inst, f, n, fields[f].data())) return e;
  I.t = hd.t;
  I.dt_dom = hd.dt;
  I.dt_main = hd.dt_next;
  I.count = hd.count;
  I.status = hd.status;
  I.err = amr_diag{hd.status, inst, -1, 0, 0, 0};
  I.done = I.status != 0 || !(I.t <= I.tEnd);
  c->topo_dirty = true;
  return AMR_OK;
  AMR_CATCH
}

int amr_fill(amr_ctx* c, int field, int kind) {
  AMR_TRY
  fill_touched(c);
  if (!c || field < 0 || field >= kNumFields || kind < 0 || kind > 2) return AMR_ERR_CONFIG;
  c->call_fail = 0;
  if (c->topo_dirty)
    if (int e = rebuild_topology(c)) return e;
  run_fill(c, make_set(c, {{field, kind}}));
  return finish_call(c);
  AMR_CATCH
}

int amr_stage(amr_ctx* c, int step) {
  AMR_TRY
  fill_touched(c);
  if (!c || step < 1 || step > c->rk) return AMR_ERR_CONFIG;
  c->call_fail = 0;
  if (c->topo_dirty)
    if (int e = rebuild_topology(c)) return e;
  if (int e = upload_params(c)) return e;
  run_stage(c, step, false);
  return finish_call(c);
  AMR_CATCH
}

int amr_ct(amr_ctx* c, int step) {
  AMR_TRY
  fill_touched(c);
  if (!c || step < 1 || step > c->rk) return AMR_ERR_CONFIG;
  c->call_fail = 0;
  if (c->topo_dirty)
    if (int e = rebuild_topology(c)) return e;
  if (int e = upload_params(c)) return e;
  run_ct(c, step);
  return finish_call(c);
  AMR_CATCH
}

int amr_cfl(amr_ctx* c, int inst, double* lambda) {
  AMR_TRY
  fill_touched(c);
  if (check_inst(c, inst) || !lambda) return AMR_ERR_CONFIG;
  c->call_fail = 0;
  if (c->topo_dirty)
    if (int e = rebuild_topology(c)) return e;
  if (int e = upload_params(c)) return e;
  if (int e = cfl_and_indicators(c, nullptr)) return e;
  if (c->call_fail) return c->call_fail;
  double lam = 1.E6;
  for (int l = c->leaf_inst_begin[inst]; l < c->leaf_inst_begin[inst + 1]; l++) lam = std_min(lam, c->h_red[2 * l]);
  *lambda = lam;
  return AMR_OK;
  AMR_CATCH
}

int amr_regrid(amr_ctx* c) {
  AMR_TRY
  fill_touched(c);
  if (!c) return AMR_ERR_CONFIG;
  c->call_fail = 0;
  if (c->topo_dirty)
    if (int e = rebuild_topology(c)) return e;
  if (int e = upload_params(c)) return e;
  if (int e = c->regrid_device ? cfl_and_regrid(c, nullptr) : cfl_and_indicators(c, nullptr)) return e;
  if (int e = c->regrid_device ? regrid_replay_all(c) : regrid_all(c)) return e;
  return finish_call(c);
  AMR_CATCH
}

int amr_schedule(amr_ctx* c, int inst, int cap, int32_t* ib3, int* n) {
  AMR_TRY
  fill_touched(c);
  if (check_inst(c, inst) || !n) return AMR_ERR_CONFIG;
  Instance& I = c->inst[inst];
  build_schedule(I);
  *n = (int)I.sched.size();
  for (int e = 0; e < *n && e < cap; e++)
    for (int q = 0; q < 3; q++) ib3[3 * e + q] = I.sched[e][q];
  return AMR_OK;
  AMR_CATCH
}

int amr_riemann_batch(int solver, int dir, double gamma, int n, const double* wL, const double* wR, double* flux,
                      int32_t* status, void* stream) {
  AMR_TRY
  if (solver < 0 || solver >= kFluxCount || (dir != 0 && dir != 1) || n < 0) return AMR_ERR_CONFIG;
  launch_riemann_batch(solver, dir, gamma, n, wL, wR, flux, status, (hipStream_t)stream);
  HIPCHK(hipGetLastError());
  return AMR_OK;
  AMR_CATCH
}

int amr_recon_batch(int NT, int nghosts, const char* limiter, double dx, int nt, const double* W, double* out,
                    void* stream) {
  AMR_TRY
  const int lim = limiter ? lim_code(limiter) : -1;
  if (lim < 0 || NT < 2 * nghosts + 4) return AMR_ERR_CONFIG;
  launch_recon_batch(NT, nghosts, lim, dx, nt, W, out, (hipStream_t)stream);
  HIPCHK(hipGetLastError());
  return AMR_OK;
  AMR_CATCH
}

int amr_set_stage_variant(amr_ctx* c, int variant) {
  AMR_TRY
  fill_touched(c);
  if (!c || variant < -1 || variant >= kStageVariants) return AMR_ERR_CONFIG;
  c->stage_variant = (variant < 0 && c->wall_khz <= 0) ? 0 : variant;
  c->graph_dirty = true;
  return AMR_OK;
  AMR_CATCH
}

int amr_stage_tuning(amr_ctx* c, int nbuckets, int32_t* chosen, double* us_per_leaf) {
  AMR_TRY
  if (!c || nbuckets < 0) return AMR_ERR_CONFIG;
  const auto* tune = c->tune_by_spec[stage_spec_of(c)];
  for (int b = 0; b < nbuckets && b < 24; b++) {
    if (chosen) chosen[b] = tune[b].chosen;
    for (int v = 0; v < kStageVariants && us_per_leaf; v++)
      us_per_leaf[b * kStageVariants + v] = tune[b].n[v] ? tune[b].us_per_leaf[v] / tune[b].n[v] : 0.0;
  }
  return AMR_OK;
  AMR_CATCH
}

int amr_set_stage_tuning(amr_ctx* c, int nbuckets, const int32_t* chosen) {
  AMR_TRY
  fill_touched(c);
  if (!c || nbuckets < 0 || nbuckets > 24 || (nbuckets && !chosen)) return AMR_ERR_CONFIG;
  for (int b = 0; b < nbuckets; b++)
    if (chosen[b] < -1 || chosen[b] >= kStageVariants) return AMR_ERR_CONFIG;
  auto* tune = c->tune_by_spec[stage_spec_of(c)];
  for (int b = 0; b < nbuckets; b++)
    if (chosen[b] >= 0) tune[b].chosen = chosen[b];
  c->graph_dirty = true;
  return AMR_OK;
  AMR_CATCH
}

int amr_set_step_kernel(amr_ctx* c, int mode) {
  AMR_TRY
  fill_touched(c);
  if (!c || mode < -1 || mode > 2 || c->in_flight) return AMR_ERR_CONFIG;
  c->step_mode = mode < 0 ? 0 : mode;
  return AMR_OK;
  AMR_CATCH
}

int amr_set_regrid_device(amr_ctx* c, int on) {
  AMR_TRY
  fill_touched(c);
  if (!c || c->in_flight || (on != 0 && on != 1)) return AMR_ERR_CONFIG;
  if (c->regrid_device != (on == 1)) {
    c->regrid_device = on == 1;
    for (auto& I : c->inst) I.dev_tree_dirty = true;
    c->topo_dirty = true;
  }
  return AMR_OK;
  AMR_CATCH
}

// The instance's current ghost-fill program by op kind: out[0..4] = copies,
// 2x2 restrictions, general restrictions, padding ops, explicit-operand
// entries (what one fill launch reads per op and per component).
int amr_debug_fill_stats(amr_ctx* c, int inst, int64_t* out) {
  AMR_TRY
  if (check_inst(c, inst) || !out) return AMR_ERR_CONFIG;
  const Instance& I = c->inst[inst];
  for (int k = 0; k < 4; k++) out[k] = I.op_kinds[k];
  out[4] = I.ngen;
  return AMR_OK;
  AMR_CATCH
}

// device regrid: 0 works on every tree in place (global memory), 1 (the
// default) stages the trees that fit in LDS
int amr_debug_regrid_lds(amr_ctx* c, int on) {
  AMR_TRY
  fill_touched(c);
  if (!c || c->in_flight) return AMR_ERR_CONFIG;
  c->regrid_lds = on != 0;
  c->topo_dirty = true;
  return AMR_OK;
  AMR_CATCH
}

int amr_debug_device_tree(amr_ctx* c, int inst, int n, int32_t* lp, int32_t* leafs, int* current) {
  AMR_TRY
  fill_touched(c);
  if (check_inst(c, inst) || !lp || !leafs || !current || c->in_flight) return AMR_ERR_CONFIG;
  Instance& I = c->inst[inst];
  if (n < 0 || n > I.maxblocks) return AMR_ERR_CONFIG;
  HIPCHK(hipStreamSynchronize(c->stream));
  const size_t off = (size_t)c->tree_off[inst];
  std::vector<uint8_t> lf(n);
  if (n) {
    HIPCHK(hipMemcpy(lp, c->dtree.lp + 8 * off, sizeof(int32_t) * 8 * n, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(lf.data(), c->dtree.leafs + off, n, hipMemcpyDeviceToHost));
  }
  for (int nb = 0; nb < n; nb++) leafs[nb] = lf[nb];
  *current = (c->regrid_device && !I.dev_tree_dirty) ? 1 : 0;
  return AMR_OK;
  AMR_CATCH
}

int amr_use_stream(amr_ctx* c, void* stream, int caller) {
  AMR_TRY
  fill_touched(c);
  if (!c || c->in_flight || c->in_group) return AMR_ERR_CONFIG;
  hipStream_t next = caller ? (hipStream_t)stream : c->own_stream;
  if (next == c->stream) return AMR_OK;
  // what the context enqueued so far comes before what it enqueues next
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  HIPCHK(hipStreamWaitEvent(next, c->ev0, 0));
  c->stream = next;
  c->graph_dirty = true;
  return AMR_OK;
  AMR_CATCH
}

int amr_set_graph(amr_ctx* c, int on) {
  AMR_TRY
  fill_touched(c);
  if (!c) return AMR_ERR_CONFIG;
  if (on) {
    // replaying one captured step over changing trees (device-side counts,
    // capacity grids) faulted the GPU in a long run; it stays off until the
    // cause is shown
    std::fprintf(stderr, "amr2d_mhd: graph replay of the step is disabled in this build\n");
    return AMR_ERR_CONFIG;
  }
  c->graph_on = on != 0;
  c->graph_dirty = true;
  return AMR_OK;
  AMR_CATCH
}

int amr_set_fill_cache(amr_ctx* c, int entries) {
  AMR_TRY
  fill_touched(c);
  if (!c || entries < 0 || entries > 64) return AMR_ERR_CONFIG;
  HIPCHK(hipStreamSynchronize(c->stream));
  c->prog_cache_max = entries;
  for (auto& I : c->inst) {
    for (auto& P : I.prog_cache) {
      hipFree(P.d_ops);
      hipFree(P.d_gen);
    }
    I.prog_cache.clear();
  }
  return AMR_OK;
  AMR_CATCH
}

int amr_profile_enable(amr_ctx* c, int mask) {
  AMR_TRY
  fill_touched(c);
  if (!c) return AMR_ERR_CONFIG;
  HIPCHK(hipStreamSynchronize(c->stream));
  flush_profile(c);
  c->prof_mask = (unsigned)mask;
  for (auto& p : c->prof) p = Profile{};
  for (auto& n : c->prof_seen) n = 0;
  return AMR_OK;
  AMR_CATCH
}

int amr_profile_sample(amr_ctx* c, int stride) {
  AMR_TRY
  if (!c || stride < 1) return AMR_ERR_CONFIG;
  c->prof_stride = stride;
  for (auto& n : c->prof_seen) n = 0;
  return AMR_OK;
  AMR_CATCH
}

int amr_profile_read(amr_ctx* c, int family, double* ms, int64_t* launches) {
  AMR_TRY
  if (!c || family < 0 || family >= PF_COUNT) return AMR_ERR_CONFIG;
  HIPCHK(hipStreamSynchronize(c->stream));
  flush_profile(c);
  if (ms) *ms = c->prof[family].ms;
  if (launches) *launches = c->prof[family].launches;
  return AMR_OK;
  AMR_CATCH
}

}  // extern "C"

// ====================================================================
// host-only test hooks (no device work): run the compiled ghost-fill
// program and the coarsen maps on host arrays, so the program compiler and
// the index tables can be checked on a machine without a GPU.
// ====================================================================
namespace {

// host mirror of the device fill evaluation (fill_value in kernels.hip)
struct HostFill {
  const double* Q;
  int nv, NT, N2, kind;
  const std::vector<FillGen>& gen;
  double elem(int64_t gs, int cell, int k) const { return Q[((size_t)gs * nv + k) * N2 + cell]; }
  double ref(int64_t r, int k) const { return elem(r >> 16, (int)(r & 0xffff), k); }
  double combine(int k, double a, double b, double c, double d) const {
    double v = 0.;
    if (kind == kFillCell) { v += a; v += b; v += c; v += d; return 0.25 * v; }
    if (kind == kFillFace) { v += a; v += k == 0 ? b : c; return 0.5 * v; }
    return a;
  }
  double general(const FillGen& g, int k) const { return combine(k, ref(g.ref[0], k), ref(g.ref[1], k), ref(g.ref[2], k), ref(g.ref[3], k)); }
  double value(const FillOp& op, int k) const {
    if (op.op == kFillCopy) return elem(op.src, op.src_cell, k);
    if (op.op == kFillRestrict)
      return combine(k, elem(op.src, op.src_cell, k), elem(op.src, op.src_cell + NT, k), elem(op.src, op.src_cell + 1, k),
                     elem(op.src, op.src_cell + 1 + NT, k));
    return general(gen[op.src], k);
  }
};

}  // namespace

extern "C" {

// Applies the fill program compiled for the given tree to Q[n][nv][NT][NT]
// in place, emulating the device chunk by chunk (read all values of a chunk,
// then write them); returns the number of live destination ops, or -1 on a
// compile error, -2 for a segment that is not whole chunks, -3 when an
// element one chunk writes is read by another chunk.  *nnodes receives the
// number of nested restriction nodes.  force_general != 0 emits every
// restriction through the explicit-operand path.
namespace {
int host_fill(int NT, int ng, int n, const int32_t* lp, const int32_t* icoord, const int32_t* leafs, double* Q, int nv,
              int kind, int* nnodes, int force_general, int* tail_chunk0) {
  amr_ctx c;
  c.NT = NT;
  c.ng = ng;
  c.nxi = NT - 2 * ng;
  c.inst.resize(1);
  Instance& I = c.inst[0];
  I.lastActive = n;
  grow_tree(I, n);
  for (int nb = 0; nb < n; nb++) {
    for (int s = 0; s < 8; s++) I.lp[nb][s] = lp[8 * nb + s];
    I.icoord[nb] = {icoord[2 * nb], icoord[2 * nb + 1]};
    I.leafs[nb] = (uint8_t)leafs[nb];
    I.gs.push_back(nb);
  }
  build_schedule(I);
  FillScratch S;
  S.force_general = force_general != 0;
  if (compile_instance(I, NT, ng, S)) return -1;
  const int N2 = NT * NT;
  HostFill hf{Q, nv, NT, N2, kind, I.gen};
  const FillOp* ops = I.ops.data();
  const size_t m = I.ops.size();
  if (m % kFillChunk) return -2;
  // the chunk invariant: no element written by one chunk is read by another
  std::vector<int32_t> chunk_of((size_t)n * N2, -1);
  for (size_t e = 0; e < m; e++)
    if (ops[e].op != kFillNop) chunk_of[(size_t)ops[e].dst_gs * N2 + ops[e].dst_cell] = (int32_t)(e / kFillChunk);
  const int off[4] = {0, NT, 1, NT + 1};
  for (size_t e = 0; e < m; e++) {
    const FillOp& op = ops[e];
    std::vector<int64_t> refs;
    if (op.op == kFillCopy) refs.push_back(src_ref(op.src, op.src_cell));
    else if (op.op == kFillRestrict)
      for (int t = 0; t < 4; t++) refs.push_back(src_ref(op.src, op.src_cell + off[t]));
    else if (op.op == kFillGeneral)
      for (int t = 0; t < 4; t++) refs.push_back(I.gen[op.src].ref[t]);
    for (int64_t r : refs) {
      const int32_t ch = chunk_of[(size_t)(r >> 16) * N2 + (r & 0xffff)];
      if (ch >= 0 && ch != (int32_t)(e / kFillChunk)) return -3;
    }
  }
  // emulate the device: per chunk, read every value, then write
  std::vector<double> buf((size_t)kFillChunk * 8, 0.0);
  const size_t first = tail_chunk0 ? (size_t)I.tail_chunk0 * kFillChunk : 0;
  if (tail_chunk0) *tail_chunk0 = I.tail_chunk0;
  for (size_t c0 = first; c0 < m; c0 += kFillChunk) {
    for (size_t e = c0; e < c0 + kFillChunk; e++)
      if (ops[e].op != kFillNop)
        for (int k = 0; k < nv; k++) buf[(e - c0) * 8 + k] = hf.value(ops[e], k);
    for (size_t e = c0; e < c0 + kFillChunk; e++)
      if (ops[e].op != kFillNop)
        for (int k = 0; k < nv; k++) Q[((size_t)ops[e].dst_gs * nv + k) * N2 + ops[e].dst_cell] = buf[(e - c0) * 8 + k];
  }
  if (nnodes) *nnodes = 0;
  int live = 0;
  for (size_t e = 0; e < m; e++) live += ops[e].op != kFillNop;
  return live;
}
}  // namespace

int amr_debug_host_fill(int NT, int ng, int n, const int32_t* lp, const int32_t* icoord, const int32_t* leafs,
                        double* Q, int nv, int kind, int* nnodes, int force_general) {
  AMR_TRY
  return host_fill(NT, ng, n, lp, icoord, leafs, Q, nv, kind, nnodes, force_general, nullptr);
  AMR_CATCH
}

// The same program applied from its linked tail on only (the chunks a
// repeated fill changes); *tail_chunk0 receives the tail's first chunk.
// Applied to the output of a whole fill, it must give what a second whole
// fill gives.
int amr_debug_host_fill_tail(int NT, int ng, int n, const int32_t* lp, const int32_t* icoord, const int32_t* leafs,
                             double* Q, int nv, int kind, int* tail_chunk0) {
  AMR_TRY
  if (!tail_chunk0) return -1;
  return host_fill(NT, ng, n, lp, icoord, leafs, Q, nv, kind, nullptr, 0, tail_chunk0);
  AMR_CATCH
}

// Compiled fill program of the given tree: up to `cap` ops copied to `ops`
// (16-byte FillOp records), the explicit-operand table size to *ngen; returns
// the op count (negative on error).
int amr_debug_compile(int NT, int ng, int n, const int32_t* lp, const int32_t* icoord, const int32_t* leafs, int cap,
                      void* ops, int* ngen) {
  AMR_TRY
  Instance I;
  I.lastActive = n;
  grow_tree(I, n);
  for (int nb = 0; nb < n; nb++) {
    for (int s = 0; s < 8; s++) I.lp[nb][s] = lp[8 * nb + s];
    I.icoord[nb] = {icoord[2 * nb], icoord[2 * nb + 1]};
    I.leafs[nb] = (uint8_t)leafs[nb];
    I.gs.push_back(nb);
  }
  build_schedule(I);
  FillScratch S;
  if (compile_instance(I, NT, ng, S)) return -1;
  const int m = (int)I.ops.size();
  if (ops) std::memcpy(ops, I.ops.data(), sizeof(FillOp) * std::min(m, cap));
  if (ngen) *ngen = (int)I.gen.size();
  return m;
  AMR_CATCH
}

// Fill program of the given tree compiled with its schedule cut into
// `ranges` ranges on up to `nthreads` threads (ranges 0: chosen by tree size
// as the engine does): up to `cap` ops to `ops`, up to `gcap` explicit-operand
// entries to `gen` and their count to *ngen; returns the op count (negative
// on error).  Any split must give the single-range program.  reuse != 0
// compiles with one scratch kept across calls (as an engine thread does),
// else with a fresh one.
int amr_debug_compile_ranges(int NT, int ng, int n, const int32_t* lp, const int32_t* icoord, const int32_t* leafs,
                             int ranges, int nthreads, int reuse, int cap, void* ops, int gcap, void* gen,
                             int* ngen) {
  AMR_TRY
  static FillScratch kept;
  Instance I;
  I.lastActive = n;
  grow_tree(I, n);
  for (int nb = 0; nb < n; nb++) {
    for (int s = 0; s < 8; s++) I.lp[nb][s] = lp[8 * nb + s];
    I.icoord[nb] = {icoord[2 * nb], icoord[2 * nb + 1]};
    I.leafs[nb] = (uint8_t)leafs[nb];
    I.gs.push_back(nb);
  }
  build_schedule(I);
  FillScratch fresh;
  FillScratch& S = reuse ? kept : fresh;
  S.force_ranges = ranges;
  if (compile_instance(I, NT, ng, S, nthreads)) return -1;
  const int m = (int)I.ops.size();
  if (ops) std::memcpy(ops, I.ops.data(), sizeof(FillOp) * std::min(m, cap));
  if (gen) std::memcpy(gen, I.gen.data(), sizeof(FillGen) * std::min<size_t>(I.gen.size(), (size_t)gcap));
  if (ngen) *ngen = (int)I.gen.size();
  return m;
  AMR_CATCH
}

// Exception barrier check: runs `ranges` ranges of the compile's thread pool
// on up to `nthreads` threads, range `throw_at` raising std::bad_alloc; the
// entry point must return AMR_ERR_HOST (and 0 when throw_at is out of range).
int amr_debug_pool_throw(int ranges, int nthreads, int throw_at) {
  AMR_TRY
  FillScratch S;
  std::atomic<int> ran{0};
  for_ranges(S, std::max(1, std::min(ranges, 255)), nthreads, [&](int r) {
    ran.fetch_add(1);
    if (r == throw_at) throw std::bad_alloc();
  });
  return ran.load() == std::max(1, std::min(ranges, 255)) ? AMR_OK : AMR_ERR_CONFIG;
  AMR_CATCH
}

// Host cost of the schedule + fill-program compile for the given tree:
// mean nanoseconds over `reps` compiles on up to `nthreads` threads that
// reuse one scratch (as the engine does).
double amr_debug_compile_ns(int NT, int ng, int n, const int32_t* lp, const int32_t* icoord, const int32_t* leafs,
                            int reps, int nthreads) {
  AMR_TRY
  Instance I;
  I.lastActive = n;
  grow_tree(I, n);
  for (int nb = 0; nb < n; nb++) {
    for (int s = 0; s < 8; s++) I.lp[nb][s] = lp[8 * nb + s];
    I.icoord[nb] = {icoord[2 * nb], icoord[2 * nb + 1]};
    I.leafs[nb] = (uint8_t)leafs[nb];
    I.gs.push_back(nb);
  }
  FillScratch S;
  build_schedule(I);
  if (compile_instance(I, NT, ng, S, nthreads)) return -1;
  const auto t0 = std::chrono::steady_clock::now();
  for (int r = 0; r < reps; r++) {
    build_schedule(I);
    compile_instance(I, NT, ng, S, nthreads);
  }
  return std::chrono::duration<double, std::nano>(std::chrono::steady_clock::now() - t0).count() / std::max(1, reps);
  AMR_CATCH_NEG
}

// coarsen source maps (see coarsen_maps): mU[NT*NT], mB[2*NT*NT]
int amr_debug_coarsen_maps(int NT, int ng, int32_t* mU, int32_t* mB) {
  AMR_TRY
  std::vector<int32_t> u, b;
  coarsen_maps(NT, ng, u, b);
  std::memcpy(mU, u.data(), u.size() * sizeof(int32_t));
  std::memcpy(mB, b.data(), b.size() * sizeof(int32_t));
  return 0;
  AMR_CATCH
}

// the locate_bounds schedule of a tree (no device work)
// Algorithmic operations per zone of one RK2 stage (mean of stages 1 and 2)
// over the instance's current leaves, counted on the host from the current
// state: out[0..4] = add/sub, mul, div, sqrt, total.
int amr_debug_stage_flops(amr_ctx* c, int inst, double* out) {
  AMR_TRY
  if (check_inst(c, inst) || !out || c->in_flight) return AMR_ERR_CONFIG;
  HIPCHK(hipStreamSynchronize(c->stream));
  const Instance& I = c->inst[inst];
  const size_t N2 = (size_t)c->NT * c->NT;
  std::vector<double> U(8 * N2), B(2 * N2);
  StageFlops acc;
  int64_t leaves = 0;
  for (int nb = 0; nb < I.lastActive; nb++) {
    if (!I.leafs[nb]) continue;
    HIPCHK(hipMemcpy(U.data(), c->F.f[kU] + 8 * N2 * I.gs[nb], 8 * N2 * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(B.data(), c->F.f[kBi] + 2 * N2 * I.gs[nb], 2 * N2 * sizeof(double), hipMemcpyDeviceToHost));
    for (int step = 1; step <= 2; step++)
      count_stage_tile(c->NT, c->ng, U.data(), B.data(), I.gamma, I.dx[I.lp[nb][0]], I.dt_dom, I.lim, I.solver, step,
                       acc);
    leaves++;
  }
  const double zones = 2.0 * (double)leaves * (c->NT - 2 * c->ng) * (c->NT - 2 * c->ng);
  if (zones <= 0) return AMR_ERR_CONFIG;
  out[0] = acc.add / zones;
  out[1] = acc.mul / zones;
  out[2] = acc.div / zones;
  out[3] = acc.sqrt / zones;
  out[4] = out[0] + out[1] + out[2] + out[3];
  return AMR_OK;
  AMR_CATCH
}

// Host-only check of the host half of the device regrid: the regrid of the
// given tree and per-block indicators (grad[nb], read for leaves) decided on
// the host (regrid_decide), then its ordered op list -- refinement parents,
// coarsening first sons, as the device records them -- replayed on a fresh
// copy of the tree (regrid_replay).  Returns 0 when both give the same tree
// (every row, coordinates, leaf flags, lastActive), slots and data-transfer
// ops, 1 when they differ, negative on an input error; the decided tree goes
// to lp_out/icoord_out/leafs_out (up to `cap` blocks) and counts[0..2] =
// refinements, coarsenings, lastActive.
int amr_debug_regrid_replay(int NT, int ng, int maxlevs, int n, const int32_t* lp, const int32_t* icoord,
                            const int32_t* leafs, const double* grad, int cap, int32_t* lp_out, int32_t* icoord_out,
                            int32_t* leafs_out, int32_t* counts) {
  AMR_TRY
  if (maxlevs < 1 || maxlevs > 7 || !lp || !icoord || !leafs || !grad || !counts) return AMR_ERR_CONFIG;
  if (validate_tree(n, maxblocks_of(maxlevs), maxlevs, NT - 2 * ng, lp, icoord, leafs)) return AMR_ERR_CONFIG;
  auto make = [&](amr_ctx& c) {
    c.NT = NT;
    c.ng = ng;
    c.inst.resize(1);
    Instance& I = c.inst[0];
    I.maxlevs = maxlevs;
    I.maxblocks = maxblocks_of(maxlevs);
    c.cap_slots = I.maxblocks;
    grow_tree(I, n);
    for (int nb = 0; nb < n; nb++) {
      for (int k = 0; k < 8; k++) I.lp[nb][k] = lp[8 * nb + k];
      I.icoord[nb] = {icoord[2 * nb], icoord[2 * nb + 1]};
      I.leafs[nb] = (uint8_t)leafs[nb];
      slot_of(&c, I, nb);
    }
    I.lastActive = n;
  };
  amr_ctx a, b;
  make(a);
  make(b);
  std::vector<double> g(grad, grad + n);
  std::vector<RegridOp> ra, ca, rb, cb;
  const int ea = regrid_decide(&a, 0, g, ra, ca);
  // the op list a device regrid would record: parents, then first sons
  std::vector<int32_t> log;
  for (const auto& o : ra) log.push_back(o.nb);
  for (const auto& o : ca) log.push_back(o.son_nb);
  RegridOut o{};
  o.nref = (int32_t)ra.size();
  o.ncoars = (int32_t)ca.size();
  const int eb = regrid_replay(&b, 0, o, log.data(), rb, cb);
  const Instance &A = a.inst[0], &B = b.inst[0];
  auto same_ops = [](const std::vector<RegridOp>& x, const std::vector<RegridOp>& y) {
    if (x.size() != y.size()) return false;
    return x.empty() || std::memcmp(x.data(), y.data(), sizeof(RegridOp) * x.size()) == 0;
  };
  bool same = ea == eb && A.lastActive == B.lastActive && A.gs == B.gs && same_ops(ra, rb) && same_ops(ca, cb);
  for (int nb = 0; same && nb < A.lastActive; nb++)
    same = A.lp[nb] == B.lp[nb] && A.icoord[nb] == B.icoord[nb] && A.leafs[nb] == B.leafs[nb];
  counts[0] = o.nref;
  counts[1] = o.ncoars;
  counts[2] = A.lastActive;
  for (int nb = 0; nb < std::min(cap, A.lastActive); nb++) {
    if (lp_out)
      for (int k = 0; k < 8; k++) lp_out[8 * nb + k] = A.lp[nb][k];
    if (icoord_out) {
      icoord_out[2 * nb] = A.icoord[nb][0];
      icoord_out[2 * nb + 1] = A.icoord[nb][1];
    }
    if (leafs_out) leafs_out[nb] = A.leafs[nb];
  }
  if (ea) return -ea;
  return same ? 0 : 1;
  AMR_CATCH
}

int amr_debug_schedule(int n, const int32_t* lp, const int32_t* leafs, int cap, int32_t* ib3) {
  AMR_TRY
  Instance I;
  I.lastActive = n;
  grow_tree(I, n);
  for (int nb = 0; nb < n; nb++) {
    for (int s = 0; s < 8; s++) I.lp[nb][s] = lp[8 * nb + s];
    I.leafs[nb] = (uint8_t)leafs[nb];
  }
  build_schedule(I);
  for (int e = 0; e < (int)I.sched.size() && e < cap; e++)
    for (int q = 0; q < 3; q++) ib3[3 * e + q] = I.sched[e][q];
  return (int)I.sched.size();
  AMR_CATCH
}

}  // extern "C"
