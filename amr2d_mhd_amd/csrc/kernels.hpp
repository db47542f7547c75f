// Launch interface of the AMR2D-MHD device kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "amr_types.hpp"

namespace amr {

struct FieldPtrs {
  double* f[kNumFields];
};

// up to four fields filled by one fill launch
struct FillSet {
  double* base[4];
  int32_t nv[4];
  int32_t kind[4];
  int32_t n;
  // optional: copy_words 8-byte words copy_src -> copy_dst done by the same
  // launch (the step's parameter upload from pinned host memory, so that it
  // costs no copy dispatch of its own); no fill op reads copy_dst
  int32_t copy_words;
  const double* copy_src;
  double* copy_dst;
};

// refine (4 sons from a parent) or coarsen (parent from 4 sons) operation
struct RegridOp {
  int32_t parent_gs;
  int32_t son_gs[4];
  int32_t inst;
  int32_t nb;      // parent (refine) or dad (coarsen) block id
  int32_t son_nb;  // first son's block id; the four sons are son_nb .. son_nb + 3
  int32_t seq;     // position among the instance's ops of this regrid, in the
                   // reference's order (all refinements, then all coarsenings)
};

bool tile_supported(int NT, int ng);

// stage kernel variants (identical results): lanes per leaf block and where
// the face fluxes live for the RK update --
//   0: 256 lanes, fluxes in LDS      1: 512 lanes, LDS      2: 768 lanes, LDS
//   3: 256 lanes, fluxes read back from ff/gg, compiled for 4 waves per SIMD
//   4: 512 lanes, read back, 4 waves per SIMD   5: 512 lanes, LDS, 4 waves per SIMD
//   6: each leaf split into 2 row strips, one workgroup per strip (fluxes in LDS)
//   7: 4 row strips per leaf
// The fewer-lane variants suit large leaf counts (throughput), the wide ones
// and the strips small leaf counts (per-block latency, filling the device).
// A non-null `span` receives the launch's device wall-clock span (see
// stamp_start/stamp_end).
constexpr int kStageVariants = 8;
// stage-kernel specialisations: generic (solver and limiter read per leaf), or
// HLLD with a MUSCL-family limiter, or HLLD with WENO5 (24x24 tiles, the only
// ones with the ghost depth WENO5 needs), fixed at compile time
// (kStageHLLCMuscl: HLLC with a MUSCL-family limiter, used for the HLLC part of
// a launch split by specialisation; variant 5 only)
enum { kStageGeneric = 0, kStageHLLDMuscl = 1, kStageHLLDWeno5 = 2, kStageHLLCMuscl = 3, kStageSpecs = 4 };
// the variant a launch split by specialisation uses unless one is fixed
constexpr int kSplitVariant = 5;
// Work-list counts on the device (counts[kCountLeaves], counts[kCountChunks]):
// a launch given `counts` may have a larger grid than its work (a captured
// step replayed over changing trees) and its surplus workgroups exit; with
// counts = null the grid is the work.
enum { kCountLeaves = 0, kCountChunks = 1 };
void launch_stage(int NT, int ng, int variant, int spec, const FieldPtrs& F, const LeafItem* leaves, int nleaf,
                  const InstParams* ip, int step, DevError* err, hipStream_t s, hipEvent_t ev_start = nullptr,
                  hipEvent_t ev_stop = nullptr, unsigned long long* span = nullptr, const int32_t* counts = nullptr);
// corner EMF of RK stage `step` of an `rk`-stage scheme (CTtype per instance)
void launch_emf(int NT, const FieldPtrs& F, const LeafItem* leaves, int nleaf, const InstParams* ip, int step,
                int rk, DevError* err, hipStream_t s, const int32_t* counts = nullptr);
// face-B update of RK stage `step` of an `rk`-stage scheme; cell sync into U
// (last stage) or Us
void launch_bupdate(int NT, const FieldPtrs& F, const LeafItem* leaves, int nleaf, const InstParams* ip, int step,
                    int rk, hipStream_t s, const int32_t* counts = nullptr);
void launch_bsync(int NT, const FieldPtrs& F, const LeafItem* leaves, int nleaf, int last, hipStream_t s,
                  const int32_t* counts = nullptr);
// ghost fill of up to four fields through a program made of per-instance
// segments of whole kFillChunk-op chunks (device pointer tables
// inst_ops/inst_gen, cprefix[ninst+1] of chunk counts), one workgroup per chunk
void launch_fill(int NT, const FillOp* const* inst_ops, const FillGen* const* inst_gen, const int32_t* cprefix,
                 int ninst, int nchunks, const FillSet& set, hipStream_t s, const int32_t* counts = nullptr);
// per-leaf CFL bound min(dx,dy)/lambda and refinement indicator, as pairs
// lam_grad[2 leaf], lam_grad[2 leaf + 1]; when W is non-null it supplies the
// primitives (initial regrid) instead of U.
void launch_cfl_criteria(int NT, int ng, const FieldPtrs& F, const double* W, const LeafItem* leaves, int nleaf,
                         const InstParams* ip, double* lam_grad, DevError* err, hipStream_t s,
                         const int32_t* counts = nullptr);
void launch_refine(int NT, int ng, const FieldPtrs& F, const RegridOp* ops, int nops, const InstParams* ip,
                   DevError* err, hipStream_t s);

// device copy of the quadtrees of a context: lp rows [block][8] and per-block
// flags; instance blocks from RegridInst::off on
struct DeviceTree {
  int32_t* lp;
  uint8_t* leafs;
  uint8_t* iref;
  uint8_t* icoarse;
  uint8_t* mark;
};
// regrid decision and tree surgery of every instance (regrid.hip), one
// workgroup per instance, from the per-leaf CFL bounds and indicators
void launch_regrid(const RegridInst* ri, int ninst, const LeafItem* leaves, const double* lam_grad,
                   const DeviceTree& tree, RegridOut* out, int32_t* oplog, hipStream_t s);
void launch_coarsen(int NT, int ng, const FieldPtrs& F, const RegridOp* ops, int nops, const int32_t* cmapU,
                    const int32_t* cmapB, const InstParams* ip, DevError* err, hipStream_t s);

// ---------------------------------------------------- persistent step kernel
// One launch runs the device work of a whole step (main.cpp:128-230): its
// phases -- ghost fills, RK stages, the CT kernels, CFL/criteria -- run in
// order over the device work lists, separated by grid-wide barriers, in a
// grid of workgroups that are all resident at once (cooperative launch).
enum { kPhaseFill = 0, kPhaseStage, kPhaseEmf, kPhaseBupdate, kPhaseBsync, kPhaseCfl };
struct StepPhase {
  int8_t kind;
  int8_t step;      // stage / face update: RK stage; cell sync: 1 = last stage
  int8_t nf;        // fill: number of fields
  int8_t pad;
  int8_t field[4];  // fill: slot fields and their combine kinds
  int8_t fkind[4];
};
constexpr int kMaxPhases = 32;

// grid barrier state: arrival count and generation in separate 128-byte
// lines; `fail` is set when a wait exceeds its bound (then every later wait
// returns at once, so the launch still ends)
struct GridSync {
  unsigned int arrive;
  unsigned int pad0[31];
  unsigned int gen;
  unsigned int pad1[31];
};

struct StepArgs {
  FieldPtrs F;
  const LeafItem* leaves;
  int nleaf;
  int nchunks;
  const FillOp* const* ops;  // fill program tables (see launch_fill)
  const FillGen* const* gen;
  const int32_t* cprefix;
  int ninst;
  int copy_words;            // parameter upload done by the first phase
  const double* copy_src;
  double* copy_dst;
  const InstParams* ip;
  double* lam_grad;
  DevError* err;
  GridSync* sync;
  int32_t* fail;
  unsigned long long* stamps;  // when non-null: device clock at the start and after every phase
  int nphases;
  StepPhase phase[kMaxPhases];
};

// lanes per workgroup of the step kernel
constexpr int kStepLanes[2] = {256, 512};
// largest co-resident grid of the step kernel for the tile, specialisation and
// lane count (0: it cannot run here)
int step_kernel_grid(int NT, int ng, int spec, int lanes_idx);
// cooperative launch of the step kernel; events (optional) bracket it on s
int launch_step(int NT, int ng, int spec, int lanes_idx, int grid, const StepArgs& A, hipStream_t s,
                hipEvent_t ev_start = nullptr, hipEvent_t ev_stop = nullptr);

// pointwise / tile-level parity entry points
void launch_riemann_batch(int solver, int dir, double gamma, int n, const double* wL, const double* wR, double* flux,
                          int32_t* status, hipStream_t s);
void launch_recon_batch(int NT, int ng, int lim, double dx, int nt, const double* W, double* out, hipStream_t s);

}  // namespace amr
