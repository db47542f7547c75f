// Shared host/device types of the AMR2D-MHD engine.
#pragma once
#include <cstdint>

namespace amr {

// limiter ids (slopelimiter.cpp names; WENO3/WENO5 select WENO2D_primvar.cpp)
enum {
  kLim1st = 0,
  kLimNoSL,
  kLimMinmod,
  kLimVanLeer,
  kLimSuperbee,
  kLimMC,
  kLimKoren,
  kLimCada,
  kLimWENO3,
  kLimWENO5,
  kLimCount
};
// Riemann solver ids (riemannS.cpp names)
enum { kFluxRusanov = 0, kFluxHLLE, kFluxHLLC, kFluxHLLD, kFluxCount };

// slot fields held in HBM for every slot of every instance, [slot][var][j][i]
enum { kU = 0, kUs, kBi, kBis, kFF, kGG, kEMF, kNumFields };
constexpr int kFieldVars[kNumFields] = {8, 8, 2, 2, 8, 8, 1};

// ghost-fill combine kinds (boundary.cpp / boundaryfc.cpp / boundaryec.cpp)
enum { kFillCell = 0, kFillFace = 1, kFillEdge = 2 };

// error codes reported by kernels and by the host driver
enum {
  kOk = 0,
  kErrPressure = 1,   // meshblock.cpp:136-142, 164-170
  kErrNanVel = 2,     // meshblock.cpp:171-176
  kErrNanFlux = 3,    // MUSCL2D.cpp:79-85, 103-108
  kErrCtFace = 4,     // MUSCL2D.cpp:87-91, 110-114
  kErrNanEdge = 5,    // celledges.cpp:17-21, 44-47
  kErrHLLC = 6,       // riemannS.cpp:269-276
  kErrDt = 7,         // main.cpp:235-238
  kErrOverflow = 8,   // admesh.cpp:42-45
  kErrRefineNan = 9,  // refineBi_CT.cpp:47
  kErrConfig = 10,
  kErrHip = 11,
  kErrCapacity = 12,
  kErrCtPsi = 15      // CT2D.cpp:60-61 (CTtype 2)
};

// one leaf block in a device work list
struct LeafItem {
  int32_t gs;    // global slot in the arena
  int32_t inst;  // instance index
  int32_t nb;    // reference block id within the instance (diagnostics)
  int32_t lev;   // refinement level
};

// per-instance parameters read by kernels
struct InstParams {
  double gamma;
  double dt;          // dt applied in the current step (meshblock::dt)
  double dx[16];      // cell size per level (dx == dy)
  int32_t limiter;
  int32_t solver;
  int32_t maxlevs;
  int32_t rk;         // RK stages: 2 (timeIntegral.cpp) or 3 (SSP-RK3)
  int32_t ct_type;    // corner EMF: 1 (CTtype 1) or 2 (CTtype 2, CT2D.cpp:28-75)
  int32_t pad;
};

// Ghost-fill program.  A reference to a value is either a pre-fill element
// (global slot << 16 | cell) or, with kNodeBit set, a nested restriction
// node evaluated earlier in the same fill.
constexpr int64_t kNodeBit = int64_t(1) << 62;
inline __host__ __device__ int64_t src_ref(int64_t gs, int cell) { return (gs << 16) | cell; }

enum { kFillCopy = 0, kFillRestrict = 1, kFillGeneral = 2, kFillNop = 3 };

// A fill program runs in chunks of kFillChunk ops, one workgroup each: the
// workgroup reads every operand of its chunk, synchronises, then writes.  The
// compiler keeps every destination that another op of the program reads
// (a stale-value read) in the same chunk as that reader, so no chunk writes
// an element another chunk reads; instance segments are padded with
// kFillNop to whole chunks.
constexpr int kFillChunk = 256;

// One destination element of a fill (16 bytes):
//   kFillCopy      value = pre-fill element (src, src_cell)
//   kFillRestrict  fine->coarse restriction of the 2x2 cells starting at
//                  (src, src_cell): (i1,j1), (i1,j1+1), (i1+1,j1), (i1+1,j1+1)
//   kFillGeneral   restriction whose operands are given by FillGen[src]
//   kFillNop       padding
// The combine rule (cell 2x2 mean, face-pair mean, edge injection) is chosen
// by the field being filled.
struct FillOp {
  int32_t dst_gs;
  uint16_t dst_cell;
  int16_t op;
  int32_t src;
  uint16_t src_cell;
  uint16_t pad;
};

// explicit operands of a restriction, in the order of FillOp's 2x2 cells
struct FillGen {
  int64_t ref[4];
};

// Device error record, one per instance.  Each of the reference's throw sites
// a kernel meets records a 64-bit key that orders it as the reference's
// sequential execution would meet it; the earliest one is kept (atomicMax of
// the complement; 0 = no error), so the record does not depend on which lane
// gets there first.  Key fields, most significant first:
//   launch ordinal (4 bits: stage step 1-3, CFL 4, refine 5, coarsen 6),
//   sequence (16: block id, or regrid-op index in the reference's order),
//   phase (4), i (8), j (8), check (4), reported block id (16), code (4).
struct DevError {
  unsigned long long nkey;
  int32_t pad[6];
};
enum { kOrdCfl = 4, kOrdRefine = 5, kOrdCoarsen = 6 };

// Device-side regrid (admesh.cpp:17-102), per instance: where its blocks sit
// in the context's tree arrays and its leaves in the leaf work list.
struct RegridInst {
  int32_t off;          // first block of the instance in the tree arrays
  int32_t maxblocks;    // defs.hpp:30
  int32_t maxlevs;
  int32_t last_active;  // meshblock::lastActive before the regrid
  int32_t leaf_begin, leaf_end;
  int32_t regrid;       // 1: the whole admesh; 0: only the CFL minimum
  int32_t lds_rows;     // rows staged in LDS (covering every block the regrid
                        // can reach), or 0: the tree is worked on in place
};
// most tree rows a regrid workgroup stages in LDS (16-bit ids and four flag
// bytes per block, 159 KB)
constexpr int kRgLdsRows = 8128;
// What the host reads back per instance: the CFL minimum and the regrid's
// decisions.  The op log holds the instance's refinements (parent ids), then
// its coarsenings (first-son ids), in the reference's order, from the
// instance's first leaf index on (at most one op per leaf).
struct RegridOut {
  double lambda;         // min over the leaves' CFL bounds, from 1e6
  int32_t nref, ncoars;
  int32_t last_active;   // after the regrid
  int32_t status;        // 0, or kErrOverflow at op nref - 1
  int32_t fail_nb, pad;
};
// sequence of the CT phase of a stage: after every block's stage errors
// (block ids stay below 2^15 for maxlevs <= 7), before the next stage's
constexpr int kSeqCt = 0x8000;

inline __host__ __device__ unsigned long long error_key(int ord, int seq, int phase, int i, int j, int chk, int nb,
                                                        int code) {
  return ((unsigned long long)(ord & 0xf) << 60) | ((unsigned long long)(seq & 0xffff) << 44) |
         ((unsigned long long)(phase & 0xf) << 40) | ((unsigned long long)(i & 0xff) << 32) |
         ((unsigned long long)(j & 0xff) << 24) | ((unsigned long long)(chk & 0xf) << 20) |
         ((unsigned long long)(nb & 0xffff) << 4) | (unsigned long long)(code & 0xf);
}

}  // namespace amr
