// FP64 HIP kernels of the AMR2D-MHD step for gfx950.
//
// Layout: every slot field is [slot][var][j][i] with NT x NT tiles
// (NT = 16 + 2*nghosts), i fastest, so a wavefront reads 64 consecutive
// doubles of one row.  One workgroup handles one leaf block; the block's
// primitive tile and its face fluxes live in LDS for the whole stage.
// No MFMA: nothing on this path is a dense contraction.
#include <hip/hip_ext.h>

#include <algorithm>
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "physics.hpp"

namespace amr {

namespace {

constexpr int kThreads = 256;

// per-instance error records: of the errors an instance's kernels meet, the
// one the reference's sequential execution meets first is kept (see
// DevError), and a failing instance never touches another instance's record
__device__ inline void report(DevError* err, int inst, unsigned long long key) {
  atomicMax(&err[inst].nkey, ~key);
}

// Kernel span from the device's constant-rate wall clock (100 MHz): span[0]
// accumulates the complement of the earliest workgroup start, span[1] the
// latest workgroup end (both start at zero; vector atomics).
__device__ inline void stamp_start(unsigned long long* span) {
  atomicMax(span, ~(unsigned long long)wall_clock64());
}
__device__ inline void stamp_end(unsigned long long* span) { atomicMax(span + 1, (unsigned long long)wall_clock64()); }

// ------------------------------------------------------------------ stage
// MUSCL2D.cpp:13-121 + the RK update of main.cpp:147-157 / 179-189 for one
// leaf: primitives of the whole tile, edge states, normal-B override from the
// faces, Riemann fluxes on [nxminb,nxmax]^2, then every cell of the tile is
// advanced (the residual is zero outside [nxminb,nxmax]^2).
//
// Variants (same arithmetic, same results): TPB lanes per leaf, and the face
// fluxes kept in LDS for the RK update (FLDS) or read back from the global
// ff/gg the block has just written (smaller LDS footprint, more resident
// workgroups per CU).  SPEC = kStageHLLDMuscl (kStageHLLDWeno5) fixes the
// solver to HLLD and the reconstruction to the MUSCL family (WENO5) at compile
// time, used when every leaf of the launch has them: the other paths'
// registers drop out.
template <int NT, int NG, bool FLDS>
struct StageLds {
  static constexpr int N2 = NT * NT;
  static constexpr int LO = NG - 1, HI = NT - 1 - NG, NF = HI - LO + 1;
  static constexpr int NFL = FLDS ? 8 * NF * (NF + 1) : 1;
  // doubles: primitive tile, x-face and y-face fluxes
  static constexpr int kDoubles = 8 * N2 + 2 * NFL;
};

// the Riemann solver of a stage specialisation (generic: the instance's)
template <int SPEC>
__device__ inline int spec_solver(int inst_solver) {
  return SPEC == kStageHLLCMuscl ? (int)kFluxHLLC : SPEC != kStageGeneric ? (int)kFluxHLLD : inst_solver;
}

// Primitives of tile cell c into sWc[k * VS] (k = 0..7) with the checks of
// meshblock.cpp:135-142 / 163-176, cell by cell: pressure, then velocity
// NaN; a NaN primitive sets *sNan where the limiter's edge check applies.
template <int NT, int NG>
__device__ inline void stage_prim_cell(const double* __restrict__ Qin, double g, int lim, int step, const LeafItem& L,
                                       DevError* err, int c, double* sWc, int VS, int* sNan) {
  constexpr int N2 = NT * NT, LO = NG - 1, HI = NT - 1 - NG;
  double u[8], w[8];
#pragma unroll
  for (int k = 0; k < 8; k++) u[k] = Qin[k * N2 + c];
  cons_to_prim(u, g, w);
#pragma unroll
  for (int k = 0; k < 8; k++) sWc[k * VS] = w[k];
  const int i = c % NT, j = c / NT;
  if (w[4] <= 0 && i > LO && i < HI + 1 && j > LO && j < HI + 1)
    report(err, L.inst, error_key(step, L.nb, 0, i, j, 0, L.nb, kErrPressure));
  if (step == 2 && (isnan(w[1]) || isnan(w[2]) || isnan(w[3])))
    report(err, L.inst, error_key(step, L.nb, 0, i, j, 1, L.nb, kErrNanVel));
  if (lim < kLimWENO3) {
    bool bad = false;
#pragma unroll
    for (int k = 0; k < 8; k++) bad |= isnan(w[k]);
    if (bad) *sNan = 1;
  }
}

// Flux fl of the x face (ydir = 0: between cells (i, j) and (i + 1, j)) or
// y face (between (i, j) and (i, j + 1)) from the primitives in sWb (cell
// (i, j) of variable k at sWb[k * VS + j * NT + i]) and the face B, with the
// solver-failure and NaN-flux checks of MUSCL2D.cpp:62-115.
template <int NT, int SPEC>
__device__ inline void stage_face(const double* sWb, int VS, const double* __restrict__ B, int lim, int solver,
                                  double g, double dxl, int i, int j, int ydir, int step, const LeafItem& L,
                                  DevError* err, double* fl) {
  constexpr int N2 = NT * NT;
  constexpr bool muscl_only = SPEC == kStageHLLDMuscl || SPEC == kStageHLLCMuscl, weno5_only = SPEC == kStageHLLDWeno5;
  const int s = ydir ? NT : 1;  // stride along the face normal
  const int c = j * NT + i;     // left cell; right cell is c + s
  double wl[8], wr[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const double* Wk = sWb + k * VS;
    if (weno5_only || (!muscl_only && lim == kLimWENO5)) {
      wl[k] = weno5_edge(Wk[c - 2 * s], Wk[c - s], Wk[c], Wk[c + s], Wk[c + 2 * s], true);
      wr[k] = weno5_edge(Wk[c - s], Wk[c], Wk[c + s], Wk[c + 2 * s], Wk[c + 3 * s], false);
    } else if (!muscl_only && !weno5_only && lim == kLimWENO3) {
      wl[k] = weno3_edge(Wk[c - s], Wk[c], Wk[c + s], true);
      wr[k] = weno3_edge(Wk[c], Wk[c + s], Wk[c + 2 * s], false);
    } else {
      const double a = Wk[c] - Wk[c - s], b = Wk[c + s] - Wk[c], d = Wk[c + 2 * s] - Wk[c + s];
      wl[k] = muscl_plus(lim, Wk[c], a, b, dxl);
      wr[k] = muscl_minus(lim, Wk[c + s], b, d, dxl);
    }
  }
  int st;
  if (!ydir) {
    wl[5] = B[0 * N2 + c + 1];
    wr[5] = B[0 * N2 + c + 1];
    st = normal_flux(solver, wl, wr, g, fl);
  } else {
    wl[6] = B[1 * N2 + c + NT];
    wr[6] = B[1 * N2 + c + NT];
    swap_xy(wl);
    swap_xy(wr);
    st = normal_flux(solver, wl, wr, g, fl);
    swap_xy(fl);
  }
  // per (i, j) the x face (solver, NaN), then the y face
  if (st) report(err, L.inst, error_key(step, L.nb, 2, i, j, ydir ? 3 : 0, L.nb, kErrHLLC));
  bool nan = false;
#pragma unroll
  for (int k = 0; k < 8; k++) nan |= isnan(fl[k]);
  if (nan) report(err, L.inst, error_key(step, L.nb, 2, i, j, ydir ? 4 : 1, L.nb, kErrNanFlux));
}

// RK update of the eight components of cell c of the tile from their
// residuals (main.cpp:147-157 / 179-189; SSP-RK3 build-defined).  Every input
// is loaded before the first store: U, Us and Uw may alias (Uw is U), so a
// store between two components' loads would hold each component's loads
// back until the previous component's store; the cell's elements are read
// and written by this lane only, so the order changes no value.
__device__ inline void stage_rk_cell(int step, int rk, double dt, const double* res, const double* U, double* Us,
                                     double* Uw, int c, int N2) {
  double u[8], us[8];
#pragma unroll
  for (int k = 0; k < 8; k++) u[k] = U[k * N2 + c];
  if (step != 1) {
#pragma unroll
    for (int k = 0; k < 8; k++) us[k] = Us[k * N2 + c];
  }
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const int e = k * N2 + c;
    if (step == 1) Us[e] = u[k] - dt * res[k];
    else if (rk == 2) Uw[e] = 0.5 * (u[k] + us[k] - dt * res[k]);
    else if (step == 2) Us[e] = 0.75 * u[k] + 0.25 * (us[k] - dt * res[k]);
    else Uw[e] = (1. / 3.) * u[k] + (2. / 3.) * (us[k] - dt * res[k]);
  }
}

// One leaf, all TPB lanes of the workgroup; sW/sF/sG/sNan: the workgroup's
// LDS (StageLds layout).  The caller separates two calls on the same LDS by a
// workgroup barrier.
template <int NT, int NG, int TPB, bool FLDS, int SPEC>
__device__ inline void stage_block(const FieldPtrs& F, const LeafItem L, const InstParams* __restrict__ ip, int step,
                                   DevError* err, double* sW, double* sF, double* sG, int* sNan) {
  using S = StageLds<NT, NG, FLDS>;
  constexpr int N2 = S::N2, LO = S::LO, HI = S::HI, NF = S::NF;
  const InstParams& P = ip[L.inst];
  const double g = P.gamma, dxl = P.dx[L.lev], dt = P.dt;
  const int lim = SPEC == kStageHLLDWeno5 ? (int)kLimWENO5 : P.limiter;
  const int solver = spec_solver<SPEC>(P.solver);
  const size_t t8 = (size_t)L.gs * 8 * N2, t2 = (size_t)L.gs * 2 * N2;
  const double* Qin = F.f[step == 1 ? kU : kUs] + t8;
  const double* __restrict__ B = F.f[step == 1 ? kBi : kBis] + t2;
  double* ff = F.f[kFF] + t8;
  double* gg = F.f[kGG] + t8;
  const int tid = threadIdx.x;
  if (tid == 0) *sNan = 0;
  __syncthreads();

  for (int c = tid; c < N2; c += TPB) stage_prim_cell<NT, NG>(Qin, g, lim, step, L, err, c, sW + c, N2, sNan);
  if (FLDS) {
    // stale edge fluxes ff[LO][j], gg[i][LO] enter the residual of ghost cells
    for (int q = tid; q < 8 * NF; q += TPB) {
      const int k = q / NF, r = q % NF;
      sF[k * NF * (NF + 1) + r * (NF + 1)] = ff[k * N2 + (LO + r) * NT + LO];
      sG[k * NF * (NF + 1) + 0 * NF + r] = gg[k * N2 + LO * NT + LO + r];
    }
  }
  __syncthreads();
  if (*sNan) {
    if (tid == 0) report(err, L.inst, error_key(step, L.nb, 1, 0, 0, 0, L.nb, kErrNanEdge));
    return;
  }

  for (int q = tid; q < 2 * NF * NF; q += TPB) {
    const int ydir = q >= NF * NF;
    const int r = ydir ? q - NF * NF : q;
    const int i = LO + r % NF, j = LO + r / NF;
    const int c = j * NT + i;
    double fl[8];
    stage_face<NT, SPEC>(sW, N2, B, lim, solver, g, dxl, i, j, ydir, step, L, err, fl);
    if (!ydir) {
#pragma unroll
      for (int k = 0; k < 8; k++) {
        ff[k * N2 + c + 1] = fl[k];
        if (FLDS) sF[k * NF * (NF + 1) + (j - LO) * (NF + 1) + (i + 1 - LO)] = fl[k];
      }
    } else {
#pragma unroll
      for (int k = 0; k < 8; k++) {
        gg[k * N2 + c + NT] = fl[k];
        if (FLDS) sG[k * NF * (NF + 1) + (j + 1 - LO) * NF + (i - LO)] = fl[k];
      }
    }
  }
  __syncthreads();  // also orders this workgroup's ff/gg stores before the reads below

  const double* U = F.f[kU] + t8;
  double* Us = F.f[kUs] + t8;
  double* Uw = F.f[kU] + t8;
  for (int c = tid; c < N2; c += TPB) {
    const int i = c % NT, j = c / NT;
    const bool in = i >= LO && i <= HI && j >= LO && j <= HI;
    double res[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      res[k] = 0.;
      if (in) {
        if (FLDS) {
          const double* Fk = sF + k * NF * (NF + 1);
          const double* Gk = sG + k * NF * (NF + 1);
          const int fi = (j - LO) * (NF + 1) + (i - LO);
          const int gi = (j - LO) * NF + (i - LO);
          res[k] = (Fk[fi + 1] - Fk[fi]) / dxl + (Gk[gi + NF] - Gk[gi]) / dxl;
        } else {
          const double* Fk = ff + k * N2;
          const double* Gk = gg + k * N2;
          res[k] = (Fk[c + 1] - Fk[c]) / dxl + (Gk[c + NT] - Gk[c]) / dxl;
        }
      }
    }
    stage_rk_cell(step, P.rk, dt, res, U, Us, Uw, c, N2);
  }
}

template <int NT, int NG, int TPB, bool FLDS, int WPE, int SPEC = kStageGeneric>
__global__ __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(WPE))) void stage_kernel(
    FieldPtrs F, const LeafItem* __restrict__ leaves, const InstParams* __restrict__ ip, int step, DevError* err,
    unsigned long long* span, const int32_t* __restrict__ counts) {
  using S = StageLds<NT, NG, FLDS>;
  __shared__ double lds[S::kDoubles];
  __shared__ int sNan;
  if (counts && (int)blockIdx.x >= counts[kCountLeaves]) return;
  if (span && threadIdx.x == 0) stamp_start(span);
  stage_block<NT, NG, TPB, FLDS, SPEC>(F, leaves[blockIdx.x], ip, step, err, lds, lds + 8 * S::N2,
                                       lds + 8 * S::N2 + S::NFL, &sNan);
  if (span) {
    __syncthreads();
    if (threadIdx.x == 0) stamp_end(span);
  }
}

// ------------------------------------------------------------ stage strips
// The same stage with each leaf split into K row strips, one workgroup per
// strip, so that small trees still fill the device and a CU holds several
// strips whose load, face and update phases overlap.  The interior rows
// [LO, HI] are divided into K nearly equal parts; the first strip also
// updates the ghost rows below LO, the last those above HI.  A strip forms
// the primitives of its rows and NG rows either side, the x faces of its
// rows and the y faces below and above them -- the y faces on its lower
// edge are also formed by the strip below, with the same operations and so
// the same values, and written to gg by that strip only.  Results equal the
// whole-leaf kernel's bit for bit.  The workgroup of item q = leaf * K +
// strip is launched with id b such that q = (b % 8) * per + b / 8: the
// strips of one leaf run on the same XCD (workgroups are dealt to the 8 XCDs
// round robin) and share the rows they both read in its L2.
template <int NT, int NG, int K>
struct StripGeom {
  static constexpr int LO = NG - 1, HI = NT - 1 - NG, NF = HI - LO + 1;
  static constexpr int kRows = (NF + K - 1) / K;            // interior rows of the largest strip
  static constexpr int kPrimRows = kRows + 2 * NG;           // primitive rows it needs
  static constexpr int kFaces = (2 * kRows + 1) * NF;        // x faces + y faces, at most
  static constexpr int kLanes = (kFaces + 63) / 64 * 64;
  // doubles: primitives, x fluxes (stale edge face + NF per row), y fluxes
  static constexpr int kDoubles = 8 * kPrimRows * NT + 8 * kRows * (NF + 1) + 8 * (kRows + 1) * NF;
  __device__ static int row0(int s) { return LO + (s * NF) / K; }
};

template <int NT, int NG, int K, int SPEC>
__global__ __launch_bounds__((StripGeom<NT, NG, K>::kLanes)) void stage_strip_kernel(
    FieldPtrs F, const LeafItem* __restrict__ leaves, int nleaf, const InstParams* __restrict__ ip, int step,
    DevError* err, unsigned long long* span, const int32_t* __restrict__ counts) {
  using Gm = StripGeom<NT, NG, K>;
  constexpr int N2 = NT * NT, LO = Gm::LO, HI = Gm::HI, NF = Gm::NF, TPB = Gm::kLanes;
  constexpr int PR = Gm::kPrimRows, XR = Gm::kRows;
  __shared__ double lds[Gm::kDoubles];
  __shared__ int sNan;
  if (counts) nleaf = counts[kCountLeaves];
  const int per = (nleaf * K + 7) / 8;
  const int item = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
  if ((int)(blockIdx.x / 8) >= per || item >= nleaf * K) return;
  if (span && threadIdx.x == 0) stamp_start(span);
  const LeafItem L = leaves[item / K];
  const int strip = item % K;
  const int jr0 = Gm::row0(strip), jr1 = Gm::row0(strip + 1) - 1;  // interior rows
  const int r0 = strip == 0 ? 0 : jr0, r1 = strip == K - 1 ? NT - 1 : jr1;  // updated rows
  const int pr0 = jr0 - NG > 0 ? jr0 - NG : 0, pr1 = jr1 + NG < NT - 1 ? jr1 + NG : NT - 1;
  const int ya = jr0 > LO + 1 ? jr0 : LO + 1;  // first computed y-face row (face jj: cells jj - 1, jj)
  const int nx = (jr1 - jr0 + 1) * NF, ny = (jr1 + 1 - ya + 1) * NF;
  double* sW = lds;                     // [k][PR][NT], row pr0 first
  double* sF = lds + 8 * PR * NT;       // [k][XR][NF + 1]: x faces LO .. HI + 1 of rows jr0 ..
  double* sG = sF + 8 * XR * (NF + 1);  // [k][XR + 1][NF]: y faces of rows jr0 .. jr1 + 1
  const double* sWb = sW - pr0 * NT;    // cell (i, j), variable k at sWb[k * PR * NT + j * NT + i]

  const InstParams& P = ip[L.inst];
  const double g = P.gamma, dxl = P.dx[L.lev], dt = P.dt;
  const int lim = SPEC == kStageHLLDWeno5 ? (int)kLimWENO5 : P.limiter;
  const int solver = spec_solver<SPEC>(P.solver);
  const size_t t8 = (size_t)L.gs * 8 * N2, t2 = (size_t)L.gs * 2 * N2;
  const double* Qin = F.f[step == 1 ? kU : kUs] + t8;
  const double* __restrict__ B = F.f[step == 1 ? kBi : kBis] + t2;
  double* ff = F.f[kFF] + t8;
  double* gg = F.f[kGG] + t8;
  const int tid = threadIdx.x;
  if (tid == 0) sNan = 0;
  __syncthreads();

  for (int c = pr0 * NT + tid; c <= pr1 * NT + NT - 1; c += TPB)
    stage_prim_cell<NT, NG>(Qin, g, lim, step, L, err, c, sW + (c - pr0 * NT), PR * NT, &sNan);
  // stale edge fluxes ff[LO][j] of the strip's rows and, in the first strip, gg[i][LO]
  for (int q = tid; q < 8 * (jr1 - jr0 + 1); q += TPB) {
    const int k = q / (jr1 - jr0 + 1), r = q % (jr1 - jr0 + 1);
    sF[k * XR * (NF + 1) + r * (NF + 1)] = ff[k * N2 + (jr0 + r) * NT + LO];
  }
  if (jr0 == LO)
    for (int q = tid; q < 8 * NF; q += TPB) {
      const int k = q / NF, r = q % NF;
      sG[k * (XR + 1) * NF + r] = gg[k * N2 + LO * NT + LO + r];
    }
  __syncthreads();
  if (sNan) {
    // the whole-leaf kernel reports this (and writes nothing) when any
    // primitive of the tile is NaN; the strip holding it reports the same key
    if (tid == 0) report(err, L.inst, error_key(step, L.nb, 1, 0, 0, 0, L.nb, kErrNanEdge));
    return;
  }

  for (int q = tid; q < nx + ny; q += TPB) {
    const int ydir = q >= nx;
    const int r = ydir ? q - nx : q;
    const int i = LO + r % NF;
    // x face (i, i + 1) of row j, or y face jj = ya + r / NF between rows j = jj - 1 and jj
    const int j = ydir ? ya + r / NF - 1 : jr0 + r / NF;
    const int c = j * NT + i;
    double fl[8];
    stage_face<NT, SPEC>(sWb, PR * NT, B, lim, solver, g, dxl, i, j, ydir, step, L, err, fl);
    if (!ydir) {
#pragma unroll
      for (int k = 0; k < 8; k++) {
        ff[k * N2 + c + 1] = fl[k];
        sF[k * XR * (NF + 1) + (j - jr0) * (NF + 1) + (i + 1 - LO)] = fl[k];
      }
    } else {
      const bool own = j >= jr0;  // the face below the strip belongs to the strip below
#pragma unroll
      for (int k = 0; k < 8; k++) {
        if (own) gg[k * N2 + c + NT] = fl[k];
        sG[k * (XR + 1) * NF + (j + 1 - jr0) * NF + (i - LO)] = fl[k];
      }
    }
  }
  __syncthreads();

  const double* U = F.f[kU] + t8;
  double* Us = F.f[kUs] + t8;
  double* Uw = F.f[kU] + t8;
  for (int c = r0 * NT + tid; c <= r1 * NT + NT - 1; c += TPB) {
    const int i = c % NT, j = c / NT;
    const bool in = i >= LO && i <= HI && j >= LO && j <= HI;
    double res[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      res[k] = 0.;
      if (in) {
        const double* Fk = sF + k * XR * (NF + 1);
        const double* Gk = sG + k * (XR + 1) * NF;
        const int fi = (j - jr0) * (NF + 1) + (i - LO);
        const int gi = (j - jr0) * NF + (i - LO);
        res[k] = (Fk[fi + 1] - Fk[fi]) / dxl + (Gk[gi + NF] - Gk[gi]) / dxl;
      }
    }
    stage_rk_cell(step, P.rk, dt, res, U, Us, Uw, c, N2);
  }
  if (span) {
    __syncthreads();
    if (threadIdx.x == 0) stamp_end(span);
  }
}

// --------------------------------------------------------------------- CT
// CT2D.cpp:28-75 (CTtype 2) gas pressure and fast speed of one cell of the
// stage state q (the reference's per-block p and C arrays, formed here per
// corner for the four cells around it, by the same expressions)
__device__ inline void ct2_pc(const double* q, int N2, int c, double g, double& p, double& C) {
  const double q0 = q[c], q1 = q[N2 + c], q2 = q[2 * N2 + c], q3 = q[3 * N2 + c], q4 = q[4 * N2 + c],
               q5 = q[5 * N2 + c], q6 = q[6 * N2 + c], q7 = q[7 * N2 + c];
  p = (g - 1) * (q4 - 0.5 * ((q1 * q1 + q2 * q2 + q3 * q3) / q0 + (q5 * q5 + q6 * q6 + q7 * q7)));
  C = sqrt((g * p + (q5 * q5 + q6 * q6 + q7 * q7)) / q0);
}

// CT2D.cpp:15-75: corner EMF from the face fluxes for i,j in [1,NT-1].
// CTtype 1: the mean of the four face values.  CTtype 2 ("upwind"): the
// y-flux and x-flux pairs weighted 1 - psi and psi, psi = dPx / (dPx + dPy)
// where the pressure jump around the corner exceeds half the smallest
// pressure (sw1) and the flow is compressive beyond a tenth of the smallest
// fast speed (sw2), else psi = 0.5; the velocity divergence keeps the
// reference's terms as written (two y-velocity quotients divide q2 by
// itself, CT2D.cpp:45).  q is the stage state (Us, or U on the last stage).
template <int NT>
__device__ inline void emf_cell(const FieldPtrs& F, const LeafItem L, const InstParams* __restrict__ ip, int last,
                                DevError* err, int step, int c) {
  constexpr int N2 = NT * NT;
  const int i = c % NT, j = c / NT;
  if (i < 1 || j < 1) return;
  const double* gg5 = F.f[kGG] + (size_t)L.gs * 8 * N2 + 5 * N2;
  const double* ff6 = F.f[kFF] + (size_t)L.gs * 8 * N2 + 6 * N2;
  const InstParams& P = ip[L.inst];
  if (P.ct_type != 2) {
    F.f[kEMF][(size_t)L.gs * N2 + c] = 0.25 * ((gg5[c - 1] + gg5[c]) - (ff6[c - NT] + ff6[c]));
    return;
  }
  const double* q = F.f[last ? kU : kUs] + (size_t)L.gs * 8 * N2;
  // cells (i, j), (i, j-1), (i-1, j), (i-1, j-1) of the reference's [i][j]
  const int c00 = c, c0m = c - NT, cm0 = c - 1, cmm = c - 1 - NT;
  double p00, p0m, pm0, pmm, C00, C0m, Cm0, Cmm;
  ct2_pc(q, N2, c00, P.gamma, p00, C00);
  ct2_pc(q, N2, c0m, P.gamma, p0m, C0m);
  ct2_pc(q, N2, cm0, P.gamma, pm0, Cm0);
  ct2_pc(q, N2, cmm, P.gamma, pmm, Cmm);
  const double beta = 0.5, delta = 0.1;
  const double dPx = 0.5 * fabs(p00 + p0m - pm0 - pmm);
  const double dPy = 0.5 * fabs(p00 + pm0 - p0m - pmm);
  const double* q0 = q;
  const double* q1 = q + N2;
  const double* q2 = q + 2 * N2;
  const double divV = 0.5 * (q1[c00] / q0[c00] + q1[c0m] / q0[c0m] - q1[cm0] / q0[cm0] - q1[cmm] / q0[cmm]) +
                      0.5 * (q2[c00] / q0[c00] + q2[cm0] / q2[cm0] - q2[c0m] / q2[c0m] - q2[cmm] / q0[cmm]);
  double tP = pmm < p0m ? pmm : p0m;
  const double tP1 = pm0 < p00 ? pm0 : p00;
  tP = tP < tP1 ? tP : tP1;
  const bool sw1 = (dPx + dPy) > beta * tP;
  double tC = Cmm < C0m ? Cmm : C0m;
  const double tC1 = Cm0 < C00 ? Cm0 : C00;
  tC = tC < tC1 ? tC : tC1;
  const bool sw2 = (-delta * tC) > divV;
  double psi = 0.5;
  if (sw1 && sw2) psi = dPx / (dPx + dPy);
  if (psi < 0 || psi > 1) report(err, L.inst, error_key(step, kSeqCt | L.nb, 0, i, j, 0, L.nb, kErrCtPsi));
  F.f[kEMF][(size_t)L.gs * N2 + c] = 0.5 * (1. - psi) * (gg5[c - 1] + gg5[c]) - 0.5 * psi * (ff6[c - NT] + ff6[c]);
}

template <int NT>
__global__ __launch_bounds__(kThreads) void emf_kernel(FieldPtrs F, const LeafItem* __restrict__ leaves,
                                                       const InstParams* __restrict__ ip, int last, DevError* err,
                                                       int step, const int32_t* __restrict__ counts) {
  if (counts && (int)blockIdx.y >= counts[kCountLeaves]) return;
  const int c = blockIdx.x * kThreads + threadIdx.x;
  if (c < NT * NT) emf_cell<NT>(F, leaves[blockIdx.y], ip, last, err, step, c);
}

// CT2D.cpp:78-92: face-B RK update from the edge EMF.  Stage input B (Bi for
// step 1, else Bis), written array Bn (Bi on the last stage, else Bis) and
// the step's starting value B0 = Bi; SSP-RK3 stages 2-3 are build-defined.
__device__ inline double face_rk(int rk, int step, double B, double Bn, double B0, double dt, double res) {
  if (step == 1) return B - dt * res;
  if (rk == 2) return 0.5 * (B + Bn - dt * res);
  if (step == 2) return 0.75 * B0 + 0.25 * (B - dt * res);
  return (1. / 3.) * B0 + (2. / 3.) * (B - dt * res);
}

template <int NT>
__device__ inline void bupdate_cell(const FieldPtrs& F, const LeafItem L, const InstParams* __restrict__ ip, int step,
                                    int rk, int c) {
  constexpr int N2 = NT * NT;
  const int i = c % NT, j = c / NT;
  const InstParams& P = ip[L.inst];
  const double dt = P.dt, dxl = P.dx[L.lev];
  const double* E = F.f[kEMF] + (size_t)L.gs * N2;
  const double* B = F.f[step == 1 ? kBi : kBis] + (size_t)L.gs * 2 * N2;
  const double* B0 = F.f[kBi] + (size_t)L.gs * 2 * N2;
  double* Bn = F.f[step == rk ? kBi : kBis] + (size_t)L.gs * 2 * N2;
  const int cjp = j < NT - 1 ? c + NT : c;
  const int cip = i < NT - 1 ? c + 1 : c;
  const double rx = (1. / dxl) * (E[cjp] - E[c]);
  const double ry = -(1. / dxl) * (E[cip] - E[c]);
  const double bx = face_rk(rk, step, B[c], Bn[c], B0[c], dt, rx);
  const double by = face_rk(rk, step, B[N2 + c], Bn[N2 + c], B0[N2 + c], dt, ry);
  Bn[c] = bx;
  Bn[N2 + c] = by;
}

template <int NT>
__global__ __launch_bounds__(kThreads) void bupdate_kernel(FieldPtrs F, const LeafItem* __restrict__ leaves,
                                                           const InstParams* __restrict__ ip, int step, int rk,
                                                           const int32_t* __restrict__ counts) {
  if (counts && (int)blockIdx.y >= counts[kCountLeaves]) return;
  const int c = blockIdx.x * kThreads + threadIdx.x;
  if (c < NT * NT) bupdate_cell<NT>(F, leaves[blockIdx.y], ip, step, rk, c);
}

// CT2D.cpp:95-123: cell-centred B from the updated faces + energy correction
template <int NT>
__device__ inline void bsync_cell(const FieldPtrs& F, const LeafItem L, int last, int c) {
  constexpr int N2 = NT * NT;
  const int i = c % NT, j = c / NT;
  const double* Bn = F.f[last ? kBi : kBis] + (size_t)L.gs * 2 * N2;
  double* q = F.f[last ? kU : kUs] + (size_t)L.gs * 8 * N2;
  const int ip = i < NT - 1 ? i + 1 : i, jp = j < NT - 1 ? j + 1 : j;
  const double bx = 0.5 * (Bn[j * NT + ip] + Bn[j * NT + ip - 1]);
  const double by = 0.5 * (Bn[N2 + jp * NT + i] + Bn[N2 + (jp - 1) * NT + i]);
  const double q5 = q[5 * N2 + c], q6 = q[6 * N2 + c];
  q[4 * N2 + c] = q[4 * N2 + c] + 0.5 * ((bx * bx + by * by) - ((q5 * q5) + (q6 * q6)));
  q[5 * N2 + c] = bx;
  q[6 * N2 + c] = by;
}

template <int NT>
__global__ __launch_bounds__(kThreads) void bsync_kernel(FieldPtrs F, const LeafItem* __restrict__ leaves, int last,
                                                         const int32_t* __restrict__ counts) {
  if (counts && (int)blockIdx.y >= counts[kCountLeaves]) return;
  const int c = blockIdx.x * kThreads + threadIdx.x;
  if (c < NT * NT) bsync_cell<NT>(F, leaves[blockIdx.y], last, c);
}

// -------------------------------------------------------------- ghost fill
// Single-pass execution of a compiled fill program.  Every destination's value
// is a function of the pre-fill state only; a workgroup evaluates all values
// of its chunk of ops, synchronises, then writes them.  The compiler places
// each op that reads an element some other op writes in the same chunk as
// that writer, so reads always see pre-fill values.
struct FillCtx {
  const double* base;
  int nv, kind, N2, NT;
  const FillGen* gen;
};

__device__ inline double fill_elem(const FillCtx& f, int64_t gs, int cell, int k) {
  return f.base[((size_t)gs * f.nv + k) * f.N2 + cell];
}

// explicit operand of a general restriction (always a pre-fill element)
__device__ inline double fill_ref(const FillCtx& f, int64_t ref, int k) {
  return fill_elem(f, ref >> 16, (int)(ref & 0xffff), k);
}

// boundary.cpp (0.25 x 2x2 sum in misc.cpp:35-48 order), boundaryfc.cpp
// (0.5 x face pair: along j for component 0, along i otherwise) and
// boundaryec.cpp (injection); v[] holds (i1,j1),(i1,j1+1),(i1+1,j1),(i1+1,j1+1)
__device__ inline double restrict_combine(int kind, int k, double v00, double v01, double v10, double v11) {
  double v = 0.;
  if (kind == kFillCell) {
    v += v00;
    v += v01;
    v += v10;
    v += v11;
    return 0.25 * v;
  }
  if (kind == kFillFace) {
    v += v00;
    v += k == 0 ? v01 : v10;
    return 0.5 * v;
  }
  return v00;
}

__device__ inline double gen_value(const FillCtx& f, const FillGen& g, int k) {
  return restrict_combine(f.kind, k, fill_ref(f, g.ref[0], k), fill_ref(f, g.ref[1], k), fill_ref(f, g.ref[2], k),
                          fill_ref(f, g.ref[3], k));
}

__device__ inline double fill_value(const FillCtx& f, const FillOp& op, int k) {
  if (op.op == kFillCopy) return fill_elem(f, op.src, op.src_cell, k);
  if (op.op == kFillRestrict) {
    const int a = op.src_cell;
    return restrict_combine(f.kind, k, fill_elem(f, op.src, a, k), fill_elem(f, op.src, a + f.NT, k),
                            fill_elem(f, op.src, a + 1, k), fill_elem(f, op.src, a + 1 + f.NT, k));
  }
  return gen_value(f, f.gen[op.src], k);
}

// The program is a concatenation of per-instance segments of whole chunks:
// chunk w belongs to the instance q with cprefix[q] <= w < cprefix[q+1].
struct FillProgram {
  const FillOp* const* ops;
  const FillGen* const* gen;
  const int32_t* cprefix;
  int ninst;
};

__device__ inline int find_segment(const FillProgram& P, int w) {
  int lo = 0, hi = P.ninst;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (P.cprefix[mid] <= w) lo = mid;
    else hi = mid;
  }
  return lo;
}

template <int NV>
__device__ inline void fill_field(const FillCtx& f, double* base, const FillOp& op, bool live) {
  double v[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = live ? fill_value(f, op, k) : 0.;
  __syncthreads();  // every read of the chunk precedes every write
  if (live) {
#pragma unroll
    for (int k = 0; k < NV; k++) base[((size_t)op.dst_gs * NV + k) * f.N2 + op.dst_cell] = v[k];
  }
}

// One chunk of the program per kFillChunk lanes (`lane` within the chunk); a
// workgroup may run several chunks side by side, and every lane of it calls
// this together (it synchronises the workgroup).  valid = false: no chunk.
__device__ inline void fill_chunk(const FillProgram& P, const FillSet& set, int NT, int w, int lane, bool valid) {
  FillOp op{};
  op.op = kFillNop;
  int q = 0;
  if (valid) {
    q = find_segment(P, w);
    op = P.ops[q][(w - P.cprefix[q]) * kFillChunk + lane];
  }
  const bool live = op.op != kFillNop;
  for (int fi = 0; fi < set.n; fi++) {
    const FillCtx f{set.base[fi], set.nv[fi], set.kind[fi], NT * NT, NT, valid ? P.gen[q] : nullptr};
    if (set.nv[fi] == 8) fill_field<8>(f, set.base[fi], op, live);
    else if (set.nv[fi] == 2) fill_field<2>(f, set.base[fi], op, live);
    else fill_field<1>(f, set.base[fi], op, live);
  }
}

// restrict_combine for a kind fixed at compile time (the same IEEE
// operations, no branch on the kind between an op's loads)
template <int KIND>
__device__ inline double restrict_kind(int k, double v00, double v01, double v10, double v11) {
  double v = 0.;
  if constexpr (KIND == kFillCell) {
    v += v00;
    v += v01;
    v += v10;
    v += v11;
    return 0.25 * v;
  } else if constexpr (KIND == kFillFace) {
    v += v00;
    v += k == 0 ? v01 : v10;
    return 0.5 * v;
  } else {
    return v00;
  }
}

// the NV components of one field of combine kind KIND for an op whose
// operands are the elements e[t] (block slot << 16 | cell): one (copy) or
// four (restriction, 2x2 order)
template <int NV, int KIND, bool COPY>
__device__ inline void fill_gather(const double* __restrict__ base, int N2, const int64_t* e, double* v) {
  if constexpr (COPY) {
    const double* p = base + (size_t)(e[0] >> 16) * NV * N2 + (e[0] & 0xffff);
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = p[(size_t)k * N2];
  } else {
    const double* p0 = base + (size_t)(e[0] >> 16) * NV * N2 + (e[0] & 0xffff);
    const double* p1 = base + (size_t)(e[1] >> 16) * NV * N2 + (e[1] & 0xffff);
    const double* p2 = base + (size_t)(e[2] >> 16) * NV * N2 + (e[2] & 0xffff);
    const double* p3 = base + (size_t)(e[3] >> 16) * NV * N2 + (e[3] & 0xffff);
#pragma unroll
    for (int k = 0; k < NV; k++)
      v[k] = restrict_kind<KIND>(k, p0[(size_t)k * N2], p1[(size_t)k * N2], p2[(size_t)k * N2], p3[(size_t)k * N2]);
  }
}

// The same for a set of up to three fields with NV0, NV1, NV2 components
// (0: absent) and combine kinds K0, K1, K2, fixed at compile time: every value
// of every field of the chunk is read before the one barrier, then all are
// written.  The op's kind and operand elements are decoded once, outside the
// loops over fields and components, so that all of an op's loads are in
// flight together (branches per component made each component's loads wait
// for the previous ones).
template <int NV0, int NV1, int NV2, int K0, int K1, int K2>
__device__ inline void fill_chunk_nv(const FillProgram& P, const FillSet& set, int NT, int w, int lane, bool valid) {
  FillOp op{};
  op.op = kFillNop;
  int q = 0;
  if (valid) {
    q = find_segment(P, w);
    op = P.ops[q][(w - P.cprefix[q]) * kFillChunk + lane];
  }
  const bool live = op.op != kFillNop;
  const int N2 = NT * NT;
  double v0[NV0 > 0 ? NV0 : 1], v1[NV1 > 0 ? NV1 : 1], v2[NV2 > 0 ? NV2 : 1];
  if (live) {
    int64_t e[4];
    if (op.op == kFillCopy) {
      e[0] = src_ref(op.src, op.src_cell);
      if constexpr (NV0 > 0) fill_gather<NV0, K0, true>(set.base[0], N2, e, v0);
      if constexpr (NV1 > 0) fill_gather<NV1, K1, true>(set.base[1], N2, e, v1);
      if constexpr (NV2 > 0) fill_gather<NV2, K2, true>(set.base[2], N2, e, v2);
    } else {
      if (op.op == kFillRestrict) {
        const int a = op.src_cell;
        e[0] = src_ref(op.src, a);
        e[1] = src_ref(op.src, a + NT);
        e[2] = src_ref(op.src, a + 1);
        e[3] = src_ref(op.src, a + 1 + NT);
      } else {
        const FillGen& g = P.gen[q][op.src];
#pragma unroll
        for (int t = 0; t < 4; t++) e[t] = g.ref[t];
      }
      if constexpr (NV0 > 0) fill_gather<NV0, K0, false>(set.base[0], N2, e, v0);
      if constexpr (NV1 > 0) fill_gather<NV1, K1, false>(set.base[1], N2, e, v1);
      if constexpr (NV2 > 0) fill_gather<NV2, K2, false>(set.base[2], N2, e, v2);
    }
  }
  __syncthreads();  // every read of the chunk precedes every write
  if (live) {
#pragma unroll
    for (int k = 0; k < NV0; k++) set.base[0][((size_t)op.dst_gs * NV0 + k) * N2 + op.dst_cell] = v0[k];
#pragma unroll
    for (int k = 0; k < NV1; k++) set.base[1][((size_t)op.dst_gs * NV1 + k) * N2 + op.dst_cell] = v1[k];
#pragma unroll
    for (int k = 0; k < NV2; k++) set.base[2][((size_t)op.dst_gs * NV2 + k) * N2 + op.dst_cell] = v2[k];
  }
}

// field sets of the step (engine.cpp) with their combine kinds: {U, Bi},
// {Bi, ff, gg}, {ff, gg}, one cell field, one face field, EMF; any other set
// takes the per-field path
template <int NV0, int NV1, int NV2, int K0, int K1, int K2>
__global__ __launch_bounds__(kThreads) void fill_kernel_nv(FillProgram P, FillSet set, int NT,
                                                           const int32_t* __restrict__ counts) {
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < set.copy_words; i += gridDim.x * kThreads)
    set.copy_dst[i] = set.copy_src[i];
  if (counts && (int)blockIdx.x >= counts[kCountChunks]) return;
  fill_chunk_nv<NV0, NV1, NV2, K0, K1, K2>(P, set, NT, blockIdx.x, threadIdx.x, true);
}

__global__ __launch_bounds__(kThreads) void fill_kernel(FillProgram P, FillSet set, int NT,
                                                        const int32_t* __restrict__ counts) {
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < set.copy_words; i += gridDim.x * kThreads)
    set.copy_dst[i] = set.copy_src[i];
  if (counts && (int)blockIdx.x >= counts[kCountChunks]) return;
  fill_chunk(P, set, NT, blockIdx.x, threadIdx.x, true);
}

// ------------------------------------------------------ CFL and criteria
// main.cpp:209-230 (per-leaf wave-speed bound, with the reference's
// |rho|+c_fx and |u|+c_fy terms) fused with criteria.cpp:6-44 (maximum of
// the max4diff/|w| ratio over p, rho, Bx, By on the interior).  One leaf, all
// TPB lanes (a power of two); sW: 8 NT^2 doubles and sRed: 2 TPB doubles of
// the workgroup's LDS.  The maxima never hold a NaN (std_max keeps its first
// operand against one), so the reduction order does not change them.
template <int NT, int NG, int TPB>
__device__ inline void cfl_block(const FieldPtrs& F, const double* __restrict__ Win, const LeafItem L, int leaf,
                                 const InstParams* __restrict__ ip, double* lam_grad, DevError* err, double* sW,
                                 double* sRed) {
  constexpr int N2 = NT * NT;
  constexpr int LO = NG, HI = NT - 1 - NG;
  const double g = ip[L.inst].gamma;
  const int tid = threadIdx.x;
  if (Win) {
    const double* W = Win + (size_t)leaf * 8 * N2;
    for (int q = tid; q < 8 * N2; q += TPB) sW[q] = W[q];
  } else {
    const double* U = F.f[kU] + (size_t)L.gs * 8 * N2;
    for (int c = tid; c < N2; c += TPB) {
      double u[8], w[8];
#pragma unroll
      for (int k = 0; k < 8; k++) u[k] = U[k * N2 + c];
      cons_to_prim(u, g, w);
#pragma unroll
      for (int k = 0; k < 8; k++) sW[k * N2 + c] = w[k];
      const int i = c % NT, j = c / NT;
      if (w[4] <= 0 && i > LO - 1 && i < HI + 1 && j > LO - 1 && j < HI + 1)
        report(err, L.inst, error_key(kOrdCfl, L.nb, 0, i, j, 0, L.nb, kErrPressure));
    }
  }
  __syncthreads();
  double lam = 0., grad = 0.;
  constexpr int NI = HI - LO + 1;
  for (int q = tid; q < NI * NI; q += TPB) {
    const int i = LO + q % NI, j = LO + q / NI, c = j * NT + i;
    double w[8];
#pragma unroll
    for (int k = 0; k < 8; k++) w[k] = sW[k * N2 + c];
    if (!Win) {
      lam = std_max(lam, fabs(w[0]) + fast_speed(w, g, 5));
      lam = std_max(lam, fabs(w[1]) + fast_speed(w, g, 6));
    }
    const int vars[4] = {4, 0, 5, 6};
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const double* Wk = sW + vars[v] * N2;
      const double x = Wk[c];
      double m = fabs(Wk[c + 1] - x);
      m = mac_max(m, fabs(Wk[c - 1] - x));
      m = mac_max(m, fabs(Wk[c + NT] - x));
      m = mac_max(m, fabs(Wk[c - NT] - x));
      grad = std_max(grad, m / mac_max(kEps, fabs(x)));
    }
  }
  sRed[tid] = lam;
  sRed[TPB + tid] = grad;
  __syncthreads();
  for (int w = TPB / 2; w > 0; w >>= 1) {
    if (tid < w) {
      sRed[tid] = std_max(sRed[tid], sRed[tid + w]);
      sRed[TPB + tid] = std_max(sRed[TPB + tid], sRed[TPB + tid + w]);
    }
    __syncthreads();
  }
  if (tid == 0) {
    const double dxl = ip[L.inst].dx[L.lev];
    lam_grad[2 * leaf] = std_min(dxl, dxl) / sRed[0];
    lam_grad[2 * leaf + 1] = sRed[TPB];
  }
}

template <int NT, int NG>
__global__ __launch_bounds__(kThreads) void cfl_criteria_kernel(FieldPtrs F, const double* __restrict__ Win,
                                                                const LeafItem* __restrict__ leaves,
                                                                const InstParams* __restrict__ ip, double* lam_grad,
                                                                DevError* err, const int32_t* __restrict__ counts) {
  __shared__ double sW[8 * NT * NT];
  __shared__ double sRed[2 * kThreads];
  if (counts && (int)blockIdx.x >= counts[kCountLeaves]) return;
  cfl_block<NT, NG, kThreads>(F, Win, leaves[blockIdx.x], blockIdx.x, ip, lam_grad, err, sW, sRed);
}

// ----------------------------------------------------- persistent step kernel
// Grid-wide barrier of a co-resident grid of nwg workgroups.  Lane 0 of each
// workgroup publishes the workgroup's writes (agent-scope release), arrives,
// and waits for the generation to move; the last to arrive resets the count
// and moves it.  The count is back at zero after every barrier, so the state
// serves every launch without a reset.  A wait is bounded by the device clock
// (100 MHz): past kBarrierTicks it sets *fail, and a set *fail makes every
// later wait return at once, so every workgroup reaches the end of the
// launch (the host then reports the failure and stops using this kernel).
constexpr unsigned long long kBarrierTicks = 200000000ull;  // 2 s

__device__ inline void grid_barrier(GridSync* b, unsigned nwg, int32_t* fail) {
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned g = __hip_atomic_load(&b->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const unsigned old = __hip_atomic_fetch_add(&b->arrive, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (old == nwg - 1) {
      __hip_atomic_store(&b->arrive, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&b->gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      const unsigned long long t0 = wall_clock64();
      while (__hip_atomic_load(&b->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) {
        if (__hip_atomic_load(fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
        if (wall_clock64() - t0 > kBarrierTicks) {
          __hip_atomic_store(fail, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

template <int NT, int NG, int TPB>
struct StepLds {
  using S = StageLds<NT, NG, true>;
  static constexpr int kCfl = 8 * NT * NT + 2 * TPB;
  static constexpr int kDoubles = S::kDoubles > kCfl ? S::kDoubles : kCfl;
};

template <int NT, int NG, int TPB, int SPEC>
__global__ __launch_bounds__(TPB) void step_kernel(StepArgs A) {
  using S = StageLds<NT, NG, true>;
  constexpr int N2 = NT * NT;
  __shared__ double lds[StepLds<NT, NG, TPB>::kDoubles];
  __shared__ int sNan;
  const int tid = threadIdx.x;
  const unsigned nwg = gridDim.x;
  const FillProgram prog{A.ops, A.gen, A.cprefix, A.ninst};
  if (A.stamps && blockIdx.x == 0 && tid == 0) A.stamps[0] = wall_clock64();
  for (int p = 0; p < A.nphases; p++) {
    const StepPhase ph = A.phase[p];
    if (ph.kind == kPhaseFill) {
      if (p == 0)  // the step's parameters (no fill op reads them)
        for (int i = blockIdx.x * TPB + tid; i < A.copy_words; i += nwg * TPB) A.copy_dst[i] = A.copy_src[i];
      // TPB / kFillChunk chunks side by side per workgroup
      constexpr int kPer = TPB / kFillChunk;
      const int lane = tid % kFillChunk, sub = tid / kFillChunk;
      for (int w0 = blockIdx.x * kPer; w0 < A.nchunks; w0 += nwg * kPer) {
        const int w = w0 + sub;
        FillOp op{};
        op.op = kFillNop;
        int q = 0;
        if (w < A.nchunks) {
          q = find_segment(prog, w);
          op = prog.ops[q][(w - prog.cprefix[q]) * kFillChunk + lane];
        }
        const bool live = op.op != kFillNop;
        const FillGen* gen = w < A.nchunks ? prog.gen[q] : nullptr;
        for (int f = 0; f < ph.nf; f++) {
          const int fld = ph.field[f];
          double* base = A.F.f[fld];
          if (fld == kBi || fld == kBis) fill_field<2>(FillCtx{base, 2, ph.fkind[f], N2, NT, gen}, base, op, live);
          else if (fld == kEMF) fill_field<1>(FillCtx{base, 1, ph.fkind[f], N2, NT, gen}, base, op, live);
          else fill_field<8>(FillCtx{base, 8, ph.fkind[f], N2, NT, gen}, base, op, live);
        }
      }
    } else if (ph.kind == kPhaseStage) {
      for (int l = blockIdx.x; l < A.nleaf; l += nwg) {
        __syncthreads();  // the previous leaf is done with the LDS
        stage_block<NT, NG, TPB, true, SPEC>(A.F, A.leaves[l], A.ip, ph.step, A.err, lds, lds + 8 * N2,
                                             lds + 8 * N2 + S::NFL, &sNan);
      }
    } else if (ph.kind == kPhaseCfl) {
      for (int l = blockIdx.x; l < A.nleaf; l += nwg) {
        __syncthreads();
        cfl_block<NT, NG, TPB>(A.F, nullptr, A.leaves[l], l, A.ip, A.lam_grad, A.err, lds, lds + 8 * N2);
      }
    } else {
      // CT kernels: one lane per cell of every leaf
      const int total = A.nleaf * N2;
      for (int e = blockIdx.x * TPB + tid; e < total; e += nwg * TPB) {
        const LeafItem L = A.leaves[e / N2];
        const int c = e % N2;
        if (ph.kind == kPhaseEmf) emf_cell<NT>(A.F, L, A.ip, ph.step == A.ip[L.inst].rk, A.err, ph.step, c);
        else if (ph.kind == kPhaseBupdate) bupdate_cell<NT>(A.F, L, A.ip, ph.step, A.ip[L.inst].rk, c);
        else bsync_cell<NT>(A.F, L, ph.step, c);
      }
    }
    if (p + 1 < A.nphases) grid_barrier(A.sync, nwg, A.fail);
    if (A.stamps && blockIdx.x == 0 && tid == 0) A.stamps[p + 1] = wall_clock64();
  }
}

// ------------------------------------------------------------------ regrid
// refine.cpp:17-36: minmod-limited quarter-slope prolongation of one value
__device__ inline double prolong(const double* Q, int NT, int i, int j, int sx, int sy) {
  const int c = j * NT + i;
  double v = Q[c];
  double r = (Q[c] - Q[c - 1]) / (Q[c + 1] - Q[c] + kEps);
  double phi = 0.25 * limiter(kLimMinmod, r, 0) * (Q[c + 1] - Q[c]);
  v = sx == 0 ? v - phi : v + phi;
  r = (Q[c] - Q[c - NT]) / (Q[c + NT] - Q[c] + kEps);
  phi = 0.25 * limiter(kLimMinmod, r, 0) * (Q[c + NT] - Q[c]);
  v = sy == 0 ? v - phi : v + phi;
  return v;
}

// refineBi_CT.cpp:4-50: quadratic face reconstruction sampled at the son face
__device__ inline double prolong_face(const double* B, int NT, int N2, int i, int j, int k, int signx, int signy) {
  const double* Bx = B;
  const double* By = B + N2;
  const int c = j * NT + i;
  const double x_jm = Bx[c - NT], x0 = Bx[c], x_jp = Bx[c + NT];
  const double x_ip = Bx[c + 1], x_ipjm = Bx[c + 1 - NT], x_ipjp = Bx[c + 1 + NT];
  const double y_im = By[c - 1], y0 = By[c], y_ip = By[c + 1];
  const double y_imjp = By[c - 1 + NT], y_jp = By[c + NT], y_ipjp = By[c + 1 + NT];
  const double DyBxm = limiter(kLimMinmod, (x0 - x_jm) / (x_jp - x0 + kEps), 0) * (x_jp - x0);
  const double DyBxp = limiter(kLimMinmod, (x_ip - x_ipjm) / (x_ipjp - x_ip + kEps), 0) * (x_ipjp - x_ip);
  const double DxBym = limiter(kLimMinmod, (y0 - y_im) / (y_ip - y0 + kEps), 0) * (y_ip - y0);
  const double DxByp = limiter(kLimMinmod, (y_jp - y_imjp) / (y_ipjp - y_jp + kEps), 0) * (y_ipjp - y_jp);
  const double ax = (x_ip - x0), by = -1. * ax;
  const double ay = 0.5 * (DyBxp + DyBxm), bx = 0.5 * (DxByp + DxBym);
  const double axy = (DyBxp - DyBxm), byy = axy / (-2.);
  const double bxy = (DxByp - DxBym), axx = bxy / (-2.);
  const double ayy = 0., bxx = 0.;
  if (k == 0) {
    const double a0 = 0.5 * (x_ip + x0) - 0.25 * axx;
    const double sx = signx == 0 ? 0 : 0.5, sy = signy == 0 ? -0.25 : 0.25;
    return a0 + ax * sx + ay * sy + axx * sx * sx + axy * sx * sy + ayy * sy * sy;
  }
  const double b0 = 0.5 * (y_jp + y0) - 0.25 * byy;
  const double sx = signx == 0 ? -0.25 : 0.25, sy = signy == 0 ? 0 : 0.5;
  return b0 + bx * sx + by * sy + bxx * sx * sx + bxy * sx * sy + byy * sy * sy;
}

// refine.cpp:38-59 + refineBi_CT.cpp:52-74 (+ the sons' U2W p<=0 check)
template <int NT, int NG>
__global__ __launch_bounds__(kThreads) void refine_kernel(FieldPtrs F, const RegridOp* __restrict__ ops,
                                                          const InstParams* __restrict__ ip, DevError* err) {
  constexpr int N2 = NT * NT, NX2 = (NT - 1) / 2;
  const RegridOp op = ops[blockIdx.y];
  const int q = blockIdx.z;  // son 0..3
  const int c = blockIdx.x * kThreads + threadIdx.x;
  if (c >= N2) return;
  const int i = c % NT, j = c / NT;
  const int sx = (i + NG) % 2, sy = (j + NG) % 2;
  const int pi = (q & 1) ? (i - NG + 2) / 2 + NX2 : (i + NG) / 2;
  const int pj = (q & 2) ? (j - NG + 2) / 2 + NX2 : (j + NG) / 2;
  const double* Up = F.f[kU] + (size_t)op.parent_gs * 8 * N2;
  double* Us = F.f[kU] + (size_t)op.son_gs[q] * 8 * N2;
  double u[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    u[k] = prolong(Up + k * N2, NT, pi, pj, sx, sy);
    Us[k * N2 + c] = u[k];
  }
  double w[8];
  cons_to_prim(u, ip[op.inst].gamma, w);
  // refine.cpp:69-83: U2W of son 1..4 in turn; then refineBi_CT.cpp:59-71
  // over (i, j, k) with the four sons innermost
  const int seq = op.seq;
  if (w[4] <= 0 && i > NG - 1 && i < NT - NG && j > NG - 1 && j < NT - NG)
    report(err, op.inst, error_key(kOrdRefine, seq, q, i, j, 0, op.son_nb + q, kErrPressure));
  const double* Bp = F.f[kBi] + (size_t)op.parent_gs * 2 * N2;
  double* Bs = F.f[kBi] + (size_t)op.son_gs[q] * 2 * N2;
  for (int k = 0; k < 2; k++) {
    const double v = prolong_face(Bp, NT, N2, pi, pj, k, sx, sy);
    if (isnan(v)) report(err, op.inst, error_key(kOrdRefine, seq, 4, i, j, 4 * k + q, op.nb, kErrRefineNan));
    Bs[k * N2 + c] = v;
  }
}

// coarsen.cpp:7-157 + coarsenBi_CT.cpp:7-155 through precomputed per-cell
// source maps (last write wins): map = son | i1 << 8 | j1 << 16, -1 = untouched
template <int NT, int NG>
__global__ __launch_bounds__(kThreads) void coarsen_kernel(FieldPtrs F, const RegridOp* __restrict__ ops,
                                                           const int32_t* __restrict__ cmapU,
                                                           const int32_t* __restrict__ cmapB,
                                                           const InstParams* __restrict__ ip, DevError* err) {
  constexpr int N2 = NT * NT;
  const RegridOp op = ops[blockIdx.y];
  const int c = blockIdx.x * kThreads + threadIdx.x;
  if (c >= N2) return;
  const int i = c % NT, j = c / NT;
  double* Ud = F.f[kU] + (size_t)op.parent_gs * 8 * N2;
  const int m = cmapU[c];
  if (m >= 0) {
    const int s = m & 0xff, i1 = (m >> 8) & 0xff, j1 = (m >> 16) & 0xff;
    const double* Uq = F.f[kU] + (size_t)op.son_gs[s] * 8 * N2;
    const int a = j1 * NT + i1;
    double u[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const double* Q = Uq + k * N2;
      double v = 0.;
      v += Q[a];
      v += Q[a + NT];
      v += Q[a + 1];
      v += Q[a + 1 + NT];
      u[k] = 0.25 * v;
      Ud[k * N2 + c] = u[k];
    }
    double w[8];
    cons_to_prim(u, ip[op.inst].gamma, w);
    if (w[4] <= 0 && i > NG - 1 && i < NT - NG && j > NG - 1 && j < NT - NG)
      report(err, op.inst, error_key(kOrdCoarsen, op.seq, 0, i, j, 0, op.nb, kErrPressure));
  }
  double* Bd = F.f[kBi] + (size_t)op.parent_gs * 2 * N2;
  for (int k = 0; k < 2; k++) {
    const int mb = cmapB[k * N2 + c];
    if (mb < 0) continue;
    const int s = mb & 0xff, i1 = (mb >> 8) & 0xff, j1 = (mb >> 16) & 0xff;
    const double* Q = F.f[kBi] + (size_t)op.son_gs[s] * 2 * N2 + k * N2;
    const int a = j1 * NT + i1;
    double v = 0.;
    v += Q[a];
    v += Q[k == 0 ? a + NT : a + 1];
    Bd[k * N2 + c] = 0.5 * v;
  }
}

// ----------------------------------------------------- parity entry points
__global__ void riemann_batch_kernel(int solver, int dir, double g, int n, const double* wL, const double* wR,
                                     double* flux, int32_t* status) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  double l[8], r[8], f[8];
  for (int k = 0; k < 8; k++) {
    l[k] = wL[8 * q + k];
    r[k] = wR[8 * q + k];
  }
  if (dir == 1) {
    swap_xy(l);
    swap_xy(r);
  }
  status[q] = normal_flux(solver, l, r, g, f);
  if (dir == 1) swap_xy(f);
  for (int k = 0; k < 8; k++) flux[8 * q + k] = f[k];
}

// edge states of whole tiles: out = [tile][4: xL xR yL yR][8][NT][NT]
// (celledges.cpp ranges for MUSCL, WENO2D_primvar.cpp ranges for WENO)
__global__ void recon_batch_kernel(int NT, int ng, int lim, double dxl, const double* W, double* out) {
  const int N2 = NT * NT;
  const int t = blockIdx.y;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N2) return;
  const int i = c % NT, j = c / NT;
  const double* Wt = W + (size_t)t * 8 * N2;
  double* o = out + (size_t)t * 4 * 8 * N2;
  const bool weno = lim == kLimWENO3 || lim == kLimWENO5;
  const int lo = weno ? ng - 1 : 0, hi = weno ? NT - ng : NT - 2;
  if (i < lo || i > hi || j < lo || j > hi) return;
  for (int k = 0; k < 8; k++) {
    const double* Wk = Wt + k * N2;
    for (int dir = 0; dir < 2; dir++) {
      if ((dir == 0 && k == 5) || (dir == 1 && k == 6)) continue;
      const int s = dir == 0 ? 1 : NT;
      double L, R;
      if (lim == kLimWENO5) {
        L = weno5_edge(Wk[c - 2 * s], Wk[c - s], Wk[c], Wk[c + s], Wk[c + 2 * s], true);
        R = weno5_edge(Wk[c - 2 * s], Wk[c - s], Wk[c], Wk[c + s], Wk[c + 2 * s], false);
      } else if (lim == kLimWENO3) {
        L = weno3_edge(Wk[c - s], Wk[c], Wk[c + s], true);
        R = weno3_edge(Wk[c - s], Wk[c], Wk[c + s], false);
      } else {
        const int idx = dir == 0 ? i : j;
        const int cm = idx > 0 ? c - s : c;  // backward difference, index 0 duplicates 1
        const double d0 = idx > 0 ? Wk[c] - Wk[cm] : Wk[c + s] - Wk[c];
        const double d1 = Wk[c + s] - Wk[c];
        L = muscl_plus(lim, Wk[c], d0, d1, dxl);
        R = muscl_minus(lim, Wk[c], d0, d1, dxl);
      }
      o[((dir * 2 + 0) * 8 + k) * N2 + c] = L;
      o[((dir * 2 + 1) * 8 + k) * N2 + c] = R;
    }
  }
}

}  // namespace

// ---------------------------------------------------------------- launchers
bool tile_supported(int NT, int ng) { return (NT == 20 && ng == 2) || (NT == 24 && ng == 4) || (NT == 12 && ng == 2); }

#define AMR_DISPATCH_TILE(NT, NG, CALL)                                    \
  do {                                                                     \
    if (NT == 20 && NG == 2) { constexpr int T = 20, G = 2; (void)G; CALL; }        \
    else if (NT == 24 && NG == 4) { constexpr int T = 24, G = 4; (void)G; CALL; }   \
    else if (NT == 12 && NG == 2) { constexpr int T = 12, G = 2; (void)G; CALL; }   \
  } while (0)

template <int T, int G, int TPB, bool FLDS, int WPE, int SPEC>
void launch_stage_variant(const FieldPtrs& F, const LeafItem* leaves, int nleaf, const InstParams* ip, int step,
                          DevError* err, hipStream_t s, hipEvent_t ev_start, hipEvent_t ev_stop,
                          unsigned long long* span, const int32_t* counts) {
  if (ev_start && ev_stop) {
    // events attached to the dispatch itself: they time the kernel, not the launch
    hipExtLaunchKernelGGL((stage_kernel<T, G, TPB, FLDS, WPE, SPEC>), dim3(nleaf), dim3(TPB), 0, s, ev_start,
                          ev_stop, 0, F, leaves, ip, step, err, span, counts);
    return;
  }
  stage_kernel<T, G, TPB, FLDS, WPE, SPEC><<<nleaf, TPB, 0, s>>>(F, leaves, ip, step, err, span, counts);
}

template <int T, int G, int K, int SPEC>
void launch_stage_strips(const FieldPtrs& F, const LeafItem* leaves, int nleaf, const InstParams* ip, int step,
                         DevError* err, hipStream_t s, hipEvent_t ev_start, hipEvent_t ev_stop,
                         unsigned long long* span, const int32_t* counts) {
  constexpr int TPB = StripGeom<T, G, K>::kLanes;
  const unsigned grid = 8u * (unsigned)((nleaf * K + 7) / 8);
  if (ev_start && ev_stop) {
    hipExtLaunchKernelGGL((stage_strip_kernel<T, G, K, SPEC>), dim3(grid), dim3(TPB), 0, s, ev_start, ev_stop, 0, F,
                          leaves, nleaf, ip, step, err, span, counts);
    return;
  }
  stage_strip_kernel<T, G, K, SPEC><<<grid, TPB, 0, s>>>(F, leaves, nleaf, ip, step, err, span, counts);
}

// WENO5 needs nghosts >= 3: the 24x24 tile only
#define AMR_DISPATCH_TILE_WENO5(NT, NG, CALL)                         \
  do {                                                                \
    if (NT == 24 && NG == 4) { constexpr int T = 24, G = 4; CALL; }   \
  } while (0)

template <int SPEC>
void launch_stage_spec(int NT, int ng, int variant, const FieldPtrs& F, const LeafItem* leaves, int nleaf,
                       const InstParams* ip, int step, DevError* err, hipStream_t s, hipEvent_t ev_start,
                       hipEvent_t ev_stop, unsigned long long* span, const int32_t* counts) {
#define AMR_STAGE_CALL (launch_stage_variant<T, G, TPB, FLDS, WPE, SPEC>(F, leaves, nleaf, ip, step, err, s, \
                                                                         ev_start, ev_stop, span, counts))
#define AMR_STAGE(TPB_, FLDS_, WPE_)                                           \
  do {                                                                         \
    constexpr int TPB = TPB_, WPE = WPE_;                                      \
    constexpr bool FLDS = FLDS_;                                               \
    if constexpr (SPEC == kStageHLLDWeno5) AMR_DISPATCH_TILE_WENO5(NT, ng, AMR_STAGE_CALL); \
    else AMR_DISPATCH_TILE(NT, ng, AMR_STAGE_CALL);                            \
  } while (0)
#define AMR_STRIPS_CALL (launch_stage_strips<T, G, K, SPEC>(F, leaves, nleaf, ip, step, err, s, ev_start, ev_stop, \
                                                            span, counts))
#define AMR_STRIPS(K_)                                                          \
  do {                                                                          \
    constexpr int K = K_;                                                       \
    if constexpr (SPEC == kStageHLLDWeno5) AMR_DISPATCH_TILE_WENO5(NT, ng, AMR_STRIPS_CALL); \
    else AMR_DISPATCH_TILE(NT, ng, AMR_STRIPS_CALL);                            \
  } while (0)
  switch (variant) {
    case 6: AMR_STRIPS(2); break;
    case 7: AMR_STRIPS(4); break;
    case 1: AMR_STAGE(512, true, 1); break;
    case 2: AMR_STAGE(768, true, 1); break;
    case 3: AMR_STAGE(256, false, 4); break;
    case 4: AMR_STAGE(512, false, 4); break;
    case 5: AMR_STAGE(512, true, 4); break;
    default: AMR_STAGE(256, true, 1); break;
  }
#undef AMR_STAGE
#undef AMR_STAGE_CALL
#undef AMR_STRIPS
#undef AMR_STRIPS_CALL
}

void launch_stage(int NT, int ng, int variant, int spec, const FieldPtrs& F, const LeafItem* leaves, int nleaf,
                  const InstParams* ip, int step, DevError* err, hipStream_t s, hipEvent_t ev_start,
                  hipEvent_t ev_stop, unsigned long long* span, const int32_t* counts) {
  if (nleaf <= 0) return;
  if (spec == kStageHLLCMuscl && variant == kSplitVariant) {
    // the HLLC part of a launch split by specialisation: one variant
    constexpr int SPEC = kStageHLLCMuscl, TPB = 512, WPE = 4;
    constexpr bool FLDS = true;
    AMR_DISPATCH_TILE(NT, ng, (launch_stage_variant<T, G, TPB, FLDS, WPE, SPEC>(F, leaves, nleaf, ip, step, err, s,
                                                                                 ev_start, ev_stop, span, counts)));
    return;
  }
  if (spec == kStageHLLCMuscl) spec = kStageGeneric;
  if (spec == kStageHLLDMuscl)
    launch_stage_spec<kStageHLLDMuscl>(NT, ng, variant, F, leaves, nleaf, ip, step, err, s, ev_start, ev_stop, span,
                                       counts);
  else if (spec == kStageHLLDWeno5 && NT == 24 && ng == 4)
    launch_stage_spec<kStageHLLDWeno5>(NT, ng, variant, F, leaves, nleaf, ip, step, err, s, ev_start, ev_stop, span,
                                       counts);
  else
    launch_stage_spec<kStageGeneric>(NT, ng, variant, F, leaves, nleaf, ip, step, err, s, ev_start, ev_stop, span,
                                     counts);
}

void launch_emf(int NT, const FieldPtrs& F, const LeafItem* leaves, int nleaf, const InstParams* ip, int step,
                int rk, DevError* err, hipStream_t s, const int32_t* counts) {
  if (nleaf <= 0) return;
  dim3 grid((NT * NT + kThreads - 1) / kThreads, nleaf);
  AMR_DISPATCH_TILE(NT, (NT == 24 ? 4 : 2),
                    (emf_kernel<T><<<grid, kThreads, 0, s>>>(F, leaves, ip, step == rk, err, step, counts)));
}

void launch_bupdate(int NT, const FieldPtrs& F, const LeafItem* leaves, int nleaf, const InstParams* ip, int step,
                    int rk, hipStream_t s, const int32_t* counts) {
  if (nleaf <= 0) return;
  dim3 grid((NT * NT + kThreads - 1) / kThreads, nleaf);
  AMR_DISPATCH_TILE(NT, (NT == 24 ? 4 : 2),
                    (bupdate_kernel<T><<<grid, kThreads, 0, s>>>(F, leaves, ip, step, rk, counts)));
}

void launch_bsync(int NT, const FieldPtrs& F, const LeafItem* leaves, int nleaf, int last, hipStream_t s,
                  const int32_t* counts) {
  if (nleaf <= 0) return;
  dim3 grid((NT * NT + kThreads - 1) / kThreads, nleaf);
  AMR_DISPATCH_TILE(NT, (NT == 24 ? 4 : 2), (bsync_kernel<T><<<grid, kThreads, 0, s>>>(F, leaves, last, counts)));
}

void launch_fill(int NT, const FillOp* const* inst_ops, const FillGen* const* inst_gen, const int32_t* cprefix,
                 int ninst, int nchunks, const FillSet& set, hipStream_t s, const int32_t* counts) {
  static_assert(kFillChunk == kThreads, "one workgroup per fill chunk");
  if (nchunks <= 0 || set.n <= 0) return;
  FillProgram P{inst_ops, inst_gen, cprefix, ninst};
  const int a = set.nv[0], b = set.n > 1 ? set.nv[1] : 0, c = set.n > 2 ? set.nv[2] : 0;
  const int ka = set.kind[0], kb = set.n > 1 ? set.kind[1] : -1, kc = set.n > 2 ? set.kind[2] : -1;
  constexpr int C = kFillCell, F = kFillFace, E = kFillEdge;
  if (set.n == 1 && a == 8 && ka == C) fill_kernel_nv<8, 0, 0, C, 0, 0><<<nchunks, kThreads, 0, s>>>(P, set, NT, counts);
  else if (set.n == 1 && a == 2 && ka == F)
    fill_kernel_nv<2, 0, 0, F, 0, 0><<<nchunks, kThreads, 0, s>>>(P, set, NT, counts);
  else if (set.n == 1 && a == 1 && ka == E)
    fill_kernel_nv<1, 0, 0, E, 0, 0><<<nchunks, kThreads, 0, s>>>(P, set, NT, counts);
  else if (set.n == 2 && a == 8 && b == 2 && ka == C && kb == F)
    fill_kernel_nv<8, 2, 0, C, F, 0><<<nchunks, kThreads, 0, s>>>(P, set, NT, counts);
  else if (set.n == 2 && a == 8 && b == 8 && ka == F && kb == F)
    fill_kernel_nv<8, 8, 0, F, F, 0><<<nchunks, kThreads, 0, s>>>(P, set, NT, counts);
  else if (set.n == 3 && a == 2 && b == 8 && c == 8 && ka == F && kb == F && kc == F)
    fill_kernel_nv<2, 8, 8, F, F, F><<<nchunks, kThreads, 0, s>>>(P, set, NT, counts);
  else fill_kernel<<<nchunks, kThreads, 0, s>>>(P, set, NT, counts);
}

void launch_cfl_criteria(int NT, int ng, const FieldPtrs& F, const double* W, const LeafItem* leaves, int nleaf,
                         const InstParams* ip, double* lam_grad, DevError* err, hipStream_t s,
                         const int32_t* counts) {
  if (nleaf <= 0) return;
  AMR_DISPATCH_TILE(NT, ng,
                    (cfl_criteria_kernel<T, G><<<nleaf, kThreads, 0, s>>>(F, W, leaves, ip, lam_grad, err, counts)));
}

void launch_refine(int NT, int ng, const FieldPtrs& F, const RegridOp* ops, int nops, const InstParams* ip,
                   DevError* err, hipStream_t s) {
  if (nops <= 0) return;
  dim3 grid((NT * NT + kThreads - 1) / kThreads, nops, 4);
  AMR_DISPATCH_TILE(NT, ng, (refine_kernel<T, G><<<grid, kThreads, 0, s>>>(F, ops, ip, err)));
}

void launch_coarsen(int NT, int ng, const FieldPtrs& F, const RegridOp* ops, int nops, const int32_t* cmapU,
                    const int32_t* cmapB, const InstParams* ip, DevError* err, hipStream_t s) {
  if (nops <= 0) return;
  dim3 grid((NT * NT + kThreads - 1) / kThreads, nops);
  AMR_DISPATCH_TILE(NT, ng, (coarsen_kernel<T, G><<<grid, kThreads, 0, s>>>(F, ops, cmapU, cmapB, ip, err)));
}

// the step kernel instance for (tile, specialisation, lanes), or null
const void* step_kernel_fn(int NT, int ng, int spec, int lanes_idx) {
#define AMR_STEP_FN(T_, G_)                                                                              \
  do {                                                                                                   \
    if (NT == T_ && ng == G_) {                                                                          \
      if (spec == kStageHLLDMuscl)                                                                       \
        return lanes_idx ? (const void*)step_kernel<T_, G_, 512, kStageHLLDMuscl>                        \
                         : (const void*)step_kernel<T_, G_, 256, kStageHLLDMuscl>;                       \
      if (spec == kStageHLLDWeno5 && T_ == 24)                                                           \
        return lanes_idx ? (const void*)step_kernel<T_, G_, 512, kStageHLLDWeno5>                        \
                         : (const void*)step_kernel<T_, G_, 256, kStageHLLDWeno5>;                       \
      return lanes_idx ? (const void*)step_kernel<T_, G_, 512, kStageGeneric>                            \
                       : (const void*)step_kernel<T_, G_, 256, kStageGeneric>;                           \
    }                                                                                                    \
  } while (0)
  AMR_STEP_FN(20, 2);
  AMR_STEP_FN(24, 4);
  AMR_STEP_FN(12, 2);
#undef AMR_STEP_FN
  return nullptr;
}

int step_kernel_grid(int NT, int ng, int spec, int lanes_idx) {
  const void* fn = step_kernel_fn(NT, ng, spec, lanes_idx);
  if (!fn || lanes_idx < 0 || lanes_idx > 1) return 0;
  int dev = 0, cus = 0, per_cu = 0, coop = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kStepLanes[lanes_idx], 0) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return coop ? per_cu * cus : 0;
}

int launch_step(int NT, int ng, int spec, int lanes_idx, int grid, const StepArgs& A, hipStream_t s,
                hipEvent_t ev_start, hipEvent_t ev_stop) {
  const void* fn = step_kernel_fn(NT, ng, spec, lanes_idx);
  if (!fn || grid <= 0) return 1;
  StepArgs a = A;
  void* args[] = {&a};
  if (ev_start) hipEventRecord(ev_start, s);
  const hipError_t e = hipLaunchCooperativeKernel(fn, dim3(grid), dim3(kStepLanes[lanes_idx]), args, 0, s);
  if (ev_stop) hipEventRecord(ev_stop, s);
  return e == hipSuccess ? 0 : 1;
}

void launch_riemann_batch(int solver, int dir, double gamma, int n, const double* wL, const double* wR, double* flux,
                          int32_t* status, hipStream_t s) {
  if (n <= 0) return;
  riemann_batch_kernel<<<(n + 127) / 128, 128, 0, s>>>(solver, dir, gamma, n, wL, wR, flux, status);
}

void launch_recon_batch(int NT, int ng, int lim, double dx, int nt, const double* W, double* out, hipStream_t s) {
  if (nt <= 0) return;
  dim3 grid((NT * NT + 127) / 128, nt);
  recon_batch_kernel<<<grid, 128, 0, s>>>(NT, ng, lim, dx, W, out);
}

}  // namespace amr
