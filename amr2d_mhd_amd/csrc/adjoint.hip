// Frozen-mesh step and its discrete adjoint on one uniform periodic N x N grid.
//
// Forward: the reference's RK2 step (main.cpp:128-247) restricted to one
// level with exact periodic halos -- cons->prim, minmod MUSCL edge states
// with the normal B from the faces, HLLE, stage update, CT (corner EMF, face
// update, cell sync with the energy fix).  The arithmetic is that of the
// engine's kernels (physics.hpp), so on a single-level mesh it reproduces the
// engine bit for bit.
//
// Adjoint: the exact reverse of that discrete map (branch-wise at the
// limiter and wave-speed switches), written as gather kernels so the result
// is deterministic: per-face reverse sweeps store their contributions to the
// four stencil cells, and per-cell kernels sum them.
//
// Layout: U[8][N][N] (row j, column i, i fastest); Bx[j][i] is the x-face
// between cells i-1 and i, By[j][i] the y-face between rows j-1 and j;
// fluxes follow the faces; E[j][i] sits at the corner (i-1/2, j-1/2).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>

#include "../../include/amr2d_mhd.h"
#include "abi_guard.hpp"
#include "adjoint_math.hpp"

using namespace amr;

namespace {

constexpr int kT = 256;

__device__ inline int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }
// i mod n for any i (tiles can be wider than small grids)
__device__ inline int wrap_any(int i, int n) { return ((i % n) + n) % n; }

struct Grid {
  int n;
  double dx, g;
  size_t n2;
};

// -------------------------------------------------------------- forward
__global__ __launch_bounds__(kT) void k_prim(Grid G, const double* __restrict__ U, double* __restrict__ W) {
  const size_t c = (size_t)blockIdx.x * kT + threadIdx.x;
  if (c >= G.n2) return;
  double u[8], w[8];
  for (int k = 0; k < 8; k++) u[k] = U[k * G.n2 + c];
  prim_of(u, G.g, w);
  for (int k = 0; k < 8; k++) W[k * G.n2 + c] = w[k];
}

// edge states of one face: x faces (dir 0) take Bx as the normal component,
// y faces (dir 1) By, and are returned in the rotated frame (swap_xy)
__device__ inline void face_states(const Grid& G, const double* W, const double* Bn, int dir, int i, int j,
                                   double* wl, double* wr, size_t* cell) {
  for (int s = 0; s < 4; s++) {
    const int o = s - 2;
    cell[s] = dir == 0 ? (size_t)j * G.n + wrap(i + o, G.n) : (size_t)wrap(j + o, G.n) * G.n + i;
  }
  for (int k = 0; k < 8; k++) {
    double v[4];
    for (int s = 0; s < 4; s++) v[s] = W[k * G.n2 + cell[s]];
    recon_minmod(v, wl[k], wr[k]);
  }
  const double b = Bn[(size_t)j * G.n + i];
  if (dir == 0) {
    wl[5] = b;
    wr[5] = b;
  } else {
    wl[6] = b;
    wr[6] = b;
    swap_xy(wl);
    swap_xy(wr);
  }
}

template <int SOLVER>
__global__ __launch_bounds__(kT) void k_flux(Grid G, const double* __restrict__ W, const double* __restrict__ Bx,
                                             const double* __restrict__ By, double* __restrict__ Fx,
                                             double* __restrict__ Fy) {
  const size_t f = (size_t)blockIdx.x * kT + threadIdx.x;
  if (f >= G.n2) return;
  const int dir = blockIdx.y;
  const int i = (int)(f % G.n), j = (int)(f / G.n);
  double wl[8], wr[8], fl[8];
  size_t cell[4];
  face_states(G, W, dir ? By : Bx, dir, i, j, wl, wr, cell);
  if (SOLVER == kFluxHLLD) {
    for (int k = 0; k < 8; k++) fl[k] = 0.;
    flux_hlld(wl, wr, G.g, fl);
  } else {
    flux_hlle(wl, wr, G.g, fl);
  }
  if (dir) swap_xy(fl);
  double* F = dir ? Fy : Fx;
  for (int k = 0; k < 8; k++) F[k * G.n2 + f] = fl[k];
}

// stage update: mode 1 out = U - dt res; mode 2 out = 0.5 (U + Us - dt res)
__global__ __launch_bounds__(kT) void k_update(Grid G, const double* __restrict__ U, const double* __restrict__ Us,
                                               const double* __restrict__ Fx, const double* __restrict__ Fy,
                                               double dt, int mode, double* __restrict__ out) {
  const size_t c = (size_t)blockIdx.x * kT + threadIdx.x;
  if (c >= G.n2) return;
  const int i = (int)(c % G.n), j = (int)(c / G.n);
  const size_t cx = (size_t)j * G.n + wrap(i + 1, G.n), cy = (size_t)wrap(j + 1, G.n) * G.n + i;
  for (int k = 0; k < 8; k++) {
    const size_t o = k * G.n2;
    const double res = (Fx[o + cx] - Fx[o + c]) / G.dx + (Fy[o + cy] - Fy[o + c]) / G.dx;
    out[o + c] = mode == 1 ? U[o + c] - dt * res : 0.5 * (U[o + c] + Us[o + c] - dt * res);
  }
}

__global__ __launch_bounds__(kT) void k_emf(Grid G, const double* __restrict__ Fx, const double* __restrict__ Fy,
                                            double* __restrict__ E) {
  const size_t c = (size_t)blockIdx.x * kT + threadIdx.x;
  if (c >= G.n2) return;
  const int i = (int)(c % G.n), j = (int)(c / G.n);
  const double* gy5 = Fy + 5 * G.n2;
  const double* fx6 = Fx + 6 * G.n2;
  E[c] = 0.25 * ((gy5[(size_t)j * G.n + wrap(i - 1, G.n)] + gy5[c]) - (fx6[(size_t)wrap(j - 1, G.n) * G.n + i] + fx6[c]));
}

// face update: mode 1 Bn = B0 - dt r; mode 2 Bn = 0.5 (Bs + B0 - dt r); the
// result also goes to (Bcx, Bcy) when given
__global__ __launch_bounds__(kT) void k_faces(Grid G, const double* __restrict__ B0x, const double* __restrict__ B0y,
                                              const double* __restrict__ Bsx, const double* __restrict__ Bsy,
                                              const double* __restrict__ E, double dt, int mode,
                                              double* __restrict__ Bnx, double* __restrict__ Bny,
                                              double* __restrict__ Bcx, double* __restrict__ Bcy) {
  const size_t c = (size_t)blockIdx.x * kT + threadIdx.x;
  if (c >= G.n2) return;
  const int i = (int)(c % G.n), j = (int)(c / G.n);
  const double rx = (1. / G.dx) * (E[(size_t)wrap(j + 1, G.n) * G.n + i] - E[c]);
  const double ry = -(1. / G.dx) * (E[(size_t)j * G.n + wrap(i + 1, G.n)] - E[c]);
  const double bx = mode == 1 ? B0x[c] - dt * rx : 0.5 * (Bsx[c] + B0x[c] - dt * rx);
  const double by = mode == 1 ? B0y[c] - dt * ry : 0.5 * (Bsy[c] + B0y[c] - dt * ry);
  Bnx[c] = bx;
  Bny[c] = by;
  if (Bcx) {
    Bcx[c] = bx;
    Bcy[c] = by;
  }
}

// cell sync; saves the pre-sync cell B (q5, q6) when `pre` is given
__global__ __launch_bounds__(kT) void k_bsync(Grid G, double* __restrict__ Q, const double* __restrict__ Bx,
                                              const double* __restrict__ By, double* __restrict__ pre) {
  const size_t c = (size_t)blockIdx.x * kT + threadIdx.x;
  if (c >= G.n2) return;
  const int i = (int)(c % G.n), j = (int)(c / G.n);
  const double bx = 0.5 * (Bx[(size_t)j * G.n + wrap(i + 1, G.n)] + Bx[c]);
  const double by = 0.5 * (By[(size_t)wrap(j + 1, G.n) * G.n + i] + By[c]);
  const double q5 = Q[5 * G.n2 + c], q6 = Q[6 * G.n2 + c];
  if (pre) {
    pre[c] = q5;
    pre[G.n2 + c] = q6;
  }
  Q[4 * G.n2 + c] = Q[4 * G.n2 + c] + 0.5 * ((bx * bx + by * by) - ((q5 * q5) + (q6 * q6)));
  Q[5 * G.n2 + c] = bx;
  Q[6 * G.n2 + c] = by;
}

// -------------------------------------------------------------- adjoint
// Tape of one step (doubles, N2 = N*N each unless stated): the primitive
// states the two flux passes read (W1 = prim(U), W2 = prim(Us), 8 N2 each),
// the stage-1 faces (Bxs, Bys), the pre-sync cell B of both stages (2 N2
// each) and the output faces (Bxo, Byo).
struct Tape {
  double *W1, *W2, *Bxs, *Bys, *pre1, *pre2, *Bxo, *Byo;
};
constexpr size_t kTapePerCell = 24;

Tape tape_at(double* base, size_t n2) {
  Tape t;
  t.W1 = base;
  t.W2 = base + 8 * n2;
  t.Bxs = base + 16 * n2;
  t.Bys = base + 17 * n2;
  t.pre1 = base + 18 * n2;
  t.pre2 = base + 20 * n2;
  t.Bxo = base + 22 * n2;
  t.Byo = base + 23 * n2;
  return t;
}

// bsync reverse.  Qb (adjoint of the synced state; read from Qin, written to
// Qout, which may alias) -> adjoint of the pre-sync state and of the cell
// averages bx, by of the faces (bb[2][N2]); E += 0.5 ((bx^2 + by^2) - (q5^2 +
// q6^2)), q5 <- bx, q6 <- by
__global__ __launch_bounds__(kT) void k_bsync_vjp(Grid G, const double* Qin, double* Qout,
                                                  const double* __restrict__ pre, const double* __restrict__ Bx,
                                                  const double* __restrict__ By, double* __restrict__ bb) {
  const size_t c = (size_t)blockIdx.x * kT + threadIdx.x;
  if (c >= G.n2) return;
  const int i = (int)(c % G.n), j = (int)(c / G.n);
  const double bx = 0.5 * (Bx[(size_t)j * G.n + wrap(i + 1, G.n)] + Bx[c]);
  const double by = 0.5 * (By[(size_t)wrap(j + 1, G.n) * G.n + i] + By[c]);
  const double e = Qin[4 * G.n2 + c];
  const double q5 = Qin[5 * G.n2 + c], q6 = Qin[6 * G.n2 + c];
  bb[c] = q5 + e * bx;
  bb[G.n2 + c] = q6 + e * by;
  if (Qout != Qin) {
    for (int k = 0; k < 8; k++)
      if (k != 5 && k != 6) Qout[k * G.n2 + c] = Qin[k * G.n2 + c];
  }
  Qout[5 * G.n2 + c] = -e * pre[c];
  Qout[6 * G.n2 + c] = -e * pre[G.n2 + c];
}

// face reverse of one stage fused with the cell-average reverse of the bsync
// that follows it.  Bb = Bin + 0.5 (bb of both adjacent cells) is the adjoint
// of the updated faces (written to Bb, which may alias Bin).  Mode 2 (Bn =
// 0.5 (Bs + B0 - dt r)): B0b = 0.5 Bb and Bsb = 0.5 Bb are assigned; mode 1
// (Bn = B0 - dt r): B0b += Bb.
__global__ __launch_bounds__(kT) void k_faces_vjp(Grid G, const double* Binx, const double* Biny,
                                                  const double* __restrict__ bb, int mode, double* Bbx, double* Bby,
                                                  double* __restrict__ B0bx, double* __restrict__ B0by,
                                                  double* __restrict__ Bsbx, double* __restrict__ Bsby) {
  const size_t c = (size_t)blockIdx.x * kT + threadIdx.x;
  if (c >= G.n2) return;
  const int i = (int)(c % G.n), j = (int)(c / G.n);
  const double bxb = Binx[c] + 0.5 * (bb[c] + bb[(size_t)j * G.n + wrap(i - 1, G.n)]);
  const double byb = Biny[c] + 0.5 * (bb[G.n2 + c] + bb[G.n2 + (size_t)wrap(j - 1, G.n) * G.n + i]);
  Bbx[c] = bxb;
  Bby[c] = byb;
  if (mode == 2) {
    B0bx[c] = 0.5 * bxb;
    B0by[c] = 0.5 * byb;
    Bsbx[c] = 0.5 * bxb;
    Bsby[c] = 0.5 * byb;
  } else {
    B0bx[c] += bxb;
    B0by[c] += byb;
  }
}

// adjoint of E at every corner from the face adjoints: rx_b = -s dt Bbx,
// ry_b = -s dt Bby (s = 1 or 0.5), rx = (1/dx)(E[j+1] - E[j]),
// ry = -(1/dx)(E[i+1] - E[i])
__global__ __launch_bounds__(kT) void k_emf_b(Grid G, const double* __restrict__ Bbx, const double* __restrict__ Bby,
                                              double dt, int mode, double* __restrict__ Eb) {
  const size_t c = (size_t)blockIdx.x * kT + threadIdx.x;
  if (c >= G.n2) return;
  const int i = (int)(c % G.n), j = (int)(c / G.n);
  const double s = mode == 1 ? -dt : -0.5 * dt;
  const double rxb_jm = s * Bbx[(size_t)wrap(j - 1, G.n) * G.n + i], rxb = s * Bbx[c];
  const double ryb_im = s * Bby[(size_t)j * G.n + wrap(i - 1, G.n)], ryb = s * Bby[c];
  Eb[c] = (1. / G.dx) * (rxb_jm - rxb) - (1. / G.dx) * (ryb_im - ryb);
}

// adjoint of the face fluxes from the stage update (residual adjoint
// resb = -s dt Qb, s = 1 or 0.5) and from the EMF (components 6 of Fx, 5 of
// Fy), fused with the pass-through of Qb to the stage inputs: mode 2
// (out = 0.5 (U + Us - dt res)) assigns Ub = Usb = 0.5 Qb; mode 1 (out = U -
// dt res) adds Qb to Ub
__global__ __launch_bounds__(kT) void k_flux_b(Grid G, const double* __restrict__ Qb, const double* __restrict__ Eb,
                                               double dt, int mode, double* __restrict__ Fbx,
                                               double* __restrict__ Fby, double* __restrict__ Ub,
                                               double* __restrict__ Usb) {
  const size_t f = (size_t)blockIdx.x * kT + threadIdx.x;
  if (f >= G.n2) return;
  const int i = (int)(f % G.n), j = (int)(f / G.n);
  const double s = mode == 1 ? -dt : -0.5 * dt;
  const size_t lx = (size_t)j * G.n + wrap(i - 1, G.n), ly = (size_t)wrap(j - 1, G.n) * G.n + i;
  for (int k = 0; k < 8; k++) {
    const size_t o = k * G.n2;
    const double q = Qb[o + f];
    Fbx[o + f] = (s * Qb[o + lx] - s * q) / G.dx;
    Fby[o + f] = (s * Qb[o + ly] - s * q) / G.dx;
    if (mode == 2) {
      Ub[o + f] = 0.5 * q;
      Usb[o + f] = 0.5 * q;
    } else {
      Ub[o + f] += q;
    }
  }
  Fby[5 * G.n2 + f] += 0.25 * (Eb[f] + Eb[(size_t)j * G.n + wrap(i + 1, G.n)]);
  Fbx[6 * G.n2 + f] -= 0.25 * (Eb[f] + Eb[(size_t)wrap(j + 1, G.n) * G.n + i]);
}

// per-face reverse of reconstruction + HLLE for the faces of one direction
// (blockIdx.y): stencil contributions S[s][k] (s = cells i-2 .. i+1 along the
// face normal) and the normal-B adjoint (accumulated into Bnb)
// reverse of reconstruction + HLLE at one face (i, j) of direction DIR: the
// adjoints (wlb, wrb) of its two edge states in the grid frame, the stencil
// cells (i-2 .. i+1 along the normal) and the normal-B adjoint
template <int DIR, int SOLVER>
__device__ inline void face_reverse(const Grid& G, int i, int j, const double* __restrict__ W,
                                    const double* __restrict__ Bn, const double* __restrict__ Fb, double* wlb,
                                    double* wrb, size_t* cell) {
  const size_t f = (size_t)j * G.n + i;
  double wl[8], wr[8], fb[8];
  face_states(G, W, Bn, DIR, i, j, wl, wr, cell);
  for (int k = 0; k < 8; k++) {
    fb[k] = Fb[k * G.n2 + f];
    wlb[k] = 0.;
    wrb[k] = 0.;
  }
  if (DIR) swap_xy(fb);  // adjoint of the output swap
  if (SOLVER == kFluxHLLD) hlld_vjp(wl, wr, G.g, fb, wlb, wrb);
  else hlle_vjp(wl, wr, G.g, fb, wlb, wrb);
  if (DIR) {
    swap_xy(wlb);  // adjoint of the input swap
    swap_xy(wrb);
  }
  // Recompute the stencil cells from opaque copies of (i, j) rather than
  // keep the four 64-bit offsets alive across the solver reverse.
  int io = i, jo = j;
  asm volatile("" : "+v"(io), "+v"(jo));
  for (int s = 0; s < 4; s++)
    cell[s] = DIR == 0 ? (size_t)jo * G.n + wrap(io + s - 2, G.n) : (size_t)wrap(jo + s - 2, G.n) * G.n + io;
}

// stencil contributions vb[s] of variable k (zero for the normal B, which the
// face takes from Bn) from the edge-state adjoints; out(k, vb)
template <int DIR, class Out>
__device__ inline void stencil_reverse(const Grid& G, const double* __restrict__ W, const size_t* cell,
                                       const double* wlb, const double* wrb, Out out) {
  constexpr int nc = DIR ? 6 : 5;
  // Re-read the stencil for the reconstruction reverse rather than keep the
  // 28 values face_states loaded alive across the HLLE reverse: the opaque
  // copy of the pointer stops the compiler from reusing those loads, which
  // would cost ~60 VGPRs and push the kernel to one wave per SIMD.
  const double* Wr = W;
  asm volatile("" : "+s"(Wr));
  for (int k = 0; k < 8; k++) {
    double vb[4] = {0., 0., 0., 0.};
    if (k != nc) {
      double v[4];
      for (int s = 0; s < 4; s++) v[s] = Wr[k * G.n2 + cell[s]];
      recon_minmod_vjp(v, wlb[k], wrb[k], vb);
    }
    out(k, vb);
  }
}

template <int DIR, int SOLVER>
__device__ inline void flux_vjp_face(const Grid& G, size_t f, const double* __restrict__ W,
                                     const double* __restrict__ Bn, const double* __restrict__ Fb,
                                     double* __restrict__ S, double* __restrict__ Bnb) {
  const int i = (int)(f % G.n), j = (int)(f / G.n);
  double wlb[8], wrb[8];
  size_t cell[4];
  face_reverse<DIR, SOLVER>(G, i, j, W, Bn, Fb, wlb, wrb, cell);
  constexpr int nc = DIR ? 6 : 5;
  Bnb[f] += wlb[nc] + wrb[nc];
  stencil_reverse<DIR>(G, W, cell, wlb, wrb, [&](int k, const double* vb) {
    for (int s = 0; s < 4; s++) S[((size_t)s * 8 + k) * G.n2 + f] = vb[s];
  });
}

template <int DIR, int SOLVER>
__global__ __launch_bounds__(kT) void k_flux_vjp(Grid G, const double* __restrict__ W, const double* __restrict__ Bn,
                                                 const double* __restrict__ Fb, double* __restrict__ S,
                                                 double* __restrict__ Bnb) {
  const size_t f = (size_t)blockIdx.x * kT + threadIdx.x;
  if (f >= G.n2) return;
  flux_vjp_face<DIR, SOLVER>(G, f, W, Bn, Fb, S, Bnb);
}

// per-cell sum of the stencil contributions, then the cons->prim reverse
// from the tape's primitive state: Qb += (dW/dQ)^T Wb
__global__ __launch_bounds__(kT) void k_gather_prim_vjp(Grid G, const double* __restrict__ W,
                                                        const double* __restrict__ Sx, const double* __restrict__ Sy,
                                                        double* __restrict__ Qb) {
  const size_t c = (size_t)blockIdx.x * kT + threadIdx.x;
  if (c >= G.n2) return;
  const int i = (int)(c % G.n), j = (int)(c / G.n);
  double wb[8], w[8], ub[8];
  for (int k = 0; k < 8; k++) {
    double a = 0.;
    for (int s = 0; s < 4; s++) {
      // this cell is stencil position s of the x-face i + 2 - s and y-face j + 2 - s
      const size_t fx = (size_t)j * G.n + wrap(i + 2 - s, G.n), fy = (size_t)wrap(j + 2 - s, G.n) * G.n + i;
      a += Sx[((size_t)s * 8 + k) * G.n2 + fx];
      a += Sy[((size_t)s * 8 + k) * G.n2 + fy];
    }
    wb[k] = a;
    w[k] = W[k * G.n2 + c];
    ub[k] = 0.;
  }
  prim_vjp_w(w, G.g, wb, ub);
  for (int k = 0; k < 8; k++) Qb[k * G.n2 + c] += ub[k];
}

// Fused face reverse + stencil sum + cons->prim reverse over a tile of
// TW x TH cells, one workgroup of TT lanes: the x faces the tile's cells
// read (TH rows of TW + 3) are reversed one per lane and leave their
// stencil contributions in LDS; after a barrier every cell sums its four;
// the y faces (TH + 3 rows of TW) follow in the same LDS; then the cell
// adds (dW/dQ)^T Wb to Qb.  Faces next to the tile edge are reversed by both
// tiles that read them; the normal-B adjoint of a face is added only by the
// tile that owns the face.  Deterministic: fixed summation order, no
// atomics.
template <int TW, int TH, int TT, int SOLVER>
__global__ __launch_bounds__(TT) void k_stage_vjp(Grid G, const double* __restrict__ W,
                                                  const double* __restrict__ Bx, const double* __restrict__ By,
                                                  const double* __restrict__ Fbx, const double* __restrict__ Fby,
                                                  double* __restrict__ Bnbx, double* __restrict__ Bnby,
                                                  double* __restrict__ Qb) {
  static_assert(TH * (TW + 3) <= TT && (TH + 3) * TW <= TT, "one face per lane and direction");
  __shared__ double S[8][4][TT];  // [variable][stencil position][face of the tile]
  // x-face sums of the tile's cells, kept in LDS across the y-face reverse
  // rather than in 16 VGPRs
  __shared__ double Sx[8][TW * TH];
  const int tid = threadIdx.x;
  const int i0 = blockIdx.x * TW, j0 = blockIdx.y * TH;
  // this lane's cell
  const int ci = tid % TW, cj = tid / TW;
  const int i = i0 + ci, j = j0 + cj;
  const bool has_cell = tid < TW * TH && i < G.n && j < G.n;
  // ---- x faces: tile row r, face column q -> face i0 - 1 + q
  if (tid < TH * (TW + 3)) {
    const int r = tid / (TW + 3), q = tid % (TW + 3);
    const int fj = j0 + r, fi = wrap_any(i0 - 1 + q, G.n);
    if (fj < G.n) {
      double wlb[8], wrb[8];
      size_t cell[4];
      face_reverse<0, SOLVER>(G, fi, fj, W, Bx, Fbx, wlb, wrb, cell);
      if (q >= 1 && q <= TW && i0 - 1 + q < G.n) Bnbx[(size_t)fj * G.n + fi] += wlb[5] + wrb[5];
      stencil_reverse<0>(G, W, cell, wlb, wrb, [&](int k, const double* vb) {
        for (int s = 0; s < 4; s++) S[k][s][tid] = vb[s];
      });
    }
  }
  __syncthreads();
  if (has_cell) {
    // this cell is stencil position s of the x face i + 2 - s: tile column ci + 3 - s
    for (int k = 0; k < 8; k++) {
      double a = 0.;
      for (int s = 0; s < 4; s++) a += S[k][s][cj * (TW + 3) + ci + 3 - s];
      Sx[k][tid] = a;
    }
  }
  __syncthreads();
  // ---- y faces: face row r -> face j0 - 1 + r, column q
  if (tid < (TH + 3) * TW) {
    const int r = tid / TW, q = tid % TW;
    const int fj = wrap_any(j0 - 1 + r, G.n), fi = i0 + q;
    if (fi < G.n) {
      double wlb[8], wrb[8];
      size_t cell[4];
      face_reverse<1, SOLVER>(G, fi, fj, W, By, Fby, wlb, wrb, cell);
      if (r >= 1 && r <= TH && j0 - 1 + r < G.n) Bnby[(size_t)fj * G.n + fi] += wlb[6] + wrb[6];
      stencil_reverse<1>(G, W, cell, wlb, wrb, [&](int k, const double* vb) {
        for (int s = 0; s < 4; s++) S[k][s][tid] = vb[s];
      });
    }
  }
  __syncthreads();
  if (has_cell) {
    const size_t c = (size_t)j * G.n + i;
    double w[8], ub[8], wb[8];
    for (int k = 0; k < 8; k++) {
      double a = Sx[k][tid];
      for (int s = 0; s < 4; s++) a += S[k][s][(cj + 3 - s) * TW + ci];
      wb[k] = a;
      w[k] = W[k * G.n2 + c];
      ub[k] = 0.;
    }
    prim_vjp_w(w, G.g, wb, ub);
    for (int k = 0; k < 8; k++) Qb[k * G.n2 + c] += ub[k];
  }
}

template <int TW, int TH, int TT, int SOLVER>
void launch_stage_vjp(const Grid& G, const double* W, const double* Bx, const double* By, const double* Fbx,
                      const double* Fby, double* Bnbx, double* Bnby, double* Qb, hipStream_t s) {
  const dim3 grid((unsigned)((G.n + TW - 1) / TW), (unsigned)((G.n + TH - 1) / TH));
  k_stage_vjp<TW, TH, TT, SOLVER><<<grid, TT, 0, s>>>(G, W, Bx, By, Fbx, Fby, Bnbx, Bnby, Qb);
}

unsigned blocks(size_t n) { return (unsigned)((n + kT - 1) / kT); }

}  // namespace

struct amr_adj {
  Grid G;
  int solver = kFluxHLLE;  // kFluxHLLE or kFluxHLLD
  int fused = 1;  // face reverse + gather in one tiled kernel: 1 16x13 tiles of 256 lanes, 2 32x13 tiles of
                  // 512 lanes; 0: per-face kernels and a gather
  // forward scratch
  double *Fx, *Fy, *E, *Us, *Uo, *Bxo, *Byo, *tape;
  // adjoint work
  double *Qb2, *Usb, *Fbx, *Fby, *Eb, *Bsbx, *Bsby, *Bbx, *Bby, *bb, *Sx, *Sy;
};

namespace {

#define ADJ_CHK(x)                                                                             \
  do {                                                                                         \
    hipError_t e_ = (x);                                                                       \
    if (e_ != hipSuccess) {                                                                    \
      std::fprintf(stderr, "amr2d_mhd adjoint: HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, \
                   __LINE__);                                                                  \
      return AMR_ERR_HIP;                                                                      \
    }                                                                                          \
  } while (0)

void launch_flux(amr_adj* a, const double* W, const double* Bx, const double* By, hipStream_t s) {
  const dim3 b2(blocks(a->G.n2), 2);
  if (a->solver == kFluxHLLD) k_flux<kFluxHLLD><<<b2, kT, 0, s>>>(a->G, W, Bx, By, a->Fx, a->Fy);
  else k_flux<kFluxHLLE><<<b2, kT, 0, s>>>(a->G, W, Bx, By, a->Fx, a->Fy);
}

// forward stage sequence, writing the step's tape
int forward(amr_adj* a, const double* U, const double* Bx, const double* By, double dt, double* Uo, double* Bxo,
            double* Byo, const Tape& T, hipStream_t s) {
  const Grid& G = a->G;
  const unsigned b1 = blocks(G.n2);
  k_prim<<<b1, kT, 0, s>>>(G, U, T.W1);
  launch_flux(a, T.W1, Bx, By, s);
  k_update<<<b1, kT, 0, s>>>(G, U, nullptr, a->Fx, a->Fy, dt, 1, a->Us);
  k_emf<<<b1, kT, 0, s>>>(G, a->Fx, a->Fy, a->E);
  k_faces<<<b1, kT, 0, s>>>(G, Bx, By, nullptr, nullptr, a->E, dt, 1, T.Bxs, T.Bys, nullptr, nullptr);
  k_bsync<<<b1, kT, 0, s>>>(G, a->Us, T.Bxs, T.Bys, T.pre1);
  k_prim<<<b1, kT, 0, s>>>(G, a->Us, T.W2);
  launch_flux(a, T.W2, T.Bxs, T.Bys, s);
  k_update<<<b1, kT, 0, s>>>(G, U, a->Us, a->Fx, a->Fy, dt, 2, Uo);
  k_emf<<<b1, kT, 0, s>>>(G, a->Fx, a->Fy, a->E);
  k_faces<<<b1, kT, 0, s>>>(G, Bx, By, T.Bxs, T.Bys, a->E, dt, 2, Bxo, Byo, T.Bxo, T.Byo);
  k_bsync<<<b1, kT, 0, s>>>(G, Uo, Bxo, Byo, T.pre2);
  ADJ_CHK(hipGetLastError());
  return 0;
}

// face reverse of one stage from its flux adjoints (a->Fbx, a->Fby): normal-B
// adjoints into Bnb*, cell adjoints added to Qb
void face_and_gather(amr_adj* a, const double* W, const double* Bx, const double* By, double* Bnbx, double* Bnby,
                     double* Qb, hipStream_t s) {
  const Grid& G = a->G;
  const bool hlld = a->solver == kFluxHLLD;
  if (a->fused == 1) {
    if (hlld) launch_stage_vjp<16, 13, 256, kFluxHLLD>(G, W, Bx, By, a->Fbx, a->Fby, Bnbx, Bnby, Qb, s);
    else launch_stage_vjp<16, 13, 256, kFluxHLLE>(G, W, Bx, By, a->Fbx, a->Fby, Bnbx, Bnby, Qb, s);
  } else if (a->fused == 2) {
    if (hlld) launch_stage_vjp<32, 13, 512, kFluxHLLD>(G, W, Bx, By, a->Fbx, a->Fby, Bnbx, Bnby, Qb, s);
    else launch_stage_vjp<32, 13, 512, kFluxHLLE>(G, W, Bx, By, a->Fbx, a->Fby, Bnbx, Bnby, Qb, s);
  } else {
    const unsigned b1 = blocks(G.n2);
    if (hlld) {
      k_flux_vjp<0, kFluxHLLD><<<b1, kT, 0, s>>>(G, W, Bx, a->Fbx, a->Sx, Bnbx);
      k_flux_vjp<1, kFluxHLLD><<<b1, kT, 0, s>>>(G, W, By, a->Fby, a->Sy, Bnby);
    } else {
      k_flux_vjp<0, kFluxHLLE><<<b1, kT, 0, s>>>(G, W, Bx, a->Fbx, a->Sx, Bnbx);
      k_flux_vjp<1, kFluxHLLE><<<b1, kT, 0, s>>>(G, W, By, a->Fby, a->Sy, Bnby);
    }
    k_gather_prim_vjp<<<b1, kT, 0, s>>>(G, W, a->Sx, a->Sy, Qb);
  }
}

// reverse sweep of one step from its tape and input faces
int reverse(amr_adj* a, const Tape& T, const double* Bx, const double* By, double dt, const double* Uob,
            const double* Bxob, const double* Byob, double* Ub, double* Bxb, double* Byb, hipStream_t s) {
  const Grid& G = a->G;
  const unsigned b1 = blocks(G.n2);
  // ---- stage 2
  k_bsync_vjp<<<b1, kT, 0, s>>>(G, Uob, a->Qb2, T.pre2, T.Bxo, T.Byo, a->bb);
  k_faces_vjp<<<b1, kT, 0, s>>>(G, Bxob, Byob, a->bb, 2, a->Bbx, a->Bby, Bxb, Byb, a->Bsbx, a->Bsby);
  k_emf_b<<<b1, kT, 0, s>>>(G, a->Bbx, a->Bby, dt, 2, a->Eb);
  k_flux_b<<<b1, kT, 0, s>>>(G, a->Qb2, a->Eb, dt, 2, a->Fbx, a->Fby, Ub, a->Usb);
  face_and_gather(a, T.W2, T.Bxs, T.Bys, a->Bsbx, a->Bsby, a->Usb, s);
  // ---- stage 1
  k_bsync_vjp<<<b1, kT, 0, s>>>(G, a->Usb, a->Usb, T.pre1, T.Bxs, T.Bys, a->bb);
  k_faces_vjp<<<b1, kT, 0, s>>>(G, a->Bsbx, a->Bsby, a->bb, 1, a->Bsbx, a->Bsby, Bxb, Byb, nullptr, nullptr);
  k_emf_b<<<b1, kT, 0, s>>>(G, a->Bsbx, a->Bsby, dt, 1, a->Eb);
  k_flux_b<<<b1, kT, 0, s>>>(G, a->Usb, a->Eb, dt, 1, a->Fbx, a->Fby, Ub, nullptr);
  face_and_gather(a, T.W1, Bx, By, Bxb, Byb, Ub, s);
  ADJ_CHK(hipGetLastError());
  return 0;
}

// ------------------------------------------------- allocation-free step
// Work memory of amr_adj_step_noalloc: the forward's face fluxes, corner EMF,
// stage state and tape (8 + 8 + 1 + 8 + kTapePerCell doubles per cell) for
// grids of up to kNoAllocMaxN^2 cells, in one file-scope device array that
// every GPU the library runs on holds until the process ends (25.7 MB).  Host
// code never passes the array's name to a kernel (a host-side stand-in):
// each device's address comes from hipGetSymbolAddress, taken once per
// device on its first call, and only that address is used.
constexpr int kNoAllocMaxN = 256;
constexpr size_t kNoAllocPerCell = 8 + 8 + 1 + 8 + kTapePerCell;
static __device__ __align__(256) double scratch[kNoAllocPerCell * kNoAllocMaxN * kNoAllocMaxN];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {nullptr};

// the current device's address of `scratch`
int scratch_address(double** out) {
  int dev = -1;
  ADJ_CHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return AMR_ERR_CONFIG;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    ADJ_CHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return AMR_ERR_HIP;
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

}  // namespace

extern "C" {

int amr_adj_step_noalloc(int n, double dx, double gamma, const char* flux, const double* U, const double* Bx,
                         const double* By, double dt, double* Uo, double* Bxo, double* Byo, void* stream) {
  AMR_TRY
  if (n < 4 || n > kNoAllocMaxN || dx <= 0 || !U || !Bx || !By || !Uo || !Bxo || !Byo) {
    if (n > kNoAllocMaxN)
      std::fprintf(stderr, "amr_adj_step_noalloc: n = %d exceeds the %d x %d cells of its static work memory\n", n,
                   kNoAllocMaxN, kNoAllocMaxN);
    return AMR_ERR_CONFIG;
  }
  int solver = kFluxHLLE;
  if (flux && std::strcmp(flux, "HLLD") == 0) solver = kFluxHLLD;
  else if (flux && std::strcmp(flux, "HLLE") != 0) return AMR_ERR_CONFIG;
  double* w = nullptr;
  if (int e = scratch_address(&w)) return e;
  // a context on the stack whose work buffers are views into the array
  amr_adj a{};
  a.G = Grid{n, dx, gamma, (size_t)n * n};
  a.solver = solver;
  const size_t n2 = a.G.n2;
  a.Fx = w;
  a.Fy = w + 8 * n2;
  a.E = w + 16 * n2;
  a.Us = w + 17 * n2;
  a.tape = w + 25 * n2;
  return forward(&a, U, Bx, By, dt, Uo, Bxo, Byo, tape_at(a.tape, n2), (hipStream_t)stream);
  AMR_CATCH
}

int amr_adj_create(int n, double dx, double gamma, const char* flux, amr_adj** out) {
  AMR_TRY
  if (!out || n < 4 || dx <= 0) return AMR_ERR_CONFIG;
  int solver = kFluxHLLE;
  if (flux && std::strcmp(flux, "HLLD") == 0) solver = kFluxHLLD;
  else if (flux && std::strcmp(flux, "HLLE") != 0) return AMR_ERR_CONFIG;
  amr_adj* a = new amr_adj();
  a->solver = solver;
  a->G = Grid{n, dx, gamma, (size_t)n * n};
  const size_t n2 = a->G.n2;
  double** bufs8[] = {&a->Fx, &a->Fy, &a->Us, &a->Uo, &a->Qb2, &a->Usb, &a->Fbx, &a->Fby};
  double** bufs1[] = {&a->E, &a->Bxo, &a->Byo, &a->Eb, &a->Bsbx, &a->Bsby, &a->Bbx, &a->Bby};
  bool ok = true;
  for (auto p : bufs8) ok = ok && hipMalloc((void**)p, 8 * n2 * sizeof(double)) == hipSuccess;
  for (auto p : bufs1) ok = ok && hipMalloc((void**)p, n2 * sizeof(double)) == hipSuccess;
  ok = ok && hipMalloc((void**)&a->tape, kTapePerCell * n2 * sizeof(double)) == hipSuccess;
  ok = ok && hipMalloc((void**)&a->bb, 2 * n2 * sizeof(double)) == hipSuccess;
  ok = ok && hipMalloc((void**)&a->Sx, 32 * n2 * sizeof(double)) == hipSuccess;
  ok = ok && hipMalloc((void**)&a->Sy, 32 * n2 * sizeof(double)) == hipSuccess;
  if (!ok) {
    amr_adj_destroy(a);
    return AMR_ERR_HIP;
  }
  *out = a;
  return AMR_OK;
  AMR_CATCH
}

int amr_adj_destroy(amr_adj* a) {
  AMR_TRY
  if (!a) return AMR_OK;
  double* all[] = {a->Fx,  a->Fy,  a->E,    a->Us,   a->Uo,  a->Bxo, a->Byo, a->tape, a->Qb2, a->Usb,
                   a->Fbx, a->Fby, a->Eb,   a->Bsbx, a->Bsby, a->Bbx, a->Bby, a->bb,   a->Sx,  a->Sy};
  for (double* p : all)
    if (p) hipFree(p);
  delete a;
  return AMR_OK;
  AMR_CATCH
}

int64_t amr_adj_tape_doubles(const amr_adj* a) { return a ? (int64_t)(kTapePerCell * a->G.n2) : 0; }

int amr_adj_step_tape(amr_adj* a, const double* U, const double* Bx, const double* By, double dt, double* Uo,
                      double* Bxo, double* Byo, double* tape, void* stream) {
  AMR_TRY
  if (!a || !U || !Bx || !By || !Uo || !Bxo || !Byo) return AMR_ERR_CONFIG;
  return forward(a, U, Bx, By, dt, Uo, Bxo, Byo, tape_at(tape ? tape : a->tape, a->G.n2), (hipStream_t)stream);
  AMR_CATCH
}

int amr_adj_step(amr_adj* a, const double* U, const double* Bx, const double* By, double dt, double* Uo, double* Bxo,
                 double* Byo, void* stream) {
  AMR_TRY
  return amr_adj_step_tape(a, U, Bx, By, dt, Uo, Bxo, Byo, nullptr, stream);
  AMR_CATCH
}

int amr_adj_vjp_tape(amr_adj* a, const double* tape, const double* Bx, const double* By, double dt,
                     const double* Uob, const double* Bxob, const double* Byob, double* Ub, double* Bxb, double* Byb,
                     void* stream) {
  AMR_TRY
  if (!a || !tape || !Bx || !By || !Uob || !Bxob || !Byob || !Ub || !Bxb || !Byb) return AMR_ERR_CONFIG;
  return reverse(a, tape_at(const_cast<double*>(tape), a->G.n2), Bx, By, dt, Uob, Bxob, Byob, Ub, Bxb, Byb,
                 (hipStream_t)stream);
  AMR_CATCH
}

int amr_adj_step_vjp(amr_adj* a, const double* U, const double* Bx, const double* By, double dt, const double* Uob,
                     const double* Bxob, const double* Byob, double* Ub, double* Bxb, double* Byb, void* stream) {
  AMR_TRY
  if (!a || !U || !Bx || !By || !Uob || !Bxob || !Byob || !Ub || !Bxb || !Byb) return AMR_ERR_CONFIG;
  hipStream_t s = (hipStream_t)stream;
  const Tape T = tape_at(a->tape, a->G.n2);
  // recompute the step into the context's own tape
  if (int e = forward(a, U, Bx, By, dt, a->Uo, a->Bxo, a->Byo, T, s)) return e;
  return reverse(a, T, Bx, By, dt, Uob, Bxob, Byob, Ub, Bxb, Byb, s);
  AMR_CATCH
}

// test hook: 1 = fused tiled face reverse on 16x13 tiles (default), 2 = on
// 32x13 tiles, 0 = per-face kernels and a separate gather (same derivative,
// different summation order)
int amr_debug_adj_set_fused(amr_adj* a, int mode) {
  AMR_TRY
  if (!a || mode < 0 || mode > 2) return AMR_ERR_CONFIG;
  a->fused = mode;
  return AMR_OK;
  AMR_CATCH
}

// host entry points of the pointwise reverse pieces (CPU tests)
void amr_debug_hlle_vjp(const double* wL, const double* wR, double gamma, const double* fb, double* wLb, double* wRb) {
  for (int k = 0; k < 8; k++) wLb[k] = wRb[k] = 0.;
  hlle_vjp(wL, wR, gamma, fb, wLb, wRb);
}

void amr_debug_hlle(const double* wL, const double* wR, double gamma, double* f) { flux_hlle(wL, wR, gamma, f); }

// HLLD reverse on the host; f receives the flux its own forward computed
void amr_debug_hlld_vjp(const double* wL, const double* wR, double gamma, const double* fb, double* wLb, double* wRb,
                        double* f) {
  for (int k = 0; k < 8; k++) wLb[k] = wRb[k] = 0.;
  hlld_vjp(wL, wR, gamma, fb, wLb, wRb, f);
}

void amr_debug_hlld(const double* wL, const double* wR, double gamma, double* f) {
  for (int k = 0; k < 8; k++) f[k] = 0.;
  flux_hlld(wL, wR, gamma, f);
}

// forward values of the reverse's own forward pieces next to those of the
// engine's: out[0..7] = side_fwd (cf, e, H) of both sides and roe_fwd (lmin,
// lmax); ref[0..7] = the same from side_state / roe_bounds
void amr_debug_side_roe(const double* wL, const double* wR, double gamma, double* out, double* ref) {
  FastT fL, fR;
  double eL, HL, eR, HR, lmin, lmax;
  side_fwd(wL, gamma, fL, eL, HL);
  side_fwd(wR, gamma, fR, eR, HR);
  RoeT rt;
  roe_fwd(wL, wR, gamma, HL, HR, rt, lmin, lmax);
  const double o[] = {fL.cf, eL, HL, fR.cf, eR, HR, lmin, lmax};
  Side L, R;
  side_state(wL, gamma, L);
  side_state(wR, gamma, R);
  double u, rlmin, rlmax;
  roe_bounds(wL, wR, gamma, L.H, R.H, u, rlmin, rlmax);
  const double r[] = {L.cf, L.e, L.H, R.cf, R.e, R.H, rlmin, rlmax};
  for (int k = 0; k < 8; k++) {
    out[k] = o[k];
    ref[k] = r[k];
  }
}

void amr_debug_recon_minmod_vjp(const double* v, double wlb, double wrb, double* vb, double* wlr) {
  for (int s = 0; s < 4; s++) vb[s] = 0.;
  recon_minmod_vjp(v, wlb, wrb, vb);
  recon_minmod(v, wlr[0], wlr[1]);
}

void amr_debug_prim_vjp(const double* u, double gamma, const double* wb, double* ub, double* w) {
  for (int k = 0; k < 8; k++) ub[k] = 0.;
  prim_vjp(u, gamma, wb, ub);
  prim_of(u, gamma, w);
}

}  // extern "C"
