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

#include "../../include/amr2d_mhd.h"
#include "adjoint_math.hpp"

using namespace amr;

namespace {

constexpr int kT = 256;

__device__ inline int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

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

__global__ __launch_bounds__(kT) void k_flux(Grid G, const double* __restrict__ W, const double* __restrict__ Bx,
                                             const double* __restrict__ By, double* __restrict__ Fx,
                                             double* __restrict__ Fy) {
  const size_t t = (size_t)blockIdx.x * kT + threadIdx.x;
  if (t >= 2 * G.n2) return;
  const int dir = t >= G.n2;
  const size_t f = dir ? t - G.n2 : t;
  const int i = (int)(f % G.n), j = (int)(f / G.n);
  double wl[8], wr[8], fl[8];
  size_t cell[4];
  face_states(G, W, dir ? By : Bx, dir, i, j, wl, wr, cell);
  flux_hlle(wl, wr, G.g, fl);
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

// face update: mode 1 Bn = B0 - dt r; mode 2 Bn = 0.5 (Bs + B0 - dt r)
__global__ __launch_bounds__(kT) void k_faces(Grid G, const double* __restrict__ B0x, const double* __restrict__ B0y,
                                              const double* __restrict__ Bsx, const double* __restrict__ Bsy,
                                              const double* __restrict__ E, double dt, int mode,
                                              double* __restrict__ Bnx, double* __restrict__ Bny) {
  const size_t c = (size_t)blockIdx.x * kT + threadIdx.x;
  if (c >= G.n2) return;
  const int i = (int)(c % G.n), j = (int)(c / G.n);
  const double rx = (1. / G.dx) * (E[(size_t)wrap(j + 1, G.n) * G.n + i] - E[c]);
  const double ry = -(1. / G.dx) * (E[(size_t)j * G.n + wrap(i + 1, G.n)] - E[c]);
  Bnx[c] = mode == 1 ? B0x[c] - dt * rx : 0.5 * (Bsx[c] + B0x[c] - dt * rx);
  Bny[c] = mode == 1 ? B0y[c] - dt * ry : 0.5 * (Bsy[c] + B0y[c] - dt * ry);
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
// bsync reverse: Qb (adjoint of the synced state) -> adjoint of the
// pre-sync state (in place) and of the cell averages bx, by (bb[2][N2])
__global__ __launch_bounds__(kT) void k_bsync_vjp(Grid G, double* __restrict__ Qb, const double* __restrict__ pre,
                                                  const double* __restrict__ Bx, const double* __restrict__ By,
                                                  double* __restrict__ bb) {
  const size_t c = (size_t)blockIdx.x * kT + threadIdx.x;
  if (c >= G.n2) return;
  const int i = (int)(c % G.n), j = (int)(c / G.n);
  const double bx = 0.5 * (Bx[(size_t)j * G.n + wrap(i + 1, G.n)] + Bx[c]);
  const double by = 0.5 * (By[(size_t)wrap(j + 1, G.n) * G.n + i] + By[c]);
  const double e = Qb[4 * G.n2 + c];
  bb[c] = Qb[5 * G.n2 + c] + e * bx;
  bb[G.n2 + c] = Qb[6 * G.n2 + c] + e * by;
  Qb[5 * G.n2 + c] = -e * pre[c];
  Qb[6 * G.n2 + c] = -e * pre[G.n2 + c];
}

// faces reverse, fused with the cell-average reverse of the following bsync:
// Bb* (adjoint of the updated faces, in place: + 0.5 of both adjacent cell
// averages), then the adjoint of B0 (accumulated into B0b*), of Bs (mode 2:
// written to Bsb*) and of E (written to Eb)
__global__ __launch_bounds__(kT) void k_faces_avg_vjp(Grid G, double* __restrict__ Bbx, double* __restrict__ Bby,
                                                      const double* __restrict__ bb) {
  const size_t c = (size_t)blockIdx.x * kT + threadIdx.x;
  if (c >= G.n2) return;
  const int i = (int)(c % G.n), j = (int)(c / G.n);
  Bbx[c] += 0.5 * (bb[c] + bb[(size_t)j * G.n + wrap(i - 1, G.n)]);
  Bby[c] += 0.5 * (bb[G.n2 + c] + bb[G.n2 + (size_t)wrap(j - 1, G.n) * G.n + i]);
}

__global__ __launch_bounds__(kT) void k_faces_vjp(Grid G, const double* __restrict__ Bbx, const double* __restrict__ Bby,
                                                  double dt, int mode, double* __restrict__ B0bx,
                                                  double* __restrict__ B0by, double* __restrict__ Bsbx,
                                                  double* __restrict__ Bsby) {
  const size_t c = (size_t)blockIdx.x * kT + threadIdx.x;
  if (c >= G.n2) return;
  const double s = mode == 1 ? 1. : 0.5;
  B0bx[c] += s * Bbx[c];
  B0by[c] += s * Bby[c];
  if (mode == 2) {
    Bsbx[c] = 0.5 * Bbx[c];
    Bsby[c] = 0.5 * Bby[c];
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
// resb = -s dt Qb, s = 1 or 0.5) and from the EMF (components 6 of Fx, 5 of Fy)
__global__ __launch_bounds__(kT) void k_flux_b(Grid G, const double* __restrict__ Qb, const double* __restrict__ Eb,
                                               double dt, int mode, double* __restrict__ Fbx, double* __restrict__ Fby) {
  const size_t f = (size_t)blockIdx.x * kT + threadIdx.x;
  if (f >= G.n2) return;
  const int i = (int)(f % G.n), j = (int)(f / G.n);
  const double s = mode == 1 ? -dt : -0.5 * dt;
  const size_t lx = (size_t)j * G.n + wrap(i - 1, G.n), ly = (size_t)wrap(j - 1, G.n) * G.n + i;
  for (int k = 0; k < 8; k++) {
    const size_t o = k * G.n2;
    Fbx[o + f] = (s * Qb[o + lx] - s * Qb[o + f]) / G.dx;
    Fby[o + f] = (s * Qb[o + ly] - s * Qb[o + f]) / G.dx;
  }
  Fby[5 * G.n2 + f] += 0.25 * (Eb[f] + Eb[(size_t)j * G.n + wrap(i + 1, G.n)]);
  Fbx[6 * G.n2 + f] -= 0.25 * (Eb[f] + Eb[(size_t)wrap(j + 1, G.n) * G.n + i]);
}

// per-face reverse of reconstruction + HLLE: stencil contributions S[s][k]
// (s = cells i-2 .. i+1 along the face normal) and the normal-B adjoint
__global__ __launch_bounds__(kT) void k_flux_vjp(Grid G, const double* __restrict__ W, const double* __restrict__ Bx,
                                                 const double* __restrict__ By, const double* __restrict__ Fbx,
                                                 const double* __restrict__ Fby, double* __restrict__ Sx,
                                                 double* __restrict__ Sy, double* __restrict__ Bnbx,
                                                 double* __restrict__ Bnby) {
  const size_t t = (size_t)blockIdx.x * kT + threadIdx.x;
  if (t >= 2 * G.n2) return;
  const int dir = t >= G.n2;
  const size_t f = dir ? t - G.n2 : t;
  const int i = (int)(f % G.n), j = (int)(f / G.n);
  double wl[8], wr[8], fb[8], wlb[8] = {0}, wrb[8] = {0};
  size_t cell[4];
  face_states(G, W, dir ? By : Bx, dir, i, j, wl, wr, cell);
  const double* Fb = dir ? Fby : Fbx;
  for (int k = 0; k < 8; k++) fb[k] = Fb[k * G.n2 + f];
  if (dir) swap_xy(fb);  // adjoint of the output swap
  hlle_vjp(wl, wr, G.g, fb, wlb, wrb);
  const int nc = dir ? 6 : 5;
  if (dir) {
    swap_xy(wlb);  // adjoint of the input swap
    swap_xy(wrb);
  }
  (dir ? Bnby : Bnbx)[f] += wlb[nc] + wrb[nc];
  double* S = dir ? Sy : Sx;
  for (int k = 0; k < 8; k++) {
    double vb[4] = {0., 0., 0., 0.};
    if (k != nc) {
      double v[4];
      for (int s = 0; s < 4; s++) v[s] = W[k * G.n2 + cell[s]];
      recon_minmod_vjp(v, wlb[k], wrb[k], vb);
    }
    for (int s = 0; s < 4; s++) S[((size_t)s * 8 + k) * G.n2 + f] = vb[s];
  }
}

// per-cell sum of the stencil contributions, then the cons->prim reverse:
// Qb += (dW/dQ)^T Wb
__global__ __launch_bounds__(kT) void k_gather_prim_vjp(Grid G, const double* __restrict__ Q,
                                                        const double* __restrict__ Sx, const double* __restrict__ Sy,
                                                        double* __restrict__ Qb) {
  const size_t c = (size_t)blockIdx.x * kT + threadIdx.x;
  if (c >= G.n2) return;
  const int i = (int)(c % G.n), j = (int)(c / G.n);
  double wb[8], u[8], ub[8];
  for (int k = 0; k < 8; k++) {
    double a = 0.;
    for (int s = 0; s < 4; s++) {
      // this cell is stencil position s of the x-face i + 2 - s and y-face j + 2 - s
      const size_t fx = (size_t)j * G.n + wrap(i + 2 - s, G.n), fy = (size_t)wrap(j + 2 - s, G.n) * G.n + i;
      a += Sx[((size_t)s * 8 + k) * G.n2 + fx];
      a += Sy[((size_t)s * 8 + k) * G.n2 + fy];
    }
    wb[k] = a;
    u[k] = Q[k * G.n2 + c];
    ub[k] = 0.;
  }
  prim_vjp(u, G.g, wb, ub);
  for (int k = 0; k < 8; k++) Qb[k * G.n2 + c] += ub[k];
}

// Qb = alpha * Rb (+ the previous Qb when accumulate)
__global__ __launch_bounds__(kT) void k_axpy(size_t n, double alpha, const double* __restrict__ Rb,
                                             double* __restrict__ Qb, int accumulate) {
  const size_t c = (size_t)blockIdx.x * kT + threadIdx.x;
  if (c >= n) return;
  Qb[c] = accumulate ? Qb[c] + alpha * Rb[c] : alpha * Rb[c];
}

unsigned blocks(size_t n) { return (unsigned)((n + kT - 1) / kT); }

}  // namespace

struct amr_adj {
  Grid G;
  // forward intermediates
  double *W, *Fx, *Fy, *E, *Us, *Bxs, *Bys, *pre1, *pre2, *Uo, *Bxo, *Byo;
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

// forward stage sequence; keeps Us (post-sync), Bxs, Bys, the pre-sync cell
// B of both stages (pre1, pre2) and, for the adjoint, the stage-2 faces
int forward(amr_adj* a, const double* U, const double* Bx, const double* By, double dt, double* Uo, double* Bxo,
            double* Byo, hipStream_t s) {
  const Grid& G = a->G;
  const unsigned b1 = blocks(G.n2), b2 = blocks(2 * G.n2);
  k_prim<<<b1, kT, 0, s>>>(G, U, a->W);
  k_flux<<<b2, kT, 0, s>>>(G, a->W, Bx, By, a->Fx, a->Fy);
  k_update<<<b1, kT, 0, s>>>(G, U, nullptr, a->Fx, a->Fy, dt, 1, a->Us);
  k_emf<<<b1, kT, 0, s>>>(G, a->Fx, a->Fy, a->E);
  k_faces<<<b1, kT, 0, s>>>(G, Bx, By, nullptr, nullptr, a->E, dt, 1, a->Bxs, a->Bys);
  k_bsync<<<b1, kT, 0, s>>>(G, a->Us, a->Bxs, a->Bys, a->pre1);
  k_prim<<<b1, kT, 0, s>>>(G, a->Us, a->W);
  k_flux<<<b2, kT, 0, s>>>(G, a->W, a->Bxs, a->Bys, a->Fx, a->Fy);
  k_update<<<b1, kT, 0, s>>>(G, U, a->Us, a->Fx, a->Fy, dt, 2, Uo);
  k_emf<<<b1, kT, 0, s>>>(G, a->Fx, a->Fy, a->E);
  k_faces<<<b1, kT, 0, s>>>(G, Bx, By, a->Bxs, a->Bys, a->E, dt, 2, Bxo, Byo);
  k_bsync<<<b1, kT, 0, s>>>(G, Uo, Bxo, Byo, a->pre2);
  ADJ_CHK(hipGetLastError());
  return 0;
}

}  // namespace

extern "C" {

int amr_adj_create(int n, double dx, double gamma, amr_adj** out) {
  if (!out || n < 4 || dx <= 0) return AMR_ERR_CONFIG;
  amr_adj* a = new amr_adj();
  a->G = Grid{n, dx, gamma, (size_t)n * n};
  const size_t n2 = a->G.n2;
  double** bufs8[] = {&a->W, &a->Fx, &a->Fy, &a->Us, &a->Uo, &a->Qb2, &a->Usb, &a->Fbx, &a->Fby};
  double** bufs1[] = {&a->E, &a->Bxs, &a->Bys, &a->Bxo, &a->Byo, &a->Eb, &a->Bsbx, &a->Bsby, &a->Bbx, &a->Bby};
  bool ok = true;
  for (auto p : bufs8) ok = ok && hipMalloc((void**)p, 8 * n2 * sizeof(double)) == hipSuccess;
  for (auto p : bufs1) ok = ok && hipMalloc((void**)p, n2 * sizeof(double)) == hipSuccess;
  ok = ok && hipMalloc((void**)&a->pre1, 2 * n2 * sizeof(double)) == hipSuccess;
  ok = ok && hipMalloc((void**)&a->pre2, 2 * n2 * sizeof(double)) == hipSuccess;
  ok = ok && hipMalloc((void**)&a->bb, 2 * n2 * sizeof(double)) == hipSuccess;
  ok = ok && hipMalloc((void**)&a->Sx, 32 * n2 * sizeof(double)) == hipSuccess;
  ok = ok && hipMalloc((void**)&a->Sy, 32 * n2 * sizeof(double)) == hipSuccess;
  if (!ok) {
    amr_adj_destroy(a);
    return AMR_ERR_HIP;
  }
  *out = a;
  return AMR_OK;
}

int amr_adj_destroy(amr_adj* a) {
  if (!a) return AMR_OK;
  double* all[] = {a->W,   a->Fx,   a->Fy,   a->E,    a->Us,   a->Bxs, a->Bys, a->pre1, a->pre2, a->Uo,  a->Bxo,
                   a->Byo, a->Qb2,  a->Usb,  a->Fbx,  a->Fby,  a->Eb,  a->Bsbx, a->Bsby, a->Bbx,  a->Bby, a->bb,
                   a->Sx,  a->Sy};
  for (double* p : all)
    if (p) hipFree(p);
  delete a;
  return AMR_OK;
}

int amr_adj_step(amr_adj* a, const double* U, const double* Bx, const double* By, double dt, double* Uo, double* Bxo,
                 double* Byo, void* stream) {
  if (!a || !U || !Bx || !By || !Uo || !Bxo || !Byo) return AMR_ERR_CONFIG;
  return forward(a, U, Bx, By, dt, Uo, Bxo, Byo, (hipStream_t)stream);
}

int amr_adj_step_vjp(amr_adj* a, const double* U, const double* Bx, const double* By, double dt, const double* Uob,
                     const double* Bxob, const double* Byob, double* Ub, double* Bxb, double* Byb, void* stream) {
  if (!a || !U || !Bx || !By || !Uob || !Bxob || !Byob || !Ub || !Bxb || !Byb) return AMR_ERR_CONFIG;
  hipStream_t s = (hipStream_t)stream;
  const Grid& G = a->G;
  const size_t n2 = G.n2;
  const unsigned b1 = blocks(n2), b2 = blocks(2 * n2), b8 = blocks(8 * n2);
  // recompute the step, keeping its intermediates
  if (int e = forward(a, U, Bx, By, dt, a->Uo, a->Bxo, a->Byo, s)) return e;
  ADJ_CHK(hipMemsetAsync(Ub, 0, 8 * n2 * sizeof(double), s));
  ADJ_CHK(hipMemsetAsync(Bxb, 0, n2 * sizeof(double), s));
  ADJ_CHK(hipMemsetAsync(Byb, 0, n2 * sizeof(double), s));
  // ---- stage 2 (reverse)
  k_axpy<<<b8, kT, 0, s>>>(8 * n2, 1., Uob, a->Qb2, 0);
  k_bsync_vjp<<<b1, kT, 0, s>>>(G, a->Qb2, a->pre2, a->Bxo, a->Byo, a->bb);
  ADJ_CHK(hipMemcpyAsync(a->Bbx, Bxob, n2 * sizeof(double), hipMemcpyDeviceToDevice, s));
  ADJ_CHK(hipMemcpyAsync(a->Bby, Byob, n2 * sizeof(double), hipMemcpyDeviceToDevice, s));
  k_faces_avg_vjp<<<b1, kT, 0, s>>>(G, a->Bbx, a->Bby, a->bb);
  k_faces_vjp<<<b1, kT, 0, s>>>(G, a->Bbx, a->Bby, dt, 2, Bxb, Byb, a->Bsbx, a->Bsby);
  k_emf_b<<<b1, kT, 0, s>>>(G, a->Bbx, a->Bby, dt, 2, a->Eb);
  k_flux_b<<<b1, kT, 0, s>>>(G, a->Qb2, a->Eb, dt, 2, a->Fbx, a->Fby);
  // U and Us each take half of the stage-2 output adjoint
  k_axpy<<<b8, kT, 0, s>>>(8 * n2, 0.5, a->Qb2, Ub, 1);
  k_axpy<<<b8, kT, 0, s>>>(8 * n2, 0.5, a->Qb2, a->Usb, 0);
  k_prim<<<b1, kT, 0, s>>>(G, a->Us, a->W);
  k_flux_vjp<<<b2, kT, 0, s>>>(G, a->W, a->Bxs, a->Bys, a->Fbx, a->Fby, a->Sx, a->Sy, a->Bsbx, a->Bsby);
  k_gather_prim_vjp<<<b1, kT, 0, s>>>(G, a->Us, a->Sx, a->Sy, a->Usb);
  // ---- stage 1 (reverse)
  k_bsync_vjp<<<b1, kT, 0, s>>>(G, a->Usb, a->pre1, a->Bxs, a->Bys, a->bb);
  k_faces_avg_vjp<<<b1, kT, 0, s>>>(G, a->Bsbx, a->Bsby, a->bb);
  k_faces_vjp<<<b1, kT, 0, s>>>(G, a->Bsbx, a->Bsby, dt, 1, Bxb, Byb, nullptr, nullptr);
  k_emf_b<<<b1, kT, 0, s>>>(G, a->Bsbx, a->Bsby, dt, 1, a->Eb);
  k_flux_b<<<b1, kT, 0, s>>>(G, a->Usb, a->Eb, dt, 1, a->Fbx, a->Fby);
  k_axpy<<<b8, kT, 0, s>>>(8 * n2, 1., a->Usb, Ub, 1);
  k_prim<<<b1, kT, 0, s>>>(G, U, a->W);
  k_flux_vjp<<<b2, kT, 0, s>>>(G, a->W, Bx, By, a->Fbx, a->Fby, a->Sx, a->Sy, Bxb, Byb);
  k_gather_prim_vjp<<<b1, kT, 0, s>>>(G, U, a->Sx, a->Sy, Ub);
  ADJ_CHK(hipGetLastError());
  return AMR_OK;
}

// host entry points of the pointwise reverse pieces (CPU tests)
void amr_debug_hlle_vjp(const double* wL, const double* wR, double gamma, const double* fb, double* wLb, double* wRb) {
  for (int k = 0; k < 8; k++) wLb[k] = wRb[k] = 0.;
  hlle_vjp(wL, wR, gamma, fb, wLb, wRb);
}

void amr_debug_hlle(const double* wL, const double* wR, double gamma, double* f) { flux_hlle(wL, wR, gamma, f); }

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
