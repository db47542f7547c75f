// Operation count of the stage arithmetic (host only).  The pointwise
// physics of the stage kernels (physics.hpp) is compiled a second time in
// this translation unit with its floating-point type replaced by Counted, a
// double that tallies every +, -, *, / and sqrt it performs and otherwise
// computes exactly what a double computes (same IEEE operations, same
// order; the file is built with -ffp-contract=off like the kernels).
// count_stage_tile then runs one RK stage of one leaf tile the way
// stage_kernel does -- primitives of the whole tile, the x and y faces of
// [LO, HI]^2 (reconstruction, normal-B override, Riemann flux), and the
// update of every cell -- on the tile's real data, so that data-dependent
// branches (limiter cases, HLLD regions) are counted as they are taken.
// The result is the algorithmic operation count of the reference's stage
// (MUSCL2D.cpp:13-121, riemannS.cpp, main.cpp:147-157 / 179-189), independent
// of how the GPU expands a division or square root into instructions.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "amr_types.hpp"
#include "flop_count.hpp"

namespace amr {

namespace {
thread_local StageFlops g_tally;  // the counts of the calling thread's current count
}

struct Counted {
  double v;
  constexpr Counted() : v(0.) {}
  constexpr Counted(double x) : v(x) {}
};
inline Counted operator+(Counted a, Counted b) { g_tally.add += 1; return Counted(a.v + b.v); }
inline Counted operator-(Counted a, Counted b) { g_tally.add += 1; return Counted(a.v - b.v); }
inline Counted operator*(Counted a, Counted b) { g_tally.mul += 1; return Counted(a.v * b.v); }
inline Counted operator/(Counted a, Counted b) { g_tally.div += 1; return Counted(a.v / b.v); }
inline Counted operator-(Counted a) { return Counted(-a.v); }  // a sign flip, not an operation
inline Counted& operator+=(Counted& a, Counted b) { a = a + b; return a; }
inline Counted& operator-=(Counted& a, Counted b) { a = a - b; return a; }
inline Counted& operator*=(Counted& a, Counted b) { a = a * b; return a; }
inline bool operator<(Counted a, Counted b) { return a.v < b.v; }
inline bool operator>(Counted a, Counted b) { return a.v > b.v; }
inline bool operator<=(Counted a, Counted b) { return a.v <= b.v; }
inline bool operator>=(Counted a, Counted b) { return a.v >= b.v; }
inline bool operator==(Counted a, Counted b) { return a.v == b.v; }
inline bool operator!=(Counted a, Counted b) { return a.v != b.v; }
inline Counted sqrt(Counted a) { g_tally.sqrt += 1; return Counted(std::sqrt(a.v)); }
inline Counted fabs(Counted a) { return Counted(std::fabs(a.v)); }
inline bool isnan(Counted a) { return std::isnan(a.v); }

}  // namespace amr

#define double Counted
#include "physics.hpp"
#undef double

namespace amr {

void count_stage_tile(int NT, int NG, const double* Qin, const double* B, double gamma, double dxl, double dt,
                      int lim, int solver, int step, StageFlops& acc) {
  const int N2 = NT * NT, LO = NG - 1, HI = NT - 1 - NG;
  g_tally = StageFlops{};
  const Counted g(gamma), dx(dxl), dtc(dt);
  std::vector<Counted> W(8 * (size_t)N2);
  for (int c = 0; c < N2; c++) {
    Counted u[8], w[8];
    for (int k = 0; k < 8; k++) u[k] = Counted(Qin[k * N2 + c]);
    cons_to_prim(u, g, w);
    for (int k = 0; k < 8; k++) W[k * (size_t)N2 + c] = w[k];
  }
  std::vector<Counted> Fx(8 * (size_t)N2), Fy(8 * (size_t)N2);
  for (int ydir = 0; ydir < 2; ydir++)
    for (int j = LO; j <= HI; j++)
      for (int i = LO; i <= HI; i++) {
        const int s = ydir ? NT : 1, c = j * NT + i;
        Counted wl[8], wr[8], fl[8];
        for (int k = 0; k < 8; k++) {
          const Counted* Wk = W.data() + k * (size_t)N2;
          if (lim == kLimWENO5) {
            wl[k] = weno5_edge(Wk[c - 2 * s], Wk[c - s], Wk[c], Wk[c + s], Wk[c + 2 * s], true);
            wr[k] = weno5_edge(Wk[c - s], Wk[c], Wk[c + s], Wk[c + 2 * s], Wk[c + 3 * s], false);
          } else if (lim == kLimWENO3) {
            wl[k] = weno3_edge(Wk[c - s], Wk[c], Wk[c + s], true);
            wr[k] = weno3_edge(Wk[c], Wk[c + s], Wk[c + 2 * s], false);
          } else {
            const Counted a = Wk[c] - Wk[c - s], b = Wk[c + s] - Wk[c], d = Wk[c + 2 * s] - Wk[c + s];
            wl[k] = muscl_plus(lim, Wk[c], a, b, dx);
            wr[k] = muscl_minus(lim, Wk[c + s], b, d, dx);
          }
        }
        if (!ydir) {
          wl[5] = wr[5] = Counted(B[c + 1]);
          normal_flux(solver, wl, wr, g, fl);
          for (int k = 0; k < 8; k++) Fx[k * (size_t)N2 + c + 1] = fl[k];
        } else {
          wl[6] = wr[6] = Counted(B[N2 + c + NT]);
          swap_xy(wl);
          swap_xy(wr);
          normal_flux(solver, wl, wr, g, fl);
          swap_xy(fl);
          for (int k = 0; k < 8; k++) Fy[k * (size_t)N2 + c + NT] = fl[k];
        }
      }
  // the RK2 update of every cell (the residual is zero outside [LO, HI]^2,
  // but every cell is written): stage 1 U - dt res, stage 2
  // 0.5 (U + Us - dt res); only the operations are counted here
  for (int c = 0; c < N2; c++) {
    const int i = c % NT, j = c / NT;
    const bool in = i >= LO && i <= HI && j >= LO && j <= HI;
    for (int k = 0; k < 8; k++) {
      Counted res(0.);
      if (in) {
        const Counted* F = Fx.data() + k * (size_t)N2;
        const Counted* G = Fy.data() + k * (size_t)N2;
        res = (F[c + 1] - F[c]) / dx + (G[c + NT] - G[c]) / dx;
      }
      const Counted q(Qin[k * N2 + c]);
      const Counted out = step == 1 ? q - dtc * res : 0.5 * (q + q - dtc * res);
      (void)out;
    }
  }
  acc.add += g_tally.add;
  acc.mul += g_tally.mul;
  acc.div += g_tally.div;
  acc.sqrt += g_tally.sqrt;
}

}  // namespace amr

// Host-only checks of the counting build: the x-normal flux of `solver` for
// one face computed with Counted (f, equal bit for bit to the double build's
// when the counting type performs the same operations) and the operations
// it took (counts[0..3] = add/sub, mul, div, sqrt).
extern "C" int amr_debug_count_flux(int solver, const double* wL, const double* wR, double gamma, double* f,
                                    double* counts) {
  using namespace amr;
  g_tally = StageFlops{};
  Counted l[8], r[8], o[8];
  for (int k = 0; k < 8; k++) {
    l[k] = Counted(wL[k]);
    r[k] = Counted(wR[k]);
  }
  const int st = normal_flux(solver, l, r, Counted(gamma), o);
  for (int k = 0; k < 8; k++) f[k] = o[k].v;
  counts[0] = g_tally.add;
  counts[1] = g_tally.mul;
  counts[2] = g_tally.div;
  counts[3] = g_tally.sqrt;
  return st;
}
