// Pointwise FP64 MHD physics for the stage kernels (device and host).
//
// Every expression keeps the operand order and association of the reference
// implementation (spaceauk/AMR2D_MHD) because the results must be bitwise
// equal to it; the file is compiled with -ffp-contract=off, so no FMA is
// formed.  References are file:line in the reference tree.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "amr_types.hpp"

namespace amr {

constexpr double kEps = 1.0E-15;  // defs.hpp:16

// defs.hpp:110-111 (MIN/MAX macros) and std::min/std::max: they differ on
// signed-zero ties and NaN, and the reference uses each at specific sites.
__host__ __device__ inline double mac_min(double a, double b) { return a < b ? a : b; }
__host__ __device__ inline double mac_max(double a, double b) { return a > b ? a : b; }
__host__ __device__ inline double std_min(double a, double b) { return (b < a) ? b : a; }
__host__ __device__ inline double std_max(double a, double b) { return (a < b) ? b : a; }

// conserved -> primitive of one cell (meshblock.cpp:119-179); the caller
// applies the reference's interior p <= 0 check to w[4]
__host__ __device__ inline void cons_to_prim(const double* u, double g, double* w) {
  w[0] = u[0];
  w[1] = u[1] / u[0];
  w[2] = u[2] / u[0];
  w[3] = u[3] / u[0];
  double p = (g - 1) * (u[4] - 0.5 * w[0] * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]));
  w[5] = u[5];
  w[6] = u[6];
  w[7] = u[7];
  w[4] = p - (g - 1) * (0.5 * (u[5] * u[5] + u[6] * u[6] + u[7] * u[7]));
}

// slopelimiter.cpp:6-54
__host__ __device__ inline double limiter(int id, double r, double eta) {
  switch (id) {
    case kLim1st: return 0.;
    case kLimNoSL: return r;
    case kLimMinmod: return mac_max(0., mac_min(1., r));
    case kLimVanLeer: return (fabs(r) + r) / (1. + fabs(r));
    case kLimSuperbee: {
      double lo = mac_min(2. * r, 1.);
      double hi = mac_min(r, 2.);
      return mac_max(0., mac_max(hi, lo));
    }
    case kLimMC: {
      double c = mac_min(2. * r, 0.5 * (1. + r));
      return mac_max(0., mac_min(c, 2.));
    }
    case kLimKoren: {
      double b = mac_min((1. + 2. * r) / 3., 2.);
      return mac_max(0., mac_min(2. * r, b));
    }
    case kLimCada: {
      double third = (2. + r) / 3.;
      double h = mac_max(-0.5 * r, mac_min(mac_min(2. * r, third), 1.6));
      h = mac_max(0., mac_min(third, h));
      if (eta <= 1. - kEps) return third;
      if (eta >= 1. + kEps) return h;
      return 0.5 * ((1. - (eta - 1.) / kEps) * (2. + r) / 3. + (1 + (eta - 1.) / kEps) * h);
    }
    default: return NAN;
  }
}

// misc.cpp:71-79, returns the fast magnetosonic speed along component `norm`
__host__ __device__ inline double fast_speed(const double* w, double g, int norm) {
  double ca = sqrt((w[5] * w[5] + w[6] * w[6] + w[7] * w[7]) / w[0]);
  double a = sqrt(g * w[4] / w[0]);
  double can = sqrt((w[norm] * w[norm]) / w[0]);
  double s0 = (ca * ca) + (a * a);
  double s1 = (ca * ca) - (a * a);
  return sqrt(0.5 * (s0 + sqrt(s1 * s1 + 4. * (a * a) * ((ca * ca) - (can * can)))));
}

__host__ __device__ inline double sign0(double v) { return v > 0 ? 1. : (v < 0 ? -1. : 0.); }

// conserved state and x-flux of one side (riemannS.cpp:60-108)
struct Side {
  double q[8], F[8], e, H, pt, cf;
};

__host__ __device__ inline void side_state(const double* w, double g, Side& s) {
  s.cf = fast_speed(w, g, 5);
  const double vn = w[1], Bn = w[5];
  const double b2 = w[5] * w[5] + w[6] * w[6] + w[7] * w[7];
  double e = (w[4] / (w[0] * (g - 1))) + 0.5 * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]) + 0.5 * b2 / w[0];
  e = e * w[0];
  s.e = e;
  s.H = (e + w[4] + 0.5 * b2) / w[0];
#pragma unroll
  for (int k = 0; k < 8; k++) s.q[k] = w[k];
  s.q[1] = w[0] * w[1];
  s.q[2] = w[0] * w[2];
  s.q[3] = w[0] * w[3];
  s.q[4] = e;
  s.pt = w[4] + 0.5 * b2;
  const double mv = w[0] * vn;
  s.F[0] = mv;
  s.F[1] = mv * w[1] + s.pt - Bn * w[5];
  s.F[2] = mv * w[2] - Bn * w[6];
  s.F[3] = mv * w[3] - Bn * w[7];
  s.F[4] = mv * s.H - Bn * (w[1] * w[5] + w[2] * w[6] + w[3] * w[7]);
  s.F[5] = vn * w[5] - Bn * w[1];
  s.F[6] = vn * w[6] - Bn * w[2];
  s.F[7] = vn * w[7] - Bn * w[3];
}

// Roe-average outer eigenvalues, misc.cpp:95-140 (including the reference's
// sqrt(rhoL + sqrt(rhoR)) normaliser).  Outputs u_roe, lambda_min, lambda_max.
__host__ __device__ inline void roe_bounds(const double* wL, const double* wR, double g, double HL, double HR,
                                           double& u, double& lmin, double& lmax) {
  const double sl = sqrt(wL[0]), sr = sqrt(wR[0]);
  const double nrm = sqrt(wL[0] + sqrt(wR[0]));
  double v[3], B[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    v[i] = (sl * wL[i + 1] + sr * wR[i + 1]) / nrm;
    B[i] = (sl * wL[i + 5] + sr * wR[i + 5]) / nrm;
  }
  const double RT = sqrt(wR[0] / wL[0]);
  const double d = sqrt(wR[0] * wL[0]);
  const double d6 = wL[6] - wR[6], d7 = wL[7] - wR[7];
  const double X = 0.5 * ((d6 * d6) + (d7 * d7)) / (sl + sr);
  const double Y = 0.5 * (wL[0] + wR[0]) / d;
  const double h = (HL + RT * HR) / (1 + RT);
  const double vsq = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  const double btsq = (B[1] * B[1]) + (B[2] * B[2]);
  const double bt2s = (g - 1. - (g - 2) * Y) * btsq;
  const double vax2 = (B[0] * B[0]) / d;
  const double hp = h - (vax2 + btsq / d);
  double a2 = ((g - 1.) * (hp - 0.5 * vsq) - (g - 2.) * X);
  a2 = mac_max(a2, kEps);
  const double ct2 = bt2s / d;
  const double tsum = vax2 + ct2 + a2;
  const double tdif = vax2 + ct2 - a2;
  const double cfast = sqrt(0.5 * (tsum + sqrt(tdif * tdif + 4. * a2 * ct2)));
  u = v[0];
  lmin = v[0] - cfast;
  lmax = v[0] + cfast;
}

// riemannS.cpp:58-116
__host__ __device__ inline void flux_rusanov(const double* wL, const double* wR, double g, double* f) {
  Side L, R;
  side_state(wL, g, L);
  side_state(wR, g, R);
  const double s = std_max(L.cf + fabs(wL[1]), R.cf + fabs(wR[1]));
#pragma unroll
  for (int k = 0; k < 8; k++) f[k] = 0.5 * (L.F[k] + R.F[k]) - 0.5 * s * (R.q[k] - L.q[k]);
}

// riemannS.cpp:118-189
__host__ __device__ inline void flux_hlle(const double* wL, const double* wR, double g, double* f) {
  Side L, R;
  side_state(wL, g, L);
  side_state(wR, g, R);
  double u, lmin, lmax;
  roe_bounds(wL, wR, g, L.H, R.H, u, lmin, lmax);
  const double SL = mac_min(std_min(wL[1] - L.cf, lmin), 0.);
  const double SR = mac_max(std_max(wR[1] + R.cf, lmax), 0.);
#pragma unroll
  for (int k = 0; k < 8; k++) f[k] = (SR * L.F[k] - SL * R.F[k] + SL * SR * (R.q[k] - L.q[k])) / (SR - SL);
}

// riemannS.cpp:192-295 (Linde 1998 alpha heuristic); returns nonzero where the
// reference throws (alpha NaN or SR == SL).
__host__ __device__ inline int flux_hllc(const double* wL, const double* wR, double g, double* f) {
  Side L, R;
  side_state(wL, g, L);
  side_state(wR, g, R);
  double u, lmin, lmax;
  roe_bounds(wL, wR, g, L.H, R.H, u, lmin, lmax);
  const double SL = std_min(wL[1] - L.cf, lmin);
  const double SR = std_max(wR[1] + R.cf, lmax);
  double jump = 0., resid = 0.;
#pragma unroll
  for (int k = 0; k < 8; k++) jump += fabs(R.q[k] - L.q[k]);
#pragma unroll
  for (int k = 0; k < 8; k++) resid += fabs((R.F[k] - L.F[k]) - u * (R.q[k] - L.q[k]));
  const double alpha = mac_max(0., 1. - resid / (jump + kEps));
  if (alpha > 1 || alpha < 0 || isnan(alpha)) return 1;
  if (SR - SL == 0) return 1;
  const double cL = (SL * ((1 - alpha) * SR + alpha * u)) / (SR - SL);
  const double cR = (SR * ((1 - alpha) * SL + alpha * u)) / (SR - SL);
  const double aR = SL / (SR - SL), aL = (SR) / (SR - SL);
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const double dq = R.q[k] - L.q[k];
    if (SL > 0.) f[k] = L.F[k];
    else if (SL <= 0. && u >= 0.) f[k] = cL * dq - aR * R.F[k] + aL * L.F[k];
    else if (u <= 0. && SR >= 0.) f[k] = cR * dq - aR * R.F[k] + aL * L.F[k];
    else if (SR < 0.) f[k] = R.F[k];
  }
  return 0;
}

__host__ __device__ inline double vdotb(const double* w) {
  double s = 0.;
  s += w[1] * w[5];
  s += w[2] * w[6];
  s += w[3] * w[7];
  return s;
}

// star state of one side, riemannS.cpp:467-497; ws[4] stays 0
__host__ __device__ inline void hlld_star(const double* w, double S, double SM, double pts, double e, double pt,
                                          double* ws, double& es) {
  ws[0] = w[0] * (S - w[1]) / (S - SM);
  ws[1] = SM;
  ws[4] = 0.;
  ws[5] = w[5];
  const double den = w[0] * (S - w[1]) * (S - SM) - (w[5] * w[5]);
  if (den != 0) {
    const double sv = (SM - w[1]) / den;
    ws[2] = w[2] - w[5] * w[6] * sv;
    ws[3] = w[3] - w[5] * w[7] * sv;
    const double sb = (w[0] * ((S - w[1]) * (S - w[1])) - (w[5] * w[5])) / den;
    ws[6] = w[6] * sb;
    ws[7] = w[7] * sb;
  } else {
    ws[2] = w[2];
    ws[3] = w[3];
    ws[6] = 0.;
    ws[7] = 0.;
  }
  es = (((S - w[1]) * e - pt * w[1] + pts * SM + w[5] * (vdotb(w) - vdotb(ws))) / (S - SM));
}

// conserved vector of a star state whose energy is supplied (misc.cpp:82-92 + override)
__host__ __device__ inline void star_cons(const double* w, double e, double* u) {
  u[0] = w[0];
  u[1] = w[0] * w[1];
  u[2] = w[0] * w[2];
  u[3] = w[0] * w[3];
  u[4] = e;
  u[5] = w[5];
  u[6] = w[6];
  u[7] = w[7];
}

// scalars of side_state only (the conserved vector and flux are formed where
// a flux region needs them, with the same expressions)
struct SideScalars {
  double e, H, pt, cf;
};

__host__ __device__ inline void side_scalars(const double* w, double g, SideScalars& s) {
  s.cf = fast_speed(w, g, 5);
  const double b2 = w[5] * w[5] + w[6] * w[6] + w[7] * w[7];
  double e = (w[4] / (w[0] * (g - 1))) + 0.5 * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]) + 0.5 * b2 / w[0];
  e = e * w[0];
  s.e = e;
  s.H = (e + w[4] + 0.5 * b2) / w[0];
  s.pt = w[4] + 0.5 * b2;
}

// side_state's x-flux from its scalars
__host__ __device__ inline void side_flux(const double* w, const SideScalars& s, double* F) {
  const double vn = w[1], Bn = w[5];
  const double mv = w[0] * vn;
  F[0] = mv;
  F[1] = mv * w[1] + s.pt - Bn * w[5];
  F[2] = mv * w[2] - Bn * w[6];
  F[3] = mv * w[3] - Bn * w[7];
  F[4] = mv * s.H - Bn * (w[1] * w[5] + w[2] * w[6] + w[3] * w[7]);
  F[5] = vn * w[5] - Bn * w[1];
  F[6] = vn * w[6] - Bn * w[2];
  F[7] = vn * w[7] - Bn * w[3];
}

// riemannS.cpp:397-549 (Miyoshi & Kusano 2005 HLLD).  The flux region is
// decided from the wave speeds and the star densities alone (the first
// component of each star state, riemannS.cpp:467); the star states, the
// double-star state, and the side flux and conserved vectors are formed only
// in the region that uses them -- fewer live registers in the stage kernel,
// the same operations per value.
__host__ __device__ inline void flux_hlld(const double* wL, const double* wR, double g, double* f) {
  SideScalars L, R;
  side_scalars(wL, g, L);
  side_scalars(wR, g, R);
  const double vL = wL[1], vR = wR[1];
  const double cmax = std_max(L.cf, R.cf);
  const double SL = std_min(vL, vR) - cmax;
  const double SR = std_max(vL, vR) + cmax;
  const double den = (SR - vR) * wR[0] - (SL - vL) * wL[0];
  const double SM = ((SR - vR) * wR[0] * vR - (SL - vL) * wL[0] * vL - R.pt + L.pt) / den;
  const double pts = ((SR - vR) * wR[0] * L.pt - (SL - vL) * wL[0] * R.pt +
                      wL[0] * wR[0] * (SR - vR) * (SL - vL) * (vR - vL)) / den;
  const double rhoLs = wL[0] * (SL - wL[1]) / (SL - SM);  // = hlld_star's ws[0]
  const double rhoRs = wR[0] * (SR - wR[1]) / (SR - SM);
  const double rl = sqrt(rhoLs), rr = sqrt(rhoRs);
  const double SLs = SM - fabs(wL[5]) / rl;
  const double SRs = SM + fabs(wR[5]) / rr;
  int region = 0;  // 0: no branch taken (NaN speeds), f keeps its zeros
  if (SL > 0) region = 1;
  else if (SL <= 0. && SLs >= 0.) region = 2;
  else if (SLs <= 0. && SM >= 0.) region = 3;
  else if (SM <= 0. && SRs >= 0.) region = 4;
  else if (SRs <= 0. && SR >= 0.) region = 5;
  else if (SR < 0.) region = 6;
  if (region == 1) {
    side_flux(wL, L, f);
    return;
  }
  if (region == 6) {
    side_flux(wR, R, f);
    return;
  }
  if (region == 0) return;
  double F[8], q[8], qs[8];
  if (region == 2) {
    double wLs[8], eLs;
    hlld_star(wL, SL, SM, pts, L.e, L.pt, wLs, eLs);
    side_flux(wL, L, F);
    star_cons(wL, L.e, q);
    star_cons(wLs, eLs, qs);
#pragma unroll
    for (int k = 0; k < 8; k++) f[k] = F[k] + SL * qs[k] - SL * q[k];
    return;
  }
  if (region == 5) {
    double wRs[8], eRs;
    hlld_star(wR, SR, SM, pts, R.e, R.pt, wRs, eRs);
    side_flux(wR, R, F);
    star_cons(wR, R.e, q);
    star_cons(wRs, eRs, qs);
#pragma unroll
    for (int k = 0; k < 8; k++) f[k] = F[k] + SR * qs[k] - SR * q[k];
    return;
  }
  // regions 3 and 4: the double-star state
  double wLs[8], wRs[8], eLs, eRs;
  hlld_star(wL, SL, SM, pts, L.e, L.pt, wLs, eLs);
  hlld_star(wR, SR, SM, pts, R.e, R.pt, wRs, eRs);
  const double rlr = sqrt(wLs[0] * wRs[0]);
  const double sg = sign0(wL[5]);
  double wss[8], qss[8];
  wss[0] = wLs[0];
  wss[1] = SM;
  wss[2] = (rl * wLs[2] + rr * wRs[2] + (wRs[6] - wLs[6]) * sg) / (rl + rr);
  wss[3] = (rl * wLs[3] + rr * wRs[3] + (wRs[7] - wLs[7]) * sg) / (rl + rr);
  wss[4] = 0.;
  wss[5] = wLs[5];
  wss[6] = (rl * wRs[6] + rr * wLs[6] + rlr * (wRs[2] - wLs[2]) * sg) / (rl + rr);
  wss[7] = (rl * wRs[7] + rr * wLs[7] + rlr * (wRs[3] - wLs[3]) * sg) / (rl + rr);
  const double dss = vdotb(wss);
  if (region == 3) {
    const double eLss = eLs - rl * (vdotb(wLs) - dss) * sg;
    side_flux(wL, L, F);
    star_cons(wL, L.e, q);
    star_cons(wLs, eLs, qs);
    star_cons(wss, eLss, qss);
#pragma unroll
    for (int k = 0; k < 8; k++) f[k] = F[k] + SLs * qss[k] - (SLs - SL) * qs[k] - SL * q[k];
  } else {
    const double eRss = eRs + rr * (vdotb(wRs) - dss) * sign0(wR[5]);
    side_flux(wR, R, F);
    star_cons(wR, R.e, q);
    star_cons(wRs, eRs, qs);
    star_cons(wss, eRss, qss);
#pragma unroll
    for (int k = 0; k < 8; k++) f[k] = F[k] + SRs * qss[k] - (SRs - SR) * qs[k] - SR * q[k];
  }
}

// x-normal flux of the selected solver; a y face is computed by the caller
// with (1,2) and (5,6) exchanged on input and output (riemannS.cpp:12-55).
// Components a solver leaves unset stay 0.  Returns nonzero where the
// reference throws.
__host__ __device__ inline int normal_flux(int solver, const double* wL, const double* wR, double g, double* f) {
#pragma unroll
  for (int k = 0; k < 8; k++) f[k] = 0.;
  switch (solver) {
    case kFluxRusanov: flux_rusanov(wL, wR, g, f); return 0;
    case kFluxHLLE: flux_hlle(wL, wR, g, f); return 0;
    case kFluxHLLC: return flux_hllc(wL, wR, g, f);
    case kFluxHLLD: flux_hlld(wL, wR, g, f); return 0;
    default: return 2;
  }
}

__host__ __device__ inline void swap_xy(double* w) {
  double t = w[1];
  w[1] = w[2];
  w[2] = t;
  t = w[5];
  w[5] = w[6];
  w[6] = t;
}

// WENO5 edge states of one cell from its 5-point stencil
// (WENO2D_primvar.cpp:95-198; beta_1's first term uses w_{i-2} as there).
// left=true: state at i+1/2 (wxL); left=false: state at i-1/2 (wxR).
__host__ __device__ inline double weno5_edge(double m2, double m1, double c, double p1, double p2, bool left) {
  const double ew = 1.E-15;
  double t0 = m2 - 2 * m1 + c, u0 = m2 - 4 * m1 + 3 * c;
  double t1 = m2 - 2 * c + p1, u1 = m1 - p1;
  double t2 = c - 2 * p1 + p2, u2 = 3 * c - 4 * p1 + p2;
  const double e0 = ew + ((13. / 12.) * (t0 * t0) + (1. / 4.) * (u0 * u0));
  const double e1 = ew + ((13. / 12.) * (t1 * t1) + (1. / 4.) * (u1 * u1));
  const double e2 = ew + ((13. / 12.) * (t2 * t2) + (1. / 4.) * (u2 * u2));
  const double s0 = e0 * e0, s1 = e1 * e1, s2 = e2 * e2;
  if (!left) {
    const double a0 = (1. / 10.) / s0, a1 = (6. / 10.) / s1, a2 = (3. / 10.) / s2;
    const double as = a0 + a1 + a2;
    return a0 / as * ((-1 * m2 + 5 * m1 + 2 * c) / 6) + a1 / as * ((2 * m1 + 5 * c - p1) / 6) +
           a2 / as * ((11 * c - 7 * p1 + 2 * p2) / 6);
  }
  const double a0 = (3. / 10.) / s0, a1 = (6. / 10.) / s1, a2 = (1. / 10.) / s2;
  const double as = a0 + a1 + a2;
  return a0 / as * ((2 * m2 - 7 * m1 + 11 * c) / 6) + a1 / as * ((-1 * m1 + 5 * c + 2 * p1) / 6) +
         a2 / as * ((2 * c + 5 * p1 - p2) / 6);
}

// WENO3 edge states (WENO2D_primvar.cpp:7-91)
__host__ __device__ inline double weno3_edge(double m1, double c, double p1, bool left) {
  const double ew = 1.E-15;
  const double e0 = ew + (c - m1) * (c - m1), e1 = ew + (p1 - c) * (p1 - c);
  const double s0 = e0 * e0, s1 = e1 * e1;
  if (!left) {
    const double a0 = (2. / 3.) / s0, a1 = (1. / 3.) / s1, as = a0 + a1;
    return a0 / as * (0.5 * m1 + 0.5 * c) + a1 / as * (1.5 * c - 0.5 * p1);
  }
  const double a0 = (1. / 3.) / s0, a1 = (2. / 3.) / s1, as = a0 + a1;
  return a0 / as * (1.5 * c - 0.5 * m1) + a1 / as * (0.5 * p1 + 0.5 * c);
}

// Limited linear edge states of one variable (celledges.cpp:29-64):
// `left` is the state at i+1/2 seen from cell i (slopes d0 = w_i - w_{i-1},
// d1 = w_{i+1} - w_i), `right` the state at i-1/2 seen from cell i.
__host__ __device__ inline double muscl_plus(int lim, double w, double d0, double d1, double dxl) {
  const double eta = (lim == kLimCada) ? ((d1 * d1) + (d0 * d0)) / ((1. * dxl) * (1. * dxl)) : 0.;
  const double phi = limiter(lim, d1 / (d0 + kEps), eta);
  return w + 0.5 * phi * d0;
}
__host__ __device__ inline double muscl_minus(int lim, double w, double d0, double d1, double dxl) {
  const double eta = (lim == kLimCada) ? ((d1 * d1) + (d0 * d0)) / ((1. * dxl) * (1. * dxl)) : 0.;
  const double phi = limiter(lim, d0 / (d1 + kEps), eta);
  return w - 0.5 * phi * d1;
}

}  // namespace amr
