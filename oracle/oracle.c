/* CPU oracle for the AMR2D_MHD per-step hot path -- TEST INFRASTRUCTURE.
 *
 * Plain-C restatement of spaceauk/AMR2D_MHD's algorithm.  Each function cites
 * the reference file:line it follows.  Floating-point expressions keep the
 * reference's operand order and association (the only way to be bitwise equal
 * to the reference compiled with -ffp-contract=off); everything else (data
 * layout, control structure, error handling) is this file's own.  Build with
 * -O2 -ffp-contract=off and no fast-math.
 *
 * Pinned by tests/test_oracle_vs_ref.py against the compiled reference
 * (oracle/_ref, built by oracle/build_ref.sh) and the fixtures in tests/golden.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define EPS 1.0E-15          /* defs.hpp:16 (DOUBLE_PREC) */
#define PI_REF 3.14159265359 /* defs.hpp:23 */
#define NBROOTS 4            /* defs.hpp:29 */
#define MAXLEV 16

/* defs.hpp:110-111 macros and the std:: templates differ on ties/NaN; both
 * are used by the reference at specific sites, so both are kept. */
static inline double MINm(double a, double b) { return a < b ? a : b; }
static inline double MAXm(double a, double b) { return a > b ? a : b; }
static inline double stdmin(double a, double b) { return (b < a) ? b : a; }
static inline double stdmax(double a, double b) { return (a < b) ? b : a; }

enum { L_1ST, L_NOSL, L_MINMOD, L_VANLEER, L_SUPERBEE, L_MC, L_KOREN, L_CADA, L_WENO3, L_WENO5, L_BAD };
enum { F_RUSA, F_HLLE, F_HLLC, F_HLLD, F_BAD };
enum { IC_SST, IC_SST1, IC_SSTY, IC_BWX, IC_BWX1, IC_BWY, IC_BWY1, IC_ROTOR, IC_ROTOR_P, IC_OT, IC_ROTOR_BS, IC_BAD };

static int lim_id(const char *s) {
  if (!strcmp(s, "1st")) return L_1ST;
  if (!strcmp(s, "noSL") || !strcmp(s, "WENO2D")) return L_NOSL;
  if (!strcmp(s, "minmod")) return L_MINMOD;
  if (!strcmp(s, "VanLeer")) return L_VANLEER;
  if (!strcmp(s, "Superbee")) return L_SUPERBEE;
  if (!strcmp(s, "MC")) return L_MC;
  if (!strcmp(s, "Koren")) return L_KOREN;
  if (!strcmp(s, "Cada3rd")) return L_CADA;
  if (!strcmp(s, "WENO3")) return L_WENO3;
  if (!strcmp(s, "WENO5")) return L_WENO5;
  return L_BAD;
}
static int flux_id(const char *s) {
  if (!strcmp(s, "RUSA")) return F_RUSA;
  if (!strcmp(s, "HLLE")) return F_HLLE;
  if (!strcmp(s, "HLLC")) return F_HLLC;
  if (!strcmp(s, "HLLD")) return F_HLLD;
  return F_BAD;
}
static int ic_id(const char *s) {
  static const char *names[] = {"SST", "SST1", "SSTy", "BWx", "BWx1", "BWy", "BWy1", "rotor", "rotor_p", "OT",
                                "rotor_bs"};
  for (int q = 0; q < IC_BAD; q++)
    if (!strcmp(s, names[q])) return q;
  return IC_BAD;
}

static const int NVAR[ORA_NFIELD] = {8, 8, 8, 8, 8, 2, 2, 1};

struct ora_dom {
  int ng, nxi, NT, maxlevs, maxblocks;
  int nxmin, nxmax, nxminb, nxmaxb, nx2; /* parameters.cpp:10-35 (square tiles) */
  double gamma, tEnd, cfl;
  double dx[MAXLEV], dy[MAXLEV];
  int lim, flux, ic;
  int lastActive;
  int32_t (*lp)[8];
  int32_t (*icoord)[2];
  unsigned char *leafs, *iref, *icoarse;
  double *f[ORA_NFIELD];
  int32_t (*ib)[3];
  int nbounds;
  double t, dt_main, dt_dom; /* main.cpp locals t, dt and meshblock::dt */
  int count;
  double *scr; /* dwdx dwdy wxL wxR wyL wyR, each [8][NT][NT] */
  int32_t err[4];
  double icp[8];
  int periodic; /* build-defined: wrapped root neighbours, exact geometry */
  int rk;       /* 2 (timeIntegral.cpp) or 3 (build-defined SSP-RK3) */
  int ct;       /* CT2D.cpp corner EMF: 1 (CTtype 1) or 2 (CTtype 2, upwind) */
};

#define TILE(d, fld, nb) ((d)->f[fld] + (size_t)(nb) * NVAR[fld] * (d)->NT * (d)->NT)
/* element (k, i, j) of a tile; arguments ordered like the reference's [i][j][k] */
#define AT(t, NT, k, i, j) (t)[((size_t)(k) * (NT) + (j)) * (NT) + (i)]

static int fail(ora_dom *d, int code, int nb, int i, int j) {
  if (d && d->err[0] == 0) {
    d->err[0] = code; d->err[1] = nb; d->err[2] = i; d->err[3] = j;
  }
  return code;
}

/* ---------------------------------------------------------------- limiters */
/* slopelimiter.cpp:6-54 */
static double limiter(int id, double r, double eta) {
  double phi = 0.;
  switch (id) {
    case L_1ST: phi = 0.; break;
    case L_NOSL: phi = r; break;
    case L_MINMOD: { double a = MINm(1., r); phi = MAXm(0., a); } break;
    case L_VANLEER: phi = (fabs(r) + r) / (1. + fabs(r)); break;
    case L_SUPERBEE: {
      double a = 2. * r;
      double b = MINm(a, 1.);
      a = MINm(r, 2.);
      double c = MAXm(a, b);
      phi = MAXm(0., c);
    } break;
    case L_MC: {
      double a = 2. * r, b = 0.5 * (1. + r);
      double c = MINm(a, b);
      a = MINm(c, 2.);
      phi = MAXm(0., a);
    } break;
    case L_KOREN: {
      double a = (1. + 2. * r) / 3.;
      double b = MINm(a, 2.);
      a = 2. * r;
      double c = MINm(a, b);
      phi = MAXm(0., c);
    } break;
    case L_CADA: {
      double a = MINm(2. * r, (2. + r) / 3.);
      double b = MINm(a, 1.6);
      double h = MAXm(-0.5 * r, b);
      a = MINm((2. + r) / 3., h);
      h = MAXm(0., a);
      if (eta <= 1. - EPS) phi = (2. + r) / 3.;
      else if (eta >= 1. + EPS) phi = h;
      else phi = 0.5 * ((1. - (eta - 1.) / EPS) * (2. + r) / 3. + (1 + (eta - 1.) / EPS) * h);
    } break;
    default: phi = NAN;
  }
  return phi;
}

double ora_limiter(const char *name, double r, double eta) { return limiter(lim_id(name), r, eta); }

/* ---------------------------------------------------------- MHD pointwise */
/* misc.cpp:71-79 */
static void wavespeeds(const double *w, double g, int norm, double *a, double *ca, double *can, double *cf) {
  *ca = sqrt((w[5] * w[5] + w[6] * w[6] + w[7] * w[7]) / w[0]);
  *a = sqrt(g * w[4] / w[0]);
  *can = sqrt((w[norm] * w[norm]) / w[0]);
  double s0 = (*ca * *ca) + (*a * *a);
  double s1 = (*ca * *ca) - (*a * *a);
  *cf = sqrt(0.5 * (s0 + sqrt(s1 * s1 + 4. * (*a * *a) * ((*ca * *ca) - (*can * *can)))));
}

void ora_wavespeeds(const double *w, double gamma, int norm, double *o) {
  wavespeeds(w, gamma, norm, &o[0], &o[1], &o[2], &o[3]);
}

/* misc.cpp:64-68 */
static inline double sgn(double v) { return v > 0 ? 1. : (v < 0 ? -1. : 0.); }
/* misc.cpp:51-61 for the (1..3).(5..7) products used by HLLD */
static inline double dot_vb(const double *w) {
  double s = 0.;
  s += w[1] * w[5];
  s += w[2] * w[6];
  s += w[3] * w[7];
  return s;
}
/* misc.cpp:82-92 */
static void prim2cons(const double *w, double *u, double g) {
  u[0] = w[0];
  u[1] = w[0] * w[1];
  u[2] = w[0] * w[2];
  u[3] = w[0] * w[3];
  u[4] = w[4] / (g - 1) + 0.5 * w[0] * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]);
  u[4] += 0.5 * (w[5] * w[5] + w[6] * w[6] + w[7] * w[7]);
  u[5] = w[5]; u[6] = w[6]; u[7] = w[7];
}

/* per-side quantities shared by all four solvers (riemannS.cpp:60-108) */
typedef struct { double q[8], F[8], e, H, pt, cf; } side_t;

static void side_state(const double *w, double g, side_t *s) {
  double a, ca, can;
  wavespeeds(w, g, 5, &a, &ca, &can, &s->cf);
  double vn = w[1], Bn = w[5];
  double e = (w[4] / (w[0] * (g - 1))) + 0.5 * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]) +
             0.5 * (w[5] * w[5] + w[6] * w[6] + w[7] * w[7]) / w[0];
  e = e * w[0];
  s->e = e;
  s->H = (e + w[4] + 0.5 * (w[5] * w[5] + w[6] * w[6] + w[7] * w[7])) / w[0];
  for (int k = 0; k < 8; k++) s->q[k] = w[k];
  s->q[1] = w[0] * w[1]; s->q[2] = w[0] * w[2]; s->q[3] = w[0] * w[3];
  s->q[4] = e;
  s->pt = w[4] + 0.5 * (w[5] * w[5] + w[6] * w[6] + w[7] * w[7]);
  s->F[0] = w[0] * vn;
  s->F[1] = w[0] * vn * w[1] + s->pt - Bn * w[5];
  s->F[2] = w[0] * vn * w[2] - Bn * w[6];
  s->F[3] = w[0] * vn * w[3] - Bn * w[7];
  s->F[4] = w[0] * vn * s->H - Bn * (w[1] * w[5] + w[2] * w[6] + w[3] * w[7]);
  s->F[5] = vn * w[5] - Bn * w[1];
  s->F[6] = vn * w[6] - Bn * w[2];
  s->F[7] = vn * w[7] - Bn * w[3];
}

/* misc.cpp:95-107 (denominator sqrt(rhoL+sqrt(rhoR)) kept as in the reference)
 * and misc.cpp:110-140; only the two outer eigenvalues are consumed. */
static void roe_speeds(const double *wL, const double *wR, double g, double HL, double HR,
                       double *v0, double *lam0, double *lam6) {
  double v[3], B[3];
  for (int i = 0; i < 3; i++) {
    v[i] = (sqrt(wL[0]) * wL[i + 1] + sqrt(wR[0]) * wR[i + 1]) / (sqrt(wL[0] + sqrt(wR[0])));
    B[i] = (sqrt(wL[0]) * wL[i + 5] + sqrt(wR[0]) * wR[i + 5]) / (sqrt(wL[0] + sqrt(wR[0])));
  }
  double RT = sqrt(wR[0] / wL[0]);
  double r = sqrt(wR[0] * wL[0]);
  double d6 = wL[6] - wR[6], d7 = wL[7] - wR[7];
  double Xfac = 0.5 * ((d6 * d6) + (d7 * d7)) / (sqrt(wL[0]) + sqrt(wR[0]));
  double Yfac = 0.5 * (wL[0] + wR[0]) / r;
  double h = (HL + RT * HR) / (1 + RT);
  double vsq = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  double btsq = (B[1] * B[1]) + (B[2] * B[2]);
  double bt_starsq = (g - 1. - (g - 2) * Yfac) * btsq;
  double vaxsq = (B[0] * B[0]) / r;
  double hp = h - (vaxsq + btsq / r);
  double twid_asq = ((g - 1.) * (hp - 0.5 * vsq) - (g - 2.) * Xfac);
  twid_asq = MAXm(twid_asq, EPS);
  double ct2 = bt_starsq / r;
  double tsum = vaxsq + ct2 + twid_asq;
  double tdif = vaxsq + ct2 - twid_asq;
  double cf2_cs2 = sqrt(tdif * tdif + 4. * twid_asq * ct2);
  double cfsq = 0.5 * (tsum + cf2_cs2);
  double cfast = sqrt(cfsq);
  *v0 = v[0];
  *lam0 = v[0] - cfast;
  *lam6 = v[0] + cfast;
}

/* riemannS.cpp:58-116 */
static void rusanov(const double *wL, const double *wR, double g, double *flux) {
  side_t L, R;
  side_state(wL, g, &L);
  side_state(wR, g, &R);
  double sl = L.cf + fabs(wL[1]), sr = R.cf + fabs(wR[1]);
  double s = stdmax(sl, sr);
  for (int k = 0; k < 8; k++) flux[k] = 0.5 * (L.F[k] + R.F[k]) - 0.5 * s * (R.q[k] - L.q[k]);
}

/* riemannS.cpp:118-189 */
static void hlle(const double *wL, const double *wR, double g, double *flux) {
  side_t L, R;
  side_state(wL, g, &L);
  side_state(wR, g, &R);
  double v0, l0, l6;
  roe_speeds(wL, wR, g, L.H, R.H, &v0, &l0, &l6);
  double SL = stdmin(wL[1] - L.cf, l0);
  SL = MINm(SL, 0.);
  double SR = stdmax(wR[1] + R.cf, l6);
  SR = MAXm(SR, 0.);
  for (int k = 0; k < 8; k++) flux[k] = (SR * L.F[k] - SL * R.F[k] + SL * SR * (R.q[k] - L.q[k])) / (SR - SL);
}

/* riemannS.cpp:192-295 (Linde's alpha heuristic) */
static int hllc(const double *wL, const double *wR, double g, double *flux) {
  side_t L, R;
  side_state(wL, g, &L);
  side_state(wR, g, &R);
  double v0, l0, l6;
  roe_speeds(wL, wR, g, L.H, R.H, &v0, &l0, &l6);
  double SL = stdmin(wL[1] - L.cf, l0);
  double SR = stdmax(wR[1] + R.cf, l6);
  double SM = v0;
  double s1 = 0., s2 = 0.;
  for (int k = 0; k < 8; k++) s1 += fabs(R.q[k] - L.q[k]);
  for (int k = 0; k < 8; k++) s2 += fabs((R.F[k] - L.F[k]) - v0 * (R.q[k] - L.q[k]));
  double interm = 1. - s2 / (s1 + EPS);
  double alpha = MAXm(0., interm);
  if (alpha > 1 || alpha < 0 || isnan(alpha)) return ORA_ERR_HLLC;
  if (SR - SL == 0) return ORA_ERR_HLLC;
  for (int k = 0; k < 8; k++) {
    double dq = R.q[k] - L.q[k];
    double FLs = ((SL * ((1 - alpha) * SR + alpha * v0)) / (SR - SL)) * dq - (SL / (SR - SL)) * R.F[k] +
                 ((SR) / (SR - SL)) * L.F[k];
    double FRs = ((SR * ((1 - alpha) * SL + alpha * v0)) / (SR - SL)) * dq - (SL / (SR - SL)) * R.F[k] +
                 ((SR) / (SR - SL)) * L.F[k];
    if (SL > 0.) flux[k] = L.F[k];
    else if (SL <= 0. && SM >= 0.) flux[k] = FLs;
    else if (SM <= 0. && SR >= 0.) flux[k] = FRs;
    else if (SR < 0.) flux[k] = R.F[k];
  }
  return 0;
}

/* one side of the HLLD star state, riemannS.cpp:467-481 (left) / 483-497 (right) */
static void hlld_star(const double *w, double S, double SM, double pts, double e, double pt, double *ws, double *es) {
  for (int k = 0; k < 8; k++) ws[k] = 0.;
  ws[0] = w[0] * (S - w[1]) / (S - SM);
  ws[1] = SM;
  ws[5] = w[5];
  double den = w[0] * (S - w[1]) * (S - SM) - (w[5] * w[5]);
  if (den != 0) {
    ws[2] = w[2] - w[5] * w[6] * ((SM - w[1]) / den);
    ws[3] = w[3] - w[5] * w[7] * ((SM - w[1]) / den);
    double num = w[0] * ((S - w[1]) * (S - w[1])) - (w[5] * w[5]);
    ws[6] = w[6] * (num / den);
    ws[7] = w[7] * (num / den);
  } else {
    ws[2] = w[2]; ws[3] = w[3];
    ws[6] = 0.; ws[7] = 0.;
  }
  *es = (((S - w[1]) * e - pt * w[1] + pts * SM + w[5] * (dot_vb(w) - dot_vb(ws))) / (S - SM));
}

/* riemannS.cpp:397-549 (Miyoshi & Kusano 2005) */
static void hlld(const double *wL, const double *wR, double g, double *flux) {
  side_t L, R;
  side_state(wL, g, &L);
  side_state(wR, g, &R);
  double vnL = wL[1], vnR = wR[1];
  double SL = stdmin(vnL, vnR) - stdmax(L.cf, R.cf);
  double SR = stdmax(vnL, vnR) + stdmax(L.cf, R.cf);
  double SM = ((SR - vnR) * wR[0] * vnR - (SL - vnL) * wL[0] * vnL - R.pt + L.pt) /
              ((SR - vnR) * wR[0] - (SL - vnL) * wL[0]);
  double pts = ((SR - vnR) * wR[0] * L.pt - (SL - vnL) * wL[0] * R.pt +
                wL[0] * wR[0] * (SR - vnR) * (SL - vnL) * (vnR - vnL)) /
               ((SR - vnR) * wR[0] - (SL - vnL) * wL[0]);
  double wLs[8], wRs[8], wLss[8], wRss[8], eLs, eRs;
  hlld_star(wL, SL, SM, pts, L.e, L.pt, wLs, &eLs);
  hlld_star(wR, SR, SM, pts, R.e, R.pt, wRs, &eRs);
  double sLs = sqrt(wLs[0]), sRs = sqrt(wRs[0]);
  double sgL = sgn(wL[5]);
  for (int k = 0; k < 8; k++) wLss[k] = 0.;
  wLss[0] = wLs[0];
  wLss[1] = SM;
  wLss[2] = (sLs * wLs[2] + sRs * wRs[2] + (wRs[6] - wLs[6]) * sgL) / (sLs + sRs);
  wLss[3] = (sLs * wLs[3] + sRs * wRs[3] + (wRs[7] - wLs[7]) * sgL) / (sLs + sRs);
  wLss[5] = wLs[5];
  wLss[6] = (sLs * wRs[6] + sRs * wLs[6] + sqrt(wLs[0] * wRs[0]) * (wRs[2] - wLs[2]) * sgL) / (sLs + sRs);
  wLss[7] = (sLs * wRs[7] + sRs * wLs[7] + sqrt(wLs[0] * wRs[0]) * (wRs[3] - wLs[3]) * sgL) / (sLs + sRs);
  double eLss = eLs - sLs * (dot_vb(wLs) - dot_vb(wLss)) * sgL;
  for (int k = 0; k < 8; k++) wRss[k] = wLss[k];
  double eRss = eRs + sRs * (dot_vb(wRs) - dot_vb(wRss)) * sgn(wR[5]);
  double qLs[8], qLss[8], qRs[8], qRss[8];
  prim2cons(wLs, qLs, g);
  prim2cons(wLss, qLss, g);
  qLs[4] = eLs; qLss[4] = eLss;
  prim2cons(wRs, qRs, g);
  prim2cons(wRss, qRss, g);
  qRs[4] = eRs; qRss[4] = eRss;
  double SLs = SM - fabs(wL[5]) / sLs;
  double SRs = SM + fabs(wR[5]) / sRs;
  for (int k = 0; k < 8; k++) {
    if (SL > 0) flux[k] = L.F[k];
    else if (SL <= 0. && SLs >= 0.) flux[k] = L.F[k] + SL * qLs[k] - SL * L.q[k];
    else if (SLs <= 0. && SM >= 0.) flux[k] = L.F[k] + SLs * qLss[k] - (SLs - SL) * qLs[k] - SL * L.q[k];
    else if (SM <= 0. && SRs >= 0.) flux[k] = R.F[k] + SRs * qRss[k] - (SRs - SR) * qRs[k] - SR * R.q[k];
    else if (SRs <= 0. && SR >= 0.) flux[k] = R.F[k] + SR * qRs[k] - SR * R.q[k];
    else if (SR < 0.) flux[k] = R.F[k];
  }
}

/* riemannS.cpp:7-56: y faces swap (1,2) and (5,6) on the way in and out;
 * an output not assigned by the solver stays 0 (fresh vectors). */
static int riemann(int fid, char dir, double g, const double *wL, const double *wR, double *flux) {
  double l[8], r[8], f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int k = 0; k < 8; k++) { l[k] = wL[k]; r[k] = wR[k]; }
  if (dir == 'y') {
    l[1] = wL[2]; l[2] = wL[1]; r[1] = wR[2]; r[2] = wR[1];
    l[5] = wL[6]; l[6] = wL[5]; r[5] = wR[6]; r[6] = wR[5];
  }
  int st = 0;
  switch (fid) {
    case F_RUSA: rusanov(l, r, g, f); break;
    case F_HLLE: hlle(l, r, g, f); break;
    case F_HLLC: st = hllc(l, r, g, f); break;
    case F_HLLD: hlld(l, r, g, f); break;
    default: return ORA_ERR_CONFIG;
  }
  for (int k = 0; k < 8; k++) flux[k] = f[k];
  if (dir == 'y') { flux[1] = f[2]; flux[2] = f[1]; flux[5] = f[6]; flux[6] = f[5]; }
  return st;
}

int ora_riemann(const char *method, char dir, double gamma, const double *wL, const double *wR, double *flux) {
  return riemann(flux_id(method), dir, gamma, wL, wR, flux);
}

/* -------------------------------------------------------- cons <-> prim */
/* meshblock.cpp:119-145 (U2W) and 147-179 (Us2W, which also checks NaN
 * velocities on every cell).  Whole tile; p<=0 raises on the interior. */
static int cons2prim(ora_dom *d, int NT, int lo, int hi, double g, const double *U, double *W, int nb, int nan_check) {
  const size_t S = (size_t)NT * NT;
  for (int i = 0; i < NT; i++)
    for (int j = 0; j < NT; j++) {
      size_t c = (size_t)j * NT + i;
      double w0 = U[c];
      double w1 = U[S + c] / U[c], w2 = U[2 * S + c] / U[c], w3 = U[3 * S + c] / U[c];
      double w4 = (g - 1) * (U[4 * S + c] - 0.5 * w0 * (w1 * w1 + w2 * w2 + w3 * w3));
      double b5 = U[5 * S + c], b6 = U[6 * S + c], b7 = U[7 * S + c];
      w4 = w4 - (g - 1) * (0.5 * (b5 * b5 + b6 * b6 + b7 * b7));
      W[c] = w0; W[S + c] = w1; W[2 * S + c] = w2; W[3 * S + c] = w3; W[4 * S + c] = w4;
      W[5 * S + c] = b5; W[6 * S + c] = b6; W[7 * S + c] = b7;
      if (w4 <= 0 && i > lo && i < hi && j > lo && j < hi) return fail(d, ORA_ERR_PRESSURE, nb, i, j);
      if (nan_check && (isnan(w1) || isnan(w2) || isnan(w3))) return fail(d, ORA_ERR_NAN_VEL, nb, i, j);
    }
  return 0;
}

/* meshblock.cpp:181-199 */
static void prim2cons_tile(int NT, double g, const double *W, double *U) {
  const size_t S = (size_t)NT * NT;
  for (size_t c = 0; c < S; c++) {
    double w0 = W[c], w1 = W[S + c], w2 = W[2 * S + c], w3 = W[3 * S + c];
    U[c] = w0;
    U[S + c] = w1 * w0;
    U[2 * S + c] = w2 * w0;
    U[3 * S + c] = w3 * w0;
    double e = (W[4 * S + c] / (g - 1)) + 0.5 * w0 * (w1 * w1 + w2 * w2 + w3 * w3);
    double b5 = W[5 * S + c], b6 = W[6 * S + c], b7 = W[7 * S + c];
    U[5 * S + c] = b5; U[6 * S + c] = b6; U[7 * S + c] = b7;
    U[4 * S + c] = e + 0.5 * (b5 * b5 + b6 * b6 + b7 * b7);
  }
}

/* ------------------------------------------------------- reconstruction */
/* celledges.cpp:5-66: backward differences (index 0 duplicates index 1),
 * then limited edge states for i,j in [0,NT-2]; Bx (x) and By (y) skipped. */
static int celledges(ora_dom *d, int nb, int NT, int lim, double dxl, const double *W, double *scr) {
  const size_t S = (size_t)NT * NT * 8;
  double *dwdx = scr, *dwdy = scr + S, *wxL = scr + 2 * S, *wxR = scr + 3 * S, *wyL = scr + 4 * S, *wyR = scr + 5 * S;
  for (int i = 0; i < NT; i++)
    for (int j = 0; j < NT; j++)
      for (int k = 0; k < 8; k++) {
        int im = i > 0 ? i - 1 : i, jm = j > 0 ? j - 1 : j;
        AT(dwdx, NT, k, i, j) = AT(W, NT, k, im + 1, j) - AT(W, NT, k, im, j);
        AT(dwdy, NT, k, i, j) = AT(W, NT, k, i, jm + 1) - AT(W, NT, k, i, jm);
        if (isnan(AT(dwdx, NT, k, i, j)) || isnan(AT(dwdy, NT, k, i, j))) return fail(d, ORA_ERR_NAN_EDGE, nb, i, j);
      }
  for (int i = 0; i < NT - 1; i++)
    for (int j = 0; j < NT - 1; j++)
      for (int k = 0; k < 8; k++) {
        double w = AT(W, NT, k, i, j), eta, r, phi;
        if (k != 5) {
          double a = AT(dwdx, NT, k, i, j), b = AT(dwdx, NT, k, i + 1, j);
          eta = (lim == L_CADA) ? ((b * b) + (a * a)) / ((1. * dxl) * (1. * dxl)) : 0.;
          r = b / (a + EPS);
          phi = limiter(lim, r, eta);
          AT(wxL, NT, k, i, j) = w + 0.5 * phi * a;
          r = a / (b + EPS);
          phi = limiter(lim, r, eta);
          AT(wxR, NT, k, i, j) = w - 0.5 * phi * b;
          if (isnan(AT(wxL, NT, k, i, j))) return fail(d, ORA_ERR_NAN_EDGE, nb, i + 1, j);
        }
        if (k != 6) {
          double a = AT(dwdy, NT, k, i, j), b = AT(dwdy, NT, k, i, j + 1);
          eta = (lim == L_CADA) ? ((b * b) + (a * a)) / ((1. * dxl) * (1. * dxl)) : 0.;
          r = b / (a + EPS);
          phi = limiter(lim, r, eta);
          AT(wyL, NT, k, i, j) = w + 0.5 * phi * a;
          r = a / (b + EPS);
          phi = limiter(lim, r, eta);
          AT(wyR, NT, k, i, j) = w - 0.5 * phi * b;
        }
      }
  return 0;
}

/* WENO2D_primvar.cpp:95-198 (5th order; beta_1's first term uses w_{i-2}) */
static void weno5_pair(double m2, double m1, double c, double p1, double p2, double *right, double *left) {
  const double epsw = 1.E-15;
  double p0n = (-1 * m2 + 5 * m1 + 2 * c) / 6;
  double p1n = (2 * m1 + 5 * c - p1) / 6;
  double p2n = (11 * c - 7 * p1 + 2 * p2) / 6;
  double t;
  t = m2 - 2 * m1 + c; double u0 = m2 - 4 * m1 + 3 * c;
  double B0 = (13. / 12.) * (t * t) + (1. / 4.) * (u0 * u0);
  t = m2 - 2 * c + p1; double u1 = m1 - p1;
  double B1 = (13. / 12.) * (t * t) + (1. / 4.) * (u1 * u1);
  t = c - 2 * p1 + p2; double u2 = 3 * c - 4 * p1 + p2;
  double B2 = (13. / 12.) * (t * t) + (1. / 4.) * (u2 * u2);
  double e0 = epsw + B0, e1 = epsw + B1, e2 = epsw + B2;
  double a0 = (1. / 10.) / (e0 * e0), a1 = (6. / 10.) / (e1 * e1), a2 = (3. / 10.) / (e2 * e2);
  double as = a0 + a1 + a2;
  *right = a0 / as * p0n + a1 / as * p1n + a2 / as * p2n;
  double p0p = (2 * m2 - 7 * m1 + 11 * c) / 6;
  double p1p = (-1 * m1 + 5 * c + 2 * p1) / 6;
  double p2p = (2 * c + 5 * p1 - p2) / 6;
  double b0 = (3. / 10.) / (e0 * e0), b1 = (6. / 10.) / (e1 * e1), b2 = (1. / 10.) / (e2 * e2);
  double bs = b0 + b1 + b2;
  *left = b0 / bs * p0p + b1 / bs * p1p + b2 / bs * p2p;
}

/* WENO2D_primvar.cpp:7-91 (3rd order) */
static void weno3_pair(double m1, double c, double p1, double *right, double *left) {
  const double epsw = 1.E-15;
  double p0n = 0.5 * m1 + 0.5 * c;
  double p1n = 1.5 * c - 0.5 * p1;
  double B0 = (c - m1) * (c - m1);
  double B1 = (p1 - c) * (p1 - c);
  double e0 = epsw + B0, e1 = epsw + B1;
  double a0 = (2. / 3.) / (e0 * e0), a1 = (1. / 3.) / (e1 * e1);
  double as = a0 + a1;
  *right = a0 / as * p0n + a1 / as * p1n;
  double p0p = 1.5 * c - 0.5 * m1;
  double p1p = 0.5 * p1 + 0.5 * c;
  double b0 = (1. / 3.) / (e0 * e0), b1 = (2. / 3.) / (e1 * e1);
  double bs = b0 + b1;
  *left = b0 / bs * p0p + b1 / bs * p1p;
}

/* WENO on primitives over i,j in [nxminb, nxmaxb] for both directions */
static void weno_tile(int NT, int ng, int lim, const double *W, double *scr) {
  const size_t S = (size_t)NT * NT * 8;
  double *wxL = scr + 2 * S, *wxR = scr + 3 * S, *wyL = scr + 4 * S, *wyR = scr + 5 * S;
  int lo = ng - 1, hi = NT - ng; /* nxminb, nxmaxb */
  for (int dir = 0; dir < 2; dir++)
    for (int i = lo; i <= hi; i++)
      for (int j = lo; j <= hi; j++)
        for (int k = 0; k < 8; k++) {
          if (dir == 0 && k == 5) continue;
          if (dir == 1 && k == 6) continue;
          int di = dir == 0 ? 1 : 0, dj = dir == 0 ? 0 : 1;
          double c = AT(W, NT, k, i, j), p1 = AT(W, NT, k, i + di, j + dj), m1 = AT(W, NT, k, i - di, j - dj);
          double rgt, lft;
          if (lim == L_WENO5) {
            double p2 = AT(W, NT, k, i + 2 * di, j + 2 * dj), m2 = AT(W, NT, k, i - 2 * di, j - 2 * dj);
            weno5_pair(m2, m1, c, p1, p2, &rgt, &lft);
          } else {
            weno3_pair(m1, c, p1, &rgt, &lft);
          }
          if (dir == 0) { AT(wxR, NT, k, i, j) = rgt; AT(wxL, NT, k, i, j) = lft; }
          else { AT(wyR, NT, k, i, j) = rgt; AT(wyL, NT, k, i, j) = lft; }
        }
}

/* MUSCL2D.cpp:13-121 on one tile: primitives, edges, CT face override of the
 * normal field, Riemann fluxes ff[i+1][j], gg[i][j+1] for i,j in [nxminb,nxmax]. */
static int muscl_tile(ora_dom *d, int nb, int NT, int ng, int lim, int fid, double g, double dxl, int step,
                      const double *U, const double *B, double *W, double *ff, double *gg, double *scr) {
  int st = cons2prim(d, NT, ng - 1, NT - ng, g, U, W, nb, step == 2);
  if (st) return st;
  if (lim == L_WENO3 || lim == L_WENO5) weno_tile(NT, ng, lim, W, scr);
  else if ((st = celledges(d, nb, NT, lim, dxl, W, scr))) return st;
  const size_t S = (size_t)NT * NT * 8;
  double *wxL = scr + 2 * S, *wxR = scr + 3 * S, *wyL = scr + 4 * S, *wyR = scr + 5 * S;
  for (int i = 0; i < NT - 1; i++)
    for (int j = 0; j < NT - 1; j++) {
      AT(wxL, NT, 5, i, j) = AT(B, NT, 0, i + 1, j);
      AT(wxR, NT, 5, i, j) = AT(B, NT, 0, i, j);
      AT(wyL, NT, 6, i, j) = AT(B, NT, 1, i, j + 1);
      AT(wyR, NT, 6, i, j) = AT(B, NT, 1, i, j);
    }
  for (int i = ng - 1; i <= NT - 1 - ng; i++)
    for (int j = ng - 1; j <= NT - 1 - ng; j++) {
      double wl[8], wr[8], fl[8];
      for (int k = 0; k < 8; k++) { wl[k] = AT(wxL, NT, k, i, j); wr[k] = AT(wxR, NT, k, i + 1, j); }
      if ((st = riemann(fid, 'x', g, wl, wr, fl))) return fail(d, st, nb, i, j);
      for (int k = 0; k < 8; k++) {
        AT(ff, NT, k, i + 1, j) = fl[k];
        if (isnan(fl[k])) return fail(d, ORA_ERR_NAN_FLUX, nb, i, j);
      }
      if (wl[5] - wr[5] != 0) return fail(d, ORA_ERR_CT_FACE, nb, i, j);
      for (int k = 0; k < 8; k++) { wl[k] = AT(wyL, NT, k, i, j); wr[k] = AT(wyR, NT, k, i, j + 1); }
      if ((st = riemann(fid, 'y', g, wl, wr, fl))) return fail(d, st, nb, i, j);
      for (int k = 0; k < 8; k++) {
        AT(gg, NT, k, i, j + 1) = fl[k];
        if (isnan(fl[k])) return fail(d, ORA_ERR_NAN_FLUX, nb, i, j);
      }
      if (wl[6] - wr[6] != 0) return fail(d, ORA_ERR_CT_FACE, nb, i, j);
    }
  return 0;
}

/* ------------------------------------------------------------- domain */
ora_dom *ora_create(const ora_config *c) {
  if (c->nghosts < 2 || c->nghosts % 2 || c->maxlevs < 1 || c->maxlevs >= MAXLEV || c->nx < 4 || c->nx % 2) return NULL;
  ora_dom *d = (ora_dom *)calloc(1, sizeof *d);
  d->ng = c->nghosts;
  d->nxi = c->nx;
  d->NT = c->nx + 2 * c->nghosts;
  d->maxlevs = c->maxlevs;
  /* defs.hpp:30 */
  double p = pow(NBROOTS, c->maxlevs - 1);
  d->maxblocks = NBROOTS > p ? NBROOTS : (int)(pow(4, c->maxlevs) * 2);
  d->nxmin = d->ng; d->nxmax = d->NT - 1 - d->ng;
  d->nxminb = d->nxmin - 1; d->nxmaxb = d->nxmax + 1;
  d->nx2 = (d->NT - 1) / 2;
  d->cfl = c->cfl;
  d->lim = lim_id(c->limiter);
  d->flux = flux_id(c->flux);
  d->ic = ic_id(c->ic);
  memcpy(d->icp, c->ic_params, sizeof d->icp);
  d->periodic = c->periodic != 0;
  d->rk = c->rk_stages == 3 ? 3 : 2;
  if (c->ct_type < 0 || c->ct_type > 2) { free(d); return NULL; }
  d->ct = c->ct_type == 2 ? 2 : 1;
  if (c->rk_stages != 0 && c->rk_stages != 2 && c->rk_stages != 3) { free(d); return NULL; }
  if (d->lim == L_BAD || d->flux == F_BAD || d->ic == IC_BAD) { free(d); return NULL; }
  /* meshblock.cpp:90-96; a periodic domain needs its exact length, so it
     uses lenx / (sqrt(nbroots) * nx) instead (build-defined) */
  const int cells = d->periodic ? d->NT - 2 * d->ng : d->NT - 2 * d->ng - 1;
  d->dx[0] = c->lenx / (sqrt(NBROOTS) * cells);
  d->dy[0] = c->lenx / (sqrt(NBROOTS) * cells);
  for (int l = 1; l < d->maxlevs; l++) { d->dx[l] = d->dx[l - 1] / 2.; d->dy[l] = d->dy[l - 1] / 2.; }
  size_t mb = (size_t)d->maxblocks;
  d->lp = calloc(mb, sizeof *d->lp);
  d->icoord = calloc(mb, sizeof *d->icoord);
  d->leafs = calloc(mb, 1); d->iref = calloc(mb, 1); d->icoarse = calloc(mb, 1);
  for (int f = 0; f < ORA_NFIELD; f++) d->f[f] = calloc(mb * NVAR[f] * d->NT * d->NT, sizeof(double));
  d->ib = calloc(4 * mb, sizeof *d->ib);
  d->scr = calloc((size_t)6 * 8 * d->NT * d->NT, sizeof(double));
  for (int f = 0; f < ORA_NFIELD; f++)
    if (!d->f[f]) { ora_destroy(d); return NULL; }
  return d;
}

void ora_destroy(ora_dom *d) {
  if (!d) return;
  for (int f = 0; f < ORA_NFIELD; f++) free(d->f[f]);
  free(d->lp); free(d->icoord); free(d->leafs); free(d->iref); free(d->icoarse); free(d->ib); free(d->scr);
  free(d);
}

int ora_nvar(int field) { return (field >= 0 && field < ORA_NFIELD) ? NVAR[field] : 0; }
int ora_tile(const ora_dom *d) { return d->NT; }

/* IC2Dtype.cpp:6-190 for one block; sets gamma/tEnd (setParam, IC2Dtype.cpp:171) */
static void initial_condition(ora_dom *d, int nb) {
  int NT = d->NT, ng = d->ng;
  double r4[4] = {0}, p4[4] = {0}, u4[4] = {0}, v4[4] = {0}, w4[4] = {0}, bx4[4] = {0}, by4[4] = {0}, bz4[4] = {0};
  double tEnd = 0.2, gamma = 2.0;
#define SET4(a, x0, x1, x2, x3) do { a[0] = x0; a[1] = x1; a[2] = x2; a[3] = x3; } while (0)
  switch (d->ic) {
    case IC_SST: SET4(p4, 0.1, 1.0, 1.0, 0.1); SET4(r4, 0.125, 1.0, 1.0, 0.125); break;
    case IC_SST1: SET4(p4, 1.0, 0.1, 0.1, 1.0); SET4(r4, 1.0, 0.125, 0.125, 1.0); break;
    case IC_SSTY: SET4(p4, 0.1, 0.1, 1.0, 1.0); SET4(r4, 0.125, 0.125, 1.0, 1.0); break;
    case IC_BWX: SET4(p4, 0.1, 1.0, 1.0, 0.1); SET4(r4, 0.125, 1.0, 1.0, 0.125);
      SET4(bx4, 0.75, 0.75, 0.75, 0.75); SET4(by4, -1.0, 1.0, 1.0, -1.0); break;
    case IC_BWX1: SET4(p4, 1.0, 0.1, 0.1, 1.0); SET4(r4, 1.0, 0.125, 0.125, 1.0);
      SET4(bx4, 0.75, 0.75, 0.75, 0.75); SET4(by4, 1.0, -1.0, -1.0, 1.0); break;
    case IC_BWY: SET4(p4, 0.1, 0.1, 1.0, 1.0); SET4(r4, 0.125, 0.125, 1.0, 1.0);
      SET4(by4, 0.75, 0.75, 0.75, 0.75); SET4(bx4, -1.0, -1.0, 1.0, 1.0); break;
    case IC_BWY1: SET4(p4, 1.0, 1.0, 0.1, 0.1); SET4(r4, 1.0, 1.0, 0.125, 0.125);
      SET4(by4, 0.75, 0.75, 0.75, 0.75); SET4(bx4, 1.0, 1.0, -1.0, -1.0); break;
    case IC_ROTOR: gamma = 1.4; tEnd = 0.295; p4[0] = 0.5; bx4[0] = 2.5 / (sqrt(4 * PI_REF)); break;
    /* the Balsara-Spicer rotor (p = 1, v0 = 2, Bx = 5/sqrt(4 pi)) in the form of
       IC2Dtype.cpp:75-111 with only those constants changed (build-defined) */
    case IC_ROTOR_BS: gamma = 1.4; tEnd = 0.295; p4[0] = 1.0; bx4[0] = 5.0 / (sqrt(4 * PI_REF)); break;
    case IC_ROTOR_P: gamma = d->icp[6]; tEnd = d->icp[7]; break;
    case IC_OT: gamma = 5. / 3.; tEnd = 0.5; break;
  }
#undef SET4
  double *W = TILE(d, ORA_W, nb);
  double dxl = d->dx[d->lp[nb][0]], dyl = d->dy[d->lp[nb][0]];
  /* cell centres (IC2Dtype.cpp:92-94); the periodic geometry starts at 0 */
  const double x0 = d->periodic ? 0. : d->dx[0], y0 = d->periodic ? 0. : d->dy[0];
  for (int i = 0; i < NT; i++) {
    double x = ((i + d->icoord[nb][0] - ng) + 0.5) * dxl - x0;
    for (int j = 0; j < NT; j++) {
      double y = ((j + d->icoord[nb][1] - ng) + 0.5) * dyl - y0;
      if (d->ic == IC_OT) {
        /* build-defined Orszag-Tang vortex on [0,1]^2 */
        const double pi = 3.141592653589793, B0 = 1. / sqrt(4. * pi);
        for (int k = 0; k < 8; k++) AT(W, NT, k, i, j) = 0.;
        AT(W, NT, 0, i, j) = 25. / (36. * pi);
        AT(W, NT, 1, i, j) = -sin(2. * pi * y);
        AT(W, NT, 2, i, j) = sin(2. * pi * x);
        AT(W, NT, 4, i, j) = 5. / (12. * pi);
        AT(W, NT, 5, i, j) = -B0 * sin(2. * pi * y);
        AT(W, NT, 6, i, j) = B0 * sin(4. * pi * x);
      } else if (d->ic == IC_ROTOR || d->ic == IC_ROTOR_BS) {
        /* spin speed v0 at r = 0.1: 1 (reference), 2 (Balsara-Spicer); scaling
           by 1 or 2 is exact, so the reference rotor's values are unchanged */
        const double v0 = d->ic == IC_ROTOR_BS ? 2. : 1.;
        double radius = sqrt(((x - 0.5) * (x - 0.5)) + ((y - 0.5) * (y - 0.5)));
        double fr = (23. - 200. * radius) / 3.;
        for (int k = 0; k < 8; k++) AT(W, NT, k, i, j) = 0.;
        AT(W, NT, 4, i, j) = p4[0];
        AT(W, NT, 5, i, j) = bx4[0];
        if (radius <= 0.1) {
          AT(W, NT, 0, i, j) = 10.;
          AT(W, NT, 1, i, j) = -v0 * (y - 0.5) / 0.1;
          AT(W, NT, 2, i, j) = v0 * (x - 0.5) / 0.1;
        } else if (radius > 0.1 && radius < 0.115) {
          AT(W, NT, 0, i, j) = (1 + 9. * fr);
          AT(W, NT, 1, i, j) = (-v0 * (y - 0.5) * fr / 0.1);
          AT(W, NT, 2, i, j) = v0 * (x - 0.5) * fr / 0.1;
        } else {
          AT(W, NT, 0, i, j) = 1.;
        }
      } else if (d->ic == IC_ROTOR_P) {
        /* build-defined parameterised rotor: (rho_in, v0, p, Bx, r0, r1) with a
           linear taper between r0 and r1 (the reference rotor's shape) */
        const double rho_in = d->icp[0], v0 = d->icp[1], r0 = d->icp[4], r1 = d->icp[5];
        double radius = sqrt(((x - 0.5) * (x - 0.5)) + ((y - 0.5) * (y - 0.5)));
        double fr = (r1 - radius) / (r1 - r0);
        for (int k = 0; k < 8; k++) AT(W, NT, k, i, j) = 0.;
        AT(W, NT, 4, i, j) = d->icp[2];
        AT(W, NT, 5, i, j) = d->icp[3];
        if (radius <= r0) {
          AT(W, NT, 0, i, j) = rho_in;
          AT(W, NT, 1, i, j) = -v0 * (y - 0.5) / r0;
          AT(W, NT, 2, i, j) = v0 * (x - 0.5) / r0;
        } else if (radius > r0 && radius < r1) {
          AT(W, NT, 0, i, j) = 1. + (rho_in - 1.) * fr;
          AT(W, NT, 1, i, j) = -v0 * (y - 0.5) * fr / r0;
          AT(W, NT, 2, i, j) = v0 * (x - 0.5) * fr / r0;
        } else {
          AT(W, NT, 0, i, j) = 1.;
        }
      } else {
        int q;
        if (x >= 0.5 && y >= 0.5) q = 0;
        else if (x < 0.5 && y >= 0.5) q = 1;
        else if (x < 0.5 && y < 0.5) q = 2;
        else q = 3;
        AT(W, NT, 0, i, j) = r4[q]; AT(W, NT, 1, i, j) = u4[q]; AT(W, NT, 2, i, j) = v4[q];
        AT(W, NT, 3, i, j) = w4[q]; AT(W, NT, 4, i, j) = p4[q];
        AT(W, NT, 5, i, j) = bx4[q]; AT(W, NT, 6, i, j) = by4[q]; AT(W, NT, 7, i, j) = bz4[q];
      }
    }
  }
  d->gamma = gamma;
  d->tEnd = tEnd;
  /* IC2Dtype.cpp:174-189: faces from cell averages; index 0 copies index 1 */
  double *B = TILE(d, ORA_BI, nb);
  for (int i = 0; i < NT; i++)
    for (int j = 0; j < NT; j++) {
      if (i < NT - 1) AT(B, NT, 0, i + 1, j) = 0.5 * (AT(W, NT, 5, i, j) + AT(W, NT, 5, i + 1, j));
      else AT(B, NT, 0, 0, j) = AT(B, NT, 0, 1, j);
      if (j < NT - 1) AT(B, NT, 1, i, j + 1) = 0.5 * (AT(W, NT, 6, i, j) + AT(W, NT, 6, i, j + 1));
      else AT(B, NT, 1, i, 0) = AT(B, NT, 1, i, 1);
    }
}

/* basegrid.cpp:4-89: 4 roots, IDs 0|1 bottom, 2|3 top; a periodic domain
   wraps the outer neighbour pointers instead of -1 (build-defined) */
static void basegrid(ora_dom *d) {
  static const int32_t lp0[4][8] = {{0, 0, 0, 0, -1, 1, -1, 2}, {0, 0, 0, 0, 0, -1, -1, 3},
                                    {0, 0, 0, 0, -1, 3, 0, -1}, {0, 0, 0, 0, 2, -1, 1, -1}};
  static const int32_t lpp[4][8] = {{0, 0, 0, 0, 1, 1, 2, 2}, {0, 0, 0, 0, 0, 0, 3, 3},
                                    {0, 0, 0, 0, 3, 3, 0, 0}, {0, 0, 0, 0, 2, 2, 1, 1}};
  int side = d->NT - 2 * d->ng;
  for (int r = 0; r < 4; r++) {
    memcpy(d->lp[r], d->periodic ? lpp[r] : lp0[r], sizeof lp0[r]);
    d->icoord[r][0] = (r & 1) ? side : 0;
    d->icoord[r][1] = (r & 2) ? side : 0;
    d->leafs[r] = 1;
    d->iref[r] = 1;
  }
  d->lastActive = NBROOTS;
  for (int nb = 0; nb < d->lastActive; nb++) {
    initial_condition(d, nb);
    prim2cons_tile(d->NT, d->gamma, TILE(d, ORA_W, nb), TILE(d, ORA_U, nb));
  }
}

/* --------------------------------------------------------- fill schedule */
/* locate_bounds.cpp:35-208 */
int ora_locate_bounds(ora_dom *d) {
  unsigned char(*done)[4] = calloc((size_t)d->lastActive, 4);
  int n = 0;
  for (int nb = 0; nb < d->lastActive; nb++) {
    if (!d->leafs[nb]) continue;
    /* side s: 0 left 1 right 2 bottom 3 top; opposite side; codes */
    static const int opp[4] = {1, 0, 3, 2};
    static const int box[4] = {-1, -2, -3, -4}, same[4] = {1, 4, 7, 10}, finer[4] = {2, 5, 8, 11};
    static const int soff[4] = {1, 0, 2, 0}, step2[4] = {2, 2, 1, 1};
    for (int s = 0; s < 4; s++) {
      if (done[nb][s]) continue;
      int nbr = d->lp[nb][4 + s];
      if (nbr == -1) {
        done[nb][s] = 1;
        d->ib[n][0] = nb; d->ib[n][1] = -1; d->ib[n][2] = box[s]; n++;
      } else if (d->lp[nbr][0] >= d->lp[nb][0]) {
        if (d->leafs[nbr]) {
          d->ib[n][0] = nb; d->ib[n][1] = nbr; d->ib[n][2] = same[s]; n++;
          done[nb][s] = 1;
          done[nbr][opp[s]] = 1;
        } else {
          done[nb][s] = 1;
          int n1 = d->lp[nbr][2] + soff[s], n2 = n1 + step2[s];
          d->ib[n][0] = nb; d->ib[n][1] = n1; d->ib[n][2] = finer[s]; n++;
          done[n1][opp[s]] = 1;
          done[n2][opp[s]] = 1;
        }
      }
    }
  }
  d->nbounds = n;
  free(done);
  return 0;
}

int ora_nbounds(const ora_dom *d, int32_t *ib3) {
  if (ib3) memcpy(ib3, d->ib, (size_t)d->nbounds * sizeof *d->ib);
  return d->nbounds;
}

/* fine->coarse value used by the three fills: boundary.cpp (0.25 * 2x2 sum,
 * misc.cpp:35-48 order), boundaryfc.cpp (0.5 * face pair: along j for k==0,
 * along i otherwise) and boundaryec.cpp (injection). */
static inline double fine_val(const double *Q, int NT, int kind, int k, int i1, int j1) {
  double v = 0.;
  switch (kind) {
    case ORA_FILL_CC:
      v += AT(Q, NT, k, i1, j1); v += AT(Q, NT, k, i1, j1 + 1);
      v += AT(Q, NT, k, i1 + 1, j1); v += AT(Q, NT, k, i1 + 1, j1 + 1);
      return 0.25 * v;
    case ORA_FILL_FC:
      if (k == 0) { v += AT(Q, NT, k, i1, j1); v += AT(Q, NT, k, i1, j1 + 1); }
      else { v += AT(Q, NT, k, i1, j1); v += AT(Q, NT, k, i1 + 1, j1); }
      return 0.5 * v;
    default:
      return AT(Q, NT, k, i1, j1);
  }
}

/* boundary.cpp:7-388, boundaryfc.cpp:8-404, boundaryec.cpp:7-378 (the three
 * share their structure; only fine_val differs) + extboundary.cpp:5-61. */
int ora_fill(ora_dom *d, int field, int kind) {
  const int NT = d->NT, ng = d->ng, nv = NVAR[field];
  const int nxmin = d->nxmin, nxmax = d->nxmax, nxminb = d->nxminb, nxmaxb = d->nxmaxb, nx2 = d->nx2;
  const int nymin = nxmin, nymax = nxmax, nyminb = nxminb, nymaxb = nxmaxb, ny2 = nx2;
  const int nxp1 = nxmaxb, nyp1 = nymaxb;
#define BLK(nb) TILE(d, field, nb)
#define Q(nb, k, i, j) AT(BLK(nb), NT, k, i, j)
#define FV(nb, k, i1, j1) fine_val(BLK(nb), NT, kind, k, i1, j1)
  for (int e = 0; e < d->nbounds; e++) {
    int nbs = d->ib[e][0], n1 = d->ib[e][1], n2, n3, n4, ii, jj;
    switch (d->ib[e][2]) {
      case 1: /* left, same level */
        for (int j = 0; j < NT; j++)
          for (int k = 0; k < nv; k++) {
            ii = nxmax - (ng - 1);
            for (int i = 0; i < ng; i++, ii++) Q(nbs, k, i, j) = Q(n1, k, ii, j);
            ii = nxmin;
            for (int i = nxp1; i < NT; i++, ii++) Q(n1, k, i, j) = Q(nbs, k, ii, j);
          }
        break;
      case 2: /* left, finer neighbour */
        n2 = n1 + 2;
        for (int j = nymin; j <= ny2; j++) {
          int ny2j = ny2 + j - 1 - (ng - 2);
          for (int i = 0; i < ng; i++) {
            int i1 = 2 * (i - ng / 2) - ng / 2 + nxmax, j1 = 2 * j - nymin;
            for (int k = 0; k < nv; k++) {
              Q(nbs, k, i, j) = FV(n1, k, i1, j1);
              Q(nbs, k, i, ny2j) = FV(n2, k, i1, j1);
            }
          }
        }
        n3 = d->lp[d->lp[n1][1]][6];
        if (n3 != -1) {
          if (d->leafs[n3]) {
            for (int k = 0; k < nv; k++) for (int i = 0; i < ng; i++) for (int j = 0; j < ng; j++) {
              ii = nxmax - ng + (i + 1); jj = nymax - ng + (j + 1);
              Q(nbs, k, i, j) = Q(n3, k, ii, jj);
            }
          } else {
            n3 = d->lp[n3][2] + 3;
            if (d->leafs[n3])
              for (int k = 0; k < nv; k++) for (int i = 0; i < ng; i++) for (int j = 0; j < ng; j++) {
                ii = nxmax - ng * 2 + 2 * i + 1; jj = nymax - ng * 2 + 2 * j + 1;
                Q(nbs, k, i, j) = FV(n3, k, ii, jj);
              }
          }
        }
        n4 = d->lp[d->lp[n2][1]][7];
        if (n4 != -1) {
          if (d->leafs[n4]) {
            for (int k = 0; k < nv; k++) for (int i = 0; i < ng; i++) for (int j = 0; j < ng; j++) {
              ii = nxmax - ng + (i + 1); jj = nymin + ng - (j + 1);
              Q(nbs, k, i, NT - 1 - j) = Q(n4, k, ii, jj);
            }
          } else {
            n4 = d->lp[n4][2] + 1;
            if (d->leafs[n4])
              for (int k = 0; k < nv; k++) for (int i = 0; i < ng; i++) for (int j = 0; j < ng; j++) {
                ii = nxmax - ng * 2 + 2 * i + 1; jj = nymin + ng * 2 - 2 * (j + 1);
                Q(nbs, k, i, NT - 1 - j) = FV(n4, k, ii, jj);
              }
          }
        }
        for (int i = 1; i <= ng; i++)
          for (int j = 0; j < NT; j++)
            for (int k = 0; k < nv; k++) {
              Q(n1, k, i + nxmax, j) = Q(nbs, k, (i + 2 * ng - 1) / 2, (j + ng) / 2);
              Q(n2, k, i + nxmax, j) = Q(nbs, k, (i + 2 * ng - 1) / 2, (j - ng + 2) / 2 + ny2);
            }
        break;
      case 4: /* right, same level */
        for (int j = 0; j < NT; j++)
          for (int k = 0; k < nv; k++) {
            ii = nxmin;
            for (int i = nxp1; i < NT; i++, ii++) Q(nbs, k, i, j) = Q(n1, k, ii, j);
            ii = nxmax - (ng - 1);
            for (int i = 0; i < ng; i++, ii++) Q(n1, k, i, j) = Q(nbs, k, ii, j);
          }
        break;
      case 5: /* right, finer neighbour */
        n2 = n1 + 2;
        for (int i = nxp1; i < NT; i++)
          for (int j = nymin; j <= ny2; j++) {
            int i1 = 2 * (i - nxmax + (ng / 2 - 1)), j1 = 2 * j - nymin, ny2j = ny2 + j - 1 - (ng - 2);
            for (int k = 0; k < nv; k++) {
              Q(nbs, k, i, j) = FV(n1, k, i1, j1);
              Q(nbs, k, i, ny2j) = FV(n2, k, i1, j1);
            }
          }
        n3 = d->lp[d->lp[n1][1]][6];
        if (n3 != -1) {
          if (d->leafs[n3]) {
            for (int k = 0; k < nv; k++) for (int i = nxmaxb; i < NT; i++) for (int j = 0; j < ng; j++) {
              ii = nxminb + (i - nxmax); jj = nymax - ng + (j + 1);
              Q(nbs, k, i, j) = Q(n3, k, ii, jj);
            }
          } else {
            n3 = d->lp[n3][2] + 2;
            if (d->leafs[n3])
              for (int k = 0; k < nv; k++) for (int i = nxmaxb; i < NT; i++) for (int j = 0; j < ng; j++) {
                ii = nxminb + 2 * (i - nxmax) - 1; jj = nymax - ng * 2 + 2 * j + 1;
                Q(nbs, k, i, j) = FV(n3, k, ii, jj);
              }
          }
        }
        n4 = d->lp[d->lp[n2][1]][7];
        if (n4 != -1) {
          if (d->leafs[n4]) {
            for (int k = 0; k < nv; k++) for (int i = nxmaxb; i < NT; i++) for (int j = 0; j < ng; j++) {
              ii = nxminb + (i - nxmax); jj = nymin + ng - (j + 1);
              Q(nbs, k, i, NT - 1 - j) = Q(n4, k, ii, jj);
            }
          } else {
            n4 = d->lp[n4][2];
            if (d->leafs[n4])
              for (int k = 0; k < nv; k++) for (int i = nxmaxb; i < NT; i++) for (int j = 0; j < ng; j++) {
                ii = nxminb + 2 * (i - nxmax) - 1; jj = nymin + ng * 2 - 2 * (j + 1);
                Q(nbs, k, i, NT - 1 - j) = FV(n4, k, ii, jj);
              }
          }
        }
        for (int i = 1; i <= ng; i++)
          for (int j = 0; j < NT; j++)
            for (int k = 0; k < nv; k++) {
              Q(n1, k, i - 1, j) = Q(nbs, k, (i + 1) / 2 + nxmax - ng / 2, (j + ng) / 2);
              Q(n2, k, i - 1, j) = Q(nbs, k, (i + 1) / 2 + nxmax - ng / 2, (j - ng + 2) / 2 + ny2);
            }
        break;
      case 7: /* bottom, same level */
        for (int i = 0; i < NT; i++)
          for (int k = 0; k < nv; k++) {
            jj = nymax - (ng - 1);
            for (int j = 0; j < ng; j++, jj++) Q(nbs, k, i, j) = Q(n1, k, i, jj);
            jj = nymin;
            for (int j = nyp1; j < NT; j++, jj++) Q(n1, k, i, j) = Q(nbs, k, i, jj);
          }
        break;
      case 8: /* bottom, finer neighbour */
        n2 = n1 + 1;
        for (int i = nxmin; i <= nx2; i++) {
          int nx2i = nx2 + i - 1 - (ng - 2);
          for (int j = 0; j < ng; j++) {
            int i1 = 2 * i - nxmin, j1 = 2 * (j - ng / 2) - ng / 2 + nymax;
            for (int k = 0; k < nv; k++) {
              Q(nbs, k, i, j) = FV(n1, k, i1, j1);
              Q(nbs, k, nx2i, j) = FV(n2, k, i1, j1);
            }
          }
        }
        n3 = d->lp[d->lp[n1][1]][4];
        if (n3 != -1) {
          if (d->leafs[n3]) {
            for (int k = 0; k < nv; k++) for (int i = 0; i < ng; i++) for (int j = 0; j < ng; j++) {
              ii = nxmax - ng + (i + 1); jj = nymax - ng + (j + 1);
              Q(nbs, k, i, j) = Q(n3, k, ii, jj);
            }
          } else {
            n3 = d->lp[n3][2] + 3;
            if (d->leafs[n3])
              for (int k = 0; k < nv; k++) for (int i = 0; i < ng; i++) for (int j = 0; j < ng; j++) {
                ii = nxmax - ng * 2 + 2 * i + 1; jj = nymax - ng * 2 + 2 * j + 1;
                Q(nbs, k, i, j) = FV(n3, k, ii, jj);
              }
          }
        }
        n4 = d->lp[d->lp[n2][1]][5];
        if (n4 != -1) {
          if (d->leafs[n4]) {
            for (int k = 0; k < nv; k++) for (int i = 0; i < ng; i++) for (int j = 0; j < ng; j++) {
              ii = nxmin + ng - (i + 1); jj = nymax - ng + (j + 1);
              Q(nbs, k, NT - 1 - i, j) = Q(n4, k, ii, jj);
            }
          } else {
            n4 = d->lp[n4][2] + 2;
            if (d->leafs[n4])
              for (int k = 0; k < nv; k++) for (int i = 0; i < ng; i++) for (int j = 0; j < ng; j++) {
                ii = nxmin + ng * 2 - 2 * (i + 1); jj = nymax - ng * 2 + 2 * j + 1;
                Q(nbs, k, NT - 1 - i, j) = FV(n4, k, ii, jj);
              }
          }
        }
        for (int j = 1; j <= ng; j++)
          for (int i = 0; i < NT; i++)
            for (int k = 0; k < nv; k++) {
              Q(n1, k, i, j + nymax) = Q(nbs, k, (i + ng) / 2, (j + 2 * ng - 1) / 2);
              Q(n2, k, i, j + nymax) = Q(nbs, k, (i - ng + 2) / 2 + nx2, (j + 2 * ng - 1) / 2);
            }
        break;
      case 10: /* top, same level */
        for (int i = 0; i < NT; i++)
          for (int k = 0; k < nv; k++) {
            jj = nymin;
            for (int j = nyp1; j < NT; j++, jj++) Q(nbs, k, i, j) = Q(n1, k, i, jj);
            jj = nymax - (ng - 1);
            for (int j = 0; j < ng; j++, jj++) Q(n1, k, i, j) = Q(nbs, k, i, jj);
          }
        break;
      case 11: /* top, finer neighbour */
        n2 = n1 + 1;
        for (int i = nxmin; i <= nx2; i++)
          for (int j = nyp1; j < NT; j++) {
            int i1 = 2 * i - nxmin, j1 = 2 * (j - nymax + (ng / 2 - 1)), nx2i = nx2 + i - 1 - (ng - 2);
            for (int k = 0; k < nv; k++) {
              Q(nbs, k, i, j) = FV(n1, k, i1, j1);
              Q(nbs, k, nx2i, j) = FV(n2, k, i1, j1);
            }
          }
        n3 = d->lp[d->lp[n1][1]][4];
        if (n3 != -1) {
          if (d->leafs[n3]) {
            for (int k = 0; k < nv; k++) for (int i = 0; i < ng; i++) for (int j = nymaxb; j < NT; j++) {
              ii = nxmax - ng + (i + 1); jj = nyminb + (j - nymax);
              Q(nbs, k, i, j) = Q(n3, k, ii, jj);
            }
          } else {
            n3 = d->lp[n3][2] + 1;
            if (d->leafs[n3])
              for (int k = 0; k < nv; k++) for (int i = 0; i < ng; i++) for (int j = nymaxb; j < NT; j++) {
                ii = nxmax - ng * 2 + 2 * i + 1; jj = nyminb + 2 * (j - nymax) - 1;
                Q(nbs, k, i, j) = FV(n3, k, ii, jj);
              }
          }
        }
        n4 = d->lp[d->lp[n2][1]][5];
        if (n4 != -1) {
          if (d->leafs[n4]) {
            for (int k = 0; k < nv; k++) for (int i = 0; i < ng; i++) for (int j = nymaxb; j < NT; j++) {
              ii = nxmin + ng - (i + 1); jj = nyminb + (j - nymax);
              Q(nbs, k, NT - 1 - i, j) = Q(n4, k, ii, jj);
            }
          } else {
            n4 = d->lp[n4][2];
            if (d->leafs[n4])
              for (int k = 0; k < nv; k++) for (int i = nxmaxb; i < NT; i++) for (int j = nymaxb; j < NT; j++) {
                ii = nxminb + 2 * (i - nxmax) - 1; jj = nyminb + 2 * (j - nymax) - 1;
                Q(nbs, k, i, j) = FV(n4, k, ii, jj);
              }
          }
        }
        for (int j = 1; j <= ng; j++)
          for (int i = 0; i < NT; i++)
            for (int k = 0; k < nv; k++) {
              Q(n1, k, i, j - 1) = Q(nbs, k, (i + ng) / 2, (j + 1) / 2 + nymax - ng / 2);
              Q(n2, k, i, j - 1) = Q(nbs, k, (i - ng + 2) / 2 + nx2, (j + 1) / 2 + nymax - ng / 2);
            }
        break;
      default: break;
    }
  }
  /* extboundary.cpp:5-61: mirror, applied after all internal entries */
  for (int e = 0; e < d->nbounds; e++) {
    int nbs = d->ib[e][0], ii, jj;
    switch (d->ib[e][2]) {
      case -1:
        ii = nxmin + (ng - 1);
        for (int i = 0; i < ng; i++, ii--)
          for (int j = 0; j < NT; j++) for (int k = 0; k < nv; k++) Q(nbs, k, i, j) = Q(nbs, k, ii, j);
        break;
      case -2:
        ii = nxmax;
        for (int i = nxp1; i < NT; i++, ii--)
          for (int j = 0; j < NT; j++) for (int k = 0; k < nv; k++) Q(nbs, k, i, j) = Q(nbs, k, ii, j);
        break;
      case -3:
        jj = nymin + (ng - 1);
        for (int j = 0; j < ng; j++, jj--)
          for (int i = 0; i < NT; i++) for (int k = 0; k < nv; k++) Q(nbs, k, i, j) = Q(nbs, k, i, jj);
        break;
      case -4:
        jj = nymax;
        for (int j = nyp1; j < NT; j++, jj--)
          for (int i = 0; i < NT; i++) for (int k = 0; k < nv; k++) Q(nbs, k, i, j) = Q(nbs, k, i, jj);
        break;
      default: break;
    }
  }
#undef BLK
#undef Q
#undef FV
  return 0;
}

/* ------------------------------------------------------------ stages */
/* main.cpp:144-159 / 176-195: MUSCL2D then the RK2 update of every cell
 * (timeIntegral.cpp:3-9); res is nonzero only on [nxminb,nxmax]^2. */
int ora_stage(ora_dom *d, int step) {
  const int NT = d->NT, lo = d->nxminb, hi = d->nxmax;
  for (int nb = 0; nb < d->lastActive; nb++) {
    if (!d->leafs[nb]) continue;
    double *U = TILE(d, ORA_U, nb), *Us = TILE(d, ORA_US, nb);
    const double *B = TILE(d, step == 1 ? ORA_BI : ORA_BIS, nb);
    double *ff = TILE(d, ORA_FF, nb), *gg = TILE(d, ORA_GG, nb);
    int lev = d->lp[nb][0];
    int st = muscl_tile(d, nb, NT, d->ng, d->lim, d->flux, d->gamma, d->dx[lev], step, step == 1 ? U : Us, B,
                        TILE(d, ORA_W, nb), ff, gg, d->scr);
    if (st) return st;
    double dxl = d->dx[lev], dyl = d->dy[lev], dt = d->dt_dom;
    for (int i = 0; i < NT; i++)
      for (int j = 0; j < NT; j++)
        for (int k = 0; k < 8; k++) {
          double res;
          if ((i >= lo && i <= hi) && (j >= lo && j <= hi))
            res = (AT(ff, NT, k, i + 1, j) - AT(ff, NT, k, i, j)) / dxl + (AT(gg, NT, k, i, j + 1) - AT(gg, NT, k, i, j)) / dyl;
          else
            res = 0.;
          if (step == 1) AT(Us, NT, k, i, j) = AT(U, NT, k, i, j) - dt * res;
          else if (d->rk == 2) AT(U, NT, k, i, j) = 0.5 * (AT(U, NT, k, i, j) + AT(Us, NT, k, i, j) - dt * res);
          /* SSP-RK3 (Shu-Osher), build-defined */
          else if (step == 2) AT(Us, NT, k, i, j) = 0.75 * AT(U, NT, k, i, j) + 0.25 * (AT(Us, NT, k, i, j) - dt * res);
          else AT(U, NT, k, i, j) = (1. / 3.) * AT(U, NT, k, i, j) + (2. / 3.) * (AT(Us, NT, k, i, j) - dt * res);
        }
  }
  return 0;
}

/* face-B stage update (CT2D.cpp:78-92) from the stage input B, the array
 * being written Bn and the step's starting value B0 = Bi; SSP-RK3 stages 2
 * and 3 are build-defined */
static double face_rk(int rk, int step, double B, double Bn, double B0, double dt, double res) {
  if (step == 1) return B - dt * res;
  if (rk == 2) return 0.5 * (B + Bn - dt * res);
  if (step == 2) return 0.75 * B0 + 0.25 * (B - dt * res);
  return (1. / 3.) * B0 + (2. / 3.) * (B - dt * res);
}

/* CT2D.cpp:28-75 (CTtype 2, "upwind"): per block, gas pressure p and fast
 * speed C of every cell of the stage state q, then each corner (i, j) in
 * [1, NT-1]^2 weighs the x-flux and y-flux averages by psi = dPx / (dPx +
 * dPy) where a strong pressure jump (sw1) meets a compressive flow (sw2),
 * else 0.5.  The velocity divergence keeps the reference's terms as written
 * (two of its y-velocity quotients divide q2 by itself, CT2D.cpp:45). */
static int ct_upwind_block(ora_dom *d, int nb, const double *q, const double *ff, const double *gg, double *E,
                           double *p, double *C) {
  const int NT = d->NT;
  const double g = d->gamma;
  for (int i = 0; i < NT; i++)
    for (int j = 0; j < NT; j++) {
      const double q0 = AT(q, NT, 0, i, j), q1 = AT(q, NT, 1, i, j), q2 = AT(q, NT, 2, i, j),
                   q3 = AT(q, NT, 3, i, j), q4 = AT(q, NT, 4, i, j), q5 = AT(q, NT, 5, i, j),
                   q6 = AT(q, NT, 6, i, j), q7 = AT(q, NT, 7, i, j);
      p[i * NT + j] = (g - 1) * (q4 - 0.5 * ((q1 * q1 + q2 * q2 + q3 * q3) / q0 + (q5 * q5 + q6 * q6 + q7 * q7)));
      C[i * NT + j] = sqrt((g * p[i * NT + j] + (q5 * q5 + q6 * q6 + q7 * q7)) / q0);
    }
#define P_(a, b) p[(a) * NT + (b)]
#define C_(a, b) C[(a) * NT + (b)]
#define Q_(k, a, b) AT(q, NT, k, a, b)
  for (int i = 1; i < NT; i++)
    for (int j = 1; j < NT; j++) {
      const double beta = 0.5, delta = 0.1;
      const double dPx = 0.5 * fabs(P_(i, j) + P_(i, j - 1) - P_(i - 1, j) - P_(i - 1, j - 1));
      const double dPy = 0.5 * fabs(P_(i, j) + P_(i - 1, j) - P_(i, j - 1) - P_(i - 1, j - 1));
      const double divV = 0.5 * (Q_(1, i, j) / Q_(0, i, j) + Q_(1, i, j - 1) / Q_(0, i, j - 1) -
                                 Q_(1, i - 1, j) / Q_(0, i - 1, j) - Q_(1, i - 1, j - 1) / Q_(0, i - 1, j - 1)) +
                          0.5 * (Q_(2, i, j) / Q_(0, i, j) + Q_(2, i - 1, j) / Q_(2, i - 1, j) -
                                 Q_(2, i, j - 1) / Q_(2, i, j - 1) - Q_(2, i - 1, j - 1) / Q_(0, i - 1, j - 1));
      double tP = P_(i - 1, j - 1) < P_(i, j - 1) ? P_(i - 1, j - 1) : P_(i, j - 1);
      const double tP1 = P_(i - 1, j) < P_(i, j) ? P_(i - 1, j) : P_(i, j);
      tP = tP < tP1 ? tP : tP1;
      const int sw1 = (dPx + dPy) > beta * tP;
      double tC = C_(i - 1, j - 1) < C_(i, j - 1) ? C_(i - 1, j - 1) : C_(i, j - 1);
      const double tC1 = C_(i - 1, j) < C_(i, j) ? C_(i - 1, j) : C_(i, j);
      tC = tC < tC1 ? tC : tC1;
      const int sw2 = (-delta * tC) > divV;
      double psi = 0.5;
      if (sw1 && sw2) psi = dPx / (dPx + dPy);
      if (psi < 0 || psi > 1) return fail(d, ORA_ERR_CT_PSI, nb, i, j);
      AT(E, NT, 0, i, j) = 0.5 * (1. - psi) * (AT(gg, NT, 5, i - 1, j) + AT(gg, NT, 5, i, j)) -
                           0.5 * psi * (AT(ff, NT, 6, i, j - 1) + AT(ff, NT, 6, i, j));
    }
#undef P_
#undef C_
#undef Q_
  return 0;
}

/* CT2D.cpp:11-124: corner EMF (CTtype 1, or CTtype 2 above), edge fill,
 * face-B update, face fill, then cell-centred B from faces with the energy
 * correction. */
int ora_ct(ora_dom *d, int step) {
  const int NT = d->NT;
  /* last stage updates Bi and U; earlier stages Bis and Us */
  const int last = step == d->rk;
  const int qf = last ? ORA_U : ORA_US, bf = step == 1 ? ORA_BI : ORA_BIS, nf = last ? ORA_BI : ORA_BIS;
  double *pc = d->ct == 2 ? (double *)malloc(sizeof(double) * 2 * NT * NT) : NULL;
  if (d->ct == 2 && !pc) return fail(d, ORA_ERR_CONFIG, -1, 0, 0);
  for (int nb = 0; nb < d->lastActive; nb++) {
    if (!d->leafs[nb]) continue;
    double *E = TILE(d, ORA_EMF, nb), *ff = TILE(d, ORA_FF, nb), *gg = TILE(d, ORA_GG, nb);
    if (d->ct == 2) {
      const int e = ct_upwind_block(d, nb, TILE(d, qf, nb), ff, gg, E, pc, pc + NT * NT);
      if (e) {
        free(pc);
        return e;
      }
      continue;
    }
    for (int i = 1; i < NT; i++)
      for (int j = 1; j < NT; j++)
        AT(E, NT, 0, i, j) = 0.25 * ((AT(gg, NT, 5, i - 1, j) + AT(gg, NT, 5, i, j)) -
                                     (AT(ff, NT, 6, i, j - 1) + AT(ff, NT, 6, i, j)));
  }
  free(pc);
  ora_fill(d, ORA_EMF, ORA_FILL_EC);
  for (int nb = 0; nb < d->lastActive; nb++) {
    if (!d->leafs[nb]) continue;
    double *E = TILE(d, ORA_EMF, nb), *B = TILE(d, bf, nb), *Bn = TILE(d, nf, nb);
    int lev = d->lp[nb][0];
    double dt = d->dt_dom;
    for (int i = 0; i < NT; i++)
      for (int j = 0; j < NT; j++) {
        int ip = i < NT - 1 ? i + 1 : i, jp = j < NT - 1 ? j + 1 : j;
        double res = (1. / d->dy[lev]) * (AT(E, NT, 0, i, jp) - AT(E, NT, 0, i, j));
        AT(Bn, NT, 0, i, j) = face_rk(d->rk, step, AT(B, NT, 0, i, j), AT(Bn, NT, 0, i, j), TILE(d, ORA_BI, nb)[i + NT * j],
                                      dt, res);
        res = -(1. / d->dx[lev]) * (AT(E, NT, 0, ip, j) - AT(E, NT, 0, i, j));
        AT(Bn, NT, 1, i, j) = face_rk(d->rk, step, AT(B, NT, 1, i, j), AT(Bn, NT, 1, i, j),
                                      TILE(d, ORA_BI, nb)[NT * NT + i + NT * j], dt, res);
      }
  }
  ora_fill(d, nf, ORA_FILL_FC);
  for (int nb = 0; nb < d->lastActive; nb++) {
    if (!d->leafs[nb]) continue;
    double *q = TILE(d, qf, nb), *Bn = TILE(d, nf, nb);
    for (int i = 0; i < NT; i++)
      for (int j = 0; j < NT; j++) {
        int ip = i < NT - 1 ? i + 1 : i, ii = ip - 1;
        int jp = j < NT - 1 ? j + 1 : j, jj = jp - 1;
        double bx = 0.5 * (AT(Bn, NT, 0, ip, j) + AT(Bn, NT, 0, ii, j));
        double by = 0.5 * (AT(Bn, NT, 1, i, jp) + AT(Bn, NT, 1, i, jj));
        double q5 = AT(q, NT, 5, i, j), q6 = AT(q, NT, 6, i, j);
        AT(q, NT, 4, i, j) = AT(q, NT, 4, i, j) + 0.5 * ((bx * bx + by * by) - ((q5 * q5) + (q6 * q6)));
        AT(q, NT, 5, i, j) = bx;
        AT(q, NT, 6, i, j) = by;
      }
  }
  return 0;
}

/* main.cpp:209-230: per leaf, U2W then the (reference's) wave-speed bound */
int ora_cfl(ora_dom *d, double *lambda_out) {
  const int NT = d->NT;
  double lambda = 1.E6;
  for (int nb = 0; nb < d->lastActive; nb++) {
    if (!d->leafs[nb]) continue;
    double *W = TILE(d, ORA_W, nb);
    int st = cons2prim(d, NT, d->nxminb, d->nxmaxb, d->gamma, TILE(d, ORA_U, nb), W, nb, 0);
    if (st) return st;
    double l1 = 0., l2 = 0., a, ca, can, cf, w[8];
    for (int i = d->nxmin; i <= d->nxmax; i++)
      for (int j = d->nxmin; j <= d->nxmax; j++) {
        for (int k = 0; k < 8; k++) w[k] = AT(W, NT, k, i, j);
        wavespeeds(w, d->gamma, 5, &a, &ca, &can, &cf);
        l1 = stdmax(l1, fabs(w[0]) + cf);
        wavespeeds(w, d->gamma, 6, &a, &ca, &can, &cf);
        l2 = stdmax(l2, fabs(w[1]) + cf);
      }
    l1 = stdmax(l1, l2);
    int lev = d->lp[nb][0];
    lambda = stdmin(lambda, stdmin(d->dx[lev], d->dy[lev]) / l1);
  }
  *lambda_out = lambda;
  return 0;
}

/* -------------------------------------------------------------- regrid */
/* refine.cpp:17-36 */
static double refine_slope(const double *Q, int NT, int k, int i, int j, int sx, int sy) {
  double res = AT(Q, NT, k, i, j);
  double r = (AT(Q, NT, k, i, j) - AT(Q, NT, k, i - 1, j)) / (AT(Q, NT, k, i + 1, j) - AT(Q, NT, k, i, j) + EPS);
  double phi = 0.25 * limiter(L_MINMOD, r, 0) * (AT(Q, NT, k, i + 1, j) - AT(Q, NT, k, i, j));
  res = sx == 0 ? res - phi : res + phi;
  r = (AT(Q, NT, k, i, j) - AT(Q, NT, k, i, j - 1)) / (AT(Q, NT, k, i, j + 1) - AT(Q, NT, k, i, j) + EPS);
  phi = 0.25 * limiter(L_MINMOD, r, 0) * (AT(Q, NT, k, i, j + 1) - AT(Q, NT, k, i, j));
  res = sy == 0 ? res - phi : res + phi;
  return res;
}

/* refineBi_CT.cpp:4-50 (son faces evaluated at sx,sy in {0,1/2} of a parent cell) */
static int refine_face(const double *B, int NT, int i, int j, int k, int signx, int signy, double *out) {
  double Bx_jm = AT(B, NT, 0, i, j - 1), Bx = AT(B, NT, 0, i, j), Bx_jp = AT(B, NT, 0, i, j + 1);
  double Bx_ip = AT(B, NT, 0, i + 1, j), Bx_ipjm = AT(B, NT, 0, i + 1, j - 1), Bx_ipjp = AT(B, NT, 0, i + 1, j + 1);
  double By_im = AT(B, NT, 1, i - 1, j), By = AT(B, NT, 1, i, j), By_ip = AT(B, NT, 1, i + 1, j);
  double By_imjp = AT(B, NT, 1, i - 1, j + 1), By_jp = AT(B, NT, 1, i, j + 1), By_ipjp = AT(B, NT, 1, i + 1, j + 1);
  double r;
  r = (Bx - Bx_jm) / (Bx_jp - Bx + EPS);
  double DyBxm = limiter(L_MINMOD, r, 0) * (Bx_jp - Bx);
  r = (Bx_ip - Bx_ipjm) / (Bx_ipjp - Bx_ip + EPS);
  double DyBxp = limiter(L_MINMOD, r, 0) * (Bx_ipjp - Bx_ip);
  r = (By - By_im) / (By_ip - By + EPS);
  double DxBym = limiter(L_MINMOD, r, 0) * (By_ip - By);
  r = (By_jp - By_imjp) / (By_ipjp - By_jp + EPS);
  double DxByp = limiter(L_MINMOD, r, 0) * (By_ipjp - By_jp);
  double ax = (Bx_ip - Bx), by = -1. * ax;
  double ay = 0.5 * (DyBxp + DyBxm), bx = 0.5 * (DxByp + DxBym);
  double axy = (DyBxp - DyBxm), byy = axy / (-2.);
  double bxy = (DxByp - DxBym), axx = bxy / (-2.);
  double a0 = 0.5 * (Bx_ip + Bx) - 0.25 * axx, b0 = 0.5 * (By_jp + By) - 0.25 * byy;
  double ayy = 0., bxx = 0., res;
  if (k == 0) {
    double sx = signx == 0 ? 0 : 0.5, sy = signy == 0 ? -0.25 : 0.25;
    res = a0 + ax * sx + ay * sy + axx * sx * sx + axy * sx * sy + ayy * sy * sy;
  } else {
    double sx = signx == 0 ? -0.25 : 0.25, sy = signy == 0 ? 0 : 0.5;
    res = b0 + bx * sx + by * sy + bxx * sx * sx + bxy * sx * sy + byy * sy * sy;
  }
  *out = res;
  return isnan(res) ? ORA_ERR_REFINE_NAN : 0;
}

/* refine.cpp:38-59 + refineBi_CT.cpp:52-74 (tile part) for one parent */
static int refine_tiles(int NT, int ng, const double *Up, const double *Bp, double *Us4[4], double *Bs4[4]) {
  int nx2 = (NT - 1) / 2, ny2 = nx2;
  for (int i = 0; i < NT; i++)
    for (int j = 0; j < NT; j++) {
      int sx = (i + ng) % 2, sy = (j + ng) % 2;
      int iA = (i + ng) / 2, iB = (i - ng + 2) / 2 + nx2;
      int jA = (j + ng) / 2, jB = (j - ng + 2) / 2 + ny2;
      for (int k = 0; k < 8; k++) {
        AT(Us4[0], NT, k, i, j) = refine_slope(Up, NT, k, iA, jA, sx, sy);
        AT(Us4[1], NT, k, i, j) = refine_slope(Up, NT, k, iB, jA, sx, sy);
        AT(Us4[2], NT, k, i, j) = refine_slope(Up, NT, k, iA, jB, sx, sy);
        AT(Us4[3], NT, k, i, j) = refine_slope(Up, NT, k, iB, jB, sx, sy);
      }
    }
  (void)Bp; (void)Bs4;
  return 0;
}

static int refine_face_tiles(int NT, int ng, const double *Bp, double *Bs4[4]) {
  int nx2 = (NT - 1) / 2, ny2 = nx2, st;
  for (int i = 0; i < NT; i++)
    for (int j = 0; j < NT; j++)
      for (int k = 0; k < 2; k++) {
        int sx = (i + ng) % 2, sy = (j + ng) % 2;
        int iA = (i + ng) / 2, iB = (i - ng + 2) / 2 + nx2;
        int jA = (j + ng) / 2, jB = (j - ng + 2) / 2 + ny2;
        if ((st = refine_face(Bp, NT, iA, jA, k, sx, sy, &AT(Bs4[0], NT, k, i, j)))) return st;
        if ((st = refine_face(Bp, NT, iB, jA, k, sx, sy, &AT(Bs4[1], NT, k, i, j)))) return st;
        if ((st = refine_face(Bp, NT, iA, jB, k, sx, sy, &AT(Bs4[2], NT, k, i, j)))) return st;
        if ((st = refine_face(Bp, NT, iB, jB, k, sx, sy, &AT(Bs4[3], NT, k, i, j)))) return st;
      }
  return 0;
}

static inline double sum2x2(const double *Q, int NT, int k, int i1, int i2, int j1, int j2) {
  double v = 0.;
  for (int i = i1; i <= i2; i++)
    for (int j = j1; j <= j2; j++) v += AT(Q, NT, k, i, j);
  return v;
}

/* coarsen.cpp:7-148 (interior + ghost ring restriction, clamped indices) */
static void coarsen_tiles(int NT, int ng, double *Qd, double *S[4], int nv) {
  int nxmin = ng, nxmax = NT - 1 - ng, nxminb = nxmin - 1, nxmaxb = nxmax + 1, nx2 = (NT - 1) / 2;
  int nymin = nxmin, nyminb = nxminb, nymaxb = nxmaxb, ny2 = nx2;
  int i1, i2, j1, j2;
#define C4(s) (0.25 * sum2x2(S[s], NT, k, i1, i2, j1, j2))
  for (int i = nxmin; i <= nx2; i++)
    for (int j = nymin; j <= ny2; j++) {
      i1 = 2 * i - nxmin; i2 = i1 + 1; j1 = 2 * j - nymin; j2 = j1 + 1;
      int nx2i = nx2 + i - 1 - (ng - 2), ny2j = ny2 + j - 1 - (ng - 2);
      for (int k = 0; k < nv; k++) {
        AT(Qd, NT, k, i, j) = C4(0);
        AT(Qd, NT, k, nx2i, j) = C4(1);
        AT(Qd, NT, k, i, ny2j) = C4(2);
        AT(Qd, NT, k, nx2i, ny2j) = C4(3);
      }
    }
  /* left */
  for (int k = 0; k < nv; k++)
    for (int i = 0; i < ng / 2; i++)
      for (int j = 0; j < ng / 2; j++) {
        i2 = nxminb - 2 * i; i1 = i2 - 1; j2 = nyminb - 2 * j; j1 = j2 - 1;
        AT(Qd, NT, k, nxminb - i, nyminb - j) = C4(0);
        j1 = nymaxb + 2 * j; j2 = j1 + 1;
        AT(Qd, NT, k, nxminb - i, nymaxb + j) = C4(2);
      }
  for (int i = 0; i <= nxminb; i++)
    for (int j = nymin; j <= ny2; j++) {
      j1 = 2 * j - nymin; j2 = j1 + 1;
      int ny2j = j + ny2 - 1 - (ng - 2);
      i2 = nxminb - 2 * i; i1 = i2 - 1;
      if (i1 < 0) { i1 = 0; i2 = 1; }
      for (int k = 0; k < nv; k++) {
        AT(Qd, NT, k, nxminb - i, j) = C4(0);
        AT(Qd, NT, k, nxminb - i, ny2j) = C4(2);
      }
    }
  /* right */
  for (int k = 0; k < nv; k++)
    for (int i = 0; i < ng / 2; i++)
      for (int j = 0; j < ng / 2; j++) {
        i1 = nxmaxb + 2 * i; i2 = i1 + 1; j2 = nyminb - 2 * j; j1 = j2 - 1;
        AT(Qd, NT, k, nxmaxb + i, nyminb - j) = C4(1);
        j1 = nymaxb + 2 * j; j2 = j1 + 1;
        AT(Qd, NT, k, nxmaxb + i, nymaxb + j) = C4(3);
      }
  for (int i = 0; i < ng; i++)
    for (int j = nymin; j <= ny2; j++) {
      j1 = 2 * j - nymin; j2 = j1 + 1;
      int ny2j = j + ny2 - 1 - (ng - 2);
      i1 = nxmaxb + 2 * i; i2 = i1 + 1;
      if (i2 > NT - 1) { i1 = NT - 2; i2 = i1 + 1; }
      for (int k = 0; k < nv; k++) {
        AT(Qd, NT, k, nxmaxb + i, j) = C4(1);
        AT(Qd, NT, k, nxmaxb + i, ny2j) = C4(3);
      }
    }
  /* bottom */
  for (int k = 0; k < nv; k++)
    for (int j = 0; j < ng / 2; j++)
      for (int i = 0; i < ng / 2; i++) {
        j2 = nyminb - 2 * j; j1 = j2 - 1; i2 = nxminb - 2 * i; i1 = i2 - 1;
        AT(Qd, NT, k, nxminb - i, nyminb - j) = C4(0);
        i1 = nxmaxb + 2 * i; i2 = i1 + 1;
        AT(Qd, NT, k, nxmaxb + i, nyminb - j) = C4(1);
      }
  for (int i = nxmin; i <= nx2; i++)
    for (int j = 0; j <= nyminb; j++) {
      i1 = 2 * i - nxmin; i2 = i1 + 1;
      int nx2i = i + nx2 - 1 - (ng - 2);
      j2 = nyminb - 2 * j; j1 = j2 - 1;
      if (j1 < 0) { j1 = 0; j2 = 1; }
      for (int k = 0; k < nv; k++) {
        AT(Qd, NT, k, i, nyminb - j) = C4(0);
        AT(Qd, NT, k, nx2i, nyminb - j) = C4(1);
      }
    }
  /* top */
  for (int k = 0; k < nv; k++)
    for (int j = 0; j < ng / 2; j++)
      for (int i = 0; i < ng / 2; i++) {
        j1 = nymaxb + 2 * j; j2 = j1 + 1; i2 = nxminb - 2 * i; i1 = i2 - 1;
        AT(Qd, NT, k, nxminb - i, nymaxb + j) = C4(2);
        i1 = nxmaxb + 2 * i; i2 = i1 + 1;
        AT(Qd, NT, k, nxmaxb + i, nymaxb + j) = C4(3);
      }
  for (int i = nxmin; i <= nx2; i++)
    for (int j = 0; j < ng; j++) {
      i1 = 2 * i - nxmin; i2 = i1 + 1;
      int nx2i = i + nx2 - 1 - (ng - 2);
      j1 = nymaxb + 2 * j; j2 = j1 + 1;
      if (j2 > NT - 1) { j1 = NT - 2; j2 = j1 + 1; }
      for (int k = 0; k < nv; k++) {
        AT(Qd, NT, k, i, nymaxb + j) = C4(2);
        AT(Qd, NT, k, nx2i, nymaxb + j) = C4(3);
      }
    }
#undef C4
}

/* coarsenBi_CT.cpp:7-155: face pairs (k==0 along j, k==1 along i) */
static void coarsen_face_tiles(int NT, int ng, double *Bd, double *S[4]) {
  int nxmin = ng, nxmax = NT - 1 - ng, nxminb = nxmin - 1, nxmaxb = nxmax + 1, nx2 = (NT - 1) / 2;
  int nymin = nxmin, nyminb = nxminb, nymaxb = nxmaxb, ny2 = nx2;
  int i1, i2, j1, j2;
#define C2(s) (0.5 * sum2x2(S[s], NT, k, i1, i2, j1, j2))
#define PAIR() do { i2 = k == 0 ? i1 : i1 + 1; j2 = k == 0 ? j1 + 1 : j1; } while (0)
  for (int i = nxmin; i <= nx2; i++)
    for (int j = nymin; j <= ny2; j++) {
      i1 = 2 * i - nxmin; j1 = 2 * j - nymin;
      int nx2i = nx2 + i - 1 - (ng - 2), ny2j = ny2 + j - 1 - (ng - 2);
      for (int k = 0; k < 2; k++) {
        PAIR();
        AT(Bd, NT, k, i, j) = C2(0);
        AT(Bd, NT, k, nx2i, j) = C2(1);
        AT(Bd, NT, k, i, ny2j) = C2(2);
        AT(Bd, NT, k, nx2i, ny2j) = C2(3);
      }
    }
  /* left */
  for (int k = 0; k < 2; k++)
    for (int i = 0; i < ng / 2; i++)
      for (int j = 0; j < ng / 2; j++) {
        i1 = nxminb - 2 * i - 1; i2 = k == 0 ? i1 : i1 + 1;
        j1 = nyminb - 2 * j - 1; j2 = k == 0 ? j1 + 1 : j1;
        AT(Bd, NT, k, nxminb - i, nyminb - j) = C2(0);
        j1 = nymaxb + 2 * j; j2 = k == 0 ? j1 + 1 : j1;
        AT(Bd, NT, k, nxminb - i, nymaxb + j) = C2(2);
      }
  for (int i = 0; i <= nxminb; i++)
    for (int j = nymin; j <= ny2; j++) {
      j1 = 2 * j - nymin;
      int ny2j = j + ny2 - 1 - (ng - 2);
      i1 = nxminb - 2 * i - 1;
      if (i1 < 0) i1 = 0;
      for (int k = 0; k < 2; k++) {
        PAIR();
        AT(Bd, NT, k, nxminb - i, j) = C2(0);
        AT(Bd, NT, k, nxminb - i, ny2j) = C2(2);
      }
    }
  /* right */
  for (int k = 0; k < 2; k++)
    for (int i = 0; i < ng / 2; i++)
      for (int j = 0; j < ng / 2; j++) {
        i1 = nxmaxb + 2 * i; i2 = k == 0 ? i1 : i1 + 1;
        j1 = nyminb - 2 * j - 1; j2 = k == 0 ? j1 + 1 : j1;
        AT(Bd, NT, k, nxmaxb + i, nyminb - j) = C2(1);
        j1 = nymaxb + 2 * j; j2 = k == 0 ? j1 + 1 : j1;
        AT(Bd, NT, k, nxmaxb + i, nymaxb + j) = C2(3);
      }
  for (int i = 0; i < ng; i++)
    for (int j = nymin; j <= ny2; j++) {
      j1 = 2 * j - nymin;
      int ny2j = j + ny2 - 1 - (ng - 2);
      i1 = nxmaxb + 2 * i;
      if (i1 + 1 > NT - 1) i1 = NT - 2;
      for (int k = 0; k < 2; k++) {
        PAIR();
        AT(Bd, NT, k, nxmaxb + i, j) = C2(1);
        AT(Bd, NT, k, nxmaxb + i, ny2j) = C2(3);
      }
    }
  /* bottom */
  for (int k = 0; k < 2; k++)
    for (int j = 0; j < ng / 2; j++)
      for (int i = 0; i < ng / 2; i++) {
        j1 = nyminb - 2 * j - 1; j2 = k == 0 ? j1 + 1 : j1;
        i1 = nxminb - 2 * i - 1; i2 = k == 0 ? i1 : i1 + 1;
        AT(Bd, NT, k, nxminb - i, nyminb - j) = C2(0);
        i1 = nxmaxb + 2 * i; i2 = k == 0 ? i1 : i1 + 1;
        AT(Bd, NT, k, nxmaxb + i, nyminb - j) = C2(1);
      }
  for (int i = nxmin; i <= nx2; i++)
    for (int j = 0; j <= nyminb; j++) {
      i1 = 2 * i - nxmin;
      int nx2i = i + nx2 - 1 - (ng - 2);
      j1 = nyminb - 2 * j - 1;
      if (j1 < 0) j1 = 0;
      for (int k = 0; k < 2; k++) {
        PAIR();
        AT(Bd, NT, k, i, nyminb - j) = C2(0);
        AT(Bd, NT, k, nx2i, nyminb - j) = C2(1);
      }
    }
  /* top */
  for (int k = 0; k < 2; k++)
    for (int j = 0; j < ng / 2; j++)
      for (int i = 0; i < ng / 2; i++) {
        j1 = nymaxb + 2 * j; j2 = k == 0 ? j1 + 1 : j1;
        i1 = nxminb - 2 * i - 1; i2 = k == 0 ? i1 : i1 + 1;
        AT(Bd, NT, k, nxminb - i, nymaxb + j) = C2(2);
        i1 = nxmaxb + 2 * i; i2 = k == 0 ? i1 : i1 + 1;
        AT(Bd, NT, k, nxmaxb + i, nymaxb + j) = C2(3);
      }
  for (int i = nxmin; i <= nx2; i++)
    for (int j = 0; j < ng; j++) {
      i1 = 2 * i - nxmin;
      int nx2i = i + nx2 - 1 - (ng - 2);
      j1 = nymaxb + 2 * j;
      if (j1 + 1 > NT - 1) j1 = NT - 2;
      for (int k = 0; k < 2; k++) {
        PAIR();
        AT(Bd, NT, k, i, nymaxb + j) = C2(2);
        AT(Bd, NT, k, nx2i, nymaxb + j) = C2(3);
      }
    }
#undef C2
#undef PAIR
}

/* updatelpup.cpp:4-118 */
static void link_sons(ora_dom *d, int nb, const int s[4]) {
  int32_t(*lp)[8] = d->lp;
  int lev = lp[nb][0] + 1;
  /* son q: sibling id q; internal links; external links inherit the parent's
   * neighbour when that is a leaf or the box, else the adjacent son of it. */
  lp[s[0]][0] = lev; lp[s[0]][1] = nb; lp[s[0]][3] = 0; lp[s[0]][5] = s[1]; lp[s[0]][7] = s[2];
  if (lp[nb][4] == -1 || d->leafs[lp[nb][4]]) lp[s[0]][4] = lp[nb][4];
  else { int n = lp[lp[nb][4]][2] + 1; lp[s[0]][4] = n; lp[n][5] = s[0]; }
  if (lp[nb][6] == -1 || d->leafs[lp[nb][6]]) lp[s[0]][6] = lp[nb][6];
  else { int n = lp[lp[nb][6]][2] + 2; lp[s[0]][6] = n; lp[n][7] = s[0]; }
  lp[s[1]][0] = lev; lp[s[1]][1] = nb; lp[s[1]][3] = 1; lp[s[1]][4] = s[0]; lp[s[1]][7] = s[3];
  if (lp[nb][5] == -1 || d->leafs[lp[nb][5]]) lp[s[1]][5] = lp[nb][5];
  else { int n = lp[lp[nb][5]][2]; lp[s[1]][5] = n; lp[n][4] = s[1]; }
  if (lp[nb][6] == -1 || d->leafs[lp[nb][6]]) lp[s[1]][6] = lp[nb][6];
  else { int n = lp[lp[nb][6]][2] + 3; lp[s[1]][6] = n; lp[n][7] = s[1]; }
  lp[s[2]][0] = lev; lp[s[2]][1] = nb; lp[s[2]][3] = 2; lp[s[2]][5] = s[3]; lp[s[2]][6] = s[0];
  if (lp[nb][4] == -1 || d->leafs[lp[nb][4]]) lp[s[2]][4] = lp[nb][4];
  else { int n = lp[lp[nb][4]][2] + 3; lp[s[2]][4] = n; lp[n][5] = s[2]; }
  if (lp[nb][7] == -1 || d->leafs[lp[nb][7]]) lp[s[2]][7] = lp[nb][7];
  else { int n = lp[lp[nb][7]][2]; lp[s[2]][7] = n; lp[n][6] = s[2]; }
  lp[s[3]][0] = lev; lp[s[3]][1] = nb; lp[s[3]][3] = 3; lp[s[3]][4] = s[2]; lp[s[3]][6] = s[1];
  if (lp[nb][5] == -1 || d->leafs[lp[nb][5]]) lp[s[3]][5] = lp[nb][5];
  else { int n = lp[lp[nb][5]][2] + 2; lp[s[3]][5] = n; lp[n][4] = s[3]; }
  if (lp[nb][7] == -1 || d->leafs[lp[nb][7]]) lp[s[3]][7] = lp[nb][7];
  else { int n = lp[lp[nb][7]][2] + 1; lp[s[3]][7] = n; lp[n][6] = s[3]; }
  d->leafs[nb] = 0;
  for (int q = 0; q < 4; q++) { d->leafs[s[q]] = 1; d->icoarse[s[q]] = 0; }
  d->iref[nb] = 0;
  d->icoarse[nb] = 0;
}

/* updatelpdown.cpp:5-54 */
static void unlink_sons(ora_dom *d, int dad, const int s[4]) {
  int32_t(*lp)[8] = d->lp;
  d->leafs[dad] = 1;
  for (int q = 0; q < 4; q++) d->leafs[s[q]] = 0;
  int n = lp[dad][4];
  if (n != -1 && !d->leafs[n]) { n = lp[lp[n][2]][5]; lp[n][5] = dad; lp[lp[n][7]][5] = dad; }
  n = lp[dad][5];
  if (n != -1 && !d->leafs[n]) { n = lp[n][2]; lp[n][4] = dad; lp[lp[n][7]][4] = dad; }
  n = lp[dad][6];
  if (n != -1 && !d->leafs[n]) { n = lp[lp[n][2]][7]; lp[n][7] = dad; lp[lp[n][5]][7] = dad; }
  n = lp[dad][7];
  if (n != -1 && !d->leafs[n]) { n = lp[n][2]; lp[n][6] = dad; lp[lp[n][5]][6] = dad; }
}

/* misc.cpp:26-32 */
static inline double max4diff(const double *W, int NT, int k, int i, int j) {
  double c = AT(W, NT, k, i, j);
  double a = fabs(AT(W, NT, k, i + 1, j) - c);
  double b = MAXm(a, fabs(AT(W, NT, k, i - 1, j) - c));
  a = MAXm(b, fabs(AT(W, NT, k, i, j + 1) - c));
  b = MAXm(a, fabs(AT(W, NT, k, i, j - 1) - c));
  return b;
}

/* criteria.cpp:6-44: returns 1 refine, 2 coarsen, 0 neither */
static int criteria_tile(const double *W, int NT, int ng, int level, int maxlevs) {
  double grad = 0;
  static const int vars[4] = {4, 0, 5, 6};
  for (int i = ng; i <= NT - 1 - ng; i++)
    for (int j = ng; j <= NT - 1 - ng; j++) {
      for (int q = 0; q < 4; q++) {
        double dm = MAXm(EPS, fabs(AT(W, NT, vars[q], i, j)));
        double v = max4diff(W, NT, vars[q], i, j);
        grad = MAXm(grad, v / dm);
      }
      if (grad >= 0.05) return level < maxlevs - 1 ? 1 : 0;
    }
  return grad <= 1.0 ? 2 : 0;
}

int ora_criteria_tile(int NT, int ng, int maxlevs, int level, const double *W, int32_t *o) {
  int r = criteria_tile(W, NT, ng, level, maxlevs);
  o[0] = r == 1;
  o[1] = r == 2;
  return 0;
}

/* admesh.cpp:17-102 with markref.cpp:6-18, critneighup.cpp:6-35,
 * critneighdown.cpp:6-33 */
int ora_regrid(ora_dom *d) {
  const int NT = d->NT, ng = d->ng;
  int32_t(*lp)[8] = d->lp;
  for (int nb = 0; nb < d->lastActive; nb++) { d->iref[nb] = 0; d->icoarse[nb] = 0; }
  for (int nb = 0; nb < d->lastActive; nb++)
    if (d->leafs[nb] && lp[nb][0] >= 0) {
      int r = criteria_tile(TILE(d, ORA_W, nb), NT, ng, lp[nb][0], d->maxlevs);
      if (r == 1) { d->iref[nb] = 1; d->icoarse[nb] = 0; }
      else if (r == 2) d->icoarse[nb] = 1;
    }
  for (int nl = d->maxlevs - 1; nl >= 0; nl--)
    for (int nb = 0; nb < d->lastActive; nb++)
      if (d->iref[nb] && lp[nb][0] == nl)
        for (int s = 4; s < 8; s++) {
          int n1 = lp[nb][s];
          if (n1 != -1 && lp[n1][0] < lp[nb][0]) d->iref[n1] = 1;
        }
  for (int nl = 0; nl < d->maxlevs - 1; nl++)
    for (int nb = 0; nb < d->lastActive; nb++) {
      if (!(lp[nb][0] == nl && d->iref[nb])) continue;
      int s[4];
      if (lp[nb][2] == 0) {
        s[0] = d->lastActive; s[1] = s[0] + 1; s[2] = s[0] + 2; s[3] = s[0] + 3;
        lp[nb][2] = s[0];
        d->lastActive += 4;
        if (d->lastActive - 1 >= d->maxblocks) return fail(d, ORA_ERR_OVERFLOW, nb, 0, 0);
      } else {
        s[0] = lp[nb][2]; s[1] = lp[s[0]][5]; s[2] = lp[s[0]][7]; s[3] = lp[s[2]][5];
      }
      double *Us4[4], *Bs4[4];
      for (int q = 0; q < 4; q++) { Us4[q] = TILE(d, ORA_U, s[q]); Bs4[q] = TILE(d, ORA_BI, s[q]); }
      refine_tiles(NT, ng, TILE(d, ORA_U, nb), NULL, Us4, NULL);
      /* refine.cpp:61-84: son coordinates and primitive refresh */
      int x0 = d->icoord[nb][0] * 2, y0 = d->icoord[nb][1] * 2, side = NT - 2 * ng;
      for (int q = 0; q < 4; q++) {
        int st = cons2prim(d, NT, d->nxminb, d->nxmaxb, d->gamma, Us4[q], TILE(d, ORA_W, s[q]), s[q], 0);
        if (st) return st;
        d->icoord[s[q]][0] = x0 + ((q & 1) ? side : 0);
        d->icoord[s[q]][1] = y0 + ((q & 2) ? side : 0);
      }
      link_sons(d, nb, s);
      int st = refine_face_tiles(NT, ng, TILE(d, ORA_BI, nb), Bs4);
      if (st) return fail(d, st, nb, 0, 0);
    }
  for (int nb = 0; nb < d->lastActive; nb++) {
    if (!d->icoarse[nb]) continue;
    for (int s = 4; s < 8; s++) {
      int n1 = lp[nb][s];
      if (n1 != -1 && !d->leafs[n1]) d->icoarse[nb] = 0;
    }
  }
  for (int nl = d->maxlevs - 1; nl >= 1; nl--)
    for (int nb = 0; nb < d->lastActive; nb++) {
      if (!(d->icoarse[nb] && lp[nb][3] == 0 && lp[nb][0] == nl)) continue;
      int dad = lp[nb][1], s[4];
      s[0] = nb; s[1] = lp[nb][5]; s[2] = lp[nb][7]; s[3] = lp[s[1]][7];
      if (!(d->icoarse[s[1]] && d->icoarse[s[2]] && d->icoarse[s[3]])) continue;
      double *Us4[4], *Bs4[4];
      for (int q = 0; q < 4; q++) { Us4[q] = TILE(d, ORA_U, s[q]); Bs4[q] = TILE(d, ORA_BI, s[q]); }
      coarsen_tiles(NT, ng, TILE(d, ORA_U, dad), Us4, 8);
      int st = cons2prim(d, NT, d->nxminb, d->nxmaxb, d->gamma, TILE(d, ORA_U, dad), TILE(d, ORA_W, dad), dad, 0);
      if (st) return st;
      unlink_sons(d, dad, s);
      coarsen_face_tiles(NT, ng, TILE(d, ORA_BI, dad), Bs4);
    }
  return 0;
}

/* ---------------------------------------------------------------- driver */
int ora_init(ora_dom *d) {
  basegrid(d);
  for (int nb = 0; nb < d->lastActive; nb++)
    if (d->leafs[nb]) prim2cons_tile(d->NT, d->gamma, TILE(d, ORA_W, nb), TILE(d, ORA_U, nb));
  d->t = 0.0; d->dt_main = 0.0; d->dt_dom = 0.0; d->count = 0;
  if (d->maxlevs > 1) return ora_regrid(d);
  return 0;
}

int ora_done(const ora_dom *d) { return !(d->t <= d->tEnd); }

/* main.cpp:128-247 */
int ora_step(ora_dom *d) {
  int st;
  ora_locate_bounds(d);
  ora_fill(d, ORA_U, ORA_FILL_CC);
  ora_fill(d, ORA_BI, ORA_FILL_FC);
  if ((st = ora_stage(d, 1))) return st;
  ora_fill(d, ORA_BI, ORA_FILL_FC);
  ora_fill(d, ORA_FF, ORA_FILL_FC);
  ora_fill(d, ORA_GG, ORA_FILL_FC);
  ora_ct(d, 1);
  ora_fill(d, ORA_US, ORA_FILL_CC);
  if ((st = ora_stage(d, 2))) return st;
  ora_fill(d, ORA_FF, ORA_FILL_FC);
  ora_fill(d, ORA_GG, ORA_FILL_FC);
  ora_ct(d, 2);
  if (d->rk == 3) {
    ora_fill(d, ORA_US, ORA_FILL_CC);
    if ((st = ora_stage(d, 3))) return st;
    ora_fill(d, ORA_FF, ORA_FILL_FC);
    ora_fill(d, ORA_GG, ORA_FILL_FC);
    ora_ct(d, 3);
  }
  ora_fill(d, ORA_U, ORA_FILL_CC);
  double lambda;
  if ((st = ora_cfl(d, &lambda))) return st;
  d->t = d->t + d->dt_main;
  d->dt_dom = d->dt_main;
  d->count = d->count + 1;
  d->dt_main = d->cfl * lambda;
  if (d->dt_main < 0 || isnan(d->dt_main)) return fail(d, ORA_ERR_DT, -1, 0, 0);
  if (d->maxlevs > 1) return ora_regrid(d);
  return 0;
}

int ora_run(ora_dom *d, int max_steps, int32_t *nleaf_log) {
  int n = 0;
  while (!ora_done(d) && (max_steps < 0 || n < max_steps)) {
    if (nleaf_log) {
      int c = 0;
      for (int nb = 0; nb < d->lastActive; nb++) c += d->leafs[nb];
      nleaf_log[n] = c;
    }
    int st = ora_step(d);
    if (st) return st;
    n++;
  }
  return 0;
}

/* ------------------------------------------------------------- accessors */
void ora_info(const ora_dom *d, int32_t *iv, double *dv) {
  int c = 0;
  for (int nb = 0; nb < d->lastActive; nb++) c += d->leafs[nb];
  iv[0] = d->lastActive; iv[1] = c; iv[2] = d->count; iv[3] = d->maxblocks;
  dv[0] = d->t; dv[1] = d->dt_dom; dv[2] = d->dt_main; dv[3] = d->gamma;
}

static const uint64_t FNV_OFF = 0xcbf29ce484222325ULL, FNV_PRIME = 0x100000001b3ULL;

uint64_t ora_hash_fields(const ora_dom *d) {
  static const int order[7] = {ORA_U, ORA_US, ORA_BI, ORA_BIS, ORA_FF, ORA_GG, ORA_EMF};
  uint64_t h = FNV_OFF;
  size_t per = (size_t)d->NT * d->NT;
  for (int q = 0; q < 7; q++) {
    const double *p = d->f[order[q]];
    size_t n = (size_t)d->lastActive * NVAR[order[q]] * per;
    for (size_t c = 0; c < n; c++) {
      uint64_t w;
      memcpy(&w, p + c, 8);
      h ^= w;
      h *= FNV_PRIME;
    }
  }
  return h;
}

uint64_t ora_hash_tree(const ora_dom *d) {
  uint64_t h = FNV_OFF;
#define MIX(v) do { h ^= (uint64_t)(int64_t)(v); h *= FNV_PRIME; } while (0)
  for (int nb = 0; nb < d->lastActive; nb++) {
    for (int q = 0; q < 8; q++) MIX(d->lp[nb][q]);
    MIX(d->icoord[nb][0]);
    MIX(d->icoord[nb][1]);
    MIX(d->leafs[nb] ? 1 : 0);
  }
#undef MIX
  return h;
}

int ora_get_field(const ora_dom *d, int field, int n, double *out) {
  if (field < 0 || field >= ORA_NFIELD || n > d->maxblocks) return ORA_ERR_CONFIG;
  memcpy(out, d->f[field], (size_t)n * NVAR[field] * d->NT * d->NT * sizeof(double));
  return 0;
}

int ora_set_field(ora_dom *d, int field, int n, const double *in) {
  if (field < 0 || field >= ORA_NFIELD || n > d->maxblocks) return ORA_ERR_CONFIG;
  memcpy(d->f[field], in, (size_t)n * NVAR[field] * d->NT * d->NT * sizeof(double));
  return 0;
}

int ora_get_tree(const ora_dom *d, int n, int32_t *lp8, int32_t *ic2, int32_t *lf) {
  for (int nb = 0; nb < n; nb++) {
    for (int q = 0; q < 8; q++) lp8[8 * nb + q] = d->lp[nb][q];
    ic2[2 * nb] = d->icoord[nb][0]; ic2[2 * nb + 1] = d->icoord[nb][1];
    lf[nb] = d->leafs[nb];
  }
  return 0;
}

int ora_set_tree(ora_dom *d, int n, const int32_t *lp8, const int32_t *ic2, const int32_t *lf) {
  if (n > d->maxblocks) return ORA_ERR_CONFIG;
  for (int nb = 0; nb < n; nb++) {
    for (int q = 0; q < 8; q++) d->lp[nb][q] = lp8[8 * nb + q];
    d->icoord[nb][0] = ic2[2 * nb]; d->icoord[nb][1] = ic2[2 * nb + 1];
    d->leafs[nb] = (unsigned char)lf[nb];
  }
  d->lastActive = n;
  return 0;
}

void ora_set_time(ora_dom *d, double t, double dom_dt, double dt_next, int count) {
  d->t = t; d->dt_dom = dom_dt; d->dt_main = dt_next; d->count = count;
}

void ora_last_error(const ora_dom *d, int32_t *o) { memcpy(o, d->err, sizeof d->err); }

/* -------------------------------------------------- block-level wrappers */
int ora_recon_tile(int NT, int ng, const char *lim, double dx, const double *W, double *xl, double *xr,
                   double *yl, double *yr) {
  int id = lim_id(lim);
  if (id == L_BAD) return ORA_ERR_CONFIG;
  size_t S = (size_t)NT * NT * 8;
  double *scr = calloc(6 * S, sizeof(double));
  int st = 0;
  if (id == L_WENO3 || id == L_WENO5) weno_tile(NT, ng, id, W, scr);
  else st = celledges(NULL, 0, NT, id, dx, W, scr);
  memcpy(xl, scr + 2 * S, S * 8); memcpy(xr, scr + 3 * S, S * 8);
  memcpy(yl, scr + 4 * S, S * 8); memcpy(yr, scr + 5 * S, S * 8);
  free(scr);
  return st;
}

int ora_muscl_tile(int NT, int ng, const char *lim, const char *flux, double gamma, double dx, int step,
                   const double *U, const double *Bi, double *ff, double *gg) {
  int lid = lim_id(lim), fid = flux_id(flux);
  if (lid == L_BAD || fid == F_BAD) return ORA_ERR_CONFIG;
  size_t S = (size_t)NT * NT * 8;
  double *scr = calloc(7 * S, sizeof(double));
  int st = muscl_tile(NULL, 0, NT, ng, lid, fid, gamma, dx, step, U, Bi, scr + 6 * S, ff, gg, scr);
  free(scr);
  return st;
}

int ora_regrid_tiles(int NT, int ng, int op, double gamma, double *U5, double *Bi5) {
  size_t su = (size_t)8 * NT * NT, sb = (size_t)2 * NT * NT;
  double *Us4[4], *Bs4[4];
  for (int q = 0; q < 4; q++) { Us4[q] = U5 + (q + 1) * su; Bs4[q] = Bi5 + (q + 1) * sb; }
  double *W = malloc(su * sizeof(double));
  int st = 0;
  if (op == 0) {
    refine_tiles(NT, ng, U5, NULL, Us4, NULL);
    for (int q = 0; q < 4 && !st; q++) st = cons2prim(NULL, NT, ng - 1, NT - ng, gamma, Us4[q], W, q + 1, 0);
    if (!st) st = refine_face_tiles(NT, ng, Bi5, Bs4);
  } else {
    coarsen_tiles(NT, ng, U5, Us4, 8);
    st = cons2prim(NULL, NT, ng - 1, NT - ng, gamma, U5, W, 0, 0);
    coarsen_face_tiles(NT, ng, Bi5, Bs4);
  }
  free(W);
  return st;
}
