// Reverse-mode (vector-Jacobian) counterparts of the pointwise pieces of the
// frozen-mesh step: cons->prim, minmod MUSCL reconstruction of one face, and
// the HLLE flux with its Roe-average wave bounds.  Each *_vjp function
// recomputes the forward values it needs and accumulates the adjoint of its
// inputs from the adjoint of its outputs.  The forward expressions are those
// of physics.hpp (the reference's operand order); the derivatives follow the
// computed expression chain, including the branch taken at every min/max and
// the 1e-15 guards, so the result is the exact derivative of the discrete
// map away from branch switches.
#pragma once
#include "physics.hpp"

namespace amr {

// Hides x from the optimizer, so that values recomputed from it are not
// merged with identical ones computed earlier and kept live in the
// meantime (a register-pressure device; no effect in the host build).
__host__ __device__ inline void opaque(double& x) {
#ifdef __HIP_DEVICE_COMPILE__
  asm volatile("" : "+v"(x));
#endif
}

// ------------------------------------------------------------ cons -> prim
__host__ __device__ inline void prim_of(const double* u, double g, double* w) {
  w[0] = u[0];
  w[1] = u[1] / u[0];
  w[2] = u[2] / u[0];
  w[3] = u[3] / u[0];
  const double p = (g - 1) * (u[4] - 0.5 * w[0] * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]));
  w[5] = u[5];
  w[6] = u[6];
  w[7] = u[7];
  w[4] = p - (g - 1) * (0.5 * (u[5] * u[5] + u[6] * u[6] + u[7] * u[7]));
}

// ub += (dw/du)^T wb, from the primitive state w = prim_of(u) (u0 = w0 and
// u5..7 = w5..7; w1..3 = u1..3 / u0)
__host__ __device__ inline void prim_vjp_w(const double* w, double g, const double* wb, double* ub) {
  const double iw0 = 1. / w[0];
  const double pb = wb[4];
  const double sv = w[1] * w[1] + w[2] * w[2] + w[3] * w[3];
  double w0b = wb[0] + pb * (-(g - 1) * 0.5 * sv);
  ub[4] += pb * (g - 1);
#pragma unroll
  for (int k = 1; k <= 3; k++) {
    const double wkb = wb[k] - pb * (g - 1) * w[0] * w[k];
    ub[k] += wkb * iw0;
    w0b -= wkb * w[k] * iw0;
  }
  ub[0] += w0b;
#pragma unroll
  for (int k = 5; k <= 7; k++) ub[k] += wb[k] - pb * (g - 1) * w[k];
}

__host__ __device__ inline void prim_vjp(const double* u, double g, const double* wb, double* ub) {
  double w[8];
  prim_of(u, g, w);
  prim_vjp_w(w, g, wb, ub);
}

// -------------------------------------------------- minmod MUSCL, one face
// Edge states of the face between cells c (left) and c+1 of a 4-cell
// stencil v = (c-1, c, c+1, c+2): muscl_plus/muscl_minus with minmod.
__host__ __device__ inline void recon_minmod(const double* v, double& wl, double& wr) {
  const double a = v[1] - v[0], b = v[2] - v[1], d = v[3] - v[2];
  wl = muscl_plus(kLimMinmod, v[1], a, b, 1.);
  wr = muscl_minus(kLimMinmod, v[2], b, d, 1.);
}

// derivative of max(0, min(1, r)) (mac_max(0, mac_min(1, r))) along the
// branch the forward takes: the value is r itself iff !(1 < r) and !(0 > r)
__host__ __device__ inline bool minmod_slope(double r) { return !(1. < r) && !(0. > r); }

// vb[0..3] += d(wl, wr)/dv ^T (wlb, wrb).  wl = v1 + 0.5 phi(rl) a with
// rl = b / (a + eps): where phi = rl, d wl/da = 0.5 rl eps / (a + eps) and
// d wl/db = 0.5 a / (a + eps); elsewhere phi is 0 or 1 and d wl/da = 0.5 phi.
// wr = v2 - 0.5 phi(rr) d with rr = b / (d + eps) likewise.  rl and rr are
// the forward's own quotients, so the branch is the one the forward took.
__host__ __device__ inline void recon_minmod_vjp(const double* v, double wlb, double wrb, double* vb) {
  const double a = v[1] - v[0], b = v[2] - v[1], d = v[3] - v[2];
  const double rl = b / (a + kEps), rr = b / (d + kEps);
  double ab, bb, db;
  if (minmod_slope(rl)) {
    const double il = 1. / (a + kEps);
    ab = 0.5 * wlb * rl * kEps * il;
    bb = 0.5 * wlb * a * il;
  } else {
    ab = 0.5 * limiter(kLimMinmod, rl, 0.) * wlb;
    bb = 0.;
  }
  if (minmod_slope(rr)) {
    const double ir = 1. / (d + kEps);
    db = -0.5 * wrb * rr * kEps * ir;
    bb -= 0.5 * wrb * d * ir;
  } else {
    db = -0.5 * limiter(kLimMinmod, rr, 0.) * wrb;
  }
  vb[0] -= ab;
  vb[1] += wlb + ab - bb;
  vb[2] += wrb + bb - db;
  vb[3] += db;
}

// ------------------------------------------------------------- HLLE flux
// The reverse keeps the forward's intermediates: fast_fwd / side_fwd /
// roe_fwd return the values of fast_speed(w, g, 5) / side_state /
// roe_bounds bit for bit (same expressions in the same order) together with
// what their reverse needs; the reverse multiplies by reciprocals formed
// once instead of dividing again.

// fast magnetosonic speed along x
struct FastT {
  double ca2, a2, can2, s1, sq, cf;
};

__host__ __device__ inline double fast_fwd(const double* w, double g, FastT& t) {
  const double ca = sqrt((w[5] * w[5] + w[6] * w[6] + w[7] * w[7]) / w[0]);
  const double a = sqrt(g * w[4] / w[0]);
  const double can = sqrt((w[5] * w[5]) / w[0]);
  t.ca2 = ca * ca;
  t.a2 = a * a;
  t.can2 = can * can;
  const double s0 = t.ca2 + t.a2;
  t.s1 = t.ca2 - t.a2;
  t.sq = sqrt(t.s1 * t.s1 + 4. * t.a2 * (t.ca2 - t.can2));
  t.cf = sqrt(0.5 * (s0 + t.sq));
  return t.cf;
}

// cfb -> wb; each of ca^2, a^2, can^2 is sqrt(q)^2 of a quotient q over w0,
// with derivative 1 in q wherever the root is positive
__host__ __device__ inline void fast_bwd(const double* w, double g, const FastT& t, double iw0, double cfb,
                                         double* wb) {
  const double tb = cfb * 0.5 / t.cf;
  const double s0b = 0.5 * tb, sqb = 0.5 * tb;
  const double innerb = t.sq > 0. ? sqb * 0.5 / t.sq : 0.;
  const double s1b = innerb * 2. * t.s1;
  const double a2b = innerb * 4. * (t.ca2 - t.can2) + s0b - s1b;
  const double ca2b = innerb * 4. * t.a2 + s0b + s1b;
  const double can2b = -innerb * 4. * t.a2;
  if (t.ca2 > 0.) {
    const double B2b = ca2b * iw0;
    wb[0] -= ca2b * t.ca2 * iw0;
    wb[5] += B2b * 2. * w[5];
    wb[6] += B2b * 2. * w[6];
    wb[7] += B2b * 2. * w[7];
  }
  if (t.a2 > 0.) {
    wb[4] += a2b * g * iw0;
    wb[0] -= a2b * t.a2 * iw0;
  }
  if (t.can2 > 0.) {
    wb[5] += can2b * 2. * w[5] * iw0;
    wb[0] -= can2b * t.can2 * iw0;
  }
}

// the fast speed, energy density and enthalpy of one side (side_state's
// cf, e and H)
__host__ __device__ inline void side_fwd(const double* w, double g, FastT& f, double& e, double& H) {
  fast_fwd(w, g, f);
  const double b2 = w[5] * w[5] + w[6] * w[6] + w[7] * w[7];
  e = (w[4] / (w[0] * (g - 1))) + 0.5 * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]) + 0.5 * b2 / w[0];
  e = e * w[0];
  H = (e + w[4] + 0.5 * b2) / w[0];
}

// component k of side_state's conserved state q and x-flux F
__host__ __device__ inline void side_qF(const double* w, double e, double H, int k, double& q, double& F) {
  const double vn = w[1], Bn = w[5];
  const double mv = w[0] * vn;
  switch (k) {
    case 0: q = w[0]; F = mv; break;
    case 1: q = w[0] * w[1]; F = mv * w[1] + (w[4] + 0.5 * (w[5] * w[5] + w[6] * w[6] + w[7] * w[7])) - Bn * w[5]; break;
    case 2: q = w[0] * w[2]; F = mv * w[2] - Bn * w[6]; break;
    case 3: q = w[0] * w[3]; F = mv * w[3] - Bn * w[7]; break;
    case 4: q = e; F = mv * H - Bn * (w[1] * w[5] + w[2] * w[6] + w[3] * w[7]); break;
    case 5: q = w[5]; F = vn * w[5] - Bn * w[1]; break;
    case 6: q = w[6]; F = vn * w[6] - Bn * w[2]; break;
    default: q = w[7]; F = vn * w[7] - Bn * w[3]; break;
  }
}

// side reverse: the adjoints of F and q are sF Nb and sQ Nb, of H Hb and of
// the fast speed cfb; accumulates into wb
__host__ __device__ inline void side_bwd(const double* w, double g, double e, double H, const FastT& f,
                                         const double* Nb, double sF, double sQ, double Hb, double cfb, double* wb) {
  const double iw0 = 1. / w[0];
  const double vn = w[1], Bn = w[5];
  const double b2 = w[5] * w[5] + w[6] * w[6] + w[7] * w[7];
  const double mv = w[0] * vn;
  const double vb_dot = w[1] * w[5] + w[2] * w[6] + w[3] * w[7];
  double Fb[8], qb[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    Fb[k] = sF * Nb[k];
    qb[k] = sQ * Nb[k];
  }
  double mvb = Fb[0] + Fb[1] * w[1] + Fb[2] * w[2] + Fb[3] * w[3] + Fb[4] * H;
  double vnb = Fb[5] * w[5] + Fb[6] * w[6] + Fb[7] * w[7];
  double Bnb = -(Fb[1] * w[5] + Fb[2] * w[6] + Fb[3] * w[7] + Fb[4] * vb_dot + Fb[5] * w[1] + Fb[6] * w[2] +
                 Fb[7] * w[3]);
  const double ptb = Fb[1];
  Hb += Fb[4] * mv;
  const double vdb = -Fb[4] * Bn;
  wb[1] += Fb[1] * mv + vdb * w[5] - Fb[5] * Bn;
  wb[2] += Fb[2] * mv + vdb * w[6] - Fb[6] * Bn;
  wb[3] += Fb[3] * mv + vdb * w[7] - Fb[7] * Bn;
  wb[5] += -Fb[1] * Bn + vdb * w[1] + Fb[5] * vn;
  wb[6] += -Fb[2] * Bn + vdb * w[2] + Fb[6] * vn;
  wb[7] += -Fb[3] * Bn + vdb * w[3] + Fb[7] * vn;
  // mv = w0 vn; pt = w4 + 0.5 b2
  wb[0] += mvb * vn;
  vnb += mvb * w[0];
  wb[4] += ptb;
  double b2b = 0.5 * ptb;
  // q: q0 = w0, q1..3 = w0 w1..3, q4 = e, q5..7 = w5..7
  wb[0] += qb[0] + qb[1] * w[1] + qb[2] * w[2] + qb[3] * w[3];
  wb[1] += qb[1] * w[0];
  wb[2] += qb[2] * w[0];
  wb[3] += qb[3] * w[0];
  double eb = qb[4];
  wb[5] += qb[5];
  wb[6] += qb[6];
  wb[7] += qb[7];
  // H = (e + w4 + 0.5 b2) / w0
  eb += Hb * iw0;
  wb[4] += Hb * iw0;
  b2b += 0.5 * Hb * iw0;
  wb[0] -= Hb * H * iw0;
  // e = e1 w0 with e1 = w4 / (w0 (g-1)) + 0.5 |v|^2 + 0.5 b2 / w0
  const double e1b = eb * w[0];
  wb[0] += eb * (e * iw0);
  wb[4] += e1b * iw0 / (g - 1);
  wb[0] -= e1b * (w[4] * iw0 / (g - 1)) * iw0;
  wb[1] += e1b * w[1];
  wb[2] += e1b * w[2];
  wb[3] += e1b * w[3];
  b2b += e1b * 0.5 * iw0;
  wb[0] -= e1b * (0.5 * b2 * iw0) * iw0;
  // b2 = w5^2 + w6^2 + w7^2
  wb[5] += b2b * 2. * w[5];
  wb[6] += b2b * 2. * w[6];
  wb[7] += b2b * 2. * w[7];
  wb[1] += vnb;
  wb[5] += Bnb;
  fast_bwd(w, g, f, iw0, cfb, wb);
}

// side reverse with q and F adjoints sQ Nb and sF Nb, plus adjoints of e
// and pt used outside q and F (HLLD star states), H and the fast speed
__host__ __device__ inline void side_bwd_vec(const double* w, double g, double e, double H, const FastT& f,
                                             const double* Nb, double sF, double sQ, double Hb, double eb_x,
                                             double ptb_x, double cfb, double* wb) {
  double Fb[8], qb[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    Fb[k] = sF * Nb[k];
    qb[k] = sQ * Nb[k];
  }
  const double iw0 = 1. / w[0];
  const double vn = w[1], Bn = w[5];
  const double b2 = w[5] * w[5] + w[6] * w[6] + w[7] * w[7];
  const double mv = w[0] * vn;
  const double vb_dot = w[1] * w[5] + w[2] * w[6] + w[3] * w[7];
  double mvb = Fb[0] + Fb[1] * w[1] + Fb[2] * w[2] + Fb[3] * w[3] + Fb[4] * H;
  double vnb = Fb[5] * w[5] + Fb[6] * w[6] + Fb[7] * w[7];
  const double Bnb = -(Fb[1] * w[5] + Fb[2] * w[6] + Fb[3] * w[7] + Fb[4] * vb_dot + Fb[5] * w[1] + Fb[6] * w[2] +
                       Fb[7] * w[3]);
  const double ptb = Fb[1] + ptb_x;
  Hb += Fb[4] * mv;
  const double vdb = -Fb[4] * Bn;
  wb[1] += Fb[1] * mv + vdb * w[5] - Fb[5] * Bn;
  wb[2] += Fb[2] * mv + vdb * w[6] - Fb[6] * Bn;
  wb[3] += Fb[3] * mv + vdb * w[7] - Fb[7] * Bn;
  wb[5] += -Fb[1] * Bn + vdb * w[1] + Fb[5] * vn;
  wb[6] += -Fb[2] * Bn + vdb * w[2] + Fb[6] * vn;
  wb[7] += -Fb[3] * Bn + vdb * w[3] + Fb[7] * vn;
  wb[0] += mvb * vn;
  vnb += mvb * w[0];
  wb[4] += ptb;
  double b2b = 0.5 * ptb;
  wb[0] += qb[0] + qb[1] * w[1] + qb[2] * w[2] + qb[3] * w[3];
  wb[1] += qb[1] * w[0];
  wb[2] += qb[2] * w[0];
  wb[3] += qb[3] * w[0];
  double eb = qb[4] + eb_x;
  wb[5] += qb[5];
  wb[6] += qb[6];
  wb[7] += qb[7];
  eb += Hb * iw0;
  wb[4] += Hb * iw0;
  b2b += 0.5 * Hb * iw0;
  wb[0] -= Hb * H * iw0;
  const double e1b = eb * w[0];
  wb[0] += eb * (e * iw0);
  wb[4] += e1b * iw0 / (g - 1);
  wb[0] -= e1b * (w[4] * iw0 / (g - 1)) * iw0;
  wb[1] += e1b * w[1];
  wb[2] += e1b * w[2];
  wb[3] += e1b * w[3];
  b2b += e1b * 0.5 * iw0;
  wb[0] -= e1b * (0.5 * b2 * iw0) * iw0;
  wb[5] += b2b * 2. * w[5];
  wb[6] += b2b * 2. * w[6];
  wb[7] += b2b * 2. * w[7];
  wb[1] += vnb;
  wb[5] += Bnb;
  fast_bwd(w, g, f, iw0, cfb, wb);
}

// Roe-average bounds (roe_bounds) with the intermediates of the reverse
struct RoeT {
  double sl, sr, nrm, v[3], B[3], RT, d, d6, d7, X, Y, h, btsq, cY, vax2, a2raw, a2, ct2, tdif, sq, cfast;
};

__host__ __device__ inline void roe_fwd(const double* wL, const double* wR, double g, double HL, double HR, RoeT& t,
                                        double& lmin, double& lmax) {
  t.sl = sqrt(wL[0]);
  t.sr = sqrt(wR[0]);
  t.nrm = sqrt(wL[0] + t.sr);
#pragma unroll
  for (int i = 0; i < 3; i++) {
    t.v[i] = (t.sl * wL[i + 1] + t.sr * wR[i + 1]) / t.nrm;
    t.B[i] = (t.sl * wL[i + 5] + t.sr * wR[i + 5]) / t.nrm;
  }
  t.RT = sqrt(wR[0] / wL[0]);
  t.d = sqrt(wR[0] * wL[0]);
  t.d6 = wL[6] - wR[6];
  t.d7 = wL[7] - wR[7];
  t.X = 0.5 * ((t.d6 * t.d6) + (t.d7 * t.d7)) / (t.sl + t.sr);
  t.Y = 0.5 * (wL[0] + wR[0]) / t.d;
  t.h = (HL + t.RT * HR) / (1 + t.RT);
  const double vsq = t.v[0] * t.v[0] + t.v[1] * t.v[1] + t.v[2] * t.v[2];
  t.btsq = (t.B[1] * t.B[1]) + (t.B[2] * t.B[2]);
  t.cY = (g - 1. - (g - 2) * t.Y);
  const double bt2s = t.cY * t.btsq;
  t.vax2 = (t.B[0] * t.B[0]) / t.d;
  const double hp = t.h - (t.vax2 + t.btsq / t.d);
  t.a2raw = ((g - 1.) * (hp - 0.5 * vsq) - (g - 2.) * t.X);
  t.a2 = mac_max(t.a2raw, kEps);
  t.ct2 = bt2s / t.d;
  const double tsum = t.vax2 + t.ct2 + t.a2;
  t.tdif = t.vax2 + t.ct2 - t.a2;
  t.sq = sqrt(t.tdif * t.tdif + 4. * t.a2 * t.ct2);
  t.cfast = sqrt(0.5 * (tsum + t.sq));
  lmin = t.v[0] - t.cfast;
  lmax = t.v[0] + t.cfast;
}

// (lminb, lmaxb) -> wL, wR, HL, HR adjoints (accumulated)
__host__ __device__ inline void roe_bwd(const double* wL, const double* wR, double g, double HR, const RoeT& t,
                                        double lminb, double lmaxb, double* wLb, double* wRb, double& HLb,
                                        double& HRb) {
  const double id = 1. / t.d, inrm = 1. / t.nrm;
  double vb[3] = {lminb + lmaxb, 0., 0.}, Bb[3] = {0., 0., 0.};
  const double cfb = lmaxb - lminb;
  const double tb = cfb * 0.5 / t.cfast;
  const double tsumb = 0.5 * tb, sqb = 0.5 * tb;
  const double inb = t.sq > 0. ? sqb * 0.5 / t.sq : 0.;
  const double tdifb = inb * 2. * t.tdif;
  const double a2b = inb * 4. * t.ct2 + tsumb - tdifb;
  const double ct2b = inb * 4. * t.a2 + tsumb + tdifb;
  double vax2b = tsumb + tdifb;
  // ct2 = bt2s / d
  const double bt2sb = ct2b * id;
  double db = -ct2b * t.ct2 * id;
  // a2 = max(a2raw, eps)
  const double a2rb = (t.a2raw > kEps) ? a2b : 0.;
  const double hpb = a2rb * (g - 1.);
  const double vsqb = -a2rb * (g - 1.) * 0.5;
  const double Xb = -a2rb * (g - 2.);
  // hp = h - (vax2 + btsq / d)
  const double hb = hpb;
  vax2b -= hpb;
  double btsqb = -hpb * id;
  db += hpb * (t.btsq * id) * id;
  // vax2 = B0^2 / d
  Bb[0] += vax2b * 2. * t.B[0] * id;
  db -= vax2b * t.vax2 * id;
  // bt2s = cY btsq
  btsqb += bt2sb * t.cY;
  const double Yb = -bt2sb * t.btsq * (g - 2);
  Bb[1] += btsqb * 2. * t.B[1];
  Bb[2] += btsqb * 2. * t.B[2];
#pragma unroll
  for (int i = 0; i < 3; i++) vb[i] += vsqb * 2. * t.v[i];
  // h = (HL + RT HR) / (1 + RT)
  const double i1 = 1. / (1 + t.RT);
  HLb += hb * i1;
  HRb += hb * t.RT * i1;
  const double RTb = (hb * HR - hb * t.h) * i1;
  // Y = 0.5 (rhoL + rhoR) / d
  wLb[0] += Yb * 0.5 * id;
  wRb[0] += Yb * 0.5 * id;
  db -= Yb * t.Y * id;
  // X = 0.5 (d6^2 + d7^2) / (sl + sr)
  const double iss = 1. / (t.sl + t.sr);
  const double d6b = Xb * t.d6 * iss, d7b = Xb * t.d7 * iss;
  double slb = -Xb * t.X * iss, srb = slb;
  wLb[6] += d6b;
  wRb[6] -= d6b;
  wLb[7] += d7b;
  wRb[7] -= d7b;
  // d = sqrt(rhoR rhoL)
  if (t.d > 0.) {
    const double pb = db * 0.5 * id;
    wRb[0] += pb * wL[0];
    wLb[0] += pb * wR[0];
  }
  // RT = sqrt(rhoR / rhoL)
  if (t.RT > 0.) {
    const double qb = RTb * 0.5 / t.RT;
    const double iL = 1. / wL[0];
    wRb[0] += qb * iL;
    wLb[0] -= qb * (wR[0] * iL) * iL;
  }
  // v_i, B_i = (sl a_L + sr a_R) / nrm
  double nrmb = 0.;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    slb += (vb[i] * wL[i + 1] + Bb[i] * wL[i + 5]) * inrm;
    srb += (vb[i] * wR[i + 1] + Bb[i] * wR[i + 5]) * inrm;
    wLb[i + 1] += vb[i] * t.sl * inrm;
    wRb[i + 1] += vb[i] * t.sr * inrm;
    wLb[i + 5] += Bb[i] * t.sl * inrm;
    wRb[i + 5] += Bb[i] * t.sr * inrm;
    nrmb -= (vb[i] * t.v[i] + Bb[i] * t.B[i]) * inrm;
  }
  // nrm = sqrt(rhoL + sr), sl = sqrt(rhoL), sr = sqrt(rhoR)
  const double inner_b = nrmb * 0.5 * inrm;
  wLb[0] += inner_b + slb * 0.5 / t.sl;
  wRb[0] += (inner_b + srb) * 0.5 / t.sr;
}

// HLLE flux (flux_hlle) reverse: fb -> wLb, wRb (accumulated)
__host__ __device__ inline void hlle_vjp(const double* wL, const double* wR, double g, const double* fb, double* wLb,
                                         double* wRb) {
  FastT fL, fR;
  double eL, HL, eR, HR;
  side_fwd(wL, g, fL, eL, HL);
  side_fwd(wR, g, fR, eR, HR);
  RoeT rt;
  double lmin, lmax;
  roe_fwd(wL, wR, g, HL, HR, rt, lmin, lmax);
  const double xl = wL[1] - fL.cf, il = std_min(xl, lmin);
  const double SL = mac_min(il, 0.);
  const double xr = wR[1] + fR.cf, ir = std_max(xr, lmax);
  const double SR = mac_max(ir, 0.);
  const double iD = 1. / (SR - SL);
  // f_k = (SR FL_k - SL FR_k + SL SR dq_k) / D
  double SLb = 0., SRb = 0., Db = 0., Nb[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    double qL, FL, qR, FR;
    side_qF(wL, eL, HL, k, qL, FL);
    side_qF(wR, eR, HR, k, qR, FR);
    const double dq = qR - qL;
    const double f = (SR * FL - SL * FR + SL * SR * dq) * iD;
    Nb[k] = fb[k] * iD;
    Db -= Nb[k] * f;
    SRb += Nb[k] * (FL + SL * dq);
    SLb += Nb[k] * (-FR + SR * dq);
  }
  SRb += Db;
  SLb -= Db;
  double cfLb = 0., cfRb = 0., lminb = 0., lmaxb = 0.;
  if (il < 0.) {  // SL = il
    if (lmin < xl) lminb += SLb;
    else {
      wLb[1] += SLb;
      cfLb -= SLb;
    }
  }
  if (ir > 0.) {  // SR = ir
    if (xr < lmax) lmaxb += SRb;
    else {
      wRb[1] += SRb;
      cfRb += SRb;
    }
  }
  double HLb = 0., HRb = 0.;
  roe_bwd(wL, wR, g, HR, rt, lminb, lmaxb, wLb, wRb, HLb, HRb);
  // FL_b = SR Nb, qL_b = -SL SR Nb; FR_b = -SL Nb, qR_b = SL SR Nb
  side_bwd(wL, g, eL, HL, fL, Nb, SR, -SL * SR, HLb, cfLb, wLb);
  side_bwd(wR, g, eR, HR, fR, Nb, -SL, SL * SR, HRb, cfRb, wRb);
}

// ------------------------------------------------------------- HLLD flux
// star state (hlld_star) reverse: (wsb, esb) -> adjoints of the side's w, S,
// SM, pts, e and pt (accumulated); ws/es are the forward's outputs
__host__ __device__ inline void hlld_star_bwd(const double* w, double S, double SM, double pts, double e, double pt,
                                              const double* ws, double es, double* wsb, double esb, double* wb,
                                              double& Sb, double& SMb, double& ptsb, double& eb, double& ptb) {
  const double a = S - w[1], b = S - SM;
  double ab = 0., bb = 0.;
  // es = (a e - pt w1 + pts SM + w5 (vdotb(w) - vdotb(ws))) / b
  const double Nb = esb / b;
  bb -= esb * es / b;
  ab += Nb * e;
  eb += Nb * a;
  ptb -= Nb * w[1];
  wb[1] -= Nb * pt;
  ptsb += Nb * SM;
  SMb += Nb * pts;
  wb[5] += Nb * (vdotb(w) - vdotb(ws));
  const double vw = Nb * w[5], vs = -Nb * w[5];
  wb[1] += vw * w[5];
  wb[5] += vw * w[1];
  wb[2] += vw * w[6];
  wb[6] += vw * w[2];
  wb[3] += vw * w[7];
  wb[7] += vw * w[3];
  wsb[1] += vs * ws[5];
  wsb[5] += vs * ws[1];
  wsb[2] += vs * ws[6];
  wsb[6] += vs * ws[2];
  wsb[3] += vs * ws[7];
  wsb[7] += vs * ws[3];
  const double den = w[0] * (S - w[1]) * (S - SM) - (w[5] * w[5]);
  if (den != 0) {
    const double sv = (SM - w[1]) / den;
    const double sb = (w[0] * ((S - w[1]) * (S - w[1])) - (w[5] * w[5])) / den;
    double denb = 0., svb = 0., sbb = 0.;
    wb[7] += wsb[7] * sb;
    sbb += wsb[7] * w[7];
    wb[6] += wsb[6] * sb;
    sbb += wsb[6] * w[6];
    const double Nsb = sbb / den;
    denb -= sbb * sb / den;
    wb[0] += Nsb * (a * a);
    ab += Nsb * w[0] * 2. * a;
    wb[5] -= Nsb * 2. * w[5];
    wb[3] += wsb[3];
    wb[5] -= wsb[3] * w[7] * sv;
    wb[7] -= wsb[3] * w[5] * sv;
    svb -= wsb[3] * w[5] * w[7];
    wb[2] += wsb[2];
    wb[5] -= wsb[2] * w[6] * sv;
    wb[6] -= wsb[2] * w[5] * sv;
    svb -= wsb[2] * w[5] * w[6];
    SMb += svb / den;
    wb[1] -= svb / den;
    denb -= svb * sv / den;
    wb[0] += denb * a * b;
    ab += denb * w[0] * b;
    bb += denb * w[0] * a;
    wb[5] -= denb * 2. * w[5];
  } else {
    wb[2] += wsb[2];
    wb[3] += wsb[3];
  }
  wb[5] += wsb[5];
  SMb += wsb[1];
  // ws0 = w0 a / b
  wb[0] += wsb[0] * a / b;
  ab += wsb[0] * w[0] / b;
  bb -= wsb[0] * ws[0] / b;
  Sb += ab + bb;
  wb[1] -= ab;
  SMb -= bb;
}

// star_cons reverse: ub -> (wb, eb) of the star state
__host__ __device__ inline void star_cons_bwd(const double* w, const double* ub, double* wb, double& eb) {
  wb[0] += ub[0] + ub[1] * w[1] + ub[2] * w[2] + ub[3] * w[3];
  wb[1] += ub[1] * w[0];
  wb[2] += ub[2] * w[0];
  wb[3] += ub[3] * w[0];
  eb += ub[4];
  wb[5] += ub[5];
  wb[6] += ub[6];
  wb[7] += ub[7];
}

// star_cons reverse with ub = sc Nb
__host__ __device__ inline void star_cons_bwd_s(const double* w, const double* Nb, double sc, double* wb,
                                                double& eb) {
  double ub[8];
#pragma unroll
  for (int k = 0; k < 8; k++) ub[k] = sc * Nb[k];
  star_cons_bwd(w, ub, wb, eb);
}

// component k of star_cons(w, e)
__host__ __device__ inline double star_cons_k(const double* w, double e, int k) {
  switch (k) {
    case 0: return w[0];
    case 1: return w[0] * w[1];
    case 2: return w[0] * w[2];
    case 3: return w[0] * w[3];
    case 4: return e;
    default: return w[k];
  }
}

// Reverse of the part of the HLLD flux every subsonic region shares: SM and
// p* over den, SL = min(vL, vR) - cmax, SR = max(vL, vR) + cmax with
// cmax = max(cfL, cfR), and both sides' state (the active side with its
// F adjoint fb and q adjoint sq fb; the other with none).  The sides' fast
// speeds, energies, enthalpies and total pressures are recomputed here
// from opaque copies of the states (the same expressions, so the same
// values) rather than held in registers across the star-state reverse.
struct HlldSides {
  double eL, HL, ptL, eR, HR, ptR;
  FastT fL, fR;
};
__host__ __device__ inline void hlld_sides(const double* wL, const double* wR, double g, HlldSides& s) {
  side_fwd(wL, g, s.fL, s.eL, s.HL);
  side_fwd(wR, g, s.fR, s.eR, s.HR);
  s.ptL = wL[4] + 0.5 * (wL[5] * wL[5] + wL[6] * wL[6] + wL[7] * wL[7]);
  s.ptR = wR[4] + 0.5 * (wR[5] * wR[5] + wR[6] * wR[6] + wR[7] * wR[7]);
}
__host__ __device__ inline void hlld_tail_bwd(const double* wL_, const double* wR_, double g, double SL, double SR,
                                              double SM, double pts, double den, bool left, const double* fb,
                                              double sq, double SLb, double SRb, double SMb, double ptsb, double eLb,
                                              double eRb, double ptLb, double ptRb, double* wLb, double* wRb) {
  double wL[8], wR[8];
  for (int k = 0; k < 8; k++) {
    wL[k] = wL_[k];
    wR[k] = wR_[k];
    opaque(wL[k]);
    opaque(wR[k]);
  }
  HlldSides s;
  hlld_sides(wL, wR, g, s);
  const double vL = wL[1], vR = wR[1];
  const double dR = SR - vR, dL = SL - vL;
  double dRb = 0., dLb = 0., denb = 0., vLb = 0., vRb = 0., wL0b = 0., wR0b = 0.;
  {
    const double Ns = SMb / den, Np = ptsb / den;
    denb -= SMb * SM / den + ptsb * pts / den;
    dRb += Ns * wR[0] * vR;
    wR0b += Ns * dR * vR;
    vRb += Ns * dR * wR[0];
    dLb -= Ns * wL[0] * vL;
    wL0b -= Ns * dL * vL;
    vLb -= Ns * dL * wL[0];
    ptRb -= Ns;
    ptLb += Ns;
    dRb += Np * wR[0] * s.ptL;
    wR0b += Np * dR * s.ptL;
    ptLb += Np * dR * wR[0];
    dLb -= Np * wL[0] * s.ptR;
    wL0b -= Np * dL * s.ptR;
    ptRb -= Np * dL * wL[0];
    const double dv = vR - vL, t = wL[0] * wR[0] * dR * dL;
    wL0b += Np * wR[0] * dR * dL * dv;
    wR0b += Np * wL[0] * dR * dL * dv;
    dRb += Np * wL[0] * wR[0] * dL * dv;
    dLb += Np * wL[0] * wR[0] * dR * dv;
    vRb += Np * t;
    vLb -= Np * t;
  }
  dRb += denb * wR[0];
  wR0b += denb * dR;
  dLb -= denb * wL[0];
  wL0b -= denb * dL;
  SRb += dRb;
  vRb -= dRb;
  SLb += dLb;
  vLb -= dLb;
  const double cmaxb = -SLb + SRb;
  if (vR < vL) vRb += SLb;
  else vLb += SLb;
  if (vL < vR) vRb += SRb;
  else vLb += SRb;
  double cfLb = 0., cfRb = 0.;
  if (s.fL.cf < s.fR.cf) cfRb += cmaxb;
  else cfLb += cmaxb;
  wLb[0] += wL0b;
  wRb[0] += wR0b;
  wLb[1] += vLb;
  wRb[1] += vRb;
  side_bwd_vec(wL, g, s.eL, s.HL, s.fL, fb, left ? 1. : 0., left ? sq : 0., 0., eLb, ptLb, cfLb, wLb);
  side_bwd_vec(wR, g, s.eR, s.HR, s.fR, fb, left ? 0. : 1., left ? 0. : sq, 0., eRb, ptRb, cfRb, wRb);
}

// HLLD flux (flux_hlld, Miyoshi & Kusano 2005 with the reference's branch
// conditions) reverse: fb -> wLb, wRb (accumulated).  The forward values are
// recomputed with the forward's own expressions; f receives the flux.  The
// region is decided first, from the wave speeds and the star densities, and
// only that region's states are formed and reversed: the supersonic regions
// (1, 6) reverse one side's flux, regions 2 and 5 one star state, regions 3
// and 4 both star states and the double-star state.  Star conserved vectors
// are formed per component where they are used.
__host__ __device__ inline void hlld_vjp(const double* wL, const double* wR, double g, const double* fb, double* wLb,
                                         double* wRb, double* f = nullptr) {
  HlldSides s;
  hlld_sides(wL, wR, g, s);
  const double vL = wL[1], vR = wR[1];
  const double cmax = std_max(s.fL.cf, s.fR.cf);
  const double SL = std_min(vL, vR) - cmax;
  const double SR = std_max(vL, vR) + cmax;
  const double den = (SR - vR) * wR[0] - (SL - vL) * wL[0];
  const double SM = ((SR - vR) * wR[0] * vR - (SL - vL) * wL[0] * vL - s.ptR + s.ptL) / den;
  const double pts = ((SR - vR) * wR[0] * s.ptL - (SL - vL) * wL[0] * s.ptR +
                      wL[0] * wR[0] * (SR - vR) * (SL - vL) * (vR - vL)) / den;
  // star densities (hlld_star's ws[0]) and the Alfven speeds they give
  const double rl = sqrt(wL[0] * (SL - wL[1]) / (SL - SM)), rr = sqrt(wR[0] * (SR - wR[1]) / (SR - SM));
  const double SLs = SM - fabs(wL[5]) / rl;
  const double SRs = SM + fabs(wR[5]) / rr;
  int region = 0;
  if (SL > 0) region = 1;
  else if (SL <= 0. && SLs >= 0.) region = 2;
  else if (SLs <= 0. && SM >= 0.) region = 3;
  else if (SM <= 0. && SRs >= 0.) region = 4;
  else if (SRs <= 0. && SR >= 0.) region = 5;
  else if (SR < 0.) region = 6;

  if (region == 0) {  // no region holds (NaN state): the forward's flux is 0
    if (f)
      for (int k = 0; k < 8; k++) f[k] = 0.;
    return;
  }
  const bool left = region <= 3;
  const double* w = left ? wL : wR;
  double* wb = left ? wLb : wRb;
  const double ee = left ? s.eL : s.eR, HH = left ? s.HL : s.HR;
  if (region == 1 || region == 6) {  // F of one side
    if (f)
      for (int k = 0; k < 8; k++) {
        double q, F;
        side_qF(w, ee, HH, k, q, F);
        f[k] = F;
      }
    side_bwd_vec(w, g, ee, HH, left ? s.fL : s.fR, fb, 1., 0., 0., 0., 0., 0., wb);
    return;
  }
  const double S = left ? SL : SR;
  double SLb = 0., SRb = 0., SMb = 0., ptsb = 0., eLb = 0., eRb = 0., ptLb = 0., ptRb = 0.;
  if (region == 2 || region == 5) {
    // f = F + S (q* - q) of the active side
    double ws[8], es;
    hlld_star(w, S, SM, pts, ee, left ? s.ptL : s.ptR, ws, es);
    double Sb = 0.;
    for (int k = 0; k < 8; k++) {
      double q, F;
      side_qF(w, ee, HH, k, q, F);
      const double qs = star_cons_k(ws, es, k);
      if (f) f[k] = F + S * qs - S * q;
      Sb += fb[k] * (qs - q);
    }
    double wsb[8] = {0., 0., 0., 0., 0., 0., 0., 0.}, esb = 0.;
    star_cons_bwd_s(ws, fb, S, wsb, esb);
    if (left) {
      hlld_star_bwd(wL, SL, SM, pts, s.eL, s.ptL, ws, es, wsb, esb, wLb, Sb, SMb, ptsb, eLb, ptLb);
      SLb = Sb;
    } else {
      hlld_star_bwd(wR, SR, SM, pts, s.eR, s.ptR, ws, es, wsb, esb, wRb, Sb, SMb, ptsb, eRb, ptRb);
      SRb = Sb;
    }
    hlld_tail_bwd(wL, wR, g, SL, SR, SM, pts, den, left, fb, -S, SLb, SRb, SMb, ptsb, eLb, eRb, ptLb, ptRb, wLb,
                  wRb);
    return;
  }
  // regions 3 and 4: both star states and the double-star state
  double wLs[8], wRs[8], eLs, eRs;
  hlld_star(wL, SL, SM, pts, s.eL, s.ptL, wLs, eLs);
  hlld_star(wR, SR, SM, pts, s.eR, s.ptR, wRs, eRs);
  const double rlr = sqrt(wLs[0] * wRs[0]);
  const double sg = sign0(wL[5]), sgR = sign0(wR[5]);
  double wss[8];
  wss[0] = wLs[0];
  wss[1] = SM;
  wss[2] = (rl * wLs[2] + rr * wRs[2] + (wRs[6] - wLs[6]) * sg) / (rl + rr);
  wss[3] = (rl * wLs[3] + rr * wRs[3] + (wRs[7] - wLs[7]) * sg) / (rl + rr);
  wss[4] = 0.;
  wss[5] = wLs[5];
  wss[6] = (rl * wRs[6] + rr * wLs[6] + rlr * (wRs[2] - wLs[2]) * sg) / (rl + rr);
  wss[7] = (rl * wRs[7] + rr * wLs[7] + rlr * (wRs[3] - wLs[3]) * sg) / (rl + rr);
  const double dss = vdotb(wss);
  const double eLss = eLs - rl * (vdotb(wLs) - dss) * sg;
  const double eRss = eRs + rr * (vdotb(wRs) - dss) * sgR;
  // region 3: f = FL + SL* q** - (SL* - SL) qL* - SL qL (region 4 the same on the right)
  const double Ss = left ? SLs : SRs;
  const double* wsa = left ? wLs : wRs;
  const double esa = left ? eLs : eRs, essa = left ? eLss : eRss;
  double Sb = 0., Ssb = 0.;
  for (int k = 0; k < 8; k++) {
    double q, F;
    side_qF(w, ee, HH, k, q, F);
    const double qs = star_cons_k(wsa, esa, k), qss = star_cons_k(wss, essa, k);
    if (f) f[k] = F + Ss * qss - (Ss - S) * qs - S * q;
    Ssb += fb[k] * (qss - qs);
    Sb += fb[k] * (qs - q);
  }
  double wLsb[8] = {0., 0., 0., 0., 0., 0., 0., 0.}, wRsb[8] = {0., 0., 0., 0., 0., 0., 0., 0.};
  double wssb[8] = {0., 0., 0., 0., 0., 0., 0., 0.};
  double eLsb = 0., eRsb = 0., eLssb = 0., eRssb = 0., rlb = 0., rrb = 0., rlrb = 0.;
  double SLsb = 0., SRsb = 0.;
  if (left) {
    star_cons_bwd_s(wss, fb, Ss, wssb, eLssb);
    star_cons_bwd_s(wLs, fb, -(Ss - S), wLsb, eLsb);
    SLsb = Ssb;
    SLb = Sb;
  } else {
    star_cons_bwd_s(wss, fb, Ss, wssb, eRssb);
    star_cons_bwd_s(wRs, fb, -(Ss - S), wRsb, eRsb);
    SRsb = Ssb;
    SRb = Sb;
  }
  // S_L* = SM - |BxL| / rl, S_R* = SM + |BxR| / rr
  SMb += SLsb + SRsb;
  wLb[5] -= SLsb * sign0(wL[5]) / rl;
  rlb += SLsb * fabs(wL[5]) / (rl * rl);
  wRb[5] += SRsb * sign0(wR[5]) / rr;
  rrb -= SRsb * fabs(wR[5]) / (rr * rr);
  // e_L** = e_L* - rl (vdotb(wLs) - dss) sg; e_R** = e_R* + rr (vdotb(wRs) - dss) sgR
  double dssb = 0., vLsb = 0., vRsb = 0.;
  eLsb += eLssb;
  rlb -= eLssb * (vdotb(wLs) - dss) * sg;
  vLsb -= eLssb * rl * sg;
  dssb += eLssb * rl * sg;
  eRsb += eRssb;
  rrb += eRssb * (vdotb(wRs) - dss) * sgR;
  vRsb += eRssb * rr * sgR;
  dssb -= eRssb * rr * sgR;
  wLsb[1] += vLsb * wLs[5];
  wLsb[5] += vLsb * wLs[1];
  wLsb[2] += vLsb * wLs[6];
  wLsb[6] += vLsb * wLs[2];
  wLsb[3] += vLsb * wLs[7];
  wLsb[7] += vLsb * wLs[3];
  wRsb[1] += vRsb * wRs[5];
  wRsb[5] += vRsb * wRs[1];
  wRsb[2] += vRsb * wRs[6];
  wRsb[6] += vRsb * wRs[2];
  wRsb[3] += vRsb * wRs[7];
  wRsb[7] += vRsb * wRs[3];
  // dss = vdotb(w**)
  wssb[1] += dssb * wss[5];
  wssb[5] += dssb * wss[1];
  wssb[2] += dssb * wss[6];
  wssb[6] += dssb * wss[2];
  wssb[3] += dssb * wss[7];
  wssb[7] += dssb * wss[3];
  // w**
  const double D = rl + rr;
  double Db = 0.;
  wLsb[0] += wssb[0];
  SMb += wssb[1];
  wLsb[5] += wssb[5];
  {
    const double N = wssb[2] / D;
    Db -= wssb[2] * wss[2] / D;
    rlb += N * wLs[2];
    wLsb[2] += N * rl;
    rrb += N * wRs[2];
    wRsb[2] += N * rr;
    wRsb[6] += N * sg;
    wLsb[6] -= N * sg;
  }
  {
    const double N = wssb[3] / D;
    Db -= wssb[3] * wss[3] / D;
    rlb += N * wLs[3];
    wLsb[3] += N * rl;
    rrb += N * wRs[3];
    wRsb[3] += N * rr;
    wRsb[7] += N * sg;
    wLsb[7] -= N * sg;
  }
  {
    const double N = wssb[6] / D;
    Db -= wssb[6] * wss[6] / D;
    rlb += N * wRs[6];
    wRsb[6] += N * rl;
    rrb += N * wLs[6];
    wLsb[6] += N * rr;
    rlrb += N * (wRs[2] - wLs[2]) * sg;
    wRsb[2] += N * rlr * sg;
    wLsb[2] -= N * rlr * sg;
  }
  {
    const double N = wssb[7] / D;
    Db -= wssb[7] * wss[7] / D;
    rlb += N * wRs[7];
    wRsb[7] += N * rl;
    rrb += N * wLs[7];
    wLsb[7] += N * rr;
    rlrb += N * (wRs[3] - wLs[3]) * sg;
    wRsb[3] += N * rlr * sg;
    wLsb[3] -= N * rlr * sg;
  }
  rlb += Db;
  rrb += Db;
  // rl = sqrt(rhoL*), rr = sqrt(rhoR*), rlr = sqrt(rhoL* rhoR*)
  if (rl > 0.) wLsb[0] += rlb * 0.5 / rl;
  if (rr > 0.) wRsb[0] += rrb * 0.5 / rr;
  if (rlr > 0.) {
    const double pb = rlrb * 0.5 / rlr;
    wLsb[0] += pb * wRs[0];
    wRsb[0] += pb * wLs[0];
  }
  hlld_star_bwd(wL, SL, SM, pts, s.eL, s.ptL, wLs, eLs, wLsb, eLsb, wLb, SLb, SMb, ptsb, eLb, ptLb);
  hlld_star_bwd(wR, SR, SM, pts, s.eR, s.ptR, wRs, eRs, wRsb, eRsb, wRb, SRb, SMb, ptsb, eRb, ptRb);
  hlld_tail_bwd(wL, wR, g, SL, SR, SM, pts, den, left, fb, -S, SLb, SRb, SMb, ptsb, eLb, eRb, ptLb, ptRb, wLb,
                wRb);
}

}  // namespace amr
