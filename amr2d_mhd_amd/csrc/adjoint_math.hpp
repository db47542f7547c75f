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

}  // namespace amr
