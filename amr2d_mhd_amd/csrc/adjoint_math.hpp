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

// ub += (dw/du)^T wb
__host__ __device__ inline void prim_vjp(const double* u, double g, const double* wb, double* ub) {
  double w[8];
  prim_of(u, g, w);
  const double pb = wb[4];
  const double sv = w[1] * w[1] + w[2] * w[2] + w[3] * w[3];
  double w0b = wb[0] + pb * (-(g - 1) * 0.5 * sv);
  ub[4] += pb * (g - 1);
  for (int k = 1; k <= 3; k++) {
    const double wkb = wb[k] + pb * (-(g - 1) * w[0] * 2. * w[k] * 0.5);
    ub[k] += wkb / u[0];
    w0b -= wkb * w[k] / u[0];
  }
  ub[0] += w0b;
  for (int k = 5; k <= 7; k++) ub[k] += wb[k] - pb * (g - 1) * u[k];
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
__host__ __device__ inline double minmod_slope(double r) { return (!(1. < r) && !(0. > r)) ? 1. : 0.; }

// vb[0..3] += d(wl, wr)/dv ^T (wlb, wrb)
__host__ __device__ inline void recon_minmod_vjp(const double* v, double wlb, double wrb, double* vb) {
  const double a = v[1] - v[0], b = v[2] - v[1], d = v[3] - v[2];
  const double rl = b / (a + kEps), rr = b / (d + kEps);
  const double pl = limiter(kLimMinmod, rl, 0.), pr = limiter(kLimMinmod, rr, 0.);
  // wl = v1 + 0.5 * pl * a
  double ab = 0.5 * pl * wlb;
  const double rlb = 0.5 * a * wlb * minmod_slope(rl);
  double bb = rlb / (a + kEps);
  ab -= rlb * rl / (a + kEps);
  vb[1] += wlb;
  // wr = v2 - 0.5 * pr * d
  double db = -0.5 * pr * wrb;
  const double rrb = -0.5 * d * wrb * minmod_slope(rr);
  bb += rrb / (d + kEps);
  db -= rrb * rr / (d + kEps);
  vb[2] += wrb;
  vb[1] += ab;
  vb[0] -= ab;
  vb[2] += bb;
  vb[1] -= bb;
  vb[3] += db;
  vb[2] -= db;
}

// ------------------------------------------------------------- HLLE flux
// fast speed along x (fast_speed(w, g, 5)) and its gradient
__host__ __device__ inline void fast_speed_vjp(const double* w, double g, double cfb, double* wb) {
  const double B2 = w[5] * w[5] + w[6] * w[6] + w[7] * w[7];
  const double ca = sqrt(B2 / w[0]);
  const double a = sqrt(g * w[4] / w[0]);
  const double can = sqrt((w[5] * w[5]) / w[0]);
  const double s0 = (ca * ca) + (a * a);
  const double s1 = (ca * ca) - (a * a);
  const double inner = s1 * s1 + 4. * (a * a) * ((ca * ca) - (can * can));
  const double sq = sqrt(inner);
  const double cf = sqrt(0.5 * (s0 + sq));
  // backward
  const double tb = cfb * 0.5 / cf;  // d cf / d (0.5 (s0 + sq))
  const double s0b = 0.5 * tb;
  const double sqb = 0.5 * tb;
  const double innerb = sq > 0. ? sqb * 0.5 / sq : 0.;
  const double s1b = innerb * 2. * s1;
  const double a2 = a * a, ca2 = ca * ca, can2 = can * can;
  double a2b = innerb * 4. * (ca2 - can2) + s0b - s1b;
  double ca2b = innerb * 4. * a2 + s0b + s1b;
  const double can2b = -innerb * 4. * a2;
  const double cab = ca2b * 2. * ca, ab = a2b * 2. * a, canb = can2b * 2. * can;
  // ca = sqrt(B2 / w0)
  if (ca > 0.) {
    const double qb = cab * 0.5 / ca;
    const double B2b = qb / w[0];
    wb[0] -= qb * (B2 / w[0]) / w[0];
    wb[5] += B2b * 2. * w[5];
    wb[6] += B2b * 2. * w[6];
    wb[7] += B2b * 2. * w[7];
  }
  // a = sqrt(g * w4 / w0)
  if (a > 0.) {
    const double qb = ab * 0.5 / a;
    wb[4] += qb * g / w[0];
    wb[0] -= qb * (g * w[4] / w[0]) / w[0];
  }
  // can = sqrt(w5^2 / w0)
  if (can > 0.) {
    const double qb = canb * 0.5 / can;
    wb[5] += qb * 2. * w[5] / w[0];
    wb[0] -= qb * ((w[5] * w[5]) / w[0]) / w[0];
  }
}

// side_state backward: adjoints of q, F, H and cf -> wb
__host__ __device__ inline void side_vjp(const double* w, double g, const Side& s, const double* qb,
                                         const double* Fb, double Hb, double cfb, double* wb) {
  const double vn = w[1], Bn = w[5];
  const double b2 = w[5] * w[5] + w[6] * w[6] + w[7] * w[7];
  const double mv = w[0] * vn;
  const double vb_dot = w[1] * w[5] + w[2] * w[6] + w[3] * w[7];
  double mvb = 0., ptb = 0., eb = 0., b2b = 0., vnb = 0., Bnb = 0.;
  // F
  mvb += Fb[0];
  mvb += Fb[1] * w[1];
  wb[1] += Fb[1] * mv;
  ptb += Fb[1];
  Bnb -= Fb[1] * w[5];
  wb[5] -= Fb[1] * Bn;
  mvb += Fb[2] * w[2];
  wb[2] += Fb[2] * mv;
  Bnb -= Fb[2] * w[6];
  wb[6] -= Fb[2] * Bn;
  mvb += Fb[3] * w[3];
  wb[3] += Fb[3] * mv;
  Bnb -= Fb[3] * w[7];
  wb[7] -= Fb[3] * Bn;
  mvb += Fb[4] * s.H;
  Hb += Fb[4] * mv;
  Bnb -= Fb[4] * vb_dot;
  const double vdb = -Fb[4] * Bn;
  wb[1] += vdb * w[5];
  wb[5] += vdb * w[1];
  wb[2] += vdb * w[6];
  wb[6] += vdb * w[2];
  wb[3] += vdb * w[7];
  wb[7] += vdb * w[3];
  vnb += Fb[5] * w[5];
  wb[5] += Fb[5] * vn;
  Bnb -= Fb[5] * w[1];
  wb[1] -= Fb[5] * Bn;
  vnb += Fb[6] * w[6];
  wb[6] += Fb[6] * vn;
  Bnb -= Fb[6] * w[2];
  wb[2] -= Fb[6] * Bn;
  vnb += Fb[7] * w[7];
  wb[7] += Fb[7] * vn;
  Bnb -= Fb[7] * w[3];
  wb[3] -= Fb[7] * Bn;
  // mv = w0 * vn
  wb[0] += mvb * vn;
  vnb += mvb * w[0];
  // pt = w4 + 0.5 b2
  wb[4] += ptb;
  b2b += 0.5 * ptb;
  // q: q0 = w0, q1..3 = w0 w1..3, q4 = e, q5..7 = w5..7
  wb[0] += qb[0];
  for (int k = 1; k <= 3; k++) {
    wb[0] += qb[k] * w[k];
    wb[k] += qb[k] * w[0];
  }
  eb += qb[4];
  wb[5] += qb[5];
  wb[6] += qb[6];
  wb[7] += qb[7];
  // H = (e + w4 + 0.5 b2) / w0
  eb += Hb / w[0];
  wb[4] += Hb / w[0];
  b2b += 0.5 * Hb / w[0];
  wb[0] -= Hb * s.H / w[0];
  // e = e1 * w0 with e1 = w4/(w0 (g-1)) + 0.5 |v|^2 + 0.5 b2 / w0
  const double e1 = s.e / w[0];
  const double e1b = eb * w[0];
  wb[0] += eb * e1;
  wb[4] += e1b / (w[0] * (g - 1));
  wb[0] -= e1b * (w[4] / (w[0] * (g - 1))) / w[0];
  wb[1] += e1b * w[1];
  wb[2] += e1b * w[2];
  wb[3] += e1b * w[3];
  b2b += e1b * 0.5 / w[0];
  wb[0] -= e1b * (0.5 * b2 / w[0]) / w[0];
  // b2 = w5^2 + w6^2 + w7^2
  wb[5] += b2b * 2. * w[5];
  wb[6] += b2b * 2. * w[6];
  wb[7] += b2b * 2. * w[7];
  wb[1] += vnb;
  wb[5] += Bnb;
  fast_speed_vjp(w, g, cfb, wb);
}

// roe_bounds backward: (u, lmin, lmax) adjoints -> wL, wR, HL, HR adjoints
__host__ __device__ inline void roe_vjp(const double* wL, const double* wR, double g, double HL, double HR,
                                        double ub, double lminb, double lmaxb, double* wLb, double* wRb, double& HLb,
                                        double& HRb) {
  const double sl = sqrt(wL[0]), sr = sqrt(wR[0]);
  const double srr = sqrt(wR[0]);
  const double nrm = sqrt(wL[0] + srr);
  double v[3], B[3];
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
  const double cY = (g - 1. - (g - 2) * Y);
  const double bt2s = cY * btsq;
  const double vax2 = (B[0] * B[0]) / d;
  const double hp = h - (vax2 + btsq / d);
  const double a2raw = ((g - 1.) * (hp - 0.5 * vsq) - (g - 2.) * X);
  const double a2 = mac_max(a2raw, kEps);
  const double ct2 = bt2s / d;
  const double tsum = vax2 + ct2 + a2;
  const double tdif = vax2 + ct2 - a2;
  const double sq = sqrt(tdif * tdif + 4. * a2 * ct2);
  const double cfast = sqrt(0.5 * (tsum + sq));
  // backward
  double vb[3] = {ub + lminb + lmaxb, 0., 0.}, Bb[3] = {0., 0., 0.};
  const double cfb = -lminb + lmaxb;
  const double tb = cfb * 0.5 / cfast;
  const double tsumb = 0.5 * tb;
  const double sqb = 0.5 * tb;
  const double inb = sq > 0. ? sqb * 0.5 / sq : 0.;
  const double tdifb = inb * 2. * tdif;
  double a2b = inb * 4. * ct2 + tsumb - tdifb;
  double ct2b = inb * 4. * a2 + tsumb + tdifb;
  double vax2b = tsumb + tdifb;
  // ct2 = bt2s / d
  double bt2sb = ct2b / d;
  double db = -ct2b * ct2 / d;
  // a2 = max(a2raw, eps)
  const double a2rb = (a2raw > kEps) ? a2b : 0.;
  double hpb = a2rb * (g - 1.);
  double vsqb = -a2rb * (g - 1.) * 0.5;
  const double Xb = -a2rb * (g - 2.);
  // hp = h - (vax2 + btsq / d)
  double hb = hpb;
  vax2b -= hpb;
  double btsqb = -hpb / d;
  db += hpb * (btsq / d) / d;
  // vax2 = B0^2 / d
  Bb[0] += vax2b * 2. * B[0] / d;
  db -= vax2b * vax2 / d;
  // bt2s = cY * btsq
  btsqb += bt2sb * cY;
  double Yb = -bt2sb * btsq * (g - 2);
  Bb[1] += btsqb * 2. * B[1];
  Bb[2] += btsqb * 2. * B[2];
  for (int i = 0; i < 3; i++) vb[i] += vsqb * 2. * v[i];
  // h = (HL + RT HR) / (1 + RT)
  HLb += hb / (1 + RT);
  HRb += hb * RT / (1 + RT);
  double RTb = hb * HR / (1 + RT) - hb * h / (1 + RT);
  // Y = 0.5 (rhoL + rhoR) / d
  wLb[0] += Yb * 0.5 / d;
  wRb[0] += Yb * 0.5 / d;
  db -= Yb * Y / d;
  // X = 0.5 (d6^2 + d7^2) / (sl + sr)
  const double d6b = Xb * d6 / (sl + sr), d7b = Xb * d7 / (sl + sr);
  double slb = -Xb * X / (sl + sr), srb = -Xb * X / (sl + sr);
  wLb[6] += d6b;
  wRb[6] -= d6b;
  wLb[7] += d7b;
  wRb[7] -= d7b;
  // d = sqrt(rhoR rhoL)
  if (d > 0.) {
    const double pb = db * 0.5 / d;
    wRb[0] += pb * wL[0];
    wLb[0] += pb * wR[0];
  }
  // RT = sqrt(rhoR / rhoL)
  if (RT > 0.) {
    const double qb = RTb * 0.5 / RT;
    wRb[0] += qb / wL[0];
    wLb[0] -= qb * (wR[0] / wL[0]) / wL[0];
  }
  // v_i, B_i = (sl a_L + sr a_R) / nrm
  double nrmb = 0.;
  for (int i = 0; i < 3; i++) {
    slb += vb[i] * wL[i + 1] / nrm + Bb[i] * wL[i + 5] / nrm;
    srb += vb[i] * wR[i + 1] / nrm + Bb[i] * wR[i + 5] / nrm;
    wLb[i + 1] += vb[i] * sl / nrm;
    wRb[i + 1] += vb[i] * sr / nrm;
    wLb[i + 5] += Bb[i] * sl / nrm;
    wRb[i + 5] += Bb[i] * sr / nrm;
    nrmb -= vb[i] * v[i] / nrm + Bb[i] * B[i] / nrm;
  }
  // nrm = sqrt(rhoL + sqrt(rhoR))
  const double inner_b = nrmb * 0.5 / nrm;
  wLb[0] += inner_b;
  wRb[0] += inner_b * 0.5 / srr;
  wLb[0] += slb * 0.5 / sl;
  wRb[0] += srb * 0.5 / sr;
}

// HLLE flux (flux_hlle) backward: fb -> wLb, wRb (accumulated)
__host__ __device__ inline void hlle_vjp(const double* wL, const double* wR, double g, const double* fb, double* wLb,
                                         double* wRb) {
  Side L, R;
  side_state(wL, g, L);
  side_state(wR, g, R);
  double u, lmin, lmax;
  roe_bounds(wL, wR, g, L.H, R.H, u, lmin, lmax);
  const double xl = wL[1] - L.cf, il = std_min(xl, lmin);
  const double SL = mac_min(il, 0.);
  const double xr = wR[1] + R.cf, ir = std_max(xr, lmax);
  const double SR = mac_max(ir, 0.);
  const double D = SR - SL;
  double SLb = 0., SRb = 0., Db = 0.;
  double qLb[8], qRb[8], FLb[8], FRb[8];
  for (int k = 0; k < 8; k++) {
    const double dq = R.q[k] - L.q[k];
    const double f = (SR * L.F[k] - SL * R.F[k] + SL * SR * dq) / D;
    const double Nb = fb[k] / D;
    Db -= fb[k] * f / D;
    FLb[k] = SR * Nb;
    FRb[k] = -SL * Nb;
    qRb[k] = SL * SR * Nb;
    qLb[k] = -qRb[k];
    SRb += Nb * (L.F[k] + SL * dq);
    SLb += Nb * (-R.F[k] + SR * dq);
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
  roe_vjp(wL, wR, g, L.H, R.H, 0., lminb, lmaxb, wLb, wRb, HLb, HRb);
  side_vjp(wL, g, L, qLb, FLb, HLb, cfLb, wLb);
  side_vjp(wR, g, R, qRb, FRb, HRb, cfRb, wRb);
}

}  // namespace amr
