"""Frozen uniform periodic mesh step in torch, for autograd -- TEST INFRASTRUCTURE.

The same restatement as uniform_step.py (see there for the reference
file:line of every piece), written with torch operations so that torch's
autograd yields the exact branch-wise derivative of the discrete step; the
GPU tests compare the hand-written adjoint kernels against it.  Only tests
use it.
"""
import torch
EPS = 1.0e-15
def roll(A, s, d): return torch.roll(A, s, dims=d)
def prim(U, g):
    w0 = U[0]; w1 = U[1] / U[0]; w2 = U[2] / U[0]; w3 = U[3] / U[0]
    p = (g - 1) * (U[4] - 0.5 * w0 * (w1 * w1 + w2 * w2 + w3 * w3))
    w4 = p - (g - 1) * (0.5 * (U[5] * U[5] + U[6] * U[6] + U[7] * U[7]))
    return torch.stack([w0, w1, w2, w3, w4, U[5], U[6], U[7]])
def minmod(r):
    m = torch.where(1.0 < r, torch.ones_like(r), r)
    return torch.where(0.0 > m, torch.zeros_like(m), m)
def recon(vm1, v0, vp1, vp2):
    a, b, d = v0 - vm1, vp1 - v0, vp2 - vp1
    return v0 + 0.5 * minmod(b / (a + EPS)) * a, vp1 - 0.5 * minmod(b / (d + EPS)) * d
def fast(w, g):
    ca = torch.sqrt((w[5] * w[5] + w[6] * w[6] + w[7] * w[7]) / w[0]); a = torch.sqrt(g * w[4] / w[0])
    can = torch.sqrt((w[5] * w[5]) / w[0]); s0 = ca * ca + a * a; s1 = ca * ca - a * a
    return torch.sqrt(0.5 * (s0 + torch.sqrt(s1 * s1 + 4. * (a * a) * ((ca * ca) - (can * can)))))
def side(w, g):
    cf = fast(w, g); vn, Bn = w[1], w[5]; b2 = w[5] * w[5] + w[6] * w[6] + w[7] * w[7]
    e = (w[4] / (w[0] * (g - 1))) + 0.5 * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]) + 0.5 * b2 / w[0]; e = e * w[0]
    H = (e + w[4] + 0.5 * b2) / w[0]
    q = torch.stack([w[0], w[0] * w[1], w[0] * w[2], w[0] * w[3], e, w[5], w[6], w[7]])
    pt = w[4] + 0.5 * b2; mv = w[0] * vn; vb = w[1] * w[5] + w[2] * w[6] + w[3] * w[7]
    F = torch.stack([mv, mv * w[1] + pt - Bn * w[5], mv * w[2] - Bn * w[6], mv * w[3] - Bn * w[7], mv * H - Bn * vb,
                     vn * w[5] - Bn * w[1], vn * w[6] - Bn * w[2], vn * w[7] - Bn * w[3]])
    return q, F, H, cf, e, pt
def roe(wL, wR, g, HL, HR):
    sl, sr = torch.sqrt(wL[0]), torch.sqrt(wR[0]); nrm = torch.sqrt(wL[0] + torch.sqrt(wR[0]))
    v = [(sl * wL[i + 1] + sr * wR[i + 1]) / nrm for i in range(3)]; B = [(sl * wL[i + 5] + sr * wR[i + 5]) / nrm for i in range(3)]
    RT = torch.sqrt(wR[0] / wL[0]); d = torch.sqrt(wR[0] * wL[0]); d6, d7 = wL[6] - wR[6], wL[7] - wR[7]
    X = 0.5 * ((d6 * d6) + (d7 * d7)) / (sl + sr); Y = 0.5 * (wL[0] + wR[0]) / d
    h = (HL + RT * HR) / (1 + RT); vsq = v[0] * v[0] + v[1] * v[1] + v[2] * v[2]; btsq = B[1] * B[1] + B[2] * B[2]
    bt2s = (g - 1. - (g - 2) * Y) * btsq; vax2 = (B[0] * B[0]) / d; hp = h - (vax2 + btsq / d)
    a2 = ((g - 1.) * (hp - 0.5 * vsq) - (g - 2.) * X); a2 = torch.where(a2 > EPS, a2, torch.full_like(a2, EPS))
    ct2 = bt2s / d; tsum = vax2 + ct2 + a2; tdif = vax2 + ct2 - a2
    cf = torch.sqrt(0.5 * (tsum + torch.sqrt(tdif * tdif + 4. * a2 * ct2)))
    return v[0] - cf, v[0] + cf
def hlle(wL, wR, g):
    qL, FL, HL, cfL = side(wL, g)[:4]; qR, FR, HR, cfR = side(wR, g)[:4]; lmin, lmax = roe(wL, wR, g, HL, HR)
    xl = wL[1] - cfL; il = torch.where(lmin < xl, lmin, xl); SL = torch.where(il < 0., il, torch.zeros_like(il))
    xr = wR[1] + cfR; ir = torch.where(xr < lmax, lmax, xr); SR = torch.where(ir > 0., ir, torch.zeros_like(ir))
    return (SR * FL - SL * FR + SL * SR * (qR - qL)) / (SR - SL)
def W_(c, a, b): return torch.where(c, a, b)
def sign0(x): return W_(x > 0, torch.ones_like(x), W_(x < 0, -torch.ones_like(x), torch.zeros_like(x)))
def vdotb(w): return w[1] * w[5] + w[2] * w[6] + w[3] * w[7]
def hlld_star(w, S, SM, pts, e, pt):
    den = w[0] * (S - w[1]) * (S - SM) - (w[5] * w[5]); nz = den != 0; dd = W_(nz, den, torch.ones_like(den))
    sv = (SM - w[1]) / dd; sb = (w[0] * ((S - w[1]) * (S - w[1])) - (w[5] * w[5])) / dd; z = torch.zeros_like(den)
    ws = torch.stack([w[0] * (S - w[1]) / (S - SM), SM, W_(nz, w[2] - w[5] * w[6] * sv, w[2]),
                      W_(nz, w[3] - w[5] * w[7] * sv, w[3]), z, w[5], W_(nz, w[6] * sb, z), W_(nz, w[7] * sb, z)])
    es = ((S - w[1]) * e - pt * w[1] + pts * SM + w[5] * (vdotb(w) - vdotb(ws))) / (S - SM)
    return ws, es
def star_cons(w, e): return torch.stack([w[0], w[0] * w[1], w[0] * w[2], w[0] * w[3], e, w[5], w[6], w[7]])
def hlld(wL, wR, g):
    qL, FL, HL, cfL, eL, ptL = side(wL, g); qR, FR, HR, cfR, eR, ptR = side(wR, g); vL, vR = wL[1], wR[1]
    cmax = W_(cfL < cfR, cfR, cfL); SL = W_(vR < vL, vR, vL) - cmax; SR = W_(vL < vR, vR, vL) + cmax
    den = (SR - vR) * wR[0] - (SL - vL) * wL[0]
    SM = ((SR - vR) * wR[0] * vR - (SL - vL) * wL[0] * vL - ptR + ptL) / den
    pts = ((SR - vR) * wR[0] * ptL - (SL - vL) * wL[0] * ptR + wL[0] * wR[0] * (SR - vR) * (SL - vL) * (vR - vL)) / den
    wLs, eLs = hlld_star(wL, SL, SM, pts, eL, ptL); wRs, eRs = hlld_star(wR, SR, SM, pts, eR, ptR)
    rl, rr, rlr = torch.sqrt(wLs[0]), torch.sqrt(wRs[0]), torch.sqrt(wLs[0] * wRs[0]); sg = sign0(wL[5]); D = rl + rr
    wss = torch.stack([wLs[0], SM, (rl * wLs[2] + rr * wRs[2] + (wRs[6] - wLs[6]) * sg) / D,
                       (rl * wLs[3] + rr * wRs[3] + (wRs[7] - wLs[7]) * sg) / D, torch.zeros_like(SM), wLs[5],
                       (rl * wRs[6] + rr * wLs[6] + rlr * (wRs[2] - wLs[2]) * sg) / D,
                       (rl * wRs[7] + rr * wLs[7] + rlr * (wRs[3] - wLs[3]) * sg) / D])
    dss = vdotb(wss); eLss = eLs - rl * (vdotb(wLs) - dss) * sg; eRss = eRs + rr * (vdotb(wRs) - dss) * sign0(wR[5])
    SLs = SM - torch.abs(wL[5]) / rl; SRs = SM + torch.abs(wR[5]) / rr
    qLs, qRs, qLss, qRss = star_cons(wLs, eLs), star_cons(wRs, eRs), star_cons(wss, eLss), star_cons(wss, eRss)
    c1 = SL > 0; c2 = ~c1 & (SL <= 0.) & (SLs >= 0.); c3 = ~c1 & ~c2 & (SLs <= 0.) & (SM >= 0.)
    c4 = ~c1 & ~c2 & ~c3 & (SM <= 0.) & (SRs >= 0.); c5 = ~c1 & ~c2 & ~c3 & ~c4 & (SRs <= 0.) & (SR >= 0.)
    c6 = ~c1 & ~c2 & ~c3 & ~c4 & ~c5 & (SR < 0.)
    out = torch.zeros_like(FL)
    for c, f in ((c6, FR), (c5, FR + SR * qRs - SR * qR), (c4, FR + SRs * qRss - (SRs - SR) * qRs - SR * qR),
                 (c3, FL + SLs * qLss - (SLs - SL) * qLs - SL * qL), (c2, FL + SL * qLs - SL * qL), (c1, FL)):
        out = W_(c, f, out)
    return out
SW = [0, 2, 1, 3, 4, 6, 5, 7]
def fluxes(W, Bx, By, g, flux="HLLE"):
    riemann = hlld if flux == "HLLD" else hlle
    wl, wr = recon(roll(W, 2, 2), roll(W, 1, 2), W, roll(W, -1, 2))
    wl = torch.cat([wl[:5], Bx[None], wl[6:]]); wr = torch.cat([wr[:5], Bx[None], wr[6:]])
    Fx = riemann(wl, wr, g)
    wl, wr = recon(roll(W, 2, 1), roll(W, 1, 1), W, roll(W, -1, 1))
    wl = torch.cat([wl[:6], By[None], wl[7:]]); wr = torch.cat([wr[:6], By[None], wr[7:]])
    Fy = riemann(wl[SW], wr[SW], g)[SW]
    return Fx, Fy
def residual(Fx, Fy, dx): return (roll(Fx, -1, 2) - Fx) / dx + (roll(Fy, -1, 1) - Fy) / dx
def emf(Fx, Fy): return 0.25 * ((roll(Fy[5], 1, 1) + Fy[5]) - (roll(Fx[6], 1, 0) + Fx[6]))
def bsync(U, Bx, By):
    bx = 0.5 * (roll(Bx, -1, 1) + Bx); by = 0.5 * (roll(By, -1, 0) + By)
    e = U[4] + 0.5 * ((bx * bx + by * by) - ((U[5] * U[5]) + (U[6] * U[6])))
    return torch.stack([U[0], U[1], U[2], U[3], e, bx, by, U[7]])
def step(U, Bx, By, dt, dx, g, flux="HLLE"):
    Fx, Fy = fluxes(prim(U, g), Bx, By, g, flux); Us = U - dt * residual(Fx, Fy, dx); E = emf(Fx, Fy)
    rx = (1. / dx) * (roll(E, -1, 0) - E); ry = -(1. / dx) * (roll(E, -1, 1) - E)
    Bxs = Bx - dt * rx; Bys = By - dt * ry; Us = bsync(Us, Bxs, Bys)
    Fx, Fy = fluxes(prim(Us, g), Bxs, Bys, g, flux); Un = 0.5 * (U + Us - dt * residual(Fx, Fy, dx)); E = emf(Fx, Fy)
    rx = (1. / dx) * (roll(E, -1, 0) - E); ry = -(1. / dx) * (roll(E, -1, 1) - E)
    Bxn = 0.5 * (Bxs + Bx - dt * rx); Byn = 0.5 * (Bys + By - dt * ry)
    return bsync(Un, Bxn, Byn), Bxn, Byn
