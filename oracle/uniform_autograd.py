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
    return q, F, H, cf
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
    qL, FL, HL, cfL = side(wL, g); qR, FR, HR, cfR = side(wR, g); lmin, lmax = roe(wL, wR, g, HL, HR)
    xl = wL[1] - cfL; il = torch.where(lmin < xl, lmin, xl); SL = torch.where(il < 0., il, torch.zeros_like(il))
    xr = wR[1] + cfR; ir = torch.where(xr < lmax, lmax, xr); SR = torch.where(ir > 0., ir, torch.zeros_like(ir))
    return (SR * FL - SL * FR + SL * SR * (qR - qL)) / (SR - SL)
SW = [0, 2, 1, 3, 4, 6, 5, 7]
def fluxes(W, Bx, By, g):
    wl, wr = recon(roll(W, 2, 2), roll(W, 1, 2), W, roll(W, -1, 2))
    wl = torch.cat([wl[:5], Bx[None], wl[6:]]); wr = torch.cat([wr[:5], Bx[None], wr[6:]])
    Fx = hlle(wl, wr, g)
    wl, wr = recon(roll(W, 2, 1), roll(W, 1, 1), W, roll(W, -1, 1))
    wl = torch.cat([wl[:6], By[None], wl[7:]]); wr = torch.cat([wr[:6], By[None], wr[7:]])
    Fy = hlle(wl[SW], wr[SW], g)[SW]
    return Fx, Fy
def residual(Fx, Fy, dx): return (roll(Fx, -1, 2) - Fx) / dx + (roll(Fy, -1, 1) - Fy) / dx
def emf(Fx, Fy): return 0.25 * ((roll(Fy[5], 1, 1) + Fy[5]) - (roll(Fx[6], 1, 0) + Fx[6]))
def bsync(U, Bx, By):
    bx = 0.5 * (roll(Bx, -1, 1) + Bx); by = 0.5 * (roll(By, -1, 0) + By)
    e = U[4] + 0.5 * ((bx * bx + by * by) - ((U[5] * U[5]) + (U[6] * U[6])))
    return torch.stack([U[0], U[1], U[2], U[3], e, bx, by, U[7]])
def step(U, Bx, By, dt, dx, g):
    Fx, Fy = fluxes(prim(U, g), Bx, By, g); Us = U - dt * residual(Fx, Fy, dx); E = emf(Fx, Fy)
    rx = (1. / dx) * (roll(E, -1, 0) - E); ry = -(1. / dx) * (roll(E, -1, 1) - E)
    Bxs = Bx - dt * rx; Bys = By - dt * ry; Us = bsync(Us, Bxs, Bys)
    Fx, Fy = fluxes(prim(Us, g), Bxs, Bys, g); Un = 0.5 * (U + Us - dt * residual(Fx, Fy, dx)); E = emf(Fx, Fy)
    rx = (1. / dx) * (roll(E, -1, 0) - E); ry = -(1. / dx) * (roll(E, -1, 1) - E)
    Bxn = 0.5 * (Bxs + Bx - dt * rx); Byn = 0.5 * (Bys + By - dt * ry)
    return bsync(Un, Bxn, Byn), Bxn, Byn
