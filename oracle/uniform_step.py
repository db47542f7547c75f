"""Frozen uniform periodic mesh step in numpy -- TEST INFRASTRUCTURE.

Restates, on one N x N periodic grid with exact halos, the per-cell update of
the reference's step (main.cpp:128-247) for the minmod-MUSCL + HLLE
configuration: cons->prim (meshblock.cpp:119-179), minmod edge states
(celledges.cpp:9-64, slopelimiter.cpp:6-54) with the normal B taken from the
faces (MUSCL2D.cpp:44-53), HLLE (riemannS.cpp:118-189 with misc.cpp:71-140,
including the RoeAvg normaliser sqrt(rhoL + sqrt(rhoR))) or HLLD
(riemannS.cpp:397-549, misc.cpp:64-68 sign with sign(0) = 0), the RK2 stage
update (main.cpp:147-157, 179-189), and CT (CT2D.cpp:15-123: corner EMF, face
update with the reciprocal, cell sync with the energy fix).  Every elementwise
expression keeps the reference's operand order, so results are bitwise those
of the HIP frozen-mesh step.  Only tests and bench's checks use it.

Arrays: U[8][N][N] (j, i with i fastest), Bx[N][N] (x-face between cells
i-1 and i of row j), By[N][N] (y-face between rows j-1 and j of column i).
"""
import numpy as np

EPS = 1.0e-15


def prim(U, g):
    W = np.empty_like(U)
    W[0] = U[0]
    W[1] = U[1] / U[0]
    W[2] = U[2] / U[0]
    W[3] = U[3] / U[0]
    p = (g - 1) * (U[4] - 0.5 * W[0] * (W[1] * W[1] + W[2] * W[2] + W[3] * W[3]))
    W[5:8] = U[5:8]
    W[4] = p - (g - 1) * (0.5 * (U[5] * U[5] + U[6] * U[6] + U[7] * U[7]))
    return W


def minmod(r):
    m = np.where(1.0 < r, 1.0, r)          # mac_min(1, r)
    return np.where(0.0 > m, 0.0, m)        # mac_max(0, m)


def recon(vm1, v0, vp1, vp2):
    """(wl, wr) at the face between v0 and vp1 (muscl_plus / muscl_minus)."""
    a, b, d = v0 - vm1, vp1 - v0, vp2 - vp1
    wl = v0 + 0.5 * minmod(b / (a + EPS)) * a
    wr = vp1 - 0.5 * minmod(b / (d + EPS)) * d
    return wl, wr


def fast_speed(w, g):
    ca = np.sqrt((w[5] * w[5] + w[6] * w[6] + w[7] * w[7]) / w[0])
    a = np.sqrt(g * w[4] / w[0])
    can = np.sqrt((w[5] * w[5]) / w[0])
    s0 = (ca * ca) + (a * a)
    s1 = (ca * ca) - (a * a)
    return np.sqrt(0.5 * (s0 + np.sqrt(s1 * s1 + 4. * (a * a) * ((ca * ca) - (can * can)))))


def side(w, g):
    cf = fast_speed(w, g)
    vn, Bn = w[1], w[5]
    b2 = w[5] * w[5] + w[6] * w[6] + w[7] * w[7]
    e = (w[4] / (w[0] * (g - 1))) + 0.5 * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]) + 0.5 * b2 / w[0]
    e = e * w[0]
    H = (e + w[4] + 0.5 * b2) / w[0]
    q = w.copy()
    q[1] = w[0] * w[1]
    q[2] = w[0] * w[2]
    q[3] = w[0] * w[3]
    q[4] = e
    pt = w[4] + 0.5 * b2
    mv = w[0] * vn
    F = np.empty_like(w)
    F[0] = mv
    F[1] = mv * w[1] + pt - Bn * w[5]
    F[2] = mv * w[2] - Bn * w[6]
    F[3] = mv * w[3] - Bn * w[7]
    vb = w[1] * w[5] + w[2] * w[6] + w[3] * w[7]
    F[4] = mv * H - Bn * vb
    F[5] = vn * w[5] - Bn * w[1]
    F[6] = vn * w[6] - Bn * w[2]
    F[7] = vn * w[7] - Bn * w[3]
    return q, F, H, cf, e, pt


def roe_bounds(wL, wR, g, HL, HR):
    sl, sr = np.sqrt(wL[0]), np.sqrt(wR[0])
    nrm = np.sqrt(wL[0] + np.sqrt(wR[0]))
    v = [(sl * wL[i + 1] + sr * wR[i + 1]) / nrm for i in range(3)]
    B = [(sl * wL[i + 5] + sr * wR[i + 5]) / nrm for i in range(3)]
    RT = np.sqrt(wR[0] / wL[0])
    d = np.sqrt(wR[0] * wL[0])
    d6, d7 = wL[6] - wR[6], wL[7] - wR[7]
    X = 0.5 * ((d6 * d6) + (d7 * d7)) / (sl + sr)
    Y = 0.5 * (wL[0] + wR[0]) / d
    h = (HL + RT * HR) / (1 + RT)
    vsq = v[0] * v[0] + v[1] * v[1] + v[2] * v[2]
    btsq = (B[1] * B[1]) + (B[2] * B[2])
    bt2s = (g - 1. - (g - 2) * Y) * btsq
    vax2 = (B[0] * B[0]) / d
    hp = h - (vax2 + btsq / d)
    a2 = ((g - 1.) * (hp - 0.5 * vsq) - (g - 2.) * X)
    a2 = np.where(a2 > EPS, a2, EPS)
    ct2 = bt2s / d
    tsum = vax2 + ct2 + a2
    tdif = vax2 + ct2 - a2
    cfast = np.sqrt(0.5 * (tsum + np.sqrt(tdif * tdif + 4. * a2 * ct2)))
    return v[0] - cfast, v[0] + cfast


def hlle(wL, wR, g):
    qL, FL, HL, cfL = side(wL, g)[:4]
    qR, FR, HR, cfR = side(wR, g)[:4]
    lmin, lmax = roe_bounds(wL, wR, g, HL, HR)
    xl = wL[1] - cfL
    il = np.where(lmin < xl, lmin, xl)       # std::min
    SL = np.where(il < 0., il, 0.)           # MIN(., 0)
    xr = wR[1] + cfR
    ir = np.where(xr < lmax, lmax, xr)       # std::max
    SR = np.where(ir > 0., ir, 0.)           # MAX(., 0)
    return (SR * FL - SL * FR + SL * SR * (qR - qL)) / (SR - SL)


def sign0(x):
    return np.where(x > 0, 1.0, np.where(x < 0, -1.0, 0.0))


def vdotb(w):
    s = w[1] * w[5]
    s = s + w[2] * w[6]
    return s + w[3] * w[7]


def hlld_star(w, S, SM, pts, e, pt):
    """Star state of one side (riemannS.cpp:467-497); ws[4] stays 0."""
    ws = np.empty_like(w)
    ws[0] = w[0] * (S - w[1]) / (S - SM)
    ws[1] = SM
    ws[4] = 0.
    ws[5] = w[5]
    den = w[0] * (S - w[1]) * (S - SM) - (w[5] * w[5])
    nz = den != 0
    dd = np.where(nz, den, 1.0)
    sv = (SM - w[1]) / dd
    ws[2] = np.where(nz, w[2] - w[5] * w[6] * sv, w[2])
    ws[3] = np.where(nz, w[3] - w[5] * w[7] * sv, w[3])
    sb = (w[0] * ((S - w[1]) * (S - w[1])) - (w[5] * w[5])) / dd
    ws[6] = np.where(nz, w[6] * sb, 0.)
    ws[7] = np.where(nz, w[7] * sb, 0.)
    es = (((S - w[1]) * e - pt * w[1] + pts * SM + w[5] * (vdotb(w) - vdotb(ws))) / (S - SM))
    return ws, es


def star_cons(w, e):
    u = w.copy()
    u[1] = w[0] * w[1]
    u[2] = w[0] * w[2]
    u[3] = w[0] * w[3]
    u[4] = e
    return u


def hlld(wL, wR, g):
    """HLLD (riemannS.cpp:397-549, Miyoshi & Kusano 2005) with the
    reference's wave-speed estimate and region conditions."""
    qL, FL, HL, cfL, eL, ptL = side(wL, g)
    qR, FR, HR, cfR, eR, ptR = side(wR, g)
    vL, vR = wL[1], wR[1]
    cmax = np.where(cfL < cfR, cfR, cfL)            # std::max
    SL = np.where(vR < vL, vR, vL) - cmax           # std::min
    SR = np.where(vL < vR, vR, vL) + cmax
    den = (SR - vR) * wR[0] - (SL - vL) * wL[0]
    SM = ((SR - vR) * wR[0] * vR - (SL - vL) * wL[0] * vL - ptR + ptL) / den
    pts = ((SR - vR) * wR[0] * ptL - (SL - vL) * wL[0] * ptR +
           wL[0] * wR[0] * (SR - vR) * (SL - vL) * (vR - vL)) / den
    wLs, eLs = hlld_star(wL, SL, SM, pts, eL, ptL)
    wRs, eRs = hlld_star(wR, SR, SM, pts, eR, ptR)
    rl, rr, rlr = np.sqrt(wLs[0]), np.sqrt(wRs[0]), np.sqrt(wLs[0] * wRs[0])
    sg = sign0(wL[5])
    wss = np.empty_like(wL)
    wss[0] = wLs[0]
    wss[1] = SM
    wss[2] = (rl * wLs[2] + rr * wRs[2] + (wRs[6] - wLs[6]) * sg) / (rl + rr)
    wss[3] = (rl * wLs[3] + rr * wRs[3] + (wRs[7] - wLs[7]) * sg) / (rl + rr)
    wss[4] = 0.
    wss[5] = wLs[5]
    wss[6] = (rl * wRs[6] + rr * wLs[6] + rlr * (wRs[2] - wLs[2]) * sg) / (rl + rr)
    wss[7] = (rl * wRs[7] + rr * wLs[7] + rlr * (wRs[3] - wLs[3]) * sg) / (rl + rr)
    dss = vdotb(wss)
    eLss = eLs - rl * (vdotb(wLs) - dss) * sg
    eRss = eRs + rr * (vdotb(wRs) - dss) * sign0(wR[5])
    SLs = SM - np.abs(wL[5]) / rl
    SRs = SM + np.abs(wR[5]) / rr
    qLs, qRs = star_cons(wLs, eLs), star_cons(wRs, eRs)
    qLss, qRss = star_cons(wss, eLss), star_cons(wss, eRss)
    c1 = SL > 0
    c2 = ~c1 & (SL <= 0.) & (SLs >= 0.)
    c3 = ~c1 & ~c2 & (SLs <= 0.) & (SM >= 0.)
    c4 = ~c1 & ~c2 & ~c3 & (SM <= 0.) & (SRs >= 0.)
    c5 = ~c1 & ~c2 & ~c3 & ~c4 & (SRs <= 0.) & (SR >= 0.)
    c6 = ~c1 & ~c2 & ~c3 & ~c4 & ~c5 & (SR < 0.)
    return np.select([c1, c2, c3, c4, c5, c6],
                     [FL, FL + SL * qLs - SL * qL, FL + SLs * qLss - (SLs - SL) * qLs - SL * qL,
                      FR + SRs * qRss - (SRs - SR) * qRs - SR * qR, FR + SR * qRs - SR * qR, FR], 0.)


def fluxes(W, Bx, By, g, flux="HLLE"):
    """x-face fluxes Fx[8][j][i] (face between i-1 and i) and y-face Fy."""
    riemann = hlld if flux == "HLLD" else hlle
    sh = lambda A, o, ax: np.roll(A, -o, axis=ax)  # A[.. idx + o ..]
    # x faces: left cell i-1, right cell i; stencil i-2 .. i+1 (axis 2)
    wl, wr = recon(sh(W, -2, 2), sh(W, -1, 2), W, sh(W, 1, 2))
    wl[5] = Bx
    wr[5] = Bx
    with np.errstate(all="ignore"):  # unused branches of the region selection
        Fx = riemann(wl, wr, g)
    # y faces: swap x/y components, then swap the flux back
    wl, wr = recon(sh(W, -2, 1), sh(W, -1, 1), W, sh(W, 1, 1))
    wl[6] = By
    wr[6] = By
    sw = [0, 2, 1, 3, 4, 6, 5, 7]
    with np.errstate(all="ignore"):
        Fy = riemann(wl[sw], wr[sw], g)[sw]
    return Fx, Fy


def residual(Fx, Fy, dx):
    return (np.roll(Fx, -1, axis=2) - Fx) / dx + (np.roll(Fy, -1, axis=1) - Fy) / dx


def emf(Fx, Fy):
    """E[j][i] at the corner (i-1/2, j-1/2)."""
    return 0.25 * ((np.roll(Fy[5], 1, axis=1) + Fy[5]) - (np.roll(Fx[6], 1, axis=0) + Fx[6]))


def bsync(U, Bx, By):
    bx = 0.5 * (np.roll(Bx, -1, axis=1) + Bx)
    by = 0.5 * (np.roll(By, -1, axis=0) + By)
    q5, q6 = U[5].copy(), U[6].copy()
    U[4] = U[4] + 0.5 * ((bx * bx + by * by) - ((q5 * q5) + (q6 * q6)))
    U[5] = bx
    U[6] = by


def step(U, Bx, By, dt, dx, g, flux="HLLE"):
    """One RK2 step with CT; returns new (U, Bx, By)."""
    # stage 1
    Fx, Fy = fluxes(prim(U, g), Bx, By, g, flux)
    Us = U - dt * residual(Fx, Fy, dx)
    E = emf(Fx, Fy)
    rx = (1. / dx) * (np.roll(E, -1, axis=0) - E)
    ry = -(1. / dx) * (np.roll(E, -1, axis=1) - E)
    Bxs = Bx - dt * rx
    Bys = By - dt * ry
    bsync(Us, Bxs, Bys)
    # stage 2
    Fx, Fy = fluxes(prim(Us, g), Bxs, Bys, g, flux)
    Un = 0.5 * (U + Us - dt * residual(Fx, Fy, dx))
    E = emf(Fx, Fy)
    rx = (1. / dx) * (np.roll(E, -1, axis=0) - E)
    ry = -(1. / dx) * (np.roll(E, -1, axis=1) - E)
    Bxn = 0.5 * (Bxs + Bx - dt * rx)
    Byn = 0.5 * (Bys + By - dt * ry)
    bsync(Un, Bxn, Byn)
    return Un, Bxn, Byn


def cfl_dt(U, g, dx, cfl):
    """CFL bound of the reference (main.cpp:209-230, with its density-for-
    velocity slip) on the whole grid."""
    W = prim(U, g)
    lam = np.maximum(np.abs(W[0]) + fast_speed(W, g),
                     np.abs(W[1]) + fast_speed(W[[0, 2, 1, 3, 4, 6, 5, 7]], g))
    return cfl * dx / lam.max()
