"""Pointwise reverse-mode pieces of the frozen-mesh adjoint (host builds of
the device functions) against central finite differences, and the numpy
frozen-mesh step against the oracle's AMR step on a single-level periodic
mesh."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
import uniform_step as US
from amr2d_mhd_amd._lib import lib


def _p(a):
    return a.ctypes.data


def _state(rng):
    w = np.zeros(8)
    w[0] = rng.uniform(0.3, 2.0)
    w[1:4] = rng.normal(0, 0.7, 3)
    w[4] = rng.uniform(0.2, 2.0)
    w[5:8] = rng.normal(0, 0.8, 3)
    return w


def test_prim_vjp_matches_fd():
    rng = np.random.default_rng(11)
    for _ in range(50):
        w = _state(rng)
        u = np.zeros(8)
        u[0] = w[0]
        u[1:4] = w[0] * w[1:4]
        u[4] = w[4] / 0.4 + 0.5 * w[0] * (w[1:4] ** 2).sum() + 0.5 * (w[5:8] ** 2).sum()
        u[5:8] = w[5:8]
        wb, ub, wout = rng.normal(size=8), np.zeros(8), np.zeros(8)
        lib().amr_debug_prim_vjp(_p(u), 1.4, _p(wb), _p(ub), _p(wout))
        d = rng.normal(size=8)
        eps = 1e-6
        wp, wm = np.zeros(8), np.zeros(8)
        tmp = np.zeros(8)
        up, um = u + eps * d, u - eps * d  # keep the perturbed inputs alive across the calls
        lib().amr_debug_prim_vjp(_p(up), 1.4, _p(wb), _p(tmp), _p(wp))
        lib().amr_debug_prim_vjp(_p(um), 1.4, _p(wb), _p(tmp), _p(wm))
        fd = ((wp - wm) / (2 * eps)) @ wb
        assert abs(fd - ub @ d) < 1e-6 * max(1.0, abs(fd))


def test_recon_minmod_vjp_matches_fd():
    rng = np.random.default_rng(12)
    checked = 0
    while checked < 200:
        v = rng.normal(size=4)
        a, b, dd = v[1] - v[0], v[2] - v[1], v[3] - v[2]
        rl, rr = b / a, b / dd
        if min(abs(rl), abs(rl - 1), abs(rr), abs(rr - 1)) < 1e-3:
            continue  # keep clear of the limiter kinks
        wlb, wrb = rng.normal(size=2)
        vb, wlr = np.zeros(4), np.zeros(2)
        lib().amr_debug_recon_minmod_vjp(_p(v), wlb, wrb, _p(vb), _p(wlr))
        d = rng.normal(size=4)
        eps = 1e-7
        p, m, tmp = np.zeros(2), np.zeros(2), np.zeros(4)
        vp, vm = v + eps * d, v - eps * d
        lib().amr_debug_recon_minmod_vjp(_p(vp), wlb, wrb, _p(tmp), _p(p))
        lib().amr_debug_recon_minmod_vjp(_p(vm), wlb, wrb, _p(tmp), _p(m))
        fd = ((p - m) / (2 * eps)) @ np.array([wlb, wrb])
        assert abs(fd - vb @ d) < 1e-6 * max(1.0, abs(fd))
        checked += 1


@pytest.mark.parametrize("gamma", [1.4, 5.0 / 3.0, 2.0])
def test_reverse_forward_pieces_equal_engine_bitwise(gamma):
    """The forward values the HLLE reverse computes for itself (fast speed,
    energy, enthalpy, Roe bounds) are those of the engine's flux code bit for
    bit, so every branch the reverse takes is the forward's."""
    rng = np.random.default_rng(int(gamma * 1000) + 5)
    out, ref = np.zeros(8), np.zeros(8)
    for trial in range(2000):
        wL, wR = _state(rng), _state(rng)
        if trial % 3 == 0:
            wR[5] = wL[5]
        if trial % 7 == 0:  # vanishing transverse field and normal field
            wL[6:8] = 0.0
            wR[5:8] = 0.0
        lib().amr_debug_side_roe(_p(wL), _p(wR), gamma, _p(out), _p(ref))
        assert np.array_equal(out.view(np.uint64), ref.view(np.uint64)), (trial, out, ref)


@pytest.mark.parametrize("gamma", [1.4, 5.0 / 3.0, 2.0])
def test_hlle_vjp_matches_fd(gamma):
    rng = np.random.default_rng(int(gamma * 100))
    n = 0
    for trial in range(400):
        wL, wR = _state(rng), _state(rng)
        wR[5] = wL[5]
        if trial % 4 == 0:  # supersonic and clamped-speed cases
            wL[1] += 3.0
            wR[1] += 3.0
        fb = rng.normal(size=8)
        wLb, wRb = np.zeros(8), np.zeros(8)
        lib().amr_debug_hlle_vjp(_p(wL), _p(wR), gamma, _p(fb), _p(wLb), _p(wRb))
        dL, dR = rng.normal(size=8), rng.normal(size=8)
        eps = 1e-6
        fp, fm = np.zeros(8), np.zeros(8)
        lp, rp, lm, rm = wL + eps * dL, wR + eps * dR, wL - eps * dL, wR - eps * dR
        lib().amr_debug_hlle(_p(lp), _p(rp), gamma, _p(fp))
        lib().amr_debug_hlle(_p(lm), _p(rm), gamma, _p(fm))
        fd = ((fp - fm) / (2 * eps)) @ fb
        ad = wLb @ dL + wRb @ dR
        # branch switches of the wave-speed bounds inside +-eps are rare; allow them
        if abs(fd - ad) > 1e-5 * max(1.0, abs(fd)):
            f0 = np.zeros(8)
            l2, r2 = wL + 2 * eps * dL, wR + 2 * eps * dR
            lib().amr_debug_hlle(_p(l2), _p(r2), gamma, _p(f0))
            one = ((f0 - fp) / eps) @ fb
            assert abs(one - fd) > 1e-3 * max(1.0, abs(fd)), f"trial {trial}: fd {fd} adjoint {ad}"
            continue
        n += 1
    assert n > 380


@pytest.mark.parametrize("gamma", [1.4, 5.0 / 3.0, 2.0])
def test_hlld_vjp_matches_fd(gamma):
    """HLLD reverse: its own forward equals the engine's flux_hlld bit for bit,
    and its directional derivatives match central differences, over states
    reaching every wave region (supersonic both ways, subsonic, vanishing
    normal field)."""
    rng = np.random.default_rng(int(gamma * 100) + 7)
    n = 0
    for trial in range(600):
        wL, wR = _state(rng), _state(rng)
        wR[5] = wL[5]
        if trial % 6 == 0:
            wL[1] += 3.0
            wR[1] += 3.0
        elif trial % 6 == 1:
            wL[1] -= 3.0
            wR[1] -= 3.0
        elif trial % 6 == 2:
            wL[5] = wR[5] = 1e-3 * wL[5]
        fb = rng.normal(size=8)
        wLb, wRb, f, f0 = np.zeros(8), np.zeros(8), np.zeros(8), np.zeros(8)
        lib().amr_debug_hlld_vjp(_p(wL), _p(wR), gamma, _p(fb), _p(wLb), _p(wRb), _p(f))
        lib().amr_debug_hlld(_p(wL), _p(wR), gamma, _p(f0))
        assert np.array_equal(f.view(np.uint64), f0.view(np.uint64)), trial
        dL, dR = rng.normal(size=8), rng.normal(size=8)
        dR[5] = dL[5]  # one normal field per face
        eps = 1e-6
        fp, fm = np.zeros(8), np.zeros(8)
        lp, rp, lm, rm = wL + eps * dL, wR + eps * dR, wL - eps * dL, wR - eps * dR
        lib().amr_debug_hlld(_p(lp), _p(rp), gamma, _p(fp))
        lib().amr_debug_hlld(_p(lm), _p(rm), gamma, _p(fm))
        fd = ((fp - fm) / (2 * eps)) @ fb
        ad = wLb @ dL + wRb @ dR
        if abs(fd - ad) > 1e-5 * max(1.0, abs(fd)):
            # a wave-region or min/max switch inside +-eps: one-sided differences disagree
            f2 = np.zeros(8)
            l2, r2 = wL + 2 * eps * dL, wR + 2 * eps * dR
            lib().amr_debug_hlld(_p(l2), _p(r2), gamma, _p(f2))
            one = ((f2 - fp) / eps) @ fb
            assert abs(one - fd) > 1e-3 * max(1.0, abs(fd)), f"trial {trial}: fd {fd} adjoint {ad}"
            continue
        n += 1
    assert n > 580


@pytest.mark.parametrize("flux", ["HLLE", "HLLD"])
def test_numpy_frozen_step_equals_oracle_amr_step(flux):
    """One level, periodic, minmod + HLLE or HLLD, RK2: the numpy frozen-mesh
    step equals the oracle's AMR step (the reference's semantics) bit for bit."""
    ng, NT, N = 2, 20, 32
    d = O.Domain(ic="OT", limiter="minmod", flux=flux, cfl=0.4, nx=16, nghosts=ng, maxlevs=1, periodic=True)
    assert d.init() == 0

    def glob():
        U, B = d.get_field(O.U), d.get_field(O.BI)
        lp, ic, lf = d.get_tree()
        G, Bx, By = np.zeros((8, N, N)), np.zeros((N, N)), np.zeros((N, N))
        for nb in range(4):
            x0, y0 = ic[nb]
            G[:, y0:y0 + 16, x0:x0 + 16] = U[nb, :, ng:NT - ng, ng:NT - ng]
            Bx[y0:y0 + 16, x0:x0 + 16] = B[nb, 0, ng:NT - ng, ng:NT - ng]
            By[y0:y0 + 16, x0:x0 + 16] = B[nb, 1, ng:NT - ng, ng:NT - ng]
        return G, Bx, By

    for s in range(8):
        U0, Bx0, By0 = glob()
        dt = d.info()["dt"]
        assert d.step() == 0
        U1, Bx1, By1 = glob()
        Un, Bxn, Byn = US.step(U0, Bx0, By0, dt, 1.0 / 32, 5.0 / 3.0, flux)
        assert np.array_equal(Un, U1) and np.array_equal(Bxn, Bx1) and np.array_equal(Byn, By1), f"step {s}"


def _minmod_branches(W):
    """Branch taken by each minmod evaluation of a step's reconstruction:
    0 (r <= 0), 1 (0 < r <= 1: r), 2 (r > 1) for both edge states of every
    x and y face and variable."""
    import uniform_step as US

    out = []
    for ax in (2, 1):
        vm1, v0, vp1, vp2 = (np.roll(W, s, axis=ax) for s in (1, 0, -1, -2))
        a, b, d = v0 - vm1, vp1 - v0, vp2 - vp1
        for r in (b / (a + US.EPS), b / (d + US.EPS)):
            out.append(np.where(r > 1.0, 2, np.where(r > 0.0, 1, 0)))
    return np.stack(out)


@pytest.mark.parametrize("flux", ["HLLE", "HLLD"])
def test_torch_restatement_forward_equals_numpy(flux):
    """The torch restatement (oracle/uniform_autograd.py, the one the GPU
    adjoint tests differentiate) against the numpy one (uniform_step.py,
    bitwise equal to the oracle's AMR step): over 6 steps they take the same
    minmod branches at every face and agree to a few ulps.  They are not
    bitwise equal because torch's CPU sqrt is not correctly rounded (about 1
    in 110 results is off by one ulp), while numpy's, the reference's and the
    GPU's are."""
    import torch

    import uniform_autograd as UA
    import uniform_step as US
    from amr2d_mhd_amd import adjoint as AD

    g, n = 5.0 / 3.0, 32
    rho, v, p, Bx, By = AD.orszag_tang_controls(n, device="cpu", modulation=0.05, noise=1e-3)
    U = AD.cons_from_controls(rho, v, p, Bx, By, g).contiguous()
    dt = AD.cfl_dt(U, g, 1.0 / n, 0.4)
    Un, Bxn, Byn = U.numpy().copy(), Bx.numpy().copy(), By.numpy().copy()
    Ut, Bxt, Byt = U.clone(), Bx.clone(), By.clone()
    for s in range(6):
        assert np.array_equal(_minmod_branches(US.prim(Un, g)), _minmod_branches(UA.prim(Ut, g).numpy())), s
        Un, Bxn, Byn = US.step(Un, Bxn, Byn, dt, 1.0 / n, g, flux)
        Ut, Bxt, Byt = UA.step(Ut, Bxt, Byt, dt, 1.0 / n, g, flux)
        for a, b in ((Ut.numpy(), Un), (Bxt.numpy(), Bxn), (Byt.numpy(), Byn)):
            assert np.allclose(a, b, rtol=2e-15, atol=2e-15 * np.abs(b).max()), (s, np.abs(a - b).max())
    assert torch.equal(torch.sqrt(torch.tensor([4.0, 9.0])), torch.tensor([2.0, 3.0]))
