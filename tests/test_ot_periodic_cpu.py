"""Orszag-Tang vortex, periodic boundaries and SSP-RK3 (build-defined: the
reference has none of them), pinned on the oracle by what the mathematics
fixes: the closed-form initial state, the periodic ghost identity, exact
mass and momentum conservation and div B under CT on a periodic uniform mesh,
and the SSP-RK3 stage combinations expressed through the reference's own
stage-1 operator."""
import math

import numpy as np
import pytest

import oracle as O

NT, NG = 24, 4
OT = dict(ic="OT", limiter="WENO5", flux="HLLD", cfl=0.4, nx=16, nghosts=4, periodic=True)


def _leaves(d):
    lp, ic, lf = d.get_tree()
    return [nb for nb in range(len(lf)) if lf[nb]], lp, ic


def test_ot_initial_state_closed_form():
    d = O.Domain(maxlevs=1, rk_stages=3, **OT)
    assert d.init() == 0 and d.info()["gamma"] == 5.0 / 3.0
    W = d.get_field(O.W)
    leaves, lp, ic = _leaves(d)
    dx = 1.0 / 32
    pi = math.pi
    for nb in leaves:
        for j in range(NT):
            for i in range(NT):
                x = (i + ic[nb, 0] - NG + 0.5) * dx
                y = (j + ic[nb, 1] - NG + 0.5) * dx
                w = W[nb, :, j, i]
                assert math.isclose(w[0], 25 / (36 * pi), rel_tol=1e-15)
                assert math.isclose(w[4], 5 / (12 * pi), rel_tol=1e-14)
                assert abs(w[1] + math.sin(2 * pi * y)) < 1e-15 and abs(w[2] - math.sin(2 * pi * x)) < 1e-15
                b0 = 1 / math.sqrt(4 * pi)
                assert abs(w[5] + b0 * math.sin(2 * pi * y)) < 1e-15 and abs(w[6] - b0 * math.sin(4 * pi * x)) < 1e-15


def test_periodic_fill_wraps():
    """On the 2x2 root mesh every edge ghost cell holds the interior value of
    the cell one domain length away.  Corner ghost cells follow the
    reference's order-dependent fill (a same-level copy spans the neighbour's
    ghost rows as they are at that moment, Q6) and are not checked here."""
    d = O.Domain(maxlevs=1, **OT)
    assert d.init() == 0
    rng = np.random.default_rng(5)
    U = d.get_field(O.U)
    U[:] = rng.uniform(0.5, 1.5, U.shape)
    d.set_field(O.U, U)
    d.locate_bounds()
    d.fill(O.U, O.FILL_CC)
    U = d.get_field(O.U)
    leaves, lp, ic = _leaves(d)
    # global periodic field of the 32x32 interior cells
    G = np.zeros((8, 32, 32))
    for nb in leaves:
        x0, y0 = ic[nb] // 1
        G[:, y0:y0 + 16, x0:x0 + 16] = U[nb, :, NG:NT - NG, NG:NT - NG]
    for nb in leaves:
        x0, y0 = ic[nb]
        for j in range(NT):
            for i in range(NT):
                corner = (i < NG or i >= NT - NG) and (j < NG or j >= NT - NG)
                if corner:
                    continue
                gi, gj = (x0 + i - NG) % 32, (y0 + j - NG) % 32
                assert np.array_equal(U[nb, :, j, i], G[:, gj, gi]), (nb, i, j)


@pytest.mark.parametrize("rk,lim,flux", [(3, "WENO5", "HLLD"), (2, "MC", "HLLC"), (3, "minmod", "HLLE")])
def test_periodic_conservation_and_divb(rk, lim, flux):
    """Uniform periodic mesh: total mass and momentum are conserved to
    round-off and the face field stays divergence-free under CT."""
    ng = 4 if lim.startswith("WENO") else 2
    nt = 16 + 2 * ng
    d = O.Domain(ic="OT", limiter=lim, flux=flux, cfl=0.4, nx=16, nghosts=ng, maxlevs=1, periodic=True, rk_stages=rk)
    assert d.init() == 0

    def totals():
        U = d.get_field(O.U)
        return sum(U[nb, :, ng:nt - ng, ng:nt - ng].sum(axis=(1, 2)) for nb in range(4))

    def divb():
        B = d.get_field(O.BI)
        m = 0.0
        for nb in range(4):
            bx, by = B[nb, 0], B[nb, 1]
            dv = (bx[ng:nt - ng, ng + 1:nt - ng + 1] - bx[ng:nt - ng, ng:nt - ng]) + \
                 (by[ng + 1:nt - ng + 1, ng:nt - ng] - by[ng:nt - ng, ng:nt - ng])
            m = max(m, float(np.abs(dv).max()))
        return m

    t0 = totals()
    assert divb() < 1e-15
    for s in range(25):
        assert d.step() == 0, (s, d.last_error())
    t1 = totals()
    assert abs(t1[0] - t0[0]) < 1e-13 * abs(t0[0])
    assert abs(t1[1] - t0[1]) < 1e-12 and abs(t1[2] - t0[2]) < 1e-12 and abs(t1[3] - t0[3]) < 1e-12
    assert divb() < 1e-13


def _clone(src, **kw):
    d = O.Domain(**kw)
    assert d.init() == 0
    d.set_tree(*src.get_tree())
    for f in range(8):
        d.set_field(f, src.get_field(f))
    inf = src.info()
    d.set_time(inf["t"], inf["dt"], inf["dt_next"], inf["count"])
    return d


def test_ssp_rk3_combinations_via_stage_one():
    """Stage 2 is 3/4 U + 1/4 S(Us) and stage 3 is 1/3 U + 2/3 S(Us), where S is
    the forward-Euler stage-1 operator of the reference (MUSCL2D + RK2ndv1
    step 1), for the cell state and, through CT, for the face field."""
    kw = dict(maxlevs=1, rk_stages=3, **OT)
    a = O.Domain(**kw)
    assert a.init() == 0
    for _ in range(3):  # leave the dt = 0 start-up steps
        assert a.step() == 0
    a.locate_bounds()
    a.fill(O.U, O.FILL_CC)
    a.fill(O.BI, O.FILL_FC)
    assert a.stage(1) == 0
    a.fill(O.BI, O.FILL_FC)
    a.fill(O.FF, O.FILL_FC)
    a.fill(O.GG, O.FILL_FC)
    a.ct(1)
    a.fill(O.US, O.FILL_CC)
    for step, (cu, cs) in ((2, (0.75, 0.25)), (3, (1.0 / 3.0, 2.0 / 3.0))):
        U0, Us, Bi0, Bis = (a.get_field(f) for f in (O.U, O.US, O.BI, O.BIS))
        # forward-Euler stage of the same operator on the intermediate state
        b = _clone(a, **kw)
        b.set_field(O.U, Us)
        b.set_field(O.BI, Bis)
        b.locate_bounds()
        assert b.stage(1) == 0 and a.stage(step) == 0
        out = O.U if step == 3 else O.US
        assert np.array_equal(a.get_field(out), cu * U0 + cs * b.get_field(O.US))
        for f, k in ((O.FF, O.FILL_FC), (O.GG, O.FILL_FC)):
            a.fill(f, k)
            b.fill(f, k)
        b.ct(1)
        a.ct(step)
        outb = O.BI if step == 3 else O.BIS
        assert np.array_equal(a.get_field(outb), cu * Bi0 + cs * b.get_field(O.BIS))
        if step == 2:
            a.fill(O.US, O.FILL_CC)
