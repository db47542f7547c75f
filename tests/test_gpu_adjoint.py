"""Frozen-mesh step on the GPU and its hand-written discrete adjoint.

Bar: the HIP step equals the numpy restatement (itself bitwise equal to the
oracle's AMR step on a single-level mesh) bit for bit; the adjoint gradient
matches central finite differences of the HIP forward and of the CPU forward
along seeded random directions; checkpoint-and-recompute does not change the
gradient."""
import numpy as np
import pytest

import uniform_step as US

pytestmark = pytest.mark.gpu

GAMMA = 5.0 / 3.0


def _setup(n, steps, vscale=1.0, flux="HLLE"):
    import torch

    from amr2d_mhd_amd import adjoint as AD

    rho, v, p, Bx, By = AD.orszag_tang_controls(n, vscale=vscale, modulation=0.05, noise=1e-3)
    mesh = AD.FrozenMesh(n, 1.0 / n, GAMMA, flux)
    U = AD.cons_from_controls(rho, v, p, Bx, By, GAMMA).contiguous()
    dts, Uc, Bxc, Byc = [], U, Bx, By
    for _ in range(steps):
        dts.append(AD.cfl_dt(Uc, GAMMA, 1.0 / n, 0.4))
        Uc, Bxc, Byc = mesh.step(Uc, Bxc, Byc, dts[-1])
    torch.cuda.synchronize()
    return mesh, (rho, v, p, Bx, By), dts


def _objective(ctrl, mesh, dts, target, every):
    from amr2d_mhd_amd import adjoint as AD

    rho, v, p, Bx, By = ctrl
    U0 = AD.cons_from_controls(rho, v, p, Bx, By, GAMMA).contiguous()
    UT, _, _ = AD.rollout(U0, Bx.contiguous(), By.contiguous(), mesh, dts, every)
    return ((UT[0] - target) ** 2).sum()


@pytest.mark.parametrize("flux", ["HLLE", "HLLD"])
def test_gpu_step_equals_numpy_bitwise(flux):
    import torch

    mesh, (rho, v, p, Bx, By), dts = _setup(32, 4, flux=flux)
    from amr2d_mhd_amd import adjoint as AD

    U = AD.cons_from_controls(rho, v, p, Bx, By, GAMMA).contiguous()
    Un, Bxn, Byn = U.cpu().numpy(), Bx.cpu().numpy(), By.cpu().numpy()
    Ug, Bxg, Byg = U, Bx, By
    for dt in dts:
        Ug, Bxg, Byg = mesh.step(Ug, Bxg, Byg, dt)
        Un, Bxn, Byn = US.step(Un, Bxn, Byn, dt, 1.0 / 32, GAMMA, flux)
        torch.cuda.synchronize()
        assert np.array_equal(Ug.cpu().numpy(), Un)
        assert np.array_equal(Bxg.cpu().numpy(), Bxn) and np.array_equal(Byg.cpu().numpy(), Byn)


def _grad_and_target(n, steps, every, flux="HLLE"):
    import torch

    from amr2d_mhd_amd import adjoint as AD

    mesh, ctrl, dts = _setup(n, steps, flux=flux)
    # target: the same set-up with the initial velocity scaled by 1.05
    rt, vt, pt, Bxt, Byt = AD.orszag_tang_controls(n, vscale=1.05, modulation=0.05, noise=1e-3)
    with torch.no_grad():
        Ut = AD.cons_from_controls(rt, vt, pt, Bxt, Byt, GAMMA).contiguous()
        UT, _, _ = AD.rollout(Ut, Bxt, Byt, mesh, dts, every)
        target = UT[0].clone()
    leaves = [t.clone().requires_grad_(True) for t in ctrl]
    J = _objective(leaves, mesh, dts, target, every)
    J.backward()
    return mesh, ctrl, dts, target, [t.grad.clone() for t in leaves], float(J)


@pytest.mark.parametrize("flux", ["HLLE", "HLLD"])
def test_adjoint_matches_central_fd(flux):
    """Away from limiter kinks: generic (seeded-noise) initial data and a
    step small enough that no limiter ratio crosses a switch point (at 1e-6
    a few of the ~1e5 minmod ratios per stage sit close enough to r = 1 to
    be crossed, which finite differences cannot resolve)."""
    import torch

    n, steps, every = 64, 30, 10
    mesh, ctrl, dts, target, grads, J0 = _grad_and_target(n, steps, every, flux)
    assert J0 > 0
    gen = torch.Generator(device="cpu").manual_seed(20261202)
    eps = 1e-8
    errs = []
    for _ in range(6):
        dirs = [torch.randn(t.shape, generator=gen, dtype=torch.float64).to(t.device) * t.abs().mean()
                for t in ctrl]
        with torch.no_grad():
            jp = _objective([t + eps * d for t, d in zip(ctrl, dirs)], mesh, dts, target, every)
            jm = _objective([t - eps * d for t, d in zip(ctrl, dirs)], mesh, dts, target, every)
        fd = float(jp - jm) / (2 * eps)
        ad = float(sum((g * d).sum() for g, d in zip(grads, dirs)))
        errs.append(abs(fd - ad) / max(abs(fd), 1e-30))
    # at eps 1e-8 the finite differences themselves carry ~1e-5 of round-off
    assert np.median(errs) < 5e-5 and max(errs) < 2e-4, errs


def test_checkpoint_interval_does_not_change_gradient():
    """Checkpointing every 1 or 5 steps and taping every step in the forward
    (every = 0) give the same gradient bit for bit."""
    n, steps = 32, 12
    _, _, _, _, g1, _ = _grad_and_target(n, steps, 1)
    _, _, _, _, g5, _ = _grad_and_target(n, steps, 5)
    _, _, _, _, g0, _ = _grad_and_target(n, steps, 0)
    for a, b, c in zip(g1, g5, g0):
        assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())
        assert np.array_equal(a.cpu().numpy(), c.cpu().numpy())


def test_tape_reverse_equals_recomputing_reverse():
    """amr_adj_vjp_tape over a tape written by amr_adj_step_tape equals
    amr_adj_step_vjp (which recomputes the step) bit for bit, and the taped
    step returns the plain step's state."""
    import torch

    from amr2d_mhd_amd import adjoint as AD

    n = 48
    mesh, (rho, v, p, Bx, By), dts = _setup(n, 2)
    U = AD.cons_from_controls(rho, v, p, Bx, By, GAMMA).contiguous()
    tape = mesh.new_tapes(1, U.device)[0]
    outs_t = mesh.step(U, Bx, By, dts[0], tape)
    outs = mesh.step(U, Bx, By, dts[0])
    gen = torch.Generator().manual_seed(5)
    seeds = [torch.randn(t.shape, generator=gen, dtype=torch.float64).cuda() for t in outs]
    g_tape = mesh.vjp(tape, Bx, By, dts[0], *seeds)
    g_re = mesh.step_vjp(U, Bx, By, dts[0], *seeds)
    torch.cuda.synchronize()
    for a, b in zip(outs_t, outs):
        assert torch.equal(a, b)
    for a, b in zip(g_tape, g_re):
        assert torch.equal(a, b)


@pytest.mark.parametrize("n,flux", [(16, "HLLE"), (45, "HLLE"), (64, "HLLE"), (45, "HLLD")])
def test_tiled_reverse_equals_per_face_reverse(n, flux):
    """The tiled face reverse (stencil sums in LDS, faces at tile edges
    reversed twice, grids narrower than a tile and not a multiple of it)
    equals the per-face kernels + gather up to the summation order."""
    import torch

    from amr2d_mhd_amd import adjoint as AD
    from amr2d_mhd_amd._lib import lib

    mesh, (rho, v, p, Bx, By), dts = _setup(n, 2, flux=flux)
    U = AD.cons_from_controls(rho, v, p, Bx, By, GAMMA).contiguous()
    outs = mesh.step(U, Bx, By, dts[0])
    gen = torch.Generator().manual_seed(11)
    seeds = [torch.randn(t.shape, generator=gen, dtype=torch.float64).cuda() for t in outs]
    res = {}
    for mode in (0, 2, 1):  # per-face + gather, 32x13 tiles, 16x13 tiles (the default, set last)
        assert lib().amr_debug_adj_set_fused(mesh.h, mode) == 0
        res[mode] = mesh.step_vjp(U, Bx, By, dts[0], *seeds)
    torch.cuda.synchronize()
    for mode in (1, 2):
        for a, b in zip(res[mode], res[0]):
            assert torch.allclose(a, b, rtol=1e-13, atol=1e-15 * float(b.abs().max())), (mode, float((a - b).abs().max()))


def test_adjoint_matches_fd_of_cpu_forward():
    """Directional derivative from the GPU adjoint against central finite
    differences of the numpy forward (the restated reference step)."""
    import torch

    from amr2d_mhd_amd import adjoint as AD

    n, steps = 32, 3
    mesh, ctrl, dts = _setup(n, steps)
    rho, v, p, Bx, By = ctrl
    target = torch.full((n, n), 0.2, dtype=torch.float64, device="cuda")
    leaves = [t.clone().requires_grad_(True) for t in ctrl]
    J = _objective(leaves, mesh, dts, target, 2)
    J.backward()

    def J_cpu(c):
        r, vv, pp, bx, by = [t.cpu().numpy() for t in c]
        U = AD.cons_from_controls(*[torch.from_numpy(np.asarray(a)) for a in (r, vv, pp, bx, by)], GAMMA).numpy()
        for dt in dts:
            U, bx, by = US.step(U, bx, by, dt, 1.0 / n, GAMMA)
        return float(((U[0] - 0.2) ** 2).sum())

    gen = torch.Generator(device="cpu").manual_seed(7)
    eps = 1e-8
    for _ in range(3):
        dirs = [torch.randn(t.shape, generator=gen, dtype=torch.float64).to(t.device) * t.abs().mean() for t in ctrl]
        fd = (J_cpu([t + eps * d for t, d in zip(ctrl, dirs)]) - J_cpu([t - eps * d for t, d in zip(ctrl, dirs)])) / (
            2 * eps)
        ad = float(sum((l.grad * d).sum() for l, d in zip(leaves, dirs)))
        assert abs(fd - ad) < 1e-5 * abs(fd), (fd, ad)


@pytest.mark.parametrize("flux", ["HLLE", "HLLD"])
def test_adjoint_equals_autograd_of_restatement(flux):
    """The hand-written adjoint of one step equals torch autograd through the
    torch restatement of the step, component by component."""
    import torch

    import uniform_autograd as UA
    from amr2d_mhd_amd import adjoint as AD

    n = 16
    rho, v, p, Bx, By = AD.orszag_tang_controls(n, modulation=0.05, noise=1e-3)
    mesh = AD.FrozenMesh(n, 1.0 / n, GAMMA, flux)
    U = AD.cons_from_controls(rho, v, p, Bx, By, GAMMA).contiguous()
    dt = AD.cfl_dt(U, GAMMA, 1.0 / n, 0.4)
    for _ in range(2):
        U, Bx, By = mesh.step(U, Bx, By, dt)
    gen = torch.Generator().manual_seed(3)
    seeds = [torch.randn(t.shape, generator=gen, dtype=torch.float64) for t in (U, Bx, By)]
    Ub, Bxb, Byb = mesh.step_vjp(U, Bx, By, dt, *[s.cuda() for s in seeds])
    Ut, Bxt, Byt = [t.detach().cpu().clone().requires_grad_(True) for t in (U, Bx, By)]
    outs = UA.step(Ut, Bxt, Byt, dt, 1.0 / n, GAMMA, flux)
    sum((o * s).sum() for o, s in zip(outs, seeds)).backward()
    for got, ref in ((Ub, Ut.grad), (Bxb, Bxt.grad), (Byb, Byt.grad)):
        assert torch.allclose(got.cpu(), ref, rtol=1e-12, atol=1e-13 * float(ref.abs().max()))


def _jvp_restatement(ctrl, dirs, dts, target, n, flux):
    """Exact directional derivative J'(x) d of the rollout along the branches
    the forward takes: forward-mode AD through the torch restatement of the
    step (oracle/uniform_autograd.py), on the GPU.  Returns (J, J'd, U_T)."""
    import torch
    import torch.autograd.forward_ad as fwAD

    import uniform_autograd as UA
    from amr2d_mhd_amd import adjoint as AD

    with torch.no_grad(), fwAD.dual_level():
        rho, v, p, Bx, By = [fwAD.make_dual(t, d) for t, d in zip(ctrl, dirs)]
        U = AD.cons_from_controls(rho, v, p, Bx, By, GAMMA)
        for dt in dts:
            U, Bx, By = UA.step(U, Bx, By, dt, 1.0 / n, GAMMA, flux)
        J = ((U[0] - target) ** 2).sum()
        Jp, Jt = fwAD.unpack_dual(J)
        return float(Jp), float(Jt), fwAD.unpack_dual(U).primal.clone()


def _fd(ctrl, dirs, eps, mesh, dts, target, every):
    import torch

    with torch.no_grad():
        jp = _objective([t + eps * d for t, d in zip(ctrl, dirs)], mesh, dts, target, every)
        jm = _objective([t - eps * d for t, d in zip(ctrl, dirs)], mesh, dts, target, every)
    return float(jp - jm) / (2 * eps)


@pytest.mark.parametrize("flux,ndirs", [("HLLE", 3), ("HLLD", 2)])
def test_adjoint_at_full_size_equals_exact_directional_derivative(flux, ndirs):
    """The adjoint workload at its own size: 512 x 512, 400 steps,
    checkpoints every 20.  For seeded random directions d, the hand-written
    adjoint's g.d equals J'(x) d computed exactly -- by forward-mode AD along
    the same branches through the torch restatement of the step -- to 1e-9
    relative, so the adjoint has no defect that shows only at scale (1024
    tiles, 20 checkpoint segments).  The restatement's rollout lands on the
    HIP rollout's final state to round-off.  (Central finite differences at
    this size do not converge to J'(x) d: with ~1e10 limiter evaluations per
    rollout, a perturbation of 1e-6..1e-8 moves some minmod ratios across
    their kinks.  AMR_ADJ_DIRS=32 prints adjoint, exact derivative and the
    FD quotients per direction.)"""
    import os

    import torch

    n, steps, every = 512, 400, 20
    mesh, ctrl, dts, target, grads, J0 = _grad_and_target(n, steps, every, flux)
    from amr2d_mhd_amd import adjoint as AD

    with torch.no_grad():
        U = AD.cons_from_controls(*ctrl, GAMMA).contiguous()
        UT, _, _ = AD.rollout(U, ctrl[3].contiguous(), ctrl[4].contiguous(), mesh, dts, every)
    ndirs = int(os.environ.get("AMR_ADJ_DIRS", ndirs))
    gen = torch.Generator(device="cpu").manual_seed(20261202)
    rows = []
    for k in range(ndirs):
        dirs = [torch.randn(t.shape, generator=gen, dtype=torch.float64).to(t.device) * t.abs().mean() for t in ctrl]
        ad = float(sum((g * d).sum() for g, d in zip(grads, dirs)))
        Jr, jv, UTr = _jvp_restatement(ctrl, dirs, dts, target, n, flux)
        if k == 0:
            assert abs(Jr - J0) <= 1e-12 * abs(J0), (Jr, J0)
            assert torch.allclose(UTr, UT, rtol=1e-10, atol=1e-12 * float(UT.abs().max()))
        rows.append((k, ad, jv, abs(ad - jv) / abs(jv)))
        if "AMR_ADJ_DIRS" in os.environ:
            f6, f8 = _fd(ctrl, dirs, 1e-6, mesh, dts, target, every), _fd(ctrl, dirs, 1e-8, mesh, dts, target, every)
            print(f"{flux} dir {k:2d}: adjoint {ad:+.15e}  exact {jv:+.15e}  rel {abs(ad - jv) / abs(jv):.2e}  "
                  f"FD1e-6 rel {abs(f6 - jv) / abs(jv):.2e}  FD1e-8 rel {abs(f8 - jv) / abs(jv):.2e}", flush=True)
    mesh.close()
    assert max(r[3] for r in rows) < 1e-9, rows
