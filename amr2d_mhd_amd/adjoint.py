"""Discrete adjoint of the frozen-mesh step, behind a torch.autograd.Function.

The forward is the reference's RK2 step (minmod MUSCL + HLLE, CT) on one
uniform periodic N x N grid at a frozen dt sequence (``amr_adj_step``); the
backward runs the hand-written adjoint kernels (``amr_adj_step_vjp``) with
checkpoint-and-recompute over time steps: the forward keeps the state every
``every`` steps, and the backward recomputes each segment from its checkpoint
before sweeping it in reverse.  Tensors are float64 on the GPU; kernels run on
torch's current stream.

State layout: U[8][N][N] (rho, rho v, E, B), Bx[N][N] (x-face between cells
i-1 and i), By[N][N] (y-face between rows j-1 and j).
"""
import ctypes as C

import torch

from ._lib import AmrError, lib


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class FrozenMesh:
    """One uniform periodic N x N grid of cell size dx."""

    def __init__(self, n, dx, gamma):
        self.n, self.dx, self.gamma = int(n), float(dx), float(gamma)
        h = C.c_void_p()
        rc = lib().amr_adj_create(self.n, self.dx, self.gamma, C.byref(h))
        if rc:
            raise AmrError(rc)
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            try:
                lib().amr_adj_destroy(self.h)
            except Exception:  # interpreter shutdown
                pass
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, *ts):
        for t in ts:
            if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
                raise ValueError("frozen-mesh tensors must be contiguous float64 CUDA tensors")

    def step(self, U, Bx, By, dt):
        """One RK2 step; returns new (U, Bx, By)."""
        self._check(U, Bx, By)
        Uo, Bxo, Byo = torch.empty_like(U), torch.empty_like(Bx), torch.empty_like(By)
        rc = lib().amr_adj_step(self.h, _ptr(U), _ptr(Bx), _ptr(By), float(dt), _ptr(Uo), _ptr(Bxo), _ptr(Byo),
                                _stream())
        if rc:
            raise AmrError(rc)
        return Uo, Bxo, Byo

    def step_vjp(self, U, Bx, By, dt, Uob, Bxob, Byob):
        """Adjoint of one step taken from (U, Bx, By): J^T (Uob, Bxob, Byob)."""
        self._check(U, Bx, By, Uob, Bxob, Byob)
        Ub, Bxb, Byb = torch.empty_like(U), torch.empty_like(Bx), torch.empty_like(By)
        rc = lib().amr_adj_step_vjp(self.h, _ptr(U), _ptr(Bx), _ptr(By), float(dt), _ptr(Uob), _ptr(Bxob),
                                    _ptr(Byob), _ptr(Ub), _ptr(Bxb), _ptr(Byb), _stream())
        if rc:
            raise AmrError(rc)
        return Ub, Bxb, Byb


class Rollout(torch.autograd.Function):
    """(U0, Bx0, By0) -> state after len(dts) frozen-dt steps."""

    @staticmethod
    def forward(ctx, U0, Bx0, By0, mesh, dts, every):
        U, Bx, By = U0.contiguous(), Bx0.contiguous(), By0.contiguous()
        ckpts = []
        for s, dt in enumerate(dts):
            if s % every == 0:
                ckpts.append((U, Bx, By))
            U, Bx, By = mesh.step(U, Bx, By, dt)
        ctx.mesh, ctx.dts, ctx.every, ctx.ckpts = mesh, list(dts), int(every), ckpts
        return U, Bx, By

    @staticmethod
    def backward(ctx, gU, gBx, gBy):
        mesh, dts, every = ctx.mesh, ctx.dts, ctx.every
        gU = torch.zeros_like(ctx.ckpts[0][0]) if gU is None else gU.contiguous()
        gBx = torch.zeros_like(ctx.ckpts[0][1]) if gBx is None else gBx.contiguous()
        gBy = torch.zeros_like(ctx.ckpts[0][2]) if gBy is None else gBy.contiguous()
        for seg in range(len(ctx.ckpts) - 1, -1, -1):
            s0 = seg * every
            s1 = min(s0 + every, len(dts))
            # recompute the inputs of every step of the segment
            states = [ctx.ckpts[seg]]
            for s in range(s0, s1 - 1):
                states.append(mesh.step(*states[-1], dts[s]))
            for s in range(s1 - 1, s0 - 1, -1):
                gU, gBx, gBy = mesh.step_vjp(*states[s - s0], dts[s], gU, gBx, gBy)
        ctx.ckpts = None
        return gU, gBx, gBy, None, None, None


def rollout(U0, Bx0, By0, mesh, dts, every=20):
    return Rollout.apply(U0, Bx0, By0, mesh, dts, every)


# ------------------------------------------------- problem set-up helpers
def cons_from_controls(rho, v, p, Bx, By, gamma):
    """U from the controls: cell density, velocity[3], pressure and the face
    field (cell-centred B = the average of the two faces, Bz = 0)."""
    bcx = 0.5 * (torch.roll(Bx, -1, dims=1) + Bx)
    bcy = 0.5 * (torch.roll(By, -1, dims=0) + By)
    zero = torch.zeros_like(rho)
    E = p / (gamma - 1) + 0.5 * rho * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) + 0.5 * (bcx * bcx + bcy * bcy)
    return torch.stack([rho, rho * v[0], rho * v[1], rho * v[2], E, bcx, bcy, zero])


def orszag_tang_controls(n, device="cuda", vscale=1.0, modulation=0.0, noise=0.0, seed=20261202):
    """Orszag-Tang initial controls on [0,1]^2 (cell and face centres).  A
    nonzero `modulation` multiplies density, pressure and the two in-plane
    velocity components by smooth 1 + m sin(.) fields: plain Orszag-Tang has
    v_x constant along x and v_y along y, and the minmod limiter is not
    differentiable where neighbouring differences vanish.  A nonzero `noise`
    further multiplies them by 1 + noise * N(0,1) and adds the discrete curl
    of a random corner potential of relative size `noise` to the face field
    (keeping its discrete divergence zero), all seeded: the symmetric samples
    of the sines are otherwise exactly equal across their extrema, which puts
    those limiter evaluations exactly on a kink."""
    pi = torch.pi
    x = (torch.arange(n, dtype=torch.float64, device=device) + 0.5) / n
    xf = torch.arange(n, dtype=torch.float64, device=device) / n
    Y, X = torch.meshgrid(x, x, indexing="ij")
    Yf_x, Xf_x = torch.meshgrid(x, xf, indexing="ij")  # x-faces: x at i/n, y at (j+1/2)/n
    Yf_y, Xf_y = torch.meshgrid(xf, x, indexing="ij")  # y-faces
    B0 = 1.0 / (4 * pi) ** 0.5
    rho = torch.full((n, n), 25.0 / (36 * pi), dtype=torch.float64, device=device)
    p = torch.full((n, n), 5.0 / (12 * pi), dtype=torch.float64, device=device)
    if modulation:
        rho = rho * (1 + modulation * torch.sin(2 * pi * (X + 2 * Y) + 0.3))
        p = p * (1 + modulation * torch.sin(2 * pi * (3 * X - Y) + 1.1))
    vx, vy = -torch.sin(2 * pi * Y), torch.sin(2 * pi * X)
    if modulation:
        vx = vx * (1 + modulation * torch.sin(2 * pi * X + 0.7))
        vy = vy * (1 + modulation * torch.sin(2 * pi * Y + 1.9))
    Bx = -B0 * torch.sin(2 * pi * Yf_x)
    By = B0 * torch.sin(4 * pi * Xf_y)
    if noise:
        gen = torch.Generator(device="cpu").manual_seed(seed)
        jit = [1 + noise * torch.randn((n, n), generator=gen, dtype=torch.float64).to(device) for _ in range(4)]
        rho, p, vx, vy = rho * jit[0], p * jit[1], vx * jit[2], vy * jit[3]
        # Az at the corners (i-1/2, j-1/2): Bx += dAz/dy, By -= dAz/dx
        Az = noise * B0 / n * torch.randn((n, n), generator=gen, dtype=torch.float64).to(device)
        Bx = Bx + (torch.roll(Az, -1, dims=0) - Az) * n
        By = By - (torch.roll(Az, -1, dims=1) - Az) * n
    v = torch.stack([vx, vy, torch.zeros_like(X)]) * vscale
    Bx, By = Bx.contiguous(), By.contiguous()
    return rho, v, p, Bx, By


def fast_speed(W, gamma, norm):
    ca2 = (W[5] * W[5] + W[6] * W[6] + W[7] * W[7]) / W[0]
    a2 = gamma * W[4] / W[0]
    can2 = W[norm] * W[norm] / W[0]
    return torch.sqrt(0.5 * (ca2 + a2 + torch.sqrt((ca2 - a2) ** 2 + 4 * a2 * (ca2 - can2))))


def cfl_dt(U, gamma, dx, cfl):
    """Proper CFL bound max(|u|+cf_x, |v|+cf_y) of a state (frozen-dt set-up)."""
    rho = U[0]
    W = torch.stack([rho, U[1] / rho, U[2] / rho, U[3] / rho,
                     (gamma - 1) * (U[4] - 0.5 * (U[1] ** 2 + U[2] ** 2 + U[3] ** 2) / rho
                                    - 0.5 * (U[5] ** 2 + U[6] ** 2 + U[7] ** 2)), U[5], U[6], U[7]])
    lam = torch.maximum(W[1].abs() + fast_speed(W, gamma, 5), W[2].abs() + fast_speed(W, gamma, 6))
    return float(cfl * dx / lam.max())
