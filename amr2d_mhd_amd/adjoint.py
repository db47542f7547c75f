"""Discrete adjoint of the frozen-mesh step, behind a torch.autograd.Function.

The forward is the reference's RK2 step (minmod MUSCL + HLLE or HLLD, CT) on one
uniform periodic N x N grid at a frozen dt sequence (``amr_adj_step``); the
backward runs the hand-written reverse kernels (``amr_adj_vjp_tape``) over
per-step tapes (the intermediates of one step, 24 N^2 doubles).  With
``every`` > 0 the forward keeps the state every ``every`` steps and the
backward recomputes each segment from its checkpoint, writing the segment's
tapes, before sweeping it in reverse; with ``every`` = 0 the forward writes
the tape of every step and the backward recomputes nothing (memory: steps x
24 N^2 doubles).  Tensors are float64 on the GPU; kernels run on torch's
current stream.

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
    """One uniform periodic N x N grid of cell size dx; flux "HLLE" or "HLLD"."""

    def __init__(self, n, dx, gamma, flux="HLLE"):
        self.n, self.dx, self.gamma, self.flux = int(n), float(dx), float(gamma), flux
        h = C.c_void_p()
        rc = lib().amr_adj_create(self.n, self.dx, self.gamma, flux.encode(), C.byref(h))
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

    def tape_doubles(self):
        return int(lib().amr_adj_tape_doubles(self.h))

    def new_tapes(self, count, device):
        return torch.empty((count, self.tape_doubles()), dtype=torch.float64, device=device)

    def step(self, U, Bx, By, dt, tape=None):
        """One RK2 step; returns new (U, Bx, By).  `tape`, a float64 CUDA
        tensor of tape_doubles() elements, receives the step's tape."""
        self._check(U, Bx, By)
        if tape is not None:
            self._check(tape)
            if tape.numel() < self.tape_doubles():
                raise ValueError("tape too small")
        Uo, Bxo, Byo = torch.empty_like(U), torch.empty_like(Bx), torch.empty_like(By)
        rc = lib().amr_adj_step_tape(self.h, _ptr(U), _ptr(Bx), _ptr(By), float(dt), _ptr(Uo), _ptr(Bxo),
                                     _ptr(Byo), _ptr(tape) if tape is not None else None, _stream())
        if rc:
            raise AmrError(rc)
        return Uo, Bxo, Byo

    def vjp(self, tape, Bx, By, dt, Uob, Bxob, Byob):
        """Adjoint of the step whose tape is `tape` and input faces (Bx, By)."""
        self._check(tape, Bx, By, Uob, Bxob, Byob)
        if tape.numel() < self.tape_doubles():
            raise ValueError("tape too small")
        Ub, Bxb, Byb = torch.empty_like(Uob), torch.empty_like(Bxob), torch.empty_like(Byob)
        rc = lib().amr_adj_vjp_tape(self.h, _ptr(tape), _ptr(Bx), _ptr(By), float(dt), _ptr(Uob), _ptr(Bxob),
                                    _ptr(Byob), _ptr(Ub), _ptr(Bxb), _ptr(Byb), _stream())
        if rc:
            raise AmrError(rc)
        return Ub, Bxb, Byb

    def step_vjp(self, U, Bx, By, dt, Uob, Bxob, Byob):
        """Adjoint of one step taken from (U, Bx, By): J^T (Uob, Bxob, Byob),
        recomputing the step first."""
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
        every = int(every)
        ckpts, tapes = [], None
        if every == 0 and len(dts):
            tapes = mesh.new_tapes(len(dts), U.device)
        for s, dt in enumerate(dts):
            if every == 0:
                ckpts.append((Bx, By))  # input faces of every step
            elif s % every == 0:
                ckpts.append((U, Bx, By))
            U, Bx, By = mesh.step(U, Bx, By, dt, tapes[s] if tapes is not None else None)
        ctx.mesh, ctx.dts, ctx.every, ctx.ckpts, ctx.tapes = mesh, list(dts), every, ckpts, tapes
        return U, Bx, By

    @staticmethod
    def backward(ctx, gU, gBx, gBy):
        mesh, dts, every = ctx.mesh, ctx.dts, ctx.every
        if not dts:
            return gU, gBx, gBy, None, None, None
        ref = ctx.ckpts[0][-1]
        gU = torch.zeros((8,) + tuple(ref.shape), dtype=ref.dtype, device=ref.device) if gU is None \
            else gU.contiguous()
        gBx = torch.zeros_like(ref) if gBx is None else gBx.contiguous()
        gBy = torch.zeros_like(ref) if gBy is None else gBy.contiguous()
        if every == 0:
            for s in range(len(dts) - 1, -1, -1):
                Bx, By = ctx.ckpts[s]
                gU, gBx, gBy = mesh.vjp(ctx.tapes[s], Bx, By, dts[s], gU, gBx, gBy)
        else:
            tapes = mesh.new_tapes(min(every, len(dts)), ref.device)
            for seg in range(len(ctx.ckpts) - 1, -1, -1):
                s0 = seg * every
                s1 = min(s0 + every, len(dts))
                # recompute the segment, keeping each step's input faces and tape
                U, Bx, By = ctx.ckpts[seg]
                faces = []
                for s in range(s0, s1):
                    faces.append((Bx, By))
                    U, Bx, By = mesh.step(U, Bx, By, dts[s], tapes[s - s0])
                for s in range(s1 - 1, s0 - 1, -1):
                    gU, gBx, gBy = mesh.vjp(tapes[s - s0], *faces[s - s0], dts[s], gU, gBx, gBy)
        ctx.ckpts = ctx.tapes = None
        return gU, gBx, gBy, None, None, None


def rollout(U0, Bx0, By0, mesh, dts, every=20):
    """Differentiable frozen-dt rollout; `every` = 0 tapes every step in the
    forward instead of checkpointing."""
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
