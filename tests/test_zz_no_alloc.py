"""amr_adj_step_noalloc (the allocation-free frozen-mesh step) against
amr_adj_step on a context, through the C ABI: bit for bit equal on two calls
in a row, the second on a smaller grid (the shared static work memory keeps
the first call's contents), for both solvers; sizes beyond its static work
memory are rejected without touching the outputs."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GAMMA = 5.0 / 3.0


def _inputs(n, seed):
    from amr2d_mhd_amd import adjoint as AD

    rho, v, p, Bx, By = AD.orszag_tang_controls(n, vscale=1.0, modulation=0.05, noise=1e-3, seed=seed)
    U = AD.cons_from_controls(rho, v, p, Bx, By, GAMMA).contiguous()
    dt = AD.cfl_dt(U, GAMMA, 1.0 / n, 0.4)
    return U, Bx.contiguous(), By.contiguous(), dt


def _with_context(L, n, flux, U, Bx, By, dt, stream):
    import torch

    h = C.c_void_p()
    assert L.amr_adj_create(n, 1.0 / n, GAMMA, flux.encode(), C.byref(h)) == 0
    Uo, Bxo, Byo = torch.empty_like(U), torch.empty_like(Bx), torch.empty_like(By)
    rc = L.amr_adj_step(h, U.data_ptr(), Bx.data_ptr(), By.data_ptr(), dt, Uo.data_ptr(), Bxo.data_ptr(),
                        Byo.data_ptr(), stream)
    torch.cuda.synchronize()
    assert rc == 0
    assert L.amr_adj_destroy(h) == 0
    return Uo.cpu().numpy(), Bxo.cpu().numpy(), Byo.cpu().numpy()


def _no_alloc(L, n, flux, U, Bx, By, dt, stream):
    import torch

    Uo, Bxo, Byo = torch.full_like(U, np.nan), torch.full_like(Bx, np.nan), torch.full_like(By, np.nan)
    rc = L.amr_adj_step_noalloc(n, 1.0 / n, GAMMA, flux.encode(), U.data_ptr(), Bx.data_ptr(), By.data_ptr(), dt,
                                Uo.data_ptr(), Bxo.data_ptr(), Byo.data_ptr(), stream)
    torch.cuda.synchronize()
    assert rc == 0
    return Uo.cpu().numpy(), Bxo.cpu().numpy(), Byo.cpu().numpy()


@pytest.mark.parametrize("flux", ["HLLE", "HLLD"])
def test_no_alloc_step_equals_context_step(flux):
    import torch

    from amr2d_mhd_amd import lib

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    L = lib()
    stream = torch.cuda.current_stream().cuda_stream
    for call, (n, seed) in enumerate([(256, 11), (64, 12)]):
        U, Bx, By, dt = _inputs(n, seed)
        ref = _with_context(L, n, flux, U, Bx, By, dt, stream)
        got = _no_alloc(L, n, flux, U, Bx, By, dt, stream)
        for name, a, b in zip(("U", "Bx", "By"), got, ref):
            assert np.array_equal(a, b), f"{flux} call {call} (n={n}): {name} differs"


def test_no_alloc_rejects_oversized_grid():
    import torch

    from amr2d_mhd_amd import lib

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    L = lib()
    n = 512
    U = torch.ones((8, n, n), dtype=torch.float64, device="cuda")
    Bx = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    By = torch.zeros_like(Bx)
    Uo = torch.full_like(U, 7.0)
    Bxo, Byo = torch.full_like(Bx, 7.0), torch.full_like(By, 7.0)
    rc = L.amr_adj_step_noalloc(n, 1.0 / n, GAMMA, b"HLLE", U.data_ptr(), Bx.data_ptr(), By.data_ptr(), 1e-4,
                                Uo.data_ptr(), Bxo.data_ptr(), Byo.data_ptr(), None)
    torch.cuda.synchronize()
    assert rc == 10
    assert bool((Uo == 7.0).all()) and bool((Bxo == 7.0).all()) and bool((Byo == 7.0).all())
