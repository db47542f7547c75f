"""MI355X-native AMR2D-MHD hot path (FP64 HIP kernels for gfx950).

Host side of the C ABI in include/amr2d_mhd.h: ``Simulation`` drives one or
many independent AMR simulations (the reference's main.cpp time loop) whose
state stays in HBM; ``riemann_batch``/``recon_batch`` expose the pointwise
kernels for parity checks.
"""
from ._lib import AmrError, FIELDS, FIELD_VARS, FILL_KIND, SOLVERS, lib, make_config  # noqa: F401
from .sim import Simulation, recon_batch, riemann_batch, run_group  # noqa: F401
