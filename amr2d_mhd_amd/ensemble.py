"""Seeded ensembles of parameterised rotors advanced together in one context.

The ensemble is a build-defined workload (the reference runs one simulation
per process, with its rotor hard-coded at IC2Dtype.cpp:75-111): instance i
is a rotor whose inner density, spin speed and field strength, and whose AMR
depth, are drawn from a counter-based generator keyed by (seed, i), so any
subset of instances can be rebuilt alone with identical parameters.

Generator: Philox-4x32 with 10 rounds (Salmon, Moraes, Dror, Shaw, "Parallel
random numbers: as easy as 1, 2, 3", SC 2011), counter (i, 0, 0, 0), key
(seed mod 2^32, seed >> 32).  The four output words give, through
u = (x + 0.5) / 2^32:
    rho_in = 5 + 15 u0          (U[5, 20])
    v0     = 1 + 2 u1           (U[1, 3])
    Bx     = (0.5 + 1.5 u2) / sqrt(4 pi)
    maxlev = 3 + floor(4 u3)    (U{3, 4, 5, 6})
Fixed: p = 5, taper from r0 = 0.1 to r1 = 0.115, gamma = 1.4, t_end = 0.2,
MC-limited MUSCL, HLLC for even i and HLLD for odd i, CFL 0.5.

The pressure is a choice of this workload.  With p = 1 the reference
algorithm itself aborts 146 of the 256 rotors within their first three
steps (the slope-limited prolongation of conserved variables at the first
refinements gives p <= 0 across the taper when the rotor's kinetic energy
is large against p); with p = 2 it aborts 57, with p = 5 one, with p = 10
eight.  An instance that aborts stops alone; the others keep running.
"""
import math

SEED = 20261202
N_INSTANCES = 256

_M0, _M1 = 0xD2511F53, 0xCD9E8D57
_W0, _W1 = 0x9E3779B9, 0xBB67AE85
_MASK = 0xFFFFFFFF


def philox4x32(ctr, key, rounds=10):
    """Philox-4x32-R block function on a 4-word counter and 2-word key."""
    c0, c1, c2, c3 = (x & _MASK for x in ctr)
    k0, k1 = (x & _MASK for x in key)
    for r in range(rounds):
        if r:
            k0 = (k0 + _W0) & _MASK
            k1 = (k1 + _W1) & _MASK
        p0 = _M0 * c0
        p1 = _M1 * c2
        c0, c1, c2, c3 = ((p1 >> 32) ^ c1 ^ k0, p1 & _MASK, (p0 >> 32) ^ c3 ^ k1, p0 & _MASK)
    return c0, c1, c2, c3


def _u01(x):
    return (x + 0.5) / 4294967296.0


def rotor_params(i, seed=SEED):
    """(rho_in, v0, Bx, maxlevs) of instance i."""
    x = philox4x32((i, 0, 0, 0), (seed & _MASK, (seed >> 32) & _MASK))
    u = [_u01(v) for v in x]
    rho_in = 5.0 + 15.0 * u[0]
    v0 = 1.0 + 2.0 * u[1]
    bx = (0.5 + 1.5 * u[2]) / math.sqrt(4.0 * math.pi)
    maxlevs = 3 + min(3, int(4.0 * u[3]))
    return rho_in, v0, bx, maxlevs


def instance_config(i, seed=SEED, p=5.0, t_end=0.2, cfl=0.5, nx=16):
    """Simulation config (the dict accepted by Simulation) of instance i."""
    rho_in, v0, bx, maxlevs = rotor_params(i, seed)
    return dict(ic="rotor_p", limiter="MC", flux="HLLC" if i % 2 == 0 else "HLLD", cfl=cfl, nx=nx, nghosts=2,
                maxlevs=maxlevs, lenx=1.0, ic_params=[rho_in, v0, p, bx, 0.1, 0.115, 1.4, t_end])


def ensemble_configs(n=N_INSTANCES, seed=SEED, **kw):
    return [instance_config(i, seed, **kw) for i in range(n)]
