"""ctypes binding of the CPU oracle (oracle/liboracle.so) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module, and only as the checker.  The product package never imports it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

U, US, W, FF, GG, BI, BIS, EMF = range(8)
FIELD_NAMES = ["U", "Us", "W", "ff", "gg", "Bi", "Bis", "EMF"]
NVAR = [8, 8, 8, 8, 8, 2, 2, 1]
FILL_CC, FILL_FC, FILL_EC = 0, 1, 2


class OraConfig(C.Structure):
    _fields_ = [("nx", C.c_int), ("nghosts", C.c_int), ("maxlevs", C.c_int), ("lenx", C.c_double),
                ("cfl", C.c_double), ("ic", C.c_char * 16), ("limiter", C.c_char * 16), ("flux", C.c_char * 8),
                ("ic_params", C.c_double * 8), ("periodic", C.c_int), ("rk_stages", C.c_int),
                ("ct_type", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int32)
        L.ora_create.restype = P
        L.ora_create.argtypes = [C.POINTER(OraConfig)]
        L.ora_destroy.argtypes = [P]
        for n in ["ora_init", "ora_step", "ora_done", "ora_locate_bounds", "ora_regrid", "ora_tile"]:
            getattr(L, n).argtypes = [P]
            getattr(L, n).restype = C.c_int
        L.ora_run.argtypes = [P, C.c_int, ip]
        L.ora_info.argtypes = [P, ip, dp]
        L.ora_hash_fields.argtypes = [P]
        L.ora_hash_fields.restype = C.c_uint64
        L.ora_hash_tree.argtypes = [P]
        L.ora_hash_tree.restype = C.c_uint64
        L.ora_get_field.argtypes = [P, C.c_int, C.c_int, dp]
        L.ora_set_field.argtypes = [P, C.c_int, C.c_int, dp]
        L.ora_get_tree.argtypes = [P, C.c_int, ip, ip, ip]
        L.ora_set_tree.argtypes = [P, C.c_int, ip, ip, ip]
        L.ora_set_time.argtypes = [P, C.c_double, C.c_double, C.c_double, C.c_int]
        L.ora_fill.argtypes = [P, C.c_int, C.c_int]
        L.ora_stage.argtypes = [P, C.c_int]
        L.ora_ct.argtypes = [P, C.c_int]
        L.ora_cfl.argtypes = [P, dp]
        L.ora_nbounds.argtypes = [P, ip]
        L.ora_last_error.argtypes = [P, ip]
        L.ora_limiter.restype = C.c_double
        L.ora_limiter.argtypes = [C.c_char_p, C.c_double, C.c_double]
        L.ora_riemann.argtypes = [C.c_char_p, C.c_char, C.c_double, dp, dp, dp]
        L.ora_wavespeeds.argtypes = [dp, C.c_double, C.c_int, dp]
        L.ora_recon_tile.argtypes = [C.c_int, C.c_int, C.c_char_p, C.c_double, dp, dp, dp, dp, dp]
        L.ora_muscl_tile.argtypes = [C.c_int, C.c_int, C.c_char_p, C.c_char_p, C.c_double, C.c_double, C.c_int,
                                     dp, dp, dp, dp]
        L.ora_regrid_tiles.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, dp, dp]
        L.ora_criteria_tile.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, dp, ip]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


class Domain:
    """One oracle simulation (the reference's meshblock + main loop)."""

    def __init__(self, ic="BWx", limiter="MC", flux="RUSA", cfl=0.4, nx=16, nghosts=2, maxlevs=4, lenx=1.0,
                 ic_params=None, periodic=False, rk_stages=2, ct_type=1):
        cfg = OraConfig(nx, nghosts, maxlevs, lenx, cfl, ic.encode(), limiter.encode(), flux.encode())
        cfg.periodic = 1 if periodic else 0
        cfg.rk_stages = int(rk_stages)
        cfg.ct_type = int(ct_type)
        if ic_params is not None:
            for q, v in enumerate(ic_params):
                cfg.ic_params[q] = float(v)
        self.h = lib().ora_create(C.byref(cfg))
        if not self.h:
            raise ValueError("bad oracle configuration")
        self.NT = nx + 2 * nghosts

    def __del__(self):
        if getattr(self, "h", None):
            lib().ora_destroy(self.h)
            self.h = None

    def init(self):
        return lib().ora_init(self.h)

    def step(self):
        return lib().ora_step(self.h)

    def done(self):
        return bool(lib().ora_done(self.h))

    def info(self):
        iv = np.zeros(4, np.int32)
        dv = np.zeros(4, np.float64)
        lib().ora_info(self.h, _ip(iv), _dp(dv))
        return dict(lastActive=int(iv[0]), nleaf=int(iv[1]), count=int(iv[2]), maxblocks=int(iv[3]),
                    t=float(dv[0]), dt=float(dv[1]), dt_next=float(dv[2]), gamma=float(dv[3]))

    def hash_fields(self):
        return int(lib().ora_hash_fields(self.h))

    def hash_tree(self):
        return int(lib().ora_hash_tree(self.h))

    def get_field(self, f, n=None):
        n = self.info()["lastActive"] if n is None else n
        out = np.zeros((n, NVAR[f], self.NT, self.NT), np.float64)
        lib().ora_get_field(self.h, f, n, _dp(out))
        return out

    def set_field(self, f, arr):
        arr = np.ascontiguousarray(arr, np.float64)
        return lib().ora_set_field(self.h, f, arr.shape[0], _dp(arr))

    def get_tree(self, n=None):
        n = self.info()["lastActive"] if n is None else n
        lp = np.zeros((n, 8), np.int32)
        ic = np.zeros((n, 2), np.int32)
        lf = np.zeros(n, np.int32)
        lib().ora_get_tree(self.h, n, _ip(lp), _ip(ic), _ip(lf))
        return lp, ic, lf

    def set_tree(self, lp, ic, lf):
        lp = np.ascontiguousarray(lp, np.int32)
        ic = np.ascontiguousarray(ic, np.int32)
        lf = np.ascontiguousarray(lf, np.int32)
        return lib().ora_set_tree(self.h, lp.shape[0], _ip(lp), _ip(ic), _ip(lf))

    def set_time(self, t, dt, dt_next, count):
        lib().ora_set_time(self.h, t, dt, dt_next, count)

    def load_state(self, st):
        """Load a reference state dump (see refstate.read_state)."""
        self.set_tree(st["lp"], st["icoord"], st["leafs"])
        for name, f in zip(FIELD_NAMES, range(8)):
            self.set_field(f, st[name])
        self.set_time(st["t"], st["dt"], 0.0, 0)

    def last_error(self):
        e = np.zeros(4, np.int32)
        lib().ora_last_error(self.h, _ip(e))
        return e.tolist()

    def locate_bounds(self):
        lib().ora_locate_bounds(self.h)
        n = lib().ora_nbounds(self.h, None)
        ib = np.zeros((n, 3), np.int32)
        lib().ora_nbounds(self.h, _ip(ib))
        return ib

    def fill(self, f, kind):
        return lib().ora_fill(self.h, f, kind)

    def stage(self, step):
        return lib().ora_stage(self.h, step)

    def ct(self, step):
        return lib().ora_ct(self.h, step)

    def cfl(self):
        lam = C.c_double(0)
        st = lib().ora_cfl(self.h, C.byref(lam))
        return st, lam.value

    def regrid(self):
        return lib().ora_regrid(self.h)


def limiter(name, r, eta):
    L = lib()
    return np.array([L.ora_limiter(name.encode(), float(a), float(b)) for a, b in zip(r, eta)])


def riemann(method, direction, gamma, wL, wR):
    L = lib()
    wL = np.ascontiguousarray(wL, np.float64)
    wR = np.ascontiguousarray(wR, np.float64)
    out = np.zeros_like(wL)
    st = np.zeros(len(wL), np.int32)
    for q in range(len(wL)):
        st[q] = L.ora_riemann(method.encode(), direction.encode(), gamma, _dp(wL[q]), _dp(wR[q]), _dp(out[q]))
    return out, st


def wavespeeds(w, gamma, norm):
    L = lib()
    w = np.ascontiguousarray(w, np.float64)
    out = np.zeros((len(w), 4))
    for q in range(len(w)):
        L.ora_wavespeeds(_dp(w[q]), gamma, norm, _dp(out[q]))
    return out


def recon_tile(W, lim, ng, dx=1.0):
    NT = W.shape[-1]
    W = np.ascontiguousarray(W, np.float64)
    outs = [np.zeros((8, NT, NT)) for _ in range(4)]
    st = lib().ora_recon_tile(NT, ng, lim.encode(), dx, _dp(W), *[_dp(o) for o in outs])
    return outs, st


def muscl_tile(U, B, lim, flux, gamma, step, ng, dx=1.0):
    NT = U.shape[-1]
    U = np.ascontiguousarray(U, np.float64)
    B = np.ascontiguousarray(B, np.float64)
    ff = np.zeros((8, NT, NT))
    gg = np.zeros((8, NT, NT))
    st = lib().ora_muscl_tile(NT, ng, lim.encode(), flux.encode(), gamma, dx, step, _dp(U), _dp(B), _dp(ff), _dp(gg))
    return ff, gg, st


def regrid_tiles(U5, B5, op, gamma, ng):
    NT = U5.shape[-1]
    U5 = np.ascontiguousarray(U5, np.float64).copy()
    B5 = np.ascontiguousarray(B5, np.float64).copy()
    st = lib().ora_regrid_tiles(NT, ng, op, gamma, _dp(U5), _dp(B5))
    return U5, B5, st


def criteria_tile(W, ng, maxlevs, level):
    NT = W.shape[-1]
    W = np.ascontiguousarray(W, np.float64)
    o = np.zeros(2, np.int32)
    lib().ora_criteria_tile(NT, ng, maxlevs, level, _dp(W), _ip(o))
    return int(o[0]), int(o[1])
