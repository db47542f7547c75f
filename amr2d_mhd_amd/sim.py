"""Python side of the AMR2D-MHD engine: a thin owner of an ``amr_ctx``.

Mirrors the reference's driver vocabulary (main.cpp): a run is configured by
the input.txt triple (IC, limiter, flux), CFL and the compile-time nghosts /
maxlevs, advances with ``step()`` (one main-loop iteration) and exposes the
meshblock arrays (U, Us, Bi, Bis, ff, gg, EMF) per slot for inspection.
"""
import ctypes as C

import numpy as np

from ._lib import FIELD_VARS, FIELDS, FILL_KIND, SOLVERS, AmrError, Diag, InstInfo, lib, make_config

# stage-kernel variants of the library (kStageVariants in csrc/kernels.hpp)
STAGE_VARIANTS = 8


def _check(ctx, rc):
    if rc != 0:
        d = Diag()
        if ctx:
            lib().amr_last_error(ctx, C.byref(d))
        raise AmrError(rc, d if ctx else None)


class Simulation:
    """One or more independent AMR simulations advanced together on the GPU."""

    def __init__(self, configs, slot_capacity=0):
        if isinstance(configs, dict):
            configs = [configs]
        cfgs = (type(make_config()) * len(configs))(*[make_config(**c) for c in configs])
        self.n = len(configs)
        self.nx = configs[0].get("nx", 16)
        self.ng = configs[0].get("nghosts", 2)
        self.NT = self.nx + 2 * self.ng
        h = C.c_void_p()
        rc = lib().amr_create(self.n, cfgs, int(slot_capacity), C.byref(h))
        if rc != 0:
            raise AmrError(rc)
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().amr_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---------------------------------------------------------------- driver
    def init(self):
        _check(self.h, lib().amr_init(self.h))

    def step(self):
        n = C.c_int(0)
        _check(self.h, lib().amr_step(self.h, C.byref(n)))
        return n.value

    def run(self, max_steps=-1):
        """Advance up to max_steps steps (-1: until every instance is done);
        returns (steps, zone-updates).  On failure the raised AmrError carries
        the steps and zone-updates completed before it."""
        steps, zu = C.c_int64(0), C.c_int64(0)
        try:
            _check(self.h, lib().amr_run(self.h, int(max_steps), C.byref(steps), C.byref(zu)))
        except AmrError as e:
            e.steps, e.zone_updates = steps.value, zu.value
            raise
        return steps.value, zu.value

    def info(self, inst=0):
        o = InstInfo()
        _check(self.h, lib().amr_info(self.h, inst, C.byref(o)))
        return {f: getattr(o, f) for f, _ in InstInfo._fields_}

    # ----------------------------------------------------------------- state
    def download(self, field, inst=0, nslots=None):
        nslots = self.info(inst)["last_active"] if nslots is None else nslots
        out = np.zeros((nslots, FIELD_VARS[field], self.NT, self.NT), np.float64)
        _check(self.h, lib().amr_download(self.h, inst, FIELDS[field], nslots, out.ctypes.data))
        return out

    def upload(self, field, arr, inst=0):
        arr = np.ascontiguousarray(arr, np.float64)
        _check(self.h, lib().amr_upload(self.h, inst, FIELDS[field], arr.shape[0], arr.ctypes.data))

    def get_tree(self, inst=0, nslots=None):
        nslots = self.info(inst)["last_active"] if nslots is None else nslots
        lp = np.zeros((nslots, 8), np.int32)
        ic = np.zeros((nslots, 2), np.int32)
        lf = np.zeros(nslots, np.int32)
        _check(self.h, lib().amr_get_tree(self.h, inst, nslots, lp.ctypes.data, ic.ctypes.data, lf.ctypes.data))
        return lp, ic, lf

    def set_tree(self, lp, ic, lf, inst=0):
        lp = np.ascontiguousarray(lp, np.int32)
        ic = np.ascontiguousarray(ic, np.int32)
        lf = np.ascontiguousarray(lf, np.int32)
        _check(self.h, lib().amr_set_tree(self.h, inst, lp.shape[0], lp.ctypes.data, ic.ctypes.data, lf.ctypes.data))

    def set_time(self, t, dt, dt_next=0.0, count=0, inst=0):
        _check(self.h, lib().amr_set_time(self.h, inst, t, dt, dt_next, count))

    def load_state(self, st, inst=0):
        """Load a full state (tree, all slot fields, t and the dt lag) given as
        a dict of arrays, as returned by state(); dt_next and count default to
        0 for states that carry only t and dt (the reference's own dumps)."""
        self.set_tree(st["lp"], st["icoord"], st["leafs"], inst)
        for f in FIELDS:
            self.upload(f, st[f], inst)
        self.set_time(st["t"], st["dt"], st.get("dt_next", 0.0), st.get("count", 0), inst)

    def state(self, inst=0):
        """Tree, every slot field of every allocated block, t, the dt applied
        by the current step, the dt computed for the next one and the step
        count: everything the following steps read."""
        lp, ic, lf = self.get_tree(inst)
        inf = self.info(inst)
        st = dict(lp=lp, icoord=ic, leafs=lf, t=inf["t"], dt=inf["dt"], dt_next=inf["dt_next"], count=inf["count"])
        for f in FIELDS:
            st[f] = self.download(f, inst)
        return st

    def save_checkpoint(self, path, inst=0):
        """Write a binary checkpoint of one instance (amr_checkpoint_save)."""
        _check(self.h, lib().amr_checkpoint_save(self.h, inst, str(path).encode()))

    def load_checkpoint(self, path, inst=0):
        """Restart one instance from a checkpoint of the same configuration."""
        _check(self.h, lib().amr_checkpoint_load(self.h, inst, str(path).encode()))

    # ---------------------------------------------------------------- phases
    def fill(self, field, kind):
        _check(self.h, lib().amr_fill(self.h, FIELDS[field], FILL_KIND[kind]))

    def stage(self, step):
        _check(self.h, lib().amr_stage(self.h, step))

    def ct(self, step):
        _check(self.h, lib().amr_ct(self.h, step))

    def cfl(self, inst=0):
        lam = C.c_double(0)
        _check(self.h, lib().amr_cfl(self.h, inst, C.byref(lam)))
        return lam.value

    def regrid(self):
        _check(self.h, lib().amr_regrid(self.h))

    def schedule(self, inst=0):
        n = C.c_int(0)
        _check(self.h, lib().amr_schedule(self.h, inst, 0, None, C.byref(n)))
        ib = np.zeros((n.value, 3), np.int32)
        _check(self.h, lib().amr_schedule(self.h, inst, n.value, ib.ctypes.data, C.byref(n)))
        return ib

    # --------------------------------------------------------------- timing
    PROFILE_FAMILIES = ("stage", "fill", "ct", "cfl", "regrid", "host_topology", "host_regrid", "host_wait",
                        "host_compile", "graph", "step", "graph_capture")

    def set_step_kernel(self, mode):
        """Persistent step kernel: 1 (256 lanes) or 2 (512 lanes) runs each step
        as one launch, 0 launches per phase, -1 restores the default."""
        _check(self.h, lib().amr_set_step_kernel(self.h, int(mode)))

    def set_regrid_device(self, on=True):
        """Decide each regrid on the device (the default) or on the host from
        downloaded per-leaf indicators; the trees and results are identical."""
        _check(self.h, lib().amr_set_regrid_device(self.h, 1 if on else 0))

    def fill_stats(self, inst=0):
        """Op mix of the instance's current ghost-fill program: copies, 2x2
        restrictions, general restrictions, padding ops, explicit-operand entries."""
        out = np.zeros(5, np.int64)
        _check(self.h, lib().amr_debug_fill_stats(self.h, inst, out.ctypes.data))
        return dict(zip(("copy", "restrict", "general", "nop", "gen_entries"), out.tolist()))

    def set_regrid_lds(self, on=True):
        """Device regrid: stage trees that fit in LDS (the default), or work on
        every tree in place in device memory (same results)."""
        _check(self.h, lib().amr_debug_regrid_lds(self.h, 1 if on else 0))

    def device_tree(self, inst=0, nslots=None):
        """The device copy of the instance's tree that the device regrid works
        on: (lp[n, 8], leafs[n], current), current False when the copy is
        stale (host regrid, or a tree set since the last step)."""
        n = self.info(inst)["last_active"] if nslots is None else nslots
        lp = np.zeros((n, 8), np.int32)
        lf = np.zeros(n, np.int32)
        cur = C.c_int(0)
        _check(self.h, lib().amr_debug_device_tree(self.h, inst, n, lp.ctypes.data, lf.ctypes.data, C.byref(cur)))
        return lp, lf, bool(cur.value)

    def use_stream(self, stream):
        """Run the context on `stream` (a torch.cuda.Stream or a raw hipStream_t
        handle as an int), or on its own stream again when `stream` is None."""
        if stream is None:
            _check(self.h, lib().amr_use_stream(self.h, None, 0))
        else:
            handle = stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream)
            _check(self.h, lib().amr_use_stream(self.h, C.c_void_p(handle), 1))

    def stage_flops(self, inst=0):
        """Operations per zone of one RK2 stage (mean of the two stages) on the
        instance's current leaves, counted on the host by the op-counting
        build of the stage arithmetic: {add, mul, div, sqrt, total}."""
        out = np.zeros(5)
        _check(self.h, lib().amr_debug_stage_flops(self.h, inst, out.ctypes.data))
        return dict(zip(("add", "mul", "div", "sqrt", "total"), out.tolist()))

    def set_graph(self, on=True):
        """Replay each step's device work as a HIP graph, or launch it directly (the default)."""
        _check(self.h, lib().amr_set_graph(self.h, 1 if on else 0))

    def set_fill_cache(self, entries):
        """Fill programs kept per instance for recurring trees (default 4; 0 = compile on every change)."""
        _check(self.h, lib().amr_set_fill_cache(self.h, int(entries)))

    def set_stage_variant(self, variant):
        """Select the stage-kernel variant (0 to STAGE_VARIANTS - 1, identical results; -1 = tuned per
        leaf-count bucket)."""
        _check(self.h, lib().amr_set_stage_variant(self.h, int(variant)))

    def stage_tuning(self, nbuckets=24):
        """Tuner decisions: {bucket: (chosen variant, [mean us per leaf of each variant])}."""
        chosen = np.zeros(nbuckets, np.int32)
        us = np.zeros((nbuckets, STAGE_VARIANTS))
        _check(self.h, lib().amr_stage_tuning(self.h, nbuckets, chosen.ctypes.data, us.ctypes.data))
        return {b: (int(chosen[b]), us[b].tolist()) for b in range(nbuckets) if chosen[b] >= 0 or us[b].any()}

    def set_stage_tuning(self, chosen):
        """Import tuner decisions {bucket: variant} (as stage_tuning() returns
        them, or plain ints): those buckets run no trial launches."""
        arr = np.full(24, -1, np.int32)
        for b, v in chosen.items():
            arr[int(b)] = v[0] if isinstance(v, (tuple, list)) else int(v)
        _check(self.h, lib().amr_set_stage_tuning(self.h, 24, arr.ctypes.data))

    def profile(self, families=PROFILE_FAMILIES, stride=1):
        """Start HIP-event timing of the given kernel families (resets totals);
        stride n > 1 times one launch in n of each family."""
        _check(self.h, lib().amr_profile_sample(self.h, int(stride)))
        mask = 0
        for f in families:
            mask |= 1 << self.PROFILE_FAMILIES.index(f)
        _check(self.h, lib().amr_profile_enable(self.h, mask))

    def profile_read(self):
        out = {}
        for q, nm in enumerate(self.PROFILE_FAMILIES):
            ms, n = C.c_double(0), C.c_int64(0)
            _check(self.h, lib().amr_profile_read(self.h, q, C.byref(ms), C.byref(n)))
            out[nm] = (ms.value, n.value)
        return out


def run_group(sims, max_steps=-1):
    """Run several Simulations round robin (amr_run_group): while the host
    completes one's step, the steps launched on the others keep the GPU busy.
    Returns (largest step count, total zone-updates); on failure the raised
    AmrError carries both."""
    arr = (C.c_void_p * len(sims))(*[s.h for s in sims])
    steps, zu = C.c_int64(0), C.c_int64(0)
    rc = lib().amr_run_group(len(sims), arr, int(max_steps), C.byref(steps), C.byref(zu))
    if rc:
        e = AmrError(rc)
        e.steps, e.zone_updates = steps.value, zu.value
        raise e
    return steps.value, zu.value


def riemann_batch(method, direction, gamma, wL, wR):
    """riemannS on the GPU for torch cuda float64 tensors wL, wR of shape [n, 8]."""
    import torch

    assert wL.is_cuda and wL.dtype == torch.float64 and wL.shape[-1] == 8
    wL = wL.contiguous()
    wR = wR.contiguous()
    flux = torch.empty_like(wL)
    status = torch.empty(wL.shape[0], dtype=torch.int32, device=wL.device)
    stream = torch.cuda.current_stream(wL.device).cuda_stream
    rc = lib().amr_riemann_batch(SOLVERS[method], 0 if direction == "x" else 1, float(gamma), wL.shape[0],
                                 wL.data_ptr(), wR.data_ptr(), flux.data_ptr(), status.data_ptr(), stream)
    if rc != 0:
        raise AmrError(rc)
    return flux, status


def recon_batch(W, limiter, nghosts, dx=1.0):
    """Edge states of primitive tiles W [nt, 8, NT, NT] (torch cuda float64) ->
    [nt, 4, 8, NT, NT] (wxL, wxR, wyL, wyR); entries outside the reference's
    computed range stay zero."""
    import torch

    W = W.contiguous()
    nt, NT = W.shape[0], W.shape[-1]
    out = torch.zeros((nt, 4, 8, NT, NT), dtype=torch.float64, device=W.device)
    stream = torch.cuda.current_stream(W.device).cuda_stream
    rc = lib().amr_recon_batch(NT, nghosts, limiter.encode(), float(dx), nt, W.data_ptr(), out.data_ptr(), stream)
    if rc != 0:
        raise AmrError(rc)
    return out
