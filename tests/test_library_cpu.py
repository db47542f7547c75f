"""CPU-side checks of the product library: it loads, exports every entry point
declared in include/*.h, and its host-side fill-program compiler and coarsen
tables reproduce the reference on random trees (no GPU needed)."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

import oracle as O
import refrun

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(REPO, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(amr_[a-z_0-9]+)\s*\(", text))
    return sorted(names)


def test_library_exports_all_declared_symbols():
    from amr2d_mhd_amd._lib import lib

    L = lib()
    declared = _declared_symbols()
    assert len(declared) >= 20
    missing = [n for n in declared if not hasattr(L, n)]
    assert not missing, f"missing exports: {missing}"


def test_python_binding_covers_header():
    from amr2d_mhd_amd import _lib

    assert set(_declared_symbols()) <= set(_lib.EXPORTED)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from amr2d_mhd_amd import _lib

    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(RuntimeError):
        _lib.lib()


need_ref = pytest.mark.skipif(not refrun.have_ref(2, 5), reason="compiled reference (oracle/_ref) not built")


@need_ref
@pytest.mark.parametrize("general", [0, 1])
@pytest.mark.parametrize("ng,lv,seed,pr,pc,rounds", [(2, 4, 1, 0.5, 0.2, 5), (2, 6, 7, 0.3, 0.3, 8),
                                                     (4, 5, 3, 0.4, 0.25, 6), (2, 5, 99, 0.25, 0.5, 10)])
def test_fill_program_compiler_matches_reference(tmp_path, ng, lv, seed, pr, pc, rounds, general):
    """The symbolic gather program (host compiler + gather/scatter semantics,
    emulated on the host) equals the reference's sequential fills."""
    from amr2d_mhd_amd._lib import lib

    rc, err = refrun.run(str(tmp_path), ng, lv, ic="BWx", lim="MC", flux="RUSA", randtree=(seed, pr, pc, rounds))
    assert rc == 0, err
    pre = refrun.read_state(str(tmp_path / "state_rt_pre.bin"))
    post = refrun.read_state(str(tmp_path / "state_rt_fill.bin"))
    n, NT = pre["lastActive"], pre["NT"]
    for fld, nv, kind in [("U", 8, 0), ("Bi", 2, 1), ("ff", 8, 1), ("gg", 8, 1), ("EMF", 1, 2)]:
        Q = np.ascontiguousarray(pre[fld]).copy()
        nn = C.c_int(0)
        nops = lib().amr_debug_host_fill(NT, ng, n, pre["lp"].ctypes.data, pre["icoord"].ctypes.data,
                                         pre["leafs"].ctypes.data, Q.ctypes.data, nv, kind, C.byref(nn), general)
        assert nops > 0
        assert np.array_equal(Q, post[fld]), fld


def _compile(lp, ic, lf, NT, ng, ranges, threads, reuse):
    from amr2d_mhd_amd._lib import lib

    L, ngen = lib(), C.c_int(0)
    args = (NT, ng, len(lf), lp.ctypes.data, ic.ctypes.data, lf.ctypes.data, ranges, threads, reuse)
    m = L.amr_debug_compile_ranges(*args, 0, None, 0, None, C.byref(ngen))
    assert m > 0
    ops = np.zeros((m, 16), np.uint8)
    gen = np.zeros((max(1, ngen.value), 32), np.uint8)
    assert L.amr_debug_compile_ranges(*args, m, ops.ctypes.data, ngen.value, gen.ctypes.data, C.byref(ngen)) == m
    return ops, gen[:ngen.value]


@need_ref
@pytest.mark.parametrize("ng,lv,seed,pr,pc,rounds", [(2, 6, 7, 0.3, 0.3, 8), (4, 5, 3, 0.4, 0.25, 6),
                                                     (2, 5, 99, 0.25, 0.5, 10), (4, 5, 12, 0.35, 0.3, 8)])
def test_split_compile_equals_single_range(tmp_path, ng, lv, seed, pr, pc, rounds):
    """The fill program compiled with the schedule cut into ranges -- each run
    on its own table, later ranges' reads of what earlier ones wrote
    substituted afterwards, several threads, tables reused across compiles
    of different splits -- is the single-range program op for op (and the
    single-range one equals the reference's fills, test above)."""
    rc, err = refrun.run(str(tmp_path), ng, lv, ic="BWx", lim="MC", flux="RUSA", randtree=(seed, pr, pc, rounds))
    assert rc == 0, err
    pre = refrun.read_state(str(tmp_path / "state_rt_pre.bin"))
    lp, ic, lf = [np.ascontiguousarray(pre[k], dtype=np.int32) for k in ("lp", "icoord", "leafs")]
    NT = pre["NT"]
    ref_ops, ref_gen = _compile(lp, ic, lf, NT, ng, 1, 1, 0)
    linked = 0
    for ranges, threads, reuse in [(2, 1, 1), (3, 2, 1), (7, 4, 1), (16, 8, 1), (1, 1, 1), (5, 5, 0), (16, 16, 1)]:
        ops, gen = _compile(lp, ic, lf, NT, ng, ranges, threads, reuse)
        assert np.array_equal(ops, ref_ops), (ranges, threads, reuse)
        assert np.array_equal(gen, ref_gen), (ranges, threads, reuse)
    # the trees exercise the stale-read grouping (ops behind the first
    # padded chunk boundary after the plain ops)
    kinds = ref_ops.view(np.int16).reshape(-1, 8)[:, 3]
    assert (kinds == 3).sum() > 0  # padding exists
    assert len(ref_ops) % 256 == 0


@pytest.mark.parametrize("ranges,threads", [(1, 1), (5, 1), (16, 4), (16, 16)])
def test_no_exception_crosses_the_abi(ranges, threads):
    """A host allocation failure inside a range of the compile's thread pool
    (helper thread or caller) comes back as AMR_ERR_HOST, the pool stays
    usable, and every range runs when nothing throws."""
    from amr2d_mhd_amd._lib import lib

    L = lib()
    for throw_at in range(ranges):
        assert L.amr_debug_pool_throw(ranges, threads, throw_at) == 13
    assert L.amr_debug_pool_throw(ranges, threads, -1) == 0


def test_schedule_matches_oracle():
    """locate_bounds order of the product equals the oracle's on evolving trees."""
    from amr2d_mhd_amd._lib import lib

    d = O.Domain(ic="BWx", limiter="MC", flux="RUSA", cfl=0.4, nx=16, nghosts=2, maxlevs=4)
    d.init()
    for s in range(30):
        d.step()
        lp, ic, lf = d.get_tree()
        ib = d.locate_bounds()
        out = np.zeros((4 * len(lf), 3), np.int32)
        n = lib().amr_debug_schedule(len(lf), lp.ctypes.data, lf.ctypes.data, len(out), out.ctypes.data)
        assert np.array_equal(out[:n], ib), s


@need_ref
@pytest.mark.parametrize("ng", [2, 4])
def test_coarsen_maps_match_reference(ng):
    """Coarsen on the device reads per-cell source tables; emulate them on the
    host and compare with the reference's coarsen/coarsenBi_CT."""
    import struct

    from amr2d_mhd_amd._lib import lib

    NT = 16 + 2 * ng
    N2 = NT * NT
    mU = np.zeros(N2, np.int32)
    mB = np.zeros(2 * N2, np.int32)
    lib().amr_debug_coarsen_maps(NT, ng, mU.ctypes.data, mB.ctypes.data)
    rng = np.random.default_rng(ng)
    U5 = rng.uniform(0.5, 1.5, (5, 8, NT, NT))
    U5[:, 1:4] *= 0.1  # keep pressure positive (the reference throws on p <= 0)
    U5[:, 4] += 5.0
    B5 = rng.uniform(-1, 1, (5, 2, NT, NT))
    payload = struct.pack("<iiid", 1, NT, 1, 1.4)
    for s in range(5):
        payload += U5[s].tobytes() + B5[s].tobytes()
    raw = refrun.vec("regrid", payload, ng=ng, lv=5)
    rU = np.frombuffer(raw[:U5[0].nbytes], np.float64).reshape(8, NT, NT)
    rB = np.frombuffer(raw[U5[0].nbytes:U5[0].nbytes + B5[0].nbytes], np.float64).reshape(2, NT, NT)
    eU, eB = U5[0].copy(), B5[0].copy()
    for c in range(N2):
        i, j = c % NT, c // NT
        m = mU[c]
        if m >= 0:
            s, i1, j1 = m & 0xFF, (m >> 8) & 0xFF, (m >> 16) & 0xFF
            for k in range(8):
                Q = U5[s + 1, k]
                v = 0.0
                v += Q[j1, i1]
                v += Q[j1 + 1, i1]
                v += Q[j1, i1 + 1]
                v += Q[j1 + 1, i1 + 1]
                eU[k, j, i] = 0.25 * v
        for k in range(2):
            m = mB[k * N2 + c]
            if m < 0:
                continue
            s, i1, j1 = m & 0xFF, (m >> 8) & 0xFF, (m >> 16) & 0xFF
            Q = B5[s + 1, k]
            v = 0.0
            v += Q[j1, i1]
            v += Q[j1 + 1, i1] if k == 0 else Q[j1, i1 + 1]
            eB[k, j, i] = 0.5 * v
    assert np.array_equal(eU, rU) and np.array_equal(eB, rB)
