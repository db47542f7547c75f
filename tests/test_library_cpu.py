"""CPU-side checks of the product library: it loads, exports every entry point
declared in include/*.h, and its host-side fill-program compiler and coarsen
tables reproduce the reference on random trees (no GPU needed)."""
import ctypes as C
import glob
import os
import re
import subprocess

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


def _prototypes():
    """name -> (return kind, [argument kinds]) of every entry point declared in
    include/*.h; kinds: i32, i64, f64, str (char pointer), ptr, void."""
    def kind(decl):
        d = " ".join(decl.replace("const ", "").split())
        if "*" in d:
            return "str" if re.match(r"char\s*\*\s*\w*$", d) else "ptr"
        t = d.split()[0]
        return {"int": "i32", "int32_t": "i32", "int64_t": "i64", "double": "f64", "void": "void"}[t]

    out = {}
    for h in glob.glob(os.path.join(REPO, "include", "*.h")):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        for ret, name, args in re.findall(r"([\w\s\*]+?)\b(amr_[a-z_0-9]+)\s*\(([^)]*)\)\s*;", text):
            a = [x.strip() for x in args.split(",") if x.strip() and x.strip() != "void"]
            out[name] = (kind(ret.strip() + (" *" if ret.strip().endswith("*") else "")), [kind(x) for x in a])
    return out


def _ctypes_kind(t):
    if t is None:
        return "void"
    if t in (C.c_int, C.c_int32):
        return "i32"
    if t is C.c_int64:
        return "i64"
    if t is C.c_double:
        return "f64"
    if t is C.c_char_p:
        return "str"
    if t is C.c_void_p or (isinstance(t, type) and issubclass(t, C._Pointer)):
        return "ptr"
    raise AssertionError(f"unexpected ctypes type {t}")


def test_ctypes_prototypes_match_header():
    """The Python binding's prototypes agree with the header's declarations
    (return and argument kinds, argument count) for every declared entry point."""
    from amr2d_mhd_amd import _lib

    protos = _prototypes()
    assert len(protos) >= 40
    for name, (ret, args) in protos.items():
        cret, cargs = _lib._SIGS[name]
        assert _ctypes_kind(cret) == ret, name
        assert [_ctypes_kind(a) for a in cargs] == args, name


def _build(tmp_path, cmd):
    p = subprocess.run(cmd, capture_output=True, text=True, cwd=REPO)
    assert p.returncode == 0, p.stderr


def test_header_is_c11_and_every_entry_point_links(tmp_path):
    """tests/abi/abi_check.c: the header compiles as strict C11, every declared
    entry point resolves against the built library, the struct layouts equal
    the Python binding's, and the host-only tree validation accepts the base
    tree and rejects a bad leaf flag."""
    from amr2d_mhd_amd import _lib

    src = os.path.join(REPO, "tests", "abi", "abi_check.c")
    listed = set(re.findall(r"ENTRY\((amr_[a-z_0-9]+)\)", open(src).read()))
    assert listed == set(_prototypes()), "abi_check.c must list every declared entry point"
    exe = str(tmp_path / "abi_check")
    libdir = os.path.join(REPO, "amr2d_mhd_amd")
    _build(tmp_path, ["gcc", "-std=c11", "-Wall", "-Wextra", "-Werror", "-pedantic", src, "-I" + os.path.join(REPO, "include"),
                      "-L" + libdir, "-l:libamr2d_mhd.so", "-Wl,-rpath," + libdir, "-o", exe])
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = {ln.split()[0]: ln.split()[1:] for ln in out.stdout.splitlines()}
    assert int(lines["entries"][0]) == len(listed)
    for cname, S in (("amr_sim_config", _lib.SimConfig), ("amr_inst_info", _lib.InstInfo), ("amr_diag", _lib.Diag)):
        v = lines[cname]
        assert int(v[0]) == C.sizeof(S), cname
        fields = {v[k]: (int(v[k + 1]), int(v[k + 2])) for k in range(1, len(v), 3)}
        assert set(fields) == {f[0] for f in S._fields_}, cname
        for f, _t in S._fields_:
            d = getattr(S, f)
            assert fields[f] == (d.offset, d.size), (cname, f)
    assert lines["validate_base"] == ["0"] and lines["validate_bad"] == ["10"]


def test_reference_loop_example_builds(tmp_path):
    """The reference's time loop bound to the C ABI
    (amr2d_mhd_amd/examples/reference_loop.cpp) compiles as C++17 without
    warnings and links against the library."""
    exe = str(tmp_path / "reference_loop")
    libdir = os.path.join(REPO, "amr2d_mhd_amd")
    _build(tmp_path, ["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror",
                      os.path.join(REPO, "amr2d_mhd_amd", "examples", "reference_loop.cpp"),
                      "-I" + os.path.join(REPO, "include"), "-L" + libdir, "-l:libamr2d_mhd.so",
                      "-Wl,-rpath," + libdir, "-o", exe])
    assert os.access(exe, os.X_OK)


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


@need_ref
@pytest.mark.parametrize("ng,lv,seed,pr,pc,rounds", [(2, 4, 1, 0.5, 0.2, 5), (2, 6, 7, 0.3, 0.3, 8),
                                                     (4, 5, 3, 0.4, 0.25, 6), (2, 5, 99, 0.25, 0.5, 10)])
def test_repeated_fill_needs_only_the_linked_tail(tmp_path, ng, lv, seed, pr, pc, rounds):
    """A step whose tree did not change starts with a fill of U and Bi that
    repeats the previous step's last fills of both; the engine runs only the
    chunks from the program's linked (stale-read) tail on.  On the host
    emulation of the device program, over the reference's random trees and
    random values: whole fill, then the tail alone, equals whole fill twice
    bit for bit; the tail is a small part of the program, and the second
    whole fill does change values (the stale reads), so skipping it
    altogether would not do."""
    from amr2d_mhd_amd._lib import lib

    rc, err = refrun.run(str(tmp_path), ng, lv, ic="BWx", lim="MC", flux="RUSA", randtree=(seed, pr, pc, rounds))
    assert rc == 0, err
    pre = refrun.read_state(str(tmp_path / "state_rt_pre.bin"))
    n, NT = pre["lastActive"], pre["NT"]
    lp, ic, lf = (np.ascontiguousarray(pre[k], dtype=np.int32) for k in ("lp", "icoord", "leafs"))
    rng = np.random.default_rng(seed)
    changed = False
    for nv, kind in [(8, 0), (2, 1), (1, 2)]:
        once = rng.standard_normal((n, nv, NT, NT))
        nops = lib().amr_debug_host_fill(NT, ng, n, lp.ctypes.data, ic.ctypes.data, lf.ctypes.data,
                                         once.ctypes.data, nv, kind, C.byref(C.c_int(0)), 0)
        assert nops > 0
        twice, tail = once.copy(), once.copy()
        assert lib().amr_debug_host_fill(NT, ng, n, lp.ctypes.data, ic.ctypes.data, lf.ctypes.data,
                                         twice.ctypes.data, nv, kind, C.byref(C.c_int(0)), 0) == nops
        t0 = C.c_int(-1)
        assert lib().amr_debug_host_fill_tail(NT, ng, n, lp.ctypes.data, ic.ctypes.data, lf.ctypes.data,
                                              tail.ctypes.data, nv, kind, C.byref(t0)) == nops
        assert np.array_equal(tail, twice)
        nchunks = lib().amr_debug_compile(NT, ng, n, lp.ctypes.data, ic.ctypes.data, lf.ctypes.data, 0, None,
                                          C.byref(C.c_int(0))) // 256
        assert 0 < t0.value and nchunks - t0.value <= nchunks // 4  # 6-13% on these trees
        changed |= not np.array_equal(twice, once)
    assert changed


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


@pytest.mark.parametrize("gamma", [1.4, 5.0 / 3.0, 2.0])
def test_hlld_host_build_equals_oracle(gamma):
    """The product's HLLD flux (physics.hpp, host build of the same source the
    kernels use) equals the oracle's bit for bit over states reaching every
    wave region: subsonic, supersonic both ways, zero and vanishing normal
    field, identical sides."""
    from amr2d_mhd_amd._lib import lib

    L = lib()
    rng = np.random.default_rng(int(gamma * 1000))
    n = 1500

    def states():
        w = np.zeros((n, 8))
        w[:, 0] = rng.uniform(0.1, 10, n)
        w[:, 1:4] = rng.normal(0, 2, (n, 3))
        w[:, 4] = rng.uniform(0.01, 10, n)
        w[:, 5:8] = rng.normal(0, 2, (n, 3))
        return w

    for case in range(6):
        wL, wR = states(), states()
        wR[:, 5] = wL[:, 5]
        if case == 1:
            wL[:, 1] += 6
            wR[:, 1] += 6
        elif case == 2:
            wL[:, 1] -= 6
            wR[:, 1] -= 6
        elif case == 3:
            wL[:, 5] = wR[:, 5] = 0.0
        elif case == 4:
            wL[:, 5] = wR[:, 5] = 1e-9 * wL[:, 5]
        elif case == 5:
            wR[:] = wL
        ref, _ = O.riemann("HLLD", "x", gamma, wL, wR)
        out = np.zeros((n, 8))
        for q in range(n):
            L.amr_debug_hlld(wL[q].ctypes.data, wR[q].ctypes.data, gamma, out[q].ctypes.data)
        same = (out.view(np.uint64) == ref.view(np.uint64)) | (np.isnan(out) & np.isnan(ref))
        assert same.all(), f"case {case}: {(~same.all(1)).sum()} states differ"


def test_operation_counting_build_is_the_real_arithmetic():
    """The op-counting build of the pointwise physics (flop_count.cpp: the
    same physics.hpp with a counting scalar type) computes every solver's flux
    bit for bit like the oracle, so the operations it counts are the ones the
    kernels perform; the branch-free solvers (Rusanov, HLLE) take the same
    number of operations on every state, HLLD the fewest in its supersonic
    regions (one side's flux) and more in the subsonic ones (star states)."""
    from amr2d_mhd_amd._lib import lib

    L = lib()
    rng = np.random.default_rng(11)
    totals = {}
    for solver, name in ((0, "RUSA"), (1, "HLLE"), (2, "HLLC"), (3, "HLLD")):
        seen = []
        for q in range(300):
            wL = np.concatenate([[rng.uniform(0.2, 5)], rng.normal(0, 1, 3), [rng.uniform(0.1, 5)], rng.normal(0, 1, 3)])
            wR = np.concatenate([[rng.uniform(0.2, 5)], rng.normal(0, 1, 3), [rng.uniform(0.1, 5)], rng.normal(0, 1, 3)])
            wR[5] = wL[5]
            if name == "HLLD" and q % 3 == 0:
                wL[1] += 8.0  # supersonic to the right: region 1
                wR[1] += 8.0
            f, c = np.zeros(8), np.zeros(4)
            L.amr_debug_count_flux(solver, wL.ctypes.data, wR.ctypes.data, 1.4, f.ctypes.data, c.ctypes.data)
            ref, _ = O.riemann(name, "x", 1.4, wL[None], wR[None])
            assert np.array_equal(f.view(np.uint64), ref[0].view(np.uint64)), (name, q)
            seen.append((q % 3 == 0, c.sum()))
        totals[name] = seen
    for name in ("RUSA", "HLLE"):
        assert len({t for _, t in totals[name]}) == 1, name
    sup = {t for s, t in totals["HLLD"] if s}
    sub = [t for s, t in totals["HLLD"] if not s]
    # (some unshifted random states are supersonic too)
    assert len(sup) == 1 and min(sub) >= max(sup) and max(sub) > max(sup)


@pytest.mark.parametrize("maxlevs,seed,p_ref", [(4, 1, 0.35), (5, 2, 0.3), (6, 3, 0.25), (5, 4, 0.6)])
def test_regrid_replay_equals_host_decision(maxlevs, seed, p_ref):
    """The host half of the device regrid: an ordered op list (refinement
    parents, then coarsening first sons, as the device records them) replayed
    on the host tree gives exactly the tree, slots and data-transfer ops of the
    host's own admesh -- over 40 successive regrids with random indicators, so
    that son groups are reused and stale blocks keep pointers."""
    from amr2d_mhd_amd._lib import lib

    d = O.Domain(ic="BWx", limiter="MC", flux="RUSA", cfl=0.4, nx=16, nghosts=2, maxlevs=maxlevs)
    assert d.init() == 0
    lp, ic, lf = [np.ascontiguousarray(a, np.int32) for a in d.get_tree()]
    rng = np.random.default_rng(seed)
    nref = ncoars = 0
    for r in range(40):
        n = len(lf)
        grad = np.where(rng.random(n) < p_ref, 1.0, 0.0)
        cap = 2 * 4 ** maxlevs
        lp2 = np.zeros((cap, 8), np.int32)
        ic2 = np.zeros((cap, 2), np.int32)
        lf2 = np.zeros(cap, np.int32)
        counts = np.zeros(3, np.int32)
        rc = lib().amr_debug_regrid_replay(20, 2, maxlevs, n, lp.ctypes.data, ic.ctypes.data, lf.ctypes.data,
                                           grad.ctypes.data, cap, lp2.ctypes.data, ic2.ctypes.data, lf2.ctypes.data,
                                           counts.ctypes.data)
        assert rc == 0, (r, rc)
        nref, ncoars = nref + counts[0], ncoars + counts[1]
        m = counts[2]
        lp, ic, lf = lp2[:m].copy(), ic2[:m].copy(), lf2[:m].copy()
    assert nref > 20 and ncoars > 5, (nref, ncoars)
