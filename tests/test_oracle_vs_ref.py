"""Pin the CPU oracle to the compiled reference (oracle/_ref, built from the
reference's own sources with -ffp-contract=off).

Every comparison is bitwise.  Tests that need the compiled reference are
skipped when it is not built; tests/golden holds fixtures produced by it
(tests/golden/make_golden.py) so the long runs stay pinned without it.
"""
import os
import struct

import numpy as np
import pytest

import oracle as O
import refrun

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")

need_ref = pytest.mark.skipif(not refrun.have_ref(2, 5), reason="compiled reference (oracle/_ref) not built")
need_ref4 = pytest.mark.skipif(not refrun.have_ref(4, 5), reason="compiled reference (oracle/_ref) not built")

LIMITERS = ["1st", "noSL", "WENO2D", "minmod", "VanLeer", "Superbee", "MC", "Koren", "Cada3rd"]


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


def physical_states(rng, n):
    w = np.zeros((n, 8))
    w[:, 0] = rng.uniform(0.1, 3.0, n)
    w[:, 1:4] = rng.normal(0, 1.0, (n, 3))
    w[:, 4] = rng.uniform(0.05, 3.0, n)
    w[:, 5:8] = rng.normal(0, 1.0, (n, 3))
    return w


def riemann_inputs(seed=11, n=512):
    rng = np.random.default_rng(seed)
    wL, wR = physical_states(rng, n), physical_states(rng, n)
    wR[:32] = wL[:32]                 # identical states
    wL[32:64, 5] = wR[32:64, 5] = 0   # zero normal field
    wL[64:96, 5:8] = wR[64:96, 5:8] = 0
    wR[96:128, 0] *= 1e3              # strong density jump
    wL[128:160, 1] = wR[128:160, 1] = 6.0   # supersonic to the right
    wL[160:192, 1] = wR[160:192, 1] = -6.0  # supersonic to the left
    wL[192:224, 4] = wR[192:224, 4] = 1e-6  # near vacuum pressure
    return wL, wR


# ----------------------------------------------------------- pointwise
@need_ref
def test_limiters_match_reference():
    r = np.concatenate([np.linspace(-5, 5, 401), [0.0, -0.0, 1.0, 2.0, 0.5, 1e-300, -1e-300, 1e300, 3.0]])
    eta = np.concatenate([np.linspace(0.0, 2.0, 401), [1.0, 1 - 1e-15, 1 + 1e-15, 0.999999, 1.000001, 0, 0, 0, 0]])
    for name in LIMITERS:
        payload = struct.pack("i", len(r)) + refrun.fixed_str(name, 32) + r.tobytes() + eta.tobytes()
        ref = np.frombuffer(refrun.vec("limiter", payload), np.float64)
        assert same(O.limiter(name, r, eta), ref), name


def test_limiter_closed_forms():
    """Textbook values (independent of the reference code)."""
    r = np.array([-1.0, 0.0, 0.25, 0.5, 1.0, 2.0, 4.0])
    eta = np.zeros_like(r)
    assert same(O.limiter("minmod", r, eta), np.clip(r, 0, 1))
    assert same(O.limiter("MC", r, eta), np.maximum(0, np.minimum(np.minimum(2 * r, 0.5 * (1 + r)), 2)))
    assert same(O.limiter("Superbee", r, eta), np.maximum(0, np.maximum(np.minimum(2 * r, 1), np.minimum(r, 2))))
    assert same(O.limiter("VanLeer", r, eta), (np.abs(r) + r) / (1 + np.abs(r)))
    assert same(O.limiter("1st", r, eta), np.zeros_like(r))
    assert same(O.limiter("noSL", r, eta), r)


@need_ref
@pytest.mark.parametrize("method", ["RUSA", "HLLE", "HLLC", "HLLD"])
def test_riemann_matches_reference(method):
    wL, wR = riemann_inputs()
    for direction in "xy":
        for gamma in (2.0, 1.4, 5.0 / 3.0):
            payload = (struct.pack("i", len(wL)) + refrun.fixed_str(method, 8) + direction.encode() +
                       struct.pack("d", gamma) + wL.tobytes() + wR.tobytes())
            raw = refrun.vec("riemann", payload)
            ref = np.frombuffer(raw[:8 * 8 * len(wL)], np.float64).reshape(-1, 8)
            rst = np.frombuffer(raw[8 * 8 * len(wL):], np.int32)
            out, st = O.riemann(method, direction, gamma, wL, wR)
            assert np.array_equal(st != 0, rst != 0), (method, direction, gamma)
            ok = st == 0
            assert same(out[ok], ref[ok]), (method, direction, gamma)


def _physical_flux(w, g):
    """Ideal-MHD x-flux of a primitive state (closed form)."""
    rho, u, v, ww, p, bx, by, bz = w
    b2 = bx * bx + by * by + bz * bz
    e = p / (g - 1) + 0.5 * rho * (u * u + v * v + ww * ww) + 0.5 * b2
    pt = p + 0.5 * b2
    return np.array([rho * u, rho * u * u + pt - bx * bx, rho * u * v - bx * by, rho * u * ww - bx * bz,
                     (e + pt) * u - bx * (u * bx + v * by + ww * bz), 0.0, u * by - bx * v, u * bz - bx * ww])


@pytest.mark.parametrize("method", ["RUSA", "HLLE", "HLLC", "HLLD"])
def test_riemann_consistency_and_upwinding(method):
    """F(u,u) = F(u) and, for a supersonic left state, F = F(left): pins the
    solvers' flux assembly against the closed-form MHD flux."""
    rng = np.random.default_rng(5)
    for _ in range(20):
        w = physical_states(rng, 1)[0]
        f, st = O.riemann(method, "x", 1.4, w[None], w[None])
        if st[0]:
            continue
        np.testing.assert_allclose(f[0], _physical_flux(w, 1.4), rtol=1e-12, atol=1e-12)
    w = np.array([1.0, 20.0, 0.3, -0.2, 1.0, 0.5, 0.4, 0.1])
    w2 = w.copy()
    w2[0], w2[4], w2[6] = 0.5, 0.7, -0.3
    f, st = O.riemann(method, "x", 1.4, w[None], w2[None])
    if method != "RUSA" and not st[0]:
        np.testing.assert_allclose(f[0], _physical_flux(w, 1.4), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("method", ["RUSA", "HLLE", "HLLD"])
def test_riemann_rotation_symmetry(method):
    """The y flux equals the x flux of the state with (u,v) and (Bx,By) swapped."""
    wL, wR = riemann_inputs(seed=3, n=64)
    fy, _ = O.riemann(method, "y", 1.4, wL, wR)
    sw = [0, 2, 1, 3, 4, 6, 5, 7]
    fx, _ = O.riemann(method, "x", 1.4, wL[:, sw], wR[:, sw])
    assert same(fy, fx[:, sw])


@need_ref
def test_wavespeeds_match_reference():
    rng = np.random.default_rng(2)
    w = physical_states(rng, 300)
    for norm in (5, 6):
        payload = struct.pack("<iid", len(w), norm, 1.4) + w.tobytes()
        ref = np.frombuffer(refrun.vec("wavespeeds", payload), np.float64).reshape(-1, 4)
        assert same(O.wavespeeds(w, 1.4, norm), ref)


# ------------------------------------------------------------- tiles
def _tiles(rng, nt, NT):
    W = rng.uniform(0.2, 2.0, (nt, 8, NT, NT))
    W[1] = np.round(W[1], 1)
    W[2, :, :, : NT // 2] = 1.0
    return W


@need_ref
@pytest.mark.parametrize("lim", ["MC", "minmod", "VanLeer", "Superbee", "Koren", "Cada3rd", "1st", "noSL", "WENO3"])
def test_recon_matches_reference_ng2(lim):
    _recon_case(lim, 2)


@need_ref4
@pytest.mark.parametrize("lim", ["WENO5", "WENO3", "MC"])
def test_recon_matches_reference_ng4(lim):
    _recon_case(lim, 4)


def _recon_case(lim, ng):
    NT = 16 + 2 * ng
    rng = np.random.default_rng(40 + ng)
    W = _tiles(rng, 3, NT)
    dx = 1.0 / 30
    payload = struct.pack("iii", len(W), NT, 1) + refrun.fixed_str(lim, 16) + refrun.fixed_str("RUSA", 8) + \
        struct.pack("dd", 1.4, dx)
    for t in range(len(W)):
        payload += W[t].tobytes() + np.zeros((2, NT, NT)).tobytes()
    raw = refrun.vec("recon", payload, ng=ng, lv=5)
    per = 4 * 8 * NT * NT * 8 + 4
    for t in range(len(W)):
        ref = np.frombuffer(raw[t * per: t * per + per - 4], np.float64).reshape(4, 8, NT, NT)
        outs, st = O.recon_tile(W[t], lim, ng, dx)
        assert same(np.stack(outs), ref), (lim, t)


def _physical_tile(rng, NT, smooth=True):
    x = np.linspace(0, 1, NT)
    X, Y = np.meshgrid(x, x)
    W = np.zeros((8, NT, NT))
    W[0] = 1.0 + 0.3 * np.sin(6 * X + 2 * Y) + 0.05 * rng.random((NT, NT))
    W[1] = 0.4 * np.cos(5 * Y)
    W[2] = 0.3 * np.sin(4 * X)
    W[3] = 0.1 * rng.random((NT, NT))
    W[4] = 0.8 + 0.2 * np.cos(3 * X * Y) + 0.05 * rng.random((NT, NT))
    W[5] = 0.5 + 0.1 * Y
    W[6] = -0.3 + 0.2 * X
    W[7] = 0.05
    g = 1.4
    U = np.zeros_like(W)
    U[0] = W[0]
    U[1:4] = W[1:4] * W[0]
    U[4] = W[4] / (g - 1) + 0.5 * W[0] * (W[1] ** 2 + W[2] ** 2 + W[3] ** 2) + 0.5 * (W[5] ** 2 + W[6] ** 2 + W[7] ** 2)
    U[5:8] = W[5:8]
    B = np.zeros((2, NT, NT))
    B[0] = W[5] + 0.01 * rng.random((NT, NT))
    B[1] = W[6] + 0.01 * rng.random((NT, NT))
    return U, B


@need_ref
@pytest.mark.parametrize("flux", ["RUSA", "HLLE", "HLLC", "HLLD"])
@pytest.mark.parametrize("step", [1, 2])
def test_muscl_block_matches_reference(flux, step):
    NT, ng = 20, 2
    rng = np.random.default_rng(100 + step)
    tiles = [_physical_tile(rng, NT) for _ in range(2)]
    payload = struct.pack("iii", len(tiles), NT, step) + refrun.fixed_str("MC", 16) + refrun.fixed_str(flux, 8) + \
        struct.pack("dd", 1.4, 1.0 / 30)
    for U, B in tiles:
        payload += U.tobytes() + B.tobytes()
    raw = refrun.vec("muscl", payload)
    per = 2 * 8 * NT * NT * 8 + 4
    for t, (U, B) in enumerate(tiles):
        blk = raw[t * per:(t + 1) * per]
        rff = np.frombuffer(blk[:8 * NT * NT * 8], np.float64).reshape(8, NT, NT)
        rgg = np.frombuffer(blk[8 * NT * NT * 8:-4], np.float64).reshape(8, NT, NT)
        rst = struct.unpack("i", blk[-4:])[0]
        ff, gg, st = O.muscl_tile(U, B, "MC", flux, 1.4, step, ng, dx=1.0 / 30)
        assert (st != 0) == (rst != 0)
        if not st:
            assert same(ff, rff) and same(gg, rgg), (flux, step, t)


@need_ref
@pytest.mark.parametrize("op", [0, 1])
def test_regrid_tiles_match_reference(op):
    NT, ng = 20, 2
    rng = np.random.default_rng(7 + op)
    U5 = np.zeros((5, 8, NT, NT))
    B5 = np.zeros((5, 2, NT, NT))
    for s in range(5):
        U5[s], B5[s] = _physical_tile(rng, NT)
    payload = struct.pack("<iiid", 1, NT, op, 1.4)
    for s in range(5):
        payload += U5[s].tobytes() + B5[s].tobytes()
    raw = refrun.vec("regrid", payload)
    rU = np.zeros_like(U5)
    rB = np.zeros_like(B5)
    off = 0
    for s in range(5):
        rU[s] = np.frombuffer(raw[off:off + U5[s].nbytes], np.float64).reshape(8, NT, NT)
        off += U5[s].nbytes
        rB[s] = np.frombuffer(raw[off:off + B5[s].nbytes], np.float64).reshape(2, NT, NT)
        off += B5[s].nbytes
    oU, oB, st = O.regrid_tiles(U5, B5, op, 1.4, ng)
    assert st == 0
    assert same(oU, rU) and same(oB, rB)


@need_ref
def test_criteria_match_reference():
    NT = 20
    rng = np.random.default_rng(9)
    payload = struct.pack("ii", 24, NT)
    cases = []
    for q in range(24):
        W = 1.0 + rng.uniform(-1, 1) * 0.03 * rng.random((8, NT, NT))
        if q % 3 == 0:
            W[5] = 0.0  # B = 0: the eps floor makes round-off decide
        if q % 4 == 1:
            W[4, 10, 10] *= 1.2
        lev = q % 5
        cases.append((lev, W))
        payload += struct.pack("i", lev) + W.tobytes()
    ref = np.frombuffer(refrun.vec("criteria", payload), np.int32).reshape(-1, 2)
    for q, (lev, W) in enumerate(cases):
        assert O.criteria_tile(W, 2, 5, lev) == tuple(ref[q]), q


# ---------------------------------------------------- fills on random trees
@need_ref
@pytest.mark.parametrize("ng,lv,seed", [(2, 4, 1), (2, 6, 7), (4, 5, 3)])
def test_fills_and_ct_on_random_trees(tmp_path, ng, lv, seed):
    rc, err = refrun.run(str(tmp_path), ng, lv, ic="BWx", lim="MC", flux="RUSA", randtree=(seed, 0.4, 0.25, 6))
    assert rc == 0, err
    pre = refrun.read_state(str(tmp_path / "state_rt_pre.bin"))
    post = refrun.read_state(str(tmp_path / "state_rt_fill.bin"))
    ct = refrun.read_state(str(tmp_path / "state_rt_ct.bin"))
    d = O.Domain(ic="BWx", limiter="MC", flux="RUSA", cfl=0.4, nx=16, nghosts=ng, maxlevs=lv)
    d.load_state(pre)
    d.set_time(0.0, pre["dt"], 0.0, 0)
    ib = d.locate_bounds()
    d.fill(O.U, O.FILL_CC)
    d.fill(O.BI, O.FILL_FC)
    d.fill(O.FF, O.FILL_FC)
    d.fill(O.GG, O.FILL_FC)
    d.fill(O.EMF, O.FILL_EC)
    for name, f in [("U", O.U), ("Bi", O.BI), ("ff", O.FF), ("gg", O.GG), ("EMF", O.EMF)]:
        assert same(d.get_field(f), post[name]), name
    assert len(ib) > 0
    d.ct(1)
    for name, f in [("Us", O.US), ("Bi", O.BI), ("Bis", O.BIS), ("EMF", O.EMF)]:
        assert same(d.get_field(f), ct[name]), name


# ---------------------------------------------------------- full runs
def _hash_run(dom, log, nsteps):
    assert dom.init() == 0
    assert dom.hash_fields() == int(log[0]["hf_post"]) and dom.hash_tree() == int(log[0]["ht_post"])
    for s in range(nsteps):
        applied = dom.info()["dt"]  # meshblock::dt used by this step (lagged, main.cpp:231)
        assert dom.step() == 0
        r = log[s + 1]
        assert applied == r["dt"], s
        assert dom.info()["t"] == r["t"], s
        assert dom.hash_fields() == int(r["hf_post"]) and dom.hash_tree() == int(r["ht_post"]), f"step {s}"


def test_brio_wu_lv4_full_run_golden():
    """Config 1 (453 steps) against the reference's per-step state hashes."""
    log = np.load(os.path.join(GOLDEN, "run_BWx_MC_RUSA_lv4.npy"))
    dom = O.Domain(ic="BWx", limiter="MC", flux="RUSA", cfl=0.4, nx=16, nghosts=2, maxlevs=4)
    _hash_run(dom, log, len(log) - 1)
    assert dom.done() and len(log) - 1 == 453


@pytest.mark.parametrize("name,kw,steps", [
    ("run_rotor_MC_HLLD_lv5", dict(ic="rotor", limiter="MC", flux="HLLD", cfl=0.5, maxlevs=5, nghosts=2), 120),
    ("run_rotor_MC_HLLC_lv4", dict(ic="rotor", limiter="MC", flux="HLLC", cfl=0.5, maxlevs=4, nghosts=2), 60),
    ("run_rotor_MC_HLLE_lv4", dict(ic="rotor", limiter="MC", flux="HLLE", cfl=0.5, maxlevs=4, nghosts=2), 60),
    ("run_BWy_minmod_RUSA_lv3", dict(ic="BWy", limiter="minmod", flux="RUSA", cfl=0.4, maxlevs=3, nghosts=2), 60),
    ("run_SST_Superbee_RUSA_lv3", dict(ic="SST", limiter="Superbee", flux="RUSA", cfl=0.4, maxlevs=3, nghosts=2), 60),
    ("run_rotor_WENO5_HLLD_lv3_ng4", dict(ic="rotor", limiter="WENO5", flux="HLLD", cfl=0.4, maxlevs=3, nghosts=4), 40),
    # CTtype 2 (CT2D.cpp:28-75), against the reference built with CTtype = 2; on the rotor the
    # upwind weight dPx / (dPx + dPy) is taken at ~0.4% of the corners, the rest use 0.5
    ("run_rotor_MC_HLLD_lv4_ct2", dict(ic="rotor", limiter="MC", flux="HLLD", cfl=0.5, maxlevs=4, nghosts=2,
                                       ct_type=2), 300),
    ("run_rotor_minmod_HLLE_lv4_ct2", dict(ic="rotor", limiter="minmod", flux="HLLE", cfl=0.5, maxlevs=4, nghosts=2,
                                           ct_type=2), 200),
    ("run_rotor_WENO5_HLLD_lv3_ng4_ct2", dict(ic="rotor", limiter="WENO5", flux="HLLD", cfl=0.4, maxlevs=3,
                                              nghosts=4, ct_type=2), 60),
])
def test_runs_golden(name, kw, steps):
    log = np.load(os.path.join(GOLDEN, name + ".npy"))
    dom = O.Domain(nx=16, **kw)
    _hash_run(dom, log, min(steps, len(log) - 1))


def test_cfl05_brio_wu_fails_like_reference():
    """The reference aborts config 1 at CFL 0.5 in step 2 (Us2W p<=0 at block 53, i=17, j=2)."""
    dom = O.Domain(ic="BWx", limiter="MC", flux="RUSA", cfl=0.5, nx=16, nghosts=2, maxlevs=4)
    dom.init()
    assert [dom.step() for _ in range(3)] == [0, 0, 1]
    assert dom.last_error() == [1, 53, 17, 2]


@pytest.mark.skipif(not os.access(os.path.join(refrun.REF_DIR, "ng2_lv5_bs", "refrun"), os.X_OK),
                    reason="patched-IC reference (oracle/_ref/ng2_lv5_bs) not built")
@pytest.mark.parametrize("cfl", [0.5, 0.3])
def test_rotor_bs_aborts_like_reference(tmp_path, cfl):
    """The rotor workload's stated initial condition (Balsara-Spicer: p = 1,
    v0 = 2, Bx = 5/sqrt(4 pi)) against the reference built with only those
    constants of IC2Dtype.cpp changed: identical state hashes through step 0,
    then the reference itself aborts in its second step -- U2W of a freshly
    refined son (refine.cpp:69-83) finds p <= 0 at block 39, i = j = 2 (the
    domain corner) -- and the oracle stops with the same code at the same
    block and cell, at any CFL."""
    import re
    import subprocess

    exe = os.path.join(refrun.REF_DIR, "ng2_lv5_bs", "refrun")
    p = subprocess.run([exe, "run", "--dir", str(tmp_path), "--verbose", "--log", str(tmp_path / "log.bin"),
                        "--input", refrun.input_line(16, 1.0, cfl, "rotor", "MC", "HLLD")],
                       capture_output=True, text=True, timeout=600)
    assert p.returncode == 10, p.stderr
    m = re.search(r"Error at U2W as zero/negative pressure! P=\S+, at i=(\d+), j=(\d+) & nb=(\d+)", p.stdout)
    assert m, p.stdout[-2000:]
    ref_i, ref_j, ref_nb = (int(g) for g in m.groups())
    log = refrun.read_log(str(tmp_path))
    assert len(log) == 2  # the initial state and step 0
    dom = O.Domain(ic="rotor_bs", limiter="MC", flux="HLLD", cfl=cfl, nx=16, nghosts=2, maxlevs=5)
    assert dom.init() == 0
    assert (dom.hash_fields(), dom.hash_tree()) == (int(log[0]["hf_post"]), int(log[0]["ht_post"]))
    assert dom.step() == 0
    assert (dom.hash_fields(), dom.hash_tree()) == (int(log[1]["hf_post"]), int(log[1]["ht_post"]))
    assert dom.step() == 1
    assert dom.last_error() == [1, ref_nb, ref_i, ref_j] == [1, 39, 2, 2]
