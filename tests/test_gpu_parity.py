"""Parity of the HIP path against the CPU oracle, through the C ABI.

Bar: bitwise equality (FP64, same evaluation order, no FMA contraction) of
every slot field of every slot, the quadtree and the time/dt sequence."""
import numpy as np
import pytest

import oracle as O
from gpu_util import FIELD_MAP, compare_states, gpu_state, lockstep, oracle_state

pytestmark = pytest.mark.gpu


def _physical_states(rng, n, bscale=1.0):
    w = np.zeros((n, 8))
    w[:, 0] = rng.uniform(0.1, 3.0, n)
    w[:, 1:4] = rng.normal(0, 1.0, (n, 3))
    w[:, 4] = rng.uniform(0.05, 3.0, n)
    w[:, 5:8] = rng.normal(0, bscale, (n, 3))
    return w


def _torch():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.mark.parametrize("method", ["RUSA", "HLLE", "HLLC", "HLLD"])
@pytest.mark.parametrize("direction", ["x", "y"])
def test_riemann_batch_bitwise(method, direction):
    torch = _torch()
    import amr2d_mhd_amd as A

    rng = np.random.default_rng(1000 + len(method) + (direction == "y"))
    n = 4096
    wL, wR = _physical_states(rng, n), _physical_states(rng, n)
    # degenerate edges: identical states, zero normal field, zero field, strong jumps
    wR[:64] = wL[:64]
    wL[64:128, 5] = 0.0
    wR[64:128, 5] = 0.0
    wL[128:192, 5:8] = 0.0
    wR[128:192, 5:8] = 0.0
    wR[192:256, 0] *= 100.0
    wL[256:320, 1] = 5.0
    wR[256:320, 1] = 5.0
    wL[320:384, 1] = -5.0
    wR[320:384, 1] = -5.0
    for gamma in (2.0, 1.4, 5.0 / 3.0):
        fo, so = O.riemann(method, direction, gamma, wL, wR)
        fg, sg = A.riemann_batch(method, direction, gamma, torch.from_numpy(wL).cuda(),
                                 torch.from_numpy(wR).cuda())
        fg, sg = fg.cpu().numpy(), sg.cpu().numpy()
        assert np.array_equal(so != 0, sg != 0)
        ok = so == 0
        same = (fg[ok] == fo[ok]) | (np.isnan(fg[ok]) & np.isnan(fo[ok]))
        assert same.all(), f"{method} {direction} gamma={gamma}: {int((~same).sum())} flux components differ"


@pytest.mark.parametrize("lim,ng", [("MC", 2), ("minmod", 2), ("VanLeer", 2), ("Superbee", 2), ("Koren", 2),
                                    ("Cada3rd", 2), ("1st", 2), ("noSL", 2), ("WENO3", 2), ("WENO5", 4)])
def test_recon_batch_bitwise(lim, ng):
    torch = _torch()
    import amr2d_mhd_amd as A

    NT = 16 + 2 * ng
    rng = np.random.default_rng(7 + ng)
    nt = 6
    W = rng.uniform(0.2, 2.0, (nt, 8, NT, NT))
    W[1] = np.round(W[1], 1)  # ties and flat regions
    W[2, :, :, :NT // 2] = 1.0
    out = A.recon_batch(torch.from_numpy(W).cuda(), lim, ng, dx=1.0 / 30).cpu().numpy()
    weno = lim.startswith("WENO")
    lo, hi = (ng - 1, NT - ng) if weno else (0, NT - 2)
    for t in range(nt):
        (xl, xr, yl, yr), st = O.recon_tile(W[t], lim, ng, dx=1.0 / 30)
        ref = np.stack([xl, xr, yl, yr])
        sl = (slice(None), slice(None), slice(lo, hi + 1), slice(lo, hi + 1))
        a, b = out[t][sl], ref[sl]
        a = a.copy()
        b = b.copy()
        a[0:2, 5] = b[0:2, 5] = 0  # x edges skip Bx, y edges skip By in the reference
        a[2:4, 6] = b[2:4, 6] = 0
        assert np.array_equal(a, b), f"{lim}: tile {t} differs in {int((a != b).sum())} entries"


def _pair(ic, lim, flux, cfl, maxlevs, nghosts=2):
    import amr2d_mhd_amd as A

    sim = A.Simulation(dict(ic=ic, limiter=lim, flux=flux, cfl=cfl, nx=16, nghosts=nghosts, maxlevs=maxlevs))
    dom = O.Domain(ic=ic, limiter=lim, flux=flux, cfl=cfl, nx=16, nghosts=nghosts, maxlevs=maxlevs)
    assert dom.init() == 0
    sim.init()
    d = compare_states(gpu_state(sim), oracle_state(dom))
    assert not d, "initial state differs: " + "; ".join(d[:4])
    return sim, dom


def test_upload_between_steps_makes_the_next_first_fill_whole():
    """After a step that leaves the tree unchanged, the next step's first fill
    of U and Bi runs only the fill programs' linked tails -- valid because U
    and Bi still hold the last fills' output.  An amr_upload in between
    breaks that: here it raises the total energy of every interior cell, the
    same on the oracle, so the next step needs every ghost cell refilled; the
    run must stay bitwise on the oracle."""
    sim, dom = _pair("rotor", "MC", "HLLD", 0.5, 4)
    lockstep(sim, dom, 20, full_every=10, label="rotor lv4 before")
    for _ in range(200):  # reach a step after which the tree is unchanged
        before = [a.copy() for a in sim.get_tree()]
        lockstep(sim, dom, 1, full_every=1, label="rotor lv4 search")
        after = sim.get_tree()
        if all(x.shape == y.shape and np.array_equal(x, y) for x, y in zip(before, after)):
            break
    else:
        pytest.fail("no step without a regrid op in 200 steps")
    ng = 2
    U = dom.get_field(O.U)
    U[:, 4, ng:-ng, ng:-ng] *= 1.001
    assert dom.set_field(O.U, U) == 0
    sim.upload("U", U)
    d = compare_states(gpu_state(sim), oracle_state(dom))
    assert not d, "uploaded state differs: " + "; ".join(d[:4])
    lockstep(sim, dom, 10, full_every=1, label="rotor lv4 after upload")


def test_brio_wu_lv4_full_run_bitwise():
    """Config 1: BWx, 4 levels, MC + Rusanov, CT, CFL 0.4, to t=0.2 (453 steps)."""
    sim, dom = _pair("BWx", "MC", "RUSA", 0.4, 4)
    n = lockstep(sim, dom, 1000, full_every=25, label="BWx lv4")
    assert n == 453 and dom.done() and sim.info()["done"]
    assert sim.info()["t"] == dom.info()["t"]


def test_rotor_lv5_hlld_bitwise():
    """Rotor (reference IC), 5 levels, MC + HLLD, CFL 0.5: first 150 steps."""
    sim, dom = _pair("rotor", "MC", "HLLD", 0.5, 5)
    lockstep(sim, dom, 150, full_every=30, label="rotor lv5")


@pytest.mark.parametrize("flux", ["HLLE", "HLLC"])
def test_rotor_lv4_other_solvers_bitwise(flux):
    sim, dom = _pair("rotor", "MC", flux, 0.5, 4)
    lockstep(sim, dom, 60, full_every=20, label=f"rotor {flux}")


@pytest.mark.parametrize("ic,lim", [("BWy", "minmod"), ("BWx1", "VanLeer"), ("SST", "Superbee")])
def test_other_ics_limiters_bitwise(ic, lim):
    sim, dom = _pair(ic, lim, "RUSA", 0.4, 3)
    lockstep(sim, dom, 60, full_every=20, label=f"{ic} {lim}")


def test_weno5_ng4_bitwise():
    """WENO5 needs nghosts=4 (24x24 tiles)."""
    sim, dom = _pair("rotor", "WENO5", "HLLD", 0.4, 3, nghosts=4)
    lockstep(sim, dom, 40, full_every=10, label="WENO5 ng4")


@pytest.mark.parametrize("regrid_device", [True, False], ids=["regrid on the device", "regrid on the host"])
def test_phases_on_oracle_state(regrid_device):
    """Each phase applied to the same mid-run rotor state on both sides, the
    regrid (admesh, with its data transfer) last."""
    import amr2d_mhd_amd as A

    dom = O.Domain(ic="rotor", limiter="MC", flux="HLLD", cfl=0.5, nx=16, nghosts=2, maxlevs=5)
    dom.init()
    for _ in range(40):
        assert dom.step() == 0
    st = oracle_state(dom)
    sim = A.Simulation(dict(ic="rotor", limiter="MC", flux="HLLD", cfl=0.5, nx=16, nghosts=2, maxlevs=5))
    sim.set_regrid_device(regrid_device)
    sim.init()
    sim.load_state(st)
    dom.locate_bounds()
    assert np.array_equal(sim.schedule(), dom.locate_bounds())
    phases = [("fill U", lambda: sim.fill("U", "cell"), lambda: dom.fill(O.U, O.FILL_CC)),
              ("fill Bi", lambda: sim.fill("Bi", "face"), lambda: dom.fill(O.BI, O.FILL_FC)),
              ("stage 1", lambda: sim.stage(1), lambda: dom.stage(1)),
              ("fill ff", lambda: sim.fill("ff", "face"), lambda: dom.fill(O.FF, O.FILL_FC)),
              ("fill gg", lambda: sim.fill("gg", "face"), lambda: dom.fill(O.GG, O.FILL_FC)),
              ("ct 1", lambda: sim.ct(1), lambda: dom.ct(1)),
              ("fill Us", lambda: sim.fill("Us", "cell"), lambda: dom.fill(O.US, O.FILL_CC)),
              ("stage 2", lambda: sim.stage(2), lambda: dom.stage(2)),
              ("ct 2", lambda: sim.ct(2), lambda: dom.ct(2))]
    for name, g, o in phases:
        g()
        assert o() == 0
        d = compare_states(gpu_state(sim), oracle_state(dom))
        assert not d, f"after {name}: " + "; ".join(d[:4])
    lam_g = sim.cfl()
    st_o, lam_o = dom.cfl()
    assert st_o == 0 and lam_g == lam_o
    tree_before = dom.get_tree()[0]
    sim.regrid()
    assert dom.regrid() == 0
    assert not np.array_equal(dom.get_tree()[0][: len(tree_before)], tree_before) or \
        len(dom.get_tree()[0]) != len(tree_before)  # the regrid changed the tree
    d = compare_states(gpu_state(sim), oracle_state(dom))
    assert not d, "after regrid: " + "; ".join(d[:4])


def test_pressure_error_reported():
    """Config 1 at CFL 0.5 fails at step 2 with p <= 0 in the reference (Us2W)."""
    import amr2d_mhd_amd as A

    dom = O.Domain(ic="BWx", limiter="MC", flux="RUSA", cfl=0.5, nx=16, nghosts=2, maxlevs=4)
    dom.init()
    codes = [dom.step() for _ in range(3)]
    sim = A.Simulation(dict(ic="BWx", limiter="MC", flux="RUSA", cfl=0.5, nx=16, nghosts=2, maxlevs=4))
    sim.init()
    sim.step()
    sim.step()
    with pytest.raises(A.AmrError) as ei:
        sim.step()
    assert codes[:2] == [0, 0] and codes[2] == 1 and ei.value.code == 1


@pytest.mark.skip(reason="graph replay is disabled: replaying one captured step over changing trees "
                         "faulted the GPU twice in a long rotor run; its cause is not yet shown, so the case "
                         "stays off the GPU (amr_set_graph(ctx, 1) returns AMR_ERR_CONFIG)")
def test_graph_replay_equals_direct_launches():
    """The step replayed as a HIP graph equals direct launches bit for bit,
    replay after replay, over hundreds of tree changes: the captured kernels
    read the leaf and chunk counts from the device, so a graph is captured
    again only when the counts outgrow (or shrink far below) its grids or the
    stage-kernel variant changes -- not at every tree change."""
    import amr2d_mhd_amd as A

    cfg = dict(ic="rotor", limiter="MC", flux="HLLD", cfl=0.5, nx=16, nghosts=2, maxlevs=5)
    g, d = A.Simulation(cfg), A.Simulation(cfg)
    for x in (g, d):  # graph replay serves per-phase launches
        x.set_step_kernel(0)
    g.set_graph(True)
    d.set_graph(False)
    g.profile(("graph", "graph_capture", "host_topology"))
    g.init()
    d.init()
    for s in range(600):
        g.step()
        d.step()
        assert g.info() == d.info(), f"step {s}"
        if s % 25 == 24:
            diff = compare_states(gpu_state(g), gpu_state(d))
            assert not diff, f"after step {s}: " + "; ".join(diff[:4])
    pr = g.profile_read()
    replays, captures, rebuilds = pr["graph"][1], pr["graph_capture"][1], pr["host_topology"][1]
    assert replays > 500, replays
    assert rebuilds > 40 and captures * 4 < rebuilds, (rebuilds, captures)


def test_fill_program_cache_equals_recompiling():
    """Reusing cached fill programs for recurring trees equals compiling the
    program at every tree change bit for bit, and the cache is hit."""
    import amr2d_mhd_amd as A

    cfg = dict(ic="rotor", limiter="MC", flux="HLLD", cfl=0.5, nx=16, nghosts=2, maxlevs=5)
    a, b = A.Simulation(cfg), A.Simulation(cfg)
    b.set_fill_cache(0)
    for s in (a, b):
        s.profile(("host_topology", "host_compile"))
        s.init()
    for k in range(900):
        a.step()
        b.step()
        if k % 150 == 149:
            diff = compare_states(gpu_state(a), gpu_state(b))
            assert not diff, f"after step {k}: " + "; ".join(diff[:4])
    assert a.info() == b.info()
    pa, pb = a.profile_read(), b.profile_read()
    rebuilds, compiled, compiled_plain = pa["host_topology"][1], pa["host_compile"][1], pb["host_compile"][1]
    assert rebuilds == pb["host_topology"][1]
    assert compiled < compiled_plain <= rebuilds, (compiled, compiled_plain, rebuilds)


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4, 5, 6, 7])
def test_stage_variant_bitwise(variant):
    """Every stage-kernel variant reproduces the oracle bit for bit."""
    import amr2d_mhd_amd as A

    sim = A.Simulation(dict(ic="rotor", limiter="MC", flux="HLLD", cfl=0.5, nx=16, nghosts=2, maxlevs=4))
    sim.set_step_kernel(0)  # the variants are those of the per-phase stage launch
    sim.set_stage_variant(variant)
    dom = O.Domain(ic="rotor", limiter="MC", flux="HLLD", cfl=0.5, nx=16, nghosts=2, maxlevs=4)
    assert dom.init() == 0
    sim.init()
    lockstep(sim, dom, 40, full_every=10, label=f"stage variant {variant}")


@pytest.mark.parametrize("variant", [6, 7])
@pytest.mark.parametrize("cfg", [
    dict(ic="rotor", limiter="MC", flux="HLLC", cfl=0.5, nx=16, nghosts=2, maxlevs=4),  # generic kernel
    dict(ic="OT", limiter="WENO5", flux="HLLD", cfl=0.4, nx=16, nghosts=4, maxlevs=4, periodic=True, rk_stages=3),
    dict(ic="BWx", limiter="Koren", flux="HLLD", cfl=0.4, nx=8, nghosts=2, maxlevs=4),  # 12x12 tiles
    dict(ic="OT", limiter="minmod", flux="HLLE", cfl=0.4, nx=16, nghosts=2, maxlevs=4, periodic=True, rk_stages=3),
], ids=["rotor-HLLC", "OT-WENO5-rk3", "BWx-nx8", "OT-minmod-HLLE-rk3"])
def test_stage_strips_bitwise(cfg, variant):
    """The leaf split into 2 or 4 row strips (one workgroup each) reproduces
    the oracle bit for bit: generic kernel, WENO5 on 24x24 tiles with RK3,
    and 12x12 tiles.  SSP-RK3's second stage reads and writes Us, which strips
    of one leaf cannot do without racing on their halo rows; it runs as the
    whole-leaf variant (the last case failed when strips ran it)."""
    import amr2d_mhd_amd as A

    sim = A.Simulation(cfg)
    sim.set_step_kernel(0)
    sim.set_stage_variant(variant)
    dom = O.Domain(**cfg)
    assert dom.init() == 0
    sim.init()
    lockstep(sim, dom, 30, full_every=10, label=f"strips {variant}")


@pytest.mark.parametrize("lim,flux,rk,maxlevs,ng", [("WENO5", "HLLD", 3, 4, 4), ("MC", "HLLC", 2, 3, 2),
                                                    ("minmod", "HLLE", 3, 4, 2)])
def test_orszag_tang_periodic_bitwise(lim, flux, rk, maxlevs, ng):
    """Orszag-Tang on the periodic domain (build-defined), RK2 and SSP-RK3."""
    import amr2d_mhd_amd as A

    cfg = dict(ic="OT", limiter=lim, flux=flux, cfl=0.4, nx=16, nghosts=ng, maxlevs=maxlevs, periodic=True,
               rk_stages=rk)
    sim = A.Simulation(cfg)
    dom = O.Domain(**cfg)
    assert dom.init() == 0
    sim.init()
    d = compare_states(gpu_state(sim), oracle_state(dom))
    assert not d, "initial state differs: " + "; ".join(d[:4])
    lockstep(sim, dom, 40, full_every=10, label=f"OT {lim} {flux} rk{rk}")


def test_rk3_mirror_boundaries_bitwise():
    """SSP-RK3 with the reference's mirror boundaries (Brio-Wu, 3 levels)."""
    import amr2d_mhd_amd as A

    cfg = dict(ic="BWx", limiter="MC", flux="RUSA", cfl=0.4, nx=16, nghosts=2, maxlevs=3, rk_stages=3)
    sim = A.Simulation(cfg)
    dom = O.Domain(**cfg)
    assert dom.init() == 0
    sim.init()
    lockstep(sim, dom, 40, full_every=10, label="BWx rk3")


def test_rotor_lv5_whole_run_equals_reference_log():
    """The bench workload against the compiled reference's own log
    (tests/golden/run_rotor_MC_HLLD_lv5_full.npy, serial build): the rotor, 5
    levels, MC + HLLD, CFL 0.5, the whole run, 4763 steps to t = 0.295.  Every step's
    applied dt, time, slot count and leaf count equal the reference's; the
    FNV-1a hashes of every slot field (U Us Bi Bis ff gg EMF) and of the
    quadtree -- the oracle's hash over the state downloaded from the GPU --
    equal the reference's at the start and every 200 steps."""
    import os

    import amr2d_mhd_amd as A

    log = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "run_rotor_MC_HLLD_lv5_full.npy"))
    cfg = dict(ic="rotor", limiter="MC", flux="HLLD", cfl=0.5, nx=16, nghosts=2, maxlevs=5)
    sim = A.Simulation(cfg)
    sim.init()
    dom = O.Domain(**cfg)  # used for its hash only
    assert dom.init() == 0

    def hashes():
        st = sim.state()
        dom.set_tree(st["lp"], st["icoord"], st["leafs"])
        for name, f in FIELD_MAP.items():
            dom.set_field(f, st[name])
        return dom.hash_fields(), dom.hash_tree()

    assert hashes() == (int(log[0]["hf_post"]), int(log[0]["ht_post"]))
    n = len(log) - 1
    assert n == 4763
    for s in range(n):
        applied = sim.info()["dt"]
        sim.step()
        r, inf = log[s + 1], sim.info()
        assert (applied, inf["t"], inf["last_active"], inf["nleaf"]) == (r["dt"], r["t"], r["la_post"], r["nleaf_post"]), s
        if (s + 1) % 200 == 0:
            assert hashes() == (int(r["hf_post"]), int(r["ht_post"])), f"state hash differs after step {s}"
    assert hashes() == (int(log[-1]["hf_post"]), int(log[-1]["ht_post"])), "final state hash differs"
    assert sim.info()["t"] >= 0.295 and log[-2]["t"] < 0.295  # the run ends at the same step
    sim.close()


def test_rotor_bs_aborts_like_reference():
    """The rotor workload's stated initial condition (rotor_bs) at 5 levels:
    the GPU state equals the oracle's at the start and after step 0, and the
    next step fails with the reference's own abort (p <= 0 in the refined son
    39 at cell (2, 2), the oracle being pinned to the patched-IC reference)."""
    import amr2d_mhd_amd as A
    from amr2d_mhd_amd import AmrError

    cfg = dict(ic="rotor_bs", limiter="MC", flux="HLLD", cfl=0.5, nx=16, nghosts=2, maxlevs=5)
    sim = A.Simulation(cfg)
    dom = O.Domain(**cfg)
    assert dom.init() == 0
    sim.init()
    d = compare_states(gpu_state(sim), oracle_state(dom))
    assert not d, "initial state differs: " + "; ".join(d[:4])
    assert dom.step() == 0
    sim.step()
    d = compare_states(gpu_state(sim), oracle_state(dom))
    assert not d, "state after step 0 differs: " + "; ".join(d[:4])
    assert dom.step() == 1 and dom.last_error() == [1, 39, 2, 2]
    with pytest.raises(AmrError) as e:
        sim.step()
    assert e.value.code == 1
    assert (e.value.diag.nb, e.value.diag.i, e.value.diag.j) == (39, 2, 2)
    assert sim.info()["status"] == 1


OT_WORKLOAD = dict(ic="OT", limiter="WENO5", flux="HLLD", cfl=0.4, nx=16, nghosts=4, periodic=True, rk_stages=3)


def test_orszag_tang_lv5_stops_where_the_oracle_does():
    """Orszag-Tang, 5 levels, WENO5 + HLLD, SSP-RK3 (build-defined), against the
    oracle's log (tests/golden/ot_WENO5_HLLD_rk3_lv5_abort.npz): every step's
    applied dt, t, slot and leaf counts and tree hash, the hash of every slot
    field every 100 steps, then the scheme's own failure -- p <= 0 in block
    567, cell (19, 17) at step 686 -- with the same code at the same place.
    (t = 0.5 is out of this scheme's reach at 5 and 6 levels.)"""
    import os

    import amr2d_mhd_amd as A
    from amr2d_mhd_amd import AmrError
    from gpu_util import follow_oracle_log

    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                             "ot_WENO5_HLLD_rk3_lv5_abort.npz"))
    rows, abort = z["steps"], [int(v) for v in z["abort"]]
    assert len(rows) == 686 and abort == [1, 567, 19, 17, 686]
    cfg = dict(OT_WORKLOAD, maxlevs=5)
    sim = A.Simulation(cfg)
    sim.init()
    dom = O.Domain(**cfg)  # hashes only
    assert dom.init() == 0
    follow_oracle_log(sim, dom, rows, "OT lv5")
    with pytest.raises(AmrError) as e:
        sim.step()
    assert [e.value.code, e.value.diag.nb, e.value.diag.i, e.value.diag.j] == abort[:4]
    sim.close()


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("case", ["BWx lv4 full run", "rotor lv5", "OT WENO5 rk3 lv4", "rotor HLLC lv4 generic",
                                  "rotor CTtype 2 lv4"])
def test_step_kernel_modes_bitwise(case, mode):
    """Per-phase launches (mode 0) and the persistent step kernel with 256 or
    512 lanes (modes 1, 2) all reproduce the oracle bit for bit, through the
    specialised (HLLD + MUSCL, HLLD + WENO5) and generic stage code; the step
    kernel really runs in modes 1 and 2."""
    import amr2d_mhd_amd as A

    cfgs = {"BWx lv4 full run": (dict(ic="BWx", limiter="MC", flux="RUSA", cfl=0.4, maxlevs=4), 1000),
            "rotor lv5": (dict(ic="rotor", limiter="MC", flux="HLLD", cfl=0.5, maxlevs=5), 150),
            "OT WENO5 rk3 lv4": (dict(ic="OT", limiter="WENO5", flux="HLLD", cfl=0.4, maxlevs=4, nghosts=4,
                                      periodic=True, rk_stages=3), 40),
            "rotor HLLC lv4 generic": (dict(ic="rotor", limiter="MC", flux="HLLC", cfl=0.5, maxlevs=4), 60),
            "rotor CTtype 2 lv4": (dict(ic="rotor", limiter="MC", flux="HLLD", cfl=0.5, maxlevs=4, ct_type=2), 60)}
    cfg, steps = cfgs[case]
    cfg = dict(dict(nx=16, nghosts=2), **cfg)
    sim = A.Simulation(cfg)
    sim.set_step_kernel(mode)
    sim.profile(("step",))
    dom = O.Domain(**cfg)
    assert dom.init() == 0
    sim.init()
    n = lockstep(sim, dom, steps, full_every=20, label=f"{case} step kernel mode {mode}")
    launched = sim.profile_read()["step"][1]
    assert launched == (0 if mode == 0 else n), (launched, n)


def test_orszag_tang_lv6_first_1000_steps_equal_oracle_log():
    """The Orszag-Tang workload at its 6 levels (WENO5 + HLLD, SSP-RK3,
    periodic, build-defined), first 1000 steps (to t = 0.158, ~4000 leaves),
    against the oracle's log (tests/golden/ot_WENO5_HLLD_rk3_lv6_1000.npz):
    every step's applied dt, t, slot and leaf counts and tree hash, and the
    hash of every slot field every 100 steps."""
    import os

    import amr2d_mhd_amd as A
    from gpu_util import follow_oracle_log

    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ot_WENO5_HLLD_rk3_lv6_1000.npz"))
    rows = z["steps"]
    assert len(rows) == 1001 and int(z["abort"][0]) == -1
    cfg = dict(OT_WORKLOAD, maxlevs=6)
    sim = A.Simulation(cfg)
    sim.init()
    dom = O.Domain(**cfg)  # hashes only
    assert dom.init() == 0
    follow_oracle_log(sim, dom, rows, "OT lv6")
    assert sim.info()["nleaf"] == int(rows[-1]["nleaf"]) > 3000
    sim.close()


def test_graph_replay_is_disabled():
    """amr_set_graph(ctx, 1) is refused while graph replay is disabled; direct
    launches (0) stay accepted."""
    import amr2d_mhd_amd as A
    from amr2d_mhd_amd import AmrError

    sim = A.Simulation(dict(ic="BWx", limiter="MC", flux="RUSA", cfl=0.4, nx=16, nghosts=2, maxlevs=3))
    with pytest.raises(AmrError) as e:
        sim.set_graph(True)
    assert e.value.code == 10
    sim.set_graph(False)


CT2_CASES = {
    "rotor MC HLLD lv4": (dict(ic="rotor", limiter="MC", flux="HLLD", cfl=0.5, nx=16, nghosts=2, maxlevs=4), 300,
                          "run_rotor_MC_HLLD_lv4_ct2"),
    "rotor minmod HLLE lv4": (dict(ic="rotor", limiter="minmod", flux="HLLE", cfl=0.5, nx=16, nghosts=2, maxlevs=4),
                              200, "run_rotor_minmod_HLLE_lv4_ct2"),
    "rotor WENO5 HLLD lv3": (dict(ic="rotor", limiter="WENO5", flux="HLLD", cfl=0.4, nx=16, nghosts=4, maxlevs=3),
                             60, "run_rotor_WENO5_HLLD_lv3_ng4_ct2"),
}


@pytest.mark.parametrize("case", list(CT2_CASES))
def test_ct_type2_equals_reference_log(case):
    """CTtype 2 (the upwind corner EMF of CT2D.cpp:28-75) against the log of
    the reference built with CTtype = 2: every step's dt, time and tree sizes,
    and the hashes of every slot field and of the tree every 20 steps and at
    the end.  On the rotor the upwind weight dPx / (dPx + dPy) is taken at a
    fraction of the corners, so both branches of the weight are exercised."""
    import os

    import amr2d_mhd_amd as A

    cfg, steps, name = CT2_CASES[case]
    log = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name + ".npy"))
    assert len(log) - 1 == steps
    sim = A.Simulation(dict(cfg, ct_type=2))
    sim.init()
    dom = O.Domain(**dict(cfg, ct_type=2))  # used for its hash only
    assert dom.init() == 0

    def hashes():
        st = sim.state()
        dom.set_tree(st["lp"], st["icoord"], st["leafs"])
        for fname, f in FIELD_MAP.items():
            dom.set_field(f, st[fname])
        return dom.hash_fields(), dom.hash_tree()

    assert hashes() == (int(log[0]["hf_post"]), int(log[0]["ht_post"]))
    for s in range(steps):
        applied = sim.info()["dt"]
        sim.step()
        r, inf = log[s + 1], sim.info()
        assert (applied, inf["t"], inf["last_active"], inf["nleaf"]) == (r["dt"], r["t"], r["la_post"], r["nleaf_post"]), s
        if (s + 1) % 20 == 0 or s + 1 == steps:
            assert hashes() == (int(r["hf_post"]), int(r["ht_post"])), f"state hash differs after step {s}"
    sim.close()


def test_ct_type2_lockstep_with_oracle_rk3_periodic():
    """CTtype 2 with the build-defined SSP-RK3 and periodic Orszag-Tang (no
    reference counterpart): lockstep with the oracle."""
    import amr2d_mhd_amd as A

    cfg = dict(ic="OT", limiter="MC", flux="HLLD", cfl=0.4, nx=16, nghosts=2, maxlevs=4, periodic=True,
               rk_stages=3, ct_type=2)
    sim = A.Simulation(cfg)
    dom = O.Domain(**cfg)
    assert dom.init() == 0
    sim.init()
    lockstep(sim, dom, 40, full_every=10, label="OT CTtype 2 rk3")


def test_ct_type_rejected_outside_range():
    import amr2d_mhd_amd as A
    from amr2d_mhd_amd import AmrError

    with pytest.raises(AmrError):
        A.Simulation(dict(ic="rotor", limiter="MC", flux="HLLD", cfl=0.5, nx=16, nghosts=2, maxlevs=4, ct_type=3))
