"""Regrid decided on the device (regrid.hip): markref, critneighup, the
refinement loop with updatelpup, critneighdown and the coarsening loop with
updatelpdown (admesh.cpp:17-102) run on a device copy of each instance's tree;
the host applies the device's ordered op list to its own copy.

Checked here: after every step the device copy equals the host tree bit for
bit -- every row up to lastActive, the neighbour pointers of stale blocks
included -- and runs with the regrid decided on the device (tree staged in
LDS, or worked on in place) and on the host (amr_set_regrid_device) are
identical in every slot field, for single
instances with mirror and periodic boundaries, RK2 and SSP-RK3, and for a
batch of instances of uneven depth.  Bitwise agreement with the reference
itself is what the lockstep and reference-log tests check, which run with
the device regrid (the default)."""
import numpy as np
import pytest

from gpu_util import compare_states, gpu_state

pytestmark = pytest.mark.gpu

ROTOR = dict(ic="rotor", limiter="MC", flux="HLLD", cfl=0.5, nx=16, nghosts=2, maxlevs=5)
OT6 = dict(ic="OT", limiter="WENO5", flux="HLLD", cfl=0.4, nx=16, nghosts=4, maxlevs=6, periodic=True, rk_stages=3)
BW5 = dict(ic="BWx", limiter="minmod", flux="HLLE", cfl=0.4, nx=16, nghosts=2, maxlevs=5)


def _trees_equal(sim, label, inst=0):
    lp, _, lf = sim.get_tree(inst)
    dlp, dlf, current = sim.device_tree(inst)
    assert current, f"{label}: the device copy of the tree is stale"
    if not np.array_equal(lp, dlp):
        rows = np.unique(np.argwhere(lp != dlp)[:, 0])
        raise AssertionError(f"{label}: lp differs in rows {rows[:8].tolist()}: host {lp[rows[0]].tolist()} "
                             f"device {dlp[rows[0]].tolist()}")
    assert np.array_equal(lf, dlf), f"{label}: leaf flags differ"


@pytest.mark.parametrize("lds", [True, False], ids=["tree in LDS", "tree in place"])
@pytest.mark.parametrize("name,cfg,steps", [("rotor lv5", ROTOR, 400), ("OT WENO5 RK3 lv6", OT6, 100),
                                            ("Brio-Wu lv5", BW5, 200)])
def test_device_tree_equals_host_tree_every_step(name, cfg, steps, lds):
    import amr2d_mhd_amd as A

    print("case:", name, "lds" if lds else "in place", flush=True)
    sim = A.Simulation(cfg)
    sim.set_regrid_lds(lds)
    sim.init()
    _trees_equal(sim, f"{name} after init")
    changes, la0 = 0, sim.info()["last_active"]
    prev = (la0, sim.info()["nleaf"])
    for s in range(steps):
        sim.step()
        inf = sim.info()
        cur = (inf["last_active"], inf["nleaf"])
        changes += cur != prev
        prev = cur
        _trees_equal(sim, f"{name} step {s}")
    assert changes >= 10 and prev[0] > la0, (changes, la0, prev)
    sim.close()


@pytest.mark.parametrize("lds", [True, False], ids=["tree in LDS", "tree in place"])
@pytest.mark.parametrize("name,cfg,steps", [("rotor lv5", ROTOR, 300), ("OT WENO5 RK3 lv6", OT6, 60)])
def test_device_and_host_regrid_give_identical_runs(name, cfg, steps, lds):
    import amr2d_mhd_amd as A

    print("case:", name, "lds" if lds else "in place", flush=True)
    dev = A.Simulation(cfg)
    dev.set_regrid_lds(lds)
    dev.init()
    host = A.Simulation(cfg)
    host.set_regrid_device(False)
    host.init()
    for s in range(steps):
        dev.step()
        host.step()
        a, b = dev.info(), host.info()
        for k in ("t", "dt", "dt_next", "last_active", "nleaf", "status"):
            assert a[k] == b[k], f"{name} step {s}: {k} {a[k]!r} (device) vs {b[k]!r} (host)"
        if (s + 1) % 50 == 0 or s == steps - 1:
            d = compare_states(gpu_state(dev), gpu_state(host))
            assert not d, f"{name} step {s}: " + "; ".join(d[:4])
    _, _, current = host.device_tree()
    assert not current  # the host-regrid context never refreshed a device copy
    dev.close()
    host.close()


def test_switching_regrid_side_mid_run_is_seamless():
    """Device -> host -> device in the middle of a rotor run: the same run as
    one that stays on the device (the device copy is refreshed from the host
    tree when the device takes over again)."""
    import amr2d_mhd_amd as A

    ref = A.Simulation(ROTOR)
    ref.init()
    sim = A.Simulation(ROTOR)
    sim.init()
    for phase, on, steps in (("device", True, 60), ("host", False, 60), ("device again", True, 80)):
        print("case: regrid on", phase, flush=True)
        sim.set_regrid_device(on)
        for _ in range(steps):
            ref.step()
            sim.step()
        d = compare_states(gpu_state(ref), gpu_state(sim))
        assert not d, f"after the {phase} phase: " + "; ".join(d[:4])
    _trees_equal(sim, "after switching back")
    sim.close()
    ref.close()


def test_batch_of_uneven_depths_device_vs_host_regrid():
    """Eight seeded rotors of depths 3 to 6 in one context: every instance's
    device tree equals its host tree every step, and the batch equals the same
    batch with the regrid decided on the host."""
    import amr2d_mhd_amd as A
    from amr2d_mhd_amd.ensemble import instance_config

    cfgs = [instance_config(i) for i in range(8)]
    assert len({c["maxlevs"] for c in cfgs}) >= 3
    dev = A.Simulation(cfgs)
    dev.init()
    host = A.Simulation(cfgs)
    host.set_regrid_device(False)
    host.init()
    for s in range(80):
        dev.step()
        host.step()
        for q in range(len(cfgs)):
            a, b = dev.info(q), host.info(q)
            assert a == b, f"instance {q} step {s}: {a} vs {b}"
            if not a["done"]:
                _trees_equal(dev, f"instance {q} step {s}", inst=q)
    for q in range(len(cfgs)):
        d = compare_states(gpu_state(dev, q), gpu_state(host, q))
        assert not d, f"instance {q}: " + "; ".join(d[:4])
    dev.close()
    host.close()


def test_ensemble_of_128_rotors_device_vs_host_regrid():
    """128 seeded rotors (depths 3 to 6, HLLC and HLLD) in one context for 400
    steps, regrid decided on the device and on the host: every instance's
    info, tree and slot fields identical."""
    import amr2d_mhd_amd as A
    from amr2d_mhd_amd.ensemble import ensemble_configs

    cfgs = ensemble_configs(128)
    runs = []
    for dev in (True, False):
        print("case: 128 rotors, regrid on the", "device" if dev else "host", flush=True)
        sim = A.Simulation(cfgs)
        sim.set_regrid_device(dev)
        sim.init()
        sim.run(400)
        runs.append(sim)
    dev, host = runs
    changed = 0
    for q in range(len(cfgs)):
        a, b = dev.info(q), host.info(q)
        assert a == b, f"instance {q}: {a} vs {b}"
        changed += a["last_active"] > 4
        d = compare_states(gpu_state(dev, q), gpu_state(host, q))
        assert not d, f"instance {q}: " + "; ".join(d[:4])
    assert changed > 100
    dev.close()
    host.close()
