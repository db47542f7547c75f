"""Batched ensembles: many independent rotors of uneven AMR depth in one context.

Bar: every instance of a batch is bitwise identical to the same instance run
alone (tree, all slot fields, t and dt), the parameterised rotor is bitwise
identical to the oracle, and an instance that hits one of the reference's
throw sites stops alone with the oracle's error code while the rest go on."""
import numpy as np
import pytest

import oracle as O
from gpu_util import compare_states, gpu_state, lockstep, oracle_state

pytestmark = pytest.mark.gpu


def _alone(cfg, steps):
    import amr2d_mhd_amd as A

    sim = A.Simulation(cfg)
    sim.init()
    for _ in range(steps):
        if sim.step() == 0:
            break
    return sim


@pytest.mark.parametrize("inst", [0, 1])
def test_rotor_p_vs_oracle_bitwise(inst):
    """Seeded parameterised rotors (HLLC for even, HLLD for odd instances)."""
    import amr2d_mhd_amd as A
    from amr2d_mhd_amd.ensemble import instance_config

    cfg = instance_config(inst)
    cfg["maxlevs"] = 4
    sim = A.Simulation(cfg)
    dom = O.Domain(**cfg)
    assert dom.init() == 0
    sim.init()
    d = compare_states(gpu_state(sim), oracle_state(dom))
    assert not d, "initial state differs: " + "; ".join(d[:4])
    lockstep(sim, dom, 80, full_every=20, label=f"rotor_p {inst}")


def test_rotor_p_at_depths_5_and_6_batched_vs_oracle():
    """Seeded parameterised rotors at their drawn depths 5 and 6 (HLLC and
    HLLD), advanced together in one context: each instance equals its own
    oracle run -- dt and t every step, tree and every slot field every 20
    steps -- for 60 steps."""
    import amr2d_mhd_amd as A
    from amr2d_mhd_amd.ensemble import instance_config

    ids = [1, 6, 12, 13]
    cfgs = [instance_config(i) for i in ids]
    assert [(c["maxlevs"], c["flux"]) for c in cfgs] == [(5, "HLLD"), (5, "HLLC"), (6, "HLLC"), (6, "HLLD")]
    batch = A.Simulation(cfgs)
    batch.init()
    doms = [O.Domain(**c) for c in cfgs]
    for d in doms:
        assert d.init() == 0
    for s in range(60):
        batch.step()
        for q, d in enumerate(doms):
            assert d.step() == 0
            inf = batch.info(q)
            assert (inf["t"], inf["dt"]) == (d.info()["t"], d.info()["dt"]), (ids[q], s)
        if (s + 1) % 20 == 0:
            for q, d in enumerate(doms):
                diff = compare_states(gpu_state(batch, q), oracle_state(d))
                assert not diff, f"instance {ids[q]} after step {s}: " + "; ".join(diff[:4])
    batch.close()


def test_stage_launches_split_by_solver_class():
    """A batch mixing HLLD + MC, HLLC + MC and HLLE + MC instances runs each
    RK stage as one launch per class (HLLD and HLLC specialised, HLLE on the
    generic kernel), and every instance equals its alone run bit for bit (an
    instance alone runs one class, unsplit)."""
    import amr2d_mhd_amd as A
    from amr2d_mhd_amd.ensemble import instance_config

    cfgs = [instance_config(i) for i in (1, 2, 3, 4)]
    assert [c["flux"] for c in cfgs] == ["HLLD", "HLLC", "HLLD", "HLLC"]
    cfgs.append(dict(instance_config(5), flux="HLLE"))
    steps = 30
    batch = A.Simulation(cfgs)
    batch.init()
    batch.profile(("stage",))
    for _ in range(steps):
        batch.step()
    ms, launches = batch.profile_read()["stage"]
    assert launches == 3 * 2 * steps, launches
    for q, cfg in enumerate(cfgs):
        alone = _alone(cfg, steps)
        ib, ia = batch.info(q), alone.info()
        assert ib["status"] == ia["status"] == 0
        assert (ib["t"], ib["dt"], ib["nleaf"]) == (ia["t"], ia["dt"], ia["nleaf"]), q
        d = compare_states(gpu_state(batch, q), gpu_state(alone))
        assert not d, f"instance {q}: " + "; ".join(d[:4])
        alone.close()
    batch.close()


def test_batch_equals_alone_bitwise():
    """Fourteen seeded instances (levels 3-6, HLLC and HLLD) advanced together."""
    import amr2d_mhd_amd as A
    from amr2d_mhd_amd.ensemble import instance_config

    ids = list(range(14))
    cfgs = [instance_config(i) for i in ids]
    assert {c["maxlevs"] for c in cfgs} == {3, 4, 5, 6}
    steps = 40
    batch = A.Simulation(cfgs)
    batch.init()
    for _ in range(steps):
        batch.step()
    for q, cfg in enumerate(cfgs):
        alone = _alone(cfg, steps)
        ib, ia = batch.info(q), alone.info()
        assert ib["status"] == ia["status"] == 0
        assert (ib["t"], ib["dt"], ib["count"], ib["nleaf"]) == (ia["t"], ia["dt"], ia["count"], ia["nleaf"])
        d = compare_states(gpu_state(batch, q), gpu_state(alone))
        assert not d, f"instance {ids[q]}: " + "; ".join(d[:4])


def test_grouped_run_equals_single_context_bitwise():
    """The same instances split into round-robin groups (amr_run_group, one
    stream per context, the steps of the groups overlapping on the GPU), and
    stepped through amr_step_begin / amr_step_end, equal one context's run."""
    import ctypes as C

    import amr2d_mhd_amd as A
    from amr2d_mhd_amd._lib import lib
    from amr2d_mhd_amd.ensemble import instance_config

    cfgs = [instance_config(i) for i in range(12)]
    steps = 50
    one = A.Simulation(cfgs)
    one.init()
    n1, z1 = one.run(steps)
    groups = [A.Simulation(cfgs[0:5]), A.Simulation(cfgs[5:12])]
    for g in groups:
        g.init()
    ng, zg = A.run_group(groups, steps)
    assert (ng, zg) == (n1, z1)
    split = A.Simulation(cfgs[5:12])
    split.init()
    active = C.c_int(0)
    for _ in range(steps):
        assert lib().amr_step_begin(split.h) == 0
        assert lib().amr_step_begin(split.h) != 0  # a step is already in flight
        assert lib().amr_step_end(split.h, C.byref(active)) == 0
    for q in range(len(cfgs)):
        g, gq = (groups[0], q) if q < 5 else (groups[1], q - 5)
        assert one.info(q) == g.info(gq)
        d = compare_states(gpu_state(one, q), gpu_state(g, gq))
        assert not d, f"instance {q}: " + "; ".join(d[:4])
        if q >= 5:
            assert split.info(q - 5) == g.info(gq)
            d = compare_states(gpu_state(split, q - 5), gpu_state(g, gq))
            assert not d, f"instance {q} (begin/end): " + "; ".join(d[:4])


def test_failing_instances_are_isolated():
    """At p = 1 the reference aborts many seeded rotors early (p <= 0 after the
    first refinements).  Each such instance must stop with the oracle's code at
    the oracle's step; the survivors must equal their alone runs."""
    import amr2d_mhd_amd as A
    from amr2d_mhd_amd.ensemble import instance_config

    steps = 4
    cfgs, expect = [], []
    for i in range(16):
        cfg = instance_config(i, p=1.0)
        cfg["maxlevs"] = min(cfg["maxlevs"], 4)
        dom = O.Domain(**cfg)
        code, at = dom.init(), 0
        for s in range(steps):
            if code:
                break
            code, at = dom.step(), s + 1
        cfgs.append(cfg)
        expect.append((code, at if code else None))
    assert any(e[0] for e in expect) and not all(e[0] for e in expect)
    batch = A.Simulation(cfgs)
    batch.init()  # instances that fail here are retired, the call succeeds
    got_step = [0 if batch.info(q)["status"] else None for q in range(len(cfgs))]
    for s in range(steps):
        batch.step()
        for q in range(len(cfgs)):
            if got_step[q] is None and batch.info(q)["status"]:
                got_step[q] = s + 1
    for q, cfg in enumerate(cfgs):
        inf = batch.info(q)
        assert (inf["status"], got_step[q]) == expect[q], f"instance {q}: gpu {inf['status']} at {got_step[q]}, " \
                                                          f"oracle {expect[q]}"
        if inf["status"] == 0:
            d = compare_states(gpu_state(batch, q), gpu_state(_alone(cfg, steps)))
            assert not d, f"survivor {q}: " + "; ".join(d[:4])


def test_single_instance_failure_still_raises():
    import amr2d_mhd_amd as A
    from amr2d_mhd_amd.ensemble import instance_config

    cfg = None
    for i in range(64):
        c = instance_config(i, p=1.0)
        c["maxlevs"] = 3
        dom = O.Domain(**c)
        if dom.init() != 0:
            cfg = c
            break
    assert cfg is not None
    sim = A.Simulation(cfg)
    with pytest.raises(A.AmrError) as ei:
        sim.init()
    assert ei.value.code == 1 and sim.info()["status"] == 1


def test_full_ensemble_runs_and_samples_match():
    """All 256 seeded instances in one context for 12 steps.  One instance per
    depth equals its alone run; every instance that aborts does so with the
    oracle's code at the oracle's step."""
    import amr2d_mhd_amd as A
    from amr2d_mhd_amd.ensemble import ensemble_configs

    cfgs = ensemble_configs()
    steps = 12
    batch = A.Simulation(cfgs)
    batch.init()
    failed_at = {q: 0 for q in range(len(cfgs)) if batch.info(q)["status"]}
    for s in range(steps):
        batch.step()
        for q in range(len(cfgs)):
            if q not in failed_at and batch.info(q)["status"]:
                failed_at[q] = s + 1
    picked = {}
    for q, c in enumerate(cfgs):
        if q not in failed_at:
            picked.setdefault(c["maxlevs"], q)
    assert len(picked) == 4
    for q in picked.values():
        alone = _alone(cfgs[q], steps)
        assert alone.info()["status"] == 0
        d = compare_states(gpu_state(batch, q), gpu_state(alone))
        assert not d, f"instance {q}: " + "; ".join(d[:4])
    assert len(failed_at) < len(cfgs) // 8
    for q, at in failed_at.items():
        dom = O.Domain(**cfgs[q])
        code, k = dom.init(), 0
        while not code and k < at:
            code, k = dom.step(), k + 1
        assert (code, k) == (batch.info(q)["status"], at), f"instance {q}: oracle {code} at {k}, gpu " \
                                                             f"{batch.info(q)['status']} at {at}"


ROTOR_LV5 = dict(ic="rotor", limiter="MC", flux="HLLD", cfl=0.5, nx=16, nghosts=2, maxlevs=5, lenx=1.0)


def test_slot_arena_exhaustion_raises():
    """A context whose slot arena is too small for the tree its instance grows
    stops that instance with AMR_ERR_CAPACITY (12) and, having no instance
    left running, raises it."""
    import amr2d_mhd_amd as A

    sim = A.Simulation(ROTOR_LV5, slot_capacity=24)
    with pytest.raises(A.AmrError) as ei:
        sim.init()
        sim.run(200)
    assert ei.value.code == 12
    assert sim.info()["status"] == 12
    sim.close()


def test_slot_arena_exhaustion_is_isolated():
    """In a batch sharing one arena, the instance that runs out of slots stops
    with AMR_ERR_CAPACITY; an instance that needs no new slot (one level:
    never refines) keeps running and equals its alone run bit for bit."""
    import amr2d_mhd_amd as A

    quiet = dict(ic="BWx", limiter="MC", flux="RUSA", cfl=0.4, nx=16, nghosts=2, maxlevs=1)
    steps = 20  # the one-level Brio-Wu run is still going after 20 steps
    batch = A.Simulation([ROTOR_LV5, quiet], slot_capacity=28)
    batch.init()
    for _ in range(steps):
        assert batch.step() >= 1
    assert batch.info(0)["status"] == 12
    assert batch.info(1)["status"] == 0
    alone = _alone(quiet, steps)
    d = compare_states(gpu_state(batch, 1), gpu_state(alone))
    assert not d, "; ".join(d[:4])
    batch.close()
    alone.close()
