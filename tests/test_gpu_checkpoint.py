"""Checkpoint / restart through the C ABI (amr_checkpoint_save/_load): a run
restarted from a checkpoint in a fresh context continues bit for bit as the
uninterrupted run (every slot field of every slot, the tree, t, dt, dt_next,
step count), for the reference's RK2 with mirror boundaries and for the
build-defined periodic SSP-RK3 with WENO5; corrupt or foreign files are
rejected with the instance left unchanged."""
import os

import numpy as np
import pytest

from gpu_util import compare_states, gpu_state

pytestmark = pytest.mark.gpu

ROTOR = dict(ic="rotor", limiter="MC", flux="HLLD", cfl=0.5, nx=16, nghosts=2, maxlevs=5)
OT = dict(ic="OT", limiter="WENO5", flux="HLLD", cfl=0.4, nx=16, nghosts=4, maxlevs=4, periodic=True, rk_stages=3)


def _same_run(a, b, label):
    ia, ib = a.info(), b.info()
    for k in ("t", "dt", "dt_next", "count", "last_active", "nleaf", "done", "status"):
        assert ia[k] == ib[k], f"{label}: {k} differs: {ia[k]!r} vs {ib[k]!r}"
    d = compare_states(gpu_state(a), gpu_state(b))
    assert not d, f"{label}: " + "; ".join(d[:4])


@pytest.mark.parametrize("cfg,before,after", [(ROTOR, 37, 63), (OT, 23, 27)])
def test_restart_from_checkpoint_is_bitwise_continuous(tmp_path, cfg, before, after):
    import amr2d_mhd_amd as A

    ref = A.Simulation(cfg)
    ref.init()
    ref.run(before + after)
    part = A.Simulation(cfg)
    part.init()
    part.run(before)
    path = tmp_path / "run.ckpt"
    part.save_checkpoint(path)
    part.close()
    resumed = A.Simulation(cfg)  # fresh context: other slot map, no program cache, untuned
    resumed.load_checkpoint(path)
    assert resumed.info()["count"] == before
    resumed.run(after)
    _same_run(ref, resumed, f"{cfg['ic']} restarted at step {before}")


def test_state_roundtrip_continues_bitwise():
    """load_state(state()) carries t and the dt lag: the resumed run equals
    the uninterrupted one."""
    import amr2d_mhd_amd as A

    ref = A.Simulation(ROTOR)
    ref.init()
    ref.run(50)
    part = A.Simulation(ROTOR)
    part.init()
    part.run(20)
    st = part.state()
    resumed = A.Simulation(ROTOR)
    resumed.load_state(st)
    resumed.run(30)
    _same_run(ref, resumed, "load_state(state())")


def test_bad_checkpoints_are_rejected(tmp_path):
    import amr2d_mhd_amd as A
    from amr2d_mhd_amd import AmrError

    src = A.Simulation(ROTOR)
    src.init()
    src.run(15)
    path = tmp_path / "good.ckpt"
    src.save_checkpoint(path)
    raw = path.read_bytes()
    (tmp_path / "short.ckpt").write_bytes(raw[: len(raw) // 2])
    flipped = bytearray(raw)
    flipped[len(raw) // 3] ^= 0x10
    (tmp_path / "flip.ckpt").write_bytes(bytes(flipped))

    dst = A.Simulation(ROTOR)
    dst.init()
    dst.run(5)
    before = gpu_state(dst)
    for name in ("short.ckpt", "flip.ckpt", "missing.ckpt"):
        with pytest.raises(AmrError) as e:
            dst.load_checkpoint(tmp_path / name)
        assert e.value.code == 14, name
    for cfg in (dict(ROTOR, flux="HLLE"), dict(ROTOR, ct_type=2)):  # another solver, another corner EMF
        other = A.Simulation(cfg)
        with pytest.raises(AmrError) as e:
            other.load_checkpoint(path)
        assert e.value.code == 10, cfg
        other.close()
    # the rejected loads left the instance as it was
    d = compare_states(gpu_state(dst), before)
    assert not d, "; ".join(d[:3])
    assert dst.info()["count"] == 5


def test_set_tree_rejects_corrupt_tree_unchanged():
    import amr2d_mhd_amd as A
    from amr2d_mhd_amd import AmrError

    sim = A.Simulation(ROTOR)
    sim.init()
    sim.run(10)
    lp, ic, lf = sim.get_tree()
    before = gpu_state(sim)
    bad = lp.copy()
    leaf = int(np.flatnonzero(lf[4:])[0]) + 4
    bad[leaf, 4] = lp.shape[0] + 7
    with pytest.raises(AmrError) as e:
        sim.set_tree(bad, ic, lf)
    assert e.value.code == 10
    with pytest.raises(AmrError):
        sim.upload("U", np.zeros((2 * 4 ** 5 + 1, 8, 20, 20)))
    d = compare_states(gpu_state(sim), before)
    assert not d, "; ".join(d[:3])
    # downloads past the used blocks read zeros and take no slot
    extra = sim.download("U", nslots=lp.shape[0] + 3)
    assert not extra[lp.shape[0]:].any()
    sim.run(3)


def test_checkpoint_reloaded_into_a_context_that_ran_on(tmp_path):
    """A checkpoint loaded back into the context that wrote it, after that
    context ran on and allocated blocks past the checkpoint's lastActive,
    continues as the uninterrupted run: the blocks past the loaded tree start
    blank, so later refinements allocate the same son groups."""
    import amr2d_mhd_amd as A

    ref = A.Simulation(ROTOR)
    ref.init()
    ref.run(150)
    sim = A.Simulation(ROTOR)
    sim.init()
    sim.run(40)
    path = tmp_path / "run.ckpt"
    sim.save_checkpoint(path)
    n_saved = sim.info()["last_active"]
    sim.run(260)
    assert sim.info()["last_active"] > n_saved
    sim.load_checkpoint(path)
    sim.run(110)
    _same_run(ref, sim, "reloaded into the context that ran on")
