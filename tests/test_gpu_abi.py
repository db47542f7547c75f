"""The C ABI from outside Python and on a caller's stream: the reference's
time loop written against include/amr2d_mhd.h (amr2d_mhd_amd/examples/
reference_loop, built by the library's Makefile) reaches the same state as
the Python binding, and a context moved onto a PyTorch stream (amr_use_stream)
computes the same run bit for bit."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from gpu_util import compare_states, gpu_state

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ROTOR4 = dict(ic="rotor", limiter="MC", flux="HLLD", cfl=0.5, nx=16, nghosts=2, maxlevs=4)


def _fnv1a(b):
    h = 1469598103934665603
    for x in b:
        h = ((h ^ x) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def test_reference_loop_example_matches_python_binding():
    """150 rotor steps at 4 levels through the C++ driver and through the
    Python binding: step count, t, dt, slot and leaf counts and the hash of
    U over the active slots agree."""
    import amr2d_mhd_amd as A

    exe = os.path.join(REPO, "amr2d_mhd_amd", "examples", "reference_loop")
    assert os.access(exe, os.X_OK), "examples/reference_loop not built (make -C amr2d_mhd_amd/csrc)"
    print("case: reference_loop rotor MC HLLD lv4 150 steps", flush=True)
    p = subprocess.run([exe, "rotor", "MC", "HLLD", "4", "2", "150"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    v = p.stdout.split()
    got = dict(zip(v[0::2], v[1::2]))
    sim = A.Simulation(ROTOR4)
    sim.init()
    for _ in range(150):
        sim.step()
    inf = sim.info()
    assert int(got["steps"]) == 150
    assert float(got["t"]) == inf["t"] and float(got["dt"]) == inf["dt"]
    assert int(got["last_active"]) == inf["last_active"] and int(got["nleaf"]) == inf["nleaf"]
    U = np.ascontiguousarray(sim.download("U", nslots=inf["last_active"]))
    assert int(got["hashU"]) == _fnv1a(U.tobytes())
    sim.close()


def test_context_on_caller_streams_is_bitwise_unchanged():
    """A rotor run moved onto a PyTorch side stream, then onto the null
    stream, then back onto its own stream, equals the run that stays on its
    own stream, state for state; switching between amr_step_begin and
    amr_step_end is refused."""
    import torch

    import amr2d_mhd_amd as A
    from amr2d_mhd_amd import AmrError
    from amr2d_mhd_amd._lib import lib

    ref = A.Simulation(ROTOR4)
    ref.init()
    sim = A.Simulation(ROTOR4)
    side = torch.cuda.Stream()
    sim.use_stream(side)
    sim.init()
    for phase, stream, steps in (("side stream", side, 40), ("null stream", 0, 30), ("own stream", None, 30)):
        print("case: use_stream", phase, flush=True)
        if phase != "side stream":
            sim.use_stream(stream)
        for _ in range(steps):
            ref.step()
            sim.step()
        torch.cuda.synchronize()
        d = compare_states(gpu_state(ref), gpu_state(sim))
        assert not d, f"{phase}: " + "; ".join(d[:4])
        assert ref.info()["t"] == sim.info()["t"]
    assert lib().amr_step_begin(sim.h) == 0
    with pytest.raises(AmrError):
        sim.use_stream(side)
    n = C.c_int(0)
    assert lib().amr_step_end(sim.h, C.byref(n)) == 0  # the step begun above
    ref.step()
    sim.use_stream(side)
    sim.step()
    ref.step()
    d = compare_states(gpu_state(ref), gpu_state(sim))
    assert not d, "after a refused switch: " + "; ".join(d[:4])
    sim.use_stream(None)
    sim.close()
    ref.close()


def test_profile_event_sampling():
    """amr_profile_sample: with stride 3 one stage launch in three is timed
    (two launches per step), with a mean close to the every-launch mean."""
    import amr2d_mhd_amd as A

    means = {}
    for stride in (1, 3):
        sim = A.Simulation(ROTOR4)
        sim.init()
        sim.run(100)
        sim.profile(("stage",), stride=stride)
        sim.run(60)
        ms, n = sim.profile_read()["stage"]
        assert n == 120 // stride, (stride, n)
        means[stride] = ms / n
        sim.close()
    assert abs(means[3] - means[1]) < 0.5 * means[1]
