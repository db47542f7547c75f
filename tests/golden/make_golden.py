"""Generate the golden fixtures in tests/golden from the compiled reference.

Each fixture is the reference harness's per-step log of a run: step, applied
dt, t, leaf counts and FNV-1a hashes of every slot field (U Us Bi Bis ff gg
EMF, [nb][var][j][i] order) and of the quadtree after each step.  Run from the
repository root after oracle/build_ref.sh:  python tests/golden/make_golden.py
"""
import os
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import refrun  # noqa: E402

RUNS = [
    ("run_BWx_MC_RUSA_lv4", dict(ng=2, lv=4, ic="BWx", lim="MC", flux="RUSA", cfl=0.4), None),
    ("run_rotor_MC_HLLD_lv5", dict(ng=2, lv=5, ic="rotor", lim="MC", flux="HLLD", cfl=0.5), 150),
    ("run_rotor_MC_HLLC_lv4", dict(ng=2, lv=4, ic="rotor", lim="MC", flux="HLLC", cfl=0.5), 60),
    ("run_rotor_MC_HLLE_lv4", dict(ng=2, lv=4, ic="rotor", lim="MC", flux="HLLE", cfl=0.5), 60),
    ("run_BWy_minmod_RUSA_lv3", dict(ng=2, lv=3, ic="BWy", lim="minmod", flux="RUSA", cfl=0.4), 60),
    ("run_SST_Superbee_RUSA_lv3", dict(ng=2, lv=3, ic="SST", lim="Superbee", flux="RUSA", cfl=0.4), 60),
    ("run_rotor_WENO5_HLLD_lv3_ng4", dict(ng=4, lv=3, ic="rotor", lim="WENO5", flux="HLLD", cfl=0.4), 40),
    # CTtype 2 (the upwind corner EMF of CT2D.cpp:28-75), reference built with CTtype = 2
    ("run_rotor_MC_HLLD_lv4_ct2", dict(ng=2, lv=4, ic="rotor", lim="MC", flux="HLLD", cfl=0.5, ct2=True), 300),
    ("run_rotor_minmod_HLLE_lv4_ct2", dict(ng=2, lv=4, ic="rotor", lim="minmod", flux="HLLE", cfl=0.5, ct2=True), 200),
    ("run_rotor_WENO5_HLLD_lv3_ng4_ct2", dict(ng=4, lv=3, ic="rotor", lim="WENO5", flux="HLLD", cfl=0.4, ct2=True),
     60),
    # the bench workload's whole run, 4763 steps to t = 0.295 (GPU test against the reference
    # directly; about 2 h of one core)
    ("run_rotor_MC_HLLD_lv5_full", dict(ng=2, lv=5, ic="rotor", lim="MC", flux="HLLD", cfl=0.5), None),
]


def one(item):
    name, kw, steps = item
    with tempfile.TemporaryDirectory() as td:
        rc, err = refrun.run(td, steps=steps, timeout=4 * 3600, **kw)
        if rc != 0:
            raise RuntimeError(f"{name}: reference failed {rc} {err}")
        log = refrun.read_log(td)
    np.save(os.path.join(HERE, name + ".npy"), log)
    return name, len(log) - 1


if __name__ == "__main__":
    only = set(sys.argv[1:])
    todo = [r for r in RUNS if not only or r[0] in only]
    with ThreadPoolExecutor(max_workers=min(7, len(todo))) as ex:
        for name, n in ex.map(one, todo):
            print(name, n, "steps")
