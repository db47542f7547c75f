"""Generate the oracle-written logs of the build-defined Orszag-Tang workload
(periodic domain, WENO5 + HLLD, SSP-RK3, CFL 0.4; the reference has no
periodic boundary, no Orszag-Tang state and no RK3, so its oracle is the
restatement in oracle/oracle.c, pinned by the closed-form checks of
tests/test_ot_periodic_cpu.py).

Each log is an .npz with
  steps: per step (0 = the initial state): step, slot and leaf counts, the dt
         the step applied, t after it, the FNV-1a hash of the quadtree, and
         every HASH_EVERY steps (and at the last step) the hash of every slot
         field (U Us Bi Bis ff gg EMF) -- 0 where not taken;
  abort: [code, block, i, j, failing step] of the run's own failure, or -1s.
Run from the repository root:  python tests/golden/make_oracle_logs.py [name ...]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle"))
import oracle as O  # noqa: E402

STEP_DTYPE = np.dtype([("step", "i4"), ("last_active", "i4"), ("nleaf", "i4"), ("pad", "i4"), ("dt", "f8"),
                       ("t", "f8"), ("hash_tree", "u8"), ("hash_fields", "u8")])
OT = dict(ic="OT", limiter="WENO5", flux="HLLD", cfl=0.4, nx=16, nghosts=4, periodic=True, rk_stages=3)
HASH_EVERY = 100
RUNS = {
    # the Orszag-Tang workload at its 6 levels: first 1000 steps (about an hour)
    "ot_WENO5_HLLD_rk3_lv6_1000": (dict(OT, maxlevs=6), 1000),
    # 5 levels, until the scheme's own p <= 0 failure (step 685)
    "ot_WENO5_HLLD_rk3_lv5_abort": (dict(OT, maxlevs=5), None),
}


def make(name):
    cfg, steps = RUNS[name]
    dom = O.Domain(**cfg)
    assert dom.init() == 0
    rows, abort = [], [-1] * 5

    def row(s, applied, fields):
        inf = dom.info()
        rows.append((s, inf["lastActive"], inf["nleaf"], 0, applied, inf["t"], dom.hash_tree(),
                     dom.hash_fields() if fields else 0))

    row(0, 0.0, True)
    s = 0
    while not dom.done() and (steps is None or s < steps):
        applied = dom.info()["dt"]
        rc = dom.step()
        if rc != 0:
            abort = list(dom.last_error()) + [s + 1]
            break
        s += 1
        row(s, applied, s % HASH_EVERY == 0)
        if s % 50 == 0:
            print(name, s, dom.info()["nleaf"], flush=True)
    if rows[-1][7] == 0:
        if abort[0] != -1:
            # the failed step has already changed the state: run again to the
            # last completed step for its field hash
            dom = O.Domain(**cfg)
            assert dom.init() == 0
            for _ in range(s):
                assert dom.step() == 0
        r = list(rows[-1])
        r[7] = dom.hash_fields()
        rows[-1] = tuple(r)
    np.savez(os.path.join(HERE, name + ".npz"), steps=np.array(rows, STEP_DTYPE), abort=np.array(abort, np.int64))
    print(name, "done", s, "steps, abort", abort, flush=True)


if __name__ == "__main__":
    for n in sys.argv[1:] or sorted(RUNS):
        make(n)
