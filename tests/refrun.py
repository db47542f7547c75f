"""Helpers that drive the compiled reference harness (oracle/_ref/<variant>/refrun)
and read its outputs.  Test infrastructure only; skipped where the compiled
reference is absent."""
import os
import subprocess
import tempfile

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DIR = os.path.join(REPO, "oracle", "_ref")

LOG_DTYPE = np.dtype([("step", "i4"), ("la_pre", "i4"), ("nleaf_pre", "i4"), ("la_post", "i4"),
                      ("nleaf_post", "i4"), ("pad", "i4"), ("dt", "f8"), ("t", "f8"), ("hf_pre", "u8"),
                      ("ht_pre", "u8"), ("hf_post", "u8"), ("ht_post", "u8")])

STATE_FIELDS = [("U", 8), ("Us", 8), ("W", 8), ("Bi", 2), ("Bis", 2), ("ff", 8), ("gg", 8), ("EMF", 1)]


def refrun_path(ng, lv, omp=False, ct2=False):
    return os.path.join(REF_DIR, f"ng{ng}_lv{lv}{'_omp' if omp else ''}{'_ct2' if ct2 else ''}", "refrun")


def have_ref(ng=2, lv=4, omp=False, ct2=False):
    return os.access(refrun_path(ng, lv, omp, ct2), os.X_OK)


def input_line(nx, lenx, cfl, ic, lim, flux):
    return f"{nx} {nx} 8|{lenx} {lenx}|{cfl}|{ic} {lim} {flux}"


def run(workdir, ng, lv, nx=16, lenx=1.0, cfl=0.4, ic="BWx", lim="MC", flux="RUSA", steps=None, dumps=(),
        trace=None, randtree=None, omp=False, timing=False, env=None, timeout=3600, ct2=False):
    exe = refrun_path(ng, lv, omp, ct2)
    args = [exe, "run", "--dir", workdir, "--input", input_line(nx, lenx, cfl, ic, lim, flux),
            "--log", os.path.join(workdir, "log.bin")]
    if steps is not None:
        args += ["--steps", str(steps)]
    if dumps:
        args += ["--dump", ",".join(str(s) for s in dumps)]
    if trace is not None:
        args += ["--trace", str(trace)]
    if randtree is not None:
        args += ["--randtree", "%d,%g,%g,%d" % randtree]
    if timing:
        args += ["--timing", os.path.join(workdir, "timing.json")]
    os.makedirs(workdir, exist_ok=True)
    p = subprocess.run(args, capture_output=True, text=True, env=env, timeout=timeout)
    return p.returncode, p.stderr


def read_log(workdir):
    return np.fromfile(os.path.join(workdir, "log.bin"), dtype=LOG_DTYPE)


def read_state(path):
    with open(path, "rb") as fp:
        hdr = np.frombuffer(fp.read(32), np.int32)
        assert hdr[0] == 0x53524D41, "bad state magic"
        NT, ng, nA = int(hdr[2]), int(hdr[3]), int(hdr[4])
        t, dt = np.frombuffer(fp.read(16), np.float64)
        lp = np.frombuffer(fp.read(4 * 8 * nA), np.int32).reshape(nA, 8).copy()
        ic = np.frombuffer(fp.read(4 * 2 * nA), np.int32).reshape(nA, 2).copy()
        lf = np.frombuffer(fp.read(4 * nA), np.int32).copy()
        st = dict(NT=NT, ng=ng, lastActive=nA, step=int(hdr[6]), maxlevs=int(hdr[7]), t=float(t), dt=float(dt),
                  lp=lp, icoord=ic, leafs=lf)
        for name, nv in STATE_FIELDS:
            st[name] = np.frombuffer(fp.read(8 * nA * nv * NT * NT), np.float64).reshape(nA, nv, NT, NT).copy()
    return st


def vec(kind, payload, ng=2, lv=5):
    """Run one reference function over a binary payload; returns output bytes."""
    exe = refrun_path(ng, lv)
    with tempfile.TemporaryDirectory() as td:
        inp, out = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(inp, "wb") as fp:
            fp.write(payload)
        p = subprocess.run([exe, "vec", kind, inp, out], capture_output=True, text=True, timeout=600)
        if p.returncode != 0:
            raise RuntimeError(f"refrun vec {kind} failed: {p.stderr}")
        with open(out, "rb") as fp:
            return fp.read()


def fixed_str(s, n):
    b = s.encode()
    assert len(b) < n
    return b + b"\0" * (n - len(b))
