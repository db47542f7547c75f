#!/usr/bin/env python3
"""Benchmark of the AMR2D-MHD hot path on one MI355X.

Workload (one GPU): the MHD rotor of the reference (IC2Dtype.cpp:75-111),
4 root blocks of 16x16, 5 AMR levels, MC-limited MUSCL + HLLD, CT, RK2,
CFL 0.5, FP64.  A "step" is one iteration of the reference time loop
(main.cpp:128-247) over the current leaf blocks: ghost fills, two RK stages,
CT, CFL reduction and regrid.  The run starts from the initial condition; W
untimed steps are taken, then exactly K steps are timed (defaults cover the
whole run to t = 0.295, 4763 steps).  The metric is leaf zone-updates per
second (one interior leaf cell advanced by one step).

Also reported:
  roofline      the stage kernel (reconstruction + Riemann + RK, the
                dominant kernel): algorithmic bytes per launch / mean launch
                time from HIP events on the engine's stream during the timed
                region, against the 8 TB/s HBM peak;
  cpu_baseline  the reference itself (compiled from its sources into
                oracle/_ref with its OpenMP loops) on the first steps of the
                same rotor run, on the host cores of this machine.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

ROTOR = dict(ic="rotor", limiter="MC", flux="HLLD", cfl=0.5, nx=16, nghosts=2, maxlevs=5, lenx=1.0)
ROTOR_STEPS = 4763  # steps of the full rotor run to t = 0.295 without FMA contraction (reference build = this engine)
# compulsory HBM traffic of one stage-kernel launch per leaf zone (FP64):
# stage 1 reads U (64 B) and face B (16 B), writes Us (64 B); stage 2 reads
# Us and U (128 B) and face B (16 B), writes U (64 B) -> mean 176 B
STAGE_BYTES_PER_ZONE = 176.0
# HBM bytes per leaf zone one stage launch actually moves, from rocprofv3 FETCH_SIZE
# (x2.0, copy-calibrated) + WRITE_SIZE (x1.0) on this workload:
# profiles/stage_traffic_rotor_lv5.md (272.7 read + 270.0 written)
STAGE_TRAFFIC_BYTES_PER_ZONE = 542.7
HBM_PEAK_GBS = 8000.0


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fp:
            for line in fp:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(steps, threads):
    """Time the compiled reference (OpenMP build) on the first `steps` steps of
    the rotor run; zone-updates counted from its own leaf log."""
    import numpy as np

    exe = os.path.join(HERE, "oracle", "_ref", "ng2_lv5_omp", "refrun")
    if not os.access(exe, os.X_OK):
        return None
    with tempfile.TemporaryDirectory() as td:
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        inp = "16 16 8|1.0 1.0|0.5|rotor MC HLLD"
        p = subprocess.run([exe, "run", "--dir", td, "--input", inp, "--steps", str(steps), "--log",
                            os.path.join(td, "log.bin"), "--timing", os.path.join(td, "timing.json")],
                           capture_output=True, text=True, env=env, timeout=900)
        if p.returncode != 0:
            return None
        with open(os.path.join(td, "timing.json")) as fp:
            tm = json.load(fp)
        dt = np.dtype([("step", "i4"), ("la_pre", "i4"), ("nleaf_pre", "i4"), ("la_post", "i4"),
                       ("nleaf_post", "i4"), ("pad", "i4"), ("dt", "f8"), ("t", "f8"), ("h", "u8", 4)])
        log = np.fromfile(os.path.join(td, "log.bin"), dtype=dt)
        zu = int((log["nleaf_pre"][1:].astype(np.int64) * 256).sum())
    return dict(value=zu / tm["loop_seconds"], unit="zone-updates/s", cores=threads, kind="reference",
                sample=f"reference OpenMP build, rotor lv5 MC+HLLD CFL 0.5, first {steps} steps "
                       f"({zu} zone-updates in {tm['loop_seconds']:.2f} s), host CPU: {cpu_model()}")


def ensemble_bench(steps, full, n_inst, plain=True, warmup=0):
    """Batched ensemble of seeded parameterised rotors against the same
    instances run one after another, each alone in its own context (plain
    launches).  Both legs time exactly `steps` steps of every instance from
    its initial condition, so their zone-updates are identical; with `full`
    the batch also runs every instance to t = 0.2."""
    import torch

    import amr2d_mhd_amd as A
    from amr2d_mhd_amd.ensemble import SEED, ensemble_configs

    cfgs = ensemble_configs(n_inst)
    batch = A.Simulation(cfgs)
    batch.init()
    if warmup:
        batch.run(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    nb, zb = batch.run(steps)
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    failed = [q for q in range(len(cfgs)) if batch.info(q)["status"]]
    batch.close()
    tp, zp = 0.0, 0
    for q, cfg in enumerate(cfgs if plain and not warmup else []):
        if q in failed:
            continue  # an aborted instance's partial step is not a zone-update in either leg
        sim = A.Simulation(cfg)
        sim.init()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, z = sim.run(steps)
        torch.cuda.synchronize()
        tp += time.perf_counter() - t0
        zp += z
        sim.close()
    out = {"instances": len(cfgs), "seed": SEED, "aborted_instances": failed,
           "workload": "seeded rotor_p ensemble: rho_in U[5,20], v0 U[1,3], Bx U[0.5,2]/sqrt(4pi), p 5, "
                       "max level U{3..6}, MC-MUSCL, HLLC (even) / HLLD (odd), CT, RK2, CFL 0.5",
           "head_to_head_steps": steps, "batched_warmup_steps": warmup,
           "batched": {"zone_updates": zb, "seconds": tb, "value": zb / tb},
           "plain_sequential": {"zone_updates": zp, "seconds": tp, "value": zp / tp if tp else None},
           "speedup": (zb / tb) / (zp / tp) if tp and zp else None}
    if full:
        batch = A.Simulation(cfgs)
        batch.init()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        n, z = batch.run(-1)
        torch.cuda.synchronize()
        tf = time.perf_counter() - t0
        ts = [batch.info(q)["t"] for q in range(len(cfgs))]
        out["full_run"] = {"steps": n, "zone_updates": z, "seconds": tf, "value": z / tf,
                           "aborted_instances": [q for q in range(len(cfgs)) if batch.info(q)["status"]],
                           "t_min": min(ts), "t_max": max(ts)}
        batch.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=ROTOR_STEPS - 57)
    ap.add_argument("--warmup", type=int, default=57)
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--ensemble", type=int, default=0, metavar="N",
                    help="also time a batched ensemble of N seeded rotors against N plain runs")
    ap.add_argument("--ens-steps", type=int, default=200)
    ap.add_argument("--ens-full", action="store_true", help="also run the batched ensemble to t = 0.2")
    ap.add_argument("--ens-warmup", type=int, default=0,
                    help="untimed batched steps before the timed ones (skips the plain leg)")
    ap.add_argument("--only-ensemble", action="store_true",
                    help="time only the batched ensemble (for profiling); prints its JSON object")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one 256 MiB device copy (counter calibration for rocprofv3 --pmc runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on exactly one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        return 2

    import torch

    import amr2d_mhd_amd as A

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.only_ensemble:
        print(json.dumps(ensemble_bench(args.ens_steps, args.ens_full, args.ensemble or 256, plain=False,
                                        warmup=args.ens_warmup)))
        return 0
    if args.copy_calibration:
        x = torch.empty(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        x.fill_(1.0)
        y.copy_(x)
        torch.cuda.synchronize()

    sim = A.Simulation(ROTOR)
    sim.init()
    if args.warmup:
        sim.run(args.warmup)
    torch.cuda.synchronize()
    sim.profile(("stage",))
    t0 = time.perf_counter()
    steps, zu = sim.run(args.steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = sim.profile_read()
    if steps != args.steps:
        print(f"bench.py: run ended after {steps} of {args.steps} timed steps", file=sys.stderr)
        return 1
    secs = t1 - t0
    value = zu / secs
    stage_ms, stage_n = prof["stage"]
    zones_per_launch = zu / steps  # every step launches the stage kernel twice over the same leaves
    achieved = STAGE_BYTES_PER_ZONE * zones_per_launch / (stage_ms / stage_n * 1e-3) / 1e9
    info = sim.info()
    out = {
        "metric": "zone-updates/s FP64 at 1 GPU (rotor, 5 AMR levels, MC+HLLD, CT)",
        "value": value,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": secs / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (reference rotor initial condition)",
        "config": {"workload": "MHD rotor, 4 root blocks of 16x16, 5 AMR levels, MC-MUSCL + HLLD, CT, "
                               "RK2, CFL 0.5, gamma 1.4, from t=0 (whole run to t=0.295 by default)",
                   "t_end_reached": info["t"], "leaves_at_end": info["nleaf"],
                   "zone_updates": zu, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": STAGE_TRAFFIC_BYTES_PER_ZONE * zones_per_launch,
                     "kernel": "stage_kernel<20,2>", "mean_launch_us": stage_ms / stage_n * 1e3,
                     "launches": stage_n, "bytes_per_zone": STAGE_BYTES_PER_ZONE,
                     "zones_per_launch": zones_per_launch},
        "cpu_baseline": None,
    }
    if args.ensemble:
        out["ensemble"] = ensemble_bench(args.ens_steps, args.ens_full, args.ensemble, warmup=args.ens_warmup)
    if not args.no_cpu:
        threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(args.cpu_steps, threads)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
