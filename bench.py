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
  valu          the same kernel's executed FP64 FLOPs per launch (counter-
                measured per zone) / mean launch time, against 78.6 TFLOP/s;
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
# profiles/stage_traffic_rotor_lv5.md (267.6 read + 270.0 written, 768-lane variant
# specialised to HLLD/MUSCL)
STAGE_TRAFFIC_BYTES_PER_ZONE = 537.6
HBM_PEAK_GBS = 8000.0
# FP64 FLOPs one stage launch executes per leaf zone: rocprofv3 counters
# (SQ_INSTS_VALU_FLOPS_FP64 + SQ_INSTS_VALU_FLOPS_FP64_TRANS) x 64 lanes per
# dispatch / (256 x leaves), median over the stage dispatches of this workload
# (profiles/fp64_flops_rotor_lv5_adjoint_512.md, 768-lane variant specialised
# to HLLD/MUSCL); FP64 division and sqrt count with their full instruction
# sequences, a wave counts every branch any of its lanes takes, and the whole
# 20x20 tile is computed
STAGE_FP64_FLOPS_PER_ZONE = 3956.0
FP64_PEAK_TFLOPS = 78.6


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


def ensemble_bench(steps, full, n_inst, plain=True, warmup=0, groups=2):
    """Batched ensemble of seeded parameterised rotors against the same
    instances run one after another, each alone in its own context (plain
    launches).  The batch is split into `groups` contexts of consecutive
    instances run round robin (amr_run_group), so one group's host work
    overlaps the other groups' device work.  Both legs time exactly `steps`
    steps of every instance from its initial condition, so their
    zone-updates are identical; with `full` the batch also runs every
    instance to t = 0.2."""
    import torch

    import amr2d_mhd_amd as A
    from amr2d_mhd_amd.ensemble import SEED, ensemble_configs

    cfgs = ensemble_configs(n_inst)
    groups = max(1, min(groups, len(cfgs)))
    bounds = [round(len(cfgs) * k / groups) for k in range(groups + 1)]

    def make_batch():
        sims = [A.Simulation(cfgs[bounds[k]:bounds[k + 1]]) for k in range(groups)]
        for b in sims:
            b.init()
        return sims

    def statuses(sims):
        return [b.info(q)["status"] for b in sims for q in range(b.n)]

    batch = make_batch()
    if warmup:
        A.run_group(batch, warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    nb, zb = A.run_group(batch, steps)
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    failed = [q for q, st in enumerate(statuses(batch)) if st]
    for b in batch:
        b.close()
    tp, zp = 0.0, 0
    for q, cfg in enumerate(cfgs if plain and not warmup else []):
        sim = A.Simulation(cfg)
        try:
            sim.init()
        except A.AmrError:
            sim.close()
            continue  # aborted in the initial regrid: no zone-updates in either leg
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        try:
            _, z = sim.run(steps)
        except A.AmrError as e:  # the instance aborts alone, as in the batch
            z = e.zone_updates
        torch.cuda.synchronize()
        tp += time.perf_counter() - t0
        zp += z
        sim.close()
    out = {"instances": len(cfgs), "seed": SEED, "aborted_instances": failed,
           "workload": "seeded rotor_p ensemble: rho_in U[5,20], v0 U[1,3], Bx U[0.5,2]/sqrt(4pi), p 5, "
                       "max level U{3..6}, MC-MUSCL, HLLC (even) / HLLD (odd), CT, RK2, CFL 0.5",
           "head_to_head_steps": steps, "batched_warmup_steps": warmup,
           "batched": {"zone_updates": zb, "seconds": tb, "value": zb / tb, "groups": groups},
           "plain_sequential": {"zone_updates": zp, "seconds": tp, "value": zp / tp if tp else None},
           "speedup": (zb / tb) / (zp / tp) if tp and zp else None}
    if full:
        batch = make_batch()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        n, z = A.run_group(batch, -1)
        torch.cuda.synchronize()
        tf = time.perf_counter() - t0
        ts = [b.info(q)["t"] for b in batch for q in range(b.n)]
        out["full_run"] = {"steps": n, "zone_updates": z, "seconds": tf, "value": z / tf,
                           "aborted_instances": [q for q, st in enumerate(statuses(batch)) if st],
                           "t_min": min(ts), "t_max": max(ts)}
        for b in batch:
            b.close()
    return out


OT = dict(ic="OT", limiter="WENO5", flux="HLLD", cfl=0.4, nx=16, nghosts=4, maxlevs=6, lenx=1.0, periodic=True,
          rk_stages=3)


def ot_bench(steps, warmup):
    """Orszag-Tang vortex (build-defined: periodic [0,1]^2, 4 roots of 16x16,
    6 levels, WENO5 + HLLD, SSP-RK3, CT, gamma 5/3): `warmup` untimed steps
    from t = 0, then `steps` timed steps (a negative count runs to t = 0.5)."""
    import torch

    import amr2d_mhd_amd as A

    sim = A.Simulation(OT)
    sim.init()
    if warmup:
        sim.run(warmup)
    torch.cuda.synchronize()
    # host-side timers only (no device events): the fill-program compiles and
    # the whole topology rebuilds they are part of
    sim.profile(("host_compile", "host_topology"))
    t0 = time.perf_counter()
    n, zu = sim.run(steps)
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    prof = sim.profile_read()
    info = sim.info()
    return {"workload": "Orszag-Tang, periodic, 4 roots of 16x16, 6 levels, WENO5 + HLLD, SSP-RK3, CT, CFL 0.4",
            "warmup": warmup, "steps": n, "zone_updates": zu, "seconds": secs, "value": zu / secs if secs else None,
            "unit": "zone-updates/s", "t_reached": info["t"], "leaves_at_end": info["nleaf"],
            "host_compile_ms": prof["host_compile"][0], "compiles": prof["host_compile"][1],
            "host_topology_ms": prof["host_topology"][0]}


def adjoint_bench(n, steps, every, fd_dirs, flux="HLLE"):
    """Frozen-mesh adjoint (build-defined workload): Orszag-Tang controls on a
    uniform periodic n x n grid (smoothly modulated and seeded-noise density,
    pressure and velocity, so no limiter evaluation starts on a kink),
    minmod-MUSCL + HLLE, CT, RK2, `steps` steps at a dt sequence frozen from
    the forward run, checkpoints every `every` steps.  J = sum (rho_T -
    rho_target)^2, target = the same set-up with the velocity scaled by 1.05.
    Reports the forward and backward cost per step, checkpointed and with
    every step's tape kept from the forward (every = 0), the per-call cost of
    the step, taped step, taped reverse and recomputing reverse and, over
    `fd_dirs` seeded random directions, the relative gap between the adjoint
    directional derivative and central finite differences at eps 1e-6 and
    1e-8."""
    import torch

    from amr2d_mhd_amd import adjoint as AD

    g = 5.0 / 3.0
    mesh = AD.FrozenMesh(n, 1.0 / n, g, flux)
    ctrl = AD.orszag_tang_controls(n, modulation=0.05, noise=1e-3)
    U = AD.cons_from_controls(*ctrl, g).contiguous()
    dts, Uc, Bxc, Byc = [], U, ctrl[3], ctrl[4]
    for _ in range(steps):
        dts.append(AD.cfl_dt(Uc, g, 1.0 / n, 0.4))
        Uc, Bxc, Byc = mesh.step(Uc, Bxc, Byc, dts[-1])
    tgt = AD.orszag_tang_controls(n, vscale=1.05, modulation=0.05, noise=1e-3)
    with torch.no_grad():
        UT, _, _ = AD.rollout(AD.cons_from_controls(*tgt, g).contiguous(), tgt[3], tgt[4], mesh, dts, every)
        target = UT[0].clone()

    def objective(c, ev=every):
        U0 = AD.cons_from_controls(*c, g).contiguous()
        UT, _, _ = AD.rollout(U0, c[3].contiguous(), c[4].contiguous(), mesh, dts, ev)
        return ((UT[0] - target) ** 2).sum()

    def timed(ev):
        ls = [t.clone().requires_grad_(True) for t in ctrl]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        J = objective(ls, ev)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        J.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        return ls, J, t1 - t0, t2 - t1

    timed(every)  # warm-up (allocator, first launches)
    leaves, J, tf, tb = timed(every)
    _, J0, tf0, tb0 = timed(0)
    # per-step kernel cost of one forward step, one taped reverse step and one
    # recomputing reverse step on the same state
    reps = 20
    tape = mesh.new_tapes(1, U.device)[0]
    gU, gBx, gBy = torch.ones_like(U), torch.ones_like(ctrl[3]), torch.ones_like(ctrl[4])

    def per_call(fn):
        fn()
        torch.cuda.synchronize()
        a0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - a0) / reps * 1e3

    step_ms = per_call(lambda: mesh.step(U, ctrl[3], ctrl[4], dts[0]))
    step_tape_ms = per_call(lambda: mesh.step(U, ctrl[3], ctrl[4], dts[0], tape))
    vjp_ms = per_call(lambda: mesh.vjp(tape, ctrl[3], ctrl[4], dts[0], gU, gBx, gBy))
    step_vjp_ms = per_call(lambda: mesh.step_vjp(U, ctrl[3], ctrl[4], dts[0], gU, gBx, gBy))
    out = {"workload": f"frozen uniform periodic {n}x{n}, Orszag-Tang controls, minmod-MUSCL + {flux}, CT, RK2, "
                       f"{steps} steps, checkpoints every {every}",
           "J": float(J), "forward_ms_per_step": tf / steps * 1e3,
           "backward_ms_per_step": tb / steps * 1e3,
           "backward_over_forward": tb / tf,
           "note": "backward recomputes every checkpoint segment, writing each step's tape, then runs the "
                   "taped reverse sweep",
           "taped_every_step": {"J": float(J0), "forward_ms_per_step": tf0 / steps * 1e3,
                                "backward_ms_per_step": tb0 / steps * 1e3, "backward_over_forward": tb0 / tf0,
                                "tape_bytes": steps * mesh.tape_doubles() * 8},
           "step_ms": step_ms, "step_tape_ms": step_tape_ms, "vjp_tape_ms": vjp_ms, "step_vjp_ms": step_vjp_ms,
           "vjp_tape_over_step": vjp_ms / step_ms}
    if fd_dirs:
        gen = torch.Generator(device="cpu").manual_seed(20261202)
        errs = {1e-6: [], 1e-8: []}
        for _ in range(fd_dirs):
            dirs = [torch.randn(t.shape, generator=gen, dtype=torch.float64).to(t.device) * t.abs().mean()
                    for t in ctrl]
            ad = float(sum((l.grad * d).sum() for l, d in zip(leaves, dirs)))
            for eps in errs:
                with torch.no_grad():
                    jp = objective([t + eps * d for t, d in zip(ctrl, dirs)])
                    jm = objective([t - eps * d for t, d in zip(ctrl, dirs)])
                fd = float(jp - jm) / (2 * eps)
                errs[eps].append(abs(fd - ad) / abs(fd))
        import statistics
        out["fd_check"] = {f"eps_{e:.0e}": {"directions": len(v), "median_rel_err": statistics.median(v),
                                            "max_rel_err": max(v)} for e, v in errs.items()}
    mesh.close()
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
    ap.add_argument("--ens-groups", type=int, default=2,
                    help="contexts the batch is split into, run round robin (1: one context)")
    ap.add_argument("--only-ensemble", action="store_true",
                    help="time only the batched ensemble (for profiling); prints its JSON object")
    ap.add_argument("--ot-steps", type=int, default=0,
                    help="also time the Orszag-Tang workload for this many steps (-1: to t = 0.5)")
    ap.add_argument("--ot-warmup", type=int, default=0)
    ap.add_argument("--adjoint", type=int, default=0, metavar="N",
                    help="also run the frozen-mesh adjoint on an N x N grid")
    ap.add_argument("--adj-steps", type=int, default=400)
    ap.add_argument("--adj-every", type=int, default=20)
    ap.add_argument("--adj-fd", type=int, default=0, help="finite-difference check directions")
    ap.add_argument("--adj-flux", default="HLLE", choices=["HLLE", "HLLD"])
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
                                        groups=args.ens_groups,
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
        "valu": {"bound": "fp64", "kernel": "stage_kernel<20,2>",
                 "achieved": STAGE_FP64_FLOPS_PER_ZONE * zones_per_launch / (stage_ms / stage_n * 1e-3) / 1e12,
                 "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                 "frac": STAGE_FP64_FLOPS_PER_ZONE * zones_per_launch / (stage_ms / stage_n * 1e-3) / 1e12
                 / FP64_PEAK_TFLOPS,
                 "flops_per_zone": STAGE_FP64_FLOPS_PER_ZONE},
        "cpu_baseline": None,
    }
    if args.ot_steps:
        out["orszag_tang"] = ot_bench(args.ot_steps, args.ot_warmup)
    if args.adjoint:
        out["adjoint"] = adjoint_bench(args.adjoint, args.adj_steps, args.adj_every, args.adj_fd, args.adj_flux)
    if args.ensemble:
        out["ensemble"] = ensemble_bench(args.ens_steps, args.ens_full, args.ensemble, warmup=args.ens_warmup,
                                         groups=args.ens_groups)
    if not args.no_cpu:
        threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(args.cpu_steps, threads)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
