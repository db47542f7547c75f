#!/usr/bin/env python3
"""Benchmark of the AMR2D-MHD hot path on one MI355X.

Metric: zone-updates/s FP64 at 1 GPU, batched vs plain; adjoint/forward step
cost.  One line of JSON carries all three:

  value         the plain method: the MHD rotor of the reference
                (IC2Dtype.cpp:75-111), 4 root blocks of 16x16, 5 AMR levels,
                MC-limited MUSCL + HLLD, CT, RK2, CFL 0.5.  A "step" is one
                iteration of the reference time loop (main.cpp:128-247) over
                the current leaf blocks: ghost fills, two RK stages, CT, CFL
                reduction and regrid.  The run starts from the initial
                condition; W untimed steps, then exactly K timed steps
                (defaults: the whole run, 4763 steps).  A zone-update is one
                interior leaf cell advanced by one step.  The stage-kernel
                tuner's decisions come from an untimed run, so no trial launch
                falls in the timed window;
  whole_run     (when the window is not the whole run) every step of the run;
  batched_vs_plain  256 seeded rotors of depth 3-6 advanced together against
                the same instances run one after another (same zone-updates);
  adjoint_over_forward  the frozen-mesh adjoint, 512x512, 400 steps,
                checkpoints every 20: backward / forward time, HLLE and HLLD;
  roofline      the stage kernel (reconstruction + Riemann + RK, the
                dominant kernel): algorithmic bytes per launch / mean launch
                time from HIP events on the engine's stream during the timed
                window, against the 8 TB/s HBM peak;
  valu          the same kernel's executed FP64 FLOPs per launch (counter-
                measured per zone) / mean launch time, against 78.6 TFLOP/s;
                op_counted: the algorithmic operations per zone (each +, -, *,
                /, sqrt of the source once, counted on the host on the run's
                own states) at the same launch time;
  cpu_baseline  the reference itself (compiled from its sources into
                oracle/_ref with its OpenMP loops) over the same steps of the
                same run, on the host's cores, in its FMA-free build and with
                its makefile's own flags.
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
# (x2.0, copy-calibrated) + WRITE_SIZE (x1.0) over the 9526 stage launches of the
# whole run, with the variants the tuner picks: profiles/stage_traffic_rotor_lv5.md
# (311.5 read + 286.5 written)
STAGE_TRAFFIC_BYTES_PER_ZONE = 598.0
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


def cpu_threads():
    """Threads for the CPU baseline: the host's CPU share for this GPU as the
    environment states it (OMP_NUM_THREADS), else every core of the affinity
    mask."""
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return env or len(os.sched_getaffinity(0))


def cpu_baseline(first, steps, threads, variant="ng2_lv5_omp"):
    """Time the compiled reference (an OpenMP build in oracle/_ref) on steps
    [first, first + steps) of the rotor run from t = 0: the harness stamps the
    end of every step of the reference's own loop, and the zone-updates of the
    window are its leaf counts x 256.  The reference keeps t and the dt lag in
    locals of its main loop, so it always starts from t = 0."""
    import numpy as np

    exe = os.path.join(HERE, "oracle", "_ref", variant, "refrun")
    if not os.access(exe, os.X_OK):
        return None
    with tempfile.TemporaryDirectory() as td:
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        inp = "16 16 8|1.0 1.0|0.5|rotor MC HLLD"
        st = os.path.join(td, "steps.bin")
        p = subprocess.run([exe, "run", "--dir", td, "--input", inp, "--steps", str(first + steps),
                            "--step-times", st], capture_output=True, text=True, env=env, timeout=900)
        if p.returncode != 0:
            return None
        rec = np.fromfile(st, np.float64).reshape(-1, 2)
    if len(rec) < first + steps:
        return None
    t0 = rec[first - 1, 0] if first > 0 else 0.0
    secs = rec[first + steps - 1, 0] - t0
    zu = int(rec[first:first + steps, 1].sum()) * 256
    return dict(value=zu / secs, unit="zone-updates/s", zone_updates=zu, seconds=secs, steps=[first, first + steps])


def cpu_baselines(first, steps, max_steps):
    """The reference on the host's cores over the GPU's timed window (its
    first `max_steps` steps when the window is longer): the FMA-free build
    (the oracle's arithmetic) and the build with the reference makefile's own
    flags."""
    threads = cpu_threads()
    n = min(steps, max_steps)
    plain = cpu_baseline(first, n, threads, "ng2_lv5_omp")
    if plain is None:
        return None
    native = cpu_baseline(first, n, threads, "ng2_lv5_omp_native")
    window = "the GPU's timed window" if n == steps else f"the first {n} steps of the GPU's timed window"
    out = dict(value=plain["value"], unit="zone-updates/s", cores=threads, kind="reference",
               sample=f"reference OpenMP build (oracle/_ref/ng2_lv5_omp: -O3 -funroll-loops -fopenmp, no FMA "
                      f"contraction), rotor lv5 MC+HLLD CFL 0.5, steps {first}..{first + n} of the run from t = 0 "
                      f"({window}): {plain['zone_updates']} zone-updates in {plain['seconds']:.3f} s on {threads} "
                      f"threads; host CPU: {cpu_model()}",
               steps=plain["steps"], nproc=os.cpu_count(), affinity_cores=len(os.sched_getaffinity(0)),
               threads_note="OMP_NUM_THREADS (the host's CPU share for one GPU) when set, else the affinity mask")
    if native is not None:
        out["reference_flags"] = dict(
            value=native["value"], unit="zone-updates/s", zone_updates=native["zone_updates"],
            seconds=native["seconds"],
            build="reference makefile flags (makefile:9-19: -O3 -funroll-loops -fopenmp, FMA contraction on) "
                  "with -march=x86-64-v3 for -march=native (built on another host)")
    return out


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
            b.set_regrid_device(REGRID_DEVICE)
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
        sim.set_regrid_device(REGRID_DEVICE)
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
           "J": float(J.detach()), "forward_ms_per_step": tf / steps * 1e3,
           "backward_ms_per_step": tb / steps * 1e3,
           "backward_over_forward": tb / tf,
           "note": "backward recomputes every checkpoint segment, writing each step's tape, then runs the "
                   "taped reverse sweep",
           "taped_every_step": {"J": float(J0.detach()), "forward_ms_per_step": tf0 / steps * 1e3,
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


def stage_op_count(at=(1000, 2500, 4000)):
    """Algorithmic FP64 operations per zone of one stage (mean of the two RK2
    stages) of the rotor workload, counted on the host by the op-counting
    build of the stage arithmetic on the run's own states at the given steps
    (untimed): the operations of the source, each division and square root
    counted once, every branch as taken."""
    import amr2d_mhd_amd as A

    sim = A.Simulation(ROTOR)
    sim.init()
    done, rows = 0, []
    for s in at:
        sim.run(s - done)
        done = s
        rows.append(sim.stage_flops())
    sim.close()
    mean = {k: sum(r[k] for r in rows) / len(rows) for k in rows[0]}
    return dict(flops_per_zone=mean["total"], add_sub=mean["add"], mul=mean["mul"], div=mean["div"], sqrt=mean["sqrt"],
                states=f"rotor states after steps {', '.join(str(s) for s in at)}")


def rotor_tuning():
    """Stage-kernel tuner decisions of the whole rotor run, from an untimed
    run (the timed runs import them, so no trial launch falls in a timed
    region)."""
    import amr2d_mhd_amd as A

    sim = A.Simulation(ROTOR)
    sim.init()
    sim.run(-1)
    tun = {b: v[0] for b, v in sim.stage_tuning().items() if v[0] >= 0}
    sim.close()
    return tun


STEP_KERNEL = 0  # engine step mode of the rotor runs (--step-kernel)
REGRID_DEVICE = True  # regrid decided on the device (--regrid)
FILL_STATS = False  # report the fill program's op mix at both ends of the window (--fill-stats)
# HIP events on one stage launch in EVENT_STRIDE (odd: stage 1 and stage 2
# launches alternate in the sample); events on every launch cost ~6% of the run
EVENT_STRIDE = 15


def rotor_run(warmup, steps, tuning):
    """`warmup` untimed steps of the rotor from t = 0, then `steps` timed
    steps (-1: to the end of the run) with the stage kernel's launches timed
    by HIP events on the engine's stream."""
    import torch

    import amr2d_mhd_amd as A

    sim = A.Simulation(ROTOR)
    sim.set_stage_tuning(tuning)
    sim.set_step_kernel(STEP_KERNEL)
    sim.set_regrid_device(REGRID_DEVICE)
    sim.init()
    if warmup:
        sim.run(warmup)
    torch.cuda.synchronize()
    nleaf0 = sim.info()["nleaf"]
    fill0 = sim.fill_stats() if FILL_STATS else None
    tuner0 = sim.stage_tuning()
    sim.profile(("stage",), stride=EVENT_STRIDE)
    t0 = time.perf_counter()
    n, zu = sim.run(steps)
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    prof = sim.profile_read()
    info = sim.info()
    fill1 = sim.fill_stats() if FILL_STATS else None
    # a trial launch changes its bucket's per-variant timings
    trial_buckets = sorted(b for b, v in sim.stage_tuning().items() if tuner0.get(b) != v)
    sim.close()
    stage_ms, stage_n = prof["stage"]
    return dict(steps=n, zone_updates=zu, seconds=secs, value=zu / secs, ms_per_step=secs / n * 1e3,
                t_reached=info["t"], leaves_at_start=nleaf0, leaves_at_end=info["nleaf"],
                stage_ms=stage_ms, stage_launches=stage_n, trial_buckets=trial_buckets,
                fill_program=dict(window_start=fill0, window_end=fill1) if FILL_STATS else None)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=ROTOR_STEPS, help="timed rotor steps (-1: to the end)")
    ap.add_argument("--warmup", type=int, default=0, help="untimed rotor steps from t = 0 before the timed ones")
    ap.add_argument("--cpu-max-steps", type=int, default=200,
                    help="the CPU baseline times at most this many steps of the GPU's window")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-whole-run", action="store_true")
    ap.add_argument("--ensemble", type=int, default=256, metavar="N",
                    help="batched ensemble of N seeded rotors against N plain runs (0: skip)")
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
    ap.add_argument("--adjoint", type=int, default=512, metavar="N",
                    help="frozen-mesh adjoint on an N x N grid, HLLE and HLLD (0: skip)")
    ap.add_argument("--adj-steps", type=int, default=400)
    ap.add_argument("--adj-every", type=int, default=20)
    ap.add_argument("--adj-fd", type=int, default=0, help="finite-difference check directions")
    ap.add_argument("--no-op-count", action="store_true",
                    help="skip the host-side operation count of the stage arithmetic")
    ap.add_argument("--step-kernel", type=int, default=0, choices=[0, 1, 2],
                    help="rotor runs with per-phase launches (0) or the persistent step kernel, 256 (1) or 512 (2) "
                         "lanes; with 1 or 2 the stage figures time whole stage phases, barrier included")
    ap.add_argument("--adj-flux", default="both", choices=["HLLE", "HLLD", "both"])
    ap.add_argument("--fill-stats", action="store_true",
                    help="report the ghost-fill program's op mix (copies, restrictions, padding) at both ends of the window")
    ap.add_argument("--regrid", default="device", choices=["device", "host"],
                    help="where the rotor and ensemble runs decide each regrid (identical results)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one 256 MiB device copy (counter calibration for rocprofv3 --pmc runs)")
    args = ap.parse_args()
    global STEP_KERNEL, REGRID_DEVICE, FILL_STATS
    STEP_KERNEL = args.step_kernel
    FILL_STATS = args.fill_stats
    REGRID_DEVICE = args.regrid == "device"
    if args.gpus != 1:
        print(f"bench.py: this build runs on exactly one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        return 2

    import torch

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.only_ensemble:
        print(json.dumps(ensemble_bench(args.ens_steps, args.ens_full, args.ensemble or 256, plain=False,
                                        groups=args.ens_groups, warmup=args.ens_warmup)))
        return 0
    if args.copy_calibration:
        x = torch.empty(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        x.fill_(1.0)
        y.copy_(x)
        torch.cuda.synchronize()

    tuning = rotor_tuning()
    steps = args.steps if args.steps >= 0 else -1
    win = rotor_run(args.warmup, steps, tuning)
    if steps >= 0 and win["steps"] != steps:
        print(f"bench.py: run ended after {win['steps']} of {steps} timed steps", file=sys.stderr)
        return 1
    steps = win["steps"]
    zones_per_launch = win["zone_updates"] / steps  # each step launches the stage kernel twice over its leaves
    mean_launch_s = win["stage_ms"] / win["stage_launches"] * 1e-3
    achieved = STAGE_BYTES_PER_ZONE * zones_per_launch / mean_launch_s / 1e9
    flops = STAGE_FP64_FLOPS_PER_ZONE * zones_per_launch / mean_launch_s / 1e12
    out = {
        "metric": "zone-updates/s FP64 at 1 GPU, batched vs plain; adjoint/forward step cost",
        "value": win["value"],
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": win["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (the reference's rotor initial condition, IC2Dtype.cpp:75-111)",
        "config": {"workload": "MHD rotor (plain method), 4 root blocks of 16x16, 5 AMR levels, MC-MUSCL + HLLD, "
                               "CT, RK2, CFL 0.5, gamma 1.4; value = zone-updates/s over steps "
                               f"{args.warmup}..{args.warmup + steps} of the run from t = 0",
                   "window_steps": [args.warmup, args.warmup + steps], "t_end_reached": win["t_reached"],
                   "leaves_at_window_start": win["leaves_at_start"], "leaves_at_end": win["leaves_at_end"],
                   "zone_updates": win["zone_updates"], "parallelism": "single GPU",
                   "step_execution": ["per-phase launches", "persistent step kernel, 256 lanes",
                                      "persistent step kernel, 512 lanes"][args.step_kernel],
                   "regrid": ("on the device (markref, critneighup/down, updatelpup/down per instance; the host "
                              "reads back the op list)") if REGRID_DEVICE else "on the host (per-leaf indicators downloaded)",
                   **({"fill_program": win["fill_program"]} if win["fill_program"] else {}),
                   "stage_tuner": "decisions imported from an untimed run; leaf-count buckets that ran "
                                  f"trial launches inside the timed window: {win['trial_buckets']}"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": STAGE_TRAFFIC_BYTES_PER_ZONE * zones_per_launch,
                     "kernel": "stage_kernel<20,2>", "mean_launch_us": mean_launch_s * 1e6,
                     "launches": win["stage_launches"], "events_on_one_launch_in": EVENT_STRIDE,
                     "bytes_per_zone": STAGE_BYTES_PER_ZONE,
                     "zones_per_launch": zones_per_launch},
        "valu": {"bound": "fp64", "kernel": "stage_kernel<20,2>", "achieved": flops, "peak": FP64_PEAK_TFLOPS,
                 "unit": "TFLOP/s", "frac": flops / FP64_PEAK_TFLOPS, "flops_per_zone": STAGE_FP64_FLOPS_PER_ZONE},
        "cpu_baseline": None,
    }
    if not args.no_op_count:
        oc = stage_op_count()
        oc_rate = oc["flops_per_zone"] * zones_per_launch / mean_launch_s / 1e12
        out["valu"]["op_counted"] = dict(oc, achieved=oc_rate, unit="TFLOP/s", frac=oc_rate / FP64_PEAK_TFLOPS,
                                         stage_flops_per_zone_update=2 * oc["flops_per_zone"],
                                         note="flops_per_zone: one stage; a zone-update (one step) runs two "
                                              "stages, plus CT and the CFL reduction (not counted)")
    if not args.no_whole_run and not (args.warmup == 0 and args.steps in (-1, ROTOR_STEPS)):
        w = rotor_run(0, -1, tuning)
        out["whole_run"] = {"workload": "the same rotor run, every step from t = 0 to t = 0.295", "steps": w["steps"],
                            "zone_updates": w["zone_updates"], "seconds": w["seconds"], "value": w["value"],
                            "ms_per_step": w["ms_per_step"], "t_reached": w["t_reached"]}
    if args.ensemble:
        out["batched_vs_plain"] = ensemble_bench(args.ens_steps, args.ens_full, args.ensemble,
                                                 warmup=args.ens_warmup, groups=args.ens_groups)
    if args.adjoint:
        fluxes = ["HLLE", "HLLD"] if args.adj_flux == "both" else [args.adj_flux]
        out["adjoint_over_forward"] = {f: adjoint_bench(args.adjoint, args.adj_steps, args.adj_every, args.adj_fd, f)
                                       for f in fluxes}
    if args.ot_steps:
        out["orszag_tang"] = ot_bench(args.ot_steps, args.ot_warmup)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baselines(args.warmup, steps, args.cpu_max_steps)
        if out["cpu_baseline"]:
            cb = out["cpu_baseline"]
            if cb["steps"] == [args.warmup, args.warmup + steps]:
                out["gpu_over_cpu_same_window"] = out["value"] / cb["value"]
            else:
                # the CPU times a bounded part of the GPU's window: time the GPU on that part too
                a, b = cb["steps"]
                g = rotor_run(a, b - a, tuning)
                out["gpu_on_cpu_window"] = {"steps": [a, b], "zone_updates": g["zone_updates"],
                                            "seconds": g["seconds"], "value": g["value"]}
                out["gpu_over_cpu_same_window"] = g["value"] / cb["value"]
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
