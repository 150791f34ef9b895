"""Benchmark: replica-MC pose evaluations/s on one MI355X.

Headline workload (`value`, `config`, `roofline`, `cpu_baseline`): the
largest single-GPU configuration, the stress shape -- a 150-atom ligand in
64 conformations against a 20,000-point protein (35 A sphere) in 32
conformations, KDE 5,000, MCS 100 rows, 32 temperatures: 65,536 replicas,
exchange between neighbouring temperatures every 50 steps.  Synthetic inputs
are written in the SDF/PDB/.ff formats from one seed and read through the
loader.  One "step" is one MC step of every replica (65,536 pose
evaluations).  The timed region runs the batched form (persistent fused
kernel over a cost-ordered work queue, autotuned workgroup shape,
hipGraph-replayed chunks + exchange) with inputs resident in HBM; `--steps`
steps after `--warmup` untimed steps (the autotuner's trial launches run in
the warm-up).

Beside it, each in its own section of the JSON line, each with its own
fraction of the FP32 vector peak:
  plain     the headline workload in the plain form (one workgroup per
            replica, one launch per chunk, record read-back) over the same
            steps;
  c2        the full replica ensemble (48 atoms x 20 conformations against
            4,000 points x 11 conformations, 16 temperatures: 3,520
            replicas), batched and plain over the same steps, with the same
            exchange rounds in both forms;
  screen    batched screening: 1,024 ligands (lognormal 12-96 atoms, 8
            conformations, 8 temperatures; ~393k replicas) against one
            resident 5,000-point protein in 6 conformations, all jobs in one
            persistent launch per chunk; versus plain launches (one job after
            another) on a sample of the ligands;
  serve     the serving loop on an 8,000-point protein in 4 conformations:
            requests of 10-120 atoms, 4-32 conformations, 8 temperatures,
            500-5,000 steps arriving as a Poisson stream at 50, 200 and 800
            per second, admitted under a slot limit that binds; p50/p99 job
            latency, the queueing share of latency (admission - submission)
            and achieved pose evaluations/s per rate, versus the same
            requests run one after another with plain launches;
  cpu_baseline  the CPU oracle (OpenMP restatement of MonteCarlo_d,
            OMP_PROC_BIND=true) on a bounded sample of the headline workload.

The roofline's `achieved` is algorithmic FLOP per launch (FLOP per pose x
replicas x steps per launch) over the average launch duration of the Monte
Carlo kernel in the timed region, measured with hipEvents on the stream the
kernel is launched on; `traffic` is HBM bytes per launch from the committed
rocprofv3 FETCH_SIZE/WRITE_SIZE summary of this same command.
"""
import os

# OpenMP reads its binding when the runtime loads (the CPU oracle's); the
# reference's run instructions bind threads (doc/instructions.txt:140-141)
os.environ.setdefault("OMP_PROC_BIND", "true")

import argparse  # noqa: E402
import json  # noqa: E402
import math  # noqa: E402
import re  # noqa: E402
import sys  # noqa: E402
import time  # noqa: E402

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "replica-MC pose evaluations/s on 1 MI355X vs plain; p50/p99 job latency"
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "mc_kernel_stress_counters.md")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def frac(tflops, perf):
    return None if tflops is None else tflops / perf.FP32_PEAK_TFLOPS


def cpu_baseline(cx, target_seconds, sample_reps):
    """The CPU oracle (the repository's C restatement of MonteCarlo_d, OpenMP
    over replicas) on a bounded sample of the workload: `sample_reps`
    replicas spread evenly over the ensemble for as many steps as fit the
    time budget, timed on the host's cores."""
    from geauxdock_cs_v2_amd import ctypes_defs as D
    from geauxdock_cs_v2_amd import perf
    from oracle import bind as O
    # host facts before the OpenMP runtime binds this thread to one CPU
    host = perf.host_info()
    threads = perf.host_threads()
    reps = cx.replicas()
    n = len(reps)
    stride = max(1, n // sample_reps)
    sub_idx = list(range(0, n, stride))[:sample_reps]
    sub = (D.ReplicaMC * len(sub_idx))(*[reps[i] for i in sub_idx])
    view = cx.view
    t0 = time.perf_counter()
    O.montecarlo(view, sub, 0, 1, 99, nthreads=threads)
    one = time.perf_counter() - t0
    steps = max(1, int(target_seconds / max(one, 1e-3)))
    t0 = time.perf_counter()
    O.montecarlo(view, sub, 1, steps, 99, nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": len(sub_idx) * steps / dt, "unit": "pose evaluations/s", "cores": threads, "kind": "port",
            "sample": f"{len(sub_idx)} of {n} replicas (every {stride}th) x {steps} steps, {dt:.1f} s, "
                      f"{threads} threads (OMP_PROC_BIND={host['omp_proc_bind']}; affinity mask "
                      f"{host['affinity_cpus']} CPUs, capped at {host['thread_cap']}, one GPU's host share) on "
                      f"{host['sockets']} x {host['model']} ({host['cores_per_socket']} cores/socket, "
                      f"{host['threads_per_core']} threads/core)",
            "host": host}


def traffic_from_summary(steps_per_launch):
    """HBM bytes per launch of the headline kernel from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary, with the steps per launch it was
    measured at; scaled to `steps_per_launch` (None if absent)."""
    try:
        with open(TRAFFIC_SUMMARY) as f:
            text = f.read()
    except OSError:
        return None, None
    m = re.search(r"^bytes_per_launch:\s*([0-9.eE+]+)", text, re.M)
    k = re.search(r"^steps_per_launch:\s*([0-9]+)", text, re.M)
    if not m or not k:
        return None, None
    b, s = float(m.group(1)), int(k.group(1))
    return b * steps_per_launch / s, s


def issue_from_summary():
    """VALU issue figures of the headline kernel from the committed rocprofv3
    instruction-counter summary (the same file as the traffic), or None."""
    try:
        with open(TRAFFIC_SUMMARY) as f:
            text = f.read()
    except OSError:
        return None
    out = {}
    for key in ("valu_per_wave_pair", "trans_per_wave_pair", "valu_issue_per_quad_cycle"):
        m = re.search(rf"^{key}:\s*([0-9.eE+]+)", text, re.M)
        if not m:
            return None
        out[key] = float(m.group(1))
    out["trans_share"] = out["trans_per_wave_pair"] / out["valu_per_wave_pair"]
    out["issue_peak_per_quad_cycle"] = 2.0
    out["source"] = os.path.relpath(TRAFFIC_SUMMARY, ROOT)
    return out


def batched_window(dev, jobs, steps, warmup, exchange, shape=None):
    """Warm-up (autotuner trials and `warmup` steps), then `steps` timed
    steps; returns (wall s, device ms, launches, steps per launch)."""
    shape = shape or {}
    if warmup > 0:
        dev.batch_run(jobs, warmup, exchange=exchange if warmup >= exchange > 0 else 0, graph=True, **shape)
    t0 = time.perf_counter()
    dev_ms = dev.batch_run(jobs, steps, exchange=exchange, graph=True, **shape)
    wall = time.perf_counter() - t0
    chunk = exchange if exchange > 0 else min(steps, 100)
    launches = math.ceil(steps / chunk)
    return wall, dev_ms, launches, steps / launches


def plain_window(dev, t, cx, steps, exchange, seed):
    """The plain form over `steps` steps: one launch per chunk (the exchange
    interval, else 100 steps), records read back after each, an exchange
    round after each full interval -- the rounds the batched form runs."""
    n_rep = cx.size["n_rep"]
    chunk = exchange if exchange > 0 else 100
    jp = dev.upload_job(t, cx, seed=seed)
    dev.montecarlo(jp, 0, 1)  # first-launch costs out of the window
    rnd = 0
    t0 = time.perf_counter()
    s1 = 1
    while s1 <= steps:
        s2 = min(chunk - (s1 - 1) % chunk, steps + 1 - s1)
        dev.montecarlo(jp, s1, s2)
        dev.records(jp, n_rep, max(s2, 100))
        s1 += s2
        if exchange > 0 and (s1 - 1) % exchange == 0:
            dev.exchange(jp, rnd)
            rnd += 1
    wall = time.perf_counter() - t0
    dev.free_job(jp)
    return n_rep * steps / wall


def describe(cx, name, exch):
    z = cx.size
    return (f"{name}: ligand {z['lig_natom']} atoms x {z['n_lig']} conf, protein {z['prt_npoint']} points x "
            f"{z['n_prt']} conf, {z['n_tmp']} temperatures, {z['n_rep']} replicas, exchange every {exch}")


def run_ensemble(args, dev, perf, synth, name, steps, warmup, seed):
    """One single-job configuration, batched then plain over the same steps."""
    cx, info = synth.load_config(name)
    n_rep = cx.size["n_rep"]
    exch = info["cfg"]["exchange"]
    f_pose = perf.flops_per_pose(cx)
    log(f"{name}: {cx.size}, {f_pose} FLOP/pose")
    t = dev.upload_target(cx)
    # the whole window once beforehand on a twin of the job (same tables and
    # seed), so first-use costs of the timed configuration (launch-table
    # growth, the clock ramp after host work) fall outside the timed window;
    # the timed steps stay the ones right after the warm-up
    tw = dev.upload_job(t, cx, seed=seed)
    batched_window(dev, [tw], steps, warmup, exch)
    dev.free_job(tw)
    j = dev.upload_job(t, cx, seed=seed)
    wall, dev_ms, launches, spl = batched_window(dev, [j], steps, warmup, exch)
    shape = dev.job_shape(j)
    plan = dev.batch_plan()  # of the window's last launch
    dev.free_job(j)
    value = n_rep * steps / wall
    launch_ms = dev_ms / launches
    tflops = f_pose * n_rep * spl / (launch_ms * 1e-3) / 1e12
    log(f"{name} batched: {steps} steps in {wall:.3f} s wall ({dev_ms:.1f} ms device, {launches} launches of "
        f"{launch_ms:.2f} ms), {value:.4g} poses/s, {tflops:.1f} TFLOP/s; {plan[0]} tasks in {plan[2]} step "
        f"slices over {plan[1]} workgroups")
    plain = None
    if args.plain:
        plain = plain_window(dev, t, cx, steps, exch, seed)
        log(f"{name} plain: {steps} steps, {plain:.4g} poses/s")
    dev.free_target(t)
    return cx, {"workload": describe(cx, name, exch), "replicas": n_rep, "steps": steps, "value": value,
                "wall_s": wall, "device_ms": dev_ms, "launches": launches, "steps_per_launch": spl,
                "launch_ms": launch_ms, "flop_per_pose": f_pose, "tflops": tflops, "frac": frac(tflops, perf),
                "shape": {"waves": shape[0], "slot_waves": shape[1], "point_slices": shape[2],
                          "replicas_per_group": shape[3], "chosen_by": "autotuner"},
                "plan": {"tasks": plan[0], "workgroups": plan[1], "step_slices": plan[2]},
                "plain": {"value": plain, "steps": steps,
                          "form": "one workgroup per replica, one launch per chunk, record read-back per chunk"},
                "speedup_vs_plain": value / plain if plain else None}


def run_screen(args, dev, perf, synth):
    info = synth.many_files("screen")
    t0 = time.perf_counter()
    cxs = synth.load_many(info, threads=perf.host_threads())
    log(f"screen: {len(cxs)} ligands loaded and staged in {time.perf_counter() - t0:.1f} s")
    n_rep = sum(c.size["n_rep"] for c in cxs)
    t = dev.upload_target(cxs[0])
    jobs = [dev.upload_job(t, c, seed=100 + i) for i, c in enumerate(cxs)]
    dev.batch_run(jobs, 10, graph=True)  # warm-up (and the shape autotuner's trials)
    t0 = time.perf_counter()
    dev_ms = dev.batch_run(jobs, args.screen_steps, graph=True)
    wall = time.perf_counter() - t0
    value = n_rep * args.screen_steps / wall
    flop = sum(perf.flops_per_pose(c) * c.size["n_rep"] for c in cxs) * args.screen_steps
    tflops = flop / (dev_ms * 1e-3) / 1e12
    waves = dev.job_shape(jobs[0])[0]
    log(f"screen batched: {args.screen_steps} steps x {n_rep} replicas in {wall:.3f} s "
        f"({dev_ms:.1f} ms device), {value:.4g} poses/s, {tflops:.1f} TFLOP/s, {waves}-wave workgroups")
    for j in jobs:
        dev.free_job(j)
    # plain launches, one ligand after another, on every 64th ligand
    sample = list(range(0, len(cxs), 64))
    plain_poses = 0
    plain_jobs = [dev.upload_job(t, cxs[i], seed=100 + i) for i in sample]
    dev.montecarlo(plain_jobs[0], 0, 1)
    t0 = time.perf_counter()
    for i, j in zip(sample, plain_jobs):
        dev.montecarlo(j, 1, args.screen_plain_steps)
        dev.records(j, cxs[i].size["n_rep"], 100)
        plain_poses += cxs[i].size["n_rep"] * args.screen_plain_steps
    plain_wall = time.perf_counter() - t0
    plain_value = plain_poses / plain_wall
    for j in plain_jobs:
        dev.free_job(j)
    dev.free_target(t)
    log(f"screen plain: {len(sample)} ligands x {args.screen_plain_steps} steps in {plain_wall:.3f} s, "
        f"{plain_value:.4g} poses/s")
    atoms = [l["n_atoms"] for l in info["ligands"]]
    return {"workload": f"screen: {len(cxs)} ligands of {min(atoms)}-{max(atoms)} atoms (median "
                        f"{sorted(atoms)[len(atoms) // 2]}) x 8 conf x 8 temperatures against a "
                        f"{cxs[0].size['prt_npoint']}-point protein x {cxs[0].size['n_prt']} conf, "
                        f"{n_rep} replicas",
            "value": value, "steps": args.screen_steps, "tflops": tflops, "frac": frac(tflops, perf),
            "waves_per_group": waves,
            "plain": {"value": plain_value, "sample": f"ligands 0, 64, ... ({len(sample)}) x "
                                                      f"{args.screen_plain_steps} steps, one after another"},
            "speedup_vs_plain": value / plain_value}


def run_serve(args, dev, perf, synth, Server):
    import numpy as np
    info = synth.many_files("serve", n=args.serve_requests)
    cxs = synth.load_many(info, threads=perf.host_threads())
    steps = [l["steps"] for l in info["ligands"]]
    poses = [c.size["n_rep"] * s for c, s in zip(cxs, steps)]
    flops = [perf.flops_per_pose(c) * p for c, p in zip(cxs, poses)]
    t = dev.upload_target(cxs[0])
    # serial plain launches of the same requests, every step of each: one
    # 100-step launch per chunk, record read-back, exchange every 100 steps
    # (the server's own cadence)
    warm = dev.upload_job(t, cxs[0], seed=499)
    dev.montecarlo(warm, 0, 1)
    dev.free_job(warm)
    plain_dur = []
    for i, c in enumerate(cxs):
        j = dev.upload_job(t, c, seed=500 + i)
        t0 = time.perf_counter()
        for k, s1 in enumerate(range(0, steps[i], 100)):
            s2 = min(100, steps[i] - s1)
            dev.montecarlo(j, s1, s2)
            dev.records(j, c.size["n_rep"], 100)
            if s2 == 100:
                dev.exchange(j, k)
        plain_dur.append(time.perf_counter() - t0)
        dev.free_job(j)
        if (i + 1) % 25 == 0:
            log(f"serve plain: {i + 1} of {len(cxs)} requests, {sum(plain_dur):.1f} s")
    plain_value = sum(poses) / sum(plain_dur)
    log(f"serve plain: {len(cxs)} requests run one after another, {sum(plain_dur):.2f} s, {plain_value:.4g} poses/s")
    rates = []
    rng = np.random.default_rng(7)
    gaps = rng.exponential(1.0, len(cxs))  # Poisson arrivals of unit rate, scaled per load
    capacity = None                           # requests/s the server sustains
    # the workload's arrival rates, highest first: the highest is far above
    # what one GPU serves, so its span gives the capacity; then loads below it
    runs = [(lam, None) for lam in sorted(synth.SERVE["rates"], reverse=True)]
    runs += [(None, load) for load in synth.SERVE["loads"]]
    for lam, load in runs:
        if lam is None:
            lam = load * capacity
        arrivals = np.cumsum(gaps) / lam
        srv = Server(dev, t, hbm_budget=args.serve_hbm_gib << 30, max_active=args.serve_max_active, chunk=100,
                     exchange=100)
        ids = []
        t0 = time.perf_counter()
        for i, (c, a) in enumerate(zip(cxs, arrivals)):
            dt = a - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            ids.append(srv.submit(c, steps[i], 900 + i))
            if (i + 1) % 50 == 0:
                log(f"serve: {i + 1} of {len(cxs)} requests submitted, {time.perf_counter() - t0:.1f} s")
        srv.drain()
        res = [srv.poll(i)[0] for i in ids]
        srv.close()
        if any(r.status != 0 for r in res):
            raise RuntimeError(f"serving at {lam:g}/s: failed requests {[r.status for r in res]}")
        lat = np.array([r.t_done_ms - r.t_submit_ms for r in res])
        queue = np.array([r.t_admit_ms - r.t_submit_ms for r in res])
        span = max(r.t_done_ms for r in res) - min(r.t_submit_ms for r in res)
        achieved = sum(poses) / (span * 1e-3)
        tflops = sum(flops) / (span * 1e-3) / 1e12
        if capacity is None:
            capacity = len(cxs) / (span * 1e-3)
        # the same arrivals served one after another by plain launches
        # (measured durations, first come first served)
        free_at, plain_lat = 0.0, []
        for a, dur in zip(arrivals, plain_dur):
            start = max(a, free_at)
            free_at = start + dur
            plain_lat.append((free_at - a) * 1e3)
        rates.append({"load": load, "rate_per_s": lam, "requests": len(ids),
                      "max_active": args.serve_max_active,
                      "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                      "queue_p50_ms": float(np.percentile(queue, 50)), "queue_p99_ms": float(np.percentile(queue, 99)),
                      "queue_share": float(queue.sum() / lat.sum()),
                      "value": achieved, "tflops": tflops, "frac": frac(tflops, perf),
                      "speedup_vs_plain": achieved / plain_value,
                      "plain_p50_ms": float(np.percentile(plain_lat, 50)),
                      "plain_p99_ms": float(np.percentile(plain_lat, 99))})
        name = f"{lam:.3g}/s" + ("" if load is None else f" (load {load:g})")
        log(f"serve {name}: p50 {rates[-1]['p50_ms']:.0f} ms, p99 {rates[-1]['p99_ms']:.0f} ms, queueing "
            f"{100 * rates[-1]['queue_share']:.0f}% of latency, {achieved:.4g} poses/s ({tflops:.1f} TFLOP/s); "
            f"serial plain p50 {rates[-1]['plain_p50_ms']:.0f} ms, p99 {rates[-1]['plain_p99_ms']:.0f} ms")
    dev.free_target(t)
    return {"workload": f"serve: {len(cxs)} requests per rate, {min(l['n_atoms'] for l in info['ligands'])}-"
                        f"{max(l['n_atoms'] for l in info['ligands'])} atoms, 4-32 conf, 8 temperatures, "
                        f"{min(steps)}-{max(steps)} steps, against an {cxs[0].size['prt_npoint']}-point protein "
                        f"x {cxs[0].size['n_prt']} conf, at most {args.serve_max_active} requests admitted at once; "
                        f"Poisson arrivals at {', '.join(f'{r:g}' for r in synth.SERVE['rates'])}/s, then at "
                        f"{', '.join(f'{l:g}' for l in synth.SERVE['loads'])} of the capacity measured at the "
                        f"highest rate",
            "capacity_per_s": capacity, "rates": rates, "plain": {"value": plain_value, "sample": "every step of each request, one request "
                                                                       "after another, plain launches"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed MC steps of the headline workload")
    ap.add_argument("--warmup", type=int, default=5, help="untimed MC steps before the timed ones")
    ap.add_argument("--sections", default="stress,c2,screen,serve",
                    help="comma list of stress (the headline), c2, screen, serve")
    ap.add_argument("--c2-steps", type=int, default=-1, help="steps of the c2 section (default: --steps)")
    ap.add_argument("--no-plain", dest="plain", action="store_false", help="skip the plain-form comparisons")
    ap.add_argument("--screen-steps", type=int, default=100)
    ap.add_argument("--screen-plain-steps", type=int, default=100)
    ap.add_argument("--serve-requests", type=int, default=24)
    ap.add_argument("--serve-max-active", type=int, default=4,
                    help="requests admitted at once (binds: the rest queue until a slot frees)")
    ap.add_argument("--serve-hbm-gib", type=int, default=64)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-replicas", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB five times (scales rocprof byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copied {args.copy_calibration} MiB x 5")
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported on this build "
              "(replica ladders shard by job across GPUs with no collective in the MC loop)", file=sys.stderr)
        sys.exit(2)
    sections = set(s.strip() for s in args.sections.split(",") if s.strip())
    if args.c2_steps < 0:
        args.c2_steps = args.steps

    from geauxdock_cs_v2_amd import perf, synth
    from geauxdock_cs_v2_amd.device import Device, Server

    dev = Device(0)
    cx5, stress = (run_ensemble(args, dev, perf, synth, "c5", args.steps, args.warmup, 5)
                   if "stress" in sections else (None, None))
    c2 = run_ensemble(args, dev, perf, synth, "c2", args.c2_steps, args.warmup, 2024)[1] if "c2" in sections else None
    screen = run_screen(args, dev, perf, synth) if "screen" in sections else None
    serve = run_serve(args, dev, perf, synth, Server) if "serve" in sections else None
    dev.close()

    cpu = None
    if not args.no_cpu and cx5 is not None:
        cpu = cpu_baseline(cx5, args.cpu_seconds, args.cpu_replicas)
        log(f"cpu oracle: {cpu['value']:.4g} poses/s ({cpu['sample']})")

    roof = None
    if stress:
        traffic, traffic_steps = traffic_from_summary(stress["steps_per_launch"])
        # pairwise vector work with no matrix-core form: bound by the FP32
        # vector ALUs, whose peak is MI355X's published FP32 figure
        roof = {"bound": "valu", "achieved": stress["tflops"], "peak": perf.FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": stress["frac"], "traffic": traffic, "kernel": "mc_kernel",
                "launch_ms": stress["launch_ms"], "launches": stress["launches"],
                "steps_per_launch": stress["steps_per_launch"],
                "unit_of_work": f"one launch of the Monte Carlo kernel: {stress['steps_per_launch']:g} steps of all "
                                f"{stress['replicas']} replicas",
                "flop_per_pose": stress["flop_per_pose"], "shape": stress["shape"], "plan": stress["plan"]}
        if traffic:
            roof["hbm_gbs"] = traffic / (stress["launch_ms"] * 1e-3) / 1e9
            roof["hbm_frac"] = roof["hbm_gbs"] / perf.HBM_PEAK_GBS
            roof["traffic_source"] = os.path.relpath(TRAFFIC_SUMMARY, ROOT)
            roof["traffic_measured_steps_per_launch"] = traffic_steps
        issue = issue_from_summary()
        if issue:
            roof["issue"] = issue
    head = stress or {"value": None, "wall_s": float("nan"), "workload": None, "replicas": None,
                      "plain": {"value": None}, "speedup_vs_plain": None}
    out = {
        "metric": METRIC, "value": head["value"], "unit": "pose evaluations/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": head["wall_s"] * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": head["workload"], "replicas": head["replicas"],
                   "form": "batched (persistent fused kernel over a cost-ordered work queue, autotuned shape, "
                           "hipGraph chunks)"},
        "plain": head["plain"],
        "speedup_vs_plain": head["speedup_vs_plain"],
        "roofline": roof,
        "cpu_baseline": cpu,
        "c2": c2,
        "screen": screen,
        "serve": serve,
    }
    if serve:
        key = lambda r: f"{r['rate_per_s']:.3g}/s" + ("" if r["load"] is None else f" (load {r['load']:g})")
        out["p50_ms"] = {key(r): r["p50_ms"] for r in serve["rates"]}
        out["p99_ms"] = {key(r): r["p99_ms"] for r in serve["rates"]}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
