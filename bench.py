"""Benchmark: replica-MC pose evaluations/s on one MI355X.

Headline workload (`value`): the full replica ensemble -- a 48-atom ligand in
20 conformations against a 4,000-point protein in 11 conformations, KDE
1,000, MCS 40 x 24, 16 temperatures (0.3-2.0): 3,520 replicas, exchange
between neighbouring temperatures every 100 steps.  Synthetic inputs are
written in the SDF/PDB/.ff formats from one seed and read through the loader.
One "step" is one MC step of every replica (3,520 pose evaluations).  The
timed region runs the batched form (replicas packed per workgroup, autotuned
shape, hipGraph-replayed 100-step chunks + exchange), inputs resident in HBM.

Beside it, each in its own section of the JSON line:
  plain     the plain form (one workgroup per replica, one launch per
            100-step chunk, record read-back per chunk) over the same steps;
  roofline  the Monte Carlo kernel's algorithmic FLOP rate: FLOP per pose x
            poses per 100-step chunk / the average chunk time of the timed
            region (hipEvents around the chunk sequence; a chunk is the
            kernel's dispatches for the two workgroup groups, which run side
            by side on two streams); traffic per chunk from the committed
            rocprofv3 counter summary;
  screen    batched screening: 1,024 ligands (lognormal 12-96 atoms, 8
            conformations, 8 temperatures; ~393k replicas) against one
            resident 5,000-point protein in 6 conformations, all jobs in one
            chunk of fused launches, on a window of steps; versus plain launches
            (one job after another, one workgroup per replica) on a sample of
            the ligands;
  serve     the serving loop on an 8,000-point protein in 4 conformations:
            requests of 10-120 atoms, 4-32 conformations, 8 temperatures,
            500-5,000 steps arriving as a Poisson stream at 50, 200 and 800
            per second; p50/p99 job latency and achieved pose evaluations/s
            per rate, versus the same requests run one after another with
            plain launches (their throughput, and the latencies the same
            arrivals would see);
  stress    the stress shape at full size (65,536 replicas, 150 atoms,
            20,000-point protein in 32 conformations) on a 20-step window,
            batched and plain;
  cpu_baseline  the CPU oracle (OpenMP restatement of MonteCarlo_d) on a
            bounded sample of the headline workload.
"""
import argparse
import json
import math
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "replica-MC pose evaluations/s on 1 MI355X vs plain; p50/p99 job latency"
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "mc_kernel_c2_traffic.md")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(cx, target_seconds):
    """The CPU oracle (the repository's C restatement of MonteCarlo_d, OpenMP
    over replicas) on a bounded sample: every 10th replica for as many steps
    as fit the time budget."""
    from geauxdock_cs_v2_amd import ctypes_defs as D
    from geauxdock_cs_v2_amd import perf
    from oracle import bind as O
    threads = perf.host_threads()
    reps = cx.replicas()
    n = len(reps)
    sub_idx = list(range(0, n, 10))
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
            "sample": f"{len(sub_idx)} of {n} replicas (every 10th) x {steps} steps, {dt:.1f} s, "
                      f"{threads} threads on {perf.cpu_model()}"}


def traffic_from_summary():
    """HBM bytes per 100-step chunk of the headline kernel, as the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary states it (None if absent)."""
    try:
        with open(TRAFFIC_SUMMARY) as f:
            m = re.search(r"^bytes_per_chunk:\s*([0-9.eE+]+)", f.read(), re.M)
        return float(m.group(1)) if m else None
    except OSError:
        return None


def run_c2(args, dev, perf, synth):
    cx, info = synth.load_config("c2")
    cfg = info["cfg"]
    n_rep = cx.size["n_rep"]
    exch = cfg["exchange"]
    chunk = exch if exch > 0 else 100
    f_pose = perf.flops_per_pose(cx)
    log(f"c2: {cx.size}, {f_pose} FLOP/pose")
    shape = {}
    if args.shape:
        w, s = (int(x) for x in args.shape.split(","))
        shape = {"waves": w, "point_slices": s}
    t = dev.upload_target(cx)
    j = dev.upload_job(t, cx, seed=2024)
    # warm-up: the autotuner's trial launches (unless a shape is given) and
    # one chunk of the real sequence
    dev.batch_run([j], args.warmup, exchange=exch if args.warmup >= exch > 0 else 0, graph=True, **shape)
    t0 = time.perf_counter()
    dev_ms = dev.batch_run([j], args.steps, exchange=exch, graph=True, **shape)
    wall = time.perf_counter() - t0
    shape_used = dev.job_shape(j)
    value = n_rep * args.steps / wall
    chunks = math.ceil(args.steps / chunk)
    # a chunk: the Monte Carlo kernel over every replica (one dispatch per
    # workgroup group, side by side), then exchange and step advance (~10 us)
    chunk_ms = dev_ms / chunks
    steps_per_chunk = args.steps / chunks
    achieved = f_pose * n_rep * steps_per_chunk / (chunk_ms * 1e-3) / 1e12
    log(f"c2 batched: {args.steps} steps in {wall:.3f} s wall ({dev_ms:.1f} ms device, {chunks} chunks "
        f"of {chunk_ms:.2f} ms), {value:.4g} poses/s, {achieved:.1f} TFLOP/s")
    dev.free_job(j)

    plain_value = None
    if args.plain_steps > 0:
        jp = dev.upload_job(t, cx, seed=2024)
        dev.montecarlo(jp, 0, 1)  # first-launch costs out of the window
        t0 = time.perf_counter()
        for k, s1 in enumerate(range(1, args.plain_steps + 1, 100)):
            s2 = min(100, args.plain_steps + 1 - s1)
            dev.montecarlo(jp, s1, s2)
            dev.records(jp, n_rep, 100)
            if exch:
                dev.exchange(jp, k)
        plain_wall = time.perf_counter() - t0
        plain_value = n_rep * args.plain_steps / plain_wall
        dev.free_job(jp)
        log(f"c2 plain: {args.plain_steps} steps in {plain_wall:.3f} s, {plain_value:.4g} poses/s")
    dev.free_target(t)

    traffic = traffic_from_summary()
    # the path is pairwise vector work with no matrix-core form: it is bound by
    # the FP32 vector ALUs, whose peak equals MI355X's published FP32 figure
    roof = {"bound": "valu", "achieved": achieved, "peak": perf.FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / perf.FP32_PEAK_TFLOPS, "traffic": traffic,
            "kernel": "mc_kernel", "chunk_ms": chunk_ms, "steps_per_chunk": steps_per_chunk,
            "unit_of_work": "one 100-step chunk of all replicas: the kernel's dispatches for the two "
                            "workgroup groups, run side by side on two streams",
            "flop_per_pose": f_pose,
            "shape": {"waves": shape_used[0], "slot_waves": shape_used[1], "point_slices": shape_used[2],
                      "replicas_per_group": shape_used[3], "chosen_by": "given" if args.shape else "autotuner"}}
    if traffic:
        roof["hbm_gbs"] = traffic / (chunk_ms * 1e-3) / 1e9
        roof["hbm_frac"] = roof["hbm_gbs"] / perf.HBM_PEAK_GBS
        roof["traffic_source"] = os.path.relpath(TRAFFIC_SUMMARY, ROOT)
    workload = (f"c2: ligand {cx.size['lig_natom']} atoms x {cx.size['n_lig']} conf, protein "
                f"{cx.size['prt_npoint']} points x {cx.size['n_prt']} conf, {cx.size['n_tmp']} temperatures, "
                f"{n_rep} replicas, exchange every {exch}")
    return cx, {"value": value, "wall": wall, "workload": workload, "n_rep": n_rep, "roofline": roof,
                "plain": {"value": plain_value, "steps": args.plain_steps,
                          "form": "one workgroup per replica, 100-step launches, record read-back per chunk"}}


def run_screen(args, dev, perf, synth):
    info = synth.many_files("screen")
    t0 = time.perf_counter()
    cxs = synth.load_many(info, threads=perf.host_threads())
    log(f"screen: {len(cxs)} ligands loaded and staged in {time.perf_counter() - t0:.1f} s")
    n_rep = sum(c.size["n_rep"] for c in cxs)
    t = dev.upload_target(cxs[0])
    jobs = [dev.upload_job(t, c, seed=100 + i) for i, c in enumerate(cxs)]
    dev.batch_run(jobs, 10, graph=True)  # warm-up
    t0 = time.perf_counter()
    dev_ms = dev.batch_run(jobs, args.screen_steps, graph=True)
    wall = time.perf_counter() - t0
    value = n_rep * args.screen_steps / wall
    flop = sum(perf.flops_per_pose(c) * c.size["n_rep"] for c in cxs) * args.screen_steps
    log(f"screen batched: {args.screen_steps} steps x {n_rep} replicas in {wall:.3f} s "
        f"({dev_ms:.1f} ms device), {value:.4g} poses/s, {flop / (dev_ms * 1e-3) / 1e12:.1f} TFLOP/s")
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
            "value": value, "steps": args.screen_steps, "tflops": flop / (dev_ms * 1e-3) / 1e12,
            "plain": {"value": plain_value, "sample": f"ligands 0, 64, ... ({len(sample)}) x "
                                                      f"{args.screen_plain_steps} steps, one after another"},
            "speedup_vs_plain": value / plain_value}


def run_serve(args, dev, perf, synth, Server):
    import numpy as np
    info = synth.many_files("serve", n=args.serve_requests)
    cxs = synth.load_many(info, threads=perf.host_threads())
    steps = [l["steps"] for l in info["ligands"]]
    poses = [c.size["n_rep"] * s for c, s in zip(cxs, steps)]
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
    plain_value = sum(poses) / sum(plain_dur)
    log(f"serve plain: {len(cxs)} requests run one after another, {sum(plain_dur):.2f} s, {plain_value:.4g} poses/s")
    rates = []
    rng = np.random.default_rng(7)
    for lam in synth.SERVE["rates"]:
        arrivals = np.cumsum(rng.exponential(1.0 / lam, len(cxs)))
        srv = Server(dev, t, hbm_budget=args.serve_hbm_gib << 30, max_active=64, chunk=100, exchange=100)
        ids = []
        t0 = time.perf_counter()
        for i, (c, a) in enumerate(zip(cxs, arrivals)):
            dt = a - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            ids.append(srv.submit(c, steps[i], 900 + i))
        srv.drain()
        res = [srv.poll(i)[0] for i in ids]
        srv.close()
        if any(r.status != 0 for r in res):
            raise RuntimeError(f"serving at {lam}/s: failed requests {[r.status for r in res]}")
        lat = np.array([r.t_done_ms - r.t_submit_ms for r in res])
        span = max(r.t_done_ms for r in res) - min(r.t_submit_ms for r in res)
        achieved = sum(poses) / (span * 1e-3)
        # the same arrivals served one after another by plain launches
        # (measured durations, first come first served)
        free_at, plain_lat = 0.0, []
        for a, dur in zip(arrivals, plain_dur):
            start = max(a, free_at)
            free_at = start + dur
            plain_lat.append((free_at - a) * 1e3)
        rates.append({"rate_per_s": lam, "requests": len(ids), "p50_ms": float(np.percentile(lat, 50)),
                      "p99_ms": float(np.percentile(lat, 99)), "value": achieved,
                      "offered": lam * float(np.mean(poses)), "speedup_vs_plain": achieved / plain_value,
                      "plain_p50_ms": float(np.percentile(plain_lat, 50)),
                      "plain_p99_ms": float(np.percentile(plain_lat, 99))})
        log(f"serve {lam:g}/s: p50 {rates[-1]['p50_ms']:.0f} ms, p99 {rates[-1]['p99_ms']:.0f} ms, "
            f"{achieved:.4g} poses/s (offered {rates[-1]['offered']:.4g}); serial plain p50 "
            f"{rates[-1]['plain_p50_ms']:.0f} ms, p99 {rates[-1]['plain_p99_ms']:.0f} ms")
    dev.free_target(t)
    return {"workload": f"serve: {len(cxs)} requests per rate, {min(l['n_atoms'] for l in info['ligands'])}-"
                        f"{max(l['n_atoms'] for l in info['ligands'])} atoms, 4-32 conf, 8 temperatures, "
                        f"{min(steps)}-{max(steps)} steps, against an {cxs[0].size['prt_npoint']}-point protein "
                        f"x {cxs[0].size['n_prt']} conf",
            "rates": rates, "plain": {"value": plain_value, "sample": "every step of each request, one request "
                                                                       "after another, plain launches"}}


def run_stress(args, dev, perf, synth):
    """The stress shape at full size: 65,536 replicas of a 150-atom ligand
    against a 20,000-point protein in 32 conformations, on a window of steps;
    batched against plain launches over the same steps."""
    cx, info = synth.load_config("c5")
    n_rep = cx.size["n_rep"]
    f_pose = perf.flops_per_pose(cx)
    steps = args.stress_steps
    t = dev.upload_target(cx)
    j = dev.upload_job(t, cx, seed=5)
    dev.batch_run([j], 1, graph=False)  # warm-up: the autotuner's trials (one step each at this size)
    t0 = time.perf_counter()
    dev_ms = dev.batch_run([j], steps, exchange=info["cfg"]["exchange"], graph=True)
    wall = time.perf_counter() - t0
    value = n_rep * steps / wall
    shape = dev.job_shape(j)
    dev.free_job(j)
    jp = dev.upload_job(t, cx, seed=5)
    dev.montecarlo(jp, 0, 1)
    t0 = time.perf_counter()
    dev.montecarlo(jp, 1, steps)
    dev.records(jp, n_rep, max(steps, 100))
    plain_wall = time.perf_counter() - t0
    dev.free_job(jp)
    dev.free_target(t)
    plain_value = n_rep * steps / plain_wall
    tflops = f_pose * n_rep * steps / (dev_ms * 1e-3) / 1e12
    log(f"stress batched: {steps} steps x {n_rep} replicas in {wall:.3f} s, {value:.4g} poses/s, {tflops:.1f} TFLOP/s; "
        f"plain {plain_value:.4g} poses/s")
    return {"workload": f"stress: ligand {cx.size['lig_natom']} atoms x {cx.size['n_lig']} conf, protein "
                        f"{cx.size['prt_npoint']} points x {cx.size['n_prt']} conf, {cx.size['n_tmp']} temperatures, "
                        f"{n_rep} replicas", "steps": steps, "value": value, "tflops": tflops,
            "shape": {"waves": shape[0], "slot_waves": shape[1], "point_slices": shape[2],
                      "replicas_per_group": shape[3]},
            "plain": {"value": plain_value, "steps": steps}, "speedup_vs_plain": value / plain_value}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--plain-steps", type=int, default=-1,
                    help="steps of the plain-form comparison run (default: the same steps as the batched run)")
    ap.add_argument("--sections", default="c2,screen,serve,stress",
                    help="comma list of c2, screen, serve, stress (c2 gives value and roofline)")
    ap.add_argument("--shape", default="", help="W,S: fixed workgroup shape for c2 (skips the autotuner)")
    ap.add_argument("--screen-steps", type=int, default=100)
    ap.add_argument("--screen-plain-steps", type=int, default=100)
    ap.add_argument("--serve-requests", type=int, default=12)
    ap.add_argument("--serve-hbm-gib", type=int, default=64)
    ap.add_argument("--stress-steps", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
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
    # without c2 (profiling runs of one section) the line has no headline value

    if args.plain_steps < 0:
        args.plain_steps = args.steps
    from geauxdock_cs_v2_amd import perf, synth
    from geauxdock_cs_v2_amd.device import Device, Server

    dev = Device(0)
    cx, c2 = run_c2(args, dev, perf, synth) if "c2" in sections else (None, None)
    screen = run_screen(args, dev, perf, synth) if "screen" in sections else None
    serve = run_serve(args, dev, perf, synth, Server) if "serve" in sections else None
    stress = run_stress(args, dev, perf, synth) if "stress" in sections else None
    dev.close()

    cpu = None if (args.no_cpu or cx is None) else cpu_baseline(cx, args.cpu_seconds)
    if cpu:
        log(f"cpu oracle: {cpu['value']:.4g} poses/s ({cpu['sample']})")

    c2 = c2 or {"value": None, "wall": float("nan"), "workload": None, "n_rep": None, "roofline": None,
                "plain": {"value": None}}
    value = c2["value"]
    plain = c2["plain"]
    out = {
        "metric": METRIC, "value": value, "unit": "pose evaluations/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": c2["wall"] * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": c2["workload"], "replicas": c2["n_rep"],
                   "form": "batched (packed replicas, autotuned shape, two workgroup groups on two streams, hipGraph chunks)"},
        "plain": plain,
        "speedup_vs_plain": value / plain["value"] if (value and plain["value"]) else None,
        "roofline": c2["roofline"],
        "screen": screen,
        "serve": serve,
        "stress": stress,
        "cpu_baseline": cpu,
    }
    if serve:
        out["p50_ms"] = {str(r["rate_per_s"]): r["p50_ms"] for r in serve["rates"]}
        out["p99_ms"] = {str(r["rate_per_s"]): r["p99_ms"] for r in serve["rates"]}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
