"""Benchmark: replica-MC pose evaluations/s on one MI355X.

Workload (default): the full replica ensemble -- a 48-atom ligand in 20
conformations against a 4,000-point protein in 11 conformations, KDE 1,000,
MCS 40 x 24, 16 temperatures (0.3-2.0): 3,520 replicas, exchange between
neighbouring temperatures every 100 steps, synthetic inputs written in the
SDF/PDB/.ff formats from one seed and read through the loader.

One "step" is one MC step of every replica (3,520 pose evaluations).  The
timed region runs the batched form (replicas packed per workgroup, autotuned
shape, hipGraph-replayed chunks + exchange).  Beside it: the plain form (one
workgroup per replica, one launch per 100-step chunk, record read-back after
each chunk) on a window of the same run, the roofline of the Monte Carlo
kernel, and the CPU oracle timed on the host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "replica-MC pose evaluations/s on 1 MI355X vs plain; p50/p99 job latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(cx, target_seconds):
    """The CPU oracle (the repository's C restatement of MonteCarlo_d, OpenMP
    over replicas) on a bounded sample: every 10th replica for as many steps
    as fit the time budget."""
    import ctypes as C
    import numpy as np
    from geauxdock_cs_v2_amd import ctypes_defs as D
    from geauxdock_cs_v2_amd import perf
    from oracle import bind as O
    threads = perf.host_threads()
    reps = cx.replicas()
    n = len(reps)
    sub_idx = list(range(0, n, 10))
    sub = (D.ReplicaMC * len(sub_idx))(*[reps[i] for i in sub_idx])
    view = cx.view
    # the oracle addresses replicas by index into the array it is given
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--plain-steps", type=int, default=300)
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

    from geauxdock_cs_v2_amd import perf, synth
    from geauxdock_cs_v2_amd.device import Device

    cx, info = synth.load_config(args.workload)
    cfg = info["cfg"]
    n_rep = cx.size["n_rep"]
    exch = cfg["exchange"]
    f_pose = perf.flops_per_pose(cx)
    log(f"workload {args.workload}: {cx.size}, {f_pose} FLOP/pose")

    dev = Device(0)
    t = dev.upload_target(cx)
    j = dev.upload_job(t, cx, seed=2024)
    # warm-up (includes the autotuner's trial launches)
    dev.batch_run([j], args.warmup, exchange=exch if args.warmup >= exch > 0 else 0, graph=True)
    t0 = time.perf_counter()
    ms = dev.batch_run([j], args.steps, exchange=exch, graph=True)
    wall = time.perf_counter() - t0
    value = n_rep * args.steps / wall
    log(f"batched: {args.steps} steps in {wall:.3f} s wall ({ms:.1f} ms device), {value:.4g} poses/s")

    # roofline of the Monte Carlo kernel: one 100-step chunk per launch
    chunk = 100
    launch_ms = dev.batch_run([j], chunk, exchange=0, graph=False)
    flop_launch = f_pose * n_rep * chunk
    achieved = flop_launch / (launch_ms * 1e-3) / 1e12
    dev.free_job(j)

    # plain form: reference structure on a window of the same workload
    plain_value = None
    if args.plain_steps > 0:
        jp = dev.upload_job(t, cx, seed=2024)
        t0 = time.perf_counter()
        for s1 in range(0, args.plain_steps, 100):
            s2 = min(100, args.plain_steps - s1)
            dev.montecarlo(jp, s1, s2)
            dev.records(jp, n_rep, 100)
            if exch and (s1 + s2) % exch == 0:
                dev.exchange(jp, (s1 + s2) // exch - 1)
        plain_wall = time.perf_counter() - t0
        plain_value = n_rep * args.plain_steps / plain_wall
        dev.free_job(jp)
        log(f"plain: {args.plain_steps} steps in {plain_wall:.3f} s, {plain_value:.4g} poses/s")
    dev.free_target(t)
    dev.close()

    cpu = None if args.no_cpu else cpu_baseline(cx, args.cpu_seconds)
    if cpu:
        log(f"cpu oracle: {cpu['value']:.4g} poses/s ({cpu['sample']})")

    out = {
        "metric": METRIC, "value": value, "unit": "pose evaluations/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{args.workload}: ligand {cx.size['lig_natom']} atoms x {cx.size['n_lig']} conf, "
                               f"protein {cx.size['prt_npoint']} points x {cx.size['n_prt']} conf, "
                               f"{cx.size['n_tmp']} temperatures, {n_rep} replicas, exchange every {exch}",
                   "replicas": n_rep, "form": "batched (packed replicas, autotuned shape, hipGraph chunks)"},
        "plain": {"value": plain_value, "steps": args.plain_steps,
                  "form": "one workgroup per replica, 100-step launches, record read-back per chunk"},
        "speedup_vs_plain": value / plain_value if plain_value else None,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": perf.FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / perf.FP32_PEAK_TFLOPS, "traffic": None,
                     "kernel": "mc_kernel", "launch_ms": launch_ms, "flop_per_pose": f_pose,
                     "fma_issue_peak": perf.FP32_FMA_TFLOPS},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
