"""Phase profile of the batched Monte Carlo kernel on one GPU.

    python -m geauxdock_cs_v2_amd.phase_profile c5 20
    python -m geauxdock_cs_v2_amd.phase_profile c2 20 --shape 12,3,4
    python -m geauxdock_cs_v2_amd.phase_profile screen 100

Loads the phase-profile build of the library (libgeauxdock_amd_phase.so,
built by build.build_phase_profile: the product kernel with clock marks at the
phase boundaries of each MC step), runs a named synthetic workload in the
batched form (warm-up of 5 steps, autotuned or explicit shape), then times
`steps` steps and prints the wave-cycles spent in each phase, the pair-loop
cycles of each wave index, and the imbalance of the pair loops within a
workgroup (W times the slowest wave over the sum of the waves, per workgroup
step).  The marks cost time of their own; the shares are what this tool is for.
"""
import argparse
import ctypes as C
import os
import sys

from . import build as B
from . import lib as L

PHASES = ["proposal and rotation", "KDE", "MCS", "pair loops", "PSP points", "cross-slice sum and barriers",
          "per-slot terms and barrier", "per-replica sums and barrier", "accept and barrier"]


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("config", help="synthetic workload name (c2, c5, ...), or screen for the 1,024-ligand batch")
    ap.add_argument("steps", type=int)
    ap.add_argument("--shape", default="", help="explicit W,SW,S (waves, slot waves, point slices)")
    args = ap.parse_args(argv)
    if not os.path.exists(B.PHASE_LIB):
        raise ImportError(f"{B.PHASE_LIB} is missing: build it with build.build_phase_profile()")
    L.LIB_PATH = B.PHASE_LIB  # every binding below loads the phase-profile build
    from . import synth
    from .device import Device
    lib = L.load()
    lib.gd_phase_profile.argtypes = [C.POINTER(C.c_ulonglong), C.c_int]
    lib.gd_phase_profile.restype = C.c_int
    lib.gd_phase_profile_wg.argtypes = [C.POINTER(C.c_ulonglong), C.c_int]
    lib.gd_phase_profile_wg.restype = C.c_int
    shape = {}
    if args.shape:
        w, sw, sl = map(int, args.shape.split(","))
        shape = dict(waves=w, slots_per_wave=sw, point_slices=sl)
    buf = (C.c_ulonglong * 32)()
    with Device(0) as dev:
        if args.config == "screen":
            from . import perf
            info = synth.many_files("screen")
            cxs = synth.load_many(info, threads=perf.host_threads())
            t = dev.upload_target(cxs[0])
            jobs = [dev.upload_job(t, c, seed=100 + i) for i, c in enumerate(cxs)]
            exch, warm = 0, 10
        else:
            cx, info = synth.load_config(args.config)
            exch, warm = info["cfg"]["exchange"], 5
            t = dev.upload_target(cx)
            jobs = [dev.upload_job(t, cx, seed=2024)]
        dev.batch_run(jobs, warm, exchange=0, graph=True, **shape)
        if lib.gd_phase_profile(buf, 1):
            raise RuntimeError("gd_phase_profile failed")
        ms = dev.batch_run(jobs, args.steps, exchange=exch, graph=True, **shape)
        if lib.gd_phase_profile(buf, 1):
            raise RuntimeError("gd_phase_profile failed")
        W = dev.job_shape(jobs[0])[0]
        plan = dev.batch_plan()
        nwg = plan[1]
        wg = (C.c_ulonglong * (2 * nwg))()
        if lib.gd_phase_profile_wg(wg, nwg):
            raise RuntimeError("gd_phase_profile_wg failed")
        print(f"{args.config}: {args.steps} steps in {ms:.2f} ms device, shape (waves, slot waves, slices, "
              f"replicas) of the first job {dev.job_shape(jobs[0])}; last launch {plan[0]} tasks in {plan[2]} "
              f"step slices over {nwg} workgroups")
    tot = sum(buf[k] for k in range(len(PHASES)))
    for k, name in enumerate(PHASES):
        print(f"  {name:32s} {100.0 * buf[k] / tot:6.2f} %")
    waves = [buf[16 + w] for w in range(W)]
    n = max(1, buf[10])
    print(f"  pair-loop imbalance (W x slowest / sum over waves, per workgroup step): {buf[9] / max(1, sum(waves)):.3f}"
          f" over {buf[10]} workgroup steps")
    print("  mean pair-loop cycles per wave index: " + " ".join(f"{w / n:.3g}" for w in waves))
    # the last launch's workgroups on the wall clock: the span from the first
    # start to the last exit, and the share of workgroup-time spent before a
    # workgroup's start (ramp) and after its exit (queue tail)
    t0 = [wg[i] for i in range(nwg)]
    t1 = [wg[nwg + i] for i in range(nwg)]
    a, b = min(t0), max(t1)
    span = max(1, b - a)
    ramp = sum(x - a for x in t0) / (nwg * span)
    tail = sum(b - x for x in t1) / (nwg * span)
    print(f"  last launch on the wall clock: {nwg} workgroups over {span} ticks; workgroup-time before the "
          f"starts {100 * ramp:.2f} %, after the exits (queue tail) {100 * tail:.2f} %")


if __name__ == "__main__":
    sys.exit(main())
