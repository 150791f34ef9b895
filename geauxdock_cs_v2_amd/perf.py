"""Algorithmic work counts and host facts for the measurements.

FLOP per pose (one replica x one MC step), counting +,-,*,/,sqrt,exp,log as
one FLOP each and compares as none:
  60 N_L N_P            pair terms (distance 8, hydrophobic 12, LJ 12,
                        electrostatics 11, contact 8, PSP 1, H-bond 8)
  + 12 sum_l K(t_l)     KDE pairs of matching type
  + 9 sum_m ncol_m + 4 M   MCS
  + 36 N_L + 60         two atom transforms, rotation, distance, normalise, accept
"""
import os
import platform
import subprocess

import numpy as np

# MI355X published FP32 vector peak: one wave64 FMA per SIMD every two
# cycles, which gfx950 issues for plain (non-packed) FP32 FMAs with vector
# operands; transcendentals (v_rsq, v_rcp, v_exp) take about four times as long
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


def flops_per_pose(cx):
    s = cx.size
    nl, npt = s["lig_natom"], s["prt_npoint"]
    lt = cx.array("lig_t3")[:nl]
    kt = cx.array("kde_t")
    kde_pairs = int(sum(np.count_nonzero(kt == t) for t in lt))
    ncol = cx.array("mcs_ncol")
    return 60 * nl * npt + 12 * kde_pairs + 9 * int(ncol.sum()) + 4 * s["mcs_nrow"] + 36 * nl + 60


def _lscpu():
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, text=True).stdout
    except OSError:
        return {}
    kv = {}
    for line in out.splitlines():
        if ":" in line:
            k, v = line.split(":", 1)
            kv[k.strip()] = v.strip()
    return kv


def cpu_model():
    return _lscpu().get("Model name") or platform.processor() or "unknown"


def affinity_cpus():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


# CPU threads one GPU's share of the measurement hosts allows (those hosts
# give each GPU 16 host cores, whatever nproc reports)
HOST_THREAD_CAP = 16


def host_threads():
    """Threads the CPU baseline uses: the affinity mask, capped at
    HOST_THREAD_CAP."""
    return max(1, min(HOST_THREAD_CAP, affinity_cpus()))


def host_info():
    """lscpu model, sockets, cores per socket and threads per core, the
    affinity mask's CPU count, the thread cap and the OpenMP binding."""
    kv = _lscpu()

    def num(key):
        try:
            return int(kv.get(key, ""))
        except ValueError:
            return None
    return {"model": kv.get("Model name") or platform.processor() or "unknown", "sockets": num("Socket(s)"),
            "cores_per_socket": num("Core(s) per socket"), "threads_per_core": num("Thread(s) per core"),
            "logical_cpus": num("CPU(s)"), "affinity_cpus": affinity_cpus(), "thread_cap": HOST_THREAD_CAP,
            "omp_proc_bind": os.environ.get("OMP_PROC_BIND"), "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}
