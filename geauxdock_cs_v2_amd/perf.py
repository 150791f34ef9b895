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


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def host_threads():
    """Threads this process may use: the affinity mask, capped at 16 (the
    CPU share of one GPU on the measurement hosts)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))
