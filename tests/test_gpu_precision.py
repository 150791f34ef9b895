"""GPU precision: the compensated FP32 path against the FP64 oracle on the
large stress shape and on the batched screening library.

Stress shape: the 20,000-point protein in a 35 A sphere with 32
conformations, a 150-atom ligand, KDE 5,000, MCS 100 rows, 32 temperatures,
exchange every 50 steps; one ligand conformation (1,024 replicas) so 5,000
steps fit a test.  At steps 0, 1,000 and 5,000 the replicas' stored energies
are re-evaluated by the FP64 oracle at their stored moves.  The bar is
REL_TOL on every normalised term (floor TERM_FLOOR)."""
import ctypes as C
import sys

import numpy as np
import pytest

from conftest import REL_TOL, rel_err
from geauxdock_cs_v2_amd import ctypes_defs as D
from geauxdock_cs_v2_amd import lib as L
from oracle import bind as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx(lib):
    h = C.c_void_p()
    L.check(lib.gd_create(0, C.byref(h)))
    yield h
    lib.gd_destroy(h)


def run(lib, ctx, jobs, steps, exchange=0, comp=1):
    o = D.BatchOpts(steps, exchange, 1, 0, 0, 0, comp, 0)
    arr = (C.c_int * len(jobs))(*jobs)
    ms = C.c_double()
    L.check(lib.gd_batch_run(ctx, arr, len(jobs), C.byref(o), C.byref(ms)))
    return ms.value


def stored_error(lib, ctx, cx, job, sample):
    """Max relative error per term of the stored replica energies against the
    FP64 oracle at the stored moves."""
    n = cx.size["n_rep"]
    reps = (D.ReplicaMC * n)()
    L.check(lib.gd_read_replicas(ctx, job, reps))
    idx = np.asarray(sample, np.int32)
    mm = np.array([list(reps[i].movematrix)[:6] for i in idx], np.float32)
    eg = np.array([list(reps[i].energy)[:10] for i in idx])
    e64 = O.energy(cx.view, idx, mm, "f64")
    return rel_err(eg, e64).max(0)


def test_stress_shape_error_sweep(lib, ctx):
    from geauxdock_cs_v2_amd import synth
    cx, _ = synth.load_config("c5", n_lig_conf=1)
    n = cx.size["n_rep"]
    sample = np.arange(0, n, n // 32)
    t, j = C.c_int(), C.c_int()
    L.check(lib.gd_upload_target(ctx, C.byref(cx.view), C.byref(t)))
    L.check(lib.gd_upload_job(ctx, t, C.byref(cx.view), 55, C.byref(j)))
    errs = {}
    done = 0
    for target in (1, 1000, 5000):  # steps taken: step 0 only, then up to 999 and 4999
        say(f"case: stress shape, {cx.size['prt_npoint']} points x {cx.size['n_prt']} conf, "
            f"{n} replicas, steps {done}..{target - 1}")
        ms = run(lib, ctx, [j.value], target - done, exchange=50 if target > 1 else 0)
        done = target
        errs[target - 1] = stored_error(lib, ctx, cx, j.value, sample)
        say(f"  step {target - 1}: {ms:.0f} ms; max rel err vs FP64:",
            np.array2string(errs[target - 1], precision=2, max_line_width=200))
    L.check(lib.gd_free_job(ctx, j))
    L.check(lib.gd_free_target(ctx, t))
    for step, e in errs.items():
        assert e.max() <= REL_TOL, (step, e)


def test_screening_subset_error(lib, ctx):
    """64 ligands of the screening library (every 16th) batched in one launch
    for 20 steps; 4 replicas of each job against FP64."""
    from geauxdock_cs_v2_amd import synth
    info = synth.many_files("screen")
    sub = dict(info, ligands=info["ligands"][::16])
    cxs = synth.load_many(sub, threads=8)
    t = C.c_int()
    L.check(lib.gd_upload_target(ctx, C.byref(cxs[0].view), C.byref(t)))
    jobs = []
    for i, cx in enumerate(cxs):
        j = C.c_int()
        L.check(lib.gd_upload_job(ctx, t, C.byref(cx.view), 300 + i, C.byref(j)))
        jobs.append(j.value)
    say(f"case: screening subset, {len(jobs)} ligands batched, 20 steps")
    run(lib, ctx, jobs, 20)
    worst = np.zeros(10)
    for cx, j in zip(cxs, jobs):
        n = cx.size["n_rep"]
        worst = np.maximum(worst, stored_error(lib, ctx, cx, j, np.arange(0, n, n // 4)))
        L.check(lib.gd_free_job(ctx, j))
    L.check(lib.gd_free_target(ctx, t))
    say("screening max rel err vs FP64:", np.array2string(worst, precision=2, max_line_width=200))
    assert worst.max() <= REL_TOL, worst
