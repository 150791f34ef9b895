"""The no-allocation batched entry point (gd_batch_run_noalloc) against
gd_batch_run: the same jobs, seeds and explicit shapes give bit-identical
replica states and McLog counters, on two calls in a row, the second with a
smaller batch (the fixed device array keeps the first call's contents)."""
import ctypes as C
import sys

import numpy as np
import pytest

from geauxdock_cs_v2_amd import ctypes_defs as D
from geauxdock_cs_v2_amd import lib as L
from geauxdock_cs_v2_amd import synth

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


def _state(lib, ctx, job, n):
    reps = (D.ReplicaMC * n)()
    L.check(lib.gd_read_replicas(ctx, job, reps))
    mc, ex = np.zeros(n, np.int32), np.zeros(n, np.int32)
    L.check(lib.gd_read_mclog(ctx, job, mc.ctypes.data_as(C.POINTER(C.c_int)), ex.ctypes.data_as(C.POINTER(C.c_int))))
    return bytes(reps), mc, ex


@pytest.mark.parametrize("case", [("c2", 20, 10, 4), ("c1", 10, 5, 4)])
def test_noalloc_matches_batch_run(lib, ctx, case):
    name, steps, exchange, slices = case
    cx, _ = synth.load_config(name)
    n = cx.size["n_rep"]
    say(f"case: {name}, {n} replicas, {steps} steps, exchange every {exchange}, no-allocation vs allocating")
    t, ja, jb = C.c_int(), C.c_int(), C.c_int()
    L.check(lib.gd_upload_target(ctx, C.byref(cx.view), C.byref(t)))
    L.check(lib.gd_upload_job(ctx, t, C.byref(cx.view), 41, C.byref(ja)))
    L.check(lib.gd_upload_job(ctx, t, C.byref(cx.view), 41, C.byref(jb)))
    o = D.BatchOpts(steps, exchange, 1, 0, 12, slices, 1, 0)
    ms = C.c_double()
    L.check(lib.gd_batch_run(ctx, (C.c_int * 1)(ja.value), 1, C.byref(o), C.byref(ms)))
    L.check(lib.gd_batch_run_noalloc(ctx, (C.c_int * 1)(jb.value), 1, C.byref(o), C.byref(ms)))
    plan = (C.c_int * 3)()
    L.check(lib.gd_batch_plan(ctx, plan))
    say(f"plan: {plan[0]} tasks, {plan[1]} workgroups, {plan[2]} step slices; "
        f"array {lib.gd_batch_noalloc_capacity()} bytes")
    a, b = _state(lib, ctx, ja, n), _state(lib, ctx, jb, n)
    for j in (ja, jb):
        L.check(lib.gd_free_job(ctx, j))
    L.check(lib.gd_free_target(ctx, t))
    assert a[0] == b[0]
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    assert a[2].sum() > 0  # exchange rounds ran
