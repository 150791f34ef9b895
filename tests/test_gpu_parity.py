"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bars: energies within REL_TOL of the FP64 oracle (floor TERM_FLOOR, see
conftest); the move proposals bit-exact (same stream, same float rounding);
accept and exchange decisions identical except where the accept uniform sits
within the energy tolerance of the threshold."""
import ctypes as C
import math
import os
import sys

import numpy as np
import pytest

from conftest import ASTEX, KAT_1A07C1, REL_TOL, TERM_FLOOR, rel_err
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


def upload(lib, ctx, cx, seed=1234):
    t, j = C.c_int(), C.c_int()
    L.check(lib.gd_upload_target(ctx, C.byref(cx.view), C.byref(t)))
    L.check(lib.gd_upload_job(ctx, t, C.byref(cx.view), seed, C.byref(j)))
    return t.value, j.value


def free(lib, ctx, t, j):
    L.check(lib.gd_free_job(ctx, j))
    L.check(lib.gd_free_target(ctx, t))


def gpu_energy(lib, ctx, job, rep_idx, move):
    rep_idx = np.ascontiguousarray(rep_idx, np.int32)
    move = np.ascontiguousarray(move, np.float32).reshape(-1, 6)
    out = np.zeros((len(rep_idx), 10), np.float32)
    L.check(lib.gd_energy(ctx, job, len(rep_idx), rep_idx.ctypes.data_as(C.POINTER(C.c_int)),
                          move.ctypes.data_as(C.POINTER(C.c_float)), out.ctypes.data_as(C.POINTER(C.c_float))))
    return out


def read_reps(lib, ctx, job, n):
    arr = (D.ReplicaMC * n)()
    L.check(lib.gd_read_replicas(ctx, job, arr))
    return arr


# decision windows around the FP64 Metropolis threshold: the GPU's total is
# within REL_TOL of FP64 (the parity bar); the oracle's sequential FP32 sums
# reach ~1e-3 relative on the total
GPU_WINDOW = REL_TOL
CPU_WINDOW = 1e-3


def accept_uniform(lib, seed, idx_rep, step):
    """The accept draw as both sides use it: (float)k / (float)RAND_MAX with
    (float)RAND_MAX == 2^31."""
    return float(np.float32(lib.gd_stream_k(seed, 0, idx_rep, step, 6)) * np.float32(2.0 ** -31))


def shifted_e64(cx, r, move):
    """FP64 energies of replica r at `move` translated by 1e-4 A along +-x,
    +-y, +-z: (6, 10)."""
    mv = np.repeat(np.asarray(move, np.float64)[None, :6], 6, 0)
    mv[:, :3] += np.concatenate([np.eye(3), -np.eye(3)]) * 1e-4
    return O.energy(cx.view, np.full(6, r, np.int32), mv.astype(np.float32), "f64")


PSP_TERMS = {3, 9}  # the pocket-specific term and the total that carries it


def psp_step_ties(cx, rep_idx, moves, eg, e64):
    """Entries whose only disagreement with FP64 is in the pocket-specific
    term (and the total): that term adds a table entry iff an atom-point
    distance is within pmf0 (energy_prt_v1.cpp:107-109), a step function that
    flips when the distance sits at pmf0 within float rounding.  Such an entry
    passes when FP64 at a pose translated by 1e-4 A along an axis gives the
    GPU's values within REL_TOL.  Returns the entries that stay unexplained."""
    errs = rel_err(eg, e64)
    bad = [b for b in range(len(eg)) if errs[b].max() > REL_TOL]
    left = []
    for b in bad:
        if not set(np.nonzero(errs[b] > REL_TOL)[0]) <= PSP_TERMS:
            left.append(b)
            continue
        ep = shifted_e64(cx, rep_idx[b], moves[b])
        if not np.any(rel_err(np.repeat(eg[b][None, :], 6, 0), ep).max(1) <= REL_TOL):
            left.append(b)
    return bad, left


def compare_step(lib, cx, seed, step, before, g, o):
    """One MC step taken by the GPU (g) and by the oracle (o) from the same
    state `before`.  Proposals are rebuilt from the stream and must be
    bit-identical to every accepted move.  Each side's accept decision is
    judged against the FP64 threshold exp((E64 - E_old) minus_beta) at the
    proposal: the GPU's may differ from u < thr64 only where u lies within
    GPU_WINDOW of the threshold (its energies are within that of FP64), the
    oracle's only within CPU_WINDOW.  Accepted GPU states carry energies
    within REL_TOL of FP64; rejected ones are unchanged.  Returns the number
    of GPU/oracle disagreements (each one a near tie on at least one side)."""
    n = len(g)
    idx = np.arange(n, dtype=np.int32)
    prop = np.zeros((n, 6), np.float32)
    scale = np.array(cx.view.mcpara.move_scale[:6], np.float32)
    for r in range(n):
        if step == 0:
            prop[r] = np.array(before[r].movematrix[:6], np.float32)
        else:
            u = np.array([lib.gd_stream_k(seed, 0, before[r].idx_rep, step, j) for j in range(6)], np.float32)
            prop[r] = scale * (u / np.float32(2147483647)) + np.array(before[r].movematrix[:6], np.float32)
    e64 = O.energy(cx.view, idx, prop, "f64")
    flips = 0
    for r in range(n):
        if step == 0:
            acc64, tol_g, tol_c, u, thr = True, 0.0, 0.0, 0.0, 1.0
        else:
            u = accept_uniform(lib, seed, before[r].idx_rep, step)
            mb = float(cx.view.temp[before[r].idx_tmp].minus_beta)
            thr = math.exp((e64[r][9] - float(before[r].energy[9])) * mb)
            acc64 = u < thr
            scale_t = max(abs(e64[r][9]), TERM_FLOOR) * abs(mb) * min(thr, 1.0)
            tol_g = GPU_WINDOW * scale_t + 1e-7
            tol_c = CPU_WINDOW * scale_t + 1e-7
        ga, oa = bool(g[r].is_accept), bool(o[r].is_accept)
        if ga != acc64 and abs(u - thr) > tol_g:
            # a pocket-specific step tie (see psp_step_ties): the decision
            # must be the FP64 one at a pose within 1e-4 A
            ep = shifted_e64(cx, r, prop[r])
            thr_p = np.exp((ep[:, 9] - float(before[r].energy[9])) * mb)
            assert np.any((u < thr_p) == ga), ("GPU decision off the FP64 threshold", step, r, u, thr, tol_g)
        if oa != acc64:
            assert abs(u - thr) <= tol_c, ("oracle decision off the FP64 threshold", step, r, u, thr, tol_c)
        flips += ga != oa
        if ga:
            assert np.array_equal(np.array(g[r].movematrix[:6], np.float32), prop[r]), (step, r)
            eg_r = np.array(g[r].energy[:10], np.float64)
            if rel_err(eg_r, e64[r]).max() > REL_TOL:
                _, left = psp_step_ties(cx, np.array([r], np.int32), prop[r:r + 1], eg_r[None], e64[r:r + 1])
                assert not left, (step, r, rel_err(eg_r, e64[r]))
            assert g[r].step == step
        else:  # rejected: state unchanged
            assert list(g[r].movematrix) == list(before[r].movematrix)
            assert list(g[r].energy) == list(before[r].energy) and g[r].step == before[r].step
        if oa:
            assert np.array_equal(np.array(o[r].movematrix[:6], np.float32), prop[r]), (step, r)
            assert o[r].step == step
    return flips


def test_kat_on_gpu(lib, ctx, kat_complex):
    say("case: KAT 1a07C1 step-0 energy on the GPU")
    _, cx = kat_complex
    t, j = upload(lib, ctx, cx)
    e = gpu_energy(lib, ctx, j, [0], np.zeros(6))[0]
    free(lib, ctx, t, j)
    assert np.all(np.abs(e - KAT_1A07C1) <= 5e-5), e


@pytest.mark.parametrize("which", ["kat11", "c1", "c2"])
def test_pose_energies_vs_fp64(lib, ctx, kat_complex11, which):
    from geauxdock_cs_v2_amd import synth
    if which == "kat11":
        cx = kat_complex11[1]
    else:
        cx, _ = synth.load_config(which)
    n_rep = cx.size["n_rep"]
    rng = np.random.default_rng(7)
    n = min(256, n_rep)
    idx = rng.choice(n_rep, n, replace=False).astype(np.int32)
    move = np.concatenate([rng.normal(0, 0.6, (n, 3)), rng.uniform(-math.pi, math.pi, (n, 3))], 1).astype(np.float32)
    move[0] = 0
    say(f"case: pose energies {which} n={n}")
    t, j = upload(lib, ctx, cx)
    eg = gpu_energy(lib, ctx, j, idx, move)
    free(lib, ctx, t, j)
    e64 = O.energy(cx.view, idx, move, "f64")
    e32 = O.energy(cx.view, idx, move, "f32")
    err_gpu = rel_err(eg, e64).max(0)
    err_ref32 = rel_err(e32, e64).max(0)
    say("max rel err vs FP64, GPU  :", np.array2string(err_gpu, precision=2))
    say("max rel err vs FP64, CPU32:", np.array2string(err_ref32, precision=2))
    assert np.all(np.isfinite(eg))
    assert err_gpu.max() <= REL_TOL, err_gpu


def test_montecarlo_chunks_match_oracle(lib, ctx):
    """Plain chunks of the C1 workload step by step: after every step the GPU
    state is written into the oracle's copy (resynchronised), the oracle takes
    the same step, and decisions, moves and energies are compared."""
    from geauxdock_cs_v2_amd import synth
    cx, _ = synth.load_config("c1")
    n = cx.size["n_rep"]
    seed = 31337
    t, j = upload(lib, ctx, cx, seed)
    flips = 0
    for step in range(0, 60):
        say(f"case: c1 plain step {step}")
        before = read_reps(lib, ctx, j, n)
        L.check(lib.gd_montecarlo(ctx, j, step, 1))
        g = read_reps(lib, ctx, j, n)
        o = (D.ReplicaMC * n)()
        C.memmove(o, before, C.sizeof(o))
        O.montecarlo(cx.view, o, step, 1, seed)
        flips += compare_step(lib, cx, seed, step, before, g, o)
    free(lib, ctx, t, j)
    say("near-tie decision flips:", flips)
    assert flips <= 2


def test_montecarlo_long_chunk_records(lib, ctx, kat_complex11):
    """A 100-step chunk of the 11-conformation 1a07C job against the oracle's
    own 100-step chunk from the same state (kernel_cpp_montecarlo_soa.cpp:
    367-389 appends the replica on every accepted step).  Every replica whose
    GPU trajectory stays on the oracle's (no near-tie decision flip) has the
    same record window entry by entry: moves and steps bit-identical, the
    index fields and is_accept equal, energies within REL_TOL of FP64 at the
    recorded move; a replica that left it matches up to its first difference.
    The window ends at the final state."""
    cx = kat_complex11[1]
    n = cx.size["n_rep"]
    seed = 5
    say("case: kat11 plain 100-step chunk records against the oracle's")
    t, j = upload(lib, ctx, cx, seed)
    start = read_reps(lib, ctx, j, n)
    L.check(lib.gd_montecarlo(ctx, j, 0, 100))
    g = read_reps(lib, ctx, j, n)
    rec = (D.ReplicaMC * (n * 100))()
    ne = np.zeros(n, np.int32)
    L.check(lib.gd_read_records(ctx, j, rec, ne.ctypes.data_as(C.POINTER(C.c_int)), 100))
    free(lib, ctx, t, j)
    o = (D.ReplicaMC * n)()
    C.memmove(o, start, C.sizeof(o))
    orec, one = O.montecarlo(cx.view, o, 0, 100, seed, cap=100)
    assert np.all(ne >= 1) and np.all(ne <= 100)
    diverged, checked = 0, 0
    moves, reps = [], []
    for r in range(n):
        same = ne[r] == one[r] and list(g[r].movematrix) == list(o[r].movematrix) and g[r].step == o[r].step
        diverged += not same
        for k in range(ne[r] if same else min(ne[r], one[r])):
            a, b = rec[r * 100 + k], orec[r * 100 + k]
            if not same and (a.step != b.step or list(a.movematrix) != list(b.movematrix)):
                break  # past the first near-tie decision flip
            assert a.step == b.step and list(a.movematrix) == list(b.movematrix), (r, k)
            assert (a.idx_rep, a.idx_lig, a.idx_prt, a.idx_tmp, a.is_accept) == \
                (b.idx_rep, b.idx_lig, b.idx_prt, b.idx_tmp, b.is_accept), (r, k)
            moves.append(np.array(a.movematrix[:6], np.float32))
            reps.append((r, k))
            checked += 1
        last = rec[r * 100 + ne[r] - 1]
        assert last.step == g[r].step and list(last.movematrix) == list(g[r].movematrix)
    # energies of every compared entry against FP64 at its move
    idx = np.array([r for r, _ in reps], np.int32)
    e64 = O.energy(cx.view, idx, np.array(moves), "f64")
    eg = np.array([list(rec[r * 100 + k].energy)[:10] for r, k in reps])
    errs = rel_err(eg, e64)
    bad, left = psp_step_ties(cx, idx, np.array(moves), eg, e64)
    ok = np.ones(len(eg), bool)
    ok[bad] = False
    say(f"record entries compared: {checked} over {n} replicas, {diverged} left the oracle's trajectory; "
        f"max rel err vs FP64 {errs[ok].max():.2e} ({len(bad)} pocket-specific step ties, "
        f"{len(bad) - len(left)} of them reproduced within 1e-4 A)")
    for b in left[:5]:
        r, k = reps[b]
        say(f"replica {r} entry {k} step {rec[r * 100 + k].step} of {ne[r]}: term {errs[b].argmax()}, "
            f"GPU {np.array2string(eg[b], precision=5)}, FP64 {np.array2string(e64[b], precision=5)}")
    assert not left
    assert len(bad) <= max(2, checked // 20000)
    assert diverged <= max(2, n // 200)


def test_batched_records_with_step_slices(lib, ctx):
    """A batched 20-step chunk of the full replica ensemble (3,520 replicas:
    more tasks than resident workgroups, so each task runs in several step
    slices handed from workgroup to workgroup) with records kept, against the
    oracle's chunk for a sample of replicas: record windows entry by entry as
    in the plain test, final states equal, and is_accept in the entries the
    value from before the chunk (the reference sets it after the chunk)."""
    from geauxdock_cs_v2_amd import synth
    cx, _ = synth.load_config("c2")
    n = cx.size["n_rep"]
    seed = 77
    t, j = upload(lib, ctx, cx, seed)
    start = read_reps(lib, ctx, j, n)
    say("case: c2 batched 20-step chunk with records and step slices")
    _batch(lib, ctx, [j], 20, comp=1, records=1)
    plan = (C.c_int * 3)()
    L.check(lib.gd_batch_plan(ctx, plan))
    say(f"plan: {plan[0]} tasks, {plan[1]} workgroups, {plan[2]} step slices")
    g = read_reps(lib, ctx, j, n)
    rec = (D.ReplicaMC * (n * 20))()
    ne = np.zeros(n, np.int32)
    L.check(lib.gd_read_records(ctx, j, rec, ne.ctypes.data_as(C.POINTER(C.c_int)), 20))
    free(lib, ctx, t, j)
    assert plan[2] > 1 and plan[1] < plan[0] * plan[2]
    sample = list(range(64))  # protein conformation 0, temperatures 0-3
    o = (D.ReplicaMC * n)()
    C.memmove(o, start, C.sizeof(o))
    rr, nn = O.montecarlo(cx.view, o, 0, 20, seed, cap=20, r0=0, r1=64)
    orec = {r: ([rr[r * 20 + k] for k in range(nn[r])], nn[r]) for r in sample}
    diverged = 0
    for r in sample:
        entries, cnt = orec[r]
        same = ne[r] == cnt and list(g[r].movematrix) == list(o[r].movematrix) and g[r].step == o[r].step
        diverged += not same
        if not same:
            continue
        assert g[r].is_accept == o[r].is_accept
        for k in range(cnt):
            a, b = rec[r * 20 + k], entries[k]
            assert a.step == b.step and list(a.movematrix) == list(b.movematrix), (r, k)
            assert (a.idx_rep, a.idx_lig, a.idx_prt, a.idx_tmp) == (b.idx_rep, b.idx_lig, b.idx_prt, b.idx_tmp)
            assert a.is_accept == b.is_accept == start[r].is_accept, (r, k)
    say(f"{len(sample) - diverged} of {len(sample)} sampled replicas on the oracle's trajectory")
    assert diverged <= 2


def test_exchange_matches_oracle(lib, ctx, kat_complex11):
    cx = kat_complex11[1]
    n = cx.size["n_rep"]
    seed = 99
    t, j = upload(lib, ctx, cx, seed)
    rng = np.random.default_rng(3)
    reps = cx.replicas()
    for r in range(n):
        reps[r].energy[9] = float(rng.normal(0, 0.05))
    L.check(lib.gd_write_replicas(ctx, j, reps))
    for rnd in range(6):
        say(f"case: exchange round {rnd}")
        acc = C.c_int()
        L.check(lib.gd_exchange(ctx, j, rnd, C.byref(acc)))
        want = O.exchange(cx.view, reps, rnd, seed)
        g = read_reps(lib, ctx, j, n)
        assert acc.value == want
        assert [g[r].idx_tmp for r in range(n)] == [reps[r].idx_tmp for r in range(n)]
    free(lib, ctx, t, j)


def _batch(lib, ctx, jobs, steps, exchange=0, graph=0, comp=0, waves=0, records=0):
    o = D.BatchOpts(steps, exchange, graph, 0, waves, 0, comp, records)
    arr = (C.c_int * len(jobs))(*jobs)
    ms = C.c_double()
    L.check(lib.gd_batch_run(ctx, arr, len(jobs), C.byref(o), C.byref(ms)))
    return ms.value


def test_batched_steps_match_oracle(lib, ctx):
    """Batched form (packed replicas, autotuned shape) step by step against
    the oracle, resynchronised as in the plain test."""
    from geauxdock_cs_v2_amd import synth
    cx, _ = synth.load_config("c1")
    n = cx.size["n_rep"]
    seed = 8
    t, j = upload(lib, ctx, cx, seed)
    flips = 0
    for step in range(0, 30):
        say(f"case: c1 batched step {step}")
        before = read_reps(lib, ctx, j, n)
        _batch(lib, ctx, [j], 1)
        g = read_reps(lib, ctx, j, n)
        o = (D.ReplicaMC * n)()
        C.memmove(o, before, C.sizeof(o))
        O.montecarlo(cx.view, o, step, 1, seed)
        flips += compare_step(lib, cx, seed, step, before, g, o)
    free(lib, ctx, t, j)
    assert flips <= 2


def test_graph_replay_is_bit_identical(lib, ctx):
    """The same batched run with and without hipGraph capture, with exchange."""
    from geauxdock_cs_v2_amd import synth
    cx, _ = synth.load_config("c1")
    n = cx.size["n_rep"]
    res = []
    for graph in (0, 1):
        say(f"case: c1 batched 300 steps exchange 100 graph={graph}")
        t, j = upload(lib, ctx, cx, 77)
        _batch(lib, ctx, [j], 300, exchange=100, graph=graph, waves=8)
        res.append(bytes(read_reps(lib, ctx, j, n)))
        free(lib, ctx, t, j)
    assert res[0] == res[1]


@pytest.mark.parametrize("slices", [1, 4])
def test_issue_priority_policies_are_bit_identical(lib, ctx, slices):
    """The pair loops' issue-priority policies (GD_BOOST_POLICY: after half
    of the siblings, after any, never, and the progress-balancing kernel of
    its own) change only which wave issues first: 20 compensated steps with
    an exchange give the same replica states, byte for byte."""
    from geauxdock_cs_v2_amd import synth
    cx, _ = synth.load_config("c2")
    n = cx.size["n_rep"]
    res = {}
    try:
        for policy in ("half", "any", "none", "progress"):
            os.environ["GD_BOOST_POLICY"] = policy
            say(f"case: c2 batched 20 steps, 12 waves x {slices} slices, policy {policy}")
            t, j = upload(lib, ctx, cx, 31)
            o = D.BatchOpts(20, 10, 0, 0, 12, slices, 1, 0)
            arr = (C.c_int * 1)(j)
            ms = C.c_double()
            L.check(lib.gd_batch_run(ctx, arr, 1, C.byref(o), C.byref(ms)))
            res[policy] = bytes(read_reps(lib, ctx, j, n))
            free(lib, ctx, t, j)
    finally:
        os.environ.pop("GD_BOOST_POLICY", None)
    for policy in ("any", "none", "progress"):
        assert res[policy] == res["half"], policy


def test_cached_batch_shapes_match_tuned(lib, ctx):
    """A batch that finds the autotuner's choice in its cache runs the shapes
    the trials timed.  The first 16 screening ligands (16-79 atoms) are a
    batch whose packing heuristic picks different slot-wave counts for jobs
    of one ligand-size bucket (24 and 48 atoms), which a per-bucket cache of
    one job's packing would replace with that job's."""
    from geauxdock_cs_v2_amd import synth
    info = dict(synth.many_files("screen"))
    info["ligands"] = info["ligands"][:16]
    cxs = synth.load_many(info, threads=8)
    t = C.c_int()
    L.check(lib.gd_upload_target(ctx, C.byref(cxs[0].view), C.byref(t)))
    jobs = []
    for i, cx in enumerate(cxs):
        j = C.c_int()
        L.check(lib.gd_upload_job(ctx, t, C.byref(cx.view), 500 + i, C.byref(j)))
        jobs.append(j.value)

    def shapes():
        out = []
        for j in jobs:
            s = (C.c_int * 4)()
            L.check(lib.gd_job_shape(ctx, j, s))
            out.append(tuple(s))
        return out

    say("case: 16 screening ligands, autotuned batch then the cached choice")
    _batch(lib, ctx, jobs, 4, comp=1)
    tuned = shapes()
    _batch(lib, ctx, jobs, 4, comp=1)
    cached = shapes()
    for j in jobs:
        L.check(lib.gd_free_job(ctx, j))
    L.check(lib.gd_free_target(ctx, t))
    say("tuned slot waves:", [s[1] for s in tuned])
    assert cached == tuned


def test_issue_priority_policies_multi_job_batch(lib, ctx):
    """A batch of 12 screening ligands (15-50 atoms, per-job packings in one
    launch) under the half-done boost and under the progress-balancing
    kernel: 10 steps give the same replica states for every job, byte for
    byte."""
    from geauxdock_cs_v2_amd import synth
    info = dict(synth.many_files("screen"))
    info["ligands"] = info["ligands"][16:28]
    cxs = synth.load_many(info, threads=8)
    res = {}
    try:
        for policy in ("half", "progress"):
            os.environ["GD_BOOST_POLICY"] = policy
            say(f"case: 12 screening ligands batched 10 steps, policy {policy}")
            t = C.c_int()
            L.check(lib.gd_upload_target(ctx, C.byref(cxs[0].view), C.byref(t)))
            jobs = []
            for i, cx in enumerate(cxs):
                j = C.c_int()
                L.check(lib.gd_upload_job(ctx, t, C.byref(cx.view), 700 + i, C.byref(j)))
                jobs.append(j.value)
            _batch(lib, ctx, jobs, 10, comp=1)
            res[policy] = [bytes(read_reps(lib, ctx, j, cx.size["n_rep"])) for j, cx in zip(jobs, cxs)]
            for j in jobs:
                L.check(lib.gd_free_job(ctx, j))
            L.check(lib.gd_free_target(ctx, t))
    finally:
        os.environ.pop("GD_BOOST_POLICY", None)
    assert res["progress"] == res["half"]


def test_compensated_accumulation_within_tolerance(lib, ctx):
    """Kahan-compensated FP32 accumulation: 20 steps, final energies
    re-evaluated by the FP64 oracle at the GPU's final moves."""
    from geauxdock_cs_v2_amd import synth
    cx, _ = synth.load_config("c2")
    n = cx.size["n_rep"]
    say("case: c2 batched compensated 20 steps")
    t, j = upload(lib, ctx, cx, 21)
    _batch(lib, ctx, [j], 20, comp=1)
    g = read_reps(lib, ctx, j, n)
    free(lib, ctx, t, j)
    idx = np.arange(0, n, 37, dtype=np.int32)
    mm = np.array([list(g[i].movematrix)[:6] for i in idx], np.float32)
    e64 = O.energy(cx.view, idx, mm, "f64")
    eg = np.array([list(g[i].energy)[:10] for i in idx])
    err = rel_err(eg, e64).max(0)
    say("compensated max rel err vs FP64:", np.array2string(err, precision=2))
    assert err.max() <= REL_TOL


def test_two_ligands_one_resident_target(lib, ctx, tmp_path):
    """Two ligand jobs with their own .ff files (own KDE, MCS and PSP tables,
    own pocket centre) docked against one resident target, advanced together
    in one batched launch: energies of both against the FP64 oracle."""
    from geauxdock_cs_v2_amd import synth
    a = synth.generate(str(tmp_path), "jA", seed=3, n_atoms=21, n_lig_conf=3, lig_sigma=0.4, n_points=900,
                       n_prt_conf=2, prt_sigma=0.4, prt_radius=18.0, n_kde=80, mcs_rows=5, mcs_cols=8)
    b = synth.generate(str(tmp_path), "jB", seed=4, n_atoms=40, n_lig_conf=4, lig_sigma=0.4, n_points=900,
                       n_prt_conf=2, prt_sigma=0.4, prt_radius=18.0, n_kde=120, mcs_rows=7, mcs_cols=9,
                       protein=a["protein"])
    cxs = [L.load_job(v["pdb"], v["sdf"], v["ff"], k, n_tmp=3, temp_high=2.0, temp_low=0.3).stage()
           for k, v in (("jA", a), ("jB", b))]
    t = C.c_int()
    L.check(lib.gd_upload_target(ctx, C.byref(cxs[0].view), C.byref(t)))
    jobs = []
    for i, cx in enumerate(cxs):
        j = C.c_int()
        L.check(lib.gd_upload_job(ctx, t, C.byref(cx.view), 100 + i, C.byref(j)))
        jobs.append(j.value)
    say("case: two ligands, one target, batched 40 steps")
    # a job listed twice would have its replicas advanced twice at once
    o = D.BatchOpts(10, 0, 0, 0, 8, 0, 1, 0)
    dup = (C.c_int * 3)(jobs[0], jobs[1], jobs[0])
    assert lib.gd_batch_run(ctx, dup, 3, C.byref(o), None) == L.GD_ERR_ARG
    _batch(lib, ctx, jobs, 40, waves=8)
    for cx, j in zip(cxs, jobs):
        n = cx.size["n_rep"]
        g = read_reps(lib, ctx, j, n)
        idx = np.arange(n, dtype=np.int32)
        mm = np.array([list(r.movematrix)[:6] for r in g], np.float32)
        e64 = O.energy(cx.view, idx, mm, "f64")
        eg = np.array([list(r.energy)[:10] for r in g])
        err = rel_err(eg, e64).max(0)
        say("two-job max rel err:", np.array2string(err, precision=2))
        assert err.max() <= REL_TOL
        L.check(lib.gd_free_job(ctx, j))
    L.check(lib.gd_free_target(ctx, t))


def test_far_field_energies_large_protein(lib, ctx):
    """An 8,000-point protein in a 35 A sphere: most pairs are in the far
    field (only LJ and 1/d electrostatics vary there).  Poses translated by up
    to 8 A move the far boundary; every term stays within tolerance of FP64."""
    from geauxdock_cs_v2_amd import synth
    info = synth.many_files("serve", n=2)
    cx = synth.load_many(info, threads=4)[1]
    n_rep = cx.size["n_rep"]
    rng = np.random.default_rng(11)
    n = min(96, n_rep)
    idx = rng.choice(n_rep, n, replace=False).astype(np.int32)
    move = np.concatenate([rng.uniform(-8.0, 8.0, (n, 3)), rng.uniform(-math.pi, math.pi, (n, 3))], 1)
    move = move.astype(np.float32)
    move[0] = 0
    say(f"case: far field, {cx.size['prt_npoint']} points, {cx.size['lig_natom']} atoms, n={n}")
    t, j = upload(lib, ctx, cx)
    eg = gpu_energy(lib, ctx, j, idx, move)
    free(lib, ctx, t, j)
    e64 = O.energy(cx.view, idx, move, "f64")
    err = rel_err(eg, e64).max(0)
    say("far-field max rel err vs FP64:", np.array2string(err, precision=2))
    assert np.all(np.isfinite(eg))
    assert err.max() <= REL_TOL, err


@pytest.mark.parametrize("shift", [2.0, 4.0, 6.0, 9.0, 11.0, 14.0])
def test_expanded_distances_near_their_bounds(lib, ctx, shift):
    """Far and very far pairs take |u|^2 + |q|^2 + u.q' for waves within
    16 A of their ladder centre (u the atom, q the point, relative to it),
    whose rounding grows with |u| + |q|.  Ligands of the serving workload
    (up to 119 atoms, about 6 A across) translated by 2-14 A in random
    directions put their waves' farthest atom on both sides of that bound;
    every term stays within tolerance of FP64."""
    from geauxdock_cs_v2_amd import synth
    info = synth.many_files("serve", n=4)
    cx = max(synth.load_many(info, threads=4), key=lambda c: c.size["lig_natom"])
    n_rep = cx.size["n_rep"]
    rng = np.random.default_rng(int(shift * 10))
    n = min(96, n_rep)
    idx = rng.choice(n_rep, n, replace=False).astype(np.int32)
    v = rng.normal(size=(n, 3))
    v = v / np.linalg.norm(v, axis=1, keepdims=True) * shift
    move = np.concatenate([v, rng.uniform(-math.pi, math.pi, (n, 3))], 1).astype(np.float32)
    say(f"case: expanded distances, shift {shift} A, {cx.size['lig_natom']} atoms, "
        f"{cx.size['prt_npoint']} points, n={n}")
    t, j = upload(lib, ctx, cx)
    eg = gpu_energy(lib, ctx, j, idx, move)
    free(lib, ctx, t, j)
    e64 = O.energy(cx.view, idx, move, "f64")
    bad, left = psp_step_ties(cx, idx, move, eg, e64)
    err = rel_err(eg, e64).max(0)
    say(f"shift {shift} A max rel err vs FP64:", np.array2string(err, precision=2, max_line_width=200),
        f"({len(bad)} pocket-specific step ties)")
    assert np.all(np.isfinite(eg))
    assert not left


def test_explicit_stream_words(lib, ctx):
    """gd_set_stream with explicit words: a job fed the Philox words of seed B
    as a buffer makes exactly the moves and decisions of a job keyed by B,
    and launches past the buffer's last step are refused."""
    from geauxdock_cs_v2_amd import synth
    cx, _ = synth.load_config("c1")
    n = cx.size["n_rep"]
    steps = 40
    seed_a, seed_b = 101, 202
    reps = cx.replicas()
    words = np.array([[[lib.gd_stream_k(seed_b, 0, reps[r].idx_rep, s, k) for k in range(7)]
                       for s in range(1, steps + 1)] for r in range(n)], np.uint32)
    t = C.c_int()
    L.check(lib.gd_upload_target(ctx, C.byref(cx.view), C.byref(t)))
    jb, jx = C.c_int(), C.c_int()
    L.check(lib.gd_upload_job(ctx, t, C.byref(cx.view), seed_b, C.byref(jb)))
    L.check(lib.gd_upload_job(ctx, t, C.byref(cx.view), seed_a, C.byref(jx)))
    L.check(lib.gd_set_stream(ctx, jx, seed_a, words.ctypes.data_as(C.POINTER(C.c_uint32)), steps))
    say("case: explicit stream words vs keyed Philox, c1 plain 0..39")
    L.check(lib.gd_montecarlo(ctx, jb, 0, steps))
    L.check(lib.gd_montecarlo(ctx, jx, 0, steps))
    assert bytes(read_reps(lib, ctx, jb, n)) == bytes(read_reps(lib, ctx, jx, n))
    # step 40 is past the buffer (it covers steps 1..40: step 40 is the last)
    L.check(lib.gd_montecarlo(ctx, jx, steps, 1))
    assert lib.gd_montecarlo(ctx, jx, steps + 1, 1) == L.GD_ERR_ARG
    bad = words.copy()
    bad[0, 0, 0] = 0x80000000
    assert lib.gd_set_stream(ctx, jx, seed_a, bad.ctypes.data_as(C.POINTER(C.c_uint32)), steps) == L.GD_ERR_ARG
    L.check(lib.gd_free_job(ctx, jb))
    L.check(lib.gd_free_job(ctx, jx))
    L.check(lib.gd_free_target(ctx, t))


@pytest.mark.parametrize("radius", [18.0, 24.0, 28.0, 32.0, 48.0, 75.0, 160.0, 250.0])
def test_far_field_detached_ligand(lib, ctx, radius):
    """Ligands translated 18-250 A from the pocket (the reference's positive
    move drift carries them out of the protein): the far field inside the
    atoms' radius range comes into play, and from about 21 A on the very far
    field (LJ without a reciprocal), whose boundary the 24 and 28 A cases put
    inside the protein's points; beyond about 32 A from every point the LJ
    softening loses its e^2 term (75 A) and beyond about 109 A the softening
    itself (160, 250 A); every term stays within tolerance of FP64."""
    from geauxdock_cs_v2_amd import synth
    cx, _ = synth.load_config("c2")
    n_rep = cx.size["n_rep"]
    rng = np.random.default_rng(int(radius))
    n = 64
    idx = rng.choice(n_rep, n, replace=False).astype(np.int32)
    v = rng.normal(size=(n, 3))
    v = v / np.linalg.norm(v, axis=1, keepdims=True) * radius
    move = np.concatenate([v, rng.uniform(-math.pi, math.pi, (n, 3))], 1).astype(np.float32)
    say(f"case: detached ligand at {radius} A, c2, n={n}")
    t, j = upload(lib, ctx, cx)
    eg = gpu_energy(lib, ctx, j, idx, move)
    free(lib, ctx, t, j)
    e64 = O.energy(cx.view, idx, move, "f64")
    err = rel_err(eg, e64).max(0)
    say(f"detached {radius} A max rel err vs FP64:", np.array2string(err, precision=2, max_line_width=200))
    assert np.all(np.isfinite(eg))
    assert err.max() <= REL_TOL, err


def test_mclog_counters(lib, ctx, kat_complex11):
    """McLog counters: accepted MC moves per replica equal the records the
    plain chunks kept (minus the step-0 entry); accepted swaps per replica
    equal the rounds in which its temperature label changed."""
    cx = kat_complex11[1]
    n = cx.size["n_rep"]
    t, j = upload(lib, ctx, cx, 41)
    moves = np.zeros(n, np.int64)
    swaps = np.zeros(n, np.int64)
    say("case: McLog counters, kat11, 3 chunks of 50 steps with exchange")
    for c in range(3):
        L.check(lib.gd_montecarlo(ctx, j, 50 * c, 50))
        rec = (D.ReplicaMC * (n * 100))()
        ne = np.zeros(n, np.int32)
        L.check(lib.gd_read_records(ctx, j, rec, ne.ctypes.data_as(C.POINTER(C.c_int)), 100))
        moves += ne - (1 if c == 0 else 0)
        before = np.array([r.idx_tmp for r in read_reps(lib, ctx, j, n)])
        L.check(lib.gd_exchange(ctx, j, c, None))
        after = np.array([r.idx_tmp for r in read_reps(lib, ctx, j, n)])
        swaps += before != after
    mc = np.zeros(n, np.int32)
    ex = np.zeros(n, np.int32)
    L.check(lib.gd_read_mclog(ctx, j, mc.ctypes.data_as(C.POINTER(C.c_int)), ex.ctypes.data_as(C.POINTER(C.c_int))))
    free(lib, ctx, t, j)
    assert np.array_equal(mc, moves)
    assert np.array_equal(ex, swaps)
    assert swaps.sum() > 0


def test_bad_replica_layouts_are_refused(lib, ctx, kat_complex11):
    """Replica tables whose ladders do not hold each temperature once, whose
    replicas sit out of their (ligand, protein conformation) place, or whose
    idx_rep values (the stream keys and explicit-word rows) are not a
    permutation of the replica indices, are refused at upload and at write,
    before any launch reads them."""
    cx = kat_complex11[1]
    n = cx.size["n_rep"]
    n_lig, n_tmp = cx.size["n_lig"], cx.size["n_tmp"]
    t = C.c_int()
    L.check(lib.gd_upload_target(ctx, C.byref(cx.view), C.byref(t)))

    def broken(kind):
        reps = cx.replicas()
        if kind == "duplicate temperature":
            reps[n_lig].idx_tmp = reps[0].idx_tmp  # ladder (0, 0): temperatures 0, 0, 2, 3
        elif kind == "protein conformation out of place":
            reps[0].idx_prt = 1
        elif kind == "replica index out of range":
            reps[n - 1].idx_rep = n  # would index explicit stream words beyond the buffer
        elif kind == "negative replica index":
            reps[1].idx_rep = -1
        elif kind == "duplicate replica index":
            reps[1].idx_rep = reps[0].idx_rep  # two replicas would share one stream
        else:  # ligand conformation out of place
            reps[0].idx_lig = 1 % n_lig if n_lig > 1 else 0
            reps[0].idx_tmp = n_tmp  # out of range when there is one conformation
        return reps

    for kind in ("duplicate temperature", "protein conformation out of place", "ligand conformation out of place",
                 "replica index out of range", "negative replica index", "duplicate replica index"):
        say(f"case: upload and write refuse a replica table with a {kind}")
        reps = broken(kind)
        view = D.Complex.from_buffer_copy(cx.view)
        view.replica = C.cast(reps, C.POINTER(D.ReplicaMC))
        j = C.c_int(-1)
        assert lib.gd_upload_job(ctx, t, C.byref(view), 3, C.byref(j)) == L.GD_ERR_ARG
        jg = C.c_int()
        L.check(lib.gd_upload_job(ctx, t, C.byref(cx.view), 3, C.byref(jg)))
        assert lib.gd_write_replicas(ctx, jg, reps) == L.GD_ERR_ARG
        # the job still runs from its own table
        L.check(lib.gd_montecarlo(ctx, jg, 0, 2))
        g = read_reps(lib, ctx, jg, n)
        assert sorted(r.idx_tmp for r in g[:: n_lig][:n_tmp]) == list(range(n_tmp))
        L.check(lib.gd_free_job(ctx, jg))
    # a refused upload left no job behind: the target frees
    L.check(lib.gd_free_target(ctx, t))


def test_null_tables_are_refused(lib, ctx, kat_complex11):
    """A complex with a null table is refused with GD_ERR_ARG by the target
    and job uploads and by serving submission, instead of being read."""
    cx = kat_complex11[1]
    t = C.c_int()
    say("case: null protein, KDE and replica tables refused")
    view = D.Complex.from_buffer_copy(cx.view)
    view.prt_cdc = None
    assert lib.gd_upload_target(ctx, C.byref(view), C.byref(t)) == L.GD_ERR_ARG
    L.check(lib.gd_upload_target(ctx, C.byref(cx.view), C.byref(t)))
    j = C.c_int()
    for field in ("kde_x", "replica", "mcs_rowptr"):
        view = D.Complex.from_buffer_copy(cx.view)
        setattr(view, field, None)
        assert lib.gd_upload_job(ctx, t, C.byref(view), 1, C.byref(j)) == L.GD_ERR_ARG, field
        assert "null" in lib.gd_last_error().decode()
    opts = D.ServerOpts(1 << 34, 4, 50, 0)
    srv = C.c_void_p()
    L.check(lib.gd_server_create(ctx, t, C.byref(opts), C.byref(srv)))
    view = D.Complex.from_buffer_copy(cx.view)
    view.temp = None
    assert lib.gd_server_submit(srv, C.byref(view), 10, 1, C.byref(j)) == L.GD_ERR_ARG
    lib.gd_server_destroy(srv)
    L.check(lib.gd_free_target(ctx, t))


def test_mclog_after_autotuned_batch(lib, ctx):
    """The autotuner's trial launches leave the job as they found it: after an
    autotuned batched chunk the McLog accept counters equal the chunk's
    records (minus the step-0 entry).  Three temperatures give this job a
    shape key of its own, so the autotuner runs its trials here."""
    raw = L.load_job(os.path.join(ASTEX, "1a07C-1.pdb"), os.path.join(ASTEX, "1a07C1.sdf"),
                     os.path.join(ASTEX, "1a07C1-0.8.ff"), "1a07C1", n_tmp=3, temp_high=2.0, temp_low=0.3)
    cx = raw.stage()
    n = cx.size["n_rep"]
    t, j = upload(lib, ctx, cx, 43)
    say("case: McLog counters after an autotuned batched chunk, 1a07C1 x 3 temperatures, 50 steps")
    _batch(lib, ctx, [j], 50, records=1)
    rec = (D.ReplicaMC * (n * 100))()
    ne = np.zeros(n, np.int32)
    L.check(lib.gd_read_records(ctx, j, rec, ne.ctypes.data_as(C.POINTER(C.c_int)), 100))
    mc = np.zeros(n, np.int32)
    L.check(lib.gd_read_mclog(ctx, j, mc.ctypes.data_as(C.POINTER(C.c_int)), None))
    free(lib, ctx, t, j)
    assert (ne > 1).any()
    assert np.array_equal(mc, ne - 1)


def test_python_docker_known_answer(lib, kat_complex):
    """The Python Docker (the reference's Docker{complex, record, run()}) on
    the 1a07C job: one 100-step chunk whose first record entry of replica 0 is
    the known-answer step-0 energies, and the final replica states read back."""
    from geauxdock_cs_v2_amd.device import Device, Docker
    say("case: Python Docker on 1a07C")
    with Device(0) as dev:
        d = Docker(dev, seed=7)
        d.complex = kat_complex[1]
        d.run()
        n = kat_complex[1].size["n_rep"]
        first = d.record[0]
        assert first.step == 0
        assert np.all(np.abs(np.array(first.energy[:10]) - KAT_1A07C1) <= 5e-5), list(first.energy)
        assert len(d.replica) == n and all(1 <= d.next_entry[r] <= 100 for r in range(n))
        assert all(d.replica[r].step == d.record[r * 100 + d.next_entry[r] - 1].step for r in range(n))


def test_docking_driver_known_answer(lib, tmp_path):
    """The native driver (the reference's 3_geauxdock over gd::Docker) on a CSV
    job list holding the 1a07C job: its first PrintCsv line is the
    known-answer energies (util_print.cpp:351) and the record dump has one
    block per replica."""
    import os
    import subprocess
    from conftest import ASTEX
    from geauxdock_cs_v2_amd.lib import PARAM_DIR, PKG_DIR
    exe = os.path.join(PKG_DIR, "gd_dock")
    assert os.path.exists(exe), "gd_dock is built by __graft_entry__.build()"
    cols = ["magicnumber_a4a6b25c", "0", "1a07C", os.path.join(ASTEX, "1a07C-1.pdb"), "1a07C1",
            os.path.join(ASTEX, "1a07C1.sdf"), os.path.join(ASTEX, "1a07C1-0.8.ff"),
            os.path.join(PARAM_DIR, "08ff_opt"), os.path.join(PARAM_DIR, "paras"),
            os.path.join(PARAM_DIR, "08_nor_a"), os.path.join(PARAM_DIR, "08_nor_b"),
            "1", "0.044", "0.036", "0.02", "0.08", "1"] + ["-2"] * 40
    jobs = tmp_path / "jobs.csv"
    jobs.write_text("header\n" + ",".join(cols) + "\n")
    say("case: gd_dock on the 1a07C job list")
    r = subprocess.run([exe, str(jobs), "--out", str(tmp_path)], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = r.stdout.splitlines()
    assert out[0] == "rep step vdw ele pmf psp hdb hpc kde lhm dst total"
    f = out[1].split()
    assert f[0] == "0" and f[1] == "0"
    assert np.all(np.abs(np.array([float(x) for x in f[2:]]) - KAT_1A07C1) <= 1.5e-4), out[1]
    dump = (tmp_path / "out_1a07C1.txt").read_text().splitlines()
    assert sum(l.startswith("replica, ") for l in dump) == 51


def test_docking_driver_serve_mode(lib, ctx, tmp_path):
    """gd_dock --serve: the CSV job list as a request stream (the role of the
    reference's MPI job farm, src/main_mpi_noschedule/server.cpp): three job
    lines of the 1a07C job (1, 4 and 8 temperatures, 100-200 steps) served
    together.  Each output line and the last job's final-state dump equal the
    same requests served in this process (a served result depends on the
    request alone, not on the process or the mix), and the final states'
    energies match FP64 at the same poses."""
    import os
    import subprocess
    from conftest import ASTEX
    from geauxdock_cs_v2_amd import records
    from geauxdock_cs_v2_amd.lib import PARAM_DIR, PKG_DIR
    exe = os.path.join(PKG_DIR, "gd_dock")
    assert os.path.exists(exe), "gd_dock is built by __graft_entry__.build()"

    def line(n_tmp, t_hi, t_lo, n_dump):
        cols = ["magicnumber_a4a6b25c", "0", "1a07C", os.path.join(ASTEX, "1a07C-1.pdb"), "1a07C1",
                os.path.join(ASTEX, "1a07C1.sdf"), os.path.join(ASTEX, "1a07C1-0.8.ff"),
                os.path.join(PARAM_DIR, "08ff_opt"), os.path.join(PARAM_DIR, "paras"),
                os.path.join(PARAM_DIR, "08_nor_a"), os.path.join(PARAM_DIR, "08_nor_b"),
                str(n_tmp), str(t_hi), str(t_lo), "0.02", "0.08", str(n_dump)] + ["-2"] * 40
        return ",".join(cols)

    lines = [line(1, 0.044, 0.036, 1), line(4, 2.0, 0.3, 2), line(8, 2.0, 0.3, 1)]
    jobs = tmp_path / "jobs.csv"
    jobs.write_text("header\n" + "\n".join(lines) + "\n")
    say("case: gd_dock --serve on three 1a07C job lines")
    r = subprocess.run([exe, str(jobs), "--out", str(tmp_path), "--serve"], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = r.stdout.splitlines()
    assert out[0] == "rep step vdw ele pmf psp hdb hpc kde lhm dst total"
    assert len(out) == 1 + len(lines), out

    # the same requests served here
    cxs = [L.load_job_line(l).stage() for l in lines]
    t = C.c_int()
    L.check(lib.gd_upload_target(ctx, C.byref(cxs[0].view), C.byref(t)))
    srv = C.c_void_p()
    L.check(lib.gd_server_create(ctx, t.value, C.byref(D.ServerOpts(1 << 34, 256, 100, 100)), C.byref(srv)))
    ids = []
    for cx in cxs:
        i = C.c_int()
        L.check(lib.gd_server_submit(srv, C.byref(cx.view), cx.view.mcpara.steps_total, 2024, C.byref(i)))
        ids.append(i.value)
    L.check(lib.gd_server_drain(srv))
    reps = []
    for cx, i in zip(cxs, ids):
        res = D.RequestResult()
        a = (D.ReplicaMC * cx.size["n_rep"])()
        L.check(lib.gd_server_poll(srv, i, C.byref(res), a))
        assert res.status == 0
        reps.append(a)
    lib.gd_server_destroy(srv)
    L.check(lib.gd_free_target(ctx, t.value))

    for k, a in enumerate(reps):
        assert out[1 + k] == records.csv_line(0, a[0].step, a[0].energy), (k, out[1 + k])
    rows = [l for l in (tmp_path / "out_1a07C1.txt").read_text().splitlines() if l.startswith("   ")]
    last = reps[-1]
    assert len(rows) == len(last)
    for rr, row in enumerate(rows):
        want = "   %5d, %5d, %3d, " % (rr, last[rr].step, last[rr].idx_tmp) + "".join(
            "%8.4f, " % x for x in list(last[rr].movematrix)[:6] + list(last[rr].energy)[:10])
        assert row == want, (rr, row, want)
    c = cxs[-1]
    n = c.size["n_rep"]
    mm = np.array([list(x.movematrix)[:6] for x in last], np.float32)
    e64 = O.energy(c.view, np.arange(n, dtype=np.int32), mm, "f64")
    eg = np.array([list(x.energy)[:10] for x in last])
    assert rel_err(eg, e64).max() <= REL_TOL


def test_largest_ligand(lib, ctx, tmp_path):
    """A ligand of 512 atoms (the reference's MAXLIG, src/common/size.h:31) in
    2 conformations against a 1,500-point protein in 2 conformations, 4
    temperatures: its atoms span 8 waves, so replicas take 8 slot waves in
    every form.  Pose energies against FP64, then plain and batched steps
    against the oracle, step by step."""
    from geauxdock_cs_v2_amd import synth
    d = str(tmp_path)
    g = synth.generate(d, "big512", seed=31, n_atoms=512, n_lig_conf=2, lig_sigma=0.4, n_points=1500,
                       n_prt_conf=2, prt_sigma=0.4, prt_radius=20.0, n_kde=400, mcs_rows=8, mcs_cols=12,
                       lig_radius=9.0)
    raw = L.load_job(g["pdb"], g["sdf"], g["ff"], "big512", n_tmp=4, temp_high=2.0, temp_low=0.3)
    cx = raw.stage()
    n = cx.size["n_rep"]
    assert cx.size["lig_natom"] == 512 and n == 16
    rng = np.random.default_rng(5)
    idx = np.arange(n, dtype=np.int32)
    move = np.concatenate([rng.normal(0, 0.6, (n, 3)), rng.uniform(-math.pi, math.pi, (n, 3))], 1).astype(np.float32)
    say("case: 512-atom ligand, pose energies")
    seed = 12
    t, j = upload(lib, ctx, cx, seed)
    eg = gpu_energy(lib, ctx, j, idx, move)
    e64 = O.energy(cx.view, idx, move, "f64")
    err = rel_err(eg, e64).max(0)
    say("512 atoms max rel err vs FP64:", np.array2string(err, precision=2, max_line_width=200))
    assert np.all(np.isfinite(eg)) and err.max() <= REL_TOL, err
    flips = 0
    jb = C.c_int()  # a second job of the same target for the batched form
    L.check(lib.gd_upload_job(ctx, t, C.byref(cx.view), seed, C.byref(jb)))
    for form, job in (("plain", j), ("batched", jb.value)):
        for step in range(0, 6):
            say(f"case: 512-atom ligand, {form} step {step}")
            before = read_reps(lib, ctx, job, n)
            if form == "plain":
                L.check(lib.gd_montecarlo(ctx, job, step, 1))
            else:
                _batch(lib, ctx, [job], 1)
            g_ = read_reps(lib, ctx, job, n)
            o = (D.ReplicaMC * n)()
            C.memmove(o, before, C.sizeof(o))
            O.montecarlo(cx.view, o, step, 1, seed)
            flips += compare_step(lib, cx, seed, step, before, g_, o)
    L.check(lib.gd_free_job(ctx, jb.value))
    free(lib, ctx, t, j)
    # every flip was checked above to be a near-tie (accept uniform within the
    # energy tolerance of the threshold); the oracle's FP32 totals over 512
    # atoms widen that window, so a few of the 192 decisions may fall in it
    say(f"512 atoms: {flips} near-tie flips in {2 * 6 * n} decisions")
    assert flips <= 12


def test_c1_full_length_trajectory(lib, ctx):
    """The C1 workload at its full length (48 replicas, 2,000 steps, exchange
    every 100 steps, as Dock runs it, src/backend_gpu/backend_gpu.cu:91-107):
    every plain step against the oracle from the same state (compare_step),
    every exchange round's labels against the oracle's.  The positive-only
    move drift carries the ligands out of the protein within the run, so the
    Monte Carlo path passes through the near, far and very far fields."""
    from geauxdock_cs_v2_amd import synth
    cx, _ = synth.load_config("c1")
    n = cx.size["n_rep"]
    seed = 21
    t, j = upload(lib, ctx, cx, seed)
    flips = 0
    rnd = 0
    say("case: c1 full length, 2000 plain steps + exchange every 100")
    for step in range(0, 2000):
        before = read_reps(lib, ctx, j, n)
        L.check(lib.gd_montecarlo(ctx, j, step, 1))
        g = read_reps(lib, ctx, j, n)
        o = (D.ReplicaMC * n)()
        C.memmove(o, before, C.sizeof(o))
        O.montecarlo(cx.view, o, step, 1, seed)
        flips += compare_step(lib, cx, seed, step, before, g, o)
        if (step + 1) % 100 == 0:
            acc = C.c_int()
            L.check(lib.gd_exchange(ctx, j, rnd, C.byref(acc)))
            want = O.exchange(cx.view, g, rnd, seed)
            after = read_reps(lib, ctx, j, n)
            assert acc.value == want, (step, rnd)
            assert [after[r].idx_tmp for r in range(n)] == [g[r].idx_tmp for r in range(n)], (step, rnd)
            rnd += 1
            if rnd % 5 == 0:
                dist = np.array([np.linalg.norm(after[r].movematrix[:3]) for r in range(n)])
                say(f"  step {step + 1}: {flips} near-tie flips, ligand offsets {dist.min():.1f}-{dist.max():.1f} A")
    free(lib, ctx, t, j)
    say(f"c1 full length: {flips} near-tie flips in {2000 * n} decisions")
    assert flips <= 20


@pytest.mark.parametrize("which,chunks,offset", [("c1", 20, 0.0), ("kat11", 4, 0.0), ("c1", 3, 70.0),
                                                 ("c1", 3, 150.0)])
def test_batched_chunks_vs_oracle(lib, ctx, kat_complex11, which, chunks, offset):
    """The batched form as the benchmark runs it (autotuned shape, workgroup
    groups, hipGraph-replayed 100-step chunks with the exchange round inside)
    chunk by chunk against the oracle run from the same state: C1 over its
    full 2,000 steps (one protein conformation: one group), and 1a07C with 11
    protein conformations and 4 temperatures (2,244 replicas: two groups, each
    exchanging its own ladders) over 400 steps; C1 also from poses moved 70
    and 150 A along the drift axis, where the ligands are 44 and 124 A from
    every protein point (the lean very far path and its two shorter LJ
    softenings).  A replica whose 100 decisions all agree ends the chunk with
    the bit-identical move, step and temperature label and energies within
    tolerance of FP64; a near-tie decision sends a replica (and its exchange
    partners) on another trajectory, which only a few may do."""
    from geauxdock_cs_v2_amd import synth
    cx = kat_complex11[1] if which == "kat11" else synth.load_config("c1")[0]
    n = cx.size["n_rep"]
    seed = 33
    t, j = upload(lib, ctx, cx, seed)
    if offset:
        reps = read_reps(lib, ctx, j, n)
        for r in range(n):
            for k in range(3):
                reps[r].movematrix[k] += offset / math.sqrt(3.0)
        L.check(lib.gd_write_replicas(ctx, j, reps))
    diverged = 0
    say(f"case: {which} batched, {chunks} chunks of 100 steps with exchange, offset {offset} A, against the oracle")
    for k in range(chunks):
        before = read_reps(lib, ctx, j, n)
        _batch(lib, ctx, [j], 100, exchange=100, graph=1, comp=1)
        g = read_reps(lib, ctx, j, n)
        o = (D.ReplicaMC * n)()
        C.memmove(o, before, C.sizeof(o))
        O.montecarlo(cx.view, o, 100 * k, 100, seed)
        O.exchange(cx.view, o, k, seed)
        same = [list(g[r].movematrix) == list(o[r].movematrix) and g[r].step == o[r].step
                and g[r].idx_tmp == o[r].idx_tmp for r in range(n)]
        diverged += n - sum(same)
        idx = np.array([r for r in range(n) if same[r]], np.int32)
        mm = np.array([list(g[r].movematrix)[:6] for r in idx], np.float32)
        e64 = O.energy(cx.view, idx, mm, "f64")
        eg = np.array([list(g[r].energy)[:10] for r in idx])
        assert rel_err(eg, e64).max() <= REL_TOL, k
        # the next chunk starts from the GPU's state (the oracle is resynchronised)
    free(lib, ctx, t, j)
    say(f"{which} batched: {diverged} replica-chunks of {chunks * n} left the oracle's trajectory")
    assert diverged <= max(12, chunks * n // 100)


@pytest.mark.parametrize("n_atoms,n_points", [(2, 7), (3, 33)])
def test_tiny_inputs(lib, ctx, tmp_path, n_atoms, n_points):
    """Ragged small inputs: 2- and 3-atom ligands against 7- and 33-point
    proteins (fewer points than point slices, one-point segments, groups of
    four with a remainder, atom types without KDE points), 4 temperatures:
    pose energies against FP64, then plain and batched steps against the
    oracle, step by step."""
    from geauxdock_cs_v2_amd import synth
    g = synth.generate(str(tmp_path), f"t{n_atoms}", seed=3, n_atoms=n_atoms, n_lig_conf=2, lig_sigma=0.4,
                       n_points=n_points, n_prt_conf=2, prt_sigma=0.4, prt_radius=12.0, n_kde=4, mcs_rows=2,
                       mcs_cols=12)
    raw = L.load_job(g["pdb"], g["sdf"], g["ff"], f"t{n_atoms}", n_tmp=4, temp_high=2.0, temp_low=0.3)
    cx = raw.stage()
    n = cx.size["n_rep"]
    assert cx.size["prt_npoint"] == n_points and cx.size["lig_natom"] == n_atoms
    rng = np.random.default_rng(9)
    idx = np.arange(n, dtype=np.int32)
    move = np.concatenate([rng.normal(0, 2.0, (n, 3)), rng.uniform(-math.pi, math.pi, (n, 3))], 1).astype(np.float32)
    say(f"case: tiny inputs {n_atoms} atoms x {n_points} points, pose energies")
    seed = 14
    t, j = upload(lib, ctx, cx, seed)
    eg = gpu_energy(lib, ctx, j, idx, move)
    e64 = O.energy(cx.view, idx, move, "f64")
    err = rel_err(eg, e64).max(0)
    assert np.all(np.isfinite(eg)) and err.max() <= REL_TOL, err
    jb = C.c_int()
    L.check(lib.gd_upload_job(ctx, t, C.byref(cx.view), seed, C.byref(jb)))
    flips = 0
    for form, job in (("plain", j), ("batched", jb.value)):
        for step in range(0, 6):
            say(f"case: tiny inputs {n_atoms} x {n_points}, {form} step {step}")
            before = read_reps(lib, ctx, job, n)
            if form == "plain":
                L.check(lib.gd_montecarlo(ctx, job, step, 1))
            else:
                _batch(lib, ctx, [job], 1)
            g_ = read_reps(lib, ctx, job, n)
            o = (D.ReplicaMC * n)()
            C.memmove(o, before, C.sizeof(o))
            O.montecarlo(cx.view, o, step, 1, seed)
            flips += compare_step(lib, cx, seed, step, before, g_, o)
    L.check(lib.gd_free_job(ctx, jb.value))
    free(lib, ctx, t, j)
    assert flips <= 2
