"""GPU tests of the serving loop (continuous batching, HBM-footprint
admission): a request's result is the same alone and in a mix, requests that
do not fit wait for room, cancelled and malformed requests fail alone."""
import ctypes as C
import sys

import numpy as np
import pytest

from conftest import REL_TOL, rel_err
from geauxdock_cs_v2_amd import ctypes_defs as D
from geauxdock_cs_v2_amd import lib as L
from geauxdock_cs_v2_amd import synth
from oracle import bind as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def setup(lib, tmp_path_factory):
    d = str(tmp_path_factory.mktemp("serve"))
    base = synth.generate(d, "sA", seed=1, n_atoms=20, n_lig_conf=3, lig_sigma=0.4, n_points=800, n_prt_conf=2,
                          prt_sigma=0.4, prt_radius=18.0, n_kde=60, mcs_rows=4, mcs_cols=8)
    jobs = {"A": base}
    for name, seed, na, nc in [("B", 2, 37, 5), ("C", 3, 11, 2), ("D", 4, 70, 2)]:
        jobs[name] = synth.generate(d, "s" + name, seed=seed, n_atoms=na, n_lig_conf=nc, lig_sigma=0.4,
                                    n_points=800, n_prt_conf=2, prt_sigma=0.4, prt_radius=18.0, n_kde=60,
                                    mcs_rows=4, mcs_cols=8, protein=base["protein"])
    cx = {}
    for k, v in jobs.items():
        raw = L.load_job(v["pdb"], v["sdf"], v["ff"], "s" + k, n_tmp=4, temp_high=2.0, temp_low=0.3)
        cx[k] = raw.stage()
    h = C.c_void_p()
    L.check(lib.gd_create(0, C.byref(h)))
    t = C.c_int()
    L.check(lib.gd_upload_target(h, C.byref(cx["A"].view), C.byref(t)))
    yield lib, h, t.value, cx
    lib.gd_free_target(h, t.value)
    lib.gd_destroy(h)


def serve(lib, h, t, reqs, budget=1 << 34, max_active=16, chunk=50, exchange=50):
    opts = D.ServerOpts(budget, max_active, chunk, exchange)
    srv = C.c_void_p()
    L.check(lib.gd_server_create(h, t, C.byref(opts), C.byref(srv)))
    ids = []
    for cx, steps, seed in reqs:
        i = C.c_int()
        L.check(lib.gd_server_submit(srv, C.byref(cx.view), steps, seed, C.byref(i)))
        ids.append(i.value)
    L.check(lib.gd_server_drain(srv))
    out = []
    for (cx, _, _), i in zip(reqs, ids):
        res = D.RequestResult()
        reps = (D.ReplicaMC * cx.size["n_rep"])()
        L.check(lib.gd_server_poll(srv, i, C.byref(res), reps))
        out.append((res.status, bytes(reps), res, reps))
    lib.gd_server_destroy(srv)
    return out


def test_alone_and_in_the_mix_are_identical(setup):
    lib, h, t, cx = setup
    say("case: serving A alone")
    alone = serve(lib, h, t, [(cx["A"], 300, 11)])
    say("case: serving A mixed with B, C, D")
    mixed = serve(lib, h, t, [(cx["B"], 200, 12), (cx["A"], 300, 11), (cx["C"], 450, 13), (cx["D"], 100, 14)])
    assert alone[0][0] == 0 and all(m[0] == 0 for m in mixed)
    assert alone[0][1] == mixed[1][1]


def test_served_energies_match_fp64(setup):
    lib, h, t, cx = setup
    say("case: serving B, energies vs FP64")
    (st, _, _, reps), = serve(lib, h, t, [(cx["B"], 120, 5)])
    assert st == 0
    c = cx["B"]
    n = c.size["n_rep"]
    idx = np.arange(n, dtype=np.int32)
    mm = np.array([list(r.movematrix)[:6] for r in reps], np.float32)
    e64 = O.energy(c.view, idx, mm, "f64")
    eg = np.array([list(r.energy)[:10] for r in reps])
    assert rel_err(eg, e64).max() <= REL_TOL
    assert all(0 <= r.step < 120 for r in reps)


def test_admission_by_footprint(setup):
    """A budget that holds one request at a time serialises the requests
    (each admitted after the previous finished) without changing results."""
    lib, h, t, cx = setup
    fp = lib.gd_job_footprint(cx["B"].size["lig_natom"], cx["B"].size["n_lig"], cx["B"].size["kde_npoint"],
                              int(cx["B"].array("mcs_rowptr")[-1]), cx["B"].size["mcs_nrow"],
                              cx["B"].size["n_tmp"], cx["B"].size["n_rep"], 0)
    say("case: serving with a one-job budget")
    tight = serve(lib, h, t, [(cx["B"], 100, 1), (cx["B"], 100, 2), (cx["B"], 100, 3)], budget=int(fp * 1.5))
    loose = serve(lib, h, t, [(cx["B"], 100, 1), (cx["B"], 100, 2), (cx["B"], 100, 3)])
    assert [x[1] for x in tight] == [x[1] for x in loose]
    r = [x[2] for x in tight]
    assert r[1].t_admit_ms >= r[0].t_done_ms - 1e-6 and r[2].t_admit_ms >= r[1].t_done_ms - 1e-6


def test_cancel_and_bad_requests_fail_alone(setup):
    lib, h, t, cx = setup
    opts = D.ServerOpts(1 << 34, 1, 50, 0)
    srv = C.c_void_p()
    L.check(lib.gd_server_create(h, t, C.byref(opts), C.byref(srv)))
    a, b = C.c_int(), C.c_int()
    say("case: serving cancel + rejected request")
    L.check(lib.gd_server_submit(srv, C.byref(cx["A"].view), 200, 1, C.byref(a)))
    L.check(lib.gd_server_submit(srv, C.byref(cx["C"].view), 200, 2, C.byref(b)))
    L.check(lib.gd_server_cancel(srv, b.value))
    # a request whose replica table breaks the ladder layout passes the size
    # checks at submission and fails alone at admission, with its reason
    reps = cx["B"].replicas()
    n_lig_b = cx["B"].size["n_lig"]
    reps[n_lig_b].idx_tmp = reps[0].idx_tmp  # ladder (0, 0) holds temperature 0 twice
    bad_view = D.Complex.from_buffer_copy(cx["B"].view)
    bad_view.replica = C.cast(reps, C.POINTER(D.ReplicaMC))
    m = C.c_int()
    L.check(lib.gd_server_submit(srv, C.byref(bad_view), 100, 3, C.byref(m)))
    # a request larger than the whole budget is refused at submission
    small = D.ServerOpts(1024, 1, 50, 0)
    srv2 = C.c_void_p()
    L.check(lib.gd_server_create(h, t, C.byref(small), C.byref(srv2)))
    c = C.c_int()
    assert lib.gd_server_submit(srv2, C.byref(cx["A"].view), 10, 1, C.byref(c)) == L.GD_ERR_NOMEM
    lib.gd_server_destroy(srv2)
    L.check(lib.gd_server_drain(srv))
    ra, rb, rm = D.RequestResult(), D.RequestResult(), D.RequestResult()
    L.check(lib.gd_server_poll(srv, a.value, C.byref(ra), None))
    L.check(lib.gd_server_poll(srv, b.value, C.byref(rb), None))
    why_b = lib.gd_last_error().decode()
    L.check(lib.gd_server_poll(srv, m.value, C.byref(rm), None))
    why_m = lib.gd_last_error().decode()
    lib.gd_server_destroy(srv)
    assert ra.status == 0
    assert rb.status == L.GD_ERR_STATE and "cancelled" in why_b, why_b
    assert rm.status == L.GD_ERR_ARG and "permutation" in why_m, why_m


def test_exchange_rounds_fall_due_whatever_the_chunk(setup):
    """A chunk ends at the request's next exchange boundary: with exchanges
    every 50 steps, chunks of 30 steps (cut at 30, 50, 80, 100, ...) give
    the same final replicas, temperature labels included, as chunks of 50;
    with no exchange the labels stay where they started."""
    lib, h, t, cx = setup
    say("case: serving chunk 30 / exchange 50 against chunk 50 / exchange 50")
    a = serve(lib, h, t, [(cx["B"], 230, 21)], chunk=30, exchange=50)[0]
    b = serve(lib, h, t, [(cx["B"], 230, 21)], chunk=50, exchange=50)[0]
    c = serve(lib, h, t, [(cx["B"], 230, 21)], chunk=30, exchange=0)[0]
    assert a[0] == b[0] == c[0] == 0
    assert a[1] == b[1]
    start = cx["B"].replicas()
    labels = lambda reps: [r.idx_tmp for r in reps]
    assert labels(c[3]) == labels(start)
    assert labels(a[3]) != labels(start)  # four rounds at 50, 100, 150, 200 moved some


def test_serving_workload_size_requests_match_fp64(lib, tmp_path_factory):
    """Two requests of the serving workload's largest kind served together:
    an 8,000-point protein x 4 conformations (35 A ball), ligands of 105
    atoms x 8 conformations and 119 atoms x 32 conformations, 8 temperatures
    (256 and 1,024 replicas), 300 and 200 steps with exchanges every 100.
    The final energies of 64 replicas of each (every 4th, resp. 16th) are
    compared with FP64 at their final moves, pocket-specific step ties
    resolved as in the parity tests."""
    from test_gpu_parity import psp_step_ties
    d = str(tmp_path_factory.mktemp("serve_big"))
    P = synth.SERVE
    prot = synth._protein_only(d, "big", 77, P["n_points"], P["n_prt_conf"], P["prt_sigma"], P["prt_radius"])
    specs = [("bigA", 105, 8, 300), ("bigB", 119, 32, 200)]
    cxs = []
    for k, (name, na, nc, _) in enumerate(specs):
        f = synth._ligand_files(d, name, 300 + k, na, nc, P["lig_sigma"], prot)
        raw = L.load_job(f["pdb"], f["sdf"], f["ff"], name, n_tmp=P["n_tmp"], temp_high=P["t_hi"],
                         temp_low=P["t_lo"])
        cxs.append(raw.stage())
    assert [c.size["n_rep"] for c in cxs] == [256, 1024]
    assert cxs[0].size["prt_npoint"] >= 8000 - 100 and cxs[0].size["n_prt"] == 4
    h = C.c_void_p()
    L.check(lib.gd_create(0, C.byref(h)))
    t = C.c_int()
    L.check(lib.gd_upload_target(h, C.byref(cxs[0].view), C.byref(t)))
    try:
        say(f"case: serving two requests of {[c.size['lig_natom'] for c in cxs]} atoms against "
            f"{cxs[0].size['prt_npoint']} points x 4")
        out = serve(lib, h, t.value, [(c, s, 31 + k) for k, (c, (_, _, _, s)) in enumerate(zip(cxs, specs))],
                    chunk=100, exchange=100)
    finally:
        lib.gd_free_target(h, t.value)
        lib.gd_destroy(h)
    for c, (st, _, res, reps), (name, _, _, steps) in zip(cxs, out, specs):
        assert st == 0
        n = c.size["n_rep"]
        idx = np.arange(0, n, n // 64, dtype=np.int32)
        mm = np.array([list(reps[i].movematrix)[:6] for i in idx], np.float32)
        eg = np.array([list(reps[i].energy)[:10] for i in idx])
        e64 = O.energy(c.view, idx, mm, "f64")
        bad, left = psp_step_ties(c, idx, mm, eg, e64)
        err = rel_err(eg, e64)
        say(f"{name}: {len(idx)} replicas, max rel err vs FP64 {err.max():.2e} "
            f"({len(bad)} pocket-specific step ties, {len(left)} unexplained)")
        assert not left
        assert all(0 <= reps[i].step < steps for i in idx)
