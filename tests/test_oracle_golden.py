"""CPU tests: the oracle pinned against the reference's golden data, the
reference's own compiled code (oracle/_ref, where it was built), closed forms
and brute force.  No GPU calls."""
import ctypes as C
import math
import os

import numpy as np
import pytest

from conftest import ASTEX, GOLDEN, KAT_1A07C1
from geauxdock_cs_v2_amd import ctypes_defs as D
from geauxdock_cs_v2_amd import lib as L
from oracle import bind as O


def _philox_kat():
    rows = []
    with open(os.path.join(GOLDEN, "philox4x32_10_kat.txt")) as f:
        for line in f:
            tok = line.split()
            vals = [int(t, 16) for t in tok[2:]]
            rows.append((vals[0:4], vals[4:6], vals[6:10]))
    return rows


@pytest.mark.parametrize("impl", ["library", "oracle"])
def test_philox_known_answers(lib, impl):
    u32 = C.c_uint32
    for ctr, key, want in _philox_kat():
        c = (u32 * 4)(*ctr)
        k = (u32 * 2)(*key)
        out = (u32 * 4)()
        if impl == "library":
            lib.gd_philox4x32_10(c, k, out)
        else:
            O.oracle().gdo_philox4x32_10(c, k, out)
        assert list(out) == want


def test_stream_definition(lib):
    """k is word j&3 of the block at counter {s, r, kind, j>>2}, shifted right
    by one; library and oracle agree; the reference's (float)k / RAND_MAX is
    exactly k rounded to float times 2^-31."""
    o = O.oracle()
    u32 = C.c_uint32
    for seed in (0, 1, 2**40 + 7):
        for r, s, j, kind in [(0, 1, 0, 0), (5, 9, 6, 0), (3, 100, 3, 1), (1000, 7, 5, 0)]:
            k = lib.gd_stream_k(seed, kind, r, s, j)
            assert k == o.gdo_stream_k(seed, kind, r, s, j)
            c = (u32 * 4)(s, r, kind, j >> 2)
            key = (u32 * 2)(seed & 0xFFFFFFFF, seed >> 32)
            out = (u32 * 4)()
            lib.gd_philox4x32_10(c, key, out)
            assert k == out[j & 3] >> 1
    ks = np.array([0, 1, 2**24 - 1, 2**24 + 1, 2**30 + 3, 2**31 - 65, 2**31 - 64, 2**31 - 1], np.int64)
    a = ks.astype(np.float32) / np.float32(2147483647)
    b = ks.astype(np.float32) * np.float32(2.0 ** -31)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("pdb", ["1a07C-1.pdb", "1a07C.pdb"])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_kat_step0_energies(lib, pdb, prec):
    raw = L.load_job(os.path.join(ASTEX, pdb), os.path.join(ASTEX, "1a07C1.sdf"),
                     os.path.join(ASTEX, "1a07C1-0.8.ff"), "1a07C1")
    cx = raw.stage()
    e = O.energy(cx.view, [0], np.zeros(6, np.float32), prec)[0]
    assert np.all(np.abs(e - KAT_1A07C1) <= 5e-5 + 1e-12), (e, KAT_1A07C1)


def test_loader_sizes_test5(kat_complex11):
    """reference src/main_unittest/test5_loadall.cpp:19-20: 11 protein
    conformations, 51 ligand conformations, 344 points, 34 atoms, 454 KDE
    points, 11 MCS rows (one temperature there; four here)."""
    _, cx = kat_complex11
    s = cx.size
    assert (s["n_prt"], s["n_lig"], s["prt_npoint"], s["lig_natom"], s["kde_npoint"], s["mcs_nrow"]) == (
        11, 51, 344, 34, 454, 11)
    assert s["n_rep"] == 11 * 51 * 4


def _a2_rows():
    """Every row of the reference's size table; each row's ligand files are
    fixtures, so a missing file fails its row instead of dropping it."""
    rows = []
    with open(os.path.join(GOLDEN, "a2_sizes.csv")) as f:
        for line in f:
            t = line.strip().split(",")
            if t and t[0]:
                rows.append((t[0], [int(x) for x in t[1:]]))
    return rows


def test_a2_table_is_complete():
    """all 18 ligand rows of reference src/main_mpi/a2.csv:2-19"""
    assert len(_a2_rows()) == 18


@pytest.mark.parametrize("lig,want", _a2_rows())
def test_loader_sizes_a2(lib, lig, want):
    """reference src/main_mpi/a2.csv:2-19 against 1a07C.pdb, one temperature."""
    for ext in (".sdf", "-0.8.ff"):
        assert os.path.exists(os.path.join(ASTEX, lig + ext)), f"fixture {lig}{ext} missing"
    raw = L.load_job(os.path.join(ASTEX, "1a07C.pdb"), os.path.join(ASTEX, lig + ".sdf"),
                     os.path.join(ASTEX, lig + "-0.8.ff"), lig)
    s = raw.stage().size
    got = [s["n_prt"], s["n_lig"], s["n_tmp"], s["n_rep"], s["prt_npoint"], s["lig_natom"], s["kde_npoint"],
           s["mcs_nrow"]]
    assert got == want


def test_job_line_parser(lib, tmp_path):
    """the reference CSV job line (src/main/1_mklist.py:97-135 column order)"""
    from geauxdock_cs_v2_amd.lib import PARAM_DIR
    cols = ["magicnumber_a4a6b25c", "0", "1a07C", os.path.join(ASTEX, "1a07C-1.pdb"), "1a07C1",
            os.path.join(ASTEX, "1a07C1.sdf"), os.path.join(ASTEX, "1a07C1-0.8.ff"),
            os.path.join(PARAM_DIR, "08ff_opt"), os.path.join(PARAM_DIR, "paras"),
            os.path.join(PARAM_DIR, "08_nor_a"), os.path.join(PARAM_DIR, "08_nor_b"),
            "1", "0.044", "0.036", "0.02", "0.08", "1"] + ["-2"] * 40
    raw = L.load_job_line(",".join(cols))
    cx = raw.stage()
    assert cx.size["n_rep"] == 51
    assert abs(cx.view.mcpara.move_scale[0] - 0.02) < 1e-7 and abs(cx.view.mcpara.move_scale[3] - 0.08) < 1e-7
    assert cx.view.mcpara.steps_total == 100
    e = O.energy(cx.view, [0], np.zeros(6, np.float32))[0]
    assert np.all(np.abs(e - KAT_1A07C1) <= 5e-5)


def test_loader_errors_are_reported(lib, tmp_path):
    with pytest.raises(L.GDError) as ei:
        L.load_job(os.path.join(ASTEX, "1a07C-1.pdb"), str(tmp_path / "missing.sdf"),
                   os.path.join(ASTEX, "1a07C1-0.8.ff"), "1a07C1")
    assert ei.value.code == L.GD_ERR_IO
    with pytest.raises(L.GDError):  # MCS rows must carry the ligand id
        L.load_job(os.path.join(ASTEX, "1a07C-1.pdb"), os.path.join(ASTEX, "1a07C1.sdf"),
                   os.path.join(ASTEX, "1a07C1-0.8.ff"), "not_this_id")


# ---------------------------------------------------------------------------
# the reference's own code, compiled from its sources
# ---------------------------------------------------------------------------
needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (no reference tree)")


def _fields():
    return ["lig_x2", "lig_y2", "lig_z2", "lig_c2", "lig_t2", "lig_x3", "lig_y3", "lig_z3", "lig_c3", "lig_t3",
            "lig_kde_begin_idx", "lig_kde_end_idx", "prt_x", "prt_y", "prt_z", "prt_hpp", "prt_ele", "prt_t",
            "prt_c", "prt_cdc", "prt_seq3r", "kde_t", "kde_x", "kde_y", "kde_z", "mcs_ncol", "mcs_tcc",
            "mcs_rowptr", "mcs_i", "mcs_x", "mcs_y", "mcs_z", "psp"]


def _ref_view_arrays(ref_view):
    holder = L.StagedComplex.__new__(L.StagedComplex)
    holder.view = ref_view
    holder._h = None
    return holder


@needs_ref
@pytest.mark.parametrize("which", ["kat", "c1"])
def test_staging_matches_reference_builders(lib, kat_complex11, which):
    from geauxdock_cs_v2_amd import synth
    if which == "kat":  # one temperature keeps 561 replicas within the reference's MAX_REP
        raw = L.load_job(os.path.join(ASTEX, "1a07C.pdb"), os.path.join(ASTEX, "1a07C1.sdf"),
                         os.path.join(ASTEX, "1a07C1-0.8.ff"), "1a07C1")
        cx = raw.stage()
    else:
        info = synth.config_files("c1")
        c = info["cfg"]
        raw = L.load_job(info["pdb"], info["sdf"], info["ff"], info["lig_id"], n_tmp=c["n_tmp"],
                         temp_high=c["t_hi"], temp_low=c["t_lo"])
        cx = raw.stage()
    temps = [cx.view.temp[i] for i in range(cx.view.size.n_tmp)]
    ref = O.RefComplex(raw.view, temps)
    rv = _ref_view_arrays(ref.view)
    for name in _fields():
        a, b = cx.array(name), rv.array(name)
        assert a.dtype == b.dtype and np.array_equal(a.view(np.uint32) if a.dtype == np.float32 else a,
                                                     b.view(np.uint32) if b.dtype == np.float32 else b), name
    assert bytes(cx.view.enepara) == bytes(ref.view.enepara)


@needs_ref
@pytest.mark.parametrize("which", ["kat", "c1"])
def test_oracle_montecarlo_bit_exact_with_reference(lib, kat_complex11, which):
    """Two 100-step chunks of every replica: final states and accept records
    bit-identical to the reference's MonteCarlo_d fed the same stream."""
    from geauxdock_cs_v2_amd import synth
    if which == "kat":  # one temperature keeps 561 replicas within the reference's MAX_REP
        raw = L.load_job(os.path.join(ASTEX, "1a07C.pdb"), os.path.join(ASTEX, "1a07C1.sdf"),
                         os.path.join(ASTEX, "1a07C1-0.8.ff"), "1a07C1")
        cx = raw.stage()
    else:
        info = synth.config_files("c1")
        c = info["cfg"]
        raw = L.load_job(info["pdb"], info["sdf"], info["ff"], info["lig_id"], n_tmp=c["n_tmp"],
                         temp_high=c["t_hi"], temp_low=c["t_lo"])
        cx = raw.stage()
    temps = [cx.view.temp[i] for i in range(cx.view.size.n_tmp)]
    ref = O.RefComplex(raw.view, temps)
    seed = 977
    ra, rb = cx.replicas(), cx.replicas()
    for s1 in (0, 100):
        rec_a, na = O.montecarlo(cx.view, ra, s1, 100, seed, cap=100)
        rec_b, nb = ref.montecarlo(rb, s1, 100, seed)
        assert bytes(ra) == bytes(rb)
        assert np.array_equal(na, nb)
        assert bytes(rec_a) == bytes(rec_b)


# ---------------------------------------------------------------------------
# closed forms and invariances
# ---------------------------------------------------------------------------
def test_step0_distance_term_is_b8(kat_complex):
    """at zero move the ligand centre is the pocket centre, so
    e_dst = a_8 * 0 + b_8 = b_8 = -1 (data/parameters/08_nor_b line 9)"""
    _, cx = kat_complex
    e = O.energy(cx.view, [0], np.zeros(6, np.float32), "f64")[0]
    assert e[8] == pytest.approx(cx.view.enepara.b_para[8], abs=0)
    t = np.array([1.5, -2.0, 0.5, 0, 0, 0], np.float32)
    e = O.energy(cx.view, [0], t, "f64")[0]
    want = cx.view.enepara.a_para[8] * math.sqrt(float(t[0]) ** 2 + float(t[1]) ** 2 + float(t[2]) ** 2) \
        + cx.view.enepara.b_para[8]
    assert e[8] == pytest.approx(want, rel=1e-6)


def test_full_turn_rotation_is_identity(kat_complex):
    """cumulative Euler angles of 2*pi give the same pose as zero angles"""
    _, cx = kat_complex
    m0 = np.zeros(6, np.float32)
    m1 = np.array([0, 0, 0, 2 * math.pi, -2 * math.pi, 2 * math.pi], np.float32)
    e0 = O.energy(cx.view, [0], m0, "f64")[0]
    e1 = O.energy(cx.view, [0], m1, "f64")[0]
    assert np.allclose(e0, e1, rtol=1e-5, atol=1e-6)


def _tiny_complex(rng, n_l=2, n_p=3, shift=(0.0, 0.0, 0.0), perm=None):
    """A hand-built staged complex: n_l ligand atoms, n_p protein points."""
    shift = np.asarray(shift, np.float64)
    ep0 = D.EnePara0()
    src = L.load_job(os.path.join(ASTEX, "1a07C-1.pdb"), os.path.join(ASTEX, "1a07C1.sdf"),
                     os.path.join(ASTEX, "1a07C1-0.8.ff"), "1a07C1")
    C.memmove(C.byref(ep0), C.byref(src.view.enepara0), C.sizeof(ep0))
    keep = {}

    def arr(name, data, ct):
        a = np.ascontiguousarray(data, dtype=np.float32 if ct is C.c_float else np.int32)
        keep[name] = a
        return a.ctypes.data_as(C.POINTER(ct))

    center = np.array([10.0, -3.0, 5.0]) + shift
    lig = rng.normal(0, 1.2, (n_l, 3))
    lt = rng.integers(0, 24, n_l)
    lc = rng.normal(0, 0.3, n_l)
    pts = center + rng.normal(0, 4.0, (n_p, 3))
    pt = rng.integers(0, 30, n_p)
    pc = rng.integers(0, 4, n_p)
    pd = rng.integers(0, 20, n_p)
    pv = rng.normal(0, 1, n_p)
    if perm is not None:
        pts, pt, pc, pd, pv = pts[perm], pt[perm], pc[perm], pd[perm], pv[perm]
    kde = center + rng.normal(0, 2.0, (2, 3))
    kt = np.array([lt[0], lt[-1]])
    mx = center + rng.normal(0, 1.0, (1, 3))
    raw = D.Raw()
    raw.n_lig, raw.lig_natom = 1, n_l
    raw.lig_x, raw.lig_y, raw.lig_z = (arr(f"l{i}", lig[:, i], C.c_float) for i in range(3))
    raw.lig_t = arr("lt", lt, C.c_int)
    raw.lig_c = arr("lc", lc, C.c_float)
    for i in range(3):
        raw.lig_center[i] = center[i]
    raw.n_prt, raw.prt_npoint, raw.prt_nres = 1, n_p, n_p
    raw.prt_x, raw.prt_y, raw.prt_z = (arr(f"p{i}", pts[:, i], C.c_float) for i in range(3))
    raw.prt_t = arr("pt", pt, C.c_int)
    raw.prt_c = arr("pc", pc, C.c_int)
    raw.prt_d = arr("pd", pd, C.c_int)
    raw.prt_r = arr("pr", np.arange(n_p), C.c_int)
    raw.prt_seq3 = arr("ps", np.arange(n_p) + 1, C.c_int)
    raw.psp_n = n_p
    raw.psp_res = arr("sr", np.arange(n_p) + 1, C.c_int)
    raw.psp_lig = arr("sl", np.full(n_p, lt[0]), C.c_int)
    raw.psp_val = arr("sv", pv, C.c_float)
    raw.kde_npoint = 2
    raw.kde_t = arr("kt", kt, C.c_int)
    raw.kde_x, raw.kde_y, raw.kde_z = (arr(f"k{i}", kde[:, i], C.c_float) for i in range(3))
    raw.mcs_nrow = 1
    raw.mcs_tcc = arr("mt", [0.7], C.c_float)
    raw.mcs_rowptr = arr("mr", [0, 1], C.c_int)
    raw.mcs_i = arr("mi", [n_l - 1], C.c_int)
    raw.mcs_x, raw.mcs_y, raw.mcs_z = (arr(f"m{i}", mx[:, i], C.c_float) for i in range(3))
    raw.enepara0 = ep0
    raw.n_tmp = 1
    raw.temp_high = raw.temp_low = 1.0
    raw.steps_n_dump = 1
    h = C.c_void_p()
    L.check(L.load().gd_stage(C.byref(raw), C.byref(h)))
    cx = L.StagedComplex(h, L.load())
    return cx, dict(lig=lig, lt=lt, lc=lc, pts=pts, pt=pt, pc=pc, pd=pd, kde=kde, kt=kt, mx=mx, center=center,
                    ep0=ep0, keep=keep, psp_val=keep["sv"], raw=raw)


def _brute_force(cx, d):
    """Energy of the tiny complex at zero move, straight from the raw
    parameter file tables (EnePara0 indexed [point type][ligand type]) in
    FP64, one pair at a time."""
    ep0 = d["ep0"]
    # float32 inputs, as staged, combined in float64
    lig = np.float32(d["lig"]).astype(np.float64) + np.float32(d["center"]).astype(np.float64)
    e = np.zeros(10)
    lj, el = ep0.lj, ep0.el
    hpc_atom = []
    for i in range(len(lig)):
        lt, lc = int(d["lt"][i]), float(d["lc"][i])
        h = 0.0
        for p in range(len(d["pts"])):
            pt, pc, pd = int(d["pt"][p]), int(d["pc"][p]), int(d["pd"][p])
            r = float(np.linalg.norm(lig[i] - np.float32(d["pts"][p]).astype(np.float64)))
            r0, eps = ep0.vdw[pt][lt][0], ep0.vdw[pt][lt][1]
            p1 = 2 * eps * (r0 * lj[2]) ** 9 / r ** 9
            p2 = 3 * eps * (r0 * lj[2]) ** 6 / r ** 6
            e[0] += (p1 - p2) / (p1 * lj[0] * (1 + lj[1] * r * r) + 1)
            s = el[1] * r
            g = el[0] + (4 - 3 * el[0]) * s * s + (2 * el[0] - 3) * s ** 3 if s < 1 else 1 / s
            ele = ep0.ele[pd + 30] if pt == 0 else ep0.ele[pt]
            e[1] += lc * ele * g
            pmf0, pmf1 = ep0.pmf[pt][lt][0], ep0.pmf[pt][lt][1]
            e[2] += pmf1 / (1 + math.exp((-0.5 * r + 6) * (r - pmf0)))
            if pc == 2 and r - pmf0 <= 0 and lt == int(d["lt"][0]):
                e[3] += float(d["psp_val"][p])
            h0, hw = ep0.hdb[pt][lt][0], ep0.hdb[pt][lt][1]
            if h0 > 0.1:
                e[4] += (1 / hw) * math.exp(-0.5 * ((r - h0) / hw) ** 2)
            cdc = (pc == 0 and pd == 12) or pc == 2
            if cdc and r * r <= 81:
                x = r * r / 81
                h += ep0.hpp[pd] * (1 - (3.5 * x - 4.5 * x * x + 2.5 * x ** 3 - 0.5 * x ** 4))
        hpl0, hpl1 = ep0.hpl[lt][0], ep0.hpl[lt][1]
        e[5] += 0.5 * ((h - hpl0) / hpl1) ** 2 - math.log(1 / (hpl1 * math.sqrt(2 * math.pi)))
    e[4] *= -1 / math.sqrt(2 * math.pi)
    kde_h = ep0.kde
    for i in range(len(lig)):
        ks = [math.exp(-0.5 / kde_h ** 2 * float(np.sum((lig[i] - np.float32(k).astype(np.float64)) ** 2)))
              for k, t in zip(d["kde"], d["kt"]) if t == d["lt"][i]]
        if ks:
            e[6] += sum(ks) / len(ks)
    e[6] /= (kde_h * math.sqrt(2 * math.pi)) ** 3
    a = len(lig) - 1
    dd = float(np.sum((lig[a] - np.float32(d["mx"][0]).astype(np.float64)) ** 2))
    e[7] = math.log(0.7 * math.sqrt(dd / 1) / 1)
    e[8] = 0.0
    e[:7] /= len(lig)
    e[:9] = np.array(ep0.a_para[:9]) * e[:9] + np.array(ep0.b_para[:9])
    e[9] = float(np.dot(np.array(ep0.w[:9]), e[:9]))
    return e


@pytest.mark.parametrize("n_l,n_p,seed", [(1, 1, 0), (2, 3, 1), (3, 2, 2), (3, 3, 3)])
def test_brute_force_tiny(lib, n_l, n_p, seed):
    rng = np.random.default_rng(seed)
    cx, d = _tiny_complex(rng, n_l, n_p)
    got = O.energy(cx.view, [0], np.zeros(6, np.float32), "f64")[0]
    want = _brute_force(cx, d)
    # the staged tables are float (p1a = 2 eps (r0 lj2)^9 in powf, 1 / hdb in
    # float): agreement to float precision of those inputs
    assert np.allclose(got, want, rtol=2e-6, atol=2e-6), (got, want)


def test_translation_and_point_order_invariance(lib):
    rng = np.random.default_rng(11)
    cx0, _ = _tiny_complex(np.random.default_rng(11), 3, 12)
    cx1, _ = _tiny_complex(np.random.default_rng(11), 3, 12, shift=(7.0, -3.0, 2.0))
    perm = np.random.default_rng(5).permutation(12)
    cx2, _ = _tiny_complex(np.random.default_rng(11), 3, 12, perm=perm)
    move = np.array([0.3, -0.2, 0.1, 0.2, -0.4, 0.7], np.float32)
    e0 = O.energy(cx0.view, [0], move, "f64")[0]
    e1 = O.energy(cx1.view, [0], move, "f64")[0]
    e2 = O.energy(cx2.view, [0], move, "f64")[0]
    assert np.allclose(e0, e1, rtol=1e-5, atol=1e-5)
    assert np.allclose(e0, e2, rtol=1e-9, atol=1e-9)
    del rng


def test_metropolis_rules(kat_complex11):
    """Step 0 is always accepted without draws; from step 1 the proposal is
    m = scale * u_j + m_old (slots 0..5) and the move is accepted iff
    u_6 < expf((E_new - E_old) * minus_beta); downhill moves always pass."""
    _, cx = kat_complex11
    lib = L.load()
    seed = 4242
    reps = cx.replicas()
    n = len(reps)
    O.montecarlo(cx.view, reps, 0, 1, seed)
    assert all(reps[i].is_accept == 1 and reps[i].step == 0 for i in range(n))
    check = np.random.default_rng(0).choice(n, 48, replace=False)
    scale = np.array(cx.view.mcpara.move_scale[:6], np.float32)
    n_down = 0
    for step in range(1, 25):
        old = [(np.float32(reps[i].energy[9]), np.array(reps[i].movematrix[:6], np.float32)) for i in check]
        O.montecarlo(cx.view, reps, step, 1, seed)
        for (e_old, m_old), i in zip(old, check):
            r = reps[i]
            u = np.array([lib.gd_stream_k(seed, 0, r.idx_rep, step, j) for j in range(7)], np.float32) \
                / np.float32(2147483647)
            m_new = scale * u[:6] + m_old
            e_new = O.energy(cx.view, [i], m_new, "f32")[0][9]
            de = np.float32(e_new - e_old)
            thr = np.exp(np.float32(de * np.float32(cx.view.temp[r.idx_tmp].minus_beta)))
            if de <= 0:
                n_down += 1
                assert r.is_accept == 1
            if abs(float(u[6]) - float(thr)) > 1e-6:
                assert r.is_accept == int(u[6] < thr)
            if r.is_accept:
                assert np.array_equal(np.array(r.movematrix[:6], np.float32), m_new)
                assert r.energy[9] == e_new and r.step == step
            else:
                assert np.array_equal(np.array(r.movematrix[:6], np.float32), m_old)
    assert n_down > 0


def test_exchange_rule_closed_forms(kat_complex11):
    """Accept iff u < exp((beta_t - beta_t+1)(E_t - E_t+1)): a colder replica
    holding the higher energy always swaps; swaps only exchange labels between
    neighbours; two rounds touch disjoint pairs."""
    _, cx = kat_complex11
    reps = cx.replicas()
    n_lig, n_prt, n_tmp = cx.size["n_lig"], cx.size["n_prt"], cx.size["n_tmp"]
    for r in range(len(reps)):
        reps[r].energy[9] = -float(reps[r].idx_tmp)  # colder temperature, higher energy
    acc = O.exchange(cx.view, reps, 0, 99)
    assert acc == n_lig * n_prt * (n_tmp // 2)
    for p in range(n_prt):
        for l in range(n_lig):
            labels = [reps[n_tmp * n_lig * p + n_lig * t + l].idx_tmp for t in range(n_tmp)]
            assert labels == [1, 0, 3, 2]
    reps2 = cx.replicas()
    for r in range(len(reps2)):
        reps2[r].energy[9] = float(reps2[r].idx_tmp) * 1e4  # hotter replica far higher: never swap
    assert O.exchange(cx.view, reps2, 1, 99) == 0


def test_exchange_acceptance_rate_matches_probability(kat_complex11):
    """With a fixed energy gap the fraction of accepted swaps over many
    ladders approaches exp(delta) (detailed balance acceptance)."""
    _, cx = kat_complex11
    reps = cx.replicas()
    mb = [cx.view.temp[t].minus_beta for t in range(4)]
    gap = 0.02
    for r in range(len(reps)):
        reps[r].energy[9] = gap * reps[r].idx_tmp
    acc = 0
    trials = 0
    for rnd in range(0, 40, 2):
        rr = (D.ReplicaMC * len(reps))()
        C.memmove(rr, reps, C.sizeof(rr))
        acc += O.exchange(cx.view, rr, rnd, 5)
        trials += cx.size["n_lig"] * cx.size["n_prt"] * 2
    # pairs (0,1) and (2,3): delta = (beta_t - beta_t+1)(E_t - E_t+1) < 0
    b = [-x for x in mb]
    p = 0.5 * (math.exp((b[0] - b[1]) * (-gap)) + math.exp((b[2] - b[3]) * (-gap)))
    rate = acc / trials
    assert abs(rate - p) < 4 * math.sqrt(p * (1 - p) / trials) + 1e-3
