"""Pins of the host loader's conformer filter (Kabsch RMSD) and of the
temperature ladder.

Kabsch: gd_superposed_rmsd replaces the reference's Fortran u3b
(src/frontend/rmsd.f:16-255, called at src/frontend/load.cpp:350-367 to keep
conformers more than 0.1 A from the native pose).  It is checked against
closed forms (rigid copies give 0; two-point sets give half the difference of
their lengths; a scaled copy gives |1 - s| times the radius of gyration; the
mirror image of a planar set is a rotation of it) and against an SVD
restatement with the reflection correction on random sets.

Ladder: the staged minus_beta of every rung is compared bit for bit with the
reference's own Loader::set_temperature (src/frontend/frontend.cpp:222-260)
compiled from its source into oracle/_ref where that tree exists, and
everywhere against the closed form T_i = T_lo (T_hi / T_lo)^(i / (n - 1)),
minus_beta = -1 / T_i, within float rounding."""
import ctypes as C
import math
import os

import numpy as np
import pytest

from conftest import ASTEX
from geauxdock_cs_v2_amd import lib as L
from oracle import bind as O


def rmsd(lib, a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    cols = [np.ascontiguousarray(m[:, k]) for m in (a, b) for k in range(3)]
    ptr = [c.ctypes.data_as(C.POINTER(C.c_float)) for c in cols]
    return lib.gd_superposed_rmsd(*ptr, len(a))


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def svd_rmsd(a, b):
    a = a - a.mean(0)
    b = b - b.mean(0)
    u, s, vt = np.linalg.svd(a.T @ b)
    s[-1] *= np.sign(np.linalg.det(u @ vt))
    return math.sqrt(max((np.sum(a * a) + np.sum(b * b) - 2 * s.sum()) / len(a), 0.0))


def test_rigid_copy_is_zero(lib):
    rng = np.random.default_rng(1)
    for n in (3, 10, 50):
        a = rng.normal(0, 3, (n, 3)) + 30.0
        b = a @ random_rotation(rng).T + rng.normal(0, 10, 3)
        assert rmsd(lib, a, b) < 2e-3 and rmsd(lib, a, a) < 1e-3


def test_two_points(lib):
    # optimally superposed, two segments share their midpoint and direction:
    # each end is off by half the length difference
    a = np.array([[0.0, 0.0, 0.0], [3.0, 0.0, 0.0]])
    b = np.array([[1.0, 2.0, 3.0], [1.0, 2.0, 8.0]])
    assert abs(rmsd(lib, a, b) - 1.0) < 1e-5


def test_scaled_copy(lib):
    rng = np.random.default_rng(2)
    a = rng.normal(0, 2, (20, 3))
    c = a - a.mean(0)
    rg = math.sqrt(np.mean(np.sum(c * c, axis=1)))
    for s in (0.5, 0.9, 1.3):
        b = (c * s) @ random_rotation(rng).T + 5.0
        assert abs(rmsd(lib, a, b) - abs(1 - s) * rg) < 1e-4 * max(1.0, rg)


def test_mirror_of_planar_set_is_a_rotation(lib):
    rng = np.random.default_rng(3)
    a = np.c_[rng.normal(0, 2, (12, 2)), np.zeros(12)]
    b = a * np.array([1.0, 1.0, -1.0])  # mirrored through its own plane
    assert rmsd(lib, a, b) < 1e-3


def test_random_sets_against_svd(lib):
    rng = np.random.default_rng(4)
    for n in (4, 7, 34, 150):
        a = rng.normal(0, 3, (n, 3))
        b = a @ random_rotation(rng).T + rng.normal(0, 0.5, (n, 3))
        if n == 7:
            b = b * np.array([1.0, 1.0, -1.0])  # chiral mirror: the reflection correction matters
        ref = svd_rmsd(a.astype(np.float32).astype(np.float64), b.astype(np.float32).astype(np.float64))
        assert abs(rmsd(lib, a, b) - ref) < 1e-4 * max(1.0, ref), (n, rmsd(lib, a, b), ref)


LADDERS = [(1, 0.3, 2.0), (2, 0.3, 2.0), (8, 0.3, 2.0), (16, 0.3, 2.0), (32, 0.3, 2.0), (5, 0.036, 0.044),
           (12, 0.044, 0.036), (3, 1.0, 1.0), (32, 0.01, 100.0)]


def staged_ladder(lib, n, t_lo, t_hi):
    raw = L.load_job(os.path.join(ASTEX, "1a07C-1.pdb"), os.path.join(ASTEX, "1a07C1.sdf"),
                     os.path.join(ASTEX, "1a07C1-0.8.ff"), "1a07C1", n_tmp=n, temp_high=t_hi, temp_low=t_lo)
    cx = raw.stage()
    mb = np.array([cx.view.temp[i].minus_beta for i in range(n)], np.float32)
    order = np.array([cx.view.temp[i].order for i in range(n)], np.int32)
    return mb, order


@pytest.mark.parametrize("n,t_lo,t_hi", LADDERS)
def test_ladder_closed_form(lib, n, t_lo, t_hi):
    mb, order = staged_ladder(lib, n, t_lo, t_hi)
    assert list(order) == list(range(n))
    for i in range(n):
        t = t_lo * (t_hi / t_lo) ** (i / (n - 1)) if n > 1 else t_lo
        # the ratio is rounded once and the rungs accumulate one rounding per step
        assert abs(-1.0 / mb[i] - t) <= (i + 2) * 2.4e-7 * t, (i, -1.0 / mb[i], t)
    assert mb[0] == np.float32(-1.0) / np.float32(t_lo)


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (no reference tree)")
@pytest.mark.parametrize("n,t_lo,t_hi", LADDERS)
def test_ladder_matches_reference_bitwise(lib, n, t_lo, t_hi):
    mb, order = staged_ladder(lib, n, t_lo, t_hi)
    rmb, rorder = O.ref_ladder(n, t_lo, t_hi)
    assert np.array_equal(mb.view(np.uint32), rmb.view(np.uint32)), (mb, rmb)
    assert np.array_equal(order, rorder)
