"""ctypes bindings of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, by __graft_entry__.smoke() (as the checker) and by the
CPU-baseline leg of bench.py.  The product never imports this module.

liboracle.so  : the repository's C restatement (gd_oracle.c)
_ref/libgdref.so : the reference's own MonteCarlo_d, table builders and
                   temperature ladder, compiled from its sources (present only
                   where it was built)
"""
import ctypes as C
import os

import numpy as np

from geauxdock_cs_v2_amd import ctypes_defs as D

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libgdref.so")

_oracle = None
_ref = None


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise ImportError(f"{ORACLE_SO} missing: run `make -C oracle`")
        lib = C.CDLL(ORACLE_SO)
        u32p = C.POINTER(C.c_uint32)
        lib.gdo_philox4x32_10.argtypes = [u32p, u32p, u32p]
        lib.gdo_philox4x32_10.restype = None
        lib.gdo_stream_k.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
        lib.gdo_stream_k.restype = C.c_uint32
        lib.gdo_energy_f32.argtypes = [C.POINTER(D.Complex), C.c_int, C.POINTER(C.c_int),
                                       C.POINTER(C.c_float), C.POINTER(C.c_float)]
        lib.gdo_energy_f32.restype = C.c_int
        lib.gdo_energy_f64.argtypes = [C.POINTER(D.Complex), C.c_int, C.POINTER(C.c_int),
                                       C.POINTER(C.c_float), C.POINTER(C.c_double)]
        lib.gdo_energy_f64.restype = C.c_int
        lib.gdo_montecarlo_f32.argtypes = [C.POINTER(D.Complex), C.POINTER(D.ReplicaMC), C.c_int, C.c_int,
                                           C.POINTER(D.ReplicaMC), C.POINTER(C.c_int), C.c_int, C.c_int,
                                           C.c_int, C.c_uint64, C.c_int]
        lib.gdo_montecarlo_f32.restype = C.c_int
        lib.gdo_exchange.argtypes = [C.POINTER(D.Complex), C.POINTER(D.ReplicaMC), C.c_int, C.c_uint64]
        lib.gdo_exchange.restype = C.c_int
        _oracle = lib
    return _oracle


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise ImportError(f"{REF_SO} missing (built only where the reference tree exists)")
        lib = C.CDLL(REF_SO)
        lib.gdref_create.argtypes = [C.POINTER(D.Raw), C.POINTER(D.Temp), C.c_int, C.POINTER(C.c_void_p)]
        lib.gdref_create.restype = C.c_int
        lib.gdref_view.argtypes = [C.c_void_p]
        lib.gdref_view.restype = C.POINTER(D.Complex)
        lib.gdref_free.argtypes = [C.c_void_p]
        lib.gdref_free.restype = None
        lib.gdref_montecarlo.argtypes = [C.c_void_p, C.POINTER(D.ReplicaMC), C.POINTER(D.ReplicaMC),
                                         C.POINTER(C.c_int), C.c_int, C.c_int, C.c_uint64, C.c_int]
        lib.gdref_montecarlo.restype = C.c_int
        lib.gdref_set_temperature.argtypes = [C.c_int, C.c_float, C.c_float, C.POINTER(C.c_float),
                                              C.POINTER(C.c_int)]
        lib.gdref_set_temperature.restype = C.c_int
        _ref = lib
    return _ref


def _f32p(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _i32p(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def energy(view, rep_idx, move, precision="f32"):
    """Oracle pose energies: (n, 10) normalised terms."""
    lib = oracle()
    rep_idx = np.ascontiguousarray(rep_idx, dtype=np.int32)
    move = np.ascontiguousarray(move, dtype=np.float32).reshape(-1, 6)
    n = len(rep_idx)
    if precision == "f32":
        out = np.zeros((n, 10), np.float32)
        rc = lib.gdo_energy_f32(C.byref(view), n, _i32p(rep_idx), _f32p(move), _f32p(out))
    else:
        out = np.zeros((n, 10), np.float64)
        rc = lib.gdo_energy_f64(C.byref(view), n, _i32p(rep_idx), _f32p(move),
                                out.ctypes.data_as(C.POINTER(C.c_double)))
    if rc != 0:
        raise RuntimeError("oracle energy failed")
    return out


def montecarlo(view, reps, s1, s2max, seed, cap=None, r0=0, r1=None, nthreads=0):
    """Run the oracle MonteCarlo_d in place on a ctypes ReplicaMC array.
    Returns (records array, next_entry numpy array)."""
    lib = oracle()
    n = len(reps)
    r1 = n if r1 is None else r1
    cap = s2max if cap is None else cap
    rec = (D.ReplicaMC * max(1, n * cap))()
    nxt = np.zeros(n, np.int32)
    rc = lib.gdo_montecarlo_f32(C.byref(view), reps, r0, r1, rec, _i32p(nxt), cap, s1, s2max, seed, nthreads)
    if rc != 0:
        raise RuntimeError("oracle montecarlo failed")
    return rec, nxt


def exchange(view, reps, rnd, seed):
    return oracle().gdo_exchange(C.byref(view), reps, rnd, seed)


def ref_ladder(n, t_low, t_high):
    """The reference's own Loader::set_temperature: (minus_beta, order) of an
    n-rung ladder between t_low and t_high (float32 arrays)."""
    mb = np.zeros(n, np.float32)
    order = np.zeros(n, np.int32)
    if ref().gdref_set_temperature(n, t_low, t_high, _f32p(mb), _i32p(order)) != 0:
        raise ValueError("ladder size out of the reference's range")
    return mb, order


class RefComplex:
    """The reference's own staged Complex for a gd_Raw (oracle/_ref)."""

    def __init__(self, raw_view, temps):
        lib = ref()
        self._lib = lib
        n = len(temps)
        arr = (D.Temp * n)(*temps)
        h = C.c_void_p()
        if lib.gdref_create(C.byref(raw_view), arr, n, C.byref(h)) != 0:
            raise ValueError("input exceeds the reference's static capacities")
        self._h = h
        self.view = lib.gdref_view(h).contents

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.gdref_free(self._h)
            self._h = None

    def montecarlo(self, reps, s1, s2max, seed, nthreads=0):
        n = len(reps)
        rec = (D.ReplicaMC * (n * 100))()
        nxt = np.zeros(n, np.int32)
        if self._lib.gdref_montecarlo(self._h, reps, rec, _i32p(nxt), s1, s2max, seed, nthreads) != 0:
            raise RuntimeError("reference montecarlo failed")
        return rec, nxt
