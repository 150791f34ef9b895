"""Python bindings of the C ABI in include/gd/geauxdock_amd.h (ctypes).

The shared library is built in-tree (geauxdock_cs_v2_amd/libgeauxdock_amd.so)
by geauxdock_cs_v2_amd.build.  Loading fails loudly when it is missing: there
is no Python or CPU fallback for any entry point.
"""
import ctypes as C
import os

import numpy as np

from . import ctypes_defs as D

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libgeauxdock_amd.so")
PARAM_DIR = os.path.join(PKG_DIR, "data", "parameters")

GD_OK, GD_ERR_ARG, GD_ERR_IO, GD_ERR_DEVICE, GD_ERR_NOMEM, GD_ERR_STATE, GD_ERR_BUSY = 0, -1, -2, -3, -4, -5, -6


class GDError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"geauxdock error {code}: {msg}")
        self.code = code


_lib = None


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


def load(path=None):
    """Load libgeauxdock_amd.so once and declare every exported signature."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError(f"{p} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(p, mode=C.RTLD_GLOBAL)
    vp, i32, u32, u64, f32 = C.c_void_p, C.c_int, C.c_uint32, C.c_uint64, C.c_float
    pp = C.POINTER(C.c_void_p)
    _sig(lib, "gd_last_error", C.c_char_p, [])
    _sig(lib, "gd_version", C.c_char_p, [])
    _sig(lib, "gd_philox4x32_10", None, [C.POINTER(u32), C.POINTER(u32), C.POINTER(u32)])
    _sig(lib, "gd_stream_k", u32, [u64, u32, u32, u32, u32])
    _sig(lib, "gd_load_job_line", i32, [C.c_char_p, pp])
    _sig(lib, "gd_load_job", i32, [C.c_char_p] * 8 + [i32, f32, f32, f32, f32, i32, pp])
    _sig(lib, "gd_raw_view", C.POINTER(D.Raw), [vp])
    _sig(lib, "gd_raw_free", None, [vp])
    _sig(lib, "gd_stage", i32, [C.POINTER(D.Raw), pp])
    _sig(lib, "gd_complex_view", C.POINTER(D.Complex), [vp])
    _sig(lib, "gd_complex_free", None, [vp])
    _sig(lib, "gd_abi_layout", C.c_size_t, [i32])
    fpp = C.POINTER(C.c_float)
    _sig(lib, "gd_superposed_rmsd", C.c_double, [fpp] * 6 + [i32])
    if hasattr(lib, "gd_create"):
        _declare_device(lib)
    if path is None:
        _lib = lib
    return lib


def _declare_device(lib):
    vp, i32, u32, u64 = C.c_void_p, C.c_int, C.c_uint32, C.c_uint64
    pp = C.POINTER(C.c_void_p)
    pi = C.POINTER(C.c_int)
    pf = C.POINTER(C.c_float)
    _sig(lib, "gd_create", i32, [i32, pp])
    _sig(lib, "gd_destroy", None, [vp])
    _sig(lib, "gd_device_info", i32, [vp, C.c_char_p, i32, pi, C.POINTER(C.c_size_t)])
    _sig(lib, "gd_upload_target", i32, [vp, C.POINTER(D.Complex), pi])
    _sig(lib, "gd_free_target", i32, [vp, i32])
    _sig(lib, "gd_upload_job", i32, [vp, i32, C.POINTER(D.Complex), u64, pi])
    _sig(lib, "gd_free_job", i32, [vp, i32])
    _sig(lib, "gd_job_footprint", C.c_size_t, [i32] * 8)
    _sig(lib, "gd_montecarlo", i32, [vp, i32, i32, i32])
    _sig(lib, "gd_exchange", i32, [vp, i32, i32, pi])
    _sig(lib, "gd_energy", i32, [vp, i32, i32, pi, pf, pf])
    _sig(lib, "gd_read_replicas", i32, [vp, i32, C.POINTER(D.ReplicaMC)])
    _sig(lib, "gd_write_replicas", i32, [vp, i32, C.POINTER(D.ReplicaMC)])
    _sig(lib, "gd_read_records", i32, [vp, i32, C.POINTER(D.ReplicaMC), pi, i32])
    if not hasattr(lib, "gd_batch_run"):
        raise ImportError("libgeauxdock_amd.so lacks the batched and serving entry points: rebuild it")
    _sig(lib, "gd_batch_run", i32, [vp, pi, i32, C.POINTER(D.BatchOpts), C.POINTER(C.c_double)])
    _sig(lib, "gd_server_create", i32, [vp, i32, C.POINTER(D.ServerOpts), pp])
    _sig(lib, "gd_server_destroy", None, [vp])
    _sig(lib, "gd_server_submit", i32, [vp, C.POINTER(D.Complex), i32, u64, pi])
    _sig(lib, "gd_server_poll", i32, [vp, i32, C.POINTER(D.RequestResult), C.POINTER(D.ReplicaMC)])
    _sig(lib, "gd_set_stream", i32, [vp, i32, u64, C.POINTER(u32), i32])
    _sig(lib, "gd_job_shape", i32, [vp, i32, pi])
    _sig(lib, "gd_batch_plan", i32, [vp, pi])
    _sig(lib, "gd_batch_run_noalloc", i32, [vp, pi, i32, C.POINTER(D.BatchOpts), C.POINTER(C.c_double)])
    _sig(lib, "gd_batch_noalloc_capacity", C.c_size_t, [])
    _sig(lib, "gd_read_mclog", i32, [vp, i32, pi, pi])
    _sig(lib, "gd_server_cancel", i32, [vp, i32])
    _sig(lib, "gd_server_drain", i32, [vp])
    _sig(lib, "gd_server_clock_ms", C.c_double, [vp])


def check(rc, lib=None):
    if rc != GD_OK:
        lib = lib or load()
        raise GDError(rc, lib.gd_last_error().decode(errors="replace"))
    return rc


# ---------------------------------------------------------------------------
# host objects
# ---------------------------------------------------------------------------
class RawJob:
    """Loader output (gd_Raw) owned by the library."""

    def __init__(self, handle, lib):
        self._h = handle
        self._lib = lib
        self.view = lib.gd_raw_view(handle).contents

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.gd_raw_free(self._h)
            self._h = None

    def stage(self):
        h = C.c_void_p()
        check(self._lib.gd_stage(C.byref(self.view), C.byref(h)), self._lib)
        return StagedComplex(h, self._lib, keep=self)


class StagedComplex:
    """Staged tables of one job (gd_Complex) owned by the library."""

    def __init__(self, handle, lib, keep=None):
        self._h = handle
        self._lib = lib
        self._keep = keep
        self.view = lib.gd_complex_view(handle).contents

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.gd_complex_free(self._h)
            self._h = None

    @property
    def size(self):
        s = self.view.size
        return {n: getattr(s, n) for n, _ in D.ComplexSize._fields_}

    def array(self, name, n=None):
        """numpy copy of an array field of the view."""
        v = self.view
        s = v.size
        lens = {
            "lig": s.n_lig * s.lig_natom, "prt": s.n_prt * s.prt_npoint, "kde": s.kde_npoint,
        }
        if n is None:
            if name.startswith("lig_"):
                n = lens["lig"]
            elif name.startswith("prt_"):
                n = lens["prt"]
            elif name.startswith("kde_"):
                n = lens["kde"]
            elif name in ("mcs_ncol", "mcs_tcc"):
                n = s.mcs_nrow
            elif name == "mcs_rowptr":
                n = s.mcs_nrow + 1
            elif name.startswith("mcs_"):
                n = v.mcs_rowptr[s.mcs_nrow]
            elif name == "psp":
                n = D.MAXTP2 * v.psp_width
            else:
                raise KeyError(name)
        return np.ctypeslib.as_array(getattr(v, name), shape=(n,)).copy()

    def replicas(self):
        n = self.view.size.n_rep
        arr = (D.ReplicaMC * n)()
        C.memmove(arr, self.view.replica, C.sizeof(D.ReplicaMC) * n)
        return arr


def load_job(prt, lig, ff, lig_id, n_tmp=1, temp_high=0.044, temp_low=0.036, tras_scale=0.02,
             rot_scale=0.08, n_dump=1, params=None, lib=None):
    """Loader::load for explicit paths; parameter files default to the packaged set."""
    lib = lib or load()
    pdir = params or PARAM_DIR
    h = C.c_void_p()
    enc = lambda s: os.fspath(s).encode()
    rc = lib.gd_load_job(enc(prt), enc(lig), enc(ff), enc(lig_id), enc(os.path.join(pdir, "08ff_opt")),
                         enc(os.path.join(pdir, "paras")), enc(os.path.join(pdir, "08_nor_a")),
                         enc(os.path.join(pdir, "08_nor_b")), int(n_tmp), float(temp_high),
                         float(temp_low), float(tras_scale), float(rot_scale), int(n_dump), C.byref(h))
    check(rc, lib)
    return RawJob(h, lib)


def load_job_line(line, lib=None):
    lib = lib or load()
    h = C.c_void_p()
    check(lib.gd_load_job_line(line.encode(), C.byref(h)), lib)
    return RawJob(h, lib)


def replicas_to_numpy(reps):
    """(n, 6) movematrix, (n, 10) energy, and (n, 4) idx_rep/lig/prt/tmp, step, is_accept arrays."""
    n = len(reps)
    mm = np.array([list(r.movematrix) for r in reps], dtype=np.float32).reshape(n, 6)
    en = np.array([list(r.energy) for r in reps], dtype=np.float32).reshape(n, 10)
    idx = np.array([(r.idx_rep, r.idx_lig, r.idx_prt, r.idx_tmp) for r in reps], dtype=np.int32).reshape(n, 4)
    step = np.array([r.step for r in reps], dtype=np.int32)
    acc = np.array([r.is_accept for r in reps], dtype=np.int32)
    return mm, en, idx, step, acc
