"""CPU tests of the C ABI boundary: the library loads, exports every entry
point the public headers declare, and the ctypes mirrors match the compiled
structure layouts.  No GPU calls."""
import ctypes as C
import os
import re
import subprocess

import pytest

from conftest import ROOT
from geauxdock_cs_v2_amd import ctypes_defs as D
from geauxdock_cs_v2_amd import lib as L

HEADERS = [os.path.join(ROOT, "include", "gd", h) for h in os.listdir(os.path.join(ROOT, "include", "gd"))
           if h.endswith(".h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(gd_[a-z0-9_]+)\s*\(", txt):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], stdout=subprocess.PIPE, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing
    for n in declared_functions():
        assert hasattr(lib, n)


def test_struct_layouts_match(lib):
    want = {0: C.sizeof(D.Complex), 1: D.Complex.enepara.offset, 2: C.sizeof(D.Raw), 3: D.Raw.enepara0.offset,
            4: C.sizeof(D.EnePara), 5: C.sizeof(D.EnePara0), 6: C.sizeof(D.ReplicaMC), 7: C.sizeof(D.Temp),
            8: D.Complex.mcpara.offset, 9: D.Raw.n_tmp.offset}
    for k, v in want.items():
        assert lib.gd_abi_layout(k) == v, k
    # the reference's ReplicaMC is 128 bytes and EnePara 18432 (64-byte aligned)
    assert C.sizeof(D.ReplicaMC) == 128 and C.sizeof(D.EnePara) == 18432


def test_errors_without_device(lib):
    """argument checks run before any HIP call"""
    assert lib.gd_stage(None, None) == L.GD_ERR_ARG
    assert b"null" in lib.gd_last_error()
    assert lib.gd_create(0, None) == L.GD_ERR_ARG


def test_version(lib):
    assert b"gfx950" in lib.gd_version()


def test_docker_header_compiles(tmp_path):
    """include/gd/docker.hpp: the reference-shaped C++ driver over the C ABI"""
    src = tmp_path / "t.cpp"
    src.write_text('#include "gd/docker.hpp"\nint main(){ gd::Docker* d = nullptr; (void)d; return 0; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", f"-I{os.path.join(ROOT, 'include')}", str(src)],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout


def test_staging_refuses_malformed_loader_output(lib, kat_complex):
    """gd_stage checks a caller-built loader output before following it: null
    tables and indices outside their tables are refused with GD_ERR_ARG (the
    library's own loader never writes them).  CPU only."""
    import numpy as np
    raw = kat_complex[0]
    assert raw.stage() is not None  # the loader's own output stages
    h = C.c_void_p()

    def refused(edit):
        v = D.Raw.from_buffer_copy(raw.view)
        keep = edit(v)
        rc = lib.gd_stage(C.byref(v), C.byref(h))
        del keep
        return rc == L.GD_ERR_ARG, lib.gd_last_error().decode()

    def null_lig_t(v):
        v.lig_t = None

    def bad_residue_index(v):
        r = np.ctypeslib.as_array(v.prt_r, shape=(v.prt_npoint,)).copy()
        r[0] = v.prt_nres
        v.prt_r = r.ctypes.data_as(C.POINTER(C.c_int))
        return r

    def bad_point_type(v):
        t = np.ctypeslib.as_array(v.prt_t, shape=(v.prt_npoint,)).copy()
        t[3] = D.MAXTP1
        v.prt_t = t.ctypes.data_as(C.POINTER(C.c_int))
        return t

    def bad_mcs_atom(v):
        n = np.ctypeslib.as_array(v.mcs_rowptr, shape=(v.mcs_nrow + 1,))[-1]
        i = np.ctypeslib.as_array(v.mcs_i, shape=(n,)).copy()
        i[-1] = v.lig_natom
        v.mcs_i = i.ctypes.data_as(C.POINTER(C.c_int))
        return i

    for edit, words in [(null_lig_t, "null"), (bad_residue_index, "residue index"),
                        (bad_point_type, "point type"), (bad_mcs_atom, "MCS atom index")]:
        ok, why = refused(edit)
        assert ok and words in why, (edit.__name__, why)


def test_phase_profile_build_exports_its_reader():
    """The phase-profile build (a tracing build beside the product library)
    exports its readers; the product library does not carry the marks."""
    from geauxdock_cs_v2_amd import build as B
    if not os.path.exists(B.PHASE_LIB):
        pytest.skip("phase-profile build not present (build.build_phase_profile)")
    out = subprocess.run(["nm", "-D", "--defined-only", B.PHASE_LIB], stdout=subprocess.PIPE, text=True).stdout
    names = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert {"gd_phase_profile", "gd_phase_profile_wg"} <= names
    prod = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], stdout=subprocess.PIPE, text=True).stdout
    assert "gd_phase_profile" not in prod
