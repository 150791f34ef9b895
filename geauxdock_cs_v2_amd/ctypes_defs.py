"""ctypes mirrors of the plain-C structures in include/gd/gd_types.h.

The layouts must match the C header byte for byte; tests/test_abi.py checks
the sizes and field offsets against the compiled library.
"""
import ctypes as C

MAXTP1, MAXTP2, MAXTP3, MAXTP4, MAXWEI = 31, 24, 50, 20, 10
TERMS = ("vdw", "ele", "pmf", "psp", "hdb", "hpc", "kde", "lhm", "dst", "total")

fp = C.POINTER(C.c_float)
ip = C.POINTER(C.c_int)


class EnePara(C.Structure):
    _fields_ = [
        ("p1a", (C.c_float * MAXTP1) * MAXTP2),
        ("p2a", (C.c_float * MAXTP1) * MAXTP2),
        ("lj0", C.c_float), ("lj1", C.c_float),
        ("el1", C.c_float), ("el0", C.c_float), ("a1", C.c_float), ("b1", C.c_float),
        ("pmf0", (C.c_float * MAXTP1) * MAXTP2),
        ("pmf1", (C.c_float * MAXTP1) * MAXTP2),
        ("hdb0", (C.c_float * MAXTP1) * MAXTP2),
        ("hdb1", (C.c_float * MAXTP1) * MAXTP2),
        ("hpp", C.c_float * MAXTP4),
        ("hpl0", C.c_float * MAXTP2), ("hpl1", C.c_float * MAXTP2), ("hpl2", C.c_float * MAXTP2),
        ("kde2", C.c_float), ("kde3", C.c_float),
        ("a_para", C.c_float * MAXWEI), ("b_para", C.c_float * MAXWEI), ("w", C.c_float * MAXWEI),
        ("_pad", C.c_char * 56),  # aligned(64): 18376 -> 18432 bytes
    ]


class EnePara0(C.Structure):
    _fields_ = [
        ("vdw", ((C.c_float * 2) * MAXTP2) * MAXTP1),
        ("ele", C.c_float * MAXTP3),
        ("pmf", ((C.c_float * 2) * MAXTP2) * MAXTP1),
        ("hpp", C.c_float * MAXTP4),
        ("hpl", (C.c_float * 2) * MAXTP2),
        ("hdb", ((C.c_float * 2) * MAXTP2) * MAXTP1),
        ("lj", C.c_float * 3), ("el", C.c_float * 2), ("kde", C.c_float),
        ("w", C.c_float * MAXWEI), ("a_para", C.c_float * MAXWEI), ("b_para", C.c_float * MAXWEI),
    ]


class Temp(C.Structure):
    _fields_ = [("t", C.c_float), ("minus_beta", C.c_float), ("order", C.c_int)]


class ReplicaMC(C.Structure):
    _fields_ = [
        ("idx_rep", C.c_int), ("idx_lig", C.c_int), ("idx_prt", C.c_int), ("idx_tmp", C.c_int),
        ("movematrix", C.c_float * 6), ("energy", C.c_float * MAXWEI),
        ("step", C.c_int), ("is_accept", C.c_int),
        ("_pad", C.c_char * 40),  # aligned(64): 88 -> 128 bytes
    ]


class McPara(C.Structure):
    _fields_ = [("steps_per_dump", C.c_int), ("steps_n_dump", C.c_int), ("steps_total", C.c_int),
                ("move_scale", C.c_float * 6)]


class ComplexSize(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("n_lig", "n_prt", "n_tmp", "n_rep", "lig_natom",
                                       "prt_npoint", "kde_npoint", "mcs_nrow")]


def _aligned(name, fields, align_before, align_total=64):
    """Build a Structure with the padding the C compiler inserts for a
    64-byte-aligned member (`align_before`) and for the struct's alignment."""
    head = []
    for f in fields:
        if f[0] == align_before:
            probe = type(name + "_probe", (C.Structure,), {"_fields_": list(head)})
            gap = (-C.sizeof(probe)) % 64
            if gap:
                head.append(("_pad_align", C.c_char * gap))
        head.append(f)
    probe = type(name + "_probe", (C.Structure,), {"_fields_": list(head)})
    tail = (-C.sizeof(probe)) % align_total
    if tail:
        head.append(("_pad_tail", C.c_char * tail))
    return type(name, (C.Structure,), {"_fields_": head})


_complex_fields = [
        ("size", ComplexSize),
        ("lig_x2", fp), ("lig_y2", fp), ("lig_z2", fp), ("lig_c2", fp), ("lig_t2", ip),
        ("lig_x3", fp), ("lig_y3", fp), ("lig_z3", fp), ("lig_c3", fp), ("lig_t3", ip),
        ("lig_kde_begin_idx", ip), ("lig_kde_end_idx", ip),
        ("lig_center", C.c_float * 3),
        ("prt_x", fp), ("prt_y", fp), ("prt_z", fp), ("prt_hpp", fp), ("prt_ele", fp),
        ("prt_t", ip), ("prt_c", ip), ("prt_cdc", ip), ("prt_seq3r", ip),
        ("prt_pocket_center", C.c_float * 3),
        ("psp_width", C.c_int), ("psp", fp),
        ("kde_t", ip), ("kde_x", fp), ("kde_y", fp), ("kde_z", fp),
        ("mcs_ncol", ip), ("mcs_tcc", fp), ("mcs_rowptr", ip), ("mcs_i", ip),
        ("mcs_x", fp), ("mcs_y", fp), ("mcs_z", fp),
        ("enepara", EnePara),
        ("temp", C.POINTER(Temp)),
        ("replica", C.POINTER(ReplicaMC)),
        ("mcpara", McPara),
]
Complex = _aligned("Complex", _complex_fields, "enepara")


class Raw(C.Structure):
    _fields_ = [
        ("n_lig", C.c_int), ("lig_natom", C.c_int),
        ("lig_x", fp), ("lig_y", fp), ("lig_z", fp), ("lig_t", ip), ("lig_c", fp),
        ("lig_center", C.c_float * 3),
        ("n_prt", C.c_int), ("prt_npoint", C.c_int), ("prt_nres", C.c_int),
        ("prt_x", fp), ("prt_y", fp), ("prt_z", fp),
        ("prt_t", ip), ("prt_c", ip), ("prt_d", ip), ("prt_r", ip), ("prt_seq3", ip),
        ("psp_n", C.c_int), ("psp_res", ip), ("psp_lig", ip), ("psp_val", fp),
        ("kde_npoint", C.c_int), ("kde_t", ip), ("kde_x", fp), ("kde_y", fp), ("kde_z", fp),
        ("mcs_nrow", C.c_int), ("mcs_tcc", fp), ("mcs_rowptr", ip), ("mcs_i", ip),
        ("mcs_x", fp), ("mcs_y", fp), ("mcs_z", fp),
        ("enepara0", EnePara0),
        ("n_tmp", C.c_int), ("temp_high", C.c_float), ("temp_low", C.c_float),
        ("move_scale", C.c_float * 6), ("steps_n_dump", C.c_int),
    ]


class BatchOpts(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("steps", "exchange_interval", "use_graph", "slots_per_wave",
                                       "waves_per_group", "point_slices", "compensated",
                                       "keep_records")]


class ServerOpts(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("max_active_jobs", C.c_int), ("chunk_steps", C.c_int),
                ("exchange_interval", C.c_int)]


class RequestResult(C.Structure):
    _fields_ = [("status", C.c_int), ("n_rep", C.c_int), ("t_submit_ms", C.c_double),
                ("t_admit_ms", C.c_double), ("t_done_ms", C.c_double)]
