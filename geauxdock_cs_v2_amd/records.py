"""The reference's result formats, for runs driven through this package.

csv_header / csv_line   PrintCsv (src/common/util_print.cpp:77-110): the
                        "rep step vdw ... total" header and one line per
                        replica state, energies to 4 decimals -- the format
                        of the 1a07C known-answer line (util_print.cpp:351).
write_records           the per-replica record dump (util_print.cpp:378-435):
                        replica indices and temperature, accepted / proposed
                        move counts, then one line per record entry (entry,
                        step, 6 move parameters, 10 energies).
"""
from . import ctypes_defs as D

TERM_NAMES = ("vdw", "ele", "pmf", "psp", "hdb", "hpc", "kde", "lhm", "dst", "total")


def csv_header():
    return "rep step " + " ".join(TERM_NAMES)


def csv_line(idx_rep, step, energy):
    return "%3d %d" % (idx_rep, step) + "".join(" %.4f" % float(e) for e in list(energy)[:10])


def write_records(path, prt_file, lig_file, cx, replicas, record, next_entry, steps_total, cap):
    """Write the record dump.  `replicas`: the final ReplicaMC array;
    `record`: the record window of the last chunk (n_rep x cap ReplicaMC, as
    gd_read_records / Device.records return it); `next_entry[r]`: entries of
    replica r; `cx`: the staged complex (sizes and temperatures)."""
    size = cx.size
    temp = cx.view.temp
    lines = [f"prt input file: {prt_file}, \t{size['n_prt']:3d} conformations",
             f"lig input file: {lig_file}, \t{size['n_lig']:3d} conformations", ""]
    for r in range(size["n_rep"]):
        rep = replicas[r]
        ne = int(next_entry[r])
        lines.append(f"replica, {r:3d}")
        lines.append(f"    prt  index, {rep.idx_prt:3d}")
        lines.append(f"    temp index, {rep.idx_tmp:3d}, minus beta, {temp[rep.idx_tmp].minus_beta:f}")
        lines.append(f"    lig  index, {rep.idx_lig:3d} ")
        lines.append(f"    number of accepted moves, {ne - 1}")
        lines.append(f"    number of proposed moves, {steps_total - 1}")
        lines.append("    The first record entry contains the initial status.")
        lines.append('    Records in format "record entry, mc move step, translation x y z, rotation x y z, '
                     'evdw, eele, epmf, epsp, ehdb, ehpc, ekde, emcs, edst, etotal"')
        for s in range(ne):
            e = record[r * cap + s]
            body = "".join("%8.4f, " % v for v in list(e.movematrix)[:6])
            body += "".join("%8.4f, " % v for v in list(e.energy)[:10])
            lines.append("   %5d, %5d, " % (s, e.step) + body)
        lines.append("")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def empty_records(n_rep, cap):
    return (D.ReplicaMC * (n_rep * cap))()
