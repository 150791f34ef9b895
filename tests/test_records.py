"""Result formats (CPU): the PrintCsv line reproduces the reference's 1a07C
known-answer line from the oracle's step-0 energies, and the record dump has
the reference's structure and round-trips its numbers at 4 decimals."""
import numpy as np

from geauxdock_cs_v2_amd import records
from oracle import bind as O

# src/common/util_print.cpp:351 (the "Ref Result" line, its label dropped)
KAT_LINE = "  0 0 0.6434 -0.0367 -0.2079 -0.1845 0.8521 -0.8880 0.0523 0.1744 -1.0000 0.7735"


def test_csv_header():
    assert records.csv_header() == "rep step vdw ele pmf psp hdb hpc kde lhm dst total"


def test_csv_line_reproduces_the_known_answer_line(kat_complex):
    _, cx = kat_complex
    e = O.energy(cx.view, np.array([0], np.int32), np.zeros((1, 6), np.float32), "f32")[0]
    assert records.csv_line(0, 0, e) == KAT_LINE


def test_record_dump_structure(kat_complex, tmp_path):
    _, cx = kat_complex
    reps = cx.replicas()
    n = cx.size["n_rep"]
    cap = 20
    rec, nxt = O.montecarlo(cx.view, reps, 0, cap, 77, cap=cap)
    path = tmp_path / "out.txt"
    records.write_records(path, "1a07C-1.pdb", "1a07C1.sdf", cx, reps, rec, nxt, cap, cap)
    lines = path.read_text().splitlines()
    assert lines[0].startswith("prt input file: 1a07C-1.pdb, ")
    assert lines[1].startswith("lig input file: 1a07C1.sdf, ")
    assert sum(l.startswith("replica, ") for l in lines) == n
    entries = [l for l in lines if l.startswith("   ") and l.count(",") == 18]
    assert len(entries) == int(nxt.sum())
    # replica 0's block: counts and its first entry (the initial state, step 0)
    i = lines.index("replica,   0")
    assert lines[i + 4] == f"    number of accepted moves, {nxt[0] - 1}"
    assert lines[i + 5] == f"    number of proposed moves, {cap - 1}"
    fields = [float(x) for x in lines[i + 8].split(",")[:-1]]
    assert fields[0] == 0 and fields[1] == rec[0].step == 0
    want = list(rec[0].movematrix)[:6] + list(rec[0].energy)[:10]
    assert np.allclose(fields[2:], want, atol=5e-5)
