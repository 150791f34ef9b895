"""CPU tests of the synthetic workload writer: the files it writes in the
reference formats read back through the loader to exactly the requested
sizes, with effective points on the sampled sites."""
import numpy as np

from geauxdock_cs_v2_amd import lib as L
from geauxdock_cs_v2_amd import synth


def test_small_workload_roundtrip(lib, tmp_path):
    out = synth.generate(str(tmp_path), "tlig", seed=5, n_atoms=17, n_lig_conf=4, lig_sigma=0.4, n_points=301,
                         n_prt_conf=3, prt_sigma=0.5, prt_radius=16.0, n_kde=50, mcs_rows=5, mcs_cols=7)
    raw = L.load_job(out["pdb"], out["sdf"], out["ff"], "tlig", n_tmp=3, temp_high=2.0, temp_low=0.3)
    v = raw.view
    assert (v.lig_natom, v.n_lig, v.prt_npoint, v.n_prt, v.kde_npoint, v.mcs_nrow) == (17, 4, 301, 3, 50, 5)
    # points in load order equal the sites of conformation 0 (3-decimal PDB)
    sites = out["protein"]["sites"]
    px = np.ctypeslib.as_array(v.prt_x, shape=(3 * 301,))[:301]
    py = np.ctypeslib.as_array(v.prt_y, shape=(3 * 301,))[:301]
    pz = np.ctypeslib.as_array(v.prt_z, shape=(3 * 301,))[:301]
    assert np.allclose(np.stack([px, py, pz], 1), sites, atol=1.5e-3)
    # minimum spacing of the protein sites
    d = np.linalg.norm(sites[:, None] - sites[None], axis=2) + np.eye(301) * 1e9
    assert d.min() >= 2.8 - 1e-6
    cx = raw.stage()
    assert cx.size["n_rep"] == 4 * 3 * 3


def test_named_workload_sizes(lib):
    cx, info = synth.load_config("c1")
    s = cx.size
    assert (s["lig_natom"], s["n_lig"], s["prt_npoint"], s["n_prt"], s["kde_npoint"], s["mcs_nrow"],
            s["n_tmp"], s["n_rep"]) == (24, 6, 1500, 1, 200, 8, 8, 48)
    t = [cx.view.temp[i].t for i in range(8)]
    assert abs(t[0] - 0.3) < 1e-6 and abs(t[-1] - 2.0) < 1e-4
    q = cx.array("lig_c2")[:24]
    assert abs(q.sum()) < 1e-5
