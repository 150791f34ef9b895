"""Synthetic docking inputs written in the reference's own file formats.

Every workload is generated from one seed and written as
  <id>.sdf  V2000 atom block + OB_ATOM_TYPES / OB_ATOMIC_CHARGES /
            ENSEMBLE_COORDS property blocks (read by src/frontend/load.cpp:90-397)
  <id>.pdb  MODEL/ENDMDL protein conformations whose coarse-grained effective
            points (load.cpp:576-970) land on sampled sites
  <id>.ff   KDE / PSP / MCS lines (load.cpp:994-1171)
and then read back through the library's loader, so the benchmarks exercise
the same parser and staging as real inputs.

Protein sites: effective points are placed on a jittered FCC lattice that
fills a sphere around the pocket centre, minus a cavity around the ligand,
with the minimum spacing the configuration asks for where that fits.  A
residue's atoms are written so the point averages of the loader reproduce
the sites: the CA atom on the CA site; the peptide-plate site of residue r
on N(r), C(r-1) and O(r-1); a single side-chain site on CB; a two-point side
chain's first site on CD and second site on CB and CG.
"""
import hashlib
import json
import math
import os

import numpy as np

LIG_TYPES = ["Br", "C.1", "C.2", "C.3", "C.ar", "C.cat", "Cl", "F", "I", "N.1", "N.2", "N.3", "N.4",
             "N.am", "N.ar", "N.pl3", "O.2", "O.3", "O.co2", "P.3", "S.2", "S.3", "S.O", "S.O2"]
RES3 = ["ALA", "ARG", "CYS", "PRO", "LEU", "ILE", "GLU", "HIS", "TRP", "THR", "MET", "ASP", "GLY", "PHE",
        "LYS", "GLN", "SER", "ASN", "TYR", "VAL"]
# side-chain effective points per residue code (load.cpp:631-698): 0, 1 or 2
SC_POINTS = [1, 2, 1, 1, 1, 1, 2, 2, 2, 1, 2, 1, 0, 2, 2, 2, 1, 1, 2, 1]
POCKET = np.array([31.25, 9.75, 42.5])


def _rsa_ball(rng, n, radius, spacing, center, max_tries=2000000):
    """Random sequential addition of n points in a ball with minimum spacing."""
    pts = np.empty((n, 3))
    k = 0
    tries = 0
    while k < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("cannot place ligand atoms with the requested spacing")
        v = rng.uniform(-radius, radius, 3)
        if v @ v > radius * radius:
            continue
        p = center + v
        if k and np.min(np.sum((pts[:k] - p) ** 2, axis=1)) < spacing * spacing:
            continue
        pts[k] = p
        k += 1
    return pts


def _kabsch_rmsd(a, b):
    a = a - a.mean(0)
    b = b - b.mean(0)
    u, s, vt = np.linalg.svd(a.T @ b)
    d = np.sign(np.linalg.det(u @ vt))
    s[-1] *= d
    e = (np.sum(a * a) + np.sum(b * b) - 2 * s.sum()) / len(a)
    return math.sqrt(max(e, 0.0))


def _fcc_sites(rng, n, radius, cavity, spacing, center):
    """n sites from a jittered FCC lattice filling radius minus cavity."""
    vol = 4.0 / 3.0 * math.pi * (radius ** 3 - cavity ** 3)
    d = (math.sqrt(2.0) * vol / n) ** (1.0 / 3.0) * 0.97
    a = d * math.sqrt(2.0)  # cubic cell edge of the FCC lattice
    m = int(math.ceil(radius / a)) + 1
    g = np.arange(-m, m + 1) * a
    base = np.array(np.meshgrid(g, g, g, indexing="ij")).reshape(3, -1).T
    offs = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]]) * a
    lat = (base[:, None, :] + offs[None]).reshape(-1, 3)
    r = np.linalg.norm(lat, axis=1)
    lat = lat[(r <= radius) & (r >= cavity)]
    if len(lat) < n:
        raise RuntimeError("lattice too small for the requested point count")
    lat = lat[rng.choice(len(lat), n, replace=False)]
    jit = max(0.0, min(0.3, (d - spacing) / 2.0 - 1e-3))
    if jit > 0:
        v = rng.normal(size=(n, 3))
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        lat = lat + v * rng.uniform(0, jit, (n, 1))
    return center + lat, min(d, d - 2 * jit) if jit > 0 else d


def _residue_plan(rng, n_points):
    """Residue codes whose effective point count is exactly n_points:
    CA per residue, PP for residues 1..n-2, side-chain points by type."""
    codes = []
    total = 0
    while True:
        c = int(rng.integers(0, 20))
        codes.append(c)
        n = len(codes)
        total = n + max(0, n - 2) + sum(SC_POINTS[x] for x in codes)
        if total >= n_points:
            break
    # trim the excess by converting residues to fewer side-chain points
    excess = total - n_points
    i = 0
    while excess > 0 and i < len(codes):
        k = SC_POINTS[codes[i]]
        if k == 2 and excess >= 1:
            codes[i] = 0 if excess == 1 else 12  # ALA (1 point) or GLY (0)
            excess -= 1 if excess == 1 else 2
        elif k == 1 and excess >= 1:
            codes[i] = 12
            excess -= 1
        i += 1
    n = len(codes)
    assert n + max(0, n - 2) + sum(SC_POINTS[x] for x in codes) == n_points
    return codes


def write_ligand(path, lig_id, coords_confs, types, charges):
    """SDF with the native pose in the atom block and the rest as ENSEMBLE_COORDS."""
    n = coords_confs.shape[1]
    lines = [lig_id, " geauxdock-amd synthetic 3D", "",
             f"{n:3d}  0  0  0  0  0  0  0  0  0999 V2000"]
    for (x, y, z), t in zip(coords_confs[0], types):
        el = LIG_TYPES[t].split(".")[0]
        lines.append(f"{x:10.4f}{y:10.4f}{z:10.4f} {el:<3s} 0  0  0  0  0")
    lines.append("M  END")
    lines += [f">  <MOLID> ({lig_id})", lig_id, ""]
    lines += [f">  <OB_ATOMIC_CHARGES> ({lig_id})", " ".join(f"{c:.10f}" for c in charges), ""]
    lines += [f">  <ENSEMBLE_TOTAL> ({lig_id})", str(len(coords_confs) - 1), ""]
    lines += [f">  <ENSEMBLE_COORDS> ({lig_id})"]
    for conf in coords_confs[1:]:
        lines.append(" ".join(f"{v:.4f}" for v in conf.reshape(-1)))
    lines += ["", f">  <OB_ATOM_TYPES> ({lig_id})", " ".join(LIG_TYPES[t] for t in types), "", "$$$$"]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def _pdb_line(serial, name, resn, resi, xyz):
    return (f"ATOM  {serial % 100000:5d} {name:4s} {resn:3s} A{resi:4d}    "
            f"{xyz[0]:8.3f}{xyz[1]:8.3f}{xyz[2]:8.3f}  1.00  0.00")


def write_protein(path, codes, site_confs):
    """site_confs: (n_conf, n_points, 3) in effective-point load order."""
    n = len(codes)
    # point order of the loader: per residue CA, then PP of the previous
    # residue (from residue 2 on), then side-chain points
    ca_idx, pp_idx, sc_idx = [0] * n, [None] * n, [[] for _ in range(n)]
    k = 0
    for r, c in enumerate(codes):
        ca_idx[r] = k
        k += 1
        if r > 1:
            pp_idx[r - 1] = k
            k += 1
        for _ in range(SC_POINTS[c]):
            sc_idx[r].append(k)
            k += 1
    out = []
    for m, sites in enumerate(site_confs):
        out.append(f"MODEL     {m + 1:4d}")
        serial = 1
        for r, c in enumerate(codes):
            resn, resi = RES3[c], r + 1
            ca = sites[ca_idx[r]]
            npos = sites[pp_idx[r]] if pp_idx[r] is not None else ca
            nxt = pp_idx[r + 1] if r + 1 < n else None
            cpos = sites[nxt] if nxt is not None else ca
            atoms = [(" N  ", npos), (" CA ", ca), (" C  ", cpos), (" O  ", cpos)]
            sc = sc_idx[r]
            if len(sc) == 1:
                atoms.append((" CB ", sites[sc[0]]))
            elif len(sc) == 2:
                atoms += [(" CB ", sites[sc[1]]), (" CG ", sites[sc[1]]), (" CD ", sites[sc[0]])]
            for name, xyz in atoms:
                out.append(_pdb_line(serial, name, resn, resi, xyz))
                serial += 1
        out.append("ENDMDL")
    out.append("END")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
    return ca_idx, pp_idx, sc_idx


def write_ff(path, lig_id, kde, psp, mcs):
    lines = [f"CENTER {POCKET[0]:.10f} {POCKET[1]:.10f} {POCKET[2]:.10f}"]
    for t, (x, y, z) in kde:
        lines.append(f"KDE {LIG_TYPES[t]} {x:.4f} {y:.4f} {z:.4f}")
    for resi, t, v in psp:
        lines.append(f"PSP {resi} {LIG_TYPES[t]} {v:.12f}")
    for tcc, cols in mcs:
        body = " ".join(f"{i + 1} {x:.4f} {y:.4f} {z:.4f}" for i, (x, y, z) in cols)
        lines.append(f"MCS {lig_id} {tcc:.6f} {len(cols)} {body}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def make_ligand(rng, n_atoms, n_conf, sigma, spacing=1.2, radius=6.0):
    nat = _rsa_ball(rng, n_atoms, radius, spacing, POCKET)
    types = rng.integers(0, 24, n_atoms)
    q = rng.normal(0.0, 0.3, n_atoms)
    q -= q.mean()
    confs = [nat]
    tries = 0
    while len(confs) < n_conf:
        c = nat + rng.normal(0.0, sigma, nat.shape)
        if _kabsch_rmsd(c, nat) > 0.15:
            confs.append(c)
        tries += 1
        if tries > 100000:  # e.g. one atom: every conformer superposes exactly
            raise RuntimeError("cannot draw distinct ligand conformers")
    return np.array(confs), types, q


def make_protein(rng, n_points, n_conf, sigma, radius, spacing=2.8, cavity=8.0):
    codes = _residue_plan(rng, n_points)
    sites, _ = _fcc_sites(rng, n_points, radius, cavity, spacing, POCKET)
    confs = [sites]
    for _ in range(n_conf - 1):
        confs.append(sites + rng.normal(0.0, sigma, sites.shape))
    return codes, np.array(confs)


def generate(outdir, lig_id, seed, n_atoms, n_lig_conf, lig_sigma, n_points, n_prt_conf, prt_sigma,
             prt_radius, n_kde, mcs_rows, mcs_cols, prt_spacing=2.8, n_psp=48, protein=None, lig_radius=6.0):
    """Write <lig_id>.sdf/.ff (and the protein .pdb unless `protein` gives an
    existing one); returns the paths.  Ligand atoms fill a ball of
    `lig_radius` around the pocket (large ligands need more than 6 A)."""
    rng = np.random.default_rng(seed)
    os.makedirs(outdir, exist_ok=True)
    confs, types, q = make_ligand(rng, n_atoms, n_lig_conf, lig_sigma, radius=lig_radius)
    sdf = os.path.join(outdir, f"{lig_id}.sdf")
    write_ligand(sdf, lig_id, confs, types, q)
    if protein is None:
        codes, sites = make_protein(rng, n_points, n_prt_conf, prt_sigma, prt_radius, prt_spacing)
        pdb = os.path.join(outdir, f"{lig_id}_prt.pdb")
        _, _, sc_idx = write_protein(pdb, codes, sites)
        protein = {"pdb": pdb, "codes": codes, "sites": sites[0], "sc_idx": sc_idx}
    # KDE points typed like the ligand's atoms, near the native pose
    pick = rng.integers(0, n_atoms, n_kde)
    kde = [(int(types[i]), confs[0][i] + rng.normal(0.0, 1.0, 3)) for i in pick]
    # PSP entries for side-chain residues nearest the pocket
    codes, sites, sc_idx = protein["codes"], protein["sites"], protein["sc_idx"]
    cand = [(float(np.linalg.norm(sites[s[0]] - POCKET)), r) for r, s in enumerate(sc_idx) if s]
    cand.sort()
    psp = []
    for _, r in cand[: max(0, n_psp // 3)]:
        for t in rng.choice(24, 3, replace=False):
            psp.append((r + 1, int(t), float(rng.normal(0.0, 1.0))))
    mcs = []
    ncol = min(mcs_cols, n_atoms)
    for _ in range(mcs_rows):
        idx = rng.choice(n_atoms, ncol, replace=False)
        cols = [(int(i), confs[0][i] + rng.normal(0.0, 0.8, 3)) for i in idx]
        mcs.append((float(rng.uniform(0.05, 1.0)), cols))
    ff = os.path.join(outdir, f"{lig_id}.ff")
    write_ff(ff, lig_id, kde, psp, mcs)
    return {"pdb": protein["pdb"], "sdf": sdf, "ff": ff, "lig_id": lig_id, "protein": protein}


# ---------------------------------------------------------------------------
# the benchmark workloads
# ---------------------------------------------------------------------------
CONFIGS = {
    # one ligand of 24 atoms x 6 conformations, 1,500-point protein, KDE 200,
    # MCS 8 rows, 8 temperatures 0.3-2.0: 48 replicas x 2,000 steps
    "c1": dict(n_atoms=24, n_lig_conf=6, lig_sigma=0.3, n_points=1500, n_prt_conf=1, prt_sigma=0.0,
               prt_radius=20.0, n_kde=200, mcs_rows=8, mcs_cols=12, n_tmp=8, t_lo=0.3, t_hi=2.0,
               steps=2000, exchange=0),
    # 48 atoms x 20 conformations, 4,000 points x 11 conformations, KDE 1,000,
    # MCS 40 x 24, 16 temperatures: 3,520 replicas x 10,000 steps, exchange / 100
    "c2": dict(n_atoms=48, n_lig_conf=20, lig_sigma=0.5, n_points=4000, n_prt_conf=11, prt_sigma=0.5,
               prt_radius=28.0, n_kde=1000, mcs_rows=40, mcs_cols=24, n_tmp=16, t_lo=0.3, t_hi=2.0,
               steps=10000, exchange=100),
    # 150 atoms x 64 conformations, 20,000 points x 32 conformations in a
    # 35 A sphere (spacing relaxed to what fits: 2.3 A), KDE 5,000, MCS 100
    # rows, 32 temperatures: 65,536 replicas x 5,000 steps, exchange / 50
    "c5": dict(n_atoms=150, n_lig_conf=64, lig_sigma=0.5, n_points=20000, n_prt_conf=32, prt_sigma=0.5,
               prt_radius=35.0, prt_spacing=2.3, n_kde=5000, mcs_rows=100, mcs_cols=24, n_tmp=32, t_lo=0.3,
               t_hi=2.0, steps=5000, exchange=50),
}


def default_dir():
    return os.environ.get("GD_SYNTH_DIR", "/tmp/gd_synth")


def config_files(name, seed=2024, outdir=None, **override):
    """Generate (once, cached by content hash) the files of a named workload."""
    cfg = dict(CONFIGS[name])
    cfg.update(override)
    key = hashlib.sha1(json.dumps([name, seed, cfg], sort_keys=True).encode()).hexdigest()[:12]
    d = os.path.join(outdir or default_dir(), f"{name}_{key}")
    meta = os.path.join(d, "meta.json")
    if os.path.exists(meta):
        with open(meta) as f:
            return json.load(f)
    gen = {k: cfg[k] for k in ("n_atoms", "n_lig_conf", "lig_sigma", "n_points", "n_prt_conf", "prt_sigma",
                               "prt_radius", "n_kde", "mcs_rows", "mcs_cols")}
    gen["prt_spacing"] = cfg.get("prt_spacing", 2.8)
    out = generate(d, f"{name}lig", seed, **gen)
    info = {"pdb": out["pdb"], "sdf": out["sdf"], "ff": out["ff"], "lig_id": out["lig_id"], "cfg": cfg,
            "seed": seed}
    with open(meta, "w") as f:
        json.dump(info, f)
    return info


def load_config(name, seed=2024, lib=None, **override):
    """Generate + load + stage a named workload; returns (StagedComplex, info)."""
    from . import lib as L
    info = config_files(name, seed, **override)
    c = info["cfg"]
    raw = L.load_job(info["pdb"], info["sdf"], info["ff"], info["lig_id"], n_tmp=c["n_tmp"],
                     temp_high=c["t_hi"], temp_low=c["t_lo"], lib=lib)
    return raw.stage(), info


# ---------------------------------------------------------------------------
# many-ligand workloads: batched screening and the serving request stream
# ---------------------------------------------------------------------------
# screening: one resident 5,000-point protein in 6 conformations; 1,024
# ligands whose heavy-atom counts are lognormal (median 32, sigma 0.5)
# clipped to 12-96, 8 conformations each, 8 temperatures: ~393k replicas
SCREEN = dict(n_ligands=1024, median_atoms=32, log_sigma=0.5, min_atoms=12, max_atoms=96, n_lig_conf=8,
              lig_sigma=0.4, n_points=5000, n_prt_conf=6, prt_sigma=0.5, prt_radius=30.0, n_tmp=8, t_lo=0.3,
              t_hi=2.0, steps=2000)
# serving: 8,000-point protein in 4 conformations; requests of 10-120 atoms
# (uniform), 4-32 conformations, 8 temperatures, 500-5,000 steps
# (log-uniform), Poisson arrivals at 50, 200 and 800 per second
SERVE = dict(n_points=8000, n_prt_conf=4, prt_sigma=0.5, prt_radius=35.0, atoms=(10, 120), confs=(4, 32),
             n_tmp=8, t_lo=0.3, t_hi=2.0, steps=(500, 5000), rates=(50.0, 200.0, 800.0),
             loads=(0.8, 0.5), lig_sigma=0.4)


def _ligand_files(d, lig_id, rng_seed, n_atoms, n_conf, lig_sigma, protein):
    """One ligand (.sdf/.ff) of a many-ligand workload against `protein`;
    KDE 8 points per atom, MCS 8 rows of up to 12 columns."""
    return generate(d, lig_id, seed=rng_seed, n_atoms=n_atoms, n_lig_conf=n_conf, lig_sigma=lig_sigma,
                    n_points=0, n_prt_conf=0, prt_sigma=0.0, prt_radius=0.0, n_kde=8 * n_atoms, mcs_rows=8,
                    mcs_cols=12, protein=protein)


def _protein_only(d, name, seed, n_points, n_conf, sigma, radius):
    rng = np.random.default_rng(seed)
    codes, sites = make_protein(rng, n_points, n_conf, sigma, radius)
    pdb = os.path.join(d, f"{name}_prt.pdb")
    _, _, sc_idx = write_protein(pdb, codes, sites)
    return {"pdb": pdb, "codes": codes, "sites": sites[0], "sc_idx": sc_idx}


def many_files(kind, seed=2024, outdir=None, n=None):
    """Files of the screening library (kind "screen") or of the serving
    request stream (kind "serve"), generated once and cached by content;
    returns {"pdb", "ligands": [{"sdf", "ff", "lig_id", "n_atoms", "n_conf",
    "steps"?}], "cfg"}."""
    cfg = dict(SCREEN if kind == "screen" else SERVE)
    if n is not None:
        cfg["n_ligands"] = n
    key = hashlib.sha1(json.dumps([kind, seed, cfg], sort_keys=True).encode()).hexdigest()[:12]
    d = os.path.join(outdir or default_dir(), f"{kind}_{key}")
    meta = os.path.join(d, "meta.json")
    if os.path.exists(meta):
        with open(meta) as f:
            return json.load(f)
    os.makedirs(d, exist_ok=True)
    rng = np.random.default_rng(seed)
    prot = _protein_only(d, kind, seed + 1, cfg["n_points"], cfg["n_prt_conf"], cfg["prt_sigma"], cfg["prt_radius"])
    ligs = []
    count = cfg.get("n_ligands", 0) if kind == "screen" else int(n or 0)
    for i in range(count):
        if kind == "screen":
            na = int(np.clip(round(cfg["median_atoms"] * math.exp(cfg["log_sigma"] * rng.normal())),
                             cfg["min_atoms"], cfg["max_atoms"]))
            nc = cfg["n_lig_conf"]
            steps = cfg["steps"]
        else:
            na = int(rng.integers(cfg["atoms"][0], cfg["atoms"][1] + 1))
            nc = int(rng.integers(cfg["confs"][0], cfg["confs"][1] + 1))
            lo, hi = math.log(cfg["steps"][0]), math.log(cfg["steps"][1])
            steps = int(round(math.exp(rng.uniform(lo, hi))))
        lig_id = f"{kind}{i:04d}"
        out = _ligand_files(d, lig_id, seed * 7919 + i, na, nc, cfg["lig_sigma"], prot)
        ligs.append({"sdf": out["sdf"], "ff": out["ff"], "lig_id": lig_id, "n_atoms": na, "n_conf": nc,
                     "steps": steps})
    info = {"pdb": prot["pdb"], "ligands": ligs, "cfg": cfg, "seed": seed}
    with open(meta, "w") as f:
        json.dump(info, f)
    return info


def load_many(info, threads=16, lib=None):
    """Load and stage every ligand of many_files() against its protein, in
    parallel (the loader releases the interpreter while it parses)."""
    import concurrent.futures as cf
    from . import lib as L
    c = info["cfg"]

    def one(lg):
        raw = L.load_job(info["pdb"], lg["sdf"], lg["ff"], lg["lig_id"], n_tmp=c["n_tmp"], temp_high=c["t_hi"],
                         temp_low=c["t_lo"], lib=lib)
        return raw.stage()

    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(one, info["ligands"]))
