// Input parsers: SDF ligands with conformer ensembles, PDB proteins reduced
// to coarse-grained effective points, the ".ff" pocket file (KDE / PSP / MCS
// lines), the force-field "paras" file and the one-column weight and
// normalisation files.
//
// Behaviour follows the reference loader (src/frontend/load.cpp) field by
// field -- fixed SDF columns, the effective-point construction, float
// accumulation order of the point averages, the conformer RMSD filter -- with
// these deliberate differences:
//   * every job starts from zeroed state (the reference accumulates the
//     ligand centre and PSP table across jobs, load.cpp:226-238, :1098);
//   * capacities are run-time sizes; the PSP residue index is not limited to
//     511 (the reference's swapped bounds checks, load.cpp:1085-1094);
//   * errors are returned as messages instead of exit(EXIT_FAILURE);
//   * unknown residue / point / ligand type names in parameter files are
//     errors instead of out-of-range table writes.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>

#include "codes.hpp"
#include "job.hpp"

namespace gd {

namespace {

std::vector<std::string> split_ws(const std::string &s)
{
    std::vector<std::string> out;
    std::istringstream is(s);
    std::string tok;
    while (is >> tok) out.push_back(tok);
    return out;
}

bool read_lines(const std::string &path, std::vector<std::string> &lines, std::string &err,
                const char *stop_at = nullptr)
{
    std::ifstream f(path.c_str());
    if (!f.is_open()) {
        err = "cannot open " + path;
        return false;
    }
    std::string line;
    while (std::getline(f, line)) {
        lines.push_back(line);
        if (stop_at && line == stop_at) break;
    }
    return true;
}

// substr that tolerates short lines, like std::string::substr past the end
std::string field(const std::string &s, size_t pos, size_t len)
{
    if (pos >= s.size()) return std::string();
    return s.substr(pos, len);
}

float to_float_atof(const std::string &s) { return (float)std::atof(s.c_str()); }

} // namespace

// ---------------------------------------------------------------------------
// SDF ligand (src/frontend/load.cpp:90-397)
// ---------------------------------------------------------------------------
bool load_ligand(const std::string &path, int max_conf, RawJob &out, std::string &err)
{
    std::vector<std::string> lines;
    if (!read_lines(path, lines, err, "$$$$")) return false;
    const int nlines = (int)lines.size();
    if (nlines <= 10) {
        err = "ligand file too small: " + path;
        return false;
    }
    const int lna = std::atoi(field(lines[3], 0, 3).c_str());
    const int lnb = std::atoi(field(lines[3], 3, 3).c_str());
    if (lna <= 0) {
        err = "ligand atom count is illegal in " + path;
        return false;
    }
    if (4 + lna > nlines) {
        err = "truncated atom block in " + path;
        return false;
    }

    std::vector<int> types(lna, GD_BADKDE);
    std::vector<float> charges(lna, 0.0f);
    std::vector<std::vector<float>> ens; // raw ensemble conformers, [conf][3*lna]
    bool have_types = false;
    for (int i = 4 + lna + lnb; i < nlines; ++i) {
        const std::string &ln = lines[i];
        if (ln.find("MOLID") != std::string::npos) {
            if (i + 1 < nlines) out.lig_name = lines[i + 1];
        } else if (ln.find("OB_ATOM_TYPES") != std::string::npos && i + 1 < nlines) {
            std::vector<std::string> tok = split_ws(lines[i + 1]);
            for (int k = 0; k < (int)tok.size() && k < lna; ++k) types[k] = ligand_code(tok[k]);
            have_types = true;
        } else if (ln.find("OB_ATOMIC_CHARGES") != std::string::npos && i + 1 < nlines) {
            std::vector<std::string> tok = split_ws(lines[i + 1]);
            for (int k = 0; k < (int)tok.size() && k < lna; ++k) charges[k] = to_float_atof(tok[k]);
        } else if (ln.find("ENSEMBLE_COORDS") != std::string::npos) {
            // one conformer per line until an empty line (load.cpp:192-210)
            while (i + 1 < nlines && !lines[i + 1].empty() && (int)ens.size() < max_conf) {
                ++i;
                std::vector<std::string> tok = split_ws(lines[i]);
                std::vector<float> c(3 * lna, 0.0f);
                for (int k = 0; k < (int)tok.size() && k < 3 * lna; ++k) c[k] = to_float_atof(tok[k]);
                ens.push_back(c);
            }
        }
    }
    if (!have_types) {
        err = "no OB_ATOM_TYPES block in " + path;
        return false;
    }
    for (int k = 0; k < lna; ++k) {
        if (types[k] < 0 || types[k] >= GD_MAXTP2) {
            err = "ligand atom type outside the 24-type table in " + path;
            return false;
        }
    }

    // native pose from the fixed-column atom block; centroid accumulated in
    // float in atom order, then divided (load.cpp:219-246)
    std::vector<float> nat(3 * lna);
    float center[3] = {0.0f, 0.0f, 0.0f};
    for (int a = 0; a < lna; ++a) {
        const std::string &ln = lines[4 + a];
        nat[3 * a + 0] = to_float_atof(field(ln, 0, 10));
        nat[3 * a + 1] = to_float_atof(field(ln, 10, 10));
        nat[3 * a + 2] = to_float_atof(field(ln, 20, 10));
        for (int d = 0; d < 3; ++d) center[d] += nat[3 * a + d];
    }
    for (int d = 0; d < 3; ++d) center[d] /= (float)lna;
    for (int a = 0; a < lna; ++a)
        for (int d = 0; d < 3; ++d) nat[3 * a + d] -= center[d];

    // conformer filter: keep ensemble members whose best-superposition RMSD
    // to the native pose exceeds 0.1 A, each centred on its own centroid
    // (load.cpp:303-395)
    std::vector<std::vector<float>> kept;
    kept.push_back(nat);
    std::vector<float> rx(lna), ry(lna), rz(lna), mx(lna), my(lna), mz(lna);
    for (int a = 0; a < lna; ++a) {
        rx[a] = nat[3 * a];
        ry[a] = nat[3 * a + 1];
        rz[a] = nat[3 * a + 2];
    }
    for (size_t e = 0; e < ens.size(); ++e) {
        std::vector<float> mob = ens[e];
        float mc[3] = {0.0f, 0.0f, 0.0f};
        for (int a = 0; a < lna; ++a)
            for (int d = 0; d < 3; ++d) mc[d] += mob[3 * a + d];
        for (int d = 0; d < 3; ++d) mc[d] /= (float)lna;
        for (int a = 0; a < lna; ++a)
            for (int d = 0; d < 3; ++d) mob[3 * a + d] -= mc[d];
        for (int a = 0; a < lna; ++a) {
            mx[a] = mob[3 * a];
            my[a] = mob[3 * a + 1];
            mz[a] = mob[3 * a + 2];
        }
        // rms1 is the summed squared deviation, as u3b returns it
        const double r = superposed_rmsd(mx.data(), my.data(), mz.data(), rx.data(), ry.data(),
                                         rz.data(), lna);
        const float rms1 = (float)(r * r * lna);
        const float rms2 = std::sqrt(rms1 / (float)lna);
        if (rms2 > 0.1) kept.push_back(mob);
        if ((int)kept.size() >= max_conf) break;
    }

    const int n_lig = (int)kept.size();
    out.lig_x.assign(n_lig * lna, 0.0f);
    out.lig_y.assign(n_lig * lna, 0.0f);
    out.lig_z.assign(n_lig * lna, 0.0f);
    for (int c = 0; c < n_lig; ++c)
        for (int a = 0; a < lna; ++a) {
            out.lig_x[c * lna + a] = kept[c][3 * a];
            out.lig_y[c * lna + a] = kept[c][3 * a + 1];
            out.lig_z[c * lna + a] = kept[c][3 * a + 2];
        }
    out.lig_t = types;
    out.lig_c = charges;
    out.view.n_lig = n_lig;
    out.view.lig_natom = lna;
    for (int d = 0; d < 3; ++d) out.view.lig_center[d] = center[d];
    return true;
}

// ---------------------------------------------------------------------------
// PDB protein -> effective points (src/frontend/load.cpp:576-970)
// ---------------------------------------------------------------------------
bool load_protein(const std::string &path, RawJob &out, std::string &err)
{
    std::vector<std::string> lines;
    if (!read_lines(path, lines, err)) return false;

    auto is_atom = [](const std::string &s) { return s.size() > 53 && s.compare(0, 6, "ATOM  ") == 0; };
    auto is_endmdl = [](const std::string &s) { return s.size() <= 53 && s.compare(0, 6, "ENDMDL") == 0; };

    // topology from the first model: CA point, peptide-plate point of the
    // previous residue from the third residue on, side-chain points
    std::vector<int> pt_t, pt_c, pt_d, pt_r, seq3;
    for (const std::string &ln : lines) {
        if (is_atom(ln)) {
            if (ln.compare(12, 4, " CA ") != 0) continue;
            const int res = residue_code3(ln.substr(17, 3));
            if (res < 0) {
                err = "unknown residue " + ln.substr(17, 3) + " in " + path;
                return false;
            }
            const int pnr = (int)seq3.size();
            seq3.push_back(std::atoi(ln.substr(22, 4).c_str()));
            pt_t.push_back(0); pt_c.push_back(0); pt_d.push_back(res); pt_r.push_back(pnr);
            if (pnr > 1) {
                pt_t.push_back(1); pt_c.push_back(1); pt_d.push_back(res); pt_r.push_back(pnr - 1);
            }
            const SideChain sc = side_chain_points(res);
            for (int k = 0; k < sc.n; ++k) {
                pt_t.push_back(sc.type[k]); pt_c.push_back(sc.cls[k]); pt_d.push_back(res); pt_r.push_back(pnr);
            }
        } else if (is_endmdl(ln)) {
            break;
        }
    }
    const int npoint = (int)pt_t.size();
    const int nres = (int)seq3.size();
    int nconf = 0;
    for (const std::string &ln : lines)
        if (is_endmdl(ln)) ++nconf;
    if (npoint <= 0 || nconf <= 0) {
        err = "no effective points or no ENDMDL-terminated model in " + path;
        return false;
    }

    // parsed ATOM records, in file order, with the model they belong to
    struct Atom { int model, resid; char name[5]; float x, y, z; };
    std::vector<Atom> atoms;
    {
        int model = 0;
        for (const std::string &ln : lines) {
            if (is_atom(ln)) {
                Atom a;
                a.model = model;
                a.resid = std::atoi(ln.substr(22, 4).c_str());
                std::memcpy(a.name, ln.data() + 12, 4);
                a.name[4] = 0;
                a.x = to_float_atof(ln.substr(30, 8));
                a.y = to_float_atof(ln.substr(38, 8));
                a.z = to_float_atof(ln.substr(46, 8));
                atoms.push_back(a);
            } else if (is_endmdl(ln)) {
                ++model;
            }
        }
    }

    out.prt_x.assign((size_t)nconf * npoint, 0.0f);
    out.prt_y.assign((size_t)nconf * npoint, 0.0f);
    out.prt_z.assign((size_t)nconf * npoint, 0.0f);

    // atoms of each residue serial, in file order: a point only ever takes
    // atoms of its own residue (and, for a peptide plate, the previous one)
    std::map<int, std::vector<int>> by_res;
    for (int i = 0; i < (int)atoms.size(); ++i) by_res[atoms[i].resid].push_back(i);
    const std::vector<int> none;
    auto atoms_of = [&](int resid) -> const std::vector<int> & {
        auto it = by_res.find(resid);
        return it == by_res.end() ? none : it->second;
    };

    // each point averages its atoms within each model; float sums in file
    // order (load.cpp:715-852)
    std::vector<float> sx(nconf), sy(nconf), sz(nconf), sn(nconf);
    std::vector<int> cand;
    for (int p = 0; p < npoint; ++p) {
        std::fill(sx.begin(), sx.end(), 0.0f);
        std::fill(sy.begin(), sy.end(), 0.0f);
        std::fill(sz.begin(), sz.end(), 0.0f);
        std::fill(sn.begin(), sn.end(), 0.0f);
        const int r = pt_r[p];
        const int resid = seq3[r];
        const int prev = (pt_t[p] == 1 && r > 0) ? seq3[r - 1] : -1;
        const PointAtomRule rule = point_atom_rule(pt_t[p]);
        const std::vector<int> &own = atoms_of(resid);
        if (rule == RULE_PP && prev != resid) {
            const std::vector<int> &pv = atoms_of(prev);
            cand.resize(own.size() + pv.size());
            std::merge(own.begin(), own.end(), pv.begin(), pv.end(), cand.begin());
        } else {
            cand = own;
        }
        for (int ai : cand) {
            const Atom &a = atoms[ai];
            if (a.model >= nconf) continue;
            const std::string nm(a.name);
            bool take = false;
            if (a.resid == resid) {
                const bool bb = nm == " N  " || nm == " CA " || nm == " C  " || nm == " O  ";
                switch (rule) {
                case RULE_CA: take = nm == " CA "; break;
                case RULE_SC_ALL: take = !bb; break;
                case RULE_SC_DISTAL: take = !bb && nm != " CB " && nm != " CG "; break;
                case RULE_SC_PROXIMAL: take = nm == " CB " || nm == " CG "; break;
                case RULE_PP: break;
                }
            }
            if (rule == RULE_PP && r > 0 && r < nres - 1) {
                if ((a.resid == resid && nm == " N  ") || (a.resid == prev && (nm == " C  " || nm == " O  ")))
                    take = true;
            }
            if (take) {
                sx[a.model] += a.x;
                sy[a.model] += a.y;
                sz[a.model] += a.z;
                sn[a.model] += 1.0f;
            }
        }
        for (int m = 0; m < nconf; ++m) {
            if (sn[m] > 0.0f) {
                out.prt_x[(size_t)m * npoint + p] = sx[m] / sn[m];
                out.prt_y[(size_t)m * npoint + p] = sy[m] / sn[m];
                out.prt_z[(size_t)m * npoint + p] = sz[m] / sn[m];
            }
        }
    }
    out.prt_t = pt_t;
    out.prt_c = pt_c;
    out.prt_d = pt_d;
    out.prt_r = pt_r;
    out.prt_seq3 = seq3;
    out.view.n_prt = nconf;
    out.view.prt_npoint = npoint;
    out.view.prt_nres = nres;
    return true;
}

// ---------------------------------------------------------------------------
// ".ff" pocket file: KDE, PSP and MCS lines (src/frontend/load.cpp:994-1171)
// ---------------------------------------------------------------------------
bool load_lhm(const std::string &path, const std::string &lig_id, RawJob &out, std::string &err)
{
    std::vector<std::string> lines;
    if (!read_lines(path, lines, err)) return false;
    const int lna = out.view.lig_natom;
    out.mcs_rowptr.assign(1, 0);
    for (const std::string &ln : lines) {
        if (ln.size() <= 3) continue;
        const std::string tag = ln.substr(0, 3);
        if (tag == "KDE") {
            std::vector<std::string> tok = split_ws(ln);
            while (tok.size() < 5) tok.push_back("");
            const int t = ligand_code(tok[1]);
            out.kde_t.push_back(t);
            out.kde_x.push_back(to_float_atof(tok[2]));
            out.kde_y.push_back(to_float_atof(tok[3]));
            out.kde_z.push_back(to_float_atof(tok[4]));
        } else if (tag == "PSP") {
            std::vector<std::string> tok = split_ws(ln);
            while (tok.size() < 4) tok.push_back("");
            const int res = std::atoi(tok[1].c_str());
            const int lt = ligand_code(tok[2]);
            if (res < 0) {
                err = "negative PSP residue index in " + path;
                return false;
            }
            if (lt >= GD_MAXTP2) continue; // no staged ligand atom can carry this type
            out.psp_res.push_back(res);
            out.psp_lig.push_back(lt);
            out.psp_val.push_back(to_float_atof(tok[3]));
        } else if (tag == "MCS") {
            std::vector<std::string> tok = split_ws(ln);
            if (tok.size() < 4 || tok[1] != lig_id) continue;
            const float tcc = to_float_atof(tok[2]);
            const int ncol = std::atoi(tok[3].c_str());
            if (ncol <= 0 || (int)tok.size() < 4 + 4 * ncol) {
                err = "malformed MCS row in " + path;
                return false;
            }
            for (int k = 0; k < ncol; ++k) {
                const int loc = std::atoi(tok[4 + 4 * k].c_str()) - 1; // 1-based in the file
                if (loc < 0 || loc >= lna) {
                    err = "MCS atom index out of range in " + path;
                    return false;
                }
                out.mcs_i.push_back(loc);
                out.mcs_x.push_back(to_float_atof(tok[5 + 4 * k]));
                out.mcs_y.push_back(to_float_atof(tok[6 + 4 * k]));
                out.mcs_z.push_back(to_float_atof(tok[7 + 4 * k]));
            }
            out.mcs_tcc.push_back(tcc);
            out.mcs_rowptr.push_back((int)out.mcs_i.size());
        }
    }
    if (out.mcs_tcc.empty()) {
        err = "found no MCS row matching ligand id \"" + lig_id + "\" in " + path;
        return false;
    }
    if (out.kde_t.empty()) {
        err = "no KDE points in " + path;
        return false;
    }
    out.view.kde_npoint = (int)out.kde_t.size();
    out.view.mcs_nrow = (int)out.mcs_tcc.size();
    out.view.psp_n = (int)out.psp_res.size();
    return true;
}

// ---------------------------------------------------------------------------
// force-field parameters (src/frontend/load.cpp:1183-1370)
// ---------------------------------------------------------------------------
bool load_enepara(const std::string &path, gd_EnePara0 &ep, std::string &err)
{
    std::vector<std::string> lines;
    if (!read_lines(path, lines, err)) return false;
    auto bad = [&](const std::string &ln) {
        err = "unknown type code in parameter line \"" + ln + "\" of " + path;
        return false;
    };
    for (const std::string &ln : lines) {
        if (ln.size() <= 3) continue;
        const std::string tag = ln.substr(0, 3);
        std::vector<std::string> tok = split_ws(ln);
        while (tok.size() < 11) tok.push_back("");
        if (tag == "VDW" || tag == "PMF" || tag == "HDB") {
            const int p = point_code(tok[1]);
            const int l = ligand_code(tok[2]);
            if (p < 0 || l >= GD_MAXTP2) return bad(ln);
            float(*tab)[GD_MAXTP2][2] = tag == "VDW" ? ep.vdw : (tag == "PMF" ? ep.pmf : ep.hdb);
            tab[p][l][0] = to_float_atof(tok[3]);
            tab[p][l][1] = to_float_atof(tok[4]);
        } else if (tag == "PLJ") {
            for (int i = 0; i < 3; ++i) ep.lj[i] = to_float_atof(tok[i + 1]);
        } else if (tag == "ELE") {
            // "XC" names a residue's CA charge (stored at 30 + residue code),
            // anything else an effective point type
            int idx;
            if (tok[1].size() >= 2 && tok[1][1] == 'C') {
                const int r = residue_code1(tok[1].substr(0, 1));
                if (r < 0) return bad(ln);
                idx = r + 30;
            } else {
                idx = point_code(tok[1]);
                if (idx < 0) return bad(ln);
            }
            ep.ele[idx] = to_float_atof(tok[2]);
        } else if (tag == "PEL") {
            for (int i = 0; i < 2; ++i) ep.el[i] = to_float_atof(tok[i + 1]);
        } else if (tag == "HPP") {
            const int r = residue_code1(tok[1]);
            if (r < 0) return bad(ln);
            ep.hpp[r] = to_float_atof(tok[2]);
        } else if (tag == "HPL") {
            const int l = ligand_code(tok[1]);
            if (l >= GD_MAXTP2) return bad(ln);
            ep.hpl[l][0] = to_float_atof(tok[2]);
            ep.hpl[l][1] = to_float_atof(tok[3]);
        } else if (tag == "WEI") {
            // nine weights follow; the tenth slot (the total) stays zero
            for (int i = 0; i < GD_MAXWEI - 1; ++i) ep.w[i] = to_float_atof(tok[i + 1]);
        } else if (tag == "KDE") {
            ep.kde = to_float_atof(tok[1]);
        }
    }

    // rescale residue hydrophobicity to [-1, 1] (load.cpp:1343-1368)
    float lo = 1e6f, hi = -1e6f;
    for (int i = 0; i < GD_MAXTP4; ++i) {
        if (ep.hpp[i] < lo) lo = ep.hpp[i];
        if (ep.hpp[i] > hi) hi = ep.hpp[i];
    }
    for (int i = 0; i < GD_MAXTP4; ++i) {
        float v = (float)((double)((ep.hpp[i] - lo) / (hi - lo)) * 2.0 - 1.0);
        if (v < -1.0f) v = -1.0f;
        if (v > 1.0f) v = 1.0f;
        ep.hpp[i] = v;
    }
    return true;
}

// one float per line, read from the first 30 characters of each line with
// stream extraction (src/frontend/load.cpp:1372-1435, 1470-1501)
bool load_column_file(const std::string &path, float *dst, int cap, std::string &err)
{
    std::vector<std::string> lines;
    if (!read_lines(path, lines, err)) return false;
    if ((int)lines.size() > cap) {
        err = "more than " + std::to_string(cap) + " lines in " + path;
        return false;
    }
    for (size_t i = 0; i < lines.size(); ++i) {
        std::istringstream is(lines[i].substr(0, 30));
        float v = 0.0f;
        is >> v;
        dst[i] = v;
    }
    return true;
}

// ---------------------------------------------------------------------------
// job line (src/frontend/frontend.cpp:70-123)
// ---------------------------------------------------------------------------
bool parse_job_line(const std::string &line, LoadArgs &a, std::string &err)
{
    std::vector<std::string> tok;
    {
        std::string cur;
        for (char ch : line) {
            if (ch == ',') {
                tok.push_back(cur);
                cur.clear();
            } else if (ch != '\n' && ch != '\r') {
                cur.push_back(ch);
            }
        }
        tok.push_back(cur);
    }
    if (tok.size() < 17) {
        err = "job line has fewer than 17 columns";
        return false;
    }
    a.data_id = std::atoi(tok[1].c_str());
    a.prt_id = tok[2];
    a.prt_path = tok[3];
    a.lig_id = tok[4];
    a.lig_path = tok[5];
    a.ff_path = tok[6];
    a.weight_path = tok[7];
    a.enepara_path = tok[8];
    a.nora_path = tok[9];
    a.norb_path = tok[10];
    a.n_tmp = std::atoi(tok[11].c_str());
    a.temp_high = (float)std::atof(tok[12].c_str());
    a.temp_low = (float)std::atof(tok[13].c_str());
    a.tras_scale = (float)std::atof(tok[14].c_str());
    a.rot_scale = (float)std::atof(tok[15].c_str());
    a.n_dump = std::atoi(tok[16].c_str());
    return true;
}

bool load_raw(const LoadArgs &args, RawJob &out, std::string &err)
{
    out = RawJob();
    if (!load_protein(args.prt_path, out, err)) return false;
    std::memset(&out.view.enepara0, 0, sizeof(out.view.enepara0));
    if (!load_enepara(args.enepara_path, out.view.enepara0, err)) return false;
    if (!load_column_file(args.weight_path, out.view.enepara0.w, GD_MAXWEI, err)) return false;
    if (!load_column_file(args.nora_path, out.view.enepara0.a_para, GD_MAXWEI, err)) return false;
    if (!load_column_file(args.norb_path, out.view.enepara0.b_para, GD_MAXWEI, err)) return false;
    if (!load_ligand(args.lig_path, args.max_conf_lig, out, err)) return false;
    if (!load_lhm(args.ff_path, args.lig_id, out, err)) return false;
    if (args.n_tmp <= 0) {
        err = "temperature count must be positive";
        return false;
    }
    out.view.n_tmp = args.n_tmp;
    out.view.temp_high = args.temp_high;
    out.view.temp_low = args.temp_low;
    for (int i = 0; i < 3; ++i) {
        out.view.move_scale[i] = args.tras_scale;
        out.view.move_scale[i + 3] = args.rot_scale;
    }
    out.view.steps_n_dump = args.n_dump;
    out.bind();
    return true;
}

void RawJob::bind()
{
    view.lig_x = lig_x.data();
    view.lig_y = lig_y.data();
    view.lig_z = lig_z.data();
    view.lig_t = lig_t.data();
    view.lig_c = lig_c.data();
    view.prt_x = prt_x.data();
    view.prt_y = prt_y.data();
    view.prt_z = prt_z.data();
    view.prt_t = prt_t.data();
    view.prt_c = prt_c.data();
    view.prt_d = prt_d.data();
    view.prt_r = prt_r.data();
    view.prt_seq3 = prt_seq3.data();
    view.psp_res = psp_res.data();
    view.psp_lig = psp_lig.data();
    view.psp_val = psp_val.data();
    view.kde_t = kde_t.data();
    view.kde_x = kde_x.data();
    view.kde_y = kde_y.data();
    view.kde_z = kde_z.data();
    view.mcs_tcc = mcs_tcc.data();
    view.mcs_rowptr = mcs_rowptr.data();
    view.mcs_i = mcs_i.data();
    view.mcs_x = mcs_x.data();
    view.mcs_y = mcs_y.data();
    view.mcs_z = mcs_z.data();
}

// ---------------------------------------------------------------------------
// best-superposition RMSD (Horn 1987 unit-quaternion method)
// ---------------------------------------------------------------------------
namespace {
// largest eigenvalue of a symmetric 4x4 matrix by cyclic Jacobi rotations
double max_eig4(double a[4][4])
{
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int i = 0; i < 4; ++i)
            for (int j = i + 1; j < 4; ++j) off += a[i][j] * a[i][j];
        if (off < 1e-30) break;
        for (int p = 0; p < 4; ++p)
            for (int q = p + 1; q < 4; ++q) {
                if (std::fabs(a[p][q]) < 1e-300) continue;
                const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 4; ++k) {
                    const double akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - s * akq;
                    a[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 4; ++k) {
                    const double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - s * aqk;
                    a[q][k] = s * apk + c * aqk;
                }
            }
    }
    double m = a[0][0];
    for (int i = 1; i < 4; ++i) m = std::max(m, a[i][i]);
    return m;
}
} // namespace

double superposed_rmsd(const float *ax, const float *ay, const float *az,
                       const float *bx, const float *by, const float *bz, int n)
{
    double ca[3] = {0, 0, 0}, cb[3] = {0, 0, 0};
    for (int i = 0; i < n; ++i) {
        ca[0] += ax[i]; ca[1] += ay[i]; ca[2] += az[i];
        cb[0] += bx[i]; cb[1] += by[i]; cb[2] += bz[i];
    }
    for (int d = 0; d < 3; ++d) {
        ca[d] /= n;
        cb[d] /= n;
    }
    double s[3][3] = {{0}}, ga = 0.0, gb = 0.0;
    for (int i = 0; i < n; ++i) {
        const double a[3] = {ax[i] - ca[0], ay[i] - ca[1], az[i] - ca[2]};
        const double b[3] = {bx[i] - cb[0], by[i] - cb[1], bz[i] - cb[2]};
        for (int r = 0; r < 3; ++r) {
            ga += a[r] * a[r];
            gb += b[r] * b[r];
            for (int c = 0; c < 3; ++c) s[r][c] += a[r] * b[c];
        }
    }
    double k[4][4];
    k[0][0] = s[0][0] + s[1][1] + s[2][2];
    k[0][1] = s[1][2] - s[2][1];
    k[0][2] = s[2][0] - s[0][2];
    k[0][3] = s[0][1] - s[1][0];
    k[1][1] = s[0][0] - s[1][1] - s[2][2];
    k[1][2] = s[0][1] + s[1][0];
    k[1][3] = s[2][0] + s[0][2];
    k[2][2] = -s[0][0] + s[1][1] - s[2][2];
    k[2][3] = s[1][2] + s[2][1];
    k[3][3] = -s[0][0] - s[1][1] + s[2][2];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < i; ++j) k[i][j] = k[j][i];
    const double lmax = max_eig4(k);
    double e = (ga + gb - 2.0 * lmax) / n;
    if (e < 0.0) e = 0.0;
    return std::sqrt(e);
}

} // namespace gd
