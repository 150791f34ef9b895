// Staging: loader output (gd_Raw) -> the tables the Monte Carlo kernels read
// (gd_Complex).  Each step restates one reference builder:
//   protein points sorted by class, cdc/seq3r/ele/hpp derived
//                         OptimizeProtein, util_optimize_soa.cpp:216-287
//   ligand x2 (file order) and x3 (sorted by type) sets
//                         OptimizeLigand, util_optimize_soa.cpp:47-205
//   KDE points sorted by type                 OptimizeKde, :343-390
//   PSP table [ligand type][residue serial]    OptimizePsp, :298-305
//   MCS rows in ELL form                       OptimizeMcs, :397-502
//   derived energy parameters                  OptimizeEnepara, :517-568
//   geometric temperature ladder         Loader::set_temperature, frontend.cpp:222-260
//   replica index layout                 Loader::set_replica, frontend.cpp:267-299
//
// The sorts use std::sort on (index, double key) records with a strict "<"
// on the key, exactly the record and comparator of the reference, so that
// the (unstable) order of equal keys comes out the same.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "job.hpp"

namespace gd {

namespace {

struct SortRec {
    int id;
    double val;
};

bool sort_rec_less(SortRec a, SortRec b) { return a.val < b.val; }

std::vector<int> order_by(const int *key, int n)
{
    std::vector<SortRec> v;
    v.reserve(n);
    for (int j = 0; j < n; ++j) v.push_back(SortRec{j, (double)key[j]});
    std::sort(v.begin(), v.end(), sort_rec_less);
    std::vector<int> ord(n);
    for (int j = 0; j < n; ++j) ord[j] = v[j].id;
    return ord;
}

void derive_enepara(const gd_EnePara0 &e0, gd_EnePara &e)
{
    std::memset(&e, 0, sizeof(e));
    const float sqrt_2_pi = sqrtf(2.0f * GD_PI);
    for (int i = 0; i < GD_MAXTP2; ++i)
        for (int j = 0; j < GD_MAXTP1; ++j) {
            const float tmp = e0.vdw[j][i][0] * e0.lj[2];
            e.p1a[i][j] = 2.0f * e0.vdw[j][i][1] * powf(tmp, 9.0f);
            e.p2a[i][j] = 3.0f * e0.vdw[j][i][1] * powf(tmp, 6.0f);
        }
    e.lj0 = e0.lj[0];
    e.lj1 = e0.lj[1];
    e.el1 = e0.el[1];
    e.el0 = e0.el[0];
    e.a1 = 4.0f - 3.0f * e0.el[0];
    e.b1 = 2.0f * e0.el[0] - 3.0f;
    for (int i = 0; i < GD_MAXTP2; ++i)
        for (int j = 0; j < GD_MAXTP1; ++j) {
            e.pmf0[i][j] = e0.pmf[j][i][0];
            e.pmf1[i][j] = e0.pmf[j][i][1];
            e.hdb0[i][j] = e0.hdb[j][i][0];
            e.hdb1[i][j] = 1.0f / e0.hdb[j][i][1];
        }
    for (int i = 0; i < GD_MAXTP4; ++i) e.hpp[i] = e0.hpp[i];
    for (int i = 0; i < GD_MAXTP2; ++i) {
        e.hpl0[i] = e0.hpl[i][0];
        e.hpl1[i] = e0.hpl[i][1];
        e.hpl2[i] = logf(1.0f / (e.hpl1[i] * sqrt_2_pi));
    }
    e.kde2 = -0.5f / (e0.kde * e0.kde);
    e.kde3 = powf(e0.kde * sqrt_2_pi, 3.0f);
    for (int i = 0; i < GD_MAXWEI; ++i) {
        e.w[i] = e0.w[i];
        e.a_para[i] = e0.a_para[i];
        e.b_para[i] = e0.b_para[i];
    }
}

} // namespace

// A loader output built by a caller rather than by Loader::load: every array
// the staging reads is present and every index it follows stays inside its
// table (the loader itself only writes such values).
static bool check_raw(const gd_Raw &raw, std::string &err)
{
    if (!raw.lig_x || !raw.lig_y || !raw.lig_z || !raw.lig_t || !raw.lig_c || !raw.prt_x || !raw.prt_y ||
        !raw.prt_z || !raw.prt_t || !raw.prt_c || !raw.prt_d || !raw.prt_r || !raw.prt_seq3 || !raw.kde_t ||
        !raw.kde_x || !raw.kde_y || !raw.kde_z || !raw.mcs_tcc || !raw.mcs_rowptr || !raw.mcs_i || !raw.mcs_x ||
        !raw.mcs_y || !raw.mcs_z || (raw.psp_n > 0 && (!raw.psp_res || !raw.psp_lig || !raw.psp_val))) {
        err = "a table of the loader output is null";
        return false;
    }
    for (int j = 0; j < raw.lig_natom; ++j)
        if (raw.lig_t[j] < 0 || raw.lig_t[j] >= GD_MAXTP2) {
            err = "ligand atom type out of range";
            return false;
        }
    for (int j = 0; j < raw.prt_npoint; ++j) {
        const int t = raw.prt_t[j], d = raw.prt_d[j], r = raw.prt_r[j];
        if (t < 0 || t >= GD_MAXTP1 - 1 || d < 0 || d >= GD_MAXTP4 || r < 0 || r >= raw.prt_nres) {
            err = "protein point type, residue code or residue index out of range";
            return false;
        }
    }
    for (int k = 0; k < raw.psp_n; ++k)
        if (raw.psp_res[k] < 0 || raw.psp_lig[k] < 0 || raw.psp_lig[k] >= GD_MAXTP2) {
            err = "PSP entry out of range";
            return false;
        }
    if (raw.mcs_rowptr[0] != 0) {
        err = "MCS row pointers must start at 0";
        return false;
    }
    for (int m = 0; m < raw.mcs_nrow; ++m)
        if (raw.mcs_rowptr[m + 1] <= raw.mcs_rowptr[m]) {
            err = "MCS rows must be non-empty and in order";
            return false;
        }
    for (int k = 0; k < raw.mcs_rowptr[raw.mcs_nrow]; ++k)
        if (raw.mcs_i[k] < 0 || raw.mcs_i[k] >= raw.lig_natom) {
            err = "MCS atom index out of range";
            return false;
        }
    return true;
}

bool stage(const gd_Raw &raw, StagedJob &out, std::string &err)
{
    out = StagedJob();
    gd_Complex &v = out.view;
    const int n_lig = raw.n_lig, lna = raw.lig_natom;
    const int n_prt = raw.n_prt, npt = raw.prt_npoint;
    const int nkde = raw.kde_npoint, nrow = raw.mcs_nrow;
    if (n_lig <= 0 || lna <= 0 || n_prt <= 0 || npt <= 0 || nkde <= 0 || nrow <= 0 || raw.n_tmp <= 0) {
        err = "empty input: every size must be positive";
        return false;
    }
    if (!check_raw(raw, err)) return false;

    // ---- protein ----
    std::vector<int> cdc(npt), seq3r(npt);
    for (int j = 0; j < npt; ++j) {
        cdc[j] = (raw.prt_c[j] == 0 && raw.prt_d[j] == 12) || raw.prt_c[j] == 2;
        seq3r[j] = raw.prt_seq3[raw.prt_r[j]];
    }
    const std::vector<int> pord = order_by(raw.prt_c, npt);
    out.prt_x.resize((size_t)n_prt * npt);
    out.prt_y.resize((size_t)n_prt * npt);
    out.prt_z.resize((size_t)n_prt * npt);
    out.prt_t.resize((size_t)n_prt * npt);
    out.prt_c.resize((size_t)n_prt * npt);
    out.prt_cdc.resize((size_t)n_prt * npt);
    out.prt_seq3r.resize((size_t)n_prt * npt);
    out.prt_hpp.resize((size_t)n_prt * npt);
    out.prt_ele.resize((size_t)n_prt * npt);
    int max_seq3r = 0;
    for (int c = 0; c < n_prt; ++c)
        for (int j = 0; j < npt; ++j) {
            const int s = pord[j];
            const size_t o = (size_t)c * npt + j, src = (size_t)c * npt + s;
            out.prt_x[o] = raw.prt_x[src];
            out.prt_y[o] = raw.prt_y[src];
            out.prt_z[o] = raw.prt_z[src];
            out.prt_t[o] = raw.prt_t[s];
            out.prt_c[o] = raw.prt_c[s];
            const int t = raw.prt_t[s], d = raw.prt_d[s];
            out.prt_ele[o] = raw.enepara0.ele[t == 0 ? d + 30 : t];
            out.prt_seq3r[o] = seq3r[s];
            out.prt_cdc[o] = cdc[s];
            out.prt_hpp[o] = raw.enepara0.hpp[d];
            max_seq3r = std::max(max_seq3r, seq3r[s]);
        }
    for (int d = 0; d < 3; ++d) v.prt_pocket_center[d] = raw.lig_center[d];

    // ---- PSP: transpose the sparse residue x ligand-type entries ----
    int width = max_seq3r + 1;
    for (int k = 0; k < raw.psp_n; ++k) width = std::max(width, raw.psp_res[k] + 1);
    v.psp_width = width;
    out.psp.assign((size_t)GD_MAXTP2 * width, 0.0f);
    for (int k = 0; k < raw.psp_n; ++k)
        out.psp[(size_t)raw.psp_lig[k] * width + raw.psp_res[k]] = raw.psp_val[k];

    // ---- KDE sorted by type ----
    const std::vector<int> kord = order_by(raw.kde_t, nkde);
    out.kde_t.resize(nkde);
    out.kde_x.resize(nkde);
    out.kde_y.resize(nkde);
    out.kde_z.resize(nkde);
    for (int j = 0; j < nkde; ++j) {
        out.kde_t[j] = raw.kde_t[kord[j]];
        out.kde_x[j] = raw.kde_x[kord[j]];
        out.kde_y[j] = raw.kde_y[kord[j]];
        out.kde_z[j] = raw.kde_z[kord[j]];
    }

    // ---- ligand ----
    const std::vector<int> lord = order_by(raw.lig_t, lna);
    const size_t nl = (size_t)n_lig * lna;
    out.lig_x2.resize(nl); out.lig_y2.resize(nl); out.lig_z2.resize(nl);
    out.lig_c2.resize(nl); out.lig_t2.resize(nl);
    out.lig_x3.resize(nl); out.lig_y3.resize(nl); out.lig_z3.resize(nl);
    out.lig_c3.resize(nl); out.lig_t3.resize(nl);
    out.lig_kde_begin_idx.assign(nl, 0);
    out.lig_kde_end_idx.assign(nl, 0);
    for (int c = 0; c < n_lig; ++c) {
        for (int j = 0; j < lna; ++j) {
            const size_t o = (size_t)c * lna + j;
            out.lig_x2[o] = raw.lig_x[o];
            out.lig_y2[o] = raw.lig_y[o];
            out.lig_z2[o] = raw.lig_z[o];
            out.lig_t2[o] = raw.lig_t[j];
            out.lig_c2[o] = raw.lig_c[j];
            const size_t s = (size_t)c * lna + lord[j];
            out.lig_x3[o] = raw.lig_x[s];
            out.lig_y3[o] = raw.lig_y[s];
            out.lig_z3[o] = raw.lig_z[s];
            out.lig_t3[o] = raw.lig_t[lord[j]];
            out.lig_c3[o] = raw.lig_c[lord[j]];
        }
        // same-type range in the sorted KDE table; the end stays 0 when the
        // range runs to the end of the table, as in the reference's state
        // machine (util_optimize_soa.cpp:166-202)
        for (int l = 0; l < lna; ++l) {
            const size_t o = (size_t)c * lna + l;
            int state = 0;
            for (int k = 0; k < nkde; ++k) {
                const int kt = out.kde_t[k];
                if (out.lig_t3[o] == kt && state == 0) {
                    state = 1;
                    out.lig_kde_begin_idx[o] = k;
                }
                if (out.lig_t3[o] != kt && state == 1) {
                    state = 2;
                    out.lig_kde_end_idx[o] = k;
                }
            }
        }
    }
    for (int d = 0; d < 3; ++d) v.lig_center[d] = raw.lig_center[d];

    // ---- MCS (ELL rows, entries in file order) ----
    out.mcs_ncol.resize(nrow);
    out.mcs_tcc.resize(nrow);
    out.mcs_rowptr.assign(raw.mcs_rowptr, raw.mcs_rowptr + nrow + 1);
    const int nent = raw.mcs_rowptr[nrow];
    out.mcs_i.assign(raw.mcs_i, raw.mcs_i + nent);
    out.mcs_x.assign(raw.mcs_x, raw.mcs_x + nent);
    out.mcs_y.assign(raw.mcs_y, raw.mcs_y + nent);
    out.mcs_z.assign(raw.mcs_z, raw.mcs_z + nent);
    for (int m = 0; m < nrow; ++m) {
        out.mcs_ncol[m] = raw.mcs_rowptr[m + 1] - raw.mcs_rowptr[m];
        out.mcs_tcc[m] = raw.mcs_tcc[m];
    }

    derive_enepara(raw.enepara0, v.enepara);

    // ---- temperatures: T_i = T_lo * r^i, r = exp(log(T_hi/T_lo)/(n-1)),
    //      accumulated by repeated multiplication in float ----
    const int n_tmp = raw.n_tmp;
    out.temp.resize(n_tmp);
    if (n_tmp == 1) {
        const float beta = 1.0f / (1.0f * raw.temp_low);
        out.temp[0].t = raw.temp_low;
        out.temp[0].order = 0;
        out.temp[0].minus_beta = 0.0f - beta;
    } else {
        const float ratio = std::exp(std::log(raw.temp_high / raw.temp_low) / (float)(n_tmp - 1));
        float a = raw.temp_low;
        for (int i = 0; i < n_tmp; ++i) {
            const float beta = 1.0f / (1.0f * a);
            out.temp[i].t = a;
            out.temp[i].order = i;
            out.temp[i].minus_beta = 0.0f - beta;
            a *= ratio;
        }
    }

    // ---- replicas r = n_tmp*n_lig*p + n_lig*t + l ----
    const int n_rep = n_lig * n_prt * n_tmp;
    out.replica.assign(n_rep, gd_ReplicaMC());
    std::memset(out.replica.data(), 0, sizeof(gd_ReplicaMC) * n_rep);
    for (int p = 0; p < n_prt; ++p)
        for (int t = 0; t < n_tmp; ++t)
            for (int l = 0; l < n_lig; ++l) {
                const int r = n_tmp * n_lig * p + n_lig * t + l;
                out.replica[r].idx_rep = r;
                out.replica[r].idx_prt = p;
                out.replica[r].idx_tmp = t;
                out.replica[r].idx_lig = l;
            }

    v.size.n_lig = n_lig;
    v.size.n_prt = n_prt;
    v.size.n_tmp = n_tmp;
    v.size.n_rep = n_rep;
    v.size.lig_natom = lna;
    v.size.prt_npoint = npt;
    v.size.kde_npoint = nkde;
    v.size.mcs_nrow = nrow;
    v.mcpara.steps_per_dump = GD_STEPS_PER_DUMP;
    v.mcpara.steps_n_dump = raw.steps_n_dump;
    v.mcpara.steps_total = raw.steps_n_dump * GD_STEPS_PER_DUMP;
    for (int i = 0; i < 6; ++i) v.mcpara.move_scale[i] = raw.move_scale[i];
    out.bind();
    return true;
}

void StagedJob::bind()
{
    view.lig_x2 = lig_x2.data(); view.lig_y2 = lig_y2.data(); view.lig_z2 = lig_z2.data();
    view.lig_c2 = lig_c2.data(); view.lig_t2 = lig_t2.data();
    view.lig_x3 = lig_x3.data(); view.lig_y3 = lig_y3.data(); view.lig_z3 = lig_z3.data();
    view.lig_c3 = lig_c3.data(); view.lig_t3 = lig_t3.data();
    view.lig_kde_begin_idx = lig_kde_begin_idx.data();
    view.lig_kde_end_idx = lig_kde_end_idx.data();
    view.prt_x = prt_x.data(); view.prt_y = prt_y.data(); view.prt_z = prt_z.data();
    view.prt_hpp = prt_hpp.data(); view.prt_ele = prt_ele.data();
    view.prt_t = prt_t.data(); view.prt_c = prt_c.data();
    view.prt_cdc = prt_cdc.data(); view.prt_seq3r = prt_seq3r.data();
    view.psp = psp.data();
    view.kde_t = kde_t.data(); view.kde_x = kde_x.data(); view.kde_y = kde_y.data(); view.kde_z = kde_z.data();
    view.mcs_ncol = mcs_ncol.data(); view.mcs_tcc = mcs_tcc.data(); view.mcs_rowptr = mcs_rowptr.data();
    view.mcs_i = mcs_i.data(); view.mcs_x = mcs_x.data(); view.mcs_y = mcs_y.data(); view.mcs_z = mcs_z.data();
    view.temp = temp.data();
    view.replica = replica.data();
}

} // namespace gd
