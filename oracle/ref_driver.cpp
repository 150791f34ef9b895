// ref_driver.cpp -- TEST INFRASTRUCTURE: drives the reference's own CPU code
// (compiled from its sources by oracle/Makefile into oracle/_ref/) so the
// oracle restatement and the library's staging can be checked against it.
//
// What it compiles from the reference, unmodified:
//   src/backend_cpu_mic/kernel_cpp_montecarlo_soa.cpp  (MonteCarlo_d and the
//     energy fragments it includes: energy_prt_v1.cpp, energy_kde_v1.cpp,
//     energy_mcs_v3_ell.cpp)
//   src/frontend/util_optimize_soa.cpp  (OptimizeLigand/Protein/Kde/Psp/Mcs/Enepara)
// The reference draws its random numbers with rand_r seeded from time()
// (kernel_cpp_montecarlo_soa.cpp:8-15).  This file maps those two names to
// functions that return 0 for the clock and the injected Philox stream for
// rand_r, so the seed rand_r sees is the replica index and every replica
// consumes its own stream, draw by draw.  The loaders (load.cpp) are not
// built: they need the Fortran u3b routine, and there is no Fortran
// compiler in this image.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include <omp.h>

#include <geauxdock.h>
#include <size.h>
#include <toggle.h>
#include <util_optimize.h>

#include "gd/gd_types.h"

extern "C" uint32_t gdo_stream_k(uint64_t seed, uint32_t kind, uint32_t r, uint32_t s, uint32_t j);

namespace {
uint64_t g_seed = 0;
std::vector<int> g_cursor; // draws consumed per replica in the current chunk

time_t injected_time(time_t *) { return 0; }

int injected_rand_r(unsigned int *s)
{
    const int r = (int)*s;
    const int c = g_cursor[r]++;
    return (int)gdo_stream_k(g_seed, 0, (uint32_t)r, (uint32_t)(c / 7 + 1), (uint32_t)(c % 7));
}
} // namespace

#define time(x) injected_time(x)
#define rand_r(s) injected_rand_r(s)
#include "kernel_cpp_montecarlo_soa.cpp"
#undef time
#undef rand_r

struct gdref_handle {
    Complex *cx = nullptr;
    // flat copies of the reference-staged tables, exported as a gd_Complex
    std::vector<float> lx2, ly2, lz2, lc2, lx3, ly3, lz3, lc3;
    std::vector<int> lt2, lt3, kb, ke;
    std::vector<float> px, py, pz, phpp, pele;
    std::vector<int> pt, pc, pcdc, pseq3r;
    std::vector<float> psp;
    std::vector<int> kt;
    std::vector<float> kx, ky, kz;
    std::vector<int> mncol, mrowptr, mi;
    std::vector<float> mtcc, mx, my, mz;
    std::vector<gd_Temp> temp;
    std::vector<gd_ReplicaMC> rep;
    gd_Complex view;
};

extern "C" {

// Build the reference Complex from loader output, running the reference's
// Optimize* builders.  temp: the temperature ladder (n_tmp entries).
int gdref_create(const gd_Raw *raw, const gd_Temp *temp, int n_tmp, gdref_handle **out)
{
    const int n_lig = raw->n_lig, lna = raw->lig_natom, n_prt = raw->n_prt, npt = raw->prt_npoint;
    const int nkde = raw->kde_npoint, nrow = raw->mcs_nrow;
    const int n_rep = n_lig * n_prt * n_tmp;
    if (n_lig > MAX_CONF_LIG || lna > MAXLIG || n_prt > MAX_CONF_PRT || npt > MAXPRO || nkde > MAXKDE ||
        nrow > MAX_MCS_ROW || n_tmp > MAX_TMP || n_rep > MAX_REP)
        return -1;
    for (int k = 0; k < raw->psp_n; ++k)
        if (raw->psp_res[k] >= MAXPRO) return -1;

    Ligand0 *lig0 = new Ligand0[n_lig];
    for (int c = 0; c < n_lig; ++c) {
        Ligand0 &L = lig0[c];
        L.lig_natom = lna;
        for (int a = 0; a < lna; ++a) {
            L.coord_orig.x[a] = raw->lig_x[c * lna + a];
            L.coord_orig.y[a] = raw->lig_y[c * lna + a];
            L.coord_orig.z[a] = raw->lig_z[c * lna + a];
            L.t[a] = raw->lig_t[a];
            L.c[a] = raw->lig_c[a];
            L.n[a] = a;
        }
        for (int d = 0; d < 3; ++d) {
            L.pocket_center[d] = raw->lig_center[d];
            L.coord_orig.center[d] = raw->lig_center[d];
        }
    }
    Protein0 *prt0 = new Protein0[n_prt]();
    for (int c = 0; c < n_prt; ++c) {
        Protein0 &P = prt0[c];
        P.prt_npoint = npt;
        P.pnr = raw->prt_nres;
        for (int j = 0; j < npt; ++j) {
            P.x[j] = raw->prt_x[(size_t)c * npt + j];
            P.y[j] = raw->prt_y[(size_t)c * npt + j];
            P.z[j] = raw->prt_z[(size_t)c * npt + j];
            P.t[j] = raw->prt_t[j];
            P.c[j] = raw->prt_c[j];
            P.d[j] = raw->prt_d[j];
            P.r[j] = raw->prt_r[j];
        }
        for (int k = 0; k < raw->prt_nres; ++k) P.seq3[k] = raw->prt_seq3[k];
    }
    EnePara0 *ep0 = new EnePara0();
    std::memcpy(ep0, &raw->enepara0, sizeof(EnePara0));
    static_assert(sizeof(EnePara0) == sizeof(gd_EnePara0), "EnePara0 layout");
    Psp0 *psp0 = new Psp0();
    for (int k = 0; k < raw->psp_n; ++k) psp0->psp[raw->psp_res[k]][raw->psp_lig[k]] = raw->psp_val[k];
    Kde0 *kde0 = new Kde0();
    kde0->kde_npoint = nkde;
    for (int k = 0; k < nkde; ++k) {
        kde0->x[k] = raw->kde_x[k];
        kde0->y[k] = raw->kde_y[k];
        kde0->z[k] = raw->kde_z[k];
        kde0->t[k] = raw->kde_t[k];
    }
    Mcs0 *mcs0 = new Mcs0[nrow]();
    for (int m = 0; m < nrow; ++m) {
        for (int j = 0; j < MAX_MCS_COL; ++j) mcs0[m].x[j] = mcs0[m].y[j] = mcs0[m].z[j] = MCS_INVALID_COORD;
        const int b = raw->mcs_rowptr[m], e = raw->mcs_rowptr[m + 1];
        mcs0[m].ncol = e - b;
        mcs0[m].tcc = raw->mcs_tcc[m];
        for (int j = b; j < e; ++j) {
            const int loc = raw->mcs_i[j];
            mcs0[m].x[loc] = raw->mcs_x[j];
            mcs0[m].y[loc] = raw->mcs_y[j];
            mcs0[m].z[loc] = raw->mcs_z[j];
            mcs0[m].x2[j - b] = raw->mcs_x[j];
            mcs0[m].y2[j - b] = raw->mcs_y[j];
            mcs0[m].z2[j - b] = raw->mcs_z[j];
            mcs0[m].idx_col[j - b] = loc;
        }
    }

    gdref_handle *h = new gdref_handle();
    h->cx = new Complex();
    Complex *cx = h->cx;
    OptimizeKde(kde0, &cx->kde);
    OptimizeLigand(lig0, &cx->kde, cx->lig, n_lig);
    OptimizePsp(psp0, &cx->psp);
    Mcs *mcs_tmp = new Mcs[MAX_MCS_ROW]();
    OptimizeMcs(mcs0, mcs_tmp, &cx->mcs_r, &cx->mcs_ell, &cx->mcs_csr, nrow);
    for (int m = 0; m < nrow; ++m) cx->mcs[m] = mcs_tmp[m];
    delete[] mcs_tmp;
    OptimizeProtein(prt0, cx->prt, ep0, n_prt);
    OptimizeEnepara(ep0, &cx->enepara);
    for (int c = 0; c < n_prt; ++c)
        for (int d = 0; d < 3; ++d) cx->prt[c].pocket_center[d] = raw->lig_center[d];
    for (int t = 0; t < n_tmp; ++t) {
        cx->temp[t].t = temp[t].t;
        cx->temp[t].minus_beta = temp[t].minus_beta;
        cx->temp[t].order = temp[t].order;
    }
    cx->size.n_lig = n_lig;
    cx->size.n_prt = n_prt;
    cx->size.n_tmp = n_tmp;
    cx->size.n_rep = n_rep;
    cx->size.lig_natom = lna;
    cx->size.prt_npoint = npt;
    cx->size.kde_npoint = nkde;
    cx->size.mcs_nrow = nrow;
    cx->mcpara.steps_per_dump = STEPS_PER_DUMP;
    cx->mcpara.steps_n_dump = raw->steps_n_dump;
    cx->mcpara.steps_total = raw->steps_n_dump * STEPS_PER_DUMP;
    for (int i = 0; i < 6; ++i) cx->mcpara.move_scale[i] = raw->move_scale[i];
    for (int p = 0; p < n_prt; ++p)
        for (int t = 0; t < n_tmp; ++t)
            for (int l = 0; l < n_lig; ++l) {
                const int r = n_tmp * n_lig * p + n_lig * t + l;
                ReplicaMC &R = cx->replica[r];
                std::memset(&R, 0, sizeof(R));
                R.idx_rep = r;
                R.idx_prt = p;
                R.idx_tmp = t;
                R.idx_lig = l;
            }
    delete[] lig0;
    delete[] prt0;
    delete ep0;
    delete psp0;
    delete kde0;
    delete[] mcs0;

    // flat export of the reference-staged tables
    h->lx2.resize(n_lig * lna); h->ly2.resize(n_lig * lna); h->lz2.resize(n_lig * lna);
    h->lc2.resize(n_lig * lna); h->lt2.resize(n_lig * lna);
    h->lx3.resize(n_lig * lna); h->ly3.resize(n_lig * lna); h->lz3.resize(n_lig * lna);
    h->lc3.resize(n_lig * lna); h->lt3.resize(n_lig * lna);
    h->kb.resize(n_lig * lna); h->ke.resize(n_lig * lna);
    for (int c = 0; c < n_lig; ++c)
        for (int a = 0; a < lna; ++a) {
            const int o = c * lna + a;
            const Ligand &L = cx->lig[c];
            h->lx2[o] = L.x2[a]; h->ly2[o] = L.y2[a]; h->lz2[o] = L.z2[a]; h->lc2[o] = L.c2[a]; h->lt2[o] = L.t2[a];
            h->lx3[o] = L.x3[a]; h->ly3[o] = L.y3[a]; h->lz3[o] = L.z3[a]; h->lc3[o] = L.c3[a]; h->lt3[o] = L.t3[a];
            h->kb[o] = L.kde_begin_idx[a]; h->ke[o] = L.kde_end_idx[a];
        }
    const size_t np = (size_t)n_prt * npt;
    h->px.resize(np); h->py.resize(np); h->pz.resize(np); h->phpp.resize(np); h->pele.resize(np);
    h->pt.resize(np); h->pc.resize(np); h->pcdc.resize(np); h->pseq3r.resize(np);
    int width = 1;
    for (int c = 0; c < n_prt; ++c)
        for (int j = 0; j < npt; ++j) {
            const size_t o = (size_t)c * npt + j;
            const Protein &P = cx->prt[c];
            h->px[o] = P.x[j]; h->py[o] = P.y[j]; h->pz[o] = P.z[j];
            h->phpp[o] = P.hpp[j]; h->pele[o] = P.ele[j];
            h->pt[o] = P.t[j]; h->pc[o] = P.c[j]; h->pcdc[o] = P.cdc[j]; h->pseq3r[o] = P.seq3r[j];
            width = std::max(width, P.seq3r[j] + 1);
        }
    for (int k = 0; k < raw->psp_n; ++k) width = std::max(width, raw->psp_res[k] + 1);
    h->psp.assign((size_t)GD_MAXTP2 * width, 0.0f);
    for (int t = 0; t < GD_MAXTP2; ++t)
        for (int i = 0; i < width; ++i) h->psp[(size_t)t * width + i] = cx->psp.psp[t][i];
    h->kt.resize(nkde); h->kx.resize(nkde); h->ky.resize(nkde); h->kz.resize(nkde);
    for (int k = 0; k < nkde; ++k) {
        h->kt[k] = cx->kde.t[k]; h->kx[k] = cx->kde.x[k]; h->ky[k] = cx->kde.y[k]; h->kz[k] = cx->kde.z[k];
    }
    h->mncol.resize(nrow); h->mtcc.resize(nrow); h->mrowptr.assign(1, 0);
    for (int m = 0; m < nrow; ++m) {
        h->mncol[m] = cx->mcs_ell.ncol[m];
        h->mtcc[m] = cx->mcs_ell.tcc[m];
        for (int j = 0; j < h->mncol[m]; ++j) {
            h->mi.push_back(cx->mcs_ell.i[m][j]);
            h->mx.push_back(cx->mcs_ell.x[m][j]);
            h->my.push_back(cx->mcs_ell.y[m][j]);
            h->mz.push_back(cx->mcs_ell.z[m][j]);
        }
        h->mrowptr.push_back((int)h->mi.size());
    }
    h->temp.assign(temp, temp + n_tmp);
    h->rep.resize(n_rep);
    for (int r = 0; r < n_rep; ++r) std::memcpy(&h->rep[r], &cx->replica[r], sizeof(gd_ReplicaMC));
    static_assert(sizeof(ReplicaMC) == sizeof(gd_ReplicaMC), "ReplicaMC layout");
    static_assert(sizeof(EnePara) == sizeof(gd_EnePara), "EnePara layout");

    gd_Complex &v = h->view;
    std::memset(&v, 0, sizeof(v));
    v.size.n_lig = n_lig; v.size.n_prt = n_prt; v.size.n_tmp = n_tmp; v.size.n_rep = n_rep;
    v.size.lig_natom = lna; v.size.prt_npoint = npt; v.size.kde_npoint = nkde; v.size.mcs_nrow = nrow;
    v.lig_x2 = h->lx2.data(); v.lig_y2 = h->ly2.data(); v.lig_z2 = h->lz2.data(); v.lig_c2 = h->lc2.data(); v.lig_t2 = h->lt2.data();
    v.lig_x3 = h->lx3.data(); v.lig_y3 = h->ly3.data(); v.lig_z3 = h->lz3.data(); v.lig_c3 = h->lc3.data(); v.lig_t3 = h->lt3.data();
    v.lig_kde_begin_idx = h->kb.data(); v.lig_kde_end_idx = h->ke.data();
    for (int d = 0; d < 3; ++d) {
        v.lig_center[d] = cx->lig[0].center[d];
        v.prt_pocket_center[d] = cx->prt[0].pocket_center[d];
    }
    v.prt_x = h->px.data(); v.prt_y = h->py.data(); v.prt_z = h->pz.data();
    v.prt_hpp = h->phpp.data(); v.prt_ele = h->pele.data();
    v.prt_t = h->pt.data(); v.prt_c = h->pc.data(); v.prt_cdc = h->pcdc.data(); v.prt_seq3r = h->pseq3r.data();
    v.psp_width = width; v.psp = h->psp.data();
    v.kde_t = h->kt.data(); v.kde_x = h->kx.data(); v.kde_y = h->ky.data(); v.kde_z = h->kz.data();
    v.mcs_ncol = h->mncol.data(); v.mcs_tcc = h->mtcc.data(); v.mcs_rowptr = h->mrowptr.data();
    v.mcs_i = h->mi.data(); v.mcs_x = h->mx.data(); v.mcs_y = h->my.data(); v.mcs_z = h->mz.data();
    std::memcpy(&v.enepara, &cx->enepara, sizeof(gd_EnePara));
    v.temp = h->temp.data();
    v.replica = h->rep.data();
    v.mcpara.steps_per_dump = cx->mcpara.steps_per_dump;
    v.mcpara.steps_n_dump = cx->mcpara.steps_n_dump;
    v.mcpara.steps_total = cx->mcpara.steps_total;
    for (int i = 0; i < 6; ++i) v.mcpara.move_scale[i] = cx->mcpara.move_scale[i];
    *out = h;
    return 0;
}

const gd_Complex *gdref_view(const gdref_handle *h) { return &h->view; }

void gdref_free(gdref_handle *h)
{
    if (!h) return;
    delete h->cx;
    delete h;
}

// Run the reference MonteCarlo_d for steps s1..s1+s2max-1 (s2max <= 100).
// rep: n_rep replica states, in/out.  rec: n_rep * 100 entries, next_entry: n_rep.
int gdref_montecarlo(gdref_handle *h, gd_ReplicaMC *rep, gd_ReplicaMC *rec, int *next_entry, int s1,
                     int s2max, uint64_t seed, int nthreads)
{
    Complex *cx = h->cx;
    const int n_rep = cx->size.n_rep;
    if (s2max > STEPS_PER_DUMP || s1 < 0) return -1;
    for (int r = 0; r < n_rep; ++r) std::memcpy(&cx->replica[r], &rep[r], sizeof(ReplicaMC));
    g_seed = seed;
    g_cursor.assign(n_rep, std::max(s1 - 1, 0) * 7);
    if (nthreads > 0) omp_set_num_threads(nthreads);
    Record *record = new Record[n_rep]();
    MonteCarlo_d(cx, record, s1, s2max);
    for (int r = 0; r < n_rep; ++r) {
        std::memcpy(&rep[r], &cx->replica[r], sizeof(ReplicaMC));
        if (next_entry) next_entry[r] = record[r].next_entry;
        if (rec)
            for (int k = 0; k < record[r].next_entry; ++k)
                std::memcpy(&rec[(size_t)r * STEPS_PER_DUMP + k], &record[r].replica[k], sizeof(ReplicaMC));
    }
    delete[] record;
    return 0;
}

} // extern "C"
