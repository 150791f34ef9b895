// Device half of the C ABI: context, resident target and job tables, the
// plain Monte Carlo chunk (one workgroup per replica), pose energies and the
// temperature exchange.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "context.hpp"

namespace gd {

#define GD_HIP(call)                                                                                 \
    do {                                                                                             \
        hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess)                                                                        \
            return fail(e_ == hipErrorOutOfMemory ? GD_ERR_NOMEM : GD_ERR_DEVICE,                     \
                        std::string(#call) + ": " + hipGetErrorString(e_));                          \
    } while (0)

Arena::~Arena()
{
    if (base) (void)hipFree(base);
}

hipError_t Arena::init(size_t n)
{
    cap = rounded(n);
    hipError_t e = hipMalloc(&base, cap);
    if (e != hipSuccess) {
        base = nullptr;
        cap = 0;
        return e;
    }
    free_blocks.clear();
    free_blocks[0] = cap;
    in_use = 0;
    return hipSuccess;
}

void *Arena::take(size_t n)
{
    n = rounded(n);
    for (auto it = free_blocks.begin(); it != free_blocks.end(); ++it) {
        if (it->second < n) continue;
        const size_t off = it->first, size = it->second;
        free_blocks.erase(it);
        if (size > n) free_blocks[off + n] = size - n;
        in_use += n;
        return (char *)base + off;
    }
    return nullptr;
}

void Arena::give(void *q, size_t n)
{
    n = rounded(n);
    size_t off = (size_t)((char *)q - (char *)base);
    in_use -= n;
    auto next = free_blocks.lower_bound(off);
    if (next != free_blocks.end() && off + n == next->first) { // merge with the block after
        n += next->second;
        next = free_blocks.erase(next);
    }
    if (next != free_blocks.begin()) { // merge with the block before
        auto prev = std::prev(next);
        if (prev->first + prev->second == off) {
            prev->second += n;
            return;
        }
    }
    free_blocks[off] = n;
}

void DevBuf::release()
{
    if (p) {
        if (arena)
            arena->give(p, bytes);
        else
            (void)hipFree(p);
    }
    p = nullptr;
    bytes = 0;
}

hipError_t DevBuf::alloc(size_t n)
{
    release();
    if (n == 0) n = 16;
    if (arena) {
        p = arena->take(n);
        if (!p) return hipErrorOutOfMemory;
        bytes = n;
        return hipSuccess;
    }
    hipError_t e = hipMalloc(&p, n);
    if (e == hipSuccess) bytes = n;
    else p = nullptr;
    return e;
}

hipError_t upload(DevBuf &b, const void *src, size_t bytes, hipStream_t s)
{
    if (b.bytes < bytes || !b.p) {
        hipError_t e = b.alloc(bytes);
        if (e != hipSuccess) return e;
    }
    if (bytes == 0) return hipSuccess;
    return hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, s);
}

// Plain form: one workgroup per replica with the reference's 256-thread
// block (TperB, src/GNUmakefile.common.mk:62-114): the ligand atoms fill
// ceil(N_L / 64) waves and the remaining waves split the protein points.
Shape plain_shape(int natom)
{
    Shape s;
    s.nrep = 1;
    s.slot_waves = (natom + 63) / 64;
    s.slices = std::max(1, 4 / s.slot_waves);
    s.waves = s.slot_waves * s.slices;
    return s;
}

// The arrays of a complex that the uploads (target: the protein tables; job:
// ligand, KDE, MCS, temperatures, replicas, PSP and the point topology that
// is checked against the target) read; a null one is refused, not followed.
int check_complex_arrays(const gd_Complex *cx, bool target, bool job)
{
    if (target && (!cx->prt_x || !cx->prt_y || !cx->prt_z || !cx->prt_hpp || !cx->prt_ele || !cx->prt_t || !cx->prt_c ||
                   !cx->prt_cdc))
        return fail(GD_ERR_ARG, "a protein table of the complex is null");
    if (job && (!cx->lig_x2 || !cx->lig_y2 || !cx->lig_z2 || !cx->lig_c2 || !cx->lig_t2 || !cx->lig_x3 || !cx->lig_y3 ||
                !cx->lig_z3 || !cx->lig_c3 || !cx->lig_t3 || !cx->kde_t || !cx->kde_x || !cx->kde_y || !cx->kde_z ||
                !cx->mcs_ncol || !cx->mcs_tcc || !cx->mcs_rowptr || !cx->mcs_i || !cx->mcs_x || !cx->mcs_y ||
                !cx->mcs_z || !cx->temp || !cx->replica || !cx->prt_t || !cx->prt_c || !cx->prt_seq3r ||
                (cx->psp_width > 0 && !cx->psp)))
        return fail(GD_ERR_ARG, "a job table of the complex is null");
    return GD_OK;
}

int check_stream_range(const JobRes &j, long s_end)
{
    if (j.kstream_steps > 0 && s_end - 1 > j.kstream_steps)
        return fail(GD_ERR_ARG, "the job's explicit stream words end at step " + std::to_string(j.kstream_steps) +
                                    ", the launch needs step " + std::to_string(s_end - 1));
    return GD_OK;
}

int ensure_records(gd_ctx *ctx, JobRes &j, int cap)
{
    if (j.rec_cap >= cap && j.rec.p) return GD_OK;
    const size_t bytes = sizeof(gd_ReplicaMC) * (size_t)j.n_rep * cap;
    GD_HIP(j.rec.alloc(bytes));
    GD_HIP(hipMemsetAsync(j.rec.p, 0, bytes, ctx->stream));
    j.rec_cap = cap;
    j.hj.rec = j.rec.as<gd_ReplicaMC>();
    j.hj.rec_cap = cap;
    GD_HIP(hipMemcpyAsync(j.self.p, &j.hj, sizeof(DevJob), hipMemcpyHostToDevice, ctx->stream));
    return GD_OK;
}

// Device coordinates live in a frame centred on the target's pocket centre:
// differences of two ~30 A absolute coordinates would lose ~2e-6 A to float
// rounding, which the narrow KDE Gaussians and the normalisation slopes
// amplify.  The shift is done in double and rounded once.
float local(float x, float origin) { return (float)((double)x - (double)origin); }

int ensure_group_streams(gd_ctx *ctx)
{
    for (hipStream_t &g : ctx->group_stream)
        if (!g) GD_HIP(hipStreamCreateWithFlags(&g, hipStreamNonBlocking));
    return GD_OK;
}

int run_tasks(gd_ctx *ctx, const TargetRes &t, const DevJob *jobs_dev, const std::vector<DevTask> &tasks,
              int waves, int max_rep, bool compensated, bool records)
{
    GD_HIP(upload(ctx->task_buf, tasks.data(), sizeof(DevTask) * tasks.size(), ctx->stream));
    GD_HIP(launch_mc(t.params.as<DevParams>(), jobs_dev, ctx->task_buf.as<DevTask>(), (int)tasks.size(), waves, max_rep,
                     compensated, records, ctx->stream, nullptr));
    return GD_OK;
}

} // namespace gd

using namespace gd;

namespace {

int check_ctx(gd_ctx *ctx)
{
    if (!ctx) return fail(GD_ERR_ARG, "null context");
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(GD_ERR_DEVICE, "hipSetDevice failed");
    return GD_OK;
}

JobRes *find_job(gd_ctx *ctx, int id)
{
    auto it = ctx->jobs.find(id);
    return it == ctx->jobs.end() ? nullptr : &it->second;
}

// The replica layout the kernels index by (Loader::set_replica,
// src/frontend/frontend.cpp:267-299): replica r = n_tmp n_lig p + n_lig t + l
// belongs to ligand conformation l and protein conformation p, and every
// (l, p) ladder holds each temperature exactly once (the exchange kernel
// finds a ladder's replicas through their temperature labels).  idx_rep keys
// the replica's random stream and indexes explicit stream words
// ([idx_rep][step][7]), so it must be a permutation of [0, n_rep): a value
// outside would read beyond the words, a duplicate would share a stream.
int check_replica_layout(const gd_ReplicaMC *rep, int n_lig, int n_prt, int n_tmp)
{
    const size_t n_rep = (size_t)n_lig * n_prt * n_tmp;
    std::vector<char> rep_seen(n_rep, 0);
    for (size_t r = 0; r < n_rep; ++r) {
        const int k = rep[r].idx_rep;
        if (k < 0 || (size_t)k >= n_rep) return fail(GD_ERR_ARG, "replica idx_rep out of range");
        if (rep_seen[k]++) return fail(GD_ERR_ARG, "replica idx_rep values are not a permutation");
    }
    std::vector<int> seen(n_tmp);
    for (int p = 0; p < n_prt; ++p)
        for (int l = 0; l < n_lig; ++l) {
            std::fill(seen.begin(), seen.end(), 0);
            for (int t = 0; t < n_tmp; ++t) {
                const gd_ReplicaMC &r = rep[(size_t)n_tmp * n_lig * p + (size_t)n_lig * t + l];
                if (r.idx_prt != p || r.idx_lig != l || r.idx_tmp < 0 || r.idx_tmp >= n_tmp)
                    return fail(GD_ERR_ARG, "replica index fields out of range or out of place");
                if (seen[r.idx_tmp]++) return fail(GD_ERR_ARG, "ladder temperatures are not a permutation");
            }
        }
    return GD_OK;
}

} // namespace

extern "C" {

int gd_create(int device, gd_ctx **out)
{
    if (!out) return fail(GD_ERR_ARG, "null argument");
    int n = 0;
    GD_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(GD_ERR_ARG, "no such device");
    GD_HIP(hipSetDevice(device));
    gd_ctx *c = new gd_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        delete c;
        return fail(GD_ERR_DEVICE, "hipGetDeviceProperties failed");
    }
    c->cu_count = prop.multiProcessorCount;
    c->hbm_bytes = prop.totalGlobalMem;
    c->name = prop.gcnArchName;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(GD_ERR_DEVICE, "hipStreamCreate failed");
    }
    *out = c;
    return GD_OK;
}

void gd_destroy(gd_ctx *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (hipStream_t g : ctx->group_stream)
        if (g) (void)hipStreamSynchronize(g);
    ctx->jobs.clear();
    ctx->targets.clear();
    ctx->task_buf.release();
    ctx->job_buf.release();
    ctx->pose_rep.release();
    ctx->pose_move.release();
    ctx->pose_out.release();
    ctx->counter.release();
    ctx->queue.release();
    ctx->ladders.release();
    ctx->bases.release();
    for (hipStream_t &g : ctx->group_stream)
        if (g) (void)hipStreamDestroy(g);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int gd_device_info(gd_ctx *ctx, char *name, int name_len, int *cu_count, size_t *hbm_bytes)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (name && name_len > 0) {
        std::strncpy(name, ctx->name.c_str(), name_len - 1);
        name[name_len - 1] = 0;
    }
    if (cu_count) *cu_count = ctx->cu_count;
    if (hbm_bytes) *hbm_bytes = ctx->hbm_bytes;
    return GD_OK;
}

int gd_upload_target(gd_ctx *ctx, const gd_Complex *cx, int *target_id)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (!cx || !target_id) return fail(GD_ERR_ARG, "null argument");
    if (int rc = check_complex_arrays(cx, true, false)) return rc;
    const int np = cx->size.prt_npoint, nconf = cx->size.n_prt;
    if (np <= 0 || nconf <= 0) return fail(GD_ERR_ARG, "empty protein");
    const gd_EnePara &ep = cx->enepara;

    for (int i = 0; i < np; ++i)
        if (cx->prt_t[i] < 0 || cx->prt_t[i] >= GD_MAXTP1) return fail(GD_ERR_ARG, "protein point type out of range");
    const float origin[3] = {cx->prt_pocket_center[0], cx->prt_pocket_center[1], cx->prt_pocket_center[2]};

    // per point: hydrophobic-contact weight (cdc ? hpp : 0); points are
    // re-sorted by (type, hydrophobic flag), so a segment shares one set of
    // type-pair parameters and one term set
    std::vector<float> hpcw(np);
    std::vector<int> flags(np);
    for (int s = 0; s < np; ++s) {
        hpcw[s] = cx->prt_cdc[s] == 1 ? cx->prt_hpp[s] : 0.0f;
        flags[s] = hpcw[s] != 0.0f ? 1 : 0;
    }
    // smallest distance of each point from the frame origin over the
    // conformations (in the float coordinates the kernel reads): within a
    // segment points go nearest first, so the points beyond any radius are a
    // suffix of the segment
    std::vector<double> rmin(np, 1e300);
    for (int c = 0; c < nconf; ++c)
        for (int i = 0; i < np; ++i) {
            const size_t k = (size_t)c * np + i;
            const double x = local(cx->prt_x[k], origin[0]), y = local(cx->prt_y[k], origin[1]),
                         z = local(cx->prt_z[k], origin[2]);
            const double r = std::sqrt(x * x + y * y + z * z);
            rmin[i] = std::min(rmin[i], r);
        }
    std::vector<int> ord(np);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) {
        if (cx->prt_t[x] != cx->prt_t[y]) return cx->prt_t[x] < cx->prt_t[y];
        if (flags[x] != flags[y]) return flags[x] > flags[y];
        return rmin[x] < rmin[y];
    });
    std::vector<int4> seg;
    for (int i = 0; i < np;) {
        const int t = cx->prt_t[ord[i]], f = flags[ord[i]];
        int j = i;
        while (j < np && cx->prt_t[ord[j]] == t && flags[ord[j]] == f) ++j;
        seg.push_back(make_int4(t, i, j, f));
        i = j;
    }
    // exponents as quadratics in d, coefficients rounded once from double
    const double log2e = 1.4426950408889634;
    const double ljc = ep.lj0 != 0.0f ? (double)ep.lj0 : 1.0; // LJ sums are kept times ljc
    std::vector<float> pp((size_t)GD_MAXTP1 * GD_MAXTP2 * PP_N, 0.0f);
    for (int tp = 0; tp < GD_MAXTP1; ++tp)
        for (int tl = 0; tl < GD_MAXTP2; ++tl) {
            float *q = &pp[((size_t)tp * GD_MAXTP2 + tl) * PP_N];
            const double pmf0 = ep.pmf0[tl][tp];
            q[PP_P1A] = (float)(ljc * ep.p1a[tl][tp]);
            q[PP_P2A] = (float)(ljc * ep.p2a[tl][tp]);
            q[PP_PMF0] = ep.pmf0[tl][tp];
            q[PP_PMF1] = ep.pmf1[tl][tp];
            q[PP_PMFB] = (float)(log2e * (6.0 + 0.5 * pmf0));
            q[PP_PMFC] = (float)(-6.0 * log2e * pmf0);
            const bool hb = ep.hdb0[tl][tp] > 0.1f; // energy_prt_v1.cpp:138
            const double h0 = ep.hdb0[tl][tp], h1 = hb ? (double)ep.hdb1[tl][tp] : 0.0;
            const double k = 0.5 * log2e * h1 * h1;
            q[PP_HK2] = (float)(-k);
            q[PP_HK1] = (float)(2.0 * k * h0);
            q[PP_HK0] = (float)(-k * h0 * h0);
            q[PP_HDB1] = hb ? ep.hdb1[tl][tp] : 0.0f;
            q[PP_HDB0] = ep.hdb0[tl][tp];
        }

    // Far field.  Beyond far_d every pair term but LJ and 1/d electrostatics
    // is constant to float precision: the hydrophobic cutoff (9 A) and the
    // electrostatic core (1/el1) are behind; the contact sigmoid's exp factor
    // is below 2^-26, so pmf1 / (1 + e) rounds to pmf1; the H-bond Gaussian is
    // below e^-30 of hdb1.  A point whose smallest radius is r is at least
    // r - rho from every atom within rho of the origin, so per segment the
    // points at r >= rho + far_d are far for all of them; ladder[seg][k] is
    // the first such point for rho + far_d <= kFarR0 + k * kFarDr.  Likewise
    // a point whose largest radius is r is at least rho' - r from every atom
    // at radius >= rho' (a ligand that has left the pocket): ladder_in[seg][k]
    // ends the segment's leading run of points with largest radius <= rung k.
    double far_d = std::max(9.0, 1.0 / (double)ep.el1);
    // Mid field: beyond mid_d the hydrophobic term is exactly 0 (u clamps to
    // 0 past 9 A), no pair is inside the electrostatic core and the H-bond
    // Gaussian is below e^-30 of hdb1; only LJ, 1/d electrostatics and the
    // contact sigmoid vary.
    double mid_d = far_d;
    for (int tp = 0; tp < GD_MAXTP1; ++tp)
        for (int tl = 0; tl < GD_MAXTP2; ++tl)
            if (ep.hdb0[tl][tp] > 0.1f && ep.hdb1[tl][tp] != 0.0f)
                mid_d = std::max(mid_d, (double)ep.hdb0[tl][tp] + std::sqrt(60.0) / std::fabs((double)ep.hdb1[tl][tp]));
    mid_d += 0.05;
    {
        const double tp_ = 26.0 * std::log(2.0);
        for (int tp = 0; tp < GD_MAXTP1; ++tp)
            for (int tl = 0; tl < GD_MAXTP2; ++tl) {
                if (ep.pmf1[tl][tp] != 0.0f) {
                    // (d/2 - 6)(d - pmf0) >= tp_, d > 12
                    const double p0 = ep.pmf0[tl][tp], bq = 6.0 + 0.5 * p0;
                    const double disc = bq * bq - 2.0 * (6.0 * p0 - tp_);
                    far_d = std::max(far_d, bq + std::sqrt(std::max(disc, 0.0)));
                }
                if (ep.hdb0[tl][tp] > 0.1f && ep.hdb1[tl][tp] != 0.0f) // ((d - h0) h1)^2 / 2 >= 30
                    far_d = std::max(far_d, (double)ep.hdb0[tl][tp] + std::sqrt(60.0) / std::fabs((double)ep.hdb1[tl][tp]));
            }
        far_d += 0.05;
    }
    // Very far field.  The LJ denominator is 1 + e, e = X (lj_a + lj_b d2),
    // X = c p1a / d^9; e falls with d (lj_a, lj_b >= 0).  Beyond vfar_d,
    // e <= 3e-3 for every type pair, so 1 - e + e^2 equals 1 / (1 + e) to
    // within e^3 < 2^-25 and the kernel needs no reciprocal there.
    // Two more tiers for ligands that have left the protein far behind:
    // beyond lean_e_d, e <= 2^-12.5, so the e^2 term is below 2^-25 and
    // n (1 - e) holds; beyond lean_soft_d, e <= 2^-25 and the pair is n.
    // soft_reach(eps): the smallest distance from `from` on at which
    // e <= eps for every type pair (1e30 when the softening does not fall
    // with d).
    auto soft_reach = [&](double eps, double from) {
        const double ljc_ = ep.lj0 != 0.0f ? (double)ep.lj0 : 1.0;
        const double a = ep.lj0 / ljc_, b = ep.lj0 / ljc_ * ep.lj1;
        if (a < 0.0 || b < 0.0) return 1e30;
        double reach = from;
        for (int tp = 0; tp < GD_MAXTP1; ++tp)
            for (int tl = 0; tl < GD_MAXTP2; ++tl) {
                const double p1 = (double)(float)(ljc_ * ep.p1a[tl][tp]);
                auto e_at = [&](double d) { return p1 * (a + b * d * d) / std::pow(d, 9); };
                if (p1 <= 0.0 || e_at(reach) <= eps) continue;
                double lo = reach, hi = 65536.0;
                for (int it = 0; it < 80; ++it) {
                    const double m = 0.5 * (lo + hi);
                    (e_at(m) > eps ? lo : hi) = m;
                }
                reach = hi;
            }
        return reach + 0.05;
    };
    const double vfar_d = soft_reach(3.0e-3, far_d);
    const double lean_e_d = vfar_d < 1e30 ? soft_reach(std::ldexp(std::sqrt(0.5), -12), vfar_d) : 1e30;
    const double lean_soft_d = vfar_d < 1e30 ? soft_reach(std::ldexp(1.0, -25), lean_e_d) : 1e30;
    // One copy of the points and ladders per ladder centre C_k = k center_dr
    // (1,1,1)/sqrt(3): within each segment the points go by their smallest
    // distance from C_k over the conformations, and the ladders count
    // distances from C_k (copy 0, C_0 = the pocket centre, is the order
    // above).  The kernel takes the copy whose centre is nearest the wave's
    // atoms; any copy gives valid bounds.
    const size_t nseg = seg.size();
    // centre spacing: the last centre lies a ligand's extent (12 A) beyond
    // the protein's radius plus the very far distance, never closer than 4 A
    float center_dr;
    double t_prot_r = 0.0;
    {
        double prot_r = 0.0;
        for (int c = 0; c < nconf; ++c)
            for (int i = 0; i < np; ++i) {
                const size_t k = (size_t)c * np + i;
                const double x = local(cx->prt_x[k], origin[0]), y = local(cx->prt_y[k], origin[1]),
                             z = local(cx->prt_z[k], origin[2]);
                prot_r = std::max(prot_r, std::sqrt(x * x + y * y + z * z));
            }
        const double reach = prot_r + std::min(vfar_d, far_d + 64.0) + 12.0;
        center_dr = (float)std::max(4.0, reach / (kCenters - 1));
        t_prot_r = prot_r;
    }
    std::vector<float2> pa((size_t)kCenters * np);
    std::vector<float> pw((size_t)kCenters * np);
    std::vector<float4> pt((size_t)kCenters * nconf * np);
    std::vector<float> pq((size_t)kCenters * nconf * np);
    std::vector<int> ladder((size_t)kCenters * nseg * kFarRungs), ladder_in((size_t)kCenters * nseg * kFarRungs);
    std::vector<double> dmin(np), dmax(np);
    std::vector<int> ordk(np);
    for (int kc = 0; kc < kCenters; ++kc) {
        const float ccf = (float)kc * center_dr * 0.57735026918962576f; // each coordinate of C_k, rounded as the kernel does
        const double cc = (double)ccf;
        for (int i = 0; i < np; ++i) {
            dmin[i] = 1e300;
            dmax[i] = 0.0;
        }
        for (int c = 0; c < nconf; ++c)
            for (int i = 0; i < np; ++i) {
                const size_t k = (size_t)c * np + i;
                const double x = (double)local(cx->prt_x[k], origin[0]) - cc, y = (double)local(cx->prt_y[k], origin[1]) - cc,
                             z = (double)local(cx->prt_z[k], origin[2]) - cc;
                const double r = std::sqrt(x * x + y * y + z * z);
                dmin[i] = std::min(dmin[i], r);
                dmax[i] = std::max(dmax[i], r);
            }
        ordk = ord;
        if (kc > 0)
            for (const int4 &g : seg)
                std::stable_sort(ordk.begin() + g.y, ordk.begin() + g.z,
                                 [&](int x, int y) { return dmin[x] < dmin[y]; });
        for (int i = 0; i < np; ++i) {
            const int s = ordk[i];
            float2 v;
            v.x = 0.5f * hpcw[s]; // the factored contact polynomial leaves a factor 1/2 to the weight
            std::memcpy(&v.y, &cx->prt_t[s], 4);
            pa[(size_t)kc * np + i] = v;
            pw[(size_t)kc * np + i] = v.x;
        }
        // q = p - C_k from the frame coordinates the kernel's atoms use (the
        // float local values), in double: q' = -2 q rounds once, |q|^2 once
        for (int c = 0; c < nconf; ++c)
            for (int i = 0; i < np; ++i) {
                const size_t s = (size_t)c * np + ordk[i];
                const double qx = (double)local(cx->prt_x[s], origin[0]) - cc,
                             qy = (double)local(cx->prt_y[s], origin[1]) - cc,
                             qz = (double)local(cx->prt_z[s], origin[2]) - cc;
                const size_t o = ((size_t)kc * nconf + c) * np + i;
                pt[o] = make_float4((float)(-2.0 * qx), (float)(-2.0 * qy), (float)(-2.0 * qz), cx->prt_ele[ordk[i]]);
                pq[o] = (float)(qx * qx + qy * qy + qz * qz);
            }
        for (size_t g = 0; g < nseg; ++g)
            for (int k = 0; k < kFarRungs; ++k) {
                const double rk = kFarR0 + k * kFarDr;
                int i = seg[g].y;
                while (i < seg[g].z && dmin[ordk[i]] < rk) ++i;
                ladder[((size_t)kc * nseg + g) * kFarRungs + k] = i;
                int m = seg[g].y;
                while (m < seg[g].z && dmax[ordk[m]] <= rk) ++m;
                ladder_in[((size_t)kc * nseg + g) * kFarRungs + k] = m;
            }
    }

    const int id = ctx->next_id++;
    TargetRes &t = ctx->targets[id];
    // a failed upload leaves no target behind
    struct Drop {
        gd_ctx *ctx;
        int id;
        bool keep = false;
        ~Drop()
        {
            if (keep) return;
            (void)hipStreamSynchronize(ctx->stream);
            ctx->targets.erase(id);
        }
    } drop{ctx, id};
    GD_HIP(upload(t.pt, pt.data(), sizeof(float4) * pt.size(), ctx->stream));
    GD_HIP(upload(t.pq, pq.data(), sizeof(float) * pq.size(), ctx->stream));
    GD_HIP(upload(t.pa, pa.data(), sizeof(float2) * pa.size(), ctx->stream));
    GD_HIP(upload(t.pw, pw.data(), sizeof(float) * pw.size(), ctx->stream));
    GD_HIP(upload(t.pp, pp.data(), sizeof(float) * pp.size(), ctx->stream));
    GD_HIP(upload(t.seg, seg.data(), sizeof(int4) * seg.size(), ctx->stream));
    GD_HIP(upload(t.ladder, ladder.data(), sizeof(int) * ladder.size(), ctx->stream));
    GD_HIP(upload(t.ladder_in, ladder_in.data(), sizeof(int) * ladder_in.size(), ctx->stream));
    t.dt.ladder = t.ladder.as<int>();
    t.dt.ladder_in = t.ladder_in.as<int>();
    t.dt.far_d = (float)far_d;
    t.dt.mid_d = (float)std::min(mid_d, far_d);
    t.dt.vfar_d = (float)std::min(vfar_d, 1e30);
    t.dt.lean_e_d = (float)std::min(lean_e_d, 1e30);
    t.dt.lean_soft_d = (float)std::min(lean_soft_d, 1e30);
    t.dt.center_dr = center_dr;
    t.dt.inv_center_dr = 1.0f / center_dr;
    t.dt.prot_r = (float)t_prot_r + 1e-3f; // rounded up: a rung at or beyond it holds every point
    t.dt.pt = t.pt.as<float4>();
    t.dt.pq = t.pq.as<float>();
    t.dt.pa = t.pa.as<float2>();
    t.dt.pw = t.pw.as<float>();
    t.dt.pp = t.pp.as<float>();
    t.dt.seg = t.seg.as<int4>();
    t.dt.npoint = np;
    t.dt.npad = np;
    t.dt.nseg = (int)seg.size();
    t.dt.nconf = nconf;
    for (int d = 0; d < 3; ++d) t.origin[d] = origin[d];
    DevScalars &s = t.sc;
    s.lj_a = (float)(ep.lj0 / ljc);
    s.lj_b = (float)(ep.lj0 / ljc * ep.lj1);
    s.inv_lj0 = (float)(1.0 / ljc);
    s.inv_el1 = 1.0f / ep.el1;
    {
        const double e1 = ep.el1;
        s.el_e0 = (float)(e1 * ep.el0);
        s.el_a = (float)((double)ep.a1 * e1 * e1 * e1);
        s.el_b = (float)((double)ep.b1 * e1 * e1 * e1 * e1);
        s.el_dc2 = (float)(1.0 / (e1 * e1));
        s.pmf_a = (float)(-0.5 * log2e);
    }
    s.kde2 = ep.kde2;
    s.kde3inv = 1.0f / ep.kde3;
    for (int i = 0; i < GD_MAXTP2; ++i) {
        s.hpl0[i] = ep.hpl0[i];
        s.hpl1[i] = ep.hpl1[i];
        s.hpl2[i] = ep.hpl2[i];
    }
    for (int i = 0; i < GD_MAXWEI; ++i) {
        s.a_para[i] = ep.a_para[i];
        s.b_para[i] = ep.b_para[i];
        s.w[i] = ep.w[i];
    }
    t.n_prt = nconf;
    t.npoint = np;
    for (const int4 &g : seg) {
        t.seg_host.push_back(g.x);
        t.seg_host.push_back(g.y);
        t.seg_host.push_back(g.z);
    }
    t.bytes = t.pt.bytes + t.pq.bytes + t.pa.bytes + t.pw.bytes + t.pp.bytes + t.seg.bytes + t.ladder.bytes + t.ladder_in.bytes +
              sizeof(DevParams);
    // the host copy stays alive until the stream synchronisation below
    const DevParams prm_host{t.dt, t.sc};
    GD_HIP(upload(t.params, &prm_host, sizeof(prm_host), ctx->stream));
    t.dev_index.assign(np, 0);
    for (int i = 0; i < np; ++i) t.dev_index[ord[i]] = i;
    t.prt_t.assign(cx->prt_t, cx->prt_t + np);
    t.prt_c.assign(cx->prt_c, cx->prt_c + np);
    GD_HIP(hipStreamSynchronize(ctx->stream));
    drop.keep = true;
    *target_id = id;
    return GD_OK;
}

int gd_free_target(gd_ctx *ctx, int target_id)
{
    if (int rc = check_ctx(ctx)) return rc;
    for (auto &kv : ctx->jobs)
        if (kv.second.target == target_id) return fail(GD_ERR_STATE, "target still has jobs");
    if (!ctx->targets.erase(target_id)) return fail(GD_ERR_STATE, "unknown target");
    return GD_OK;
}

size_t gd_job_footprint(int lig_natom, int n_lig, int kde_npoint, int mcs_entries, int mcs_nrow, int n_tmp,
                        int n_rep, int record_cap)
{
    const size_t a = (size_t)lig_natom;
    return 2 * sizeof(float4) * a * n_lig + (sizeof(int) + sizeof(int2)) * a + sizeof(float4) * kde_npoint +
           sizeof(int) * (mcs_nrow + 1) + sizeof(float4) * mcs_entries + sizeof(float) * mcs_nrow +
           sizeof(float) * n_tmp + sizeof(gd_ReplicaMC) * (size_t)n_rep * (1 + record_cap) +
           3 * sizeof(int) * (size_t)n_rep /* next_entry, McLog counters */ + sizeof(DevJob) +
           17 * 256 // allocation granularity slack (one block per table)
           + 64 * 1024; // PSP point rows (100 bytes per point, up to ~650 points)
}

int gd_upload_job(gd_ctx *ctx, int target_id, const gd_Complex *cx, uint64_t seed, int *job_id)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (!cx || !job_id) return fail(GD_ERR_ARG, "null argument");
    if (int rc = check_complex_arrays(cx, false, true)) return rc;
    auto tit = ctx->targets.find(target_id);
    if (tit == ctx->targets.end()) return fail(GD_ERR_STATE, "unknown target");
    const gd_ComplexSize &z = cx->size;
    if (z.lig_natom <= 0 || z.lig_natom > 1024) return fail(GD_ERR_ARG, "ligand atom count must be 1..1024");
    if (z.n_lig <= 0 || z.n_tmp <= 0 || z.n_tmp > 64 || z.kde_npoint <= 0 || z.mcs_nrow <= 0)
        return fail(GD_ERR_ARG, "bad job sizes (n_tmp must be 1..64)");
    if (z.n_prt != tit->second.n_prt) return fail(GD_ERR_ARG, "job and target protein conformation counts differ");
    if (z.prt_npoint != tit->second.npoint) return fail(GD_ERR_ARG, "job and target point counts differ");
    if (z.n_rep != z.n_lig * z.n_prt * z.n_tmp) return fail(GD_ERR_ARG, "n_rep != n_lig * n_prt * n_tmp");
    const int na = z.lig_natom;
    if (int rc = check_replica_layout(cx->replica, z.n_lig, z.n_prt, z.n_tmp)) return rc;

    std::vector<float4> lig3((size_t)z.n_lig * na), lig2((size_t)z.n_lig * na);
    for (size_t i = 0; i < lig3.size(); ++i) {
        lig3[i] = make_float4(cx->lig_x3[i], cx->lig_y3[i], cx->lig_z3[i], cx->lig_c3[i]);
        lig2[i] = make_float4(cx->lig_x2[i], cx->lig_y2[i], cx->lig_z2[i], 0.0f);
    }
    std::vector<int> lt(cx->lig_t3, cx->lig_t3 + na);
    for (int v : lt)
        if (v < 0 || v >= GD_MAXTP2) return fail(GD_ERR_ARG, "ligand atom type out of range");
    std::vector<int2> krange(na);
    for (int i = 0; i < na; ++i) {
        int b = -1, e = -1;
        for (int k = 0; k < z.kde_npoint; ++k)
            if (cx->kde_t[k] == lt[i]) {
                if (b < 0) b = k;
                e = k + 1;
            }
        if (b < 0) b = e = 0;
        for (int k = b; k < e; ++k)
            if (cx->kde_t[k] != lt[i]) return fail(GD_ERR_ARG, "KDE table is not sorted by type");
        krange[i] = make_int2(b, e);
    }
    std::vector<float4> kde(z.kde_npoint);
    const float *org = tit->second.origin;
    for (int k = 0; k < z.kde_npoint; ++k)
        kde[k] = make_float4(local(cx->kde_x[k], org[0]), local(cx->kde_y[k], org[1]), local(cx->kde_z[k], org[2]), 0.0f);
    if (cx->mcs_rowptr[0] != 0) return fail(GD_ERR_ARG, "MCS row pointers must start at 0");
    const int nent = cx->mcs_rowptr[z.mcs_nrow];
    std::vector<int> mrow(cx->mcs_rowptr, cx->mcs_rowptr + z.mcs_nrow + 1);
    std::vector<float4> ment(std::max(nent, 1));
    for (int k = 0; k < nent; ++k) {
        const int ai = cx->mcs_i[k];
        if (ai < 0 || ai >= na) return fail(GD_ERR_ARG, "MCS atom index out of range");
        float w;
        std::memcpy(&w, &ai, 4);
        ment[k] = make_float4(local(cx->mcs_x[k], org[0]), local(cx->mcs_y[k], org[1]), local(cx->mcs_z[k], org[2]), w);
    }
    for (int m = 0; m < z.mcs_nrow; ++m)
        if (mrow[m + 1] <= mrow[m]) return fail(GD_ERR_ARG, "empty MCS row");
    std::vector<float> mtcc(cx->mcs_tcc, cx->mcs_tcc + z.mcs_nrow);
    std::vector<float> mbeta(z.n_tmp);
    for (int t = 0; t < z.n_tmp; ++t) mbeta[t] = cx->temp[t].minus_beta;

    // the job's PSP points (the PSP table comes with the ligand's .ff file):
    // class-2 points whose residue has a nonzero entry, in device order
    const TargetRes &tr = tit->second;
    const int np = z.prt_npoint;
    for (int s = 0; s < np; ++s)
        if (cx->prt_t[s] != tr.prt_t[s] || cx->prt_c[s] != tr.prt_c[s])
            return fail(GD_ERR_ARG, "job protein topology differs from the resident target");
    std::vector<int> psp_pt;
    std::vector<float> psp_row;
    for (int s = 0; s < np; ++s) {
        if (cx->prt_c[s] != 2) continue;
        const int res = cx->prt_seq3r[s];
        if (res < 0 || res >= cx->psp_width) return fail(GD_ERR_ARG, "residue serial outside the PSP table");
        bool any = false;
        for (int t = 0; t < GD_MAXTP2; ++t) any |= cx->psp[(size_t)t * cx->psp_width + res] != 0.0f;
        if (!any) continue;
        psp_pt.push_back(tr.dev_index[s]);
        for (int t = 0; t < GD_MAXTP2; ++t) psp_row.push_back(cx->psp[(size_t)t * cx->psp_width + res]);
    }
    const int npsp = (int)psp_pt.size();
    if (psp_pt.empty()) {
        psp_pt.push_back(0);
        psp_row.assign(GD_MAXTP2, 0.0f);
    }

    const int id = ctx->next_id++;
    JobRes &j = ctx->jobs[id];
    if (ctx->job_arena)
        for (DevBuf *b : {&j.lig3, &j.lig2, &j.lt, &j.krange, &j.kde, &j.mrow, &j.ment, &j.mtcc, &j.mbeta, &j.psp_pt,
                          &j.psp_row, &j.rep, &j.rec, &j.next_entry, &j.self, &j.kstream, &j.mclog})
            b->arena = ctx->job_arena;
    // a failed upload leaves no job behind (the device buffers free with it)
    struct Drop {
        gd_ctx *ctx;
        int id;
        bool keep = false;
        ~Drop()
        {
            if (keep) return;
            (void)hipStreamSynchronize(ctx->stream);
            ctx->jobs.erase(id);
        }
    } drop{ctx, id};
    j.target = target_id;
    hipStream_t s = ctx->stream;
    GD_HIP(upload(j.lig3, lig3.data(), sizeof(float4) * lig3.size(), s));
    GD_HIP(upload(j.lig2, lig2.data(), sizeof(float4) * lig2.size(), s));
    GD_HIP(upload(j.lt, lt.data(), sizeof(int) * lt.size(), s));
    GD_HIP(upload(j.krange, krange.data(), sizeof(int2) * krange.size(), s));
    GD_HIP(upload(j.kde, kde.data(), sizeof(float4) * kde.size(), s));
    GD_HIP(upload(j.mrow, mrow.data(), sizeof(int) * mrow.size(), s));
    GD_HIP(upload(j.ment, ment.data(), sizeof(float4) * ment.size(), s));
    GD_HIP(upload(j.mtcc, mtcc.data(), sizeof(float) * mtcc.size(), s));
    GD_HIP(upload(j.mbeta, mbeta.data(), sizeof(float) * mbeta.size(), s));
    GD_HIP(upload(j.psp_pt, psp_pt.data(), sizeof(int) * psp_pt.size(), s));
    GD_HIP(upload(j.psp_row, psp_row.data(), sizeof(float) * psp_row.size(), s));
    GD_HIP(upload(j.rep, cx->replica, sizeof(gd_ReplicaMC) * z.n_rep, s));
    GD_HIP(j.next_entry.alloc(sizeof(int) * z.n_rep));
    GD_HIP(hipMemsetAsync(j.next_entry.p, 0, sizeof(int) * z.n_rep, s));
    GD_HIP(j.self.alloc(sizeof(DevJob)));
    DevJob &h = j.hj;
    h.lig3 = j.lig3.as<float4>();
    h.lig2 = j.lig2.as<float4>();
    h.lt = j.lt.as<int>();
    h.krange = j.krange.as<int2>();
    h.kde = j.kde.as<float4>();
    h.mrow = j.mrow.as<int>();
    h.ment = j.ment.as<float4>();
    h.mtcc = j.mtcc.as<float>();
    h.mbeta = j.mbeta.as<float>();
    h.psp_pt = j.psp_pt.as<int>();
    h.psp_row = j.psp_row.as<float>();
    h.npsp = npsp;
    h.rep = j.rep.as<gd_ReplicaMC>();
    h.rec = nullptr;
    h.next_entry = j.next_entry.as<int>();
    h.natom = na;
    h.nrow = z.mcs_nrow;
    h.rec_cap = 0;
    h.n_rep = z.n_rep;
    h.n_lig = z.n_lig;
    h.n_prt = z.n_prt;
    h.n_tmp = z.n_tmp;
    for (int d = 0; d < 3; ++d) {
        h.center[d] = local(cx->lig_center[d], org[d]);
        h.pocket[d] = local(cx->prt_pocket_center[d], org[d]);
    }
    for (int k = 0; k < 6; ++k) h.scale[k] = cx->mcpara.move_scale[k];
    h.key0 = (uint32_t)seed;
    h.key1 = (uint32_t)(seed >> 32);
    h.kstream = nullptr;
    h.kstream_steps = 0;
    GD_HIP(j.mclog.alloc(sizeof(int) * 2 * (size_t)z.n_rep));
    GD_HIP(hipMemsetAsync(j.mclog.p, 0, j.mclog.bytes, s));
    h.mc_acc = j.mclog.as<int>();
    h.ex_acc = h.mc_acc + z.n_rep;
    GD_HIP(hipMemcpyAsync(j.self.p, &h, sizeof(DevJob), hipMemcpyHostToDevice, s));
    j.n_rep = z.n_rep;
    j.n_lig = z.n_lig;
    j.n_prt = z.n_prt;
    j.n_tmp = z.n_tmp;
    j.natom = na;
    j.nrow = z.mcs_nrow;
    j.seed = seed;
    j.rep_prt.resize(z.n_rep);
    for (int r = 0; r < z.n_rep; ++r) j.rep_prt[r] = cx->replica[r].idx_prt;
    j.bytes = j.lig3.bytes + j.lig2.bytes + j.lt.bytes + j.krange.bytes + j.kde.bytes + j.mrow.bytes + j.ment.bytes +
              j.mtcc.bytes + j.mbeta.bytes + j.psp_pt.bytes + j.psp_row.bytes + j.rep.bytes + j.next_entry.bytes +
              j.self.bytes + j.mclog.bytes;
    GD_HIP(hipStreamSynchronize(s));
    drop.keep = true;
    *job_id = id;
    return GD_OK;
}

int gd_set_stream(gd_ctx *ctx, int job_id, uint64_t seed, const uint32_t *k_words, int n_steps)
{
    if (int rc = check_ctx(ctx)) return rc;
    JobRes *j = find_job(ctx, job_id);
    if (!j) return fail(GD_ERR_STATE, "unknown job");
    if (k_words && n_steps <= 0) return fail(GD_ERR_ARG, "explicit stream words need n_steps > 0");
    hipStream_t s = ctx->stream;
    DevJob &h = j->hj;
    h.key0 = (uint32_t)seed;
    h.key1 = (uint32_t)(seed >> 32);
    j->seed = seed;
    if (k_words) {
        const size_t n = (size_t)j->n_rep * n_steps * 7;
        for (size_t i = 0; i < n; ++i)
            if (k_words[i] > 0x7fffffffu) return fail(GD_ERR_ARG, "stream word above RAND_MAX (2^31 - 1)");
        GD_HIP(upload(j->kstream, k_words, n * sizeof(uint32_t), s));
        h.kstream = j->kstream.as<uint32_t>();
        h.kstream_steps = n_steps;
        j->kstream_steps = n_steps;
    } else {
        j->kstream.release();
        h.kstream = nullptr;
        h.kstream_steps = 0;
        j->kstream_steps = 0;
    }
    GD_HIP(hipMemcpyAsync(j->self.p, &h, sizeof(DevJob), hipMemcpyHostToDevice, s));
    GD_HIP(hipStreamSynchronize(s));
    return GD_OK;
}

int gd_read_mclog(gd_ctx *ctx, int job_id, int *mc_accepts, int *exchange_accepts)
{
    if (int rc = check_ctx(ctx)) return rc;
    JobRes *j = find_job(ctx, job_id);
    if (!j) return fail(GD_ERR_STATE, "unknown job");
    const size_t n = (size_t)j->n_rep;
    if (mc_accepts) GD_HIP(hipMemcpyAsync(mc_accepts, j->mclog.p, sizeof(int) * n, hipMemcpyDeviceToHost, ctx->stream));
    if (exchange_accepts)
        GD_HIP(hipMemcpyAsync(exchange_accepts, j->mclog.as<int>() + n, sizeof(int) * n, hipMemcpyDeviceToHost,
                              ctx->stream));
    GD_HIP(hipStreamSynchronize(ctx->stream));
    return GD_OK;
}

int gd_free_job(gd_ctx *ctx, int job_id)
{
    if (int rc = check_ctx(ctx)) return rc;
    (void)hipStreamSynchronize(ctx->stream);
    if (!ctx->jobs.erase(job_id)) return fail(GD_ERR_STATE, "unknown job");
    return GD_OK;
}

int gd_montecarlo(gd_ctx *ctx, int job_id, int s1, int s2max)
{
    if (int rc = check_ctx(ctx)) return rc;
    JobRes *j = find_job(ctx, job_id);
    if (!j) return fail(GD_ERR_STATE, "unknown job");
    if (s1 < 0 || s2max <= 0) return fail(GD_ERR_ARG, "bad step range");
    if (int rc = check_stream_range(*j, (long)s1 + s2max)) return rc;
    if (int rc = ensure_records(ctx, *j, std::max(s2max, GD_STEPS_PER_DUMP))) return rc;
    const TargetRes &t = ctx->targets.at(j->target);
    const Shape sh = plain_shape(j->natom);
    std::vector<DevTask> tasks(j->n_rep);
    for (int r = 0; r < j->n_rep; ++r)
        tasks[r] = DevTask{0, r, 1, j->rep_prt[r], s1, s2max, sh.slices, sh.slot_waves};
    if (int rc = run_tasks(ctx, t, j->self.as<DevJob>(), tasks, sh.waves, 1, true, true)) return rc;
    GD_HIP(hipStreamSynchronize(ctx->stream));
    return GD_OK;
}

int gd_exchange(gd_ctx *ctx, int job_id, int round, int *n_accept)
{
    if (int rc = check_ctx(ctx)) return rc;
    JobRes *j = find_job(ctx, job_id);
    if (!j) return fail(GD_ERR_STATE, "unknown job");
    if (round < 0) return fail(GD_ERR_ARG, "negative exchange round");
    if (!ctx->counter.p) GD_HIP(ctx->counter.alloc(sizeof(int)));
    GD_HIP(hipMemsetAsync(ctx->counter.p, 0, sizeof(int), ctx->stream));
    GD_HIP(launch_exchange(j->self.as<DevJob>(), 0, j->n_lig * j->n_prt, j->n_lig, j->n_tmp, round,
                           ctx->counter.as<int>(), ctx->stream, nullptr));
    int cnt = 0;
    GD_HIP(hipMemcpyAsync(&cnt, ctx->counter.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    GD_HIP(hipStreamSynchronize(ctx->stream));
    if (n_accept) *n_accept = cnt;
    return GD_OK;
}

int gd_energy(gd_ctx *ctx, int job_id, int n, const int *rep_idx, const float *move, float *e_out)
{
    if (int rc = check_ctx(ctx)) return rc;
    JobRes *j = find_job(ctx, job_id);
    if (!j) return fail(GD_ERR_STATE, "unknown job");
    if (n < 0 || (n > 0 && (!rep_idx || !move || !e_out))) return fail(GD_ERR_ARG, "bad pose arguments");
    if (n == 0) return GD_OK;
    for (int i = 0; i < n; ++i)
        if (rep_idx[i] < 0 || rep_idx[i] >= j->n_rep) return fail(GD_ERR_ARG, "replica index out of range");
    const TargetRes &t = ctx->targets.at(j->target);
    const Shape sh = plain_shape(j->natom);
    std::vector<DevTask> tasks(n);
    for (int i = 0; i < n; ++i) tasks[i] = DevTask{0, i, 1, j->rep_prt[rep_idx[i]], 0, 1, sh.slices, sh.slot_waves};
    hipStream_t s = ctx->stream;
    GD_HIP(upload(ctx->task_buf, tasks.data(), sizeof(DevTask) * n, s));
    GD_HIP(upload(ctx->pose_rep, rep_idx, sizeof(int) * n, s));
    GD_HIP(upload(ctx->pose_move, move, sizeof(float) * 6 * n, s));
    if (ctx->pose_out.bytes < sizeof(float) * GD_MAXWEI * n) GD_HIP(ctx->pose_out.alloc(sizeof(float) * GD_MAXWEI * n));
    GD_HIP(launch_energy(t.params.as<DevParams>(), j->self.as<DevJob>(), ctx->task_buf.as<DevTask>(), n, sh.waves, true,
                         ctx->pose_rep.as<int>(), ctx->pose_move.as<float>(), ctx->pose_out.as<float>(), s));
    GD_HIP(hipMemcpyAsync(e_out, ctx->pose_out.p, sizeof(float) * GD_MAXWEI * n, hipMemcpyDeviceToHost, s));
    GD_HIP(hipStreamSynchronize(s));
    return GD_OK;
}

int gd_read_replicas(gd_ctx *ctx, int job_id, gd_ReplicaMC *out)
{
    if (int rc = check_ctx(ctx)) return rc;
    JobRes *j = find_job(ctx, job_id);
    if (!j || !out) return fail(GD_ERR_ARG, "unknown job or null output");
    GD_HIP(hipMemcpyAsync(out, j->rep.p, sizeof(gd_ReplicaMC) * j->n_rep, hipMemcpyDeviceToHost, ctx->stream));
    GD_HIP(hipStreamSynchronize(ctx->stream));
    return GD_OK;
}

int gd_write_replicas(gd_ctx *ctx, int job_id, const gd_ReplicaMC *in)
{
    if (int rc = check_ctx(ctx)) return rc;
    JobRes *j = find_job(ctx, job_id);
    if (!j || !in) return fail(GD_ERR_ARG, "unknown job or null input");
    if (int rc = check_replica_layout(in, j->n_lig, j->n_prt, j->n_tmp)) return rc;
    GD_HIP(hipMemcpyAsync(j->rep.p, in, sizeof(gd_ReplicaMC) * j->n_rep, hipMemcpyHostToDevice, ctx->stream));
    GD_HIP(hipStreamSynchronize(ctx->stream));
    return GD_OK;
}

int gd_read_records(gd_ctx *ctx, int job_id, gd_ReplicaMC *out, int *next_entry, int cap)
{
    if (int rc = check_ctx(ctx)) return rc;
    JobRes *j = find_job(ctx, job_id);
    if (!j) return fail(GD_ERR_STATE, "unknown job");
    if (!j->rec.p) return fail(GD_ERR_STATE, "no records yet");
    std::vector<int> ne(j->n_rep);
    GD_HIP(hipMemcpyAsync(ne.data(), j->next_entry.p, sizeof(int) * j->n_rep, hipMemcpyDeviceToHost, ctx->stream));
    GD_HIP(hipStreamSynchronize(ctx->stream));
    if (next_entry) std::copy(ne.begin(), ne.end(), next_entry);
    if (out) {
        if (cap < j->rec_cap) {
            for (int r = 0; r < j->n_rep; ++r)
                if (ne[r] > cap) return fail(GD_ERR_ARG, "record capacity too small");
            for (int r = 0; r < j->n_rep; ++r)
                GD_HIP(hipMemcpyAsync(out + (size_t)r * cap, j->rec.as<gd_ReplicaMC>() + (size_t)r * j->rec_cap,
                                      sizeof(gd_ReplicaMC) * std::max(ne[r], 0), hipMemcpyDeviceToHost, ctx->stream));
        } else {
            GD_HIP(hipMemcpy2DAsync(out, sizeof(gd_ReplicaMC) * cap, j->rec.p, sizeof(gd_ReplicaMC) * j->rec_cap,
                                    sizeof(gd_ReplicaMC) * j->rec_cap, j->n_rep, hipMemcpyDeviceToHost, ctx->stream));
        }
        GD_HIP(hipStreamSynchronize(ctx->stream));
    }
    return GD_OK;
}

} // extern "C"
