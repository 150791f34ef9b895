// Batched form: every replica of several jobs against one resident target in
// one fused launch per chunk of steps, with replicas packed R per workgroup
// so their ligand atoms fill the 64-lane waves, the protein points split
// across `slices` waves, and workgroup shapes picked by a timing autotuner.
// The chunk sequence (Monte Carlo chunk, exchange round, step advance) can be
// captured once in a hipGraph and replayed.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "context.hpp"

namespace gd {

#define GD_HIP(call)                                                                                 \
    do {                                                                                             \
        hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess) return fail(GD_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace {

__global__ void advance_kernel(int *step_base, int dstep, int *round_base)
{
    if (threadIdx.x == 0) {
        *step_base += dstep;
        *round_base += 1;
    }
}

std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, int>, Shape> g_tune_cache;

} // namespace

// Replica packing of one job for a workgroup of W waves: slot waves SW
// (dividing W) hold R = floor(64 SW / N_L) replicas, the other factor of W
// slices the protein points.
static Shape packed_shape(int natom, int W, int SW, int max_rep_block)
{
    Shape s;
    s.waves = W;
    s.slot_waves = SW;
    s.slices = W / SW;
    s.nrep = std::max(1, std::min({(64 * SW) / natom, max_rep_block, 64}));
    return s;
}

// Deterministic choice for a given W.  Score = lane use x slot fill: lane
// use counts the lanes the job's replicas occupy over every workgroup it
// launches (partly filled last groups included); slot fill is
// tasks / (slots * rounds), rounds = ceil(tasks / slots), for `slots`
// workgroups resident at once.  Within 2% of the best score, the most slot
// waves win (fewest point slices: least cross-slice work per step).
Shape heuristic_shape(int natom, int W, int max_rep_block, int n_prt, int slots)
{
    slots = std::max(1, slots);
    Shape best = packed_shape(natom, W, W, max_rep_block); // natom > 64 W: caller rejects
    std::vector<std::pair<Shape, double>> cs;
    double top = -1.0;
    for (int SW = 1; SW <= W; ++SW) {
        if (W % SW || 64 * SW < natom) continue;
        const Shape s = packed_shape(natom, W, SW, max_rep_block);
        const long per_prt = (max_rep_block + s.nrep - 1) / s.nrep;
        const long tasks = (long)n_prt * per_prt;
        const double use = (double)max_rep_block * natom / ((double)per_prt * 64.0 * SW);
        const long rounds = (tasks + slots - 1) / slots;
        const double fill = (double)tasks / ((double)rounds * slots);
        cs.emplace_back(s, use * fill);
        top = std::max(top, use * fill);
    }
    int best_sw = -1;
    for (const auto &c : cs)
        if (c.second >= top * 0.98 && c.first.slot_waves > best_sw) {
            best_sw = c.first.slot_waves;
            best = c.first;
        }
    return best;
}

void build_tasks(const JobRes &j, int job_slot, const Shape &sh, int s1, int s2max, std::vector<DevTask> &out)
{
    const int block = j.n_tmp * j.n_lig; // replicas of one protein conformation are contiguous
    for (int p = 0; p < j.n_prt; ++p)
        for (int r = 0; r < block; r += sh.nrep) {
            DevTask t;
            t.job = job_slot;
            t.rep0 = p * block + r;
            t.nrep = std::min(sh.nrep, block - r);
            t.prt = p;
            t.s1 = s1;
            t.s2max = s2max;
            t.slices = sh.slices;
            t.slot_waves = sh.slot_waves;
            out.push_back(t);
        }
}

// Autotuner: time `trial` steps of every candidate workgroup shape on the
// job's own data (state saved and restored around the trials) and keep the
// fastest; results are cached per (N_L, N_P, n_rep, n_prt).
int autotune_shape(gd_ctx *ctx, const TargetRes &t, JobRes &j, int s1, bool compensated, Shape &out)
{
    const auto key = std::make_tuple(j.natom, t.npoint, j.n_rep, j.n_prt);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            out = it->second;
            return GD_OK;
        }
    }
    std::vector<Shape> cands;
    const int block = j.n_tmp * j.n_lig;
    // workgroups of 4-16 waves; 6 and 12 allow three slot waves (192 lanes,
    // which ligands of 48, 64 or 96 atoms fill without idle lanes), and two
    // 12-wave groups fill a CU at the kernel's 6 waves per SIMD
    for (int W : {4, 6, 8, 12, 16})
        for (int SW = 1; SW <= W; ++SW) {
            if (W % SW || 64 * SW < j.natom) continue;
            const Shape s = packed_shape(j.natom, W, SW, block);
            bool dup = false;
            for (const Shape &c : cands)
                dup |= c.waves == s.waves && c.nrep == s.nrep && c.slices == s.slices;
            if (!dup) cands.push_back(s);
        }
    if (cands.empty()) return fail(GD_ERR_ARG, "ligand too large for a 16-wave workgroup");
    DevBuf backup;
    GD_HIP(backup.alloc(j.rep.bytes));
    GD_HIP(hipMemcpyAsync(backup.p, j.rep.p, j.rep.bytes, hipMemcpyDeviceToDevice, ctx->stream));
    hipEvent_t e0, e1;
    GD_HIP(hipEventCreate(&e0));
    GD_HIP(hipEventCreate(&e1));
    // two timed steps per candidate; one, untimed warm-up dropped, once a
    // step holds more than ~1e11 pair evaluations (a few 100 ms)
    const double pairs = (double)j.n_rep * j.natom * t.npoint;
    const bool large = pairs > 1e11;
    const int trial = large ? 1 : 2;
    float best_ms = 1e30f;
    Shape best = cands[0];
    for (const Shape &s : cands) {
        std::vector<DevTask> tasks;
        build_tasks(j, 0, s, s1, trial, tasks);
        // warm the code object and caches once, then time
        for (int rep = large ? 1 : 0; rep < 2; ++rep) {
            GD_HIP(hipEventRecord(e0, ctx->stream));
            if (int rc = run_tasks(ctx, t, j.self.as<DevJob>(), tasks, s.waves, s.nrep, compensated, false)) return rc;
            GD_HIP(hipEventRecord(e1, ctx->stream));
            GD_HIP(hipMemcpyAsync(j.rep.p, backup.p, j.rep.bytes, hipMemcpyDeviceToDevice, ctx->stream));
            GD_HIP(hipEventSynchronize(e1));
        }
        float ms = 0.f;
        GD_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best_ms) {
            best_ms = ms;
            best = s;
        }
    }
    GD_HIP(hipStreamSynchronize(ctx->stream));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_cache[key] = best;
    }
    out = best;
    return GD_OK;
}

} // namespace gd

using namespace gd;

extern "C" int gd_job_shape(gd_ctx *ctx, int job_id, int *shape)
{
    if (!ctx || !shape) return fail(GD_ERR_ARG, "null argument");
    auto it = ctx->jobs.find(job_id);
    if (it == ctx->jobs.end()) return fail(GD_ERR_STATE, "unknown job");
    const Shape &s = it->second.last_shape;
    if (s.waves <= 0) return fail(GD_ERR_STATE, "the job has not run in a batched launch");
    shape[0] = s.waves;
    shape[1] = s.slot_waves;
    shape[2] = s.slices;
    shape[3] = s.nrep;
    return GD_OK;
}

extern "C" int gd_batch_run(gd_ctx *ctx, const int *job_ids, int n_jobs, const gd_batch_opts *opts,
                            double *elapsed_ms)
{
    if (!ctx || !job_ids || n_jobs <= 0 || !opts) return fail(GD_ERR_ARG, "null or empty argument");
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(GD_ERR_DEVICE, "hipSetDevice failed");
    if (opts->steps <= 0 || opts->exchange_interval < 0) return fail(GD_ERR_ARG, "bad step counts");
    std::vector<JobRes *> js;
    int target = -1;
    for (int i = 0; i < n_jobs; ++i) {
        auto it = ctx->jobs.find(job_ids[i]);
        if (it == ctx->jobs.end()) return fail(GD_ERR_STATE, "unknown job");
        if (target >= 0 && it->second.target != target) return fail(GD_ERR_ARG, "batched jobs must share a target");
        target = it->second.target;
        js.push_back(&it->second);
    }
    const TargetRes &t = ctx->targets.at(target);
    const bool comp = opts->compensated != 0;
    const int chunk = opts->exchange_interval > 0 ? opts->exchange_interval : std::min(opts->steps, GD_STEPS_PER_DUMP);
    const bool records = opts->keep_records != 0;
    hipStream_t s = ctx->stream;

    // workgroup shapes: explicit, autotuned (single job) or the packing
    // heuristic under one common wave count (several jobs)
    int W = opts->waves_per_group;
    std::vector<Shape> shapes(n_jobs);
    if (opts->slots_per_wave > 0 || opts->point_slices > 0) {
        if (W <= 0) return fail(GD_ERR_ARG, "explicit shapes need waves_per_group");
        for (int i = 0; i < n_jobs; ++i) {
            const int SW = std::max(1, W / std::max(1, opts->point_slices));
            if (W % SW || 64 * SW < js[i]->natom) return fail(GD_ERR_ARG, "shape does not fit the ligand");
            shapes[i] = packed_shape(js[i]->natom, W, SW, js[i]->n_tmp * js[i]->n_lig);
            if (opts->slots_per_wave > 0)
                shapes[i].nrep = std::max(1, std::min(shapes[i].nrep, opts->slots_per_wave * SW / js[i]->natom));
        }
    } else if (n_jobs == 1 && W <= 0) {
        if (int rc = autotune_shape(ctx, t, *js[0], 0, comp, shapes[0])) return rc;
        W = shapes[0].waves;
    } else {
        if (W <= 0) W = 12;
        // each job's share of the workgroups resident at once (two per CU)
        const int slots = std::max(1, 2 * std::max(ctx->cu_count, 1) / n_jobs);
        for (int i = 0; i < n_jobs; ++i) {
            if (js[i]->natom > 64 * W) return fail(GD_ERR_ARG, "ligand too large for the workgroup");
            shapes[i] = heuristic_shape(js[i]->natom, W, js[i]->n_tmp * js[i]->n_lig, js[i]->n_prt, slots);
        }
    }
    int max_rep = 1;
    for (const Shape &sh : shapes) max_rep = std::max(max_rep, sh.nrep);
    for (int i = 0; i < n_jobs; ++i) js[i]->last_shape = shapes[i];

    // device job table for this batch
    std::vector<DevJob> table(n_jobs);
    for (int i = 0; i < n_jobs; ++i) {
        if (records)
            if (int rc = ensure_records(ctx, *js[i], chunk)) return rc;
        table[i] = js[i]->hj;
    }
    DevBuf jobs_dev, tasks_dev, bases;
    GD_HIP(upload(jobs_dev, table.data(), sizeof(DevJob) * n_jobs, s));
    GD_HIP(bases.alloc(2 * sizeof(int)));
    int *step_base = bases.as<int>();
    int *round_base = step_base + 1;

    // tasks in decreasing cost order so heavy workgroups start first
    auto make_tasks = [&](int s2max, std::vector<DevTask> &tasks) {
        tasks.clear();
        for (int i = 0; i < n_jobs; ++i) build_tasks(*js[i], i, shapes[i], 0, s2max, tasks);
        std::stable_sort(tasks.begin(), tasks.end(), [&](const DevTask &x, const DevTask &y) {
            return (long)x.nrep * table[x.job].natom > (long)y.nrep * table[y.job].natom;
        });
    };
    std::vector<int> s1_of(n_jobs), round_of(n_jobs);
    for (int i = 0; i < n_jobs; ++i) {
        s1_of[i] = js[i]->steps_done;
        round_of[i] = js[i]->exchange_rounds;
        if (int rc = check_stream_range(*js[i], s1_of[i], (long)s1_of[i] + opts->steps)) return rc;
    }
    std::vector<DevTask> tasks;
    make_tasks(chunk, tasks);
    for (DevTask &tk : tasks) tk.s1 = s1_of[tk.job];
    GD_HIP(upload(tasks_dev, tasks.data(), sizeof(DevTask) * tasks.size(), s));
    const int zero2[2] = {0, 0};
    GD_HIP(hipMemcpyAsync(bases.p, zero2, sizeof(zero2), hipMemcpyHostToDevice, s));

    auto enqueue_period = [&](int s2) -> int {
        GD_HIP(launch_mc(t.dt, t.sc, jobs_dev.as<DevJob>(), tasks_dev.as<DevTask>(), (int)tasks.size(), W, max_rep,
                         comp, records, s, step_base));
        if (opts->exchange_interval > 0)
            for (int i = 0; i < n_jobs; ++i)
                GD_HIP(launch_exchange(jobs_dev.as<DevJob>() + i, js[i]->n_lig * js[i]->n_prt, js[i]->n_lig,
                                       js[i]->n_tmp, round_of[i], nullptr, s, round_base));
        hipLaunchKernelGGL(advance_kernel, dim3(1), dim3(64), 0, s, step_base, s2, round_base);
        GD_HIP(hipGetLastError());
        return GD_OK;
    };

    const int n_full = opts->steps / chunk;
    const int rest = opts->steps - n_full * chunk;
    hipEvent_t e0, e1;
    GD_HIP(hipEventCreate(&e0));
    GD_HIP(hipEventCreate(&e1));
    GD_HIP(hipEventRecord(e0, s));
    if (n_full > 0) {
        if (opts->use_graph) {
            hipGraph_t graph = nullptr;
            hipGraphExec_t exec = nullptr;
            GD_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            int rc = enqueue_period(chunk);
            hipError_t ce = hipStreamEndCapture(s, &graph);
            if (rc) return rc;
            GD_HIP(ce);
            GD_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
            for (int c = 0; c < n_full; ++c) GD_HIP(hipGraphLaunch(exec, s));
            GD_HIP(hipStreamSynchronize(s));
            (void)hipGraphExecDestroy(exec);
            (void)hipGraphDestroy(graph);
        } else {
            for (int c = 0; c < n_full; ++c)
                if (int rc = enqueue_period(chunk)) return rc;
        }
    }
    if (rest > 0) {
        std::vector<DevTask> tail;
        make_tasks(rest, tail);
        for (DevTask &tk : tail) tk.s1 = s1_of[tk.job];
        GD_HIP(upload(tasks_dev, tail.data(), sizeof(DevTask) * tail.size(), s));
        tasks.swap(tail);
        GD_HIP(launch_mc(t.dt, t.sc, jobs_dev.as<DevJob>(), tasks_dev.as<DevTask>(), (int)tasks.size(), W, max_rep,
                         comp, records, s, step_base));
    }
    GD_HIP(hipEventRecord(e1, s));
    GD_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    GD_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (elapsed_ms) *elapsed_ms = ms;
    for (int i = 0; i < n_jobs; ++i) {
        js[i]->steps_done += opts->steps;
        if (opts->exchange_interval > 0) js[i]->exchange_rounds += n_full;
    }
    return GD_OK;
}
