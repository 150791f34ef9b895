// Batched form: every replica of several jobs against one resident target in
// fused launches per chunk of steps, with replicas packed R per workgroup so
// their ligand atoms fill the 64-lane waves, the protein points split across
// `slices` waves, and workgroup shapes picked by a timing autotuner.  The
// tasks run as two workgroup groups, one launch each per chunk on streams of
// their own; a group's chunk sequence (Monte Carlo chunk, exchange round,
// step advance) can be captured once in a hipGraph and replayed.
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

// tasks of protein conformation p of a job: its consecutive replicas, R per task
static void build_unit_tasks(const JobRes &j, int job_slot, const Shape &sh, int p, int s1, int s2max,
                             std::vector<DevTask> &out)
{
    const int block = j.n_tmp * j.n_lig; // replicas of one protein conformation are contiguous
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

void build_tasks(const JobRes &j, int job_slot, const Shape &sh, int s1, int s2max, std::vector<DevTask> &out)
{
    for (int p = 0; p < j.n_prt; ++p) build_unit_tasks(j, job_slot, sh, p, s1, s2max, out);
}

namespace {

// Workgroup groups.  A launch ends when its slowest workgroup does, so when
// its task count is not a multiple of the workgroups resident at once the
// last round leaves CUs idle until the next launch (C2: 880 tasks over 512
// resident workgroups idles 14% of the GPU).  The batch is therefore split
// into groups of whole (job, protein conformation) units, each launched on
// a stream of its own with its own task table, step counters and graph:
// one group's tail overlaps the other's next launch.  Every temperature
// ladder lies within one unit, so the exchange rounds split with the
// groups, and the results do not depend on the split.
struct Group {
    struct Range {
        int job, p0, p1; // protein conformations p0 .. p1 - 1 of batch job `job`
    };
    std::vector<Range> ranges;
    std::vector<DevTask> tasks, tail;
    DevBuf tasks_dev, bases; // bases: step and exchange-round counters advanced in the graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    Group() = default;
    Group(const Group &) = delete;
    Group(Group &&o) noexcept
        : ranges(std::move(o.ranges)), tasks(std::move(o.tasks)), tail(std::move(o.tail)),
          tasks_dev(std::move(o.tasks_dev)), bases(std::move(o.bases)), graph(o.graph), exec(o.exec)
    {
        o.graph = nullptr;
        o.exec = nullptr;
    }
    ~Group()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
    int *step_base() const { return bases.as<int>(); }
    int *round_base() const { return bases.as<int>() + 1; }
};

// waits for the group streams when it goes out of scope: declared after the
// groups, it runs first, so no graph or buffer is released while in use
struct GroupsIdle {
    gd_ctx *ctx;
    ~GroupsIdle()
    {
        for (hipStream_t g : ctx->group_stream)
            if (g) (void)hipStreamSynchronize(g);
    }
};

struct Event {
    hipEvent_t e = nullptr;
    Event() = default;
    Event(const Event &) = delete;
    ~Event()
    {
        if (e) (void)hipEventDestroy(e);
    }
};

// The units (job i, conformation p) in order, cut into at most G runs of
// about equal task counts (a unit goes to the run holding its midpoint).
std::vector<Group> make_groups(const std::vector<JobRes *> &js, const std::vector<Shape> &shapes, int G)
{
    long total = 0;
    for (size_t i = 0; i < js.size(); ++i)
        total += (long)js[i]->n_prt * ((js[i]->n_tmp * js[i]->n_lig + shapes[i].nrep - 1) / shapes[i].nrep);
    std::vector<Group> gr(std::max(1, G));
    long before = 0;
    for (size_t i = 0; i < js.size(); ++i) {
        const long c = (js[i]->n_tmp * js[i]->n_lig + shapes[i].nrep - 1) / shapes[i].nrep;
        for (int p = 0; p < js[i]->n_prt; ++p) {
            const int g = std::min((int)gr.size() - 1, (int)(((double)before + 0.5 * c) * gr.size() / total));
            auto &rs = gr[g].ranges;
            if (!rs.empty() && rs.back().job == (int)i && rs.back().p1 == p)
                rs.back().p1 = p + 1;
            else
                rs.push_back({(int)i, p, p + 1});
            before += c;
        }
    }
    std::vector<Group> out;
    for (Group &g : gr)
        if (!g.ranges.empty()) out.push_back(std::move(g));
    return out;
}

// a group's tasks for s2max steps from the batch jobs' own step counts,
// in decreasing cost order so heavy workgroups start first
void group_tasks(const Group &g, const std::vector<JobRes *> &js, const std::vector<Shape> &shapes,
                 const std::vector<int> &s1_of, int s2max, std::vector<DevTask> &out)
{
    out.clear();
    for (const Group::Range &r : g.ranges)
        for (int p = r.p0; p < r.p1; ++p)
            build_unit_tasks(*js[r.job], r.job, shapes[r.job], p, s1_of[r.job], s2max, out);
    std::stable_sort(out.begin(), out.end(), [&](const DevTask &x, const DevTask &y) {
        return (long)x.nrep * js[x.job]->natom > (long)y.nrep * js[y.job]->natom;
    });
}

// order the group streams after the work on s (fork), or s after the group
// streams (join)
int fork_groups(gd_ctx *ctx, hipStream_t s, size_t n, hipEvent_t ev)
{
    GD_HIP(hipEventRecord(ev, s));
    for (size_t g = 0; g < n; ++g) GD_HIP(hipStreamWaitEvent(ctx->group_stream[g], ev, 0));
    return GD_OK;
}

int join_groups(gd_ctx *ctx, hipStream_t s, size_t n, Event *ev)
{
    for (size_t g = 0; g < n; ++g) {
        GD_HIP(hipEventRecord(ev[g].e, ctx->group_stream[g]));
        GD_HIP(hipStreamWaitEvent(s, ev[g].e, 0));
    }
    return GD_OK;
}

} // namespace

// Autotuner: time a short run of consecutive launches of every candidate
// workgroup shape on the job's own data (state saved and restored around the
// trials), launched in workgroup groups as the batched runs are, and keep the
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
    if (int rc = ensure_group_streams(ctx)) return rc;
    hipStream_t s = ctx->stream;
    // the trials change the replica states and the acceptance counters
    DevBuf backup, backup_log;
    GD_HIP(backup.alloc(j.rep.bytes));
    GD_HIP(backup_log.alloc(j.mclog.bytes));
    GD_HIP(hipMemcpyAsync(backup.p, j.rep.p, j.rep.bytes, hipMemcpyDeviceToDevice, s));
    GD_HIP(hipMemcpyAsync(backup_log.p, j.mclog.p, j.mclog.bytes, hipMemcpyDeviceToDevice, s));
    Event e0, e1, ef, ej[gd_ctx::kGroups];
    for (Event *e : {&e0, &e1, &ef}) GD_HIP(hipEventCreate(&e->e));
    for (Event &e : ej) GD_HIP(hipEventCreateWithFlags(&e.e, hipEventDisableTiming));
    // Each candidate is timed as the real chunk sequence runs: three
    // consecutive 2-step launches per group, alternating over the group
    // streams, so that one group's tail overlaps the other's next launch (a
    // single launch per group times the tails the sequence hides, and ranks
    // C2's best shapes within noise of each other; three launches rank them
    // as full 10,000-step runs do).  Once a step holds more than ~1e11 pair
    // evaluations (a few 100 ms) one 1-step launch per group, no warm-up.
    const double pairs = (double)j.n_rep * j.natom * t.npoint;
    const bool large = pairs > 1e11;
    const int trial = large ? 1 : 2;
    const int trial_chunks = large ? 1 : 3;
    float best_ms = 1e30f;
    Shape best = cands[0];
    std::vector<JobRes *> one{&j};
    const std::vector<int> s1_of{s1};
    for (const Shape &c : cands) {
        const std::vector<Shape> shapes{c};
        std::vector<Group> gr = make_groups(one, shapes, gd_ctx::kGroups);
        GroupsIdle idle{ctx};
        for (Group &g : gr) {
            group_tasks(g, one, shapes, s1_of, trial, g.tasks);
            GD_HIP(upload(g.tasks_dev, g.tasks.data(), sizeof(DevTask) * g.tasks.size(), s));
        }
        // warm the code object and caches with one launch per group, then time
        for (int rep = large ? 1 : 0; rep < 2; ++rep) {
            GD_HIP(hipEventRecord(e0.e, s));
            if (int rc = fork_groups(ctx, s, gr.size(), ef.e)) return rc;
            for (int k = 0; k < (rep ? trial_chunks : 1); ++k)
                for (size_t g = 0; g < gr.size(); ++g)
                    GD_HIP(launch_mc(t.dt, t.sc, j.self.as<DevJob>(), gr[g].tasks_dev.as<DevTask>(),
                                     (int)gr[g].tasks.size(), c.waves, c.nrep, compensated, false,
                                     ctx->group_stream[g], nullptr));
            if (int rc = join_groups(ctx, s, gr.size(), ej)) return rc;
            GD_HIP(hipEventRecord(e1.e, s));
            GD_HIP(hipMemcpyAsync(j.rep.p, backup.p, j.rep.bytes, hipMemcpyDeviceToDevice, s));
            GD_HIP(hipMemcpyAsync(j.mclog.p, backup_log.p, j.mclog.bytes, hipMemcpyDeviceToDevice, s));
            GD_HIP(hipEventSynchronize(e1.e));
        }
        float ms = 0.f;
        GD_HIP(hipEventElapsedTime(&ms, e0.e, e1.e));
        if (ms < best_ms) {
            best_ms = ms;
            best = c;
        }
    }
    GD_HIP(hipStreamSynchronize(s));
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
        if (std::find(js.begin(), js.end(), &it->second) != js.end())
            return fail(GD_ERR_ARG, "a job is listed twice"); // its replicas would be advanced twice at once
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
        if (W > 16 || (opts->point_slices > 0 && W % opts->point_slices))
            return fail(GD_ERR_ARG, "waves_per_group must be at most 16 and a multiple of point_slices");
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
        if (W > 16) return fail(GD_ERR_ARG, "waves_per_group must be at most 16");
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
    if (int rc = ensure_group_streams(ctx)) return rc;
    DevBuf jobs_dev;
    GD_HIP(upload(jobs_dev, table.data(), sizeof(DevJob) * n_jobs, s));
    std::vector<int> s1_of(n_jobs), round_of(n_jobs);
    for (int i = 0; i < n_jobs; ++i) {
        s1_of[i] = js[i]->steps_done;
        round_of[i] = js[i]->exchange_rounds;
        if (int rc = check_stream_range(*js[i], (long)s1_of[i] + opts->steps)) return rc;
    }
    std::vector<Group> gr = make_groups(js, shapes, gd_ctx::kGroups);
    GroupsIdle idle{ctx};
    const int zero2[2] = {0, 0};
    for (Group &g : gr) {
        group_tasks(g, js, shapes, s1_of, chunk, g.tasks);
        GD_HIP(upload(g.tasks_dev, g.tasks.data(), sizeof(DevTask) * g.tasks.size(), s));
        GD_HIP(g.bases.alloc(2 * sizeof(int)));
        GD_HIP(hipMemcpyAsync(g.bases.p, zero2, sizeof(zero2), hipMemcpyHostToDevice, s));
    }

    // one chunk of a group: Monte Carlo launch, exchange round over the
    // group's ladders, step advance (no memset, so it can be captured)
    auto enqueue_period = [&](Group &g, hipStream_t gs, int s2) -> int {
        GD_HIP(launch_mc(t.dt, t.sc, jobs_dev.as<DevJob>(), g.tasks_dev.as<DevTask>(), (int)g.tasks.size(), W, max_rep,
                         comp, records, gs, g.step_base()));
        if (opts->exchange_interval > 0)
            for (const Group::Range &r : g.ranges) {
                const JobRes &j = *js[r.job];
                GD_HIP(launch_exchange(jobs_dev.as<DevJob>() + r.job, r.p0 * j.n_lig, (r.p1 - r.p0) * j.n_lig, j.n_lig,
                                       j.n_tmp, round_of[r.job], nullptr, gs, g.round_base()));
            }
        hipLaunchKernelGGL(advance_kernel, dim3(1), dim3(64), 0, gs, g.step_base(), s2, g.round_base());
        GD_HIP(hipGetLastError());
        return GD_OK;
    };

    const int n_full = opts->steps / chunk;
    const int rest = opts->steps - n_full * chunk;
    if (n_full > 0 && opts->use_graph)
        for (size_t k = 0; k < gr.size(); ++k) {
            hipStream_t gs = ctx->group_stream[k];
            GD_HIP(hipStreamBeginCapture(gs, hipStreamCaptureModeThreadLocal));
            const int rc = enqueue_period(gr[k], gs, chunk);
            const hipError_t ce = hipStreamEndCapture(gs, &gr[k].graph);
            if (rc) return rc;
            GD_HIP(ce);
            GD_HIP(hipGraphInstantiate(&gr[k].exec, gr[k].graph, nullptr, nullptr, 0));
        }
    Event e0, e1, ef, ej[gd_ctx::kGroups];
    for (Event *e : {&e0, &e1, &ef}) GD_HIP(hipEventCreate(&e->e));
    for (Event &e : ej) GD_HIP(hipEventCreateWithFlags(&e.e, hipEventDisableTiming));
    GD_HIP(hipEventRecord(e0.e, s));
    if (int rc = fork_groups(ctx, s, gr.size(), ef.e)) return rc;
    for (int c = 0; c < n_full; ++c)
        for (size_t k = 0; k < gr.size(); ++k) {
            if (opts->use_graph)
                GD_HIP(hipGraphLaunch(gr[k].exec, ctx->group_stream[k]));
            else if (int rc = enqueue_period(gr[k], ctx->group_stream[k], chunk))
                return rc;
        }
    if (rest > 0)
        for (size_t k = 0; k < gr.size(); ++k) {
            Group &g = gr[k];
            hipStream_t gs = ctx->group_stream[k];
            group_tasks(g, js, shapes, s1_of, rest, g.tail);
            // same task count as the chunk's table, so no reallocation while
            // the group's earlier launches still read it
            GD_HIP(upload(g.tasks_dev, g.tail.data(), sizeof(DevTask) * g.tail.size(), gs));
            GD_HIP(launch_mc(t.dt, t.sc, jobs_dev.as<DevJob>(), g.tasks_dev.as<DevTask>(), (int)g.tail.size(), W,
                             max_rep, comp, records, gs, g.step_base()));
        }
    if (int rc = join_groups(ctx, s, gr.size(), ej)) return rc;
    GD_HIP(hipEventRecord(e1.e, s));
    GD_HIP(hipEventSynchronize(e1.e));
    float ms = 0.f;
    GD_HIP(hipEventElapsedTime(&ms, e0.e, e1.e));
    if (elapsed_ms) *elapsed_ms = ms;
    for (int i = 0; i < n_jobs; ++i) {
        js[i]->steps_done += opts->steps;
        if (opts->exchange_interval > 0) js[i]->exchange_rounds += n_full;
    }
    return GD_OK;
}
