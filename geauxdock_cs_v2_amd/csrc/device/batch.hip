// Batched form: every replica of several jobs against one resident target in
// one persistent launch per chunk of steps.  Replicas are packed R per
// workgroup so their ligand atoms fill the 64-lane waves, the protein points
// are split across `slices` waves, and workgroup shapes and energy layouts
// are picked by a timing autotuner (per job, or per ligand-size bucket when
// several jobs share a launch).  The launch's grid is the number of
// workgroups resident at once; they walk a cost-ordered work queue of
// (task, step slice) items (see mc_kernel), so the last round of a chunk
// spreads over the whole GPU.  A chunk (queue reset, Monte Carlo launch,
// exchange round over every job's ladders, step advance) is captured once in
// a hipGraph and replayed.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
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
// single job: (N_L, N_P, n_rep, n_prt) -> shape
std::map<std::tuple<int, int, int, int>, Shape> g_tune_cache;
// several jobs: (N_L bucket, N_P, waves) -> slot waves of the bucket's jobs,
// 0 where each job keeps the packing heuristic's choice for itself
std::map<std::tuple<int, int, int>, int> g_bucket_cache;
// several jobs: (N_P, bucket signature) -> the launch's waves per workgroup,
// energy layout (for jobs with more than one point slice) and priority policy
struct BatchChoice {
    int waves, energy, boost;
};
std::map<std::pair<int, std::vector<std::pair<int, int>>>, BatchChoice> g_waves_cache;

struct Event {
    hipEvent_t e = nullptr;
    Event() = default;
    Event(const Event &) = delete;
    ~Event()
    {
        if (e) (void)hipEventDestroy(e);
    }
};

// graph of one chunk, released with the scope
struct Graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~Graph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

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
    s.layout = 0;
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
        DevTask t{};
        t.job = job_slot;
        t.rep0 = p * block + r;
        t.nrep = std::min(sh.nrep, block - r);
        t.prt = p;
        t.s1 = s1;
        t.s2max = s2max;
        t.slices = sh.slices;
        t.slot_waves = sh.slot_waves;
        t.layout = sh.layout;
        out.push_back(t);
    }
}

void build_tasks(const JobRes &j, int job_slot, const Shape &sh, int s1, int s2max, std::vector<DevTask> &out)
{
    for (int p = 0; p < j.n_prt; ++p) build_unit_tasks(j, job_slot, sh, p, s1, s2max, out);
}

int resident_slots(gd_ctx *ctx, int waves, int max_rep)
{
    return std::max(1, ctx->cu_count) * mc_blocks_per_cu(waves, max_rep);
}

// Step slices per task.  With ntask tasks of equal cost on `slots` resident
// workgroups a chunk takes ceil(ntask K / slots) / K task durations for K
// slices per task (ntask / slots at best); each slice adds a hand-over (the
// replica states through L2 and an acquire), counted as 0.5% of a slice.
int choose_slices(int ntask, int slots, int steps)
{
    if (ntask <= slots || steps <= 1) return 1;
    int best = 1;
    double best_cost = 1e30;
    for (int k = 1; k <= std::min(8, steps); ++k) {
        const double rounds = std::ceil((double)ntask * k / slots) / k;
        const double cost = rounds * (1.0 + 0.005 * (k - 1));
        if (cost < best_cost - 1e-12) {
            best_cost = cost;
            best = k;
        }
    }
    return best;
}

namespace {

// all tasks of the batch for s2max steps from each job's own step count, in
// decreasing cost order (pair evaluations per step) so heavy workgroups
// start first
void batch_tasks(const std::vector<JobRes *> &js, const std::vector<Shape> &shapes, const std::vector<int> &s1_of,
                 int s2max, std::vector<DevTask> &out)
{
    out.clear();
    for (size_t i = 0; i < js.size(); ++i) build_tasks(*js[i], (int)i, shapes[i], s1_of[i], s2max, out);
    std::stable_sort(out.begin(), out.end(), [&](const DevTask &x, const DevTask &y) {
        return (long)x.nrep * js[x.job]->natom > (long)y.nrep * js[y.job]->natom;
    });
}

int max_rep_of(const std::vector<Shape> &shapes)
{
    int m = 1;
    for (const Shape &s : shapes) m = std::max(m, s.nrep);
    return m;
}

// the launch runs the progress-policy kernel when every job's shape asks for it
bool progress_of(const std::vector<Shape> &shapes)
{
    for (const Shape &sh : shapes)
        if (!(sh.layout & LAYOUT_BOOST_PROGRESS)) return false;
    return !shapes.empty();
}

// the context's work queue, with room for ntask tasks
int ensure_queue(gd_ctx *ctx, int ntask, DevQueue &q)
{
    const size_t need = sizeof(int) * ((size_t)ntask + 1);
    if (ctx->queue.bytes < need) GD_HIP(ctx->queue.alloc(need));
    q.counter = ctx->queue.as<int>();
    q.done = ctx->queue.as<int>() + 1;
    q.ntask = ntask;
    q.nslice = 1;
    return GD_OK;
}

// One persistent launch of `tasks_dev` (ntask tasks, shape waves/max_rep,
// slices chosen for s2max steps) over the work queue q (room for ntask
// tasks): queue reset, Monte Carlo kernel.
int enqueue_mc_on(gd_ctx *ctx, const TargetRes &t, const DevJob *jobs_dev, const DevTask *tasks_dev, int ntask,
                  int s2max, int waves, int max_rep, bool comp, bool records, bool progress, hipStream_t s,
                  const int *step_base, DevQueue q)
{
    q.ntask = ntask;
    const int slots = resident_slots(ctx, waves, max_rep);
    q.nslice = choose_slices(ntask, slots, s2max);
    ctx->last_plan[0] = ntask;
    ctx->last_plan[1] = std::min(slots, ntask * q.nslice);
    ctx->last_plan[2] = q.nslice;
    GD_HIP(launch_queue_reset(q, s));
    GD_HIP(launch_mc_queue(t.params.as<DevParams>(), jobs_dev, tasks_dev, q, slots, waves, max_rep, comp, records,
                           progress, s, step_base));
    return GD_OK;
}

// the same over the context's own work queue, grown to ntask tasks
int enqueue_mc(gd_ctx *ctx, const TargetRes &t, const DevJob *jobs_dev, const DevTask *tasks_dev, int ntask,
               int s2max, int waves, int max_rep, bool comp, bool records, bool progress, hipStream_t s,
               const int *step_base)
{
    DevQueue q;
    if (int rc = ensure_queue(ctx, ntask, q)) return rc;
    return enqueue_mc_on(ctx, t, jobs_dev, tasks_dev, ntask, s2max, waves, max_rep, comp, records, progress, s,
                         step_base, q);
}

// GD_TUNE_LOG set in the environment: each autotuner trial is printed to
// stderr (the first job's shape, the layout bits, steps and time)
static void tune_log(const char *form, const Shape &s, int steps, float ms)
{
    static const bool on = std::getenv("GD_TUNE_LOG") != nullptr;
    if (on)
        std::fprintf(stderr, "gd tune %s: waves %d slot waves %d slices %d replicas %d layout %d, %d steps %.3f ms\n",
                     form, s.waves, s.slot_waves, s.slices, s.nrep, s.layout, steps, ms);
}

// Time `launches` consecutive persistent launches of `steps` steps of the
// batch with the given shapes, after one untimed launch (code object,
// caches); the replica states and McLog counters are restored afterwards.
int time_shapes(gd_ctx *ctx, const TargetRes &t, const std::vector<JobRes *> &js, const DevJob *jobs_dev,
                const std::vector<Shape> &shapes, int steps, int launches, bool warm, bool comp, float &ms)
{
    hipStream_t s = ctx->stream;
    std::vector<int> s1_of(js.size());
    for (size_t i = 0; i < js.size(); ++i) s1_of[i] = js[i]->steps_done;
    std::vector<DevTask> tasks;
    batch_tasks(js, shapes, s1_of, steps, tasks);
    const int W = shapes[0].waves;
    const int max_rep = max_rep_of(shapes);
    const bool progress = progress_of(shapes);
    DevBuf tasks_dev;
    GD_HIP(upload(tasks_dev, tasks.data(), sizeof(DevTask) * tasks.size(), s));
    Event e0, e1;
    GD_HIP(hipEventCreate(&e0.e));
    GD_HIP(hipEventCreate(&e1.e));
    if (warm)
        if (int rc = enqueue_mc(ctx, t, jobs_dev, tasks_dev.as<DevTask>(), (int)tasks.size(), steps, W, max_rep,
                                comp, false, progress, s, nullptr))
            return rc;
    GD_HIP(hipEventRecord(e0.e, s));
    for (int k = 0; k < launches; ++k)
        if (int rc = enqueue_mc(ctx, t, jobs_dev, tasks_dev.as<DevTask>(), (int)tasks.size(), steps, W, max_rep,
                                comp, false, progress, s, nullptr))
            return rc;
    GD_HIP(hipEventRecord(e1.e, s));
    GD_HIP(hipEventSynchronize(e1.e));
    GD_HIP(hipEventElapsedTime(&ms, e0.e, e1.e));
    return GD_OK;
}

// the replica states and McLog counters of the batch's jobs, saved on the
// device and put back when the trials are over
struct StateBackup {
    std::vector<DevBuf> rep, log;
    int save(gd_ctx *ctx, const std::vector<JobRes *> &js)
    {
        rep.resize(js.size());
        log.resize(js.size());
        for (size_t i = 0; i < js.size(); ++i) {
            GD_HIP(rep[i].alloc(js[i]->rep.bytes));
            GD_HIP(log[i].alloc(js[i]->mclog.bytes));
            GD_HIP(hipMemcpyAsync(rep[i].p, js[i]->rep.p, js[i]->rep.bytes, hipMemcpyDeviceToDevice, ctx->stream));
            GD_HIP(hipMemcpyAsync(log[i].p, js[i]->mclog.p, js[i]->mclog.bytes, hipMemcpyDeviceToDevice, ctx->stream));
        }
        return GD_OK;
    }
    int restore(gd_ctx *ctx, const std::vector<JobRes *> &js)
    {
        for (size_t i = 0; i < js.size(); ++i) {
            GD_HIP(hipMemcpyAsync(js[i]->rep.p, rep[i].p, js[i]->rep.bytes, hipMemcpyDeviceToDevice, ctx->stream));
            GD_HIP(hipMemcpyAsync(js[i]->mclog.p, log[i].p, js[i]->mclog.bytes, hipMemcpyDeviceToDevice, ctx->stream));
        }
        GD_HIP(hipStreamSynchronize(ctx->stream));
        return GD_OK;
    }
};

// trial length: a few steps per launch, one step once a step holds more than
// ~1e11 pair evaluations (a few 100 ms); never more than the run's own
// `cap` steps, which the explicit stream words (if any) were checked to cover
int trial_steps(const std::vector<JobRes *> &js, const TargetRes &t, int cap)
{
    double pairs = 0.0;
    bool explicit_words = false;
    for (const JobRes *j : js) {
        pairs += (double)j->n_rep * j->natom * t.npoint;
        explicit_words |= j->kstream_steps > 0;
    }
    // Small jobs run 20 steps per trial launch, so the trial splits its tasks
    // into the step slices a 20- or 100-step chunk would (4-step trials gave
    // c2 one step per work item, where per-item costs decide the ranking of
    // layouts and shapes); about 10 ms per trial launch at the kernel's
    // rate.  Jobs with explicit stream words stay within the call's steps.
    const int want = pairs > 1e11 ? 1 : pairs > 1e10 ? 4 : 20;
    return std::max(1, explicit_words ? std::min(cap, want) : want);
}

// ligand-size bucket of the multi-job autotuner: jobs whose atom counts fall
// in one bucket share a packing
int natom_bucket(int natom)
{
    static const int edges[] = {16, 24, 32, 40, 48, 64, 80, 96, 128, 192, 256, 384, 512, 768, 1024};
    for (int e : edges)
        if (natom <= e) return e;
    return 1 << 30;
}

} // namespace

// Autotuner (one job): time every candidate workgroup shape on the job's own
// data (state saved and restored around the trials) as persistent launches,
// then the KDE/MCS layouts on the fastest shape; the result is cached per
// (N_L, N_P, n_rep, n_prt).
int autotune_shape(gd_ctx *ctx, const TargetRes &t, JobRes &j, bool compensated, int max_steps, Shape &out)
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
            for (const Shape &c : cands) dup |= c.waves == s.waves && c.nrep == s.nrep && c.slices == s.slices;
            if (!dup) cands.push_back(s);
        }
    if (cands.empty()) return fail(GD_ERR_ARG, "ligand too large for a 16-wave workgroup");
    std::vector<JobRes *> one{&j};
    const DevJob *jobs_dev = j.self.as<DevJob>();
    const int steps = trial_steps(one, t, max_steps);
    const bool large = (double)j.n_rep * j.natom * t.npoint > 1e11;
    StateBackup bk;
    if (int rc = bk.save(ctx, one)) return rc;
    float best_ms = 1e30f;
    Shape best = cands[0];
    for (const Shape &c : cands) {
        float ms = 0.f;
        if (int rc = time_shapes(ctx, t, one, jobs_dev, {c}, steps, large ? 1 : 2, !large, compensated, ms)) return rc;
        if (int rc = bk.restore(ctx, one)) return rc;
        tune_log("job", c, steps, ms);
        if (ms < best_ms) {
            best_ms = ms;
            best = c;
        }
    }
    // energy layouts on the chosen shape (a single-slice shape has one layout)
    if (best.slices > 1)
        for (int layout : {(int)LAYOUT_KDE_SLICE0, (int)LAYOUT_MCS_SLICE0, LAYOUT_KDE_SLICE0 | LAYOUT_MCS_SLICE0}) {
            Shape c = best;
            c.layout = layout;
            float ms = 0.f;
            if (int rc = time_shapes(ctx, t, one, jobs_dev, {c}, steps, large ? 1 : 2, !large, compensated, ms))
                return rc;
            if (int rc = bk.restore(ctx, one)) return rc;
            tune_log("job", c, steps, ms);
            if (ms < best_ms * 0.995f) { // a layout change must win by more than the noise
                best_ms = ms;
                best = c;
            }
        }
    // issue-priority policies of the pair loops on the chosen shape and
    // layout, against the shape timed again beside them (the 1-step trials of
    // a large job vary by about 0.5 % between the shape loop and this stage,
    // and the policies differ by about 1 %): two launches each, four of a
    // large job's one step
    {
        const Shape base = best;
        const int reps = large ? 4 : 2;
        float ref = 0.f;
        if (int rc = time_shapes(ctx, t, one, jobs_dev, {base}, steps, reps, !large, compensated, ref)) return rc;
        if (int rc = bk.restore(ctx, one)) return rc;
        tune_log("job", base, steps, ref);
        float bar = ref * 0.995f;
        for (int policy : {(int)LAYOUT_BOOST_ANY, (int)LAYOUT_BOOST_NONE, (int)LAYOUT_BOOST_PROGRESS}) {
            Shape c = base;
            c.layout = (base.layout & LAYOUT_ENERGY_BITS) | policy;
            float ms = 0.f;
            if (int rc = time_shapes(ctx, t, one, jobs_dev, {c}, steps, reps, !large, compensated, ms)) return rc;
            if (int rc = bk.restore(ctx, one)) return rc;
            tune_log("job", c, steps, ms);
            if (ms < bar) {
                bar = ms;
                best = c;
            }
        }
    }
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_cache[key] = best;
    }
    out = best;
    return GD_OK;
}

// Autotuner (several jobs in one launch): the wave count W is shared by the
// launch, the packing per ligand-size bucket.  First W from {8, 12, 16} with
// the packing heuristic (each job's own choice), then, bucket by bucket in
// decreasing share of the work (buckets above 5% of it), every slot-wave
// count that divides W for all of the bucket's jobs, then the energy layouts
// and the priority policies for all jobs; every trial times the whole batch.
// Results are cached per (N_P, bucket signature) for W, layout and policy and
// per (bucket, N_P, W) for the packings (a bucket no trial changed keeps the
// heuristic's per-job choice), so a later batch runs the shapes the trials
// chose.
int autotune_batch(gd_ctx *ctx, const TargetRes &t, const std::vector<JobRes *> &js, const DevJob *jobs_dev,
                   bool compensated, int max_steps, std::vector<Shape> &shapes)
{
    const size_t n = js.size();
    std::map<int, double> work; // bucket -> pair evaluations per step
    std::map<int, int> count;
    std::vector<int> bucket(n);
    double total = 0.0;
    for (size_t i = 0; i < n; ++i) {
        bucket[i] = natom_bucket(js[i]->natom);
        const double w = (double)js[i]->n_rep * js[i]->natom * t.npoint;
        work[bucket[i]] += w;
        count[bucket[i]] += 1;
        total += w;
    }
    std::vector<std::pair<int, int>> sig(count.begin(), count.end());
    int max_natom = 0;
    for (const JobRes *j : js) max_natom = std::max(max_natom, j->natom);
    const auto heuristic_all = [&](int W, std::vector<Shape> &sh) {
        const int share = std::max(1, resident_slots(ctx, W, 8) / (int)n);
        sh.resize(n);
        for (size_t i = 0; i < n; ++i)
            sh[i] = heuristic_shape(js[i]->natom, W, js[i]->n_tmp * js[i]->n_lig, js[i]->n_prt, share);
    };
    // everything cached: no trials
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto wit = g_waves_cache.find({t.npoint, sig});
        if (wit != g_waves_cache.end()) {
            const BatchChoice ch = wit->second;
            const int W = ch.waves;
            std::vector<Shape> sh;
            heuristic_all(W, sh);
            bool all = true;
            for (size_t i = 0; i < n && all; ++i) {
                auto b = g_bucket_cache.find(std::make_tuple(bucket[i], t.npoint, W));
                if (b == g_bucket_cache.end()) {
                    all = false;
                } else {
                    if (b->second > 0) sh[i] = packed_shape(js[i]->natom, W, b->second, js[i]->n_tmp * js[i]->n_lig);
                    sh[i].layout = (sh[i].slices > 1 ? ch.energy : 0) | ch.boost;
                }
            }
            if (all) {
                shapes = sh;
                return GD_OK;
            }
        }
    }
    const int steps = trial_steps(js, t, max_steps);
    StateBackup bk;
    if (int rc = bk.save(ctx, js)) return rc;
    auto trial = [&](const std::vector<Shape> &sh, float &ms) -> int {
        float a = 0.f;
        if (int rc = time_shapes(ctx, t, js, jobs_dev, sh, steps, 1, true, compensated, a)) return rc;
        if (int rc = bk.restore(ctx, js)) return rc;
        tune_log("batch", sh[0], steps, a);
        ms = a;
        return GD_OK;
    };
    float best_ms = 1e30f;
    std::vector<Shape> best;
    // a larger W must win by more than the noise (0.5 %): on near-ties the
    // smaller workgroup, with fewer waves per SIMD waiting at the step's
    // barriers, is kept (the 1,024-ligand screen: 8 and 12 waves within 0.3 %
    // in 4-step trials and 95.5 against 95.3 TFLOP/s over 100 steps)
    for (int W : {8, 12, 16}) {
        if (max_natom > 64 * W) continue;
        std::vector<Shape> sh;
        heuristic_all(W, sh);
        float ms = 0.f;
        if (int rc = trial(sh, ms)) return rc;
        if (ms < (best.empty() ? 1e30f : best_ms * 0.995f)) {
            best_ms = ms;
            best = sh;
        }
    }
    if (best.empty()) return fail(GD_ERR_ARG, "ligand too large for a 16-wave workgroup");
    const int W = best[0].waves;
    std::map<int, int> bucket_sw; // buckets whose slot-wave count a trial chose
    int energy = 0, boost = 0;
    std::vector<std::pair<double, int>> order;
    for (const auto &kv : work) order.emplace_back(-kv.second, kv.first);
    std::sort(order.begin(), order.end());
    for (const auto &o : order) {
        const int b = o.second;
        if (-o.first < 0.05 * total) continue;
        int bmax = 0;
        for (size_t i = 0; i < n; ++i)
            if (bucket[i] == b) bmax = std::max(bmax, js[i]->natom);
        for (int SW = 1; SW <= W; ++SW) {
            if (W % SW || 64 * SW < bmax) continue;
            std::vector<Shape> sh = best;
            bool changed = false;
            for (size_t i = 0; i < n; ++i)
                if (bucket[i] == b) {
                    changed |= sh[i].slot_waves != SW;
                    sh[i] = packed_shape(js[i]->natom, W, SW, js[i]->n_tmp * js[i]->n_lig);
                    sh[i].layout = best[i].layout;
                }
            if (!changed) continue;
            float ms = 0.f;
            if (int rc = trial(sh, ms)) return rc;
            if (ms < best_ms * 0.995f) {
                best_ms = ms;
                best = sh;
                bucket_sw[b] = SW;
            }
        }
    }
    for (int layout : {(int)LAYOUT_KDE_SLICE0, (int)LAYOUT_MCS_SLICE0, LAYOUT_KDE_SLICE0 | LAYOUT_MCS_SLICE0}) {
        std::vector<Shape> sh = best;
        for (Shape &s : sh) s.layout = s.slices > 1 ? layout : 0;
        float ms = 0.f;
        if (int rc = trial(sh, ms)) return rc;
        if (ms < best_ms * 0.995f) {
            best_ms = ms;
            best = sh;
            energy = layout;
        }
    }
    // issue-priority policies of the pair loops, for all jobs at once,
    // against the shapes timed again beside them
    {
        const std::vector<Shape> base = best;
        float ref = 0.f;
        if (int rc = trial(base, ref)) return rc;
        float bar = ref * 0.995f;
        for (int policy : {(int)LAYOUT_BOOST_ANY, (int)LAYOUT_BOOST_NONE, (int)LAYOUT_BOOST_PROGRESS}) {
            std::vector<Shape> sh = base;
            for (Shape &s : sh) s.layout = (s.layout & LAYOUT_ENERGY_BITS) | policy;
            float ms = 0.f;
            if (int rc = trial(sh, ms)) return rc;
            if (ms < bar) {
                bar = ms;
                best = sh;
                boost = policy;
            }
        }
    }
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_waves_cache[{t.npoint, sig}] = BatchChoice{W, energy, boost};
        for (size_t i = 0; i < n; ++i) {
            const auto it = bucket_sw.find(bucket[i]);
            g_bucket_cache[std::make_tuple(bucket[i], t.npoint, W)] = it == bucket_sw.end() ? 0 : it->second;
        }
    }
    shapes = best;
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

extern "C" int gd_batch_plan(gd_ctx *ctx, int *plan)
{
    if (!ctx || !plan) return fail(GD_ERR_ARG, "null argument");
    for (int k = 0; k < 3; ++k) plan[k] = ctx->last_plan[k];
    return GD_OK;
}

namespace {

// the batch's jobs (listed once each, one target) and its step checks
int batch_jobs(gd_ctx *ctx, const int *job_ids, int n_jobs, const gd_batch_opts *opts, std::vector<JobRes *> &js,
               int &target)
{
    if (!ctx || !job_ids || n_jobs <= 0 || !opts) return fail(GD_ERR_ARG, "null or empty argument");
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(GD_ERR_DEVICE, "hipSetDevice failed");
    if (opts->steps <= 0 || opts->exchange_interval < 0) return fail(GD_ERR_ARG, "bad step counts");
    target = -1;
    for (int i = 0; i < n_jobs; ++i) {
        auto it = ctx->jobs.find(job_ids[i]);
        if (it == ctx->jobs.end()) return fail(GD_ERR_STATE, "unknown job");
        if (std::find(js.begin(), js.end(), &it->second) != js.end())
            return fail(GD_ERR_ARG, "a job is listed twice"); // its replicas would be advanced twice at once
        if (target >= 0 && it->second.target != target) return fail(GD_ERR_ARG, "batched jobs must share a target");
        target = it->second.target;
        js.push_back(&it->second);
    }
    for (JobRes *j : js)
        if (int rc = check_stream_range(*j, (long)j->steps_done + opts->steps)) return rc;
    return GD_OK;
}

int batch_chunk(const gd_batch_opts *opts)
{
    return opts->exchange_interval > 0 ? opts->exchange_interval : std::min(opts->steps, GD_STEPS_PER_DUMP);
}

// Workgroup shapes: explicit in opts, the packing heuristic for a given wave
// count, or (no wave count) autotuned when `tune`, else the autotuner's
// cached choice where there is one and the heuristic at 12 waves otherwise.
int choose_shapes(gd_ctx *ctx, const TargetRes &t, const std::vector<JobRes *> &js, const gd_batch_opts *opts,
                  const DevJob *jobs_dev, bool tune, std::vector<Shape> &shapes)
{
    const int n_jobs = (int)js.size();
    const bool comp = opts->compensated != 0;
    int W = opts->waves_per_group;
    shapes.assign(n_jobs, Shape{});
    auto heuristic = [&](int w) -> int {
        if (w > 16) return fail(GD_ERR_ARG, "waves_per_group must be at most 16");
        const int share = std::max(1, resident_slots(ctx, w, 8) / n_jobs);
        for (int i = 0; i < n_jobs; ++i) {
            if (js[i]->natom > 64 * w) return fail(GD_ERR_ARG, "ligand too large for the workgroup");
            shapes[i] = heuristic_shape(js[i]->natom, w, js[i]->n_tmp * js[i]->n_lig, js[i]->n_prt, share);
        }
        return GD_OK;
    };
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
        return GD_OK;
    }
    if (W > 0) return heuristic(W);
    if (tune) {
        if (n_jobs == 1) return autotune_shape(ctx, t, *js[0], comp, opts->steps, shapes[0]);
        return autotune_batch(ctx, t, js, jobs_dev, comp, opts->steps, shapes);
    }
    if (int rc = heuristic(12)) return rc;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (n_jobs == 1) {
        auto it = g_tune_cache.find(std::make_tuple(js[0]->natom, t.npoint, js[0]->n_rep, js[0]->n_prt));
        if (it != g_tune_cache.end()) shapes[0] = it->second;
    }
    return GD_OK;
}

// GD_BOOST_POLICY in the environment (half, any, none or progress) sets the
// issue-priority policy of every job's shape, whatever chose the shape (the
// autotuner's cache keeps its own choice); unset or another value: no change.
int boost_override()
{
    const char *v = std::getenv("GD_BOOST_POLICY");
    if (!v) return -1;
    const std::string p(v);
    return p == "half"       ? 0
           : p == "any"      ? (int)LAYOUT_BOOST_ANY
           : p == "none"     ? (int)LAYOUT_BOOST_NONE
           : p == "progress" ? (int)LAYOUT_BOOST_PROGRESS
                             : -1;
}

int batch_shapes(gd_ctx *ctx, const TargetRes &t, const std::vector<JobRes *> &js, const gd_batch_opts *opts,
                 const DevJob *jobs_dev, bool tune, std::vector<Shape> &shapes)
{
    if (int rc = choose_shapes(ctx, t, js, opts, jobs_dev, tune, shapes)) return rc;
    const int boost = boost_override();
    if (boost >= 0)
        for (Shape &sh : shapes) sh.layout = (sh.layout & LAYOUT_ENERGY_BITS) | boost;
    return GD_OK;
}

// Device memory a batched run uses beyond the jobs' own tables.
struct BatchMem {
    DevJob *jobs = nullptr;              // [n_jobs] job descriptors
    DevTask *tasks = nullptr;            // [ntask] a full chunk's tasks
    DevTask *tail = nullptr;             // [ntask] the last, shorter chunk's tasks
    DevQueue queue{nullptr, nullptr, 0, 1}; // counter and done[ntask]
    int4 *ladders = nullptr;             // [nlad] exchange ladders
    int *bases = nullptr;                // [2] step and exchange-round counters
    bool progress = false;               // the progress-policy kernel (progress_of)
};

// The chunk sequence of a batched run, once its tables are in `mem`: per
// full chunk a queue reset, the persistent Monte Carlo launch, an exchange
// round over every ladder and the step advance (captured in a hipGraph when
// use_graph), then the remaining steps.
int batch_launches(gd_ctx *ctx, const TargetRes &t, const gd_batch_opts *opts, int chunk, int W, int max_rep,
                   int ntask, int nlad, const BatchMem &mem, bool use_graph, double *elapsed_ms)
{
    hipStream_t s = ctx->stream;
    const bool comp = opts->compensated != 0, records = opts->keep_records != 0;
    auto enqueue_period = [&]() -> int {
        if (int rc = enqueue_mc_on(ctx, t, mem.jobs, mem.tasks, ntask, chunk, W, max_rep, comp, records,
                                   mem.progress, s, mem.bases, mem.queue))
            return rc;
        if (nlad > 0) GD_HIP(launch_exchange_list(mem.jobs, mem.ladders, nlad, s, mem.bases + 1));
        hipLaunchKernelGGL(advance_kernel, dim3(1), dim3(64), 0, s, mem.bases, chunk, mem.bases + 1);
        GD_HIP(hipGetLastError());
        return GD_OK;
    };
    const int n_full = opts->steps / chunk;
    const int rest = opts->steps - n_full * chunk;
    Graph g;
    if (n_full > 0 && use_graph) {
        GD_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_period();
        const hipError_t ce = hipStreamEndCapture(s, &g.graph);
        if (rc) return rc;
        GD_HIP(ce);
        GD_HIP(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
    }
    const auto h0 = std::chrono::steady_clock::now();
    Event e0, e1;
    if (use_graph) {
        GD_HIP(hipEventCreate(&e0.e));
        GD_HIP(hipEventCreate(&e1.e));
        GD_HIP(hipEventRecord(e0.e, s));
    }
    for (int c = 0; c < n_full; ++c) {
        if (use_graph)
            GD_HIP(hipGraphLaunch(g.exec, s));
        else if (int rc = enqueue_period())
            return rc;
    }
    if (rest > 0)
        if (int rc = enqueue_mc_on(ctx, t, mem.jobs, mem.tail, ntask, rest, W, max_rep, comp, records,
                                   mem.progress, s, mem.bases, mem.queue))
            return rc;
    float ms = 0.f;
    if (use_graph) {
        GD_HIP(hipEventRecord(e1.e, s));
        GD_HIP(hipEventSynchronize(e1.e));
        GD_HIP(hipEventElapsedTime(&ms, e0.e, e1.e));
    } else {
        GD_HIP(hipStreamSynchronize(s));
        ms = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
    }
    if (elapsed_ms) *elapsed_ms = ms;
    return GD_OK;
}

// host tables of a batched run: job descriptors, a full chunk's tasks, the
// last chunk's tasks, exchange ladders
void batch_tables(const std::vector<JobRes *> &js, const std::vector<Shape> &shapes, const gd_batch_opts *opts,
                  int chunk, std::vector<DevJob> &table, std::vector<DevTask> &tasks, std::vector<DevTask> &tail,
                  std::vector<int4> &lad)
{
    const int n_jobs = (int)js.size();
    std::vector<int> s1_of(n_jobs);
    table.resize(n_jobs);
    for (int i = 0; i < n_jobs; ++i) {
        s1_of[i] = js[i]->steps_done;
        table[i] = js[i]->hj;
    }
    batch_tasks(js, shapes, s1_of, chunk, tasks);
    const int rest = opts->steps % chunk;
    tail.clear();
    if (rest > 0) batch_tasks(js, shapes, s1_of, rest, tail);
    lad.clear();
    if (opts->exchange_interval > 0)
        for (int i = 0; i < n_jobs; ++i)
            if (js[i]->n_tmp > 1)
                for (int L = 0; L < js[i]->n_lig * js[i]->n_prt; ++L)
                    lad.push_back(make_int4(i, L, js[i]->exchange_rounds, 0));
}

void batch_done(const std::vector<JobRes *> &js, const std::vector<Shape> &shapes, const gd_batch_opts *opts,
                int chunk)
{
    for (size_t i = 0; i < js.size(); ++i) {
        js[i]->last_shape = shapes[i];
        js[i]->steps_done += opts->steps;
        if (opts->exchange_interval > 0) js[i]->exchange_rounds += opts->steps / chunk;
    }
}

} // namespace

extern "C" int gd_batch_run(gd_ctx *ctx, const int *job_ids, int n_jobs, const gd_batch_opts *opts,
                            double *elapsed_ms)
{
    std::vector<JobRes *> js;
    int target = -1;
    if (int rc = batch_jobs(ctx, job_ids, n_jobs, opts, js, target)) return rc;
    const TargetRes &t = ctx->targets.at(target);
    const int chunk = batch_chunk(opts);
    hipStream_t s = ctx->stream;
    if (opts->keep_records)
        for (JobRes *j : js)
            if (int rc = ensure_records(ctx, *j, chunk)) return rc;
    std::vector<DevJob> table;
    std::vector<DevTask> tasks, tail;
    std::vector<int4> lad;
    // the job table first: the multi-job autotuner launches from it
    DevBuf jobs_dev;
    {
        std::vector<DevJob> t0(n_jobs);
        for (int i = 0; i < n_jobs; ++i) t0[i] = js[i]->hj;
        GD_HIP(upload(jobs_dev, t0.data(), sizeof(DevJob) * n_jobs, s));
    }
    std::vector<Shape> shapes;
    if (int rc = batch_shapes(ctx, t, js, opts, jobs_dev.as<DevJob>(), true, shapes)) return rc;
    const int W = shapes[0].waves, max_rep = max_rep_of(shapes);
    batch_tables(js, shapes, opts, chunk, table, tasks, tail, lad);
    const int ntask = (int)tasks.size();
    DevBuf tasks_dev, tail_dev;
    GD_HIP(upload(tasks_dev, tasks.data(), sizeof(DevTask) * ntask, s));
    if (!tail.empty()) GD_HIP(upload(tail_dev, tail.data(), sizeof(DevTask) * tail.size(), s));
    if (!lad.empty()) GD_HIP(upload(ctx->ladders, lad.data(), sizeof(int4) * lad.size(), s));
    if (ctx->bases.bytes < 2 * sizeof(int)) GD_HIP(ctx->bases.alloc(2 * sizeof(int)));
    const int zero2[2] = {0, 0};
    GD_HIP(hipMemcpyAsync(ctx->bases.p, zero2, sizeof(zero2), hipMemcpyHostToDevice, s));
    BatchMem mem;
    if (int rc = ensure_queue(ctx, ntask, mem.queue)) return rc; // sized before any capture
    mem.jobs = jobs_dev.as<DevJob>();
    mem.tasks = tasks_dev.as<DevTask>();
    mem.tail = tail.empty() ? nullptr : tail_dev.as<DevTask>();
    mem.ladders = ctx->ladders.as<int4>();
    mem.bases = ctx->bases.as<int>();
    mem.progress = progress_of(shapes);
    if (int rc = batch_launches(ctx, t, opts, chunk, W, max_rep, ntask, (int)lad.size(), mem, opts->use_graph != 0,
                                elapsed_ms))
        return rc;
    batch_done(js, shapes, opts, chunk);
    return GD_OK;
}

// ---------------------------------------------------------------------------
// The batched run without device allocation.  Its tables (job descriptors,
// tasks, the work queue, exchange ladders, step and round counters) live in
// a file-scope device array of kScratchBytes; each GPU holds its own copy
// until the process ends, at the address hipGetSymbolAddress reports with
// that device current (the array's name in host code is only a stand-in).
// Calls that overlap on one GPU share the array and race.
// ---------------------------------------------------------------------------
namespace {

constexpr size_t kScratchBytes = (size_t)8 << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
std::map<int, void *> g_scratch_ptr; // device -> its copy of scratch

// the device address of `device`'s scratch, looked up once per device
int scratch_address(int device, void **out)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(device);
    if (it == g_scratch_ptr.end()) {
        void *p = nullptr;
        GD_HIP(hipSetDevice(device));
        GD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(GD_ERR_DEVICE, "hipGetSymbolAddress returned a null scratch address");
        it = g_scratch_ptr.emplace(device, p).first;
    }
    *out = it->second;
    return GD_OK;
}

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

} // namespace

extern "C" size_t gd_batch_noalloc_capacity(void) { return kScratchBytes; }

extern "C" int gd_batch_run_noalloc(gd_ctx *ctx, const int *job_ids, int n_jobs, const gd_batch_opts *opts,
                                    double *elapsed_ms)
{
    std::vector<JobRes *> js;
    int target = -1;
    if (int rc = batch_jobs(ctx, job_ids, n_jobs, opts, js, target)) return rc;
    const TargetRes &t = ctx->targets.at(target);
    const int chunk = batch_chunk(opts);
    if (opts->keep_records)
        for (JobRes *j : js)
            if (!j->rec.p || j->rec_cap < chunk)
                return fail(GD_ERR_ARG, "keep_records needs each job's record window to hold a chunk already "
                                        "(gd_batch_run or gd_montecarlo with records allocates it)");
    // shapes without trial launches: explicit, the autotuner's cached choice
    // or the heuristic
    std::vector<Shape> shapes;
    if (int rc = batch_shapes(ctx, t, js, opts, nullptr, false, shapes)) return rc;
    const int W = shapes[0].waves, max_rep = max_rep_of(shapes);
    std::vector<DevJob> table;
    std::vector<DevTask> tasks, tail;
    std::vector<int4> lad;
    batch_tables(js, shapes, opts, chunk, table, tasks, tail, lad);
    const size_t ntask = tasks.size();
    // scratch layout: counters | queue | jobs | tasks | tail | ladders
    const size_t o_bases = 0, o_queue = 256, o_jobs = o_queue + align256(sizeof(int) * (ntask + 1));
    const size_t o_tasks = o_jobs + align256(sizeof(DevJob) * table.size());
    const size_t o_tail = o_tasks + align256(sizeof(DevTask) * ntask);
    const size_t o_lad = o_tail + align256(sizeof(DevTask) * tail.size());
    const size_t need = o_lad + align256(sizeof(int4) * lad.size());
    if (need > kScratchBytes)
        return fail(GD_ERR_ARG, "the batch needs " + std::to_string(need) + " bytes of launch tables; the "
                                "no-allocation entry point holds " + std::to_string(kScratchBytes));
    void *base = nullptr;
    if (int rc = scratch_address(ctx->device, &base)) return rc;
    char *b = (char *)base;
    hipStream_t s = ctx->stream;
    const int zero2[2] = {0, 0};
    GD_HIP(hipMemcpyAsync(b + o_bases, zero2, sizeof(zero2), hipMemcpyHostToDevice, s));
    GD_HIP(hipMemcpyAsync(b + o_jobs, table.data(), sizeof(DevJob) * table.size(), hipMemcpyHostToDevice, s));
    GD_HIP(hipMemcpyAsync(b + o_tasks, tasks.data(), sizeof(DevTask) * ntask, hipMemcpyHostToDevice, s));
    if (!tail.empty())
        GD_HIP(hipMemcpyAsync(b + o_tail, tail.data(), sizeof(DevTask) * tail.size(), hipMemcpyHostToDevice, s));
    if (!lad.empty()) GD_HIP(hipMemcpyAsync(b + o_lad, lad.data(), sizeof(int4) * lad.size(), hipMemcpyHostToDevice, s));
    BatchMem mem;
    mem.bases = (int *)(b + o_bases);
    mem.queue = DevQueue{(int *)(b + o_queue), (int *)(b + o_queue) + 1, (int)ntask, 1}; // cleared by a kernel per chunk
    mem.jobs = (DevJob *)(b + o_jobs);
    mem.tasks = (DevTask *)(b + o_tasks);
    mem.tail = tail.empty() ? nullptr : (DevTask *)(b + o_tail);
    mem.ladders = (int4 *)(b + o_lad);
    mem.progress = progress_of(shapes);
    // no graph: instantiating one may allocate
    if (int rc = batch_launches(ctx, t, opts, chunk, W, max_rep, (int)ntask, (int)lad.size(), mem, false, elapsed_ms))
        return rc;
    batch_done(js, shapes, opts, chunk);
    return GD_OK;
}
