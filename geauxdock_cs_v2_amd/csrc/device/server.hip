// Serving loop: continuous batching of docking requests against one
// resident protein target on one GPU.
//
// Requests are queued by gd_server_submit.  A worker thread admits queued
// requests in arrival order while the device-memory budget and the slot count
// allow (admission by HBM footprint) and places each in one of two lanes.  A
// lane advances all of its requests by one chunk of Monte Carlo steps in a
// single persistent launch on its own stream (the resident workgroups walk
// the lane's cost-ordered task queue); the two lanes run side by side and
// are relaunched independently, so when one lane's queue runs dry its
// workgroups give way to the other lane's.  Requests that arrive while the
// lanes are busy join an idle lane at its next chunk boundary, each at its
// own step counter; a chunk ends at the request's next exchange boundary.
// Finished requests are read back and their tables returned to the server's
// device arena (no hipFree, whose device-wide wait would hold the worker
// until the other lane's launch ends), and the freed room lets queued
// requests in.
//
// Every request keeps the workgroup shape its ligand size alone decides and
// its own keyed random stream, so its result is the same whether it runs
// alone or mixed with others.  Failures (bad input, device memory, a launch
// error) fail the requests involved and the loop goes on.
//
// While a server runs it owns the context: the caller issues no other calls
// on that gd_ctx until gd_server_destroy.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "context.hpp"
#include "../host/job.hpp"

namespace gd {
namespace {

constexpr int kServerWaves = 12;
constexpr int kLanes = gd_ctx::kGroups;
// a request's shape depends on the request alone (so its result does not
// depend on the mix): the shape that would fill the GPU best were it alone

// host copy of the job tables of a submitted complex
struct JobCopy {
    StagedJob s;
    void take(const gd_Complex *cx)
    {
        const gd_ComplexSize &z = cx->size;
        const size_t nl = (size_t)z.n_lig * z.lig_natom;
        s.lig_x2.assign(cx->lig_x2, cx->lig_x2 + nl);
        s.lig_y2.assign(cx->lig_y2, cx->lig_y2 + nl);
        s.lig_z2.assign(cx->lig_z2, cx->lig_z2 + nl);
        s.lig_c2.assign(cx->lig_c2, cx->lig_c2 + nl);
        s.lig_t2.assign(cx->lig_t2, cx->lig_t2 + nl);
        s.lig_x3.assign(cx->lig_x3, cx->lig_x3 + nl);
        s.lig_y3.assign(cx->lig_y3, cx->lig_y3 + nl);
        s.lig_z3.assign(cx->lig_z3, cx->lig_z3 + nl);
        s.lig_c3.assign(cx->lig_c3, cx->lig_c3 + nl);
        s.lig_t3.assign(cx->lig_t3, cx->lig_t3 + nl);
        s.kde_t.assign(cx->kde_t, cx->kde_t + z.kde_npoint);
        s.kde_x.assign(cx->kde_x, cx->kde_x + z.kde_npoint);
        s.kde_y.assign(cx->kde_y, cx->kde_y + z.kde_npoint);
        s.kde_z.assign(cx->kde_z, cx->kde_z + z.kde_npoint);
        const int nent = cx->mcs_rowptr[z.mcs_nrow];
        s.mcs_ncol.assign(cx->mcs_ncol, cx->mcs_ncol + z.mcs_nrow);
        s.mcs_tcc.assign(cx->mcs_tcc, cx->mcs_tcc + z.mcs_nrow);
        s.mcs_rowptr.assign(cx->mcs_rowptr, cx->mcs_rowptr + z.mcs_nrow + 1);
        s.mcs_i.assign(cx->mcs_i, cx->mcs_i + nent);
        s.mcs_x.assign(cx->mcs_x, cx->mcs_x + nent);
        s.mcs_y.assign(cx->mcs_y, cx->mcs_y + nent);
        s.mcs_z.assign(cx->mcs_z, cx->mcs_z + nent);
        s.temp.assign(cx->temp, cx->temp + z.n_tmp);
        s.replica.assign(cx->replica, cx->replica + z.n_rep);
        // protein topology of the first conformation (checked against the
        // target) and the job's own PSP table
        const int np = z.prt_npoint;
        s.prt_t.assign(cx->prt_t, cx->prt_t + np);
        s.prt_c.assign(cx->prt_c, cx->prt_c + np);
        s.prt_seq3r.assign(cx->prt_seq3r, cx->prt_seq3r + np);
        s.psp.assign(cx->psp, cx->psp + (size_t)GD_MAXTP2 * cx->psp_width);
        s.view = *cx; // sizes, centres, psp_width, enepara, mcpara
        s.bind();
        s.view.prt_x = s.view.prt_y = s.view.prt_z = nullptr; // coordinates live in the target
        s.view.prt_hpp = s.view.prt_ele = nullptr;
        s.view.prt_cdc = nullptr;
    }
};

// one chunk launch of a lane: its requests, their tables and step counts
struct Lane {
    std::vector<int> ids;
    std::vector<int> s2, exch; // steps of the chunk, exchange round after it
    std::vector<DevJob> table;
    std::vector<DevTask> tasks;
    DevBuf jobs_dev, tasks_dev, queue; // queue: the persistent launch's counter and done[]
    hipEvent_t done = nullptr;
    hipError_t err = hipSuccess;
    bool busy = false;
};

// Copy to a device buffer that grows by doubling: a lane's tables are
// replaced rarely, so the device-wide wait of the old buffer's hipFree is
// rare too.
hipError_t upload_grow(DevBuf &b, const void *src, size_t bytes, hipStream_t s)
{
    if (!b.p || b.bytes < bytes) {
        const hipError_t e = b.alloc(std::max(bytes, 2 * b.bytes));
        if (e != hipSuccess) return e;
    }
    return bytes ? hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, s) : hipSuccess;
}

struct Request {
    int id = 0;
    JobCopy job;
    int steps = 0;
    uint64_t seed = 0;
    int status = GD_ERR_BUSY;
    int job_id = -1;
    int done = 0;
    int rounds = 0;
    int lane = 0;
    bool cancel = false;
    size_t footprint = 0;
    double t_submit = 0, t_admit = -1, t_done = -1;
    std::vector<gd_ReplicaMC> result;
    std::string error; // why the request failed or was cancelled
};

} // namespace
} // namespace gd

using namespace gd;

struct gd_server {
    gd_ctx *ctx = nullptr;
    int target = -1;
    gd_server_opts opts{};
    std::mutex mu;
    std::condition_variable cv;
    std::deque<int> pending;
    std::vector<int> active;
    std::map<int, Request> reqs;
    int next_id = 1;
    size_t used = 0;
    bool stop = false;
    std::thread worker;
    std::chrono::steady_clock::time_point t0;
    Lane lanes[kLanes];
    // request tables live in this arena (capacity: the HBM budget plus
    // alignment slack), so finishing a request frees no device memory through
    // hipFree while the other lane runs
    Arena arena;

    double now() const
    {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    void loop();
    void admit_locked();
    void finish_locked(Request &q, int status, const std::string &why = std::string());
    double lane_load_locked(int k) const;
    void rebalance_locked();
    void prepare_locked(int k);
    void launch(int k);
    void complete_locked(int k);
    bool any_busy() const
    {
        for (const Lane &l : lanes)
            if (l.busy) return true;
        return false;
    }
};

// remaining pair evaluations of the lane's requests
double gd_server::lane_load_locked(int k) const
{
    double w = 0.0;
    for (int id : active) {
        const Request &q = reqs.at(id);
        if (q.lane == k) {
            const JobRes &j = ctx->jobs.at(q.job_id);
            w += (double)j.n_rep * j.natom * (q.steps - q.done);
        }
    }
    return w;
}

void gd_server::finish_locked(Request &q, int status, const std::string &why)
{
    if (q.job_id >= 0) {
        gd_free_job(ctx, q.job_id);
        q.job_id = -1;
        used -= q.footprint;
    }
    q.status = status;
    if (status != GD_OK) q.error = why;
    q.t_done = now();
    q.job = JobCopy(); // release host tables
}

void gd_server::admit_locked()
{
    while (!pending.empty() && (int)active.size() < opts.max_active_jobs) {
        Request &q = reqs.at(pending.front());
        if (q.cancel) {
            pending.pop_front();
            finish_locked(q, GD_ERR_STATE, "cancelled before admission");
            continue;
        }
        if (used + q.footprint > opts.hbm_budget) break; // wait for room, keep arrival order
        pending.pop_front();
        int jid = -1;
        ctx->job_arena = &arena; // this upload's tables come from the server's arena
        const int rc = gd_upload_job(ctx, target, &q.job.s.view, q.seed, &jid);
        ctx->job_arena = nullptr;
        if (rc == GD_ERR_NOMEM && !active.empty()) {
            // the arena is fragmented: wait for a running request to finish
            pending.push_front(q.id);
            break;
        }
        if (rc != GD_OK) {
            finish_locked(q, rc, std::string("admission: ") + gd_last_error());
            continue;
        }
        q.job_id = jid;
        used += q.footprint;
        q.t_admit = now();
        q.lane = lane_load_locked(1) < lane_load_locked(0) ? 1 : 0;
        active.push_back(q.id);
    }
}

// Before the idle lanes are gathered: an idle lane left with no requests
// takes about half the load of the idle lane with the most requests, if that
// lane has two or more (both idle, so no request in flight moves; a lane
// never gives away its last request).  Moving a request changes no result.
void gd_server::rebalance_locked()
{
    int count[kLanes] = {};
    for (int id : active) ++count[reqs.at(id).lane];
    for (int y = 0; y < kLanes; ++y) {
        if (lanes[y].busy || count[y] != 0) continue;
        int k = -1;
        for (int x = 0; x < kLanes; ++x)
            if (x != y && !lanes[x].busy && count[x] >= 2 && (k < 0 || count[x] > count[k])) k = x;
        if (k < 0) continue;
        const double half = 0.5 * lane_load_locked(k);
        double moved = 0.0;
        for (int id : active) {
            Request &q = reqs.at(id);
            if (q.lane != k || moved >= half || count[k] <= 1) continue;
            const JobRes &j = ctx->jobs.at(q.job_id);
            const double w = (double)j.n_rep * j.natom * (q.steps - q.done);
            if (moved + w > half && moved > 0.0) continue;
            q.lane = y;
            moved += w;
            --count[k];
            ++count[y];
        }
    }
}

// Gather the chunk of idle lane k.
void gd_server::prepare_locked(int k)
{
    Lane &L = lanes[k];
    L.ids.clear();
    L.s2.clear();
    L.exch.clear();
    L.table.clear();
    L.tasks.clear();
    L.err = hipSuccess;
    for (int id : active) {
        Request &q = reqs.at(id);
        if (q.lane != k) continue;
        JobRes &j = ctx->jobs.at(q.job_id);
        int s2 = std::min(opts.chunk_steps, q.steps - q.done);
        // a chunk ends at the request's next exchange boundary, so every
        // round falls due whatever the chunk length
        if (opts.exchange_interval > 0) s2 = std::min(s2, opts.exchange_interval - q.done % opts.exchange_interval);
        const int after = q.done + s2;
        L.ids.push_back(id);
        L.s2.push_back(s2);
        L.exch.push_back(opts.exchange_interval > 0 && after % opts.exchange_interval == 0 ? 1 : 0);
        build_tasks(j, (int)L.table.size(),
                    heuristic_shape(j.natom, kServerWaves, j.n_tmp * j.n_lig, j.n_prt, 2 * ctx->cu_count), q.done, s2,
                    L.tasks);
        L.table.push_back(j.hj);
    }
    std::stable_sort(L.tasks.begin(), L.tasks.end(), [&](const DevTask &x, const DevTask &y) {
        return (long)x.nrep * L.table[x.job].natom * x.s2max > (long)y.nrep * L.table[y.job].natom * y.s2max;
    });
}

// Enqueue lane k's chunk on its stream (outside the lock: the lane's tables
// belong to the worker thread while it is busy).
void gd_server::launch(int k)
{
    Lane &L = lanes[k];
    hipStream_t gs = ctx->group_stream[k];
    int max_rep = 1, min_s2 = 1 << 30;
    for (const DevTask &t : L.tasks) {
        max_rep = std::max(max_rep, t.nrep);
        min_s2 = std::min(min_s2, t.s2max);
    }
    const TargetRes &t = ctx->targets.at(target);
    const int ntask = (int)L.tasks.size();
    hipError_t e = upload_grow(L.jobs_dev, L.table.data(), sizeof(DevJob) * L.table.size(), gs);
    if (e == hipSuccess) e = upload_grow(L.tasks_dev, L.tasks.data(), sizeof(DevTask) * L.tasks.size(), gs);
    if (e == hipSuccess && L.queue.bytes < sizeof(int) * ((size_t)ntask + 1))
        e = L.queue.alloc(std::max(sizeof(int) * ((size_t)ntask + 1), 2 * L.queue.bytes));
    if (e == hipSuccess) {
        // a persistent launch over the lane's tasks: the lanes' launches
        // share the resident workgroups, the second one filling in as the
        // first one's queue runs dry
        const int slots = resident_slots(ctx, kServerWaves, max_rep);
        DevQueue qd{L.queue.as<int>(), L.queue.as<int>() + 1, ntask, choose_slices(ntask, slots, min_s2)};
        e = launch_queue_reset(qd, gs);
        if (e == hipSuccess)
            e = launch_mc_queue(t.params.as<DevParams>(), L.jobs_dev.as<DevJob>(), L.tasks_dev.as<DevTask>(), qd, slots, kServerWaves,
                                max_rep, true, false, false, gs, nullptr);
    }
    for (size_t i = 0; i < L.ids.size() && e == hipSuccess; ++i) {
        if (!L.exch[i]) continue;
        int rounds, n_lig, n_prt, n_tmp;
        {
            std::lock_guard<std::mutex> lk(mu);
            const Request &q = reqs.at(L.ids[i]);
            const JobRes &j = ctx->jobs.at(q.job_id);
            rounds = q.rounds, n_lig = j.n_lig, n_prt = j.n_prt, n_tmp = j.n_tmp;
        }
        e = launch_exchange(L.jobs_dev.as<DevJob>() + i, 0, n_lig * n_prt, n_lig, n_tmp, rounds, nullptr, gs, nullptr);
    }
    if (e == hipSuccess) e = hipEventRecord(L.done, gs);
    if (e != hipSuccess) (void)hipStreamSynchronize(gs);
    L.err = e;
    L.busy = true;
}

// Bookkeeping of lane k's finished chunk.
void gd_server::complete_locked(int k)
{
    Lane &L = lanes[k];
    const hipError_t e = L.err;
    for (size_t i = 0; i < L.ids.size(); ++i) {
        Request &q = reqs.at(L.ids[i]);
        if (e != hipSuccess) {
            finish_locked(q, GD_ERR_DEVICE, std::string("serving launch failed: ") + hipGetErrorString(e));
            continue;
        }
        q.done += L.s2[i];
        q.rounds += L.exch[i];
        if (q.cancel) {
            finish_locked(q, GD_ERR_STATE, "cancelled");
        } else if (q.done >= q.steps) {
            q.result.resize(ctx->jobs.at(q.job_id).n_rep);
            const int rc = gd_read_replicas(ctx, q.job_id, q.result.data());
            finish_locked(q, rc, rc == GD_OK ? std::string() : std::string("read-back: ") + gd_last_error());
        }
    }
    std::vector<int> still;
    for (int id : active)
        if (reqs.at(id).job_id >= 0) still.push_back(id);
    active.swap(still);
    L.ids.clear();
    L.busy = false;
    cv.notify_all();
}

void gd_server::loop()
{
    (void)hipSetDevice(ctx->device);
    bool ok = ensure_group_streams(ctx) == GD_OK;
    for (Lane &l : lanes) ok = ok && hipEventCreateWithFlags(&l.done, hipEventDisableTiming) == hipSuccess;
    while (true) {
        bool launch_k[kLanes] = {};
        {
            std::unique_lock<std::mutex> lk(mu);
            if (!ok) { // no streams: every request fails
                for (int id : active) finish_locked(reqs.at(id), GD_ERR_DEVICE, "the serving streams could not be created");
                active.clear();
                while (!pending.empty()) {
                    finish_locked(reqs.at(pending.front()), GD_ERR_DEVICE, "the serving streams could not be created");
                    pending.pop_front();
                }
                cv.notify_all();
                cv.wait(lk, [this] { return stop || !pending.empty(); });
                if (stop) break;
                continue;
            }
            admit_locked();
            cv.notify_all();
            while (!stop && active.empty() && (pending.empty() || used + reqs.at(pending.front()).footprint > opts.hbm_budget)) {
                cv.wait(lk);
                admit_locked();
            }
            if (stop) break;
            rebalance_locked();
            for (int k = 0; k < kLanes; ++k)
                if (!lanes[k].busy) {
                    prepare_locked(k);
                    launch_k[k] = !lanes[k].ids.empty();
                }
        }
        for (int k = 0; k < kLanes; ++k)
            if (launch_k[k]) launch(k);
        // wait for a lane to finish, or for a request an idle lane can take
        while (any_busy()) {
            bool finished = false;
            for (int k = 0; k < kLanes; ++k) {
                Lane &L = lanes[k];
                if (!L.busy) continue;
                const hipError_t q = L.err == hipSuccess ? hipEventQuery(L.done) : L.err;
                if (q == hipErrorNotReady) continue;
                if (q != hipSuccess) L.err = q;
                std::lock_guard<std::mutex> lk(mu);
                complete_locked(k);
                finished = true;
            }
            if (finished) break;
            {
                std::lock_guard<std::mutex> lk(mu);
                bool idle = false;
                for (const Lane &l : lanes) idle |= !l.busy;
                if (!stop && idle && !pending.empty() && used + reqs.at(pending.front()).footprint <= opts.hbm_budget &&
                    (int)active.size() < opts.max_active_jobs)
                    break;
            }
            std::this_thread::sleep_for(std::chrono::microseconds(50));
        }
    }
    for (int k = 0; k < kLanes; ++k)
        if (ctx->group_stream[k]) (void)hipStreamSynchronize(ctx->group_stream[k]);
    for (Lane &l : lanes)
        if (l.done) (void)hipEventDestroy(l.done);
}

extern "C" {

int gd_server_create(gd_ctx *ctx, int target_id, const gd_server_opts *opts, gd_server **out)
{
    if (!ctx || !opts || !out) return fail(GD_ERR_ARG, "null argument");
    if (!ctx->targets.count(target_id)) return fail(GD_ERR_STATE, "unknown target");
    if (opts->chunk_steps <= 0 || opts->max_active_jobs <= 0 || opts->hbm_budget == 0 || opts->exchange_interval < 0)
        return fail(GD_ERR_ARG, "bad server options");
    gd_server *s = new gd_server();
    s->ctx = ctx;
    s->target = target_id;
    s->opts = *opts;
    // the budget plus each admitted request's alignment slack
    const size_t slack = (size_t)opts->max_active_jobs * 64 * 1024;
    if (s->arena.init(opts->hbm_budget + slack) != hipSuccess) {
        delete s;
        return fail(GD_ERR_NOMEM, "cannot reserve the serving budget in device memory");
    }
    s->t0 = std::chrono::steady_clock::now();
    s->worker = std::thread([s] { s->loop(); });
    *out = s;
    return GD_OK;
}

void gd_server_destroy(gd_server *srv)
{
    if (!srv) return;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
    }
    srv->cv.notify_all();
    if (srv->worker.joinable()) srv->worker.join();
    for (auto &kv : srv->reqs)
        if (kv.second.job_id >= 0) gd_free_job(srv->ctx, kv.second.job_id);
    delete srv; // every job from its arena is gone
}

int gd_server_submit(gd_server *srv, const gd_Complex *cx, int steps, uint64_t seed, int *req_id)
{
    if (!srv || !cx || !req_id || steps <= 0) return fail(GD_ERR_ARG, "bad submit arguments");
    if (int rc = check_complex_arrays(cx, false, true)) return rc;
    const gd_ComplexSize &z = cx->size;
    if (z.lig_natom <= 0 || z.lig_natom > 64 * kServerWaves || z.n_rep <= 0 || z.mcs_nrow <= 0 || z.kde_npoint <= 0)
        return fail(GD_ERR_ARG, "bad request sizes");
    const size_t fp = gd_job_footprint(z.lig_natom, z.n_lig, z.kde_npoint, cx->mcs_rowptr[z.mcs_nrow], z.mcs_nrow,
                                       z.n_tmp, z.n_rep, 0);
    std::lock_guard<std::mutex> lk(srv->mu);
    if (fp > srv->opts.hbm_budget) return fail(GD_ERR_NOMEM, "request larger than the device-memory budget");
    const int id = srv->next_id++;
    Request &q = srv->reqs[id];
    q.id = id;
    q.job.take(cx);
    q.steps = steps;
    q.seed = seed;
    q.footprint = fp;
    q.t_submit = srv->now();
    srv->pending.push_back(id);
    srv->cv.notify_all();
    *req_id = id;
    return GD_OK;
}

int gd_server_poll(gd_server *srv, int req_id, gd_request_result *res, gd_ReplicaMC *final_replicas)
{
    if (!srv) return fail(GD_ERR_ARG, "null server");
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->reqs.find(req_id);
    if (it == srv->reqs.end()) return fail(GD_ERR_STATE, "unknown request");
    const Request &q = it->second;
    if (res) {
        res->status = q.status;
        res->n_rep = (int)q.result.size();
        res->t_submit_ms = q.t_submit;
        res->t_admit_ms = q.t_admit;
        res->t_done_ms = q.t_done;
    }
    if (q.status == GD_OK && final_replicas && !q.result.empty())
        std::memcpy(final_replicas, q.result.data(), sizeof(gd_ReplicaMC) * q.result.size());
    // a failed or cancelled request's reason becomes this thread's last error
    if (q.status != GD_OK && q.status != GD_ERR_BUSY) (void)fail(q.status, q.error);
    return GD_OK;
}

int gd_server_cancel(gd_server *srv, int req_id)
{
    if (!srv) return fail(GD_ERR_ARG, "null server");
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->reqs.find(req_id);
    if (it == srv->reqs.end()) return fail(GD_ERR_STATE, "unknown request");
    it->second.cancel = true;
    srv->cv.notify_all();
    return GD_OK;
}

int gd_server_drain(gd_server *srv)
{
    if (!srv) return fail(GD_ERR_ARG, "null server");
    std::unique_lock<std::mutex> lk(srv->mu);
    srv->cv.wait(lk, [srv] { return srv->pending.empty() && srv->active.empty(); });
    return GD_OK;
}

double gd_server_clock_ms(gd_server *srv) { return srv ? srv->now() : 0.0; }

} // extern "C"
