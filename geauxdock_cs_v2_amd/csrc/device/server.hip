// Serving loop: continuous batching of docking requests against one
// resident protein target on one GPU.
//
// Requests are queued by gd_server_submit.  A worker thread admits queued
// requests in arrival order while the device-memory budget and the slot count
// allow (admission by HBM footprint), then advances every admitted request by
// one chunk of Monte Carlo steps in a single fused launch -- requests that
// arrived while others were on the GPU join at the next chunk boundary, each
// at its own step counter.  Finished requests are read back, their device
// memory freed, and the freed room lets queued requests in.
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

struct Request {
    int id = 0;
    JobCopy job;
    int steps = 0;
    uint64_t seed = 0;
    int status = GD_ERR_BUSY;
    int job_id = -1;
    int done = 0;
    int rounds = 0;
    bool cancel = false;
    size_t footprint = 0;
    double t_submit = 0, t_admit = -1, t_done = -1;
    std::vector<gd_ReplicaMC> result;
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
    DevBuf jobs_dev, tasks_dev;

    double now() const
    {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    void loop();
    void admit_locked();
    void finish_locked(Request &q, int status);
};

void gd_server::finish_locked(Request &q, int status)
{
    if (q.job_id >= 0) {
        gd_free_job(ctx, q.job_id);
        q.job_id = -1;
        used -= q.footprint;
    }
    q.status = status;
    q.t_done = now();
    q.job = JobCopy(); // release host tables
}

void gd_server::admit_locked()
{
    while (!pending.empty() && (int)active.size() < opts.max_active_jobs) {
        Request &q = reqs.at(pending.front());
        if (q.cancel) {
            pending.pop_front();
            finish_locked(q, GD_ERR_STATE);
            continue;
        }
        if (used + q.footprint > opts.hbm_budget) break; // wait for room, keep arrival order
        pending.pop_front();
        int jid = -1;
        const int rc = gd_upload_job(ctx, target, &q.job.s.view, q.seed, &jid);
        if (rc != GD_OK) {
            finish_locked(q, rc);
            continue;
        }
        q.job_id = jid;
        used += q.footprint;
        q.t_admit = now();
        active.push_back(q.id);
    }
}

void gd_server::loop()
{
    (void)hipSetDevice(ctx->device);
    std::vector<DevJob> table;
    std::vector<DevTask> tasks;
    std::vector<int> ids;
    while (true) {
        {
            std::unique_lock<std::mutex> lk(mu);
            admit_locked();
            cv.notify_all();
            while (active.empty() && !stop && (pending.empty() || used + reqs.at(pending.front()).footprint > opts.hbm_budget)) {
                cv.wait(lk);
                admit_locked();
            }
            if (stop) return;
            ids = active;
            table.clear();
            tasks.clear();
            for (size_t i = 0; i < ids.size(); ++i) {
                Request &q = reqs.at(ids[i]);
                JobRes &j = ctx->jobs.at(q.job_id);
                table.push_back(j.hj);
                const int s2 = std::min(opts.chunk_steps, q.steps - q.done);
                build_tasks(j, (int)i, heuristic_shape(j.natom, kServerWaves, j.n_tmp * j.n_lig, j.n_prt, 2 * ctx->cu_count), q.done, s2, tasks);
            }
        }
        int max_rep = 1;
        for (const DevTask &t : tasks) max_rep = std::max(max_rep, t.nrep);
        std::stable_sort(tasks.begin(), tasks.end(), [&](const DevTask &x, const DevTask &y) {
            return (long)x.nrep * table[x.job].natom * x.s2max > (long)y.nrep * table[y.job].natom * y.s2max;
        });
        const TargetRes &t = ctx->targets.at(target);
        hipError_t e = upload(jobs_dev, table.data(), sizeof(DevJob) * table.size(), ctx->stream);
        if (e == hipSuccess) e = upload(tasks_dev, tasks.data(), sizeof(DevTask) * tasks.size(), ctx->stream);
        if (e == hipSuccess)
            e = launch_mc(t.dt, t.sc, jobs_dev.as<DevJob>(), tasks_dev.as<DevTask>(), (int)tasks.size(), kServerWaves,
                          max_rep, true, false, ctx->stream, nullptr);
        std::vector<int> exch(ids.size(), 0);
        if (e == hipSuccess && opts.exchange_interval > 0) {
            std::lock_guard<std::mutex> lk(mu);
            for (size_t i = 0; i < ids.size() && e == hipSuccess; ++i) {
                Request &q = reqs.at(ids[i]);
                const int after = q.done + std::min(opts.chunk_steps, q.steps - q.done);
                if (after % opts.exchange_interval == 0) {
                    JobRes &j = ctx->jobs.at(q.job_id);
                    e = launch_exchange(jobs_dev.as<DevJob>() + i, 0, j.n_lig * j.n_prt, j.n_lig, j.n_tmp, q.rounds,
                                        nullptr, ctx->stream, nullptr);
                    exch[i] = 1;
                }
            }
        }
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);

        std::lock_guard<std::mutex> lk(mu);
        std::vector<int> still;
        for (size_t i = 0; i < ids.size(); ++i) {
            Request &q = reqs.at(ids[i]);
            if (e != hipSuccess) {
                fail(GD_ERR_DEVICE, std::string("serving launch failed: ") + hipGetErrorString(e));
                finish_locked(q, GD_ERR_DEVICE);
                continue;
            }
            q.done += std::min(opts.chunk_steps, q.steps - q.done);
            q.rounds += exch[i];
            if (q.cancel) {
                finish_locked(q, GD_ERR_STATE);
            } else if (q.done >= q.steps) {
                q.result.resize(ctx->jobs.at(q.job_id).n_rep);
                const int rc = gd_read_replicas(ctx, q.job_id, q.result.data());
                finish_locked(q, rc);
            } else {
                still.push_back(q.id);
            }
        }
        active.swap(still);
        cv.notify_all();
    }
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
    delete srv;
}

int gd_server_submit(gd_server *srv, const gd_Complex *cx, int steps, uint64_t seed, int *req_id)
{
    if (!srv || !cx || !req_id || steps <= 0) return fail(GD_ERR_ARG, "bad submit arguments");
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
