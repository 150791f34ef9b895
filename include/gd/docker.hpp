// docker.hpp -- the reference backend's Docker interface over the C ABI.
//
// Replaces class Docker of src/backend_gpu/backend_gpu.h:10-23 (and the CPU
// twin src/backend_cpu_mic/backend_cpu_mic.h:8-16): a caller fills
// `complex` (here from gd_load_job_line + gd_stage instead of
// Loader::build_complex, src/frontend/frontend.cpp:467-513), calls run(),
// and reads the accept records of the last chunk and the final replicas.
// run() follows Dock (src/backend_gpu/backend_gpu.cu:60-116): upload once,
// one MonteCarlo_d chunk of steps_per_dump steps at a time, record read-back
// after each chunk, and the replica-exchange hook between chunks filled in.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "geauxdock_amd.h"

namespace gd {

class Docker {
  public:
    const gd_Complex *complex = nullptr;  // staged job (not owned)
    std::vector<gd_ReplicaMC> record;     // [n_rep][record_cap] of the last chunk
    std::vector<int> next_entry;          // accepted entries per replica in `record`
    std::vector<gd_ReplicaMC> replica;    // final replica states after run()
    int record_cap = GD_STEPS_PER_DUMP;
    uint64_t seed = 2024;
    bool exchange = true;

    explicit Docker(int device = 0)
    {
        check(gd_create(device, &ctx_));
    }
    ~Docker()
    {
        if (ctx_) gd_destroy(ctx_);
    }
    Docker(const Docker &) = delete;
    Docker &operator=(const Docker &) = delete;

    void run()
    {
        if (!complex) throw std::runtime_error("Docker::run: no complex");
        const gd_ComplexSize &z = complex->size;
        const gd_McPara &mc = complex->mcpara;
        if (mc.steps_per_dump <= 0 || mc.steps_total < 0) throw std::runtime_error("Docker::run: bad step counts");
        // the job and target are freed on every way out of run()
        struct Held {
            gd_ctx *ctx;
            int target = -1, job = -1;
            ~Held()
            {
                if (job >= 0) gd_free_job(ctx, job);
                if (target >= 0) gd_free_target(ctx, target);
            }
        } held{ctx_};
        check(gd_upload_target(ctx_, complex, &held.target));
        check(gd_upload_job(ctx_, held.target, complex, seed, &held.job));
        const int job = held.job;
        const int cap = mc.steps_per_dump > record_cap ? mc.steps_per_dump : record_cap;
        record.assign((size_t)z.n_rep * cap, gd_ReplicaMC());
        next_entry.assign(z.n_rep, 0);
        int round = 0;
        for (int s1 = 0; s1 < mc.steps_total; s1 += mc.steps_per_dump) {
            const int s2 = (mc.steps_total - s1) < mc.steps_per_dump ? (mc.steps_total - s1) : mc.steps_per_dump;
            check(gd_montecarlo(ctx_, job, s1, s2));
            check(gd_read_records(ctx_, job, record.data(), next_entry.data(), cap));
            if (exchange && z.n_tmp > 1) check(gd_exchange(ctx_, job, round++, nullptr));
        }
        replica.assign(z.n_rep, gd_ReplicaMC());
        check(gd_read_replicas(ctx_, job, replica.data()));
        record_cap = cap;
    }

  private:
    gd_ctx *ctx_ = nullptr;
    static void check(int rc)
    {
        if (rc != GD_OK) throw std::runtime_error(std::string("geauxdock: ") + gd_last_error());
    }
};

} // namespace gd
