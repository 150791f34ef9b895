// Device context: resident targets and jobs, their device buffers and the
// host mirrors the launchers need.
#pragma once

#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "device.hpp"
#include "gd/geauxdock_amd.h"

namespace gd {

int fail(int code, const std::string &msg);

// A device-memory arena: one allocation carved into 256-byte-aligned blocks
// by first fit, with neighbouring free blocks merged.  Taking and giving back
// blocks calls no HIP function, so a thread that owns running streams can
// release memory without the device-wide wait of hipFree.
struct Arena {
    void *base = nullptr;
    size_t cap = 0, in_use = 0;
    std::map<size_t, size_t> free_blocks; // offset -> size
    Arena() = default;
    Arena(const Arena &) = delete;
    Arena &operator=(const Arena &) = delete;
    ~Arena();
    hipError_t init(size_t bytes);
    void *take(size_t n); // null when no block is large enough
    void give(void *p, size_t n);
    static size_t rounded(size_t n) { return (std::max<size_t>(n, 1) + 255) & ~(size_t)255; }
};

// owning device buffer, from hipMalloc or from an arena
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    Arena *arena = nullptr; // set before alloc: blocks come from (and return to) it
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes), arena(o.arena) { o.p = nullptr; o.bytes = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept
    {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            arena = o.arena;
            o.p = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void release();
    hipError_t alloc(size_t n);
    template <class T> T *as() const { return (T *)p; }
};

struct TargetRes {
    DevBuf pt, pq, pa, pw, pp, seg, ladder, ladder_in;
    DevBuf params; // DevParams {dt, sc} in device memory, read by the kernels
    DevTarget dt{};
    DevScalars sc{};
    int n_prt = 0, npoint = 0;
    float origin[3] = {0.f, 0.f, 0.f}; // frame origin (pocket centre) of all device coordinates
    std::vector<int> seg_host; // (type, begin, end) triples
    std::vector<int> dev_index;     // device position of each point (staged class order)
    std::vector<int> prt_t, prt_c;  // staged point types and classes, to check jobs against
    size_t bytes = 0;
};

// Workgroup shape of one task: R replicas x N_L atoms in slot_waves waves,
// times `slices` protein-point slices.
struct Shape {
    int nrep, slot_waves, slices, waves;
    int layout = 0; // TaskLayout bits
};

struct JobRes {
    int target = -1;
    DevBuf lig3, lig2, lt, krange, kde, mrow, ment, mtcc, mbeta, psp_pt, psp_row, rep, rec, next_entry, self, kstream,
        mclog;
    DevJob hj{};       // host copy of the descriptor
    int n_rep = 0, n_lig = 0, n_prt = 0, n_tmp = 0, natom = 0, nrow = 0;
    std::vector<int> rep_prt; // protein conformation of every replica (never changes)
    int rec_cap = 0;
    uint64_t seed = 0;
    int kstream_steps = 0; // steps covered by explicit stream words (0: Philox)
    size_t bytes = 0;
    int steps_done = 0;      // steps advanced by batched runs
    int exchange_rounds = 0; // exchange rounds done by batched runs
    Shape last_shape{0, 0, 0, 0}; // workgroup shape of the last batched launch
};

Shape plain_shape(int natom);
// replica packing for a W-wave workgroup: best lane use x fill of `slots`
// resident workgroups, then the fewest point slices
Shape heuristic_shape(int natom, int W, int max_rep_block, int n_prt, int slots);
// tasks of one job: consecutive replicas of each protein conformation, R per task
void build_tasks(const JobRes &j, int job_slot, const Shape &sh, int s1, int s2max, std::vector<DevTask> &out);

} // namespace gd

struct gd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int cu_count = 0;
    size_t hbm_bytes = 0;
    std::string name;
    std::map<int, gd::TargetRes> targets;
    std::map<int, gd::JobRes> jobs;
    int next_id = 1;
    // launch scratch reused across calls
    gd::DevBuf task_buf, job_buf, pose_rep, pose_move, pose_out, counter;
    // the batched form's work queue (counter + done per task), exchange
    // ladder list and graph step/round counters
    gd::DevBuf queue, ladders, bases;
    // work queue of the last persistent launch: tasks, grid, step slices
    int last_plan[3] = {0, 0, 0};
    // when set (by a server, around its own uploads), job uploads take their
    // device tables from this arena instead of hipMalloc
    gd::Arena *job_arena = nullptr;
    // streams of the batched form's workgroup groups (created on first use,
    // default priority)
    static constexpr int kGroups = 2;
    hipStream_t group_stream[kGroups] = {nullptr, nullptr};
};

namespace gd {
hipError_t upload(DevBuf &b, const void *src, size_t bytes, hipStream_t s);
int ensure_records(gd_ctx *ctx, JobRes &j, int cap);
// fails unless the job's explicit stream words (if any) cover every step
// before s_end (steps from 1 on draw words; step 0 draws none)
int check_stream_range(const JobRes &j, long s_end);
int check_complex_arrays(const gd_Complex *cx, bool target, bool job);
int run_tasks(gd_ctx *ctx, const TargetRes &t, const DevJob *jobs_dev, const std::vector<DevTask> &tasks,
              int waves, int max_rep, bool compensated, bool records);
// workgroups of shape (waves, max_rep) resident on the whole GPU at once
int resident_slots(gd_ctx *ctx, int waves, int max_rep);
// step slices per task for a persistent launch of ntask tasks on `slots`
// resident workgroups, each task running `steps` steps
int choose_slices(int ntask, int slots, int steps);
// creates the context's group streams on first use
int ensure_group_streams(gd_ctx *ctx);
} // namespace gd
