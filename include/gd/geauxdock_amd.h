/*
 * geauxdock_amd.h -- C ABI of the MI355X replica Monte Carlo docking library
 * (libgeauxdock_amd.so).
 *
 * Every entry point returns 0 on success or a negative GD_ERR_* code; the
 * message of the last failure on the calling thread is gd_last_error().
 * Nothing here exits the process (the reference prints and calls
 * exit(EXIT_FAILURE), e.g. src/frontend/frontend.cpp:381-402).
 *
 * Boundary replaced.  The reference's backend is the class
 *   class Docker { Complex *complex; Record *record; Docker(); ~Docker(); void run(); }
 * (src/backend_gpu/backend_gpu.h:10-23, src/backend_cpu_mic/backend_cpu_mic.h:8-16),
 * whose run() loops the chunk kernel MonteCarlo_d(complex, record, s1, s2max)
 * (src/backend_gpu/kernel_cuda_l1_montecarlo.cu:29-33,
 *  src/backend_cpu_mic/kernel_cpp_montecarlo_soa.cpp:32-33) and reaches the
 * empty replica-exchange hook between chunks (src/backend_gpu/backend_gpu.cu:106).
 * Those calls map to gd_upload_target + gd_upload_job, gd_montecarlo,
 * gd_exchange and gd_read_records below; include/gd/docker.hpp wraps them
 * back into a Docker class.
 */
#ifndef GEAUXDOCK_AMD_H
#define GEAUXDOCK_AMD_H

#include <stddef.h>
#include <stdint.h>

#include "gd_types.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
    GD_OK = 0,
    GD_ERR_ARG = -1,     /* bad argument or size */
    GD_ERR_IO = -2,      /* input file missing or malformed */
    GD_ERR_DEVICE = -3,  /* HIP runtime error */
    GD_ERR_NOMEM = -4,   /* device memory budget exceeded */
    GD_ERR_STATE = -5,   /* call out of order / unknown handle */
    GD_ERR_BUSY = -6     /* serving: request not finished yet */
};

/* message of the calling thread's last failure (the reference prints and
 * exits instead, e.g. src/frontend/frontend.cpp:381-402) */
const char *gd_last_error(void);
/* library version string (no counterpart in the reference) */
const char *gd_version(void);

/* ------------------------------------------------------------------ */
/* random streams                                                      */
/* ------------------------------------------------------------------ */
/* Philox4x32-10 block (Salmon et al., SC'11; the reference vendors
 * Random123, src/module/Random123, whose kat_vectors:27-29 pin it):
 * out = philox(ctr, key). */
void gd_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* The injected stream: 31-bit integer k in [0, 2^31-1] for
 *   Monte Carlo draws (kind 0): replica idx_rep r, step s >= 1, slot j
 *     slots 0..5 = move components, slot 6 = Metropolis accept;
 *   exchange draws (kind 1): exchange round s, ladder r, pair slot j.
 * The uniform used is (float)k / (float)RAND_MAX, exactly as the reference's
 * MyRand_d (src/backend_cpu_mic/kernel_cpp_montecarlo_soa.cpp:8-15)
 * converts rand_r(); on glibc (float)RAND_MAX == 2^31, so u = (float)k * 2^-31. */
uint32_t gd_stream_k(uint64_t seed, uint32_t kind, uint32_t r, uint32_t s, uint32_t j);

/* ------------------------------------------------------------------ */
/* host loader and staging (no GPU needed)                             */
/* ------------------------------------------------------------------ */
typedef struct gd_raw_handle gd_raw_handle;
typedef struct gd_complex_handle gd_complex_handle;

/* Loader::load for one CSV job line (src/frontend/frontend.cpp:415-457):
 * reads protein, parameters, ligand and .ff files named in the line. */
int gd_load_job_line(const char *csv_line, gd_raw_handle **out);
/* the same with explicit file paths and run parameters */
int gd_load_job(const char *prt_path, const char *lig_path, const char *ff_path, const char *lig_id,
                const char *weight_path, const char *enepara_path, const char *nora_path,
                const char *norb_path, int n_tmp, float temp_high, float temp_low,
                float tras_scale, float rot_scale, int n_dump, gd_raw_handle **out);
const gd_Raw *gd_raw_view(const gd_raw_handle *h);
void gd_raw_free(gd_raw_handle *h);

/* Optimize* + set_temperature + set_replica (src/frontend/util_optimize_soa.cpp,
 * src/frontend/frontend.cpp:222-299): raw -> staged tables. */
int gd_stage(const gd_Raw *raw, gd_complex_handle **out);
const gd_Complex *gd_complex_view(const gd_complex_handle *h);
void gd_complex_free(gd_complex_handle *h);

/* structure sizes and offsets as compiled (0: sizeof gd_Complex,
 * 1: offsetof(gd_Complex, enepara), 2: sizeof gd_Raw, 3: offsetof(gd_Raw,
 * enepara0), 4: sizeof gd_EnePara, 5: sizeof gd_EnePara0, 6: sizeof
 * gd_ReplicaMC, 7: sizeof gd_Temp, 8: offsetof(gd_Complex, mcpara),
 * 9: offsetof(gd_Raw, n_tmp)); lets bindings check their mirrors */
size_t gd_abi_layout(int which);

/* Kabsch RMSD after optimal superposition, the conformer filter's measure
 * (u3b of src/frontend/rmsd.f:16-255, called at src/frontend/load.cpp:360);
 * exported for tests */
double gd_superposed_rmsd(const float *ax, const float *ay, const float *az, const float *bx,
                          const float *by, const float *bz, int n);

/* ------------------------------------------------------------------ */
/* device context                                                      */
/* ------------------------------------------------------------------ */
typedef struct gd_ctx gd_ctx;

/* Docker::Docker / ~Docker device setup and teardown
 * (src/backend_gpu/backend_gpu.cu:123-167) */
int gd_create(int device, gd_ctx **out);
void gd_destroy(gd_ctx *ctx);
/* HIP device properties of the context's GPU (no counterpart in the reference) */
int gd_device_info(gd_ctx *ctx, char *name, int name_len, int *cu_count, size_t *hbm_bytes);

/* Protein conformations, EnePara and Psp of a complex, kept resident on the
 * device for any number of jobs (Docker H2D of Complex, backend_gpu.cu:83,
 * split so the target is uploaded once). */
int gd_upload_target(gd_ctx *ctx, const gd_Complex *cx, int *target_id);
/* releases a target's device tables (the device copy's cudaFree in ~Docker,
 * backend_gpu.cu:151-158) */
int gd_free_target(gd_ctx *ctx, int target_id);

/* Ligand conformations, KDE, MCS, temperatures, replicas and move scales of
 * a complex, docked against a resident target.  `seed` keys the job's
 * injected random stream.  The replica table must have the layout
 * Loader::set_replica builds (src/frontend/frontend.cpp:267-299): replica
 * n_tmp*n_lig*p + n_lig*t + l of ligand conformation l and protein
 * conformation p, each (l, p) ladder holding every temperature once;
 * otherwise GD_ERR_ARG (gd_write_replicas checks the same). */
int gd_upload_job(gd_ctx *ctx, int target_id, const gd_Complex *cx, uint64_t seed, int *job_id);
/* The job's injected stream (replaces the reference's time-seeded rand_r,
 * kernel_cpp_montecarlo_soa.cpp:8-18, 233-234, 376): `seed` re-keys the
 * Philox stream (MC draws and exchange uniforms); with k_words non-null the
 * MC draws instead come from the caller's words k in [0, RAND_MAX],
 * u = (float)k / RAND_MAX as rand_r's, laid out [n_rep][n_steps][7] (per
 * replica idx_rep, per step 1..n_steps: move slots 0-5, then the accept
 * draw; step 0 draws nothing).  Launches that need steps beyond n_steps
 * fail with GD_ERR_ARG.  k_words null returns the job to Philox. */
int gd_set_stream(gd_ctx *ctx, int job_id, uint64_t seed, const uint32_t *k_words, int n_steps);
/* releases a job's device tables (as gd_free_target) */
int gd_free_job(gd_ctx *ctx, int job_id);
/* device bytes a job of these sizes takes (admission by HBM footprint; no
 * counterpart in the reference, whose Complex has static capacities) */
size_t gd_job_footprint(int lig_natom, int n_lig, int kde_npoint, int mcs_entries, int mcs_nrow,
                        int n_tmp, int n_rep, int record_cap);

/* MonteCarlo_d chunk, plain form: steps s1 .. s1+s2max-1 of every replica
 * of the job, one workgroup per replica, then a synchronous record read-back
 * window (kernel_cuda_l1_montecarlo.cu:29-551, backend_gpu.cu:91-107).
 * Records of the chunk: up to s2max accepted states per replica. */
int gd_montecarlo(gd_ctx *ctx, int job_id, int s1, int s2max);

/* Temperature replica exchange between neighbouring temperatures of every
 * (ligand conformation, protein conformation) ladder; round `round` pairs
 * (t, t+1) with t = round % 2, 2 + round % 2, ...; acceptance
 * min(1, exp((beta_t - beta_t+1)(E_t - E_t+1))) in FP64 against the kind-1
 * stream.  Accepted pairs swap temperature labels (ReplicaMC.idx_tmp).
 * n_accept (optional) receives the number of accepted swaps. */
int gd_exchange(gd_ctx *ctx, int job_id, int round, int *n_accept);

/* Pose energies for parity: for pose i, replica rep_idx[i] with movematrix
 * move[6i..6i+5]; writes the 10 normalised terms (vdw ele pmf psp hdb hpc
 * kde lhm dst total) to e_out[10i..]. */
int gd_energy(gd_ctx *ctx, int job_id, int n, const int *rep_idx, const float *move, float *e_out);

int gd_read_replicas(gd_ctx *ctx, int job_id, gd_ReplicaMC *out);
/* McLog counters (src/common/geauxdock.h:359-368, declared but never filled
 * in the reference): per replica since upload, accepted MC moves (steps >= 1,
 * every form) and accepted temperature swaps; either array may be null. */
int gd_read_mclog(gd_ctx *ctx, int job_id, int *mc_accepts, int *exchange_accepts);
int gd_write_replicas(gd_ctx *ctx, int job_id, const gd_ReplicaMC *in);
/* records of the last chunk: out[r * cap + k] for k < next_entry[r] */
int gd_read_records(gd_ctx *ctx, int job_id, gd_ReplicaMC *out, int *next_entry, int cap);

/* ------------------------------------------------------------------ */
/* batched form: all replicas of several jobs in fused launches per chunk */
/* ------------------------------------------------------------------ */
typedef struct gd_batch_opts {
    int steps;             /* MC steps per job */
    int exchange_interval; /* steps between exchange rounds, 0 = none */
    int use_graph;         /* capture the chunk sequence in a hipGraph */
    int slots_per_wave;    /* autotuned lane packing: 0 = auto */
    int waves_per_group;   /* 0 = auto */
    int point_slices;      /* 0 = auto */
    int compensated;       /* Kahan-compensated per-replica accumulation */
    int keep_records;      /* keep the accept records of the last chunk */
} gd_batch_opts;

int gd_batch_run(gd_ctx *ctx, const int *job_ids, int n_jobs, const gd_batch_opts *opts,
                 double *elapsed_ms);
/* gd_batch_run without device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The launch
 * tables (job descriptors, tasks, work queue, exchange ladders, step and
 * round counters) live in a fixed device array of gd_batch_noalloc_capacity()
 * bytes per GPU, held until the process ends; a batch that needs more fails
 * with GD_ERR_ARG before anything runs.  No hipMalloc, hipFree, graph or
 * autotuner trial: shapes are explicit, the autotuner's cached choice or the
 * packing heuristic; use_graph is ignored and elapsed_ms is host time;
 * keep_records needs each job's record window to hold a chunk already.  The
 * array keeps the last call's contents and calls that overlap on one GPU
 * race: serialise them. */
int gd_batch_run_noalloc(gd_ctx *ctx, const int *job_ids, int n_jobs, const gd_batch_opts *opts,
                         double *elapsed_ms);
/* bytes of the device array gd_batch_run_noalloc's tables live in */
size_t gd_batch_noalloc_capacity(void);
/* Workgroup shape of the job's last batched launch (no counterpart in the
 * reference, whose block shape is a build constant, TperB): shape[0..3] =
 * waves per workgroup, slot waves, point slices, replicas per workgroup. */
int gd_job_shape(gd_ctx *ctx, int job_id, int *shape);
/* Work queue of the context's last persistent Monte Carlo launch (no
 * counterpart in the reference, whose grid-stride loop over replicas,
 * src/backend_gpu/kernel_cuda_l1_montecarlo.cu:45, has no queue):
 * plan[0..2] = tasks, workgroups launched (the resident count, at most one
 * per item), step slices per task. */
int gd_batch_plan(gd_ctx *ctx, int *plan);

/* ------------------------------------------------------------------ */
/* serving: continuous batching of docking requests on one GPU          */
/* ------------------------------------------------------------------ */
typedef struct gd_server gd_server;
typedef struct gd_server_opts {
    size_t hbm_budget;     /* bytes of device memory jobs may take */
    int max_active_jobs;   /* slots */
    int chunk_steps;       /* MC steps per batched launch */
    int exchange_interval; /* 0 = none */
} gd_server_opts;

typedef struct gd_request_result {
    int status;            /* 0 done, GD_ERR_BUSY pending, other <0 failed/cancelled */
    int n_rep;
    double t_submit_ms, t_admit_ms, t_done_ms; /* server clock */
} gd_request_result;

int gd_server_create(gd_ctx *ctx, int target_id, const gd_server_opts *opts, gd_server **out);
void gd_server_destroy(gd_server *srv);
/* queue a request: the complex's job tables, steps, seed; returns an id */
int gd_server_submit(gd_server *srv, const gd_Complex *cx, int steps, uint64_t seed, int *req_id);
/* a request's status and times, and its final replicas once done; for a
 * failed or cancelled request gd_last_error() on the calling thread then
 * holds the reason */
int gd_server_poll(gd_server *srv, int req_id, gd_request_result *res, gd_ReplicaMC *final_replicas);
/* cancel a queued or running request; it finishes with GD_ERR_STATE */
int gd_server_cancel(gd_server *srv, int req_id);
/* run the serving loop until every queued request is finished */
int gd_server_drain(gd_server *srv);
/* the server clock (ms since gd_server_create) of gd_request_result's times */
double gd_server_clock_ms(gd_server *srv);

#ifdef __cplusplus
}
#endif

#endif /* GEAUXDOCK_AMD_H */
