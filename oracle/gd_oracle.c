/*
 * gd_oracle.c -- CPU oracle of the replica Monte Carlo docking path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the GPU library
 * is compared against; it is loaded only by the test suite, by the package's
 * smoke check and by the CPU-baseline leg of bench.py.  Nothing in the
 * product links or calls it.
 *
 * Contents (each restates the reference path named beside it):
 *   gdo_energy_f32/_f64   pose energy, float and double instantiations of
 *                         mc_body.h (the terms of MonteCarlo_d,
 *                         src/backend_cpu_mic/kernel_cpp_montecarlo_soa.cpp:219-359)
 *   gdo_montecarlo_f32    the chunk kernel MonteCarlo_d
 *                         (kernel_cpp_montecarlo_soa.cpp:32-420) with the
 *                         reference's time-seeded rand_r replaced by the
 *                         injected Philox stream (gdo_stream_k)
 *   gdo_exchange          temperature replica exchange; the reference only has
 *                         an empty hook (src/backend_cpu_mic/backend_cpu_mic.cpp:92),
 *                         so this is the textbook parallel-tempering rule
 *   gdo_philox4x32_10     Philox4x32-10, Salmon et al. SC'11
 *
 * Pins (tests/): the step-0 known-answer energies of the 1a07C1 complex
 * (doc/instructions.txt:172), bit-exact agreement with the reference's own
 * MonteCarlo_d compiled from its sources (oracle/_ref), the Random123
 * philox4x32-10 known-answer vectors, closed forms (identity rotation,
 * empty-term normalisation, detailed balance of the exchange), and FP64
 * brute force on tiny inputs.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "gd/gd_types.h"

#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------- Philox4x32-10 ---------------- */
void gdo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t x0 = ctr[0], x1 = ctr[1], x2 = ctr[2], x3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    int round;
    for (round = 0; round < 10; ++round) {
        const uint64_t a = (uint64_t)0xD2511F53u * x0;
        const uint64_t b = (uint64_t)0xCD9E8D57u * x2;
        const uint32_t n0 = (uint32_t)(b >> 32) ^ x1 ^ k0;
        const uint32_t n2 = (uint32_t)(a >> 32) ^ x3 ^ k1;
        x1 = (uint32_t)b;
        x3 = (uint32_t)a;
        x0 = n0;
        x2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = x0;
    out[1] = x1;
    out[2] = x2;
    out[3] = x3;
}

uint32_t gdo_stream_k(uint64_t seed, uint32_t kind, uint32_t r, uint32_t s, uint32_t j)
{
    const uint32_t ctr[4] = {s, r, kind, j >> 2};
    const uint32_t key[2] = {(uint32_t)(seed & 0xffffffffu), (uint32_t)(seed >> 32)};
    uint32_t out[4];
    gdo_philox4x32_10(ctr, key, out);
    return out[j & 3] >> 1;
}

/* (float)k / RAND_MAX as the reference writes it, RAND_MAX = 2147483647 */
static float draw_f(uint64_t seed, int r, int step, int slot)
{
    const uint32_t k = gdo_stream_k(seed, 0, (uint32_t)r, (uint32_t)step, (uint32_t)slot);
    return (float)(int)k / (float)2147483647;
}

/* ---------------- precision instantiations ---------------- */
#define REAL float
#define SFX _f32
#define RSIN sinf
#define RCOS cosf
#define REXP expf
#define RSQRT sqrtf
#define RLOG logf
#define RPI GD_PI
#include "mc_body.h"
#undef REAL
#undef SFX
#undef RSIN
#undef RCOS
#undef REXP
#undef RSQRT
#undef RLOG
#undef RPI

#define REAL double
#define SFX _f64
#define RSIN sin
#define RCOS cos
#define REXP exp
#define RSQRT sqrt
#define RLOG log
#define RPI 3.1415926535
#include "mc_body.h"
#undef REAL
#undef SFX
#undef RSIN
#undef RCOS
#undef REXP
#undef RSQRT
#undef RLOG
#undef RPI

static int check_view(const gd_Complex *cx)
{
    if (!cx) return -1;
    if (cx->size.lig_natom <= 0 || cx->size.prt_npoint <= 0 || cx->size.kde_npoint <= 0 ||
        cx->size.mcs_nrow <= 0 || cx->size.n_rep <= 0)
        return -1;
    return 0;
}

/* energies of n poses: pose i = replica rep_idx[i] (for its ligand and
 * protein conformation) at cumulative move move[6i..6i+5] */
int gdo_energy_f32(const gd_Complex *cx, int n, const int *rep_idx, const float *move, float *e_out)
{
    int i;
    if (check_view(cx) || n < 0) return -1;
#pragma omp parallel for schedule(dynamic, 1)
    for (i = 0; i < n; ++i) {
        const gd_ReplicaMC *rp = &cx->replica[rep_idx[i]];
        float *scratch = (float *)malloc(sizeof(float) * 6 * cx->size.lig_natom);
        float m[6];
        int j;
        for (j = 0; j < 6; ++j) m[j] = move[6 * i + j];
        pose_energy__f32(cx, rp->idx_lig, rp->idx_prt, m, e_out + 10 * (size_t)i, scratch);
        free(scratch);
    }
    return 0;
}

int gdo_energy_f64(const gd_Complex *cx, int n, const int *rep_idx, const float *move, double *e_out)
{
    int i;
    if (check_view(cx) || n < 0) return -1;
#pragma omp parallel for schedule(dynamic, 1)
    for (i = 0; i < n; ++i) {
        const gd_ReplicaMC *rp = &cx->replica[rep_idx[i]];
        double *scratch = (double *)malloc(sizeof(double) * 6 * cx->size.lig_natom);
        double m[6];
        int j;
        for (j = 0; j < 6; ++j) m[j] = (double)move[6 * i + j];
        pose_energy__f64(cx, rp->idx_lig, rp->idx_prt, m, e_out + 10 * (size_t)i, scratch);
        free(scratch);
    }
    return 0;
}

/*
 * MonteCarlo_d: steps s1 .. s1+s2max-1 for replicas [r0, r1) of rep[].
 * Records: rec[r * cap + k], k < next_entry[r] (window reset per call).
 * Draw order per step >= 1: move slots 0..5, then the accept slot 6;
 * step 0 draws nothing and is always accepted.
 */
int gdo_montecarlo_f32(const gd_Complex *cx, gd_ReplicaMC *rep, int r0, int r1, gd_ReplicaMC *rec,
                       int *next_entry, int cap, int s1, int s2max, uint64_t seed, int nthreads)
{
    int r;
    if (check_view(cx) || r0 < 0 || r1 > cx->size.n_rep || s1 < 0 || s2max < 0) return -1;
    if (rec && cap < s2max) return -1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(static)
    for (r = r0; r < r1; ++r) {
        gd_ReplicaMC *rp = &rep[r];
        float *scratch = (float *)malloc(sizeof(float) * 6 * cx->size.lig_natom);
        int is_accept_s = rp->is_accept;
        int s2;
        if (next_entry) next_entry[r] = 0;
        for (s2 = 0; s2 < s2max; ++s2) {
            const int step = s1 + s2;
            float movematrix[6], e_s[GD_MAXWEI];
            int i;
            for (i = 0; i < 6; ++i) {
                const float amount = step != 0 ? draw_f(seed, rp->idx_rep, step, i) : 0.0f;
                movematrix[i] = cx->mcpara.move_scale[i] * amount + rp->movematrix[i];
            }
            pose_energy__f32(cx, rp->idx_lig, rp->idx_prt, movematrix, e_s, scratch);
            if (step != 0) {
                const float delta_energy = e_s[GD_MAXWEI - 1] - rp->energy[GD_MAXWEI - 1];
                const float mybeta = cx->temp[rp->idx_tmp].minus_beta;
                is_accept_s = (draw_f(seed, rp->idx_rep, step, 6) < expf(delta_energy * mybeta));
            } else {
                is_accept_s = 1;
            }
            if (is_accept_s == 1) {
                for (i = 0; i < GD_MAXWEI; ++i) rp->energy[i] = e_s[i];
                for (i = 0; i < 6; ++i) rp->movematrix[i] = movematrix[i];
                rp->step = step;
                if (rec && next_entry) {
                    rec[(size_t)r * cap + next_entry[r]] = *rp;
                    next_entry[r] += 1;
                }
            }
        }
        rp->is_accept = is_accept_s;
        free(scratch);
    }
    return 0;
}

/*
 * One exchange round over every ladder (ligand conformation l, protein
 * conformation p).  Round `round` pairs temperature indices (t, t+1) with
 * t = round % 2 + 2i.  Pair i of ladder L = p * n_lig + l accepts when
 *   u < exp((beta_t - beta_t+1) * (E_a - E_b)),  u = k * 2^-31 (double),
 * k = stream_k(seed, 1, L, round, i), E = energy[9] of the replica currently
 * holding temperature t (a) and t+1 (b), beta = -minus_beta.  Accepted
 * pairs swap idx_tmp.  Returns the number of accepted swaps.
 */
int gdo_exchange(const gd_Complex *cx, gd_ReplicaMC *rep, int round, uint64_t seed)
{
    const int n_lig = cx->size.n_lig, n_prt = cx->size.n_prt, n_tmp = cx->size.n_tmp;
    int p, l, t, accepted = 0;
    int *holder = (int *)malloc(sizeof(int) * (n_tmp > 0 ? n_tmp : 1));
    for (p = 0; p < n_prt; ++p)
        for (l = 0; l < n_lig; ++l) {
            const int ladder = p * n_lig + l;
            for (t = 0; t < n_tmp; ++t) {
                const int r = n_tmp * n_lig * p + n_lig * t + l;
                holder[rep[r].idx_tmp] = r;
            }
            for (t = round % 2; t + 1 < n_tmp; t += 2) {
                const int a = holder[t], b = holder[t + 1];
                const double beta_a = -(double)cx->temp[t].minus_beta;
                const double beta_b = -(double)cx->temp[t + 1].minus_beta;
                const double delta = (beta_a - beta_b) *
                                     ((double)rep[a].energy[GD_MAXWEI - 1] - (double)rep[b].energy[GD_MAXWEI - 1]);
                const uint32_t k = gdo_stream_k(seed, 1, (uint32_t)ladder, (uint32_t)round, (uint32_t)(t / 2));
                const double u = (double)k * (1.0 / 2147483648.0);
                if (u < exp(delta)) {
                    rep[a].idx_tmp = t + 1;
                    rep[b].idx_tmp = t;
                    ++accepted;
                }
            }
        }
    free(holder);
    return accepted;
}
