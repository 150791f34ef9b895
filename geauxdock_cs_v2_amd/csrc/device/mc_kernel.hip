// Fused replica Monte Carlo kernel for gfx950 (CDNA4, wave64).
//
// One workgroup advances one task: R replicas of one job that share a
// protein conformation, for s2max steps (the chunk MonteCarlo_d runs,
// src/backend_cpu_mic/kernel_cpp_montecarlo_soa.cpp:32-420).  Replica state
// lives in LDS for the whole chunk.
//
// Lane mapping.  The R replicas' ligand atoms are laid out as R * N_L
// "slots"; slot wave sw holds slots [64 sw, 64 sw + 64).  The workgroup has
// slot_waves * slices waves: wave w serves slot wave w % slot_waves and
// protein-point slice w / slot_waves.  Every lane of a wave walks the same
// protein points, so point data is wave-uniform; each lane keeps its own
// ligand atom.  Points are sorted by type, so the six type-pair parameters
// of a lane are loaded once per type segment, not per pair.
//
// One step, per slot: proposal move from the injected stream, rotation and
// translation of the atom, the six pair terms over the slice's points
// (src/backend_cpu_mic/energy_prt_v1.cpp), KDE over the atom type's KDE run
// (energy_kde_v1.cpp) and MCS rows (energy_mcs_v3_ell.cpp), both spread over
// the slices so that no wave waits on another's gather chain; cross-slice
// sum and per-atom hydrophobic restraint; then a per-replica sum,
// normalisation, weighted total, Metropolis accept and record append
// (kernel_cpp_montecarlo_soa.cpp:333-390) by one thread per replica.
#include <algorithm>

#include "device.hpp"

namespace gd {

namespace {

__device__ __forceinline__ void philox10(uint32_t &c0, uint32_t &c1, uint32_t &c2, uint32_t &c3, uint32_t k0,
                                         uint32_t k1)
{
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        const uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// (float)k / RAND_MAX with k = word >> 1; (float)RAND_MAX == 2^31
__device__ __forceinline__ float uniform31(uint32_t w) { return (float)(w >> 1) * 0x1p-31f; }

__device__ __forceinline__ float fast_exp(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }

// sin and cos of a cumulative Euler angle.  Reduction by Cody and Waite
// (Software Manual for the Elementary Functions, 1980): k = rint(x / (pi/2)),
// r = x - k pi/2 with pi/2 split into three floats P1 + P2 + P3 (P1 the float
// nearest pi/2, P2 and P3 the float-rounded remainders).  fma(-k, P1, x) is
// exact for |k| < 2^23 (k P1 and x share the grid 2^-23 min(1, ulp x) and the
// difference is below 1 in magnitude), the P2 and P3 steps round once each,
// so r carries about one ulp for the angles a run reaches (|x| up to 10^6
// rad; the moves add at most rot_scale per step).  On |r| <= pi/4: Taylor
// polynomials of sin through r^9 (next term below 2.5e-9) and of cos through
// r^10 (below 1.2e-10); quadrant from k mod 4.  No tables and no wide-argument
// branch (the library routine's reduction tables held six VGPRs of
// constants across the whole step).
__device__ __forceinline__ void sincos_cw(float x, float &s, float &c)
{
    const float kf = rintf(x * 0.636619772367581343f);
    float r = fmaf(-kf, 1.57079637050628662f, x);
    r = fmaf(-kf, -4.37113882867379e-08f, r);
    r = fmaf(-kf, -1.71512451000588e-15f, r);
    const int q = (int)kf;
    const float r2 = r * r;
    const float ps = fmaf(r2, fmaf(r2, fmaf(r2, 1.0f / 362880.0f, -1.0f / 5040.0f), 1.0f / 120.0f), -1.0f / 6.0f);
    const float sn = fmaf(r * r2, ps, r);
    const float pc = fmaf(r2, fmaf(r2, fmaf(r2, fmaf(r2, -1.0f / 3628800.0f, 1.0f / 40320.0f), -1.0f / 720.0f),
                                   1.0f / 24.0f),
                          -0.5f);
    const float cs = fmaf(r2, pc, 1.0f);
    // x = k pi/2 + r: quadrant k mod 4 rotates (sin r, cos r)
    const bool odd = q & 1, neg_s = q & 2, neg_c = (q + 1) & 2;
    const float ss = odd ? cs : sn, cc = odd ? sn : cs;
    s = neg_s ? -ss : ss;
    c = neg_c ? -cc : cc;
}

// scale * u + m rounded twice, as the reference computes the proposal
// (kernel_cpp_montecarlo_soa.cpp:234); contraction into an FMA would change
// the last bit of the cumulative move
__device__ __forceinline__ float move_update(float scale, float u, float m)
{
#pragma clang fp contract(off)
    return scale * u + m;
}

template <bool KAHAN>
__device__ __forceinline__ void acc_add(float &s, float &c, float v)
{
    if (KAHAN) {
        const float y = v - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    } else {
        s += v;
    }
}

struct Slot {
    float x, y, z; // transformed atom, relative to the ladder centre of the point table in use
    int tl;        // ligand type
};

constexpr int kBlock = 64; // points per plain FP32 partial sum
constexpr float kExpandR = 16.0f; // A: largest distance of a wave's atoms from its ladder centre for expanded distances
constexpr int kVals = 9;   // per-slot terms: vdw, ele hi, pmf, psp, hdb, hpc, kde, mcs, ele lo
constexpr int kRed = 9;    // per-slot partials of a slice: vdw, ele, pmf, psp, hdb, hpc, ele compensation, kde, mcs
static_assert(kRed <= kVals, "the cross-slice partials share the per-slot terms' LDS region");
constexpr int kIst = 5;    // ints of replica state per replica in LDS

} // namespace

namespace {

struct PairParams {
    float p1a, p2a, pmf1, pmfb, pmfc, hk2, hk1, hk0, hdb1;
};

// The scalars the pair loops use (DevScalars), read once per step.
struct PairConst {
    float lj_a, lj_b, pmf_a;
    float el_e0, el_a, el_b, el_dc2;
};

struct Sums {
    float v, e, p, h, q; // vdw (lj0-scaled), ele (el1-scaled, without the atom charge), pmf, hdb, hydrophobic
    float ec;            // compensation of e: the sum is e - ec
};

// Compensated (Kahan) add of v to the pair (s, c), whose value is s - c.
__device__ __forceinline__ void kahan_add(float &s, float &c, float v)
{
    const float y = v - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

// The target tables are read-only while a kernel runs: reading them through
// the constant address space lets the compiler fetch wave-uniform points with
// scalar loads into SGPRs instead of 64 identical vector loads.
typedef const __attribute__((address_space(4))) float cfloat;
typedef const __attribute__((address_space(4))) int cint;

// A descriptor (DevParams, DevJob) read through the constant address space
// from a point the compiler cannot move: the empty asm makes the pointer
// opaque, so each field is fetched with a scalar load where the code uses it
// instead of being hoisted to the kernel's start and held in SGPRs across
// the whole launch (which spilled them into VGPR lanes).
template <class T>
__device__ __forceinline__ const __attribute__((address_space(4))) T &uref(const T *p)
{
    auto q = (const __attribute__((address_space(4))) T *)p;
    asm volatile("" : "+s"(q));
    return *q;
}

__device__ __forceinline__ float4 ld_point(const cfloat *pt, int p)
{
    return make_float4(pt[4 * p], pt[4 * p + 1], pt[4 * p + 2], pt[4 * p + 3]);
}

// Electrostatics (:79-85), times el1: el1 g1 is 1/d outside the core and
// el1 (el0 + a1 s^2 + b1 s^3), s = el1 d, inside s < 1 (the core meets 1/s
// with equal value and slope at s = 1).  Every pair adds ele / d; a pair
// inside the core then adds ele (core - 1/d).
__device__ __forceinline__ float ele_core_delta(float d2, float rinv, float ele, const PairConst &sc)
{
    const float d = d2 * rinv;
    const float core = fmaf(d2, fmaf(sc.el_b, d, sc.el_a), sc.el_e0);
    return d2 < sc.el_dc2 ? ele * (core - rinv) : 0.0f;
}

// Soft-core Lennard-Jones (:62-71), times c (c = lj0, or 1 where lj0 = 0):
// with X = c pp1 = c p1a / d^9, c (pp1 - pp2) / pp4 = (X - c p2a / d^6) /
// (1 + X L), L = (lj0 / c)(1 + lj1 d2); numerator and denominator times d^9:
// (c p1a - c p2a d^3) / (d^9 + c p1a L), seven operations and one reciprocal
// from d (the powers of 1/d take three more).  c p1a > 0 for every type pair
// of the parameter table and L > 0, so the denominator is positive.
__device__ __forceinline__ void lj_term(float d2, float d, const PairParams &q, const PairConst &sc, Sums &s)
{
    const float d3 = d2 * d;
    const float d9 = (d3 * d3) * d3;
    const float den = fmaf(q.p1a, fmaf(sc.lj_b, d2, sc.lj_a), d9);
    s.v = fmaf(fmaf(-q.p2a, d3, q.p1a), __builtin_amdgcn_rcpf(den), s.v);
}

// Terms of one pair once d2 and 1/d are known, electrostatics aside: the
// electrostatic sum cancels by up to 1e6 over a pose, so its terms are added
// four at a time with compensation by the caller.
template <bool HPC, bool HDB>
__device__ __forceinline__ void pair_terms(float d2, float rinv, float hw, const PairParams &q,
                                           const PairConst &sc, Sums &s)
{
    const float d = d2 * rinv;
    // hydrophobic contact inside 9 A (:44-50): with t = d2 / 81,
    // 1 - 3.5t + 4.5t^2 - 2.5t^3 + 0.5t^4 = (1-t)^3 (1 - t/2) = u^3 (1 + u) / 2
    // for u = 1 - t; u clamps to 0 beyond the cutoff, the weight holds the 1/2
    if (HPC) {
        const float u = __builtin_amdgcn_fmed3f(fmaf(-d2, 1.0f / 81.0f, 1.0f), 0.0f, 1.0f);
        s.q = fmaf(hw, (u * u) * fmaf(u, u, u), s.q);
    }
    lj_term(d2, d, q, sc, s);
    // contact potential pmf1 / (1 + exp((6 - d/2)(d - pmf0))) (:95-98)
    const float xe = fmaf(sc.pmf_a, d2, fmaf(q.pmfb, d, q.pmfc));
    s.p = fmaf(q.pmf1, __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(xe)), s.p);
    // hydrogen bond hdb1 exp(-((d - hdb0) hdb1)^2 / 2) (:138-143)
    if (HDB) {
        const float xh = fmaf(q.hk2, d2, fmaf(q.hk1, d, q.hk0));
        s.h = fmaf(q.hdb1, __builtin_amdgcn_exp2f(xh), s.h);
    }
}

// Pair terms of one atom against points [pb, pe) of one type segment
// (src/backend_cpu_mic/energy_prt_v1.cpp:30-146) except the pocket-specific
// term, which the job's short PSP point list adds separately.  HPC: every
// point of the segment is a hydrophobic-contact point (cdc with nonzero hpp);
// the other points live in other segments, so the loop has no per-point
// branch.  HDB: some lane of the wave has an H-bond with this point type;
// otherwise every lane's hdb1 is 0 and the term is 0 exactly.  Points go in
// groups of four with one wave-wide test for electrostatic-core pairs
// (closer than 1 / el1: clashes, rare), whose corrections are skipped
// otherwise.
template <bool HPC, bool HDB, bool CORE = true>
__device__ __forceinline__ void pair_block(const float4 *ptc_, const float *pw_, int pb, int pe, const Slot &me,
                                           const PairParams &q, const PairConst &sc, Sums &s)
{
    const cfloat *ptc = (const cfloat *)ptc_;
    const cfloat *pw = (const cfloat *)pw_; // hydrophobic weights
    int p = pb;
    for (; p + 4 <= pe; p += 4) {
        // one base address per group: the four points are one 64-byte
        // scalar load, their weights one 16-byte load
        const cfloat *g4 = ptc + 4 * p;
        const cfloat *w4 = pw + p;
        float4 P[4];
        float d2[4], rinv[4], w[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            P[j] = make_float4(g4[4 * j], g4[4 * j + 1], g4[4 * j + 2], g4[4 * j + 3]);
            w[j] = HPC ? w4[j] : 0.0f;
            const float dx = fmaf(0.5f, P[j].x, me.x), dy = fmaf(0.5f, P[j].y, me.y), dz = fmaf(0.5f, P[j].z, me.z);
            d2[j] = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
            rinv[j] = __builtin_amdgcn_rsqf(d2[j]);
        }
        float ge = 0.0f; // the group's electrostatics
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            pair_terms<HPC, HDB>(d2[j], rinv[j], w[j], q, sc, s);
            ge = fmaf(P[j].w, rinv[j], ge);
        }
        // core corrections where some lane of the wave has a core pair in the group
        if (CORE) {
            const bool near = (d2[0] < sc.el_dc2) | (d2[1] < sc.el_dc2) | (d2[2] < sc.el_dc2) | (d2[3] < sc.el_dc2);
            if (__any(near)) {
#pragma unroll
                for (int j = 0; j < 4; ++j) ge += ele_core_delta(d2[j], rinv[j], P[j].w, sc);
            }
        }
        kahan_add(s.e, s.ec, ge);
    }
    for (; p < pe; ++p) {
        const float4 P4 = ld_point(ptc, p);
        const float dx = fmaf(0.5f, P4.x, me.x), dy = fmaf(0.5f, P4.y, me.y), dz = fmaf(0.5f, P4.z, me.z);
        const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        const float rinv = __builtin_amdgcn_rsqf(d2);
        pair_terms<HPC, HDB>(d2, rinv, HPC ? pw[p] : 0.0f, q, sc, s);
        kahan_add(s.e, s.ec, fmaf(P4.w, rinv, CORE ? ele_core_delta(d2, rinv, P4.w, sc) : 0.0f));
    }
}

// Far field: points at least far_d from every atom of the wave (see
// DevTarget::far_d).  Only LJ and 1/d electrostatics vary there; the contact
// term is pmf1 per point, the hydrophobic, core and H-bond terms are 0.
// Squared distance of the atom me from point P (q' = P.xyz, see
// DevTarget::pt).  EXPAND: |u|^2 + |q|^2 + u.q', one add and three FMAs
// instead of three subtractions and three multiply-adds; its rounding error
// is about 2^-22 (|u| + |q|)^2, so it is used only where d is large against
// |u| + |q| (far and very far pairs of waves near their ladder centre, and
// the lean path).
template <bool EXPAND>
__device__ __forceinline__ float dist2(const float4 &P, float qq, const Slot &me, float uu)
{
    if (EXPAND) return fmaf(me.x, P.x, fmaf(me.y, P.y, fmaf(me.z, P.z, uu + qq)));
    const float dx = fmaf(0.5f, P.x, me.x), dy = fmaf(0.5f, P.y, me.y), dz = fmaf(0.5f, P.z, me.z);
    return fmaf(dx, dx, fmaf(dy, dy, dz * dz));
}

template <bool EXPAND>
__device__ __forceinline__ void pair_block_far(const float4 *ptc_, const float *pq_, int pb, int pe, const Slot &me,
                                               float uu, const PairParams &q, const PairConst &sc, Sums &s)
{
    const cfloat *ptc = (const cfloat *)ptc_;
    const cfloat *pq = (const cfloat *)pq_;
    int p = pb;
    for (; p + 4 <= pe; p += 4) {
        float ge = 0.0f;
        const cfloat *q4 = pq + p;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float4 P4 = ld_point(ptc, p + j);
            const float d2 = dist2<EXPAND>(P4, EXPAND ? q4[j] : 0.0f, me, uu);
            const float rinv = __builtin_amdgcn_rsqf(d2);
            lj_term(d2, d2 * rinv, q, sc, s);
            ge = fmaf(P4.w, rinv, ge);
        }
        kahan_add(s.e, s.ec, ge);
    }
    for (; p < pe; ++p) {
        const float4 P4 = ld_point(ptc, p);
        const float d2 = dist2<EXPAND>(P4, EXPAND ? pq[p] : 0.0f, me, uu);
        const float rinv = __builtin_amdgcn_rsqf(d2);
        lj_term(d2, d2 * rinv, q, sc, s);
        kahan_add(s.e, s.ec, P4.w * rinv);
    }
    s.p = fmaf(q.pmf1, (float)(pe - pb), s.p);
}

// Very far field: points at least vfar_d from every atom of the wave, where
// the LJ softening e = X (lj_a + lj_b d2) is at most 3e-3 for every type
// pair (DevTarget::vfar_d): n / (1 + e) = n (1 - e + e^2) to float rounding,
// so the pair needs one transcendental (1/d) instead of two.  TIER 1: every
// pair beyond lean_e_d (e <= 2^-12.5), where n (1 - e) is as close; TIER 2:
// beyond lean_soft_d (e <= 2^-25), where n alone is.
// LJ of one very far pair into s.v (TIER 2: its attractive part into v2);
// returns 1/d for the electrostatics
template <int TIER>
__device__ __forceinline__ float vfar_lj(float d2, const PairParams &q, const PairConst &sc, Sums &s, float &v2)
{
    const float rinv = __builtin_amdgcn_rsqf(d2);
    const float r2 = rinv * rinv;
    const float r3 = r2 * rinv;
    const float r6 = r3 * r3;
    if (TIER == 0) {
        const float X = q.p1a * (r6 * r3);
        const float n = fmaf(-q.p2a, r6, X);
        const float e = X * fmaf(sc.lj_b, d2, sc.lj_a);
        s.v += fmaf(-n, fmaf(-e, e, e), n); // n - n (e - e^2)
    } else if (TIER == 1) {
        const float X = q.p1a * (r6 * r3);
        const float n = fmaf(-q.p2a, r6, X);
        s.v = fmaf(n, fmaf(-X, fmaf(sc.lj_b, d2, sc.lj_a), 1.0f), s.v); // n (1 - e)
    } else {
        // n = X - c p2a r^6 as two sums: at this range the repulsive part is
        // a small fraction of the attractive one (below 1e-4 for the shipped
        // parameters), so they do not cancel
        s.v = fmaf(q.p1a, r6 * r3, s.v);
        v2 = fmaf(q.p2a, r6, v2);
    }
    return rinv;
}

template <int TIER, bool EXPAND>
__device__ __forceinline__ void pair_block_vfar(const float4 *ptc_, const float *pq_, int pb, int pe, const Slot &me,
                                                float uu, const PairParams &q, const PairConst &sc, Sums &s)
{
    const cfloat *ptc = (const cfloat *)ptc_;
    const cfloat *pq = (const cfloat *)pq_;
    int p = pb;
    float v2 = 0.0f;
    for (; p + 4 <= pe; p += 4) {
        float ge = 0.0f;
        const cfloat *q4 = pq + p;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float4 P4 = ld_point(ptc, p + j);
            const float d2 = dist2<EXPAND>(P4, EXPAND ? q4[j] : 0.0f, me, uu);
            ge = fmaf(P4.w, vfar_lj<TIER>(d2, q, sc, s, v2), ge);
        }
        kahan_add(s.e, s.ec, ge);
    }
    // the last points of the block, with the same arithmetic
    for (; p < pe; ++p) {
        const float4 P4 = ld_point(ptc, p);
        const float d2 = dist2<EXPAND>(P4, EXPAND ? pq[p] : 0.0f, me, uu);
        kahan_add(s.e, s.ec, P4.w * vfar_lj<TIER>(d2, q, sc, s, v2));
    }
    if (TIER == 2) s.v -= v2;
    s.p = fmaf(q.pmf1, (float)(pe - pb), s.p);
}

__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
    return v;
}

__device__ __forceinline__ float wave_min(float v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off));
    return v;
}

} // namespace

#ifdef GD_PHASE_PROFILE
// Phase profile (the library built with -DGD_PHASE_PROFILE, a tracing build
// beside the product library): wave-cycles spent in each phase of a step,
// each wave index's pair-loop cycles, and, per workgroup step, W times the
// slowest wave's pair-loop cycles.  Summed in LDS per workgroup and added
// here when the workgroup exits; read by gd_phase_profile.
constexpr int kPhases = 9;
__device__ unsigned long long g_phase[16];      // [kPhases] phases, [9] W x slowest, [10] workgroup steps
__device__ unsigned long long g_phase_wave[16]; // pair-loop cycles per wave index
// wall-clock start and exit of each workgroup of the last launch (the queue
// tail: how long resident workgroups sit idle while the last items run)
constexpr int kWgMax = 8192;
__device__ unsigned long long g_wg_time[2][kWgMax];
#define GD_PHASE(k)                                                                                                    \
    do {                                                                                                               \
        const unsigned long long t_ = clock64();                                                                       \
        if (lane == 0) atomicAdd(&s_ph[k], t_ - ph_t);                                                                 \
        ph_t = t_;                                                                                                     \
    } while (0)
#else
#define GD_PHASE(k) \
    do {            \
    } while (0)
#endif

size_t mc_lds_bytes(int waves, int nrep)
{
    const int lanes = waves * 64;
    // one region for the cross-slice partials (kRed floats per slot and
    // slice beyond the first) and, after a barrier, the per-slot terms (kVals
    // floats per slot); then rsum, st, prop (16 + 16 + 8 floats per replica)
    // and kIst ints per replica
    return sizeof(float) * ((size_t)lanes * kVals + (size_t)nrep * 40) + sizeof(int) * (size_t)nrep * kIst;
}

// Issue priority at a segment of the pair loops.  Without PROG: priority 3
// once the count of siblings done with the step's pair loops (fin_now, read
// at the segment's start) reaches fin_thr.  With PROG: the wave adds its
// segment's n points to the workgroup's count and takes priority 3 while
// W x its own points started before the segment (own) is below the count
// read at the segment's start (unsigned arithmetic: the difference is exact
// while the waves stay within 2^31 points of each other), else priority 0.
template <bool PROG>
__device__ __forceinline__ void seg_priority(int *fin, int fin_now, int fin_thr, int W, unsigned own, int n)
{
    if constexpr (PROG) {
        if ((threadIdx.x & 63) == 0)
            __hip_atomic_fetch_add(fin, n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if ((int)((unsigned)__builtin_amdgcn_readfirstlane(fin_now) - (unsigned)W * own) > 0)
            __builtin_amdgcn_s_setprio(3);
        else
            __builtin_amdgcn_s_setprio(0);
    } else {
        if (__builtin_amdgcn_readfirstlane(fin_now) >= fin_thr) __builtin_amdgcn_s_setprio(3);
    }
}

// At least 6 waves per SIMD (at most 80 VGPRs): two 12-wave workgroups share
// a CU.  RECORDS: append every accepted step to the job's record window (the
// plain chunk keeps them; batched runs usually do not, and the
// instantiations have distinct names in profiles).  PROG: the issue-priority
// policy that balances the waves' progress through the pair loops (below),
// instead of the policy the task's layout bits select.
//
// Work distribution.  Without a queue (q.counter null) workgroup b runs task
// b for all of its steps.  With a queue the grid is the number of
// workgroups resident at once and each workgroup loops: it takes the next
// item u from the atomic counter, runs step slice k = u / ntask of task
// t = u % ntask (tasks in decreasing cost order, so heavy ones start first;
// slice k covers steps [k s2max / nslice, (k+1) s2max / nslice) of the task)
// and publishes the slice with a release store to done[t].  Slice k of a
// task waits (acquire) for done[t] == k, i.e. for the previous slice's
// replica states.  Items are taken in increasing order and a waiting
// workgroup only waits on an item taken earlier by a running workgroup,
// which itself waits only on earlier items, so the waits end.  Step slices
// let the last round of a launch spread over every resident workgroup
// instead of leaving CUs idle behind a partly filled round of whole tasks.
template <bool KAHAN, int MODE, bool RECORDS, bool PROG = false>
__global__ void __launch_bounds__(1024, 6) mc_kernel(const DevParams *__restrict__ prm, const DevJob *__restrict__ jobs,
                                                  const DevTask *__restrict__ tasks,
                                                  const int *__restrict__ step_base,
                                                  const int *__restrict__ pose_rep,
                                                  const float *__restrict__ pose_move, float *__restrict__ pose_out,
                                                  DevQueue q)
{
    extern __shared__ float lds[];
    __shared__ int s_item;
#ifdef GD_PHASE_PROFILE
    __shared__ unsigned long long s_ph[16], s_pw[16], s_pl[16];
    for (int i = threadIdx.x; i < 16; i += blockDim.x) s_ph[i] = s_pw[i] = 0;
    if (threadIdx.x == 0 && blockIdx.x < kWgMax) g_wg_time[0][blockIdx.x] = wall_clock64();
    __syncthreads();
#endif
    // waves of the item that have finished a step's pair loops; with PROG
    // the points the item's waves have started in them (modulo 2^32)
    __shared__ int s_fin;
    const int n_items = q.counter ? q.ntask * q.nslice : 0;
    for (int pass = 0;; ++pass) {
    int ti = blockIdx.x, kslice = 0, nslice = 1;
    if (q.counter) {
        if (threadIdx.x == 0) s_item = atomicAdd(q.counter, 1);
        __syncthreads();
        const int u = __builtin_amdgcn_readfirstlane(s_item);
        if (u >= n_items) break;
        ti = u % q.ntask;
        kslice = u / q.ntask;
        nslice = q.nslice;
        if (kslice > 0) {
            // the previous slice of this task: its replica states and records
            if (threadIdx.x == 0)
                while (__hip_atomic_load(q.done + ti, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < kslice)
                    __builtin_amdgcn_s_sleep(4);
            __syncthreads();
        }
    } else if (pass > 0) {
        break;
    }
    const DevTask task = tasks[ti];
    const int W = blockDim.x >> 6;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int SW = task.slot_waves;
    const int S = task.slices;
    const int sw = wave % SW;
    const int slice = wave / SW;
    const int NL = uref(jobs + task.job).natom;
    const int R = task.nrep;
    const int nslots = R * NL;
    const int slot = sw * 64 + lane;
    const bool active = slot < nslots;
    const int g = active ? slot / NL : 0;
    const int a = active ? slot - g * NL : 0;
    const int lanes = W * 64;

    float *red = lds;                          // [(S-1) * SW*64 * kRed], then reused:
    float *vals = lds;                         // [SW*64 * kVals]
    float *rsum = lds + (size_t)lanes * kVals; // [R][16] compensated term sums
    float *st = rsum + (size_t)R * 16;     // [R][16] movematrix[6], energy[10]
    float *prop = st + (size_t)R * 16;     // [R][8] proposed move, accept uniform
    int *ist = (int *)(prop + (size_t)R * 8); // [R][kIst] next_entry, is_accept, step, idx_tmp, accepted moves

    // this item's steps: slice kslice of the task's s2max steps
    const int sbeg = (int)(((long long)task.s2max * kslice) / nslice);
    const int send = (int)(((long long)task.s2max * (kslice + 1)) / nslice);
    const int s1 = task.s1 + (step_base ? *step_base : 0);

    if (MODE == MODE_MC) {
        const auto &J = uref(jobs + task.job);
        for (int i = threadIdx.x; i < R; i += blockDim.x) {
            const gd_ReplicaMC &rp = J.rep[task.rep0 + i];
            for (int k = 0; k < 6; ++k) st[i * 16 + k] = rp.movematrix[k];
            for (int k = 0; k < GD_MAXWEI; ++k) st[i * 16 + 6 + k] = rp.energy[k];
            ist[i * kIst + 0] = RECORDS && kslice > 0 ? J.next_entry[task.rep0 + i] : 0;
            ist[i * kIst + 1] = rp.is_accept;
            ist[i * kIst + 2] = rp.step;
            ist[i * kIst + 3] = rp.idx_tmp;
            ist[i * kIst + 4] = 0;
        }
    }
    if (threadIdx.x == 0) s_fin = 0;
    __syncthreads();

    // this lane's replica and atom
    int rid, pose = 0;
    if (MODE == MODE_MC) {
        rid = task.rep0 + g;
    } else {
        pose = task.rep0;
        rid = pose_rep[pose];
    }
    int idx_lig, tl;
    uint32_t idx_rep;
    float4 L3;
    {
        const auto &J = uref(jobs + task.job);
        idx_lig = J.rep[rid].idx_lig;
        idx_rep = (uint32_t)J.rep[rid].idx_rep;
        L3 = J.lig3[(size_t)idx_lig * NL + a];
        tl = J.lt[a];
    }
    const float lc = L3.w;
    const int prt = task.prt;
    int lo, hi, seg_b, seg_e;
    {
        const auto &tg = uref(prm).tg;
        const int np = tg.npoint;
        lo = (int)(((long long)np * slice) / S);
        hi = (int)(((long long)np * (slice + 1)) / S);
        // the segments that overlap this slice's points [lo, hi): the
        // segments cover the points in order, so they form one run [seg_b, seg_e)
        seg_b = 0;
        while (seg_b < tg.nseg && tg.seg[seg_b].z <= lo) ++seg_b;
        seg_e = seg_b;
        while (seg_e < tg.nseg && tg.seg[seg_e].y < hi) ++seg_e;
    }

    // Issue priority.  The SIMD's arbiter favours older waves, and the
    // waves of a persistent workgroup keep their ages for the whole launch:
    // on a SIMD shared by waves w, w + 4, w + 8 of a 12-wave workgroup, wave
    // w finishes its pair loops first and w + 8 last, while the step's
    // barrier waits for the last (measured on the stress shape: 4.9, 7.0 and
    // 9.4 million cycles for the three ranks, equal work).  A wave still in
    // its pair loops once fin_need of its siblings have finished them
    // raises its priority (checked per segment); each step starts at 0.
    // PROG instead: at each segment a wave adds the segment's points to the
    // count and takes priority 3 while it has started fewer points than the
    // workgroup's mean, else 0 (every wave walks a slice of the same size,
    // so equal counts mean equal shares of the step done).
    const int fin_need = (task.layout & LAYOUT_BOOST_NONE) ? (1 << 30)
                         : (task.layout & LAYOUT_BOOST_ANY) ? 1
                                                            : max(1, W / 2);
    for (int s2 = sbeg; s2 < send; ++s2) {
        const int step = s1 + s2;
#ifdef GD_PHASE_PROFILE
        unsigned long long ph_t = clock64();
#endif
        __builtin_amdgcn_s_setprio(0);
        const int fin_thr = W * (s2 - sbeg) + fin_need;
        const unsigned own0 = (unsigned)(s2 - sbeg) * (unsigned)(hi - lo);
        // the descriptors, read anew each step (see uref)
        const auto &J = uref(jobs + task.job);
        const auto &P = uref(prm);
        const auto &tg = P.tg;
        const PairConst sc{P.sc.lj_a, P.sc.lj_b, P.sc.pmf_a, P.sc.el_e0, P.sc.el_a, P.sc.el_b, P.sc.el_dc2};
        float m[6];
        float uacc = 0.0f;
        if (MODE == MODE_MC) {
            for (int k = 0; k < 6; ++k) m[k] = st[g * 16 + k];
            if (step != 0) {
                float u[7];
                if (J.kstream) { // explicit words, checked by the host to cover the step
                    const uint32_t *w = J.kstream + ((size_t)idx_rep * J.kstream_steps + (step - 1)) * 7;
                    for (int k = 0; k < 7; ++k) u[k] = (float)w[k] * 0x1p-31f;
                } else {
                    uint32_t b0 = (uint32_t)step, b1 = idx_rep, b2 = 0u, b3 = 0u;
                    uint32_t c0 = (uint32_t)step, c1 = idx_rep, c2 = 0u, c3 = 1u;
                    philox10(b0, b1, b2, b3, J.key0, J.key1);
                    philox10(c0, c1, c2, c3, J.key0, J.key1);
                    u[0] = uniform31(b0), u[1] = uniform31(b1), u[2] = uniform31(b2), u[3] = uniform31(b3);
                    u[4] = uniform31(c0), u[5] = uniform31(c1), u[6] = uniform31(c2);
                }
                for (int k = 0; k < 6; ++k) m[k] = move_update(J.scale[k], u[k], m[k]);
                uacc = u[6];
            }
        } else {
            for (int k = 0; k < 6; ++k) m[k] = pose_move[(size_t)pose * 6 + k];
        }
        if (a == 0 && slice == 0 && active) {
            for (int k = 0; k < 6; ++k) prop[g * 8 + k] = m[k];
            prop[g * 8 + 6] = uacc;
        }

        // Euler rotation of the cumulative angles
        // (kernel_cpp_montecarlo_soa.cpp:241-255)
        float s1a, c1a, s2a, c2a, s3a, c3a;
        sincos_cw(m[3], s1a, c1a);
        sincos_cw(m[4], s2a, c2a);
        sincos_cw(m[5], s3a, c3a);
        const float r00 = c1a * c2a, r01 = c1a * s2a * s3a - c3a * s1a, r02 = s1a * s3a + c1a * c3a * s2a;
        const float r10 = c2a * s1a, r11 = c1a * c3a + s1a * s2a * s3a, r12 = c3a * s1a * s2a - c1a * s3a;
        const float r20 = -s2a, r21 = c2a * s3a, r22 = c2a * c3a;
        const float tx = m[0] + J.center[0], ty = m[1] + J.center[1], tz = m[2] + J.center[2];
        Slot me;
        me.x = r00 * L3.x + r01 * L3.y + r02 * L3.z + tx;
        me.y = r10 * L3.x + r11 * L3.y + r12 * L3.z + ty;
        me.z = r20 * L3.x + r21 * L3.y + r22 * L3.z + tz;
        me.tl = tl;

        GD_PHASE(0); // proposal and rotation
        // KDE and MCS first: the rotation is then dead before the pair loops,
        // which need the registers.
        // KDE over the atom type's run of the type-sorted table
        // (energy_kde_v1.cpp:1-77): slice s takes points kr.x + s, kr.x + s + S, ...
        // MCS rows over the file-order atom set (energy_mcs_v3_ell.cpp):
        // worker a * S + s of the replica takes rows a * S + s, + S * N_L, ...,
        // so the rows of a small table still fall on every slice
        float ks = 0.f, mcs = 0.f;
        const int2 kr = J.krange[a];
        const bool kde0 = task.layout & LAYOUT_KDE_SLICE0, mcs0 = task.layout & LAYOUT_MCS_SLICE0;
        if (active && (!kde0 || slice == 0)) {
            const float kde2 = P.sc.kde2;
            const int KS = kde0 ? 1 : S;
            int k = kr.x + (kde0 ? 0 : slice);
            for (; k + KS < kr.y; k += 2 * KS) {
                const float4 K0 = J.kde[k], K1 = J.kde[k + KS];
                const float ax = me.x - K0.x, ay = me.y - K0.y, az = me.z - K0.z;
                const float bx = me.x - K1.x, by = me.y - K1.y, bz = me.z - K1.z;
                ks += fast_exp(kde2 * fmaf(ax, ax, fmaf(ay, ay, az * az)));
                ks += fast_exp(kde2 * fmaf(bx, bx, fmaf(by, by, bz * bz)));
            }
            if (k < kr.y) {
                const float4 K0 = J.kde[k];
                const float ax = me.x - K0.x, ay = me.y - K0.y, az = me.z - K0.z;
                ks += fast_exp(kde2 * fmaf(ax, ax, fmaf(ay, ay, az * az)));
            }
        }
        GD_PHASE(1); // KDE
        if (active && (!mcs0 || slice == 0)) {
            const int MS = mcs0 ? 1 : S;
            // the conformation's row offset is formed here each step from
            // idx_lig (opaque, so the product is not kept as a 64-bit value
            // across the step); below 2^31: n_lig x natom table entries
            int il = idx_lig;
            asm volatile("" : "+v"(il));
            const float4 *lig2 = J.lig2 + il * NL;
            for (int row = a * MS + (mcs0 ? 0 : slice); row < J.nrow; row += MS * NL) {
                const int rb = J.mrow[row], re = J.mrow[row + 1];
                float s = 0.f;
#pragma unroll 4
                for (int k2 = rb; k2 < re; ++k2) {
                    const float4 E = J.ment[k2];
                    const float4 X = lig2[__float_as_int(E.w)];
                    const float ax = r00 * X.x + r01 * X.y + r02 * X.z + tx;
                    const float ay = r10 * X.x + r11 * X.y + r12 * X.z + ty;
                    const float az = r20 * X.x + r21 * X.y + r22 * X.z + tz;
                    const float dx = ax - E.x, dy = ay - E.y, dz = az - E.z;
                    s += fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                }
                mcs += J.mtcc[row] * sqrtf(s / (float)(re - rb));
            }
        }

        GD_PHASE(2); // MCS
        // ---- pair terms over this slice's points ----
        float evdw = 0.f, eele = 0.f, epmf = 0.f, epsp = 0.f, ehdb = 0.f, ehpc = 0.f;
        float cvdw = 0.f, cele = 0.f, cpmf = 0.f, cpsp = 0.f, chdb = 0.f, chpc = 0.f;
        // far field: with the wave's atoms at radii [rho_in, rho_out], from
        // rung kf of a segment's ladder on points are at least far_d beyond
        // every atom, and up to ladder_in's rung ki they are at least far_d
        // inside every atom (a ligand that has left the pocket)
        // ladder centre: the point of the drift axis (1,1,1)/sqrt(3) nearest
        // the middle of the wave's atoms (their projections' mid-range);
        // beyond the last centre, centre 0 (the protein is then behind the
        // atoms, and the inner ladder bounds every point)
        const float ax = (me.x + me.y + me.z) * 0.57735026918962576f;
        const float kpos = rintf(0.5f * (wave_max(ax) + wave_min(ax)) * tg.inv_center_dr);
        const int kc = __builtin_amdgcn_readfirstlane(kpos > (float)(kCenters - 1) ? 0 : (int)fmaxf(kpos, 0.0f));
        const float cc = (float)kc * tg.center_dr * 0.57735026918962576f;
        const float4 *__restrict__ ptc = tg.pt + ((size_t)kc * tg.nconf + prt) * tg.npad;
        const float *__restrict__ pqc = tg.pq + ((size_t)kc * tg.nconf + prt) * tg.npad;
        const float *__restrict__ pak = tg.pw + (size_t)kc * tg.npad;
        const float ux = me.x - cc, uy = me.y - cc, uz = me.z - cc;
        const float r2 = fmaf(ux, ux, fmaf(uy, uy, uz * uz));
        // the atom relative to the ladder centre, as the point table holds the points
        const Slot mu{ux, uy, uz, me.tl};
        const float rho_out = sqrtf(wave_max(r2)), rho_in = sqrtf(wave_min(r2));
        // far and very far pairs (d >= far_d) take the expanded distance when
        // the wave lies within kExpandR of its ladder centre: (|u| + |q|) / d
        // then stays below (2 kExpandR + far_d) / far_d
        const bool expand = rho_out <= kExpandR;
        const int kf = __builtin_amdgcn_readfirstlane(max(0, (int)ceilf((rho_out + tg.far_d - kFarR0) * (1.0f / kFarDr))));
        // the same rungs for the mid field (mid_d <= far_d, so km <= kf and kmi >= ki)
        const int km = __builtin_amdgcn_readfirstlane(max(0, (int)ceilf((rho_out + tg.mid_d - kFarR0) * (1.0f / kFarDr))));
        const int kmi = __builtin_amdgcn_readfirstlane(
            min(kFarRungs - 1, (int)floorf((rho_in - tg.mid_d - kFarR0) * (1.0f / kFarDr))));
        const int ki = __builtin_amdgcn_readfirstlane(
            min(kFarRungs - 1, (int)floorf((rho_in - tg.far_d - kFarR0) * (1.0f / kFarDr))));
        // the same rungs for the very far field (vfar_d >= far_d, so kv >= kf
        // and kvi <= ki; vfar_d may be huge: clamp before converting)
        const int kv = __builtin_amdgcn_readfirstlane(
            (int)fminf(fmaxf(ceilf((rho_out + tg.vfar_d - kFarR0) * (1.0f / kFarDr)), 0.0f), (float)kFarRungs));
        const int kvi = __builtin_amdgcn_readfirstlane(
            (int)fmaxf(fminf(floorf((rho_in - tg.vfar_d - kFarR0) * (1.0f / kFarDr)), (float)(kFarRungs - 1)), -1.0f));
        // A wave whose atoms are all beyond the protein by vfar_d: from centre
        // 0, rung kvi of the inner ladder then holds every point of every
        // segment, so all pairs are very far and a lean loop walks the slice's
        // points without the per-segment splits or the PSP points (none of
        // which is within pmf0): the same arithmetic in the same order
        const bool lean = kc == 0 && kvi >= 0 && kFarR0 + (float)kvi * kFarDr >= tg.prot_r;
        // the points lie within prot_r of centre 0, so on the lean path every
        // pair is at least rho_in - prot_r apart: beyond lean_e_d and
        // lean_soft_d the softening takes fewer terms
        const float gap = rho_in - tg.prot_r;
        const int tier = __builtin_amdgcn_readfirstlane(gap >= tg.lean_soft_d ? 2 : (gap >= tg.lean_e_d ? 1 : 0));
        const cint *ladder = (const cint *)(tg.ladder + (size_t)kc * tg.nseg * kFarRungs);
        const cint *ladder_in = (const cint *)(tg.ladder_in + (size_t)kc * tg.nseg * kFarRungs);
        if (lean) {
            // the next segment's bounds and parameters load while this
            // segment's points run
            int4 sgn = seg_b < seg_e ? tg.seg[seg_b] : make_int4(0, 0, 0, 0);
            float4 pAn = *(const float4 *)(tg.pp + ((size_t)sgn.x * GD_MAXTP2 + me.tl) * PP_N);
            for (int si = seg_b; si < seg_e; ++si) {
                const int4 sg = sgn;
                const int fin_now = __hip_atomic_load(&s_fin, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                const float4 pA = pAn;
                if (si + 1 < seg_e) {
                    sgn = tg.seg[si + 1];
                    pAn = *(const float4 *)(tg.pp + ((size_t)sgn.x * GD_MAXTP2 + me.tl) * PP_N);
                }
                const int b = max(sg.y, lo), e = min(sg.z, hi);
                if (b >= e) continue;
                seg_priority<PROG>(&s_fin, fin_now, fin_thr, W, own0 + (unsigned)(b - lo), e - b);
                const PairParams q{pA.x, pA.y, pA.w, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                for (int pb = b; pb < e; pb += kBlock) {
                    const int pe = min(pb + kBlock, e);
                    Sums bs{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                    // lean: |q| <= prot_r and d >= |u| - prot_r >= vfar_d, so
                    // (|u| + |q|) / d <= (2 prot_r + vfar_d) / vfar_d
                    if (tier == 2)
                        pair_block_vfar<2, true>(ptc, pqc, pb, pe, mu, r2, q, sc, bs);
                    else if (tier == 1)
                        pair_block_vfar<1, true>(ptc, pqc, pb, pe, mu, r2, q, sc, bs);
                    else
                        pair_block_vfar<0, true>(ptc, pqc, pb, pe, mu, r2, q, sc, bs);
                    acc_add<KAHAN>(evdw, cvdw, bs.v);
                    acc_add<KAHAN>(eele, cele, bs.e);
                    acc_add<KAHAN>(eele, cele, -bs.ec);
                    acc_add<KAHAN>(epmf, cpmf, bs.p);
                }
            }
        } else {
        for (int si = seg_b; si < seg_e; ++si) {
            const int4 sg = tg.seg[si];
            // (the count is read here and tested after the run bounds, so its
            // LDS latency overlaps the ladder loads)
            const int fin_now = __hip_atomic_load(&s_fin, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            const int b = max(sg.y, lo), e = min(sg.z, hi);
            if (b >= e) continue;
            const int ie = ki >= 0 ? ladder_in[si * kFarRungs + ki] : sg.y;
            const int ob = kf < kFarRungs ? ladder[si * kFarRungs + kf] : sg.z;
            const int ive = kvi >= 0 ? ladder_in[si * kFarRungs + kvi] : sg.y;
            const int ovb = kv < kFarRungs ? ladder[si * kFarRungs + kv] : sg.z;
            const int imi = kmi >= 0 ? ladder_in[si * kFarRungs + kmi] : sg.y;
            const int om = km < kFarRungs ? ladder[si * kFarRungs + km] : sg.z;
            // very far [b, v1), far [v1, m1), mid [m1, n1), near [n1, n2),
            // mid [n2, mid), far [mid, v2), very far [v2, e); the inner and
            // outer runs cannot share a point (ive <= ie <= imi <= om <= ob
            // <= ovb), the clamps keep the order
            const int v1 = min(max(ive, b), e);
            const int m1 = min(max(ie, v1), e);
            const int n1 = min(max(imi, m1), e);
            const int n2 = min(max(om, n1), e);
            const int mid = min(max(ob, n2), e);
            const int v2 = min(max(ovb, mid), e);
            seg_priority<PROG>(&s_fin, fin_now, fin_thr, W, own0 + (unsigned)(b - lo), e - b);
            const float *P = tg.pp + ((size_t)sg.x * GD_MAXTP2 + me.tl) * PP_N;
            const float4 pA = *(const float4 *)P;
            const float4 pB = *(const float4 *)(P + 4);
            const float4 pC = *(const float4 *)(P + 8);
            const PairParams q{pA.x, pA.y, pA.w, pB.x, pB.y, pB.z, pB.w, pC.x, pC.y};
            const bool hdb = __any(q.hdb1 != 0.0f); // wave-uniform
            // plain FP32 sums over blocks of at most kBlock points, block sums
            // added with compensation: the electrostatic sum cancels by up to
            // 1e4 over a pose, which long running FP32 sums cannot hold
            for (int pb = n1; pb < n2; pb += kBlock) {
                const int pe = min(pb + kBlock, n2);
                Sums bs{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                if (sg.w) { // segment flag: hydrophobic-contact points
                    if (hdb)
                        pair_block<true, true>(ptc, pak, pb, pe, mu, q, sc, bs);
                    else
                        pair_block<true, false>(ptc, pak, pb, pe, mu, q, sc, bs);
                } else {
                    if (hdb)
                        pair_block<false, true>(ptc, pak, pb, pe, mu, q, sc, bs);
                    else
                        pair_block<false, false>(ptc, pak, pb, pe, mu, q, sc, bs);
                }
                acc_add<KAHAN>(evdw, cvdw, bs.v);
                acc_add<KAHAN>(eele, cele, bs.e);
                acc_add<KAHAN>(eele, cele, -bs.ec);
                acc_add<KAHAN>(epmf, cpmf, bs.p);
                acc_add<KAHAN>(ehdb, chdb, bs.h);
                acc_add<KAHAN>(ehpc, chpc, bs.q);
            }
            for (int part = 0; part < 6; ++part) {
                // mid [m1, n1), [n2, mid); far [v1, m1), [mid, v2); very
                // far [b, v1), [v2, e)
                const int fb = part == 0 ? m1 : part == 1 ? n2 : part == 2 ? v1 : part == 3 ? mid : part == 4 ? b : v2;
                const int fe = part == 0 ? n1 : part == 1 ? mid : part == 2 ? m1 : part == 3 ? v2 : part == 4 ? v1 : e;
                for (int pb = fb; pb < fe; pb += kBlock) {
                    const int pe = min(pb + kBlock, fe);
                    Sums bs{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                    if (part < 2)
                        pair_block<false, false, false>(ptc, pak, pb, pe, mu, q, sc, bs);
                    else if (part < 4) {
                        if (expand)
                            pair_block_far<true>(ptc, pqc, pb, pe, mu, r2, q, sc, bs);
                        else
                            pair_block_far<false>(ptc, pqc, pb, pe, mu, r2, q, sc, bs);
                    } else {
                        if (expand)
                            pair_block_vfar<0, true>(ptc, pqc, pb, pe, mu, r2, q, sc, bs);
                        else
                            pair_block_vfar<0, false>(ptc, pqc, pb, pe, mu, r2, q, sc, bs);
                    }
                    acc_add<KAHAN>(evdw, cvdw, bs.v);
                    acc_add<KAHAN>(eele, cele, bs.e);
                    acc_add<KAHAN>(eele, cele, -bs.ec);
                    acc_add<KAHAN>(epmf, cpmf, bs.p);
                }
            }
        }
#ifdef GD_PHASE_PROFILE
        if (lane == 0) s_pl[wave] = clock64() - ph_t;
#endif
        GD_PHASE(3); // pair loops
        // pocket-specific term: the job's class-2 points whose residue has a
        // PSP row add psp[t_l][residue] where d <= pmf0 (energy_prt_v1.cpp:107-109);
        // d is computed exactly as in the pair loop
        {
            const auto &J = uref(jobs + task.job);
            const auto &tg = uref(prm).tg;
            const float4 *__restrict__ ptc0 = tg.pt + (size_t)prt * tg.npad; // ladder centre 0: the pocket-centre order
            float ss = 0.f;
            for (int k = 0; k < J.npsp; ++k) {
                const int p = J.psp_pt[k];
                if (p < lo || p >= hi) continue; // the slice that owns the point adds it
                const float4 P4 = ptc0[p]; // centre 0: q' = -2 p
                const int tp = __float_as_int(tg.pa[p].y);
                const float dx = fmaf(0.5f, P4.x, me.x), dy = fmaf(0.5f, P4.y, me.y), dz = fmaf(0.5f, P4.z, me.z);
                const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                const float d = d2 * __builtin_amdgcn_rsqf(d2);
                const float pmf0 = tg.pp[((size_t)tp * GD_MAXTP2 + me.tl) * PP_N + PP_PMF0];
                if (d - pmf0 <= 0.0f) ss += J.psp_row[k * GD_MAXTP2 + me.tl];
            }
            acc_add<KAHAN>(epsp, cpsp, ss);
        }
        } // general path

        if (!PROG && (threadIdx.x & 63) == 0)
            __hip_atomic_fetch_add(&s_fin, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        GD_PHASE(4); // PSP points
        // ---- cross-slice sum ----
        if (S > 1 && slice > 0) {
            float *r = red + ((size_t)(slice - 1) * SW * 64 + slot) * kRed;
            r[0] = evdw;
            r[1] = eele;
            r[2] = epmf;
            r[3] = epsp;
            r[4] = ehdb;
            r[5] = ehpc;
            r[6] = cele; // the electrostatic sum travels with its compensation
            r[7] = ks;
            r[8] = mcs;
        }
        __syncthreads();
        if (slice == 0 && active) {
            for (int s = 1; s < S; ++s) {
                const float *r = red + ((size_t)(s - 1) * SW * 64 + slot) * kRed;
                acc_add<KAHAN>(evdw, cvdw, r[0]);
                acc_add<KAHAN>(eele, cele, r[1]);
                acc_add<KAHAN>(eele, cele, -r[6]);
                acc_add<KAHAN>(epmf, cpmf, r[2]);
                acc_add<KAHAN>(epsp, cpsp, r[3]);
                acc_add<KAHAN>(ehdb, chdb, r[4]);
                acc_add<KAHAN>(ehpc, chpc, r[5]);
                ks += r[7];
                mcs += r[8];
            }
        }
        __syncthreads(); // the partials region is reused for the per-slot terms
#ifdef GD_PHASE_PROFILE
        if (threadIdx.x == 0) {
            unsigned long long mx = 0;
            for (int w = 0; w < W; ++w) {
                mx = s_pl[w] > mx ? s_pl[w] : mx;
                s_pw[w] += s_pl[w];
            }
            s_ph[9] += mx * W;
            s_ph[10] += 1;
        }
#endif
        GD_PHASE(5); // cross-slice sum and its barriers
        if (slice == 0 && active) {
            const auto &sp = uref(prm).sc;
            // hydrophobic restraint of the atom (energy_prt_v1.cpp:154-155)
            const float hpc2 = (ehpc - sp.hpl0[tl]) / sp.hpl1[tl];
            const float hpc_atom = 0.5f * hpc2 * hpc2 - sp.hpl2[tl];
            const float kde_atom = kr.y > kr.x ? ks / (float)(kr.y - kr.x) : 0.0f;
            float *v = vals + (size_t)slot * kVals;
            v[0] = evdw * sp.inv_lj0;
            // electrostatics lc / el1 (eele - cele) as hi + lo: the atoms'
            // values are up to 1e6 times their sum, so the product is split
            // exactly (FMA remainders) and both parts are summed
            {
                const float k = lc * sp.inv_el1;
                const float kr = fmaf(lc, sp.inv_el1, -k);
                const float hi = k * eele;
                v[1] = hi;
                v[8] = fmaf(k, eele, -hi) + fmaf(kr, eele, -k * cele);
            }
            v[2] = epmf;
            v[3] = epsp;
            v[4] = ehdb;
            v[5] = hpc_atom;
            v[6] = kde_atom;
            v[7] = mcs;
        }
        __syncthreads();

        GD_PHASE(6); // per-slot terms and barrier
        // ---- per-replica sums: thread (g, k) adds term k over the atoms ----
        for (int t = threadIdx.x; t < R * kVals; t += blockDim.x) {
            const int gg = t / kVals, k = t - gg * kVals;
            float s = 0.f, c = 0.f;
            const float *v = vals + (size_t)gg * NL * kVals + k;
            for (int i = 0; i < NL; ++i) acc_add<KAHAN>(s, c, v[(size_t)i * kVals]);
            rsum[gg * 16 + k] = s - c;
        }
        __syncthreads();

        GD_PHASE(7); // per-replica sums and barrier
        // ---- normalise, total, accept (kernel_cpp_montecarlo_soa.cpp:333-390) ----
        for (int gg = threadIdx.x; gg < R; gg += blockDim.x) {
            const auto &J = uref(jobs + task.job);
            const auto &sc = uref(prm).sc;
            float e[GD_MAXWEI];
            const float *pm = prop + gg * 8;
            for (int k = 0; k < 6; ++k) e[k] = rsum[gg * 16 + k];
            e[GD_E_ELE] += rsum[gg * 16 + 8];
            e[GD_E_HDB] = e[GD_E_HDB] * (-1.0f / sqrtf(2.0f * GD_PI));
            e[GD_E_KDE] = rsum[gg * 16 + 6] * sc.kde3inv;
            e[GD_E_LHM] = logf(rsum[gg * 16 + 7] / (float)J.nrow);
            // |centre + m - pocket|: the reference adds m to the ~30 A centre
            // first and loses the small offset to rounding; subtracting the
            // two centres first is exact when they coincide (they do: both
            // are the job's native ligand centroid)
            float dst = 0.f;
            for (int k = 0; k < 3; ++k) {
                const float dd = (J.center[k] - J.pocket[k]) + pm[k];
                dst += dd * dd;
            }
            e[GD_E_DST] = sqrtf(dst);
            for (int k = 0; k < 7; ++k) e[k] = e[k] / (float)NL;
            for (int k = 0; k < GD_MAXWEI - 1; ++k) e[k] = sc.a_para[k] * e[k] + sc.b_para[k];
            float etotal = 0.f;
            for (int k = 0; k < GD_MAXWEI - 1; ++k) etotal += sc.w[k] * e[k];
            e[GD_E_TOTAL] = etotal;

            if (MODE == MODE_ENERGY) {
                for (int k = 0; k < GD_MAXWEI; ++k) pose_out[(size_t)pose * GD_MAXWEI + k] = e[k];
            } else {
                const int stp = s1 + s2;
                float *cur = st + gg * 16;
                int acc;
                if (stp != 0) {
                    const float delta = etotal - cur[6 + GD_E_TOTAL];
                    acc = pm[6] < expf(delta * J.mbeta[ist[gg * kIst + 3]]);
                } else {
                    acc = 1;
                }
                ist[gg * kIst + 1] = acc;
                if (acc && stp != 0) ist[gg * kIst + 4] += 1;
                if (acc) {
                    for (int k = 0; k < 6; ++k) cur[k] = pm[k];
                    for (int k = 0; k < GD_MAXWEI; ++k) cur[6 + k] = e[k];
                    ist[gg * kIst + 2] = stp;
                    if (RECORDS) {
                        const int r = task.rep0 + gg;
                        const int ne = ist[gg * kIst + 0];
                        gd_ReplicaMC *rc = J.rec + (size_t)r * J.rec_cap + ne;
                        const gd_ReplicaMC &src = J.rep[r];
                        rc->idx_rep = src.idx_rep;
                        rc->idx_lig = src.idx_lig;
                        rc->idx_prt = src.idx_prt;
                        rc->idx_tmp = ist[gg * kIst + 3];
                        for (int k = 0; k < 6; ++k) rc->movematrix[k] = cur[k];
                        for (int k = 0; k < GD_MAXWEI; ++k) rc->energy[k] = cur[6 + k];
                        rc->step = stp;
                        rc->is_accept = src.is_accept;
                        ist[gg * kIst + 0] = ne + 1;
                    }
                }
            }
        }
        __syncthreads();
        GD_PHASE(8); // accept and barrier
    }

    if (MODE == MODE_MC) {
        const auto &J = uref(jobs + task.job);
        for (int i = threadIdx.x; i < R; i += blockDim.x) {
            gd_ReplicaMC &rp = J.rep[task.rep0 + i];
            for (int k = 0; k < 6; ++k) rp.movematrix[k] = st[i * 16 + k];
            for (int k = 0; k < GD_MAXWEI; ++k) rp.energy[k] = st[i * 16 + 6 + k];
            rp.step = ist[i * kIst + 2];
            // the reference sets is_accept once, after the chunk's last step
            // (kernel_cpp_montecarlo_soa.cpp:415), and the record entries of
            // the chunk carry the value from before it: earlier step slices
            // leave it in place for the later slices' records
            if (kslice == nslice - 1) rp.is_accept = ist[i * kIst + 1];
            if (RECORDS) J.next_entry[task.rep0 + i] = ist[i * kIst + 0];
            J.mc_acc[task.rep0 + i] += ist[i * kIst + 4]; // this workgroup owns the replica
        }
    }
    // every wave's stores of the replica states are done before the slice
    // is published and before the LDS state is reused by the next item
    __syncthreads();
    if (q.counter && threadIdx.x == 0)
        __hip_atomic_store(q.done + ti, kslice + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } // item loop
#ifdef GD_PHASE_PROFILE
    __syncthreads();
    if (threadIdx.x < 16) {
        atomicAdd(&g_phase[threadIdx.x], s_ph[threadIdx.x]);
        atomicAdd(&g_phase_wave[threadIdx.x], s_pw[threadIdx.x]);
    }
    if (threadIdx.x == 0 && blockIdx.x < kWgMax) g_wg_time[1][blockIdx.x] = wall_clock64();
#endif
}

// Temperature replica exchange of ladder L (ligand conformation l, protein
// conformation p, L = p n_lig + l) at round rnd, see gd_exchange in
// geauxdock_amd.h.  Returns the accepted swaps.
__device__ int exchange_ladder(const DevJob &J, int L, int rnd)
{
    const int n_lig = J.n_lig, n_tmp = J.n_tmp;
    const int p = L / n_lig, l = L - p * n_lig;
    int holder[64];
    for (int t = 0; t < n_tmp; ++t) {
        const int r = n_tmp * n_lig * p + n_lig * t + l;
        holder[J.rep[r].idx_tmp] = r;
    }
    int accepted = 0;
    for (int t = rnd & 1; t + 1 < n_tmp; t += 2) {
        const int ra = holder[t], rb = holder[t + 1];
        const double beta_a = -(double)J.mbeta[t];
        const double beta_b = -(double)J.mbeta[t + 1];
        const double delta = (beta_a - beta_b) * ((double)J.rep[ra].energy[GD_E_TOTAL] - (double)J.rep[rb].energy[GD_E_TOTAL]);
        const uint32_t j = (uint32_t)(t >> 1);
        uint32_t c0 = (uint32_t)rnd, c1 = (uint32_t)L, c2 = 1u, c3 = j >> 2;
        philox10(c0, c1, c2, c3, J.key0, J.key1);
        const uint32_t wsel = (j & 3) == 0 ? c0 : ((j & 3) == 1 ? c1 : ((j & 3) == 2 ? c2 : c3));
        const double u = (double)(wsel >> 1) * (1.0 / 2147483648.0);
        if (u < exp(delta)) {
            J.rep[ra].idx_tmp = t + 1;
            J.rep[rb].idx_tmp = t;
            ++J.ex_acc[ra]; // the ladder's replicas belong to this thread alone
            ++J.ex_acc[rb];
            ++accepted;
        }
    }
    return accepted;
}

// One thread per ladder of one job, over a ladder range.
__global__ void exchange_kernel(const DevJob *__restrict__ job, int ladder0, int n_ladder, int round,
                                const int *__restrict__ round_base, int *__restrict__ accept_count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_ladder) return;
    const int rnd = round + (round_base ? *round_base : 0);
    const int accepted = exchange_ladder(*job, ladder0 + i, rnd);
    if (accept_count && accepted) atomicAdd(accept_count, accepted);
}

// One thread per entry {job, ladder, round offset} of a ladder list spanning
// several jobs.
__global__ void exchange_list_kernel(const DevJob *__restrict__ jobs, const int4 *__restrict__ ladders, int n,
                                     const int *__restrict__ round_base)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int4 e = ladders[i];
    (void)exchange_ladder(jobs[e.x], e.y, e.z + (round_base ? *round_base : 0));
}

namespace {

template <bool KAHAN, bool RECORDS, bool PROG = false>
hipError_t launch_mc_impl(const DevParams *prm, const DevJob *jobs, const DevTask *tasks, const DevQueue &q, int grid,
                          int waves, int max_rep, hipStream_t stream, const int *step_base)
{
    const size_t lds = mc_lds_bytes(waves, max_rep);
    hipLaunchKernelGGL((mc_kernel<KAHAN, MODE_MC, RECORDS, PROG>), dim3(grid), dim3(waves * 64), lds, stream, prm,
                       jobs, tasks, step_base, nullptr, nullptr, nullptr, q);
    return hipGetLastError();
}

// the progress policy is built for the batched form's kernel (compensated
// sums, no records); the other forms keep the layout bits' policy
hipError_t launch_mc_any(const DevParams *prm, const DevJob *jobs, const DevTask *tasks, const DevQueue &q, int grid,
                         int waves, int max_rep, bool compensated, bool keep_records, bool progress,
                         hipStream_t stream, const int *step_base)
{
    if (grid <= 0) return hipSuccess;
    if (compensated && !keep_records && progress)
        return launch_mc_impl<true, false, true>(prm, jobs, tasks, q, grid, waves, max_rep, stream, step_base);
    if (compensated && keep_records)
        return launch_mc_impl<true, true>(prm, jobs, tasks, q, grid, waves, max_rep, stream, step_base);
    if (compensated)
        return launch_mc_impl<true, false>(prm, jobs, tasks, q, grid, waves, max_rep, stream, step_base);
    if (keep_records)
        return launch_mc_impl<false, true>(prm, jobs, tasks, q, grid, waves, max_rep, stream, step_base);
    return launch_mc_impl<false, false>(prm, jobs, tasks, q, grid, waves, max_rep, stream, step_base);
}

__global__ void queue_reset_kernel(int *counter, int *done, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *counter = 0;
    if (i < n) done[i] = 0;
}

} // namespace

hipError_t launch_mc(const DevParams *prm, const DevJob *jobs, const DevTask *tasks, int ntask, int waves,
                     int max_rep, bool compensated, bool keep_records, hipStream_t stream, const int *step_base)
{
    const DevQueue none{nullptr, nullptr, ntask, 1};
    return launch_mc_any(prm, jobs, tasks, none, ntask, waves, max_rep, compensated, keep_records, false, stream,
                         step_base);
}

hipError_t launch_mc_queue(const DevParams *prm, const DevJob *jobs, const DevTask *tasks, const DevQueue &q,
                           int grid, int waves, int max_rep, bool compensated, bool keep_records, bool progress,
                           hipStream_t stream, const int *step_base)
{
    if (!q.counter || !q.done || q.ntask <= 0 || q.nslice <= 0) return hipErrorInvalidValue;
    return launch_mc_any(prm, jobs, tasks, q, std::min(grid, q.ntask * q.nslice), waves, max_rep, compensated,
                         keep_records, progress, stream, step_base);
}

hipError_t launch_queue_reset(const DevQueue &q, hipStream_t stream)
{
    const int n = std::max(q.ntask, 1);
    hipLaunchKernelGGL(queue_reset_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, q.counter, q.done, q.ntask);
    return hipGetLastError();
}

int mc_blocks_per_cu(int waves, int max_rep)
{
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, mc_kernel<true, MODE_MC, false>, waves * 64,
                                                     mc_lds_bytes(waves, max_rep)) != hipSuccess ||
        blocks <= 0)
        blocks = std::max(1, 24 / waves); // 6 waves per SIMD, four SIMDs
    return blocks;
}

hipError_t launch_energy(const DevParams *prm, const DevJob *jobs, const DevTask *tasks, int ntask, int waves,
                         bool compensated, const int *pose_rep, const float *pose_move, float *pose_out,
                         hipStream_t stream)
{
    if (ntask <= 0) return hipSuccess;
    const size_t lds = mc_lds_bytes(waves, 1);
    const dim3 grid(ntask), block(waves * 64);
    if (compensated)
        hipLaunchKernelGGL((mc_kernel<true, MODE_ENERGY, false>), grid, block, lds, stream, prm, jobs, tasks,
                           nullptr, pose_rep, pose_move, pose_out, DevQueue{nullptr, nullptr, ntask, 1});
    else
        hipLaunchKernelGGL((mc_kernel<false, MODE_ENERGY, false>), grid, block, lds, stream, prm, jobs, tasks,
                           nullptr, pose_rep, pose_move, pose_out, DevQueue{nullptr, nullptr, ntask, 1});
    return hipGetLastError();
}

hipError_t launch_exchange(const DevJob *job_dev, int ladder0, int n_ladder, int n_lig, int n_tmp, int round,
                           int *accept_count, hipStream_t stream, const int *round_base)
{
    (void)n_lig;
    if (n_ladder <= 0 || n_tmp < 2) return hipSuccess;
    const int bs = 64;
    hipLaunchKernelGGL(exchange_kernel, dim3((n_ladder + bs - 1) / bs), dim3(bs), 0, stream, job_dev, ladder0, n_ladder,
                       round, round_base, accept_count);
    return hipGetLastError();
}

hipError_t launch_exchange_list(const DevJob *jobs, const int4 *ladders, int n, hipStream_t stream,
                                const int *round_base)
{
    if (n <= 0) return hipSuccess;
    const int bs = 64;
    hipLaunchKernelGGL(exchange_list_kernel, dim3((n + bs - 1) / bs), dim3(bs), 0, stream, jobs, ladders, n,
                       round_base);
    return hipGetLastError();
}

} // namespace gd

#ifdef GD_PHASE_PROFILE
// Phase profile counters (out[0..15] phases, out[16..31] per wave index);
// reset clears them.  Exported by the phase-profile build only.
extern "C" int gd_phase_profile(unsigned long long *out, int reset)
{
    if (!out) return -1;
    if (hipMemcpyFromSymbol(out, HIP_SYMBOL(gd::g_phase), sizeof(unsigned long long) * 16) != hipSuccess) return -3;
    if (hipMemcpyFromSymbol(out + 16, HIP_SYMBOL(gd::g_phase_wave), sizeof(unsigned long long) * 16) != hipSuccess)
        return -3;
    if (reset) {
        static const unsigned long long z[16] = {};
        if (hipMemcpyToSymbol(HIP_SYMBOL(gd::g_phase), z, sizeof z) != hipSuccess) return -3;
        if (hipMemcpyToSymbol(HIP_SYMBOL(gd::g_phase_wave), z, sizeof z) != hipSuccess) return -3;
    }
    return 0;
}

// Wall-clock start (t[0..n)) and exit (t[n..2n)) of workgroups 0..n-1 of the
// last launch, in wall-clock ticks (hipDeviceAttributeWallClockRate).
// Exported by the phase-profile build only.
extern "C" int gd_phase_profile_wg(unsigned long long *t, int n)
{
    if (!t || n <= 0 || n > gd::kWgMax) return -1;
    const size_t bytes = sizeof(unsigned long long) * (size_t)n;
    if (hipMemcpyFromSymbol(t, HIP_SYMBOL(gd::g_wg_time), bytes, 0) != hipSuccess) return -3;
    if (hipMemcpyFromSymbol(t + n, HIP_SYMBOL(gd::g_wg_time), bytes, sizeof(unsigned long long) * gd::kWgMax) !=
        hipSuccess)
        return -3;
    return 0;
}
#endif
