// Device-side data layout shared by the kernels and the host code that
// uploads and launches them.
//
// HBM layout (one resident target, any number of jobs):
//   target  pt   float4[kCenters][nconf][npad]  {-2 (p - C_k), ele} of every
//                                     effective point p (frame centred on the
//                                     pocket centre; C_k the ladder centre of
//                                     copy k, C_0 = 0), re-sorted by (point
//                                     type, hydrophobic flag) so a run of
//                                     points (a segment) shares one set of
//                                     type-pair parameters and terms; copy k
//                                     orders each segment by distance from C_k
//           pq   float[kCenters][nconf][npad] |p - C_k|^2
//           pa   float2[kCenters][npad] {hpc weight = (cdc ? hpp : 0) / 2, point type}
//           pw   float[kCenters][npad] the hpc weights of pa alone
//           seg  int4[nseg]           {point type, begin, end, hydrophobic flag}
//           ladder int[kCenters][nseg][kFarRungs] first point of each segment at
//                                     or beyond each distance rung from the
//                                     centre (far field)
//           ladder_in int[kCenters][nseg][kFarRungs] end of each segment's leading
//                                     points whose largest distance is within the rung
//           pp   float[31][24][PP_N]  type-pair parameters, see PairParam
//   job     lig3 float4[nlig][natom]  {x3, y3, z3, c3} type-sorted atoms
//           lig2 float4[nlig][natom]  {x2, y2, z2, 0} file-order atoms (MCS)
//           lt   int[natom]           t3
//           krange int2[natom]        same-type range in the sorted KDE table
//           kde  float4[nkde]         {x, y, z, 0}
//           mrow int[nrow+1], ment float4[nent] {x, y, z, file atom index}, mtcc
//           psp_pt int[npsp], psp_row float[npsp][24]: the job's class-2 points
//                                     whose residue has PSP entries (the PSP
//                                     table comes with the ligand's .ff file)
//           rep  gd_ReplicaMC[n_rep], rec gd_ReplicaMC[n_rep][rec_cap], next_entry
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "gd/gd_types.h"

namespace gd {

// per (point type, ligand type) parameters, PP_N floats.  The contact and
// H-bond exponents are stored as quadratics in d so a pair needs two FMAs
// for each: log2(e) (6 - d/2)(d - pmf0) = pmf_a d^2 + PMFB d + PMFC and
// -log2(e) ((d - hdb0) hdb1)^2 / 2 = HK2 d^2 + HK1 d + HK0.
enum PairParam {
    PP_P1A = 0,  // LJ repulsive prefactor p1a, times c (DevScalars)
    PP_P2A = 1,  // LJ attractive prefactor p2a, times c
    PP_PMF0 = 2, // contact distance
    PP_PMF1 = 3, // contact depth
    PP_PMFB = 4, // log2(e) (6 + pmf0 / 2)
    PP_PMFC = 5, // -6 log2(e) pmf0
    PP_HK2 = 6,  // -log2(e) hdb1^2 / 2
    PP_HK1 = 7,  // log2(e) hdb1^2 hdb0
    PP_HK0 = 8,  // -log2(e) hdb1^2 hdb0^2 / 2
    PP_HDB1 = 9, // H-bond inverse width, 0 where the pair has no H-bond term
    PP_HDB0 = 10, // H-bond distance
    PP_N = 12
};

struct DevScalars {
    // LJ sums are kept times c = lj0 (1 where lj0 = 0); p1a and p2a in the
    // pair table carry the factor: pp4 = 1 + X (lj_a + lj_b d2), X = c pp1
    float lj_a, lj_b, inv_lj0;
    float inv_el1, kde2, kde3inv;
    // electrostatics inside d < 1 / el1, scaled by el1 (the sum is scaled
    // back once per atom): el1 (el0 + a1 s^2 + b1 s^3), s = el1 d, is
    // el_e0 + el_a d^2 + el_b d^3
    float el_e0, el_a, el_b, el_dc2;
    float pmf_a; // -log2(e) / 2
    float hpl0[GD_MAXTP2], hpl1[GD_MAXTP2], hpl2[GD_MAXTP2];
    float a_para[GD_MAXWEI], b_para[GD_MAXWEI], w[GD_MAXWEI];
};

// Ladder centres: the far-field ladders are built around kCenters points of
// the (1,1,1) axis through the pocket centre, DevTarget::center_dr apart;
// the reference's moves add non-negative increments to all three
// translation components, so a drifting ligand stays near that axis.
constexpr int kCenters = 16;

// Far-field radius ladder: rung k stands for radius kFarR0 + k * kFarDr.
constexpr int kFarRungs = 128;
constexpr float kFarR0 = 8.0f;
constexpr float kFarDr = 0.5f;

struct DevTarget {
    // pt, pa and the ladders come in kCenters copies: copy k orders each
    // segment's points by their distance from ladder centre k (copy 0: the
    // pocket centre); the segments themselves are the same in every copy
    // copy k holds q' = -2 (p - C_k): the direct difference x - p is
    // fma(1/2, q', x - C_k) exactly, and |x - p|^2 = |u|^2 + |q|^2 + u.q'
    // for u = x - C_k (see pq)
    const float4 *pt;  // [kCenters][nconf][npad] {q'x, q'y, q'z, ele}
    const float *pq;   // [kCenters][nconf][npad] |p - C_k|^2
    const float2 *pa;  // [kCenters][npad] {hydrophobic weight / 2, point type (int bits)}
    const float *pw;   // [kCenters][npad] hydrophobic weight / 2 alone, contiguous for wide scalar loads
    const float *pp;   // [31][24][PP_N] type-pair parameters
    const int4 *seg;   // [nseg] {point type, begin, end, hydrophobic flag}
    const int *ladder; // [kCenters][nseg][kFarRungs] first point of the segment at distance >= rung from the centre
    const int *ladder_in; // [kCenters][nseg][kFarRungs] end of the segment's leading points with largest distance <= rung
    float mid_d;       // distance beyond which the hydrophobic, core and H-bond terms vanish (mid_d <= far_d)
    float far_d;       // distance beyond which only LJ and 1/d electrostatics vary
    float vfar_d;      // distance beyond which 1 / (1 + e) of the soft-core LJ is 1 - e + e^2 to float rounding
    float lean_e_d;    // distance beyond which it is 1 - e (e <= 2^-12.5)
    float lean_soft_d; // distance beyond which it is 1 (e <= 2^-25)
    // spacing of the ladder centres: they reach the protein's radius plus
    // the very far distance and a ligand's extent; a wave beyond the last
    // takes centre 0, whose inner ladder then puts every point far
    float center_dr, inv_center_dr;
    float prot_r; // largest distance of a point from the pocket centre over the conformations
    int npoint, npad, nseg, nconf;
};

// The target's tables and scalars as the Monte Carlo kernel reads them: one
// copy in device memory per uploaded target.  The kernel reads its fields
// with scalar loads where it needs them (instead of receiving them as
// by-value arguments the compiler keeps in registers for the whole launch).
struct DevParams {
    DevTarget tg;
    DevScalars sc;
};

struct DevJob {
    const float4 *lig3;
    const float4 *lig2;
    const int *lt;
    const int2 *krange;
    const float4 *kde;
    const int *mrow;
    const float4 *ment;
    const float *mtcc;
    const float *mbeta;
    const int *psp_pt;     // [npsp] device point index of each PSP point
    const float *psp_row;  // [npsp][24] PSP value per ligand type
    gd_ReplicaMC *rep;
    gd_ReplicaMC *rec;
    int *next_entry;
    int natom, nrow, rec_cap, n_rep;
    int n_lig, n_prt, n_tmp, npsp;
    float center[3]; // ligand centre, target frame
    float pocket[3]; // the job's pocket centre, target frame
    float scale[6];
    uint32_t key0, key1;
    // explicit move/accept words k (u = k / RAND_MAX) for steps 1..kstream_steps,
    // [idx_rep][step - 1][7]; null: the keyed Philox stream
    const uint32_t *kstream;
    int kstream_steps;
    // McLog counters per replica since upload: accepted MC moves (steps >= 1)
    // and accepted temperature swaps
    int *mc_acc;
    int *ex_acc;
};

// A task is the set of replicas one workgroup advances: replicas
// [rep0, rep0 + nrep) of job `job`, all in protein conformation `prt`,
// stepping s1 .. s1 + s2max - 1.  In energy mode rep0 indexes the pose list.
struct DevTask {
    int job, rep0, nrep, prt;
    int s1, s2max, slices, slot_waves;
    int layout; // TaskLayout bits
};

// Where the KDE and MCS sums of a task run (autotuner candidates; the
// results differ only by summation order).  Spread: slice s of a replica's
// workers takes KDE points kr.x + s, kr.x + s + S, ... and MCS rows
// a S + s, a S + s + S N_L, ... (every slice shares the gathers); slice 0:
// the first point slice alone takes the atom's whole KDE run and rows
// a, a + N_L, ... (the other slices skip them).
enum TaskLayout {
    LAYOUT_KDE_SLICE0 = 1,
    LAYOUT_MCS_SLICE0 = 2,
    LAYOUT_BOOST_ANY = 4,
    LAYOUT_BOOST_NONE = 8,
    LAYOUT_BOOST_PROGRESS = 16
};
// Issue priority of the pair loops (also autotuner candidates; the results do
// not depend on it): a wave still in its pair loops raises its priority once
// half of the workgroup's waves have finished theirs (default), once any has
// (LAYOUT_BOOST_ANY), or never (LAYOUT_BOOST_NONE); or, in a launch whose
// tasks all carry LAYOUT_BOOST_PROGRESS, a wave takes priority while it has
// started fewer of its points than the workgroup's mean (a kernel of its own).
constexpr int LAYOUT_ENERGY_BITS = LAYOUT_KDE_SLICE0 | LAYOUT_MCS_SLICE0;
constexpr int LAYOUT_BOOST_BITS = LAYOUT_BOOST_ANY | LAYOUT_BOOST_NONE | LAYOUT_BOOST_PROGRESS;

enum KernelMode { MODE_MC = 0, MODE_ENERGY = 1 };

// Work queue of a persistent launch (see mc_kernel): items u = k * ntask + t
// run step slice k of task t.  counter (next item) and done[ntask] (slices
// finished per task) start at zero; a null counter launches one workgroup
// per task instead.
struct DevQueue {
    int *counter;
    int *done;
    int ntask, nslice;
};

// LDS bytes a workgroup of `waves` waves needs for `nrep` replicas
size_t mc_lds_bytes(int waves, int nrep);

// Launch the fused Monte Carlo kernel over `ntask` tasks, one workgroup of
// `waves` waves each (every task's slot_waves * slices == waves).
// step_base (device int, may be null) is added to every task's s1, so a
// captured graph can advance the step counter without new kernel arguments.
hipError_t launch_mc(const DevParams *prm, const DevJob *jobs, const DevTask *tasks, int ntask, int waves,
                     int max_rep, bool compensated, bool keep_records, hipStream_t stream, const int *step_base);

// Persistent form: `grid` workgroups walk the queue q (q.ntask tasks in
// q.nslice step slices each).  The queue's counter and done[] must be zero
// at launch (launch_queue_reset clears them on the stream, with a kernel, so
// the sequence can be captured in a graph).  progress: the kernel built with
// the progress-balancing issue-priority policy (LAYOUT_BOOST_PROGRESS; used
// with compensated sums and no records, else ignored).
hipError_t launch_mc_queue(const DevParams *prm, const DevJob *jobs, const DevTask *tasks, const DevQueue &q,
                           int grid, int waves, int max_rep, bool compensated, bool keep_records, bool progress,
                           hipStream_t stream, const int *step_base);
hipError_t launch_queue_reset(const DevQueue &q, hipStream_t stream);

// Workgroups of the Monte Carlo kernel resident on one CU for this shape
// (occupancy of the compensated batched instantiation; the others match).
int mc_blocks_per_cu(int waves, int max_rep);

// Energy mode: poses i = task.rep0 (nrep == 1), replica pose_rep[i], move pose_move[6i..]
hipError_t launch_energy(const DevParams *prm, const DevJob *jobs, const DevTask *tasks, int ntask, int waves,
                         bool compensated, const int *pose_rep, const float *pose_move, float *pose_out,
                         hipStream_t stream);

// One exchange round over ladders ladder0 .. ladder0 + n_ladder - 1 of a job
// (ladder p * n_lig + l: ligand conformation l, protein conformation p).
// round_base (device int, may be null) is added to `round`.
hipError_t launch_exchange(const DevJob *job_dev, int ladder0, int n_ladder, int n_lig, int n_tmp, int round,
                           int *accept_count, hipStream_t stream, const int *round_base);

// One exchange round over the ladders of several jobs in one launch: entry i
// of `ladders` is {job, ladder, round offset, 0}: ladder `ladder` of jobs[job]
// at round `round offset` + *round_base.
hipError_t launch_exchange_list(const DevJob *jobs, const int4 *ladders, int n, hipStream_t stream,
                                const int *round_base);

} // namespace gd
