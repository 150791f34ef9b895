/*
 * gd_types.h -- plain-C data model of the replica Monte Carlo docking path.
 *
 * The structures keep the field names and meanings of the reference data
 * model (src/common/geauxdock.h, src/common/geauxdock_soa.h) but every
 * capacity is a run-time size: arrays are pointers plus counts instead of
 * the fixed MAXLIG/MAXPRO/MAX_REP arrays of the reference.  Only the energy
 * parameter tables keep fixed shapes, because their shapes are set by the
 * atom/point type tables (24 ligand SYBYL types x 31 protein point types),
 * not by the size of an input.
 *
 * All structures are bitwise copyable and contain no C++ types, so they can
 * cross the C ABI, ctypes and the HIP runtime unchanged.
 */
#ifndef GD_TYPES_H
#define GD_TYPES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* type-table shapes (src/common/size.h:63-83) */
#define GD_MAXTP1 31 /* protein effective point types (30 used) */
#define GD_MAXTP2 24 /* ligand SYBYL atom types */
#define GD_MAXTP3 50 /* electrostatic table: point types 0..29, residues 30..49 */
#define GD_MAXTP4 20 /* residue types */
#define GD_MAXWEI 10 /* energy terms: 9 terms + total */
#define GD_BADKDE 50 /* code of a ligand type outside the 24-type table */

/* constants of the energy function (src/common/size.h:99,111,117,120) */
#define GD_PI 3.1415926535f
#define GD_MCS_INVALID_COORD 0.01234f
#define GD_STEPS_PER_DUMP 100

/* energy term indices (src/common/geauxdock.h:130-146) */
enum {
    GD_E_VDW = 0, GD_E_ELE = 1, GD_E_PMF = 2, GD_E_PSP = 3, GD_E_HDB = 4,
    GD_E_HPC = 5, GD_E_KDE = 6, GD_E_LHM = 7, GD_E_DST = 8, GD_E_TOTAL = 9
};

/* EnePara (src/common/geauxdock_soa.h:154-187): derived energy parameters,
 * tables indexed [ligand type][protein point type]; 64-byte aligned like the
 * reference struct (18432 bytes). */
typedef struct __attribute__((aligned(64))) gd_EnePara {
    float p1a[GD_MAXTP2][GD_MAXTP1];
    float p2a[GD_MAXTP2][GD_MAXTP1];
    float lj0, lj1;
    float el1, el0, a1, b1;
    float pmf0[GD_MAXTP2][GD_MAXTP1];
    float pmf1[GD_MAXTP2][GD_MAXTP1];
    float hdb0[GD_MAXTP2][GD_MAXTP1];
    float hdb1[GD_MAXTP2][GD_MAXTP1];
    float hpp[GD_MAXTP4];
    float hpl0[GD_MAXTP2];
    float hpl1[GD_MAXTP2];
    float hpl2[GD_MAXTP2];
    float kde2, kde3;
    float a_para[GD_MAXWEI];
    float b_para[GD_MAXWEI];
    float w[GD_MAXWEI];
} gd_EnePara;

/* EnePara0 (src/common/geauxdock.h:278-294): raw parameters as read from
 * the "paras", weight and normalisation files. */
typedef struct gd_EnePara0 {
    float vdw[GD_MAXTP1][GD_MAXTP2][2];
    float ele[GD_MAXTP3];
    float pmf[GD_MAXTP1][GD_MAXTP2][2];
    float hpp[GD_MAXTP4];
    float hpl[GD_MAXTP2][2];
    float hdb[GD_MAXTP1][GD_MAXTP2][2];
    float lj[3];
    float el[2];
    float kde;
    float w[GD_MAXWEI];
    float a_para[GD_MAXWEI];
    float b_para[GD_MAXWEI];
} gd_EnePara0;

/* Temp (src/common/geauxdock.h:307-312) */
typedef struct gd_Temp {
    float t;
    float minus_beta;
    int order;
} gd_Temp;

/* ReplicaMC (src/common/geauxdock.h:321-334): 128 bytes, 64-byte aligned,
 * same field order and offsets as the reference. */
typedef struct __attribute__((aligned(64))) gd_ReplicaMC {
    int idx_rep;
    int idx_lig;
    int idx_prt;
    int idx_tmp;
    float movematrix[6]; /* translation x y z, rotation x y z (cumulative) */
    float energy[GD_MAXWEI];
    int step;
    int is_accept;
} gd_ReplicaMC;

/* McPara (src/common/geauxdock.h:346-356) without the output path strings */
typedef struct gd_McPara {
    int steps_per_dump;
    int steps_n_dump;
    int steps_total;
    float move_scale[6];
} gd_McPara;

/* ComplexSize (src/common/geauxdock.h:112-125) */
typedef struct gd_ComplexSize {
    int n_lig;
    int n_prt;
    int n_tmp;
    int n_rep;
    int lig_natom;
    int prt_npoint;
    int kde_npoint;
    int mcs_nrow;
} gd_ComplexSize;

/*
 * gd_Complex: the staged tables of one docking job, i.e. what the reference
 * keeps in struct Complex (src/common/geauxdock.h:392-424) after
 * Loader::build_complex.  Arrays are owned by whoever built the view (the
 * library's loader, or the caller).  Per-conformation arrays are stored
 * conformation-major: element (conf, i) lives at [conf * n + i].
 *
 *  Ligand   (geauxdock_soa.h:8-44)  x2..c2: file atom order (MCS),
 *                                   x3..c3: sorted by SYBYL type (PRT, KDE)
 *  Protein  (geauxdock_soa.h:52-72) effective points sorted by class c
 *  Psp      (geauxdock_soa.h:77-81) psp[lig_type * psp_width + residue serial]
 *  Kde      (geauxdock_soa.h:85-93) sorted by type
 *  Mcs_ELL  (geauxdock_soa.h:120-128) rows flattened: row m holds entries
 *           [mcs_rowptr[m], mcs_rowptr[m+1]) with 0-based atom index mcs_i
 */
typedef struct gd_Complex {
    gd_ComplexSize size;

    /* Ligand[n_lig], each lig_natom atoms */
    float *lig_x2, *lig_y2, *lig_z2, *lig_c2;
    int *lig_t2;
    float *lig_x3, *lig_y3, *lig_z3, *lig_c3;
    int *lig_t3;
    int *lig_kde_begin_idx, *lig_kde_end_idx;
    float lig_center[3];

    /* Protein[n_prt], each prt_npoint points */
    float *prt_x, *prt_y, *prt_z;
    float *prt_hpp, *prt_ele;
    int *prt_t, *prt_c, *prt_cdc, *prt_seq3r;
    float prt_pocket_center[3];

    /* Psp */
    int psp_width;
    float *psp; /* [GD_MAXTP2 * psp_width] */

    /* Kde */
    int *kde_t;
    float *kde_x, *kde_y, *kde_z;

    /* Mcs (ELL rows) */
    int *mcs_ncol;
    float *mcs_tcc;
    int *mcs_rowptr; /* [mcs_nrow + 1] */
    int *mcs_i;
    float *mcs_x, *mcs_y, *mcs_z;

    gd_EnePara enepara;
    gd_Temp *temp;          /* [n_tmp] */
    gd_ReplicaMC *replica;  /* [n_rep], initial state */
    gd_McPara mcpara;
} gd_Complex;

/*
 * gd_Raw: loader output before staging -- the reference's Ligand0, Protein0,
 * Psp0, Kde0, Mcs0 and EnePara0 (src/common/geauxdock.h:174-294) with
 * run-time sizes.  Staging (gd_stage) turns it into a gd_Complex the same
 * way the reference's Optimize* builders do (src/frontend/util_optimize_soa.cpp).
 */
typedef struct gd_Raw {
    /* Ligand0[n_lig]: coordinates relative to each conformation's centroid,
     * file atom order; types and charges shared by all conformations */
    int n_lig, lig_natom;
    float *lig_x, *lig_y, *lig_z; /* [n_lig * lig_natom] */
    int *lig_t;                   /* [lig_natom] */
    float *lig_c;                 /* [lig_natom] */
    float lig_center[3];          /* native-pose centroid = pocket centre */

    /* Protein0[n_prt]: effective points in load order */
    int n_prt, prt_npoint, prt_nres;
    float *prt_x, *prt_y, *prt_z; /* [n_prt * prt_npoint] */
    int *prt_t, *prt_c, *prt_d, *prt_r; /* [prt_npoint] type, class, residue code, residue index */
    int *prt_seq3;                /* [prt_nres] residue serial numbers */

    /* Psp0 as sparse entries psp0[residue][lig_type] = val */
    int psp_n;
    int *psp_res, *psp_lig;
    float *psp_val;

    /* Kde0 in file order */
    int kde_npoint;
    int *kde_t;
    float *kde_x, *kde_y, *kde_z;

    /* Mcs0 rows in file order (entries in file order) */
    int mcs_nrow;
    float *mcs_tcc;
    int *mcs_rowptr; /* [mcs_nrow + 1] */
    int *mcs_i;      /* 0-based ligand atom index */
    float *mcs_x, *mcs_y, *mcs_z;

    gd_EnePara0 enepara0;

    /* run parameters from the job line (src/frontend/frontend.cpp:70-123) */
    int n_tmp;
    float temp_high, temp_low;
    float move_scale[6];
    int steps_n_dump;
} gd_Raw;

#ifdef __cplusplus
}
#endif

#endif /* GD_TYPES_H */
