/*
 * Body of the oracle's energy function, included once per precision by
 * gd_oracle.c with REAL / SFX / RSIN / RCOS / REXP / RSQRT / RLOG / RPI set.
 *
 * Restates the pose energy of the reference CPU backend term by term and in
 * the same operation order, so the float instantiation reproduces the
 * reference's float rounding:
 *   move + Euler rotation     src/backend_cpu_mic/kernel_cpp_montecarlo_soa.cpp:227-278
 *   PRT pair terms            src/backend_cpu_mic/energy_prt_v1.cpp:5-156
 *   hdb scale, term vector    kernel_cpp_montecarlo_soa.cpp:184, 303-309
 *   KDE                       src/backend_cpu_mic/energy_kde_v1.cpp:1-77
 *   MCS (ELL)                 src/backend_cpu_mic/energy_mcs_v3_ell.cpp:1-57
 *   DST, normalise, total     kernel_cpp_montecarlo_soa.cpp:333-359
 */

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)
#define K(x) ((REAL)(x))

/* rotation matrix of the cumulative Euler angles m[3..5] */
static void FN(rotation_)(const REAL *m, REAL rot[3][3])
{
    const REAL sin1 = RSIN(m[3]);
    const REAL cos1 = RCOS(m[3]);
    const REAL sin2 = RSIN(m[4]);
    const REAL cos2 = RCOS(m[4]);
    const REAL sin3 = RSIN(m[5]);
    const REAL cos3 = RCOS(m[5]);
    rot[0][0] = cos1 * cos2;
    rot[0][1] = cos1 * sin2 * sin3 - cos3 * sin1;
    rot[0][2] = sin1 * sin3 + cos1 * cos3 * sin2;
    rot[1][0] = cos2 * sin1;
    rot[1][1] = cos1 * cos3 + sin1 * sin2 * sin3;
    rot[1][2] = cos3 * sin1 * sin2 - cos1 * sin3;
    rot[2][0] = -1 * sin2;
    rot[2][1] = cos2 * sin3;
    rot[2][2] = cos2 * cos3;
}

/*
 * Energy of ligand conformation il in protein conformation ip at cumulative
 * move m[6]; writes the 10 normalised terms to e_s.  lx*: caller scratch of
 * 6 * lig_natom REALs.
 */
static void FN(pose_energy_)(const gd_Complex *cx, int il, int ip, const REAL *m, REAL *e_s, REAL *scratch)
{
    const gd_EnePara *ep = &cx->enepara;
    const int lna = cx->size.lig_natom;
    const int npt = cx->size.prt_npoint;
    const int nkde = cx->size.kde_npoint;
    const int nrow = cx->size.mcs_nrow;
    const size_t lo = (size_t)il * lna;
    const size_t po = (size_t)ip * npt;
    REAL *lx2 = scratch, *ly2 = scratch + lna, *lz2 = scratch + 2 * lna;
    REAL *lx3 = scratch + 3 * lna, *ly3 = scratch + 4 * lna, *lz3 = scratch + 5 * lna;
    REAL rot[3][3];
    int l, p, k, i;

    FN(rotation_)(m, rot);
    {
        const REAL c0 = cx->lig_center[0], c1 = cx->lig_center[1], c2 = cx->lig_center[2];
        for (l = 0; l < lna; ++l) {
            const REAL x2 = cx->lig_x2[lo + l], y2 = cx->lig_y2[lo + l], z2 = cx->lig_z2[lo + l];
            const REAL x3 = cx->lig_x3[lo + l], y3 = cx->lig_y3[lo + l], z3 = cx->lig_z3[lo + l];
            lx2[l] = rot[0][0] * x2 + rot[0][1] * y2 + rot[0][2] * z2 + m[0] + c0;
            ly2[l] = rot[1][0] * x2 + rot[1][1] * y2 + rot[1][2] * z2 + m[1] + c1;
            lz2[l] = rot[2][0] * x2 + rot[2][1] * y2 + rot[2][2] * z2 + m[2] + c2;
            lx3[l] = rot[0][0] * x3 + rot[0][1] * y3 + rot[0][2] * z3 + m[0] + c0;
            ly3[l] = rot[1][0] * x3 + rot[1][1] * y3 + rot[1][2] * z3 + m[1] + c1;
            lz3[l] = rot[2][0] * x3 + rot[2][1] * y3 + rot[2][2] * z3 + m[2] + c2;
        }
    }

    /* ---- PRT ---- */
    REAL evdw = 0, eele = 0, epmf = 0, epsp = 0, ehdb = 0, ehpc = 0;
    const REAL lj0 = ep->lj0, lj1 = ep->lj1, el0 = ep->el0, el1 = ep->el1, a1 = ep->a1, b1 = ep->b1;
    for (l = 0; l < lna; ++l) {
        REAL ehpc1 = 0;
        const int lt = cx->lig_t3[lo + l];
        const REAL lc = cx->lig_c3[lo + l];
        for (p = 0; p < npt; ++p) {
            const int pt = cx->prt_t[po + p];
            const REAL dx = lx3[l] - (REAL)cx->prt_x[po + p];
            const REAL dy = ly3[l] - (REAL)cx->prt_y[po + p];
            const REAL dz = lz3[l] - (REAL)cx->prt_z[po + p];
            const REAL d2 = dx * dx + dy * dy + dz * dz + K(0.0);
            const REAL d4 = d2 * d2;
            const REAL d = RSQRT(d2);

            /* hydrophobic contact, 81 = 9 A squared cutoff */
            if (cx->prt_cdc[po + p] == 1 && d2 <= K(81.0)) {
                ehpc1 += (REAL)cx->prt_hpp[po + p] *
                         (K(1.0) - (K(3.5) / K(81.0) * d2 -
                                    K(4.5) / K(81.0) / K(81.0) * d4 +
                                    K(2.5) / K(81.0) / K(81.0) / K(81.0) * d4 * d2 -
                                    K(0.5) / K(81.0) / K(81.0) / K(81.0) / K(81.0) * d4 * d4));
            }

            /* soft-core Lennard-Jones */
            {
                const REAL pp1 = (REAL)ep->p1a[lt][pt] / (d4 * d4 * d);
                const REAL pp2 = (REAL)ep->p2a[lt][pt] / (d4 * d2);
                const REAL pp4 = pp1 * lj0 * (K(1.0) + lj1 * d2) + K(1.0);
                evdw += (pp1 - pp2) / pp4;
            }

            /* electrostatics with a cubic core inside s < 1 */
            {
                const REAL s1 = el1 * d;
                REAL g1;
                if (s1 < K(1.0))
                    g1 = el0 + a1 * s1 * s1 + b1 * s1 * s1 * s1;
                else
                    g1 = K(1.0) / s1;
                eele += lc * (REAL)cx->prt_ele[po + p] * g1;
            }

            /* contact potential and pocket-specific term */
            {
                const REAL dmp = d - (REAL)ep->pmf0[lt][pt];
                epmf += (REAL)ep->pmf1[lt][pt] / (K(1.0) + REXP((K(-0.5) * d + K(6.0)) * dmp));
                if (cx->prt_c[po + p] == 2 && dmp <= 0) {
                    const int i1 = cx->prt_seq3r[po + p];
                    epsp += (REAL)cx->psp[(size_t)lt * cx->psp_width + i1];
                }
            }

            /* hydrogen bond */
            {
                const REAL h0 = ep->hdb0[lt][pt];
                if (h0 > K(0.1)) {
                    const REAL h1 = ep->hdb1[lt][pt];
                    const REAL h3 = (d - h0) * h1;
                    ehdb += h1 * REXP(K(-0.5) * h3 * h3);
                }
            }
        }
        {
            const REAL hpc2 = (ehpc1 - (REAL)ep->hpl0[lt]) / (REAL)ep->hpl1[lt];
            ehpc += K(0.5) * hpc2 * hpc2 - (REAL)ep->hpl2[lt];
        }
    }

    e_s[0] = evdw;
    e_s[1] = eele;
    e_s[2] = epmf;
    e_s[3] = epsp;
    e_s[4] = ehdb * (K(-1.0) / RSQRT(K(2.0) * RPI));
    e_s[5] = ehpc;

    /* ---- KDE ---- */
    {
        REAL ekde = 0;
        const REAL kde2 = ep->kde2;
        for (l = 0; l < lna; ++l) {
            REAL ekde1 = 0;
            int ekde2 = 0;
            const int lt = cx->lig_t3[lo + l];
            for (k = 0; k < nkde; ++k) {
                if (lt == cx->kde_t[k]) {
                    const REAL dx = lx3[l] - (REAL)cx->kde_x[k];
                    const REAL dy = ly3[l] - (REAL)cx->kde_y[k];
                    const REAL dz = lz3[l] - (REAL)cx->kde_z[k];
                    const REAL kd2 = dx * dx + dy * dy + dz * dz;
                    ekde1 += REXP(kde2 * kd2);
                    ekde2++;
                }
            }
            if (ekde2 != 0) ekde += (ekde1 / (REAL)ekde2);
        }
        e_s[6] = ekde * (K(1.0) / (REAL)ep->kde3);
    }

    /* ---- MCS (ELL rows over the file-order atom set) ---- */
    {
        REAL elhm = 0;
        for (i = 0; i < nrow; ++i) {
            REAL elhm1 = 0;
            const int ncol = cx->mcs_ncol[i];
            const int b = cx->mcs_rowptr[i];
            for (k = 0; k < ncol; ++k) {
                const int a = cx->mcs_i[b + k];
                const REAL dx = lx2[a] - (REAL)cx->mcs_x[b + k];
                const REAL dy = ly2[a] - (REAL)cx->mcs_y[b + k];
                const REAL dz = lz2[a] - (REAL)cx->mcs_z[b + k];
                elhm1 += dx * dx + dy * dy + dz * dz;
            }
            elhm += (REAL)cx->mcs_tcc[i] * RSQRT(elhm1 / (REAL)ncol);
        }
        e_s[7] = RLOG(elhm / nrow);
    }

    /* ---- DST ---- */
    {
        REAL dst = 0;
        for (i = 0; i < 3; ++i) {
            const REAL dd = (REAL)cx->lig_center[i] + m[i] - (REAL)cx->prt_pocket_center[i];
            dst += dd * dd;
        }
        e_s[8] = RSQRT(dst);
    }

    /* ---- normalise and weight ---- */
    for (i = 0; i < 7; ++i) e_s[i] = e_s[i] / lna;
    for (i = 0; i < GD_MAXWEI - 1; ++i) e_s[i] = (REAL)ep->a_para[i] * e_s[i] + (REAL)ep->b_para[i];
    {
        REAL etotal = 0;
        for (i = 0; i < GD_MAXWEI - 1; ++i) etotal += (REAL)ep->w[i] * e_s[i];
        e_s[GD_MAXWEI - 1] = etotal;
    }
}

#undef FN
#undef K
#undef CAT
#undef CAT2
