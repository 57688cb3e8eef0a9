/*
 * ORACLE (test infrastructure only).  Generic body of the source, ADE and MO
 * restatements, included by mp_oracle.c once per (storage ST, compute CT).
 *
 * Follows src/fdtd/Kernels.jl:154-305.  Required macros: ST, CT, SFX2(name), CFMA.
 */

/* inject_soft_3d_kernel! (Kernels.jl:154-174). */
static void SFX2(orc_inject_soft)(const mp_yee *g, int code, int64_t o, CT val) {
    ST *E = NULL, *D = NULL, *H = NULL;
    const ST *ie = NULL;
    switch (code) {
    case 1: E = (ST *)g->ex; D = (ST *)g->dx; ie = (const ST *)g->inv_eps_x; break;
    case 2: E = (ST *)g->ey; D = (ST *)g->dy; ie = (const ST *)g->inv_eps_y; break;
    case 3: E = (ST *)g->ez; D = (ST *)g->dz; ie = (const ST *)g->inv_eps_z; break;
    case 4: H = (ST *)g->hx; break;
    case 5: H = (ST *)g->hy; break;
    case 6: H = (ST *)g->hz; break;
    default: return;
    }
    if (E) {
        E[o] = (ST)((CT)E[o] + val);
        D[o] = (ST)((CT)D[o] + val / (CT)ie[o]);
    } else {
        H[o] = (ST)((CT)H[o] + val);
    }
}

/* inject_mode_x_3d_kernel! (Kernels.jl:176-192) on the x-plane i. */
static void SFX2(orc_inject_mode_x)(const mp_yee *g, int code, int64_t i, const CT *profile,
                                    CT value, CT hcorr) {
    const int64_t nx = g->nx, ny = g->ny, nz = g->nz;
    for (int64_t k = 0; k < nz; ++k) {
        for (int64_t j = 0; j < ny; ++j) {
            const int64_t o = i + nx * (j + ny * k);
            CT amp = value * profile[j + ny * k];
            if (code == 2) {
                ST *Ey = (ST *)g->ey, *Dy = (ST *)g->dy, *Hz = (ST *)g->hz;
                const ST *ie = (const ST *)g->inv_eps_y;
                Ey[o] = (ST)((CT)Ey[o] + amp);
                Dy[o] = (ST)((CT)Dy[o] + amp / (CT)ie[o]);
                Hz[o] = (ST)((CT)Hz[o] + hcorr * amp);
            } else if (code == 3) {
                ST *Ez = (ST *)g->ez, *Dz = (ST *)g->dz, *Hy = (ST *)g->hy;
                const ST *ie = (const ST *)g->inv_eps_z;
                Ez[o] = (ST)((CT)Ez[o] + amp);
                Dz[o] = (ST)((CT)Dz[o] + amp / (CT)ie[o]);
                Hy[o] = (ST)((CT)Hy[o] - hcorr * amp);
            }
        }
    }
}

/* ade_patch_kernel! (Kernels.jl:194-231) over the whole list. */
/* P/J storage kind sk: CT, or BF16/binary16 holding P*sc and J*jsc (read back with
 * isc = 1/sc, ijsc = 1/jsc, FP32 products). */
static void SFX2(orc_ade_patch)(ST *E_arr, mp_ade *s, CT inv_alpha, int sk, float sc, float isc,
                                float jsc, float ijsc) {
    const int64_t N = s->n;
    const int np = s->npoles;
    void *P = s->P, *J = s->J;
    CT *E_old = (CT *)s->E_old;
    const CT *fill = (const CT *)s->fill, *inv_eps = (const CT *)s->inv_eps;
    const CT *poles = (const CT *)s->poles;
#pragma omp parallel for schedule(static)
    for (int64_t n = 0; n < N; ++n) {
        const int64_t idx = s->idx[n];
        CT Eprev = E_old[n];
        CT f = fill[n];
        CT ie = inv_eps[n];
        CT sum_psi = (CT)0, sum_beta = (CT)0;
        for (int p = 0; p < np; ++p) {
            CT C1 = poles[3 * p], C2 = poles[3 * p + 1], C3 = poles[3 * p + 2];
            CT beta = C3 * f;
            CT psi = CFMA(C1, (CT)orc_sld(P, n + p * N, sk, isc),
                          CFMA(C2, (CT)orc_sld(J, n + p * N, sk, ijsc), beta * Eprev));
            sum_psi += psi;
            sum_beta += beta;
        }
        CT den = (CT)1 + sum_beta * ie;
        CT Etemp = (CT)E_arr[idx];
        if (isfinite(Etemp) && isfinite(sum_psi) && isfinite(den) && fabs((double)den) >= (double)(CT)1e-30) {
            CT Efinal = (Etemp - sum_psi * ie) / den;
            E_arr[idx] = (ST)Efinal;
            E_old[n] = Efinal;
            for (int p = 0; p < np; ++p) {
                CT C1 = poles[3 * p], C2 = poles[3 * p + 1], C3 = poles[3 * p + 2];
                CT beta = C3 * f;
                CT Pold = (CT)orc_sld(P, n + p * N, sk, isc);
                CT Jold = (CT)orc_sld(J, n + p * N, sk, ijsc);
                CT psi = CFMA(C1, Pold, CFMA(C2, Jold, beta * Eprev));
                CT Pnew = psi + beta * Efinal;
                orc_sst(P, n + p * N, sk, Pnew, sc);
                orc_sst(J, n + p * N, sk, (Pnew - Pold) * inv_alpha - Jold, jsc);
            }
        }
    }
}

/* mo_gyration_kernel! (Kernels.jl:233-305) with the pair (a, b) = (Ey, Ez) in the
 * reference; generic here. */
static void SFX2(orc_mo_gyration)(ST *Ea, ST *Eb, mp_mo *s, const ST *iea, const ST *ieb,
                                  const CT *m_TM, const CT *m_RE, CT Q_TM, CT Q_RE,
                                  CT inv_alpha) {
    const int64_t N = s->n;
    const int np = s->npoles;
    CT *Pa = (CT *)s->Pa_from_b, *Ja = (CT *)s->Ja_from_b;
    CT *Pb = (CT *)s->Pb_from_a, *Jb = (CT *)s->Jb_from_a;
    CT *Eoa = (CT *)s->E_old_a, *Eob = (CT *)s->E_old_b;
    const CT *fill = (const CT *)s->fill;
    const CT *poles = (const CT *)s->poles;
#pragma omp parallel for schedule(static)
    for (int64_t n = 0; n < N; ++n) {
        const int64_t idx = s->idx[n];
        const int32_t mpos = s->mpos[n];
        if (mpos < 0) continue;
        CT f = fill[n];
        CT qeff = Q_TM * m_TM[mpos] + Q_RE * m_RE[mpos];
        CT Ea_drive = (CT)Ea[idx];
        CT Eb_drive = (CT)Eb[idx];
        CT Ea_old = Eoa[n];
        CT Eb_old = Eob[n];
        CT sum_psi_a = (CT)0, sum_psi_b = (CT)0, sumC3 = (CT)0;
        for (int p = 0; p < np; ++p) {
            CT C1 = poles[3 * p], C2 = poles[3 * p + 1], C3 = poles[3 * p + 2];
            CT beta = C3 * f * qeff;
            CT Pa_old = Pa[n + p * N], Ja_old = Ja[n + p * N];
            sum_psi_a += CFMA(C1, Pa_old, CFMA(C2, Ja_old, beta * Eb_old));
            CT Pb_old = Pb[n + p * N], Jb_old = Jb[n + p * N];
            sum_psi_b += CFMA(C1, Pb_old, CFMA(C2, Jb_old, beta * Ea_old));
            sumC3 += C3;
        }
        CT ica = (CT)iea[idx];
        CT icb = (CT)ieb[idx];
        CT Bmo = f * qeff * sumC3;
        CT half = (CT)0.5;
        CT aa = half * ica * Bmo;
        CT ab = half * icb * Bmo;
        CT RHSa = Ea_drive - ica * sum_psi_a - aa * Eb_drive;
        CT RHSb = Eb_drive + icb * sum_psi_b + ab * Ea_drive;
        CT den = (CT)1 + aa * ab;
        CT Eb_final = (RHSb + ab * RHSa) / den;
        CT Ea_final = RHSa - aa * Eb_final;
        if (isfinite(Ea_final) && isfinite(Eb_final)) {
            Ea[idx] = (ST)Ea_final;
            Eb[idx] = (ST)Eb_final;
            Eoa[n] = Ea_final;
            Eob[n] = Eb_final;
            CT Ea_mid = half * (Ea_drive + Ea_final);
            CT Eb_mid = half * (Eb_drive + Eb_final);
            for (int p = 0; p < np; ++p) {
                CT C1 = poles[3 * p], C2 = poles[3 * p + 1], C3 = poles[3 * p + 2];
                CT beta = C3 * f * qeff;
                CT Pa_old = Pa[n + p * N], Ja_old = Ja[n + p * N];
                CT Pa_new = CFMA(C1, Pa_old, CFMA(C2, Ja_old, beta * Eb_old)) + beta * Eb_mid;
                Pa[n + p * N] = Pa_new;
                Ja[n + p * N] = (Pa_new - Pa_old) * inv_alpha - Ja_old;
                CT Pb_old = Pb[n + p * N], Jb_old = Jb[n + p * N];
                CT Pb_new = CFMA(C1, Pb_old, CFMA(C2, Jb_old, beta * Ea_old)) + beta * Ea_mid;
                Pb[n + p * N] = Pb_new;
                Jb[n + p * N] = (Pb_new - Pb_old) * inv_alpha - Jb_old;
            }
        }
    }
}

/* pabs_accumulate_kernel! (Kernels.jl:313-355); FP64 throughout. */
static void SFX2(orc_pabs)(mp_film *f, const mp_yee *g, const mp_ade *ax, const mp_ade *ay,
                           const mp_ade *az, int has_ade_work, double pabs_factor, double dt,
                           int sk, float ijsc) {
    const ST *Ex = (const ST *)g->ex, *Ey = (const ST *)g->ey, *Ez = (const ST *)g->ez;
    const int64_t N = f->n;
#pragma omp parallel for schedule(static)
    for (int64_t n = 0; n < N; ++n) {
        const int64_t li = f->cells[n];
        double ex = (double)Ex[li], ey = (double)Ey[li], ez = (double)Ez[li];
        double pabs;
        if (has_ade_work) {
            double jx = 0.0, jy = 0.0, jz = 0.0;
            const int32_t px = f->pos_x[n], py = f->pos_y[n], pz = f->pos_z[n];
            const int np = ax->npoles;
            for (int p = 0; p < np; ++p) {
                if (px >= 0) jx += orc_sld(ax->J, px + p * ax->n, sk, ijsc);
                if (py >= 0) jy += orc_sld(ay->J, py + p * ay->n, sk, ijsc);
                if (pz >= 0) jz += orc_sld(az->J, pz + p * az->n, sk, ijsc);
            }
            pabs = jl_max(ex * jx + ey * jy + ez * jz, 0.0);
        } else {
            pabs = pabs_factor * (ex * ex + ey * ey + ez * ez);
        }
        double pdt = pabs * dt;
        f->U_abs[n] += pdt;
        f->window_sum[n] += pdt;
        if (pabs > f->peak[n]) f->peak[n] = pabs;
    }
}
