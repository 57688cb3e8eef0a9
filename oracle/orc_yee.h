/*
 * ORACLE (test infrastructure only).  Generic body of the Yee + CFS-CPML restatement,
 * included by mp_oracle.c once per (storage ST, compute CT, psi storage PT) combination.
 * Nothing in the product links this code; it is the CPU checker for libmphip.
 *
 * Follows src/fdtd/Kernels.jl:9-152 (reference paths, 1-based there, 0-based here).
 * Julia's muladd is written as fma; every other expression keeps Julia's evaluation
 * order and is compiled with -ffp-contract=off so no other contraction happens.
 *
 * Required macros: ST, CT, PT, SFX(name), CFMA (fma for CT), PSI_LD(ptr, off) and
 * PSI_ST(ptr, off, v) (psi storage; the reduced kinds use the scales psc/pisc in scope).
 */

/* _pml_slab (Kernels.jl:9-14), 0-based: returns region 0 (none), 1 (lo) or 2 (hi) and the
 * 0-based slab index in *js. */
static inline int SFX(pml_slab)(int64_t idx, int64_t N, int64_t npml, int64_t *js) {
    if (npml <= 0) { *js = 0; return 0; }
    if (idx < npml) { *js = idx; return 1; }
    if (idx >= N - npml) { *js = idx - (N - npml); return 2; }
    *js = 0;
    return 0;
}

/* _cpml_term_{x,y,z}! (Kernels.jl:16-56): psi <- b*psi + a*deriv (muladd), only inside
 * a slab.  `off` is the linear index into the slab given the slab coordinate. */
static inline CT SFX(cpml_term)(PT *lo, PT *hi, int region, int64_t off, const CT *b,
                                const CT *a, int64_t idx, CT deriv, float psc, float pisc) {
    (void)psc;
    (void)pisc;
    if (region == 1) {
        CT p = CFMA(b[idx], PSI_LD(lo, off), a[idx] * deriv);
        PSI_ST(lo, off, p);
        return p;
    } else if (region == 2) {
        CT p = CFMA(b[idx], PSI_LD(hi, off), a[idx] * deriv);
        PSI_ST(hi, off, p);
        return p;
    }
    return (CT)0;
}

/* yee_H_3d_kernel! (Kernels.jl:60-101) over 0 <= i < nx-1, 0 <= j < ny-1, 0 <= k < nz-1. */
static void SFX(orc_yee_H)(const mp_yee *g, mp_cpml *c, CT dt_mu0, float psc, float pisc) {
    const int64_t nx = g->nx, ny = g->ny, nz = g->nz;
    const int64_t sy = nx, sz = nx * ny;
    const ST *Ex = (const ST *)g->ex, *Ey = (const ST *)g->ey, *Ez = (const ST *)g->ez;
    ST *Hx = (ST *)g->hx, *Hy = (ST *)g->hy, *Hz = (ST *)g->hz;
    const CT *idcx = (const CT *)g->inv_d_cell_x, *idcy = (const CT *)g->inv_d_cell_y,
             *idcz = (const CT *)g->inv_d_cell_z;
    const CT *xk = (const CT *)c->inv_kappa[0], *xa = (const CT *)c->a[0], *xb = (const CT *)c->b[0];
    const CT *yk = (const CT *)c->inv_kappa[1], *ya = (const CT *)c->a[1], *yb = (const CT *)c->b[1];
    const CT *zk = (const CT *)c->inv_kappa[2], *za = (const CT *)c->a[2], *zb = (const CT *)c->b[2];
    const int64_t wx = c->npml_x, wy = c->npml_y, wz = c->npml_z;
    PT **ps = (PT **)c->psi_h;
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < nz - 1; ++k) {
        for (int64_t j = 0; j < ny - 1; ++j) {
            for (int64_t i = 0; i < nx - 1; ++i) {
                const int64_t o = i + sy * j + sz * k;
                CT dEz_dy = ((CT)Ez[o + sy] - (CT)Ez[o]) * idcy[j];
                CT dEy_dz = ((CT)Ey[o + sz] - (CT)Ey[o]) * idcz[k];
                CT dEx_dz = ((CT)Ex[o + sz] - (CT)Ex[o]) * idcz[k];
                CT dEz_dx = ((CT)Ez[o + 1] - (CT)Ez[o]) * idcx[i];
                CT dEy_dx = ((CT)Ey[o + 1] - (CT)Ey[o]) * idcx[i];
                CT dEx_dy = ((CT)Ex[o + sy] - (CT)Ex[o]) * idcy[j];
                int64_t jsx, jsy, jsz;
                int rx = SFX(pml_slab)(i, nx, wx, &jsx);
                int ry = SFX(pml_slab)(j, ny, wy, &jsy);
                int rz = SFX(pml_slab)(k, nz, wz, &jsz);
                const int64_t ox = jsx + wx * (j + ny * k);   /* x-slab (w, ny, nz) */
                const int64_t oy = i + nx * (jsy + wy * k);   /* y-slab (nx, w, nz) */
                const int64_t oz = i + nx * (j + ny * jsz);   /* z-slab (nx, ny, w) */

                CT pHxy = SFX(cpml_term)(ps[0], ps[1], ry, oy, yb, ya, j, dEz_dy, psc, pisc);
                CT pHxz = SFX(cpml_term)(ps[2], ps[3], rz, oz, zb, za, k, dEy_dz, psc, pisc);
                Hx[o] = (ST)((CT)Hx[o] - dt_mu0 * ((dEz_dy * yk[j] + pHxy) - (dEy_dz * zk[k] + pHxz)));

                CT pHyz = SFX(cpml_term)(ps[4], ps[5], rz, oz, zb, za, k, dEx_dz, psc, pisc);
                CT pHyx = SFX(cpml_term)(ps[6], ps[7], rx, ox, xb, xa, i, dEz_dx, psc, pisc);
                Hy[o] = (ST)((CT)Hy[o] - dt_mu0 * ((dEx_dz * zk[k] + pHyz) - (dEz_dx * xk[i] + pHyx)));

                CT pHzx = SFX(cpml_term)(ps[8], ps[9], rx, ox, xb, xa, i, dEy_dx, psc, pisc);
                CT pHzy = SFX(cpml_term)(ps[10], ps[11], ry, oy, yb, ya, j, dEx_dy, psc, pisc);
                Hz[o] = (ST)((CT)Hz[o] - dt_mu0 * ((dEy_dx * xk[i] + pHzx) - (dEx_dy * yk[j] + pHzy)));
            }
        }
    }
}

/* yee_E_3d_kernel! (Kernels.jl:103-152) over 1 <= i < nx, 1 <= j < ny, 1 <= k < nz. */
static void SFX(orc_yee_E)(const mp_yee *g, mp_cpml *c, CT dt, float psc, float pisc) {
    const int64_t nx = g->nx, ny = g->ny, nz = g->nz;
    const int64_t sy = nx, sz = nx * ny;
    ST *Ex = (ST *)g->ex, *Ey = (ST *)g->ey, *Ez = (ST *)g->ez;
    const ST *Hx = (const ST *)g->hx, *Hy = (const ST *)g->hy, *Hz = (const ST *)g->hz;
    ST *Dx = (ST *)g->dx, *Dy = (ST *)g->dy, *Dz = (ST *)g->dz;
    const ST *iex = (const ST *)g->inv_eps_x, *iey = (const ST *)g->inv_eps_y,
             *iez = (const ST *)g->inv_eps_z;
    const CT *iddx = (const CT *)g->inv_d_dual_x, *iddy = (const CT *)g->inv_d_dual_y,
             *iddz = (const CT *)g->inv_d_dual_z;
    const CT *xk = (const CT *)c->inv_kappa[0], *xa = (const CT *)c->a[0], *xb = (const CT *)c->b[0];
    const CT *yk = (const CT *)c->inv_kappa[1], *ya = (const CT *)c->a[1], *yb = (const CT *)c->b[1];
    const CT *zk = (const CT *)c->inv_kappa[2], *za = (const CT *)c->a[2], *zb = (const CT *)c->b[2];
    const int64_t wx = c->npml_x, wy = c->npml_y, wz = c->npml_z;
    PT **ps = (PT **)c->psi_d;
#pragma omp parallel for schedule(static)
    for (int64_t k = 1; k < nz; ++k) {
        for (int64_t j = 1; j < ny; ++j) {
            for (int64_t i = 1; i < nx; ++i) {
                const int64_t o = i + sy * j + sz * k;
                CT dHz_dy = ((CT)Hz[o] - (CT)Hz[o - sy]) * iddy[j];
                CT dHy_dz = ((CT)Hy[o] - (CT)Hy[o - sz]) * iddz[k];
                CT dHx_dz = ((CT)Hx[o] - (CT)Hx[o - sz]) * iddz[k];
                CT dHz_dx = ((CT)Hz[o] - (CT)Hz[o - 1]) * iddx[i];
                CT dHy_dx = ((CT)Hy[o] - (CT)Hy[o - 1]) * iddx[i];
                CT dHx_dy = ((CT)Hx[o] - (CT)Hx[o - sy]) * iddy[j];
                int64_t jsx, jsy, jsz;
                int rx = SFX(pml_slab)(i, nx, wx, &jsx);
                int ry = SFX(pml_slab)(j, ny, wy, &jsy);
                int rz = SFX(pml_slab)(k, nz, wz, &jsz);
                const int64_t ox = jsx + wx * (j + ny * k);
                const int64_t oy = i + nx * (jsy + wy * k);
                const int64_t oz = i + nx * (j + ny * jsz);

                CT pDxy = SFX(cpml_term)(ps[0], ps[1], ry, oy, yb, ya, j, dHz_dy, psc, pisc);
                CT pDxz = SFX(cpml_term)(ps[2], ps[3], rz, oz, zb, za, k, dHy_dz, psc, pisc);
                CT cx_term = (dHz_dy * yk[j] + pDxy) - (dHy_dz * zk[k] + pDxz);

                CT pDyz = SFX(cpml_term)(ps[4], ps[5], rz, oz, zb, za, k, dHx_dz, psc, pisc);
                CT pDyx = SFX(cpml_term)(ps[6], ps[7], rx, ox, xb, xa, i, dHz_dx, psc, pisc);
                CT cy_term = (dHx_dz * zk[k] + pDyz) - (dHz_dx * xk[i] + pDyx);

                CT pDzx = SFX(cpml_term)(ps[8], ps[9], rx, ox, xb, xa, i, dHy_dx, psc, pisc);
                CT pDzy = SFX(cpml_term)(ps[10], ps[11], ry, oy, yb, ya, j, dHx_dy, psc, pisc);
                CT cz_term = (dHy_dx * xk[i] + pDzx) - (dHx_dy * yk[j] + pDzy);

                CT dx_new = (CT)Dx[o] + dt * cx_term;
                CT dy_new = (CT)Dy[o] + dt * cy_term;
                CT dz_new = (CT)Dz[o] + dt * cz_term;
                Dx[o] = (ST)dx_new;
                Dy[o] = (ST)dy_new;
                Dz[o] = (ST)dz_new;
                Ex[o] = (ST)(dx_new * (CT)iex[o]);
                Ey[o] = (ST)(dy_new * (CT)iey[o]);
                Ez[o] = (ST)(dz_new * (CT)iez[o]);
            }
        }
    }
}
