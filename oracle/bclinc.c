/* bclinc.c -- CPU oracle of bclinc_, test infrastructure only.  One loop nest
 * per reference kernel, in launch order, with the reference's operation
 * order (compiled with -ffp-contract=off).  See bclinc.h for the stages. */
#include "bclinc.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define P2(a, j, i) (a)[(size_t)(j) * imt + (i)]
#define P3(a, k, j, i) (a)[((size_t)(k) * jmt + (j)) * imt + (i)]

static double *zalloc(size_t n) {
    double *p = (double *)calloc(n ? n : 1, sizeof(double));
    if (!p) abort();
    return p;
}

/* bclinc12_cu (bclinc.cpp:28-62): surface stress and bottom drag */
static void bclinc12(const orc_bclinc_grid *g, const orc_bclinc_fields *f) {
    const int imt = g->imt, jmt = g->jmt;
    for (int j = 1; j < jmt - 1; ++j)
        for (int i = 1; i < imt - 1; ++i) {
            const int kmb = P2(g->kmu, j, i) - 1;
            if (kmb < 0) continue;
            const double ub = P3(f->up, kmb, j, i), vb = P3(f->vp, kmb, j, i);
            const double temp = g->c0f * sqrt(ub * ub + vb * vb);
            P2(f->sbcx, j, i) = P2(f->su, j, i) * g->od0;
            P2(f->sbcy, j, i) = P2(f->sv, j, i) * g->od0;
            P2(f->bbcx, j, i) = temp * (ub * g->cag + P2(g->snlat, j, i) * vb * g->sag);
            P2(f->bbcy, j, i) = temp * (vb * g->cag - P2(g->snlat, j, i) * ub * g->sag);
        }
}

/* bclinc11_cu (bclinc.cpp:64-104): the hydrostatic pressure at T points
 * (wka1) and the free-surface term (wka2, copied to wka) */
static void bclinc11(const orc_bclinc_grid *g, const orc_bclinc_fields *f, int isc, double *wka1,
                     double *wka2) {
    const int imt = g->imt, jmt = g->jmt, km = g->km;
    double *wkk = zalloc((size_t)km + 1);
    for (int j = 0; j < jmt; ++j)
        for (int i = 0; i < imt; ++i) {
            double aa = 0;
            if (isc != 0) aa = 0.5;
            P2(f->h0bf, j, i) *= g->onbb;
            const double work = aa * P2(f->h0bf, j, i) + (1.0 - aa) * P2(f->h0bl, j, i);
            wkk[0] = (P2(f->psa, j, i) * g->od0 + work * 9.806e0) * P3(f->vit, 0, j, i);
            for (int k = 0; k < km; ++k)
                wkk[k + 1] = wkk[k] - P3(f->gg, k, j, i) * g->dzp[k] * P3(f->vit, k, j, i);
            for (int k = 0; k < km; ++k) {
                P3(wka1, k, j, i) = 0.5 * (wkk[k] + wkk[k + 1]);
                P3(wka2, k, j, i) = (1.0 + P2(g->ohbt, j, i) * g->zkt[k]) * work * P3(f->vit, k, j, i);
                P3(f->wka, k, j, i) = P3(wka2, k, j, i);
            }
        }
    free(wkk);
}

/* the first part of bclinc2_cu / bclinc6_cu (bclinc.cpp:116-177, 429-480):
 * pressure gradient and Coriolis, semi-implicit */
static void bclinc_grad(const orc_bclinc_grid *g, const orc_bclinc_fields *f, int isc, const double *wka1,
                        const double *wka2) {
    const int imt = g->imt, jmt = g->jmt, km = g->km;
#pragma omp parallel for
    for (int k = 0; k < km; ++k)
        for (int j = 0; j < jmt - 1; ++j)
            for (int i = 1; i < imt; ++i) {
                const int kmb = P2(g->kmu, j, i) - 1;
                double gradx = P2(g->fcor, j, i) * P3(f->vp, k, j, i);
                double grady = -P2(g->fcor, j, i) * P3(f->up, k, j, i);
                if (kmb >= k) {
                    const double ggu = 0.25 * (P3(f->gg, k, j, i) + P3(f->gg, k, j, i - 1) +
                                               P3(f->gg, k, j + 1, i) + P3(f->gg, k, j + 1, i - 1));
                    gradx += P2(g->dxur, j, i) * 0.5 *
                             (P3(wka1, k, j + 1, i) - P3(wka1, k, j, i - 1) - P3(wka1, k, j + 1, i - 1) +
                              P3(wka1, k, j, i));
                    grady += P2(g->dyur, j, i) * 0.5 *
                             (P3(wka1, k, j + 1, i) - P3(wka1, k, j, i - 1) + P3(wka1, k, j + 1, i - 1) -
                              P3(wka1, k, j, i));
                    gradx -= ggu * P2(g->dxur, j, i) * 0.5 *
                             (P3(wka2, k, j + 1, i) - P3(wka2, k, j, i - 1) - P3(wka2, k, j + 1, i - 1) +
                              P3(wka2, k, j, i));
                    grady -= ggu * P2(g->dyur, j, i) * 0.5 *
                             (P3(wka2, k, j + 1, i) - P3(wka2, k, j, i - 1) + P3(wka2, k, j + 1, i - 1) -
                              P3(wka2, k, j, i));
                }
                grady -= P3(f->dlv, k, j, i);
                gradx -= P3(f->dlu, k, j, i);
                double wk1, wk2;
                if (isc == 0) {
                    wk1 = -(P2(g->epea, j, i) * grady + P2(g->epeb, j, i) * gradx);
                    wk2 = -(P2(g->epea, j, i) * gradx - P2(g->epeb, j, i) * grady);
                } else {
                    wk1 = -(P2(g->epla, j, i) * grady + P2(g->eplb, j, i) * gradx);
                    wk2 = -(P2(g->epla, j, i) * gradx - P2(g->eplb, j, i) * grady);
                }
                P3(f->dlv, k, j, i) = wk1 * P3(f->viv, k, j, i);
                P3(f->dlu, k, j, i) = wk2 * P3(f->viv, k, j, i);
            }
}

/* the implicit vertical viscosity of one U column (bclinc3_cu :197-259,
 * bclinc4_cu :301-374, bclinc71_cu :498-576, bclinc72_cu :589-667): x holds
 * the explicit update at levels 0..kz-1 and receives the solution there */
static void tri(const orc_bclinc_grid *g, const orc_bclinc_fields *f, int j, int i, int kz, double dt,
                double *x, double sbc, double bbc) {
    const int imt = g->imt, jmt = g->jmt, km = g->km;
    const double *akmu = f->akmu, *odzt = g->odzt, *odzp = g->odzp;
    double a8[km + 1], b8[km + 1], c8[km + 1], d8[km + 1], e8[km + 1], f8[km + 1];
    for (int k = 1; k < kz; ++k) {
        a8[k] = P3(akmu, k - 1, j, i) * odzt[k] * odzp[k] * dt * 0.5;
        d8[k] = P3(x, k, j, i);
    }
    for (int k = 1; k < kz - 1; ++k) {
        c8[k] = P3(akmu, k, j, i) * odzt[k + 1] * odzp[k] * dt * 0.5;
        b8[k] = 1.0 + a8[k] + c8[k];
        e8[k] = 0.0;
        f8[k] = 0.0;
    }
    /* B. C. at top */
    a8[0] = odzp[0] * dt * 0.5;
    c8[0] = P3(akmu, 0, j, i) * odzt[1] * odzp[0] * dt * 0.5;
    b8[0] = 1.0 + c8[0];
    d8[0] = P3(x, 0, j, i);
    e8[0] = 0.0;
    f8[0] = 0.0;
    /* B. C. at bottom */
    b8[kz - 1] = 1.0 + a8[kz - 1];
    c8[kz - 1] = odzp[kz - 1] * dt * 0.5;
    e8[kz] = 0.0;
    f8[kz] = 0.0;
    d8[kz - 1] = P3(x, kz - 1, j, i) - bbc * odzp[kz - 1] * dt * 0.5;
    /* invert */
    for (int k = kz - 1; k >= 0; --k) {
        const double g0 = 1.0 / (b8[k] - c8[k] * e8[k + 1]);
        e8[k] = a8[k] * g0;
        f8[k] = (d8[k] + c8[k] * f8[k + 1]) * g0;
    }
    /* B. C. at surface */
    P3(x, 0, j, i) = (e8[0] * sbc + f8[0]) * P3(f->viv, 0, j, i);
    for (int k = 1; k < kz; ++k)
        P3(x, k, j, i) = (e8[k] * P3(x, k - 1, j, i) + f8[k]) * P3(f->viv, k, j, i);
}

/* first-guess x = y + dl * dt over the interior columns, all levels */
static void explicit_step(const orc_bclinc_grid *g, double *x, const double *y, const double *dl, double dt) {
    const int imt = g->imt, jmt = g->jmt, km = g->km;
    for (int k = 0; k < km; ++k)
        for (int j = 2; j < jmt - 2; ++j)
            for (int i = 2; i < imt - 2; ++i) P3(x, k, j, i) = P3(y, k, j, i) + P3(dl, k, j, i) * dt;
}

static void solve_interior(const orc_bclinc_grid *g, const orc_bclinc_fields *f, double dt, double *x,
                           const double *sbc, const double *bbc) {
    const int imt = g->imt, jmt = g->jmt;
#pragma omp parallel for
    for (int j = 2; j < jmt - 2; ++j)
        for (int i = 2; i < imt - 2; ++i) {
            const int kz = P2(g->kmu, j, i);
            if (kz > 0) tri(g, f, j, i, kz, dt, x, P2(sbc, j, i), P2(bbc, j, i));
        }
}

/* (x - depth mean of x + barotropic) * viv, column by column, with the sum
 * in bclinc4_cu / bclinc5_cu's order (dzp * ohbu * x * viv) or in
 * bclinc81_cu / bclinc82_cu's (dzp * ohbu * viv * x) */
static double depth_sum(const orc_bclinc_grid *g, const double *x, const double *viv, int j, int i, int order81) {
    const int imt = g->imt, jmt = g->jmt, km = g->km;
    double work = 0;
    for (int k = 0; k < km; ++k) {
        if (order81)
            work += g->dzp[k] * P2(g->ohbu, j, i) * P3(viv, k, j, i) * P3(x, k, j, i);
        else
            work += g->dzp[k] * P2(g->ohbu, j, i) * P3(x, k, j, i) * P3(viv, k, j, i);
    }
    return work;
}

void orc_bclinc_stage(const orc_bclinc_grid *g, const orc_bclinc_fields *f, int isc, int stage) {
    const int imt = g->imt, jmt = g->jmt, km = g->km;
    const size_t n3 = (size_t)imt * jmt * km;
    if (stage == 0) {
        double *wka1 = zalloc(n3), *wka2 = zalloc(n3);
        bclinc12(g, f);
        bclinc11(g, f, isc, wka1, wka2);
        bclinc_grad(g, f, isc, wka1, wka2);
        if (isc < 1) {
            /* bclinc71_cu, bclinc72_cu */
            explicit_step(g, f->u, f->up, f->dlu, g->dtc);
            solve_interior(g, f, g->dtc, f->u, f->sbcx, f->bbcx);
            explicit_step(g, f->v, f->vp, f->dlv, g->dtc);
            solve_interior(g, f, g->dtc, f->v, f->sbcy, f->bbcy);
        } else {
            /* the rest of bclinc2_cu (:179-184) and bclinc3_cu */
            explicit_step(g, f->wka, f->vp, f->dlv, g->dtc2);
            solve_interior(g, f, g->dtc2, f->wka, f->sbcy, f->bbcy);
        }
        free(wka1);
        free(wka2);
    } else if (stage == 1) {
        if (isc < 1) {
            /* bclinc81_cu, bclinc82_cu (:670-726) on the halo-updated u, v */
            for (int pass = 0; pass < 2; ++pass) {
                double *x = pass ? f->v : f->u, *tf = pass ? f->vtf : f->utf;
                const double *b = pass ? f->vb : f->ub;
#pragma omp parallel for
                for (int j = 0; j < jmt; ++j)
                    for (int i = 0; i < imt; ++i) {
                        const double work = depth_sum(g, x, f->viv, j, i, 1);
                        for (int k = 0; k < km; ++k) {
                            P3(x, k, j, i) = (P3(x, k, j, i) - work + P2(b, j, i)) * P3(f->viv, k, j, i);
                            P3(tf, k, j, i) += P3(x, k, j, i);
                        }
                    }
            }
        } else {
            /* bclinc4_cu (:266-379) */
#pragma omp parallel for
            for (int j = 0; j < jmt; ++j)
                for (int i = 0; i < imt; ++i) {
                    const double work = depth_sum(g, f->wka, f->viv, j, i, 0);
                    for (int k = 0; k < km; ++k) {
                        P3(f->wka, k, j, i) = (P3(f->wka, k, j, i) - work + P2(f->vb, j, i)) * P3(f->viv, k, j, i);
                        P3(f->vp, k, j, i) = g->afc2 * P3(f->v, k, j, i) + g->afc1 * (P3(f->vp, k, j, i) + P3(f->wka, k, j, i));
                        P3(f->v, k, j, i) = P3(f->wka, k, j, i);
                    }
                }
            explicit_step(g, f->wka, f->up, f->dlu, g->dtc2);
            solve_interior(g, f, g->dtc2, f->wka, f->sbcx, f->bbcx);
        }
    } else if (stage == 2 && isc >= 1) {
        /* bclinc5_cu (:381-416) */
#pragma omp parallel for
        for (int j = 0; j < jmt; ++j)
            for (int i = 0; i < imt; ++i) {
                const double work = depth_sum(g, f->wka, f->viv, j, i, 0);
                for (int k = 0; k < km; ++k) {
                    const double w = (P3(f->wka, k, j, i) - work + P2(f->ub, j, i)) * P3(f->viv, k, j, i);
                    P3(f->up, k, j, i) = g->afc2 * P3(f->u, k, j, i) + g->afc1 * (P3(f->up, k, j, i) + w);
                    P3(f->u, k, j, i) = w;
                    P3(f->utf, k, j, i) += w;
                    P3(f->vtf, k, j, i) += P3(f->v, k, j, i);
                }
            }
    }
}
