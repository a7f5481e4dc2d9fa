/* readyt.c -- CPU oracle of readyt_ (readyt.cpp:511-598), test infrastructure
 * only.  Each kernel of the reference becomes one loop nest over the cells
 * its guard admits, in launch order, with the reference's operation order
 * (compiled with -ffp-contract=off).  See readyt.h for the layout. */
#include "readyt.h"

#include <string.h>

#include "oracle.h"

#define P2(a, j, i) (a)[(size_t)(j) * imt + (i)]
#define P3(a, k, j, i) (a)[((size_t)(k) * jmt + (j)) * imt + (i)]

void orc_thermal(double tt, double ss, double ppa, double vit, double od0, double *alpha,
                 double *beta) {
    /* thermal_cu (readyt.cpp:272-294), McDougall (1987) fits as in
     * thermalexpansion.F90:41-57: tt in C, ss1 in psu - 35, tm1 in dbar */
    const double tt1 = tt, tt2 = tt1 * tt1, tt3 = tt2 * tt1;
    const double ss1 = ss * 1000.0e0, ss2 = ss1 * ss1;
    const double tm1 = -ppa / od0 / 10000.e0 * vit, tm2 = tm1 * tm1, tm3 = tm2 * tm1;
    const double b =
        (0.785567e-3 - 0.301985e-5 * tt1 + 0.555579e-7 * tt2 - 0.415613e-9 * tt3 +
         ss1 * (-0.356603e-6 + 0.788212e-8 * tt1 + 0.408195e-10 * tm1 - 0.602281e-15 * tm2) +
         ss2 * (0.515032e-8) + tm1 * (-0.121555e-7 + 0.192867e-9 * tt1 - 0.2131127e-11 * tt2) +
         tm2 * (0.176621e-12 - 0.175379e-14 * tt1) + tm3 * (0.12155e-17)) *
        vit;
    const double a =
        (0.665157e-1 + 0.170907e-1 * tt1 - 0.203814e-3 * tt2 + 0.298357e-5 * tt3 -
         0.255019e-7 * tt3 * tt1 +
         ss1 * (0.378110e-2 - 0.846960e-4 * tt1 - 0.164759e-6 * tm1 - 0.251520e-11 * tm2) +
         ss2 * (-0.678662e-5) + tm1 * (0.380374e-4 - 0.933746e-6 * tt1 + 0.791325e-8 * tt2) +
         0.512857e-12 * tm2 * tt2 - 0.302285e-13 * tm3) *
        b * vit;
    *alpha = a;
    *beta = b;
}

void orc_readyt(const orc_readyt_grid *g, const orc_readyt_fields *f, int ist) {
    const int imt = g->imt, jmt = g->jmt, km = g->km, kmm1 = km - 1;
    const size_t n2 = (size_t)imt * jmt, n3 = n2 * km;
    const double od0 = g->od0, grav = ORC_READYT_G, p5 = 0.5;
    const double *C = g->eosc;
    const double *atT = f->at, *atS = f->at + n3;
    const double *abT = f->atb + n2, *abS = f->atb + n3 + 2 * n2;   /* level k at plane k+1 */

    /* clears before the launches (readyt.cpp:536-543) */
    memset(f->alpha, 0, n3 * sizeof(double));
    memset(f->beta, 0, n3 * sizeof(double));
    memset(f->rit, 0, n2 * kmm1 * sizeof(double));
    memset(f->ric, 0, n2 * kmm1 * sizeof(double));
    memset(f->rict, 0, n2 * kmm1 * sizeof(double));
    memset(f->ricdttms, 0, n2 * kmm1 * sizeof(double));
    memset(f->ricdt, 0, n2 * kmm1 * sizeof(double));
    memset(f->akt, 0, 2 * n3 * sizeof(double));

    /* readyt_cu_1 (readyt.cpp:106-158): time levels, T/S on the U grid */
    for (int j = 0; j < jmt; ++j)
        for (int i = 0; i < imt; ++i) {
            P2(f->h0l, j, i) = P2(f->h0f, j, i);
            P2(f->h0f, j, i) = P2(f->h0, j, i);
        }
#pragma omp parallel for
    for (int k = 0; k < km; ++k)
        for (int j = 0; j < jmt; ++j)
            for (int i = 0; i < imt; ++i) {
                P3(f->utl, k, j, i) = P3(f->utf, k, j, i);
                P3(f->vtl, k, j, i) = P3(f->vtf, k, j, i);
                P3(f->utf, k, j, i) = P3(f->u, k, j, i);
                P3(f->vtf, k, j, i) = P3(f->v, k, j, i);
                if (j < jmt - 1 && i >= 1) {
                    P3(f->dlu, k, j, i) = P2(g->au0, j, i) * P3(abT, k, j, i) +
                                          P2(g->aus, j, i) * P3(abT, k, j + 1, i) +
                                          P2(g->auw, j, i) * P3(abT, k, j, i - 1) +
                                          P2(g->ausw, j, i) * P3(abT, k, j + 1, i - 1);
                    P3(f->dlv, k, j, i) = P2(g->au0, j, i) * P3(abS, k, j, i) +
                                          P2(g->aus, j, i) * P3(abS, k, j + 1, i) +
                                          P2(g->auw, j, i) * P3(abS, k, j, i - 1) +
                                          P2(g->ausw, j, i) * P3(abS, k, j + 1, i - 1);
                } else {
                    P3(f->dlu, k, j, i) = 0.0;
                    P3(f->dlv, k, j, i) = 0.0;
                }
            }

    /* dens_cu (readyt.cpp:37-77): Richardson-number buoyancy on the U grid
     * (ric) and on the T grid (rict), densities of both levels referred to
     * level k+1 */
#pragma omp parallel for
    for (int k = 0; k < kmm1; ++k)
        for (int j = 0; j < jmt - 1; ++j)
            for (int i = 1; i < imt; ++i) {
                double tup = P3(f->dlu, k, j, i) - g->to[k + 1];
                double sup = P3(f->dlv, k, j, i) - g->so[k + 1];
                double tlo = P3(f->dlu, k + 1, j, i) - g->to[k + 1];
                double slo = P3(f->dlv, k + 1, j, i) - g->so[k + 1];
                double rhoup = orc_dens(C, km, k + 1, tup, sup);
                double rholo = orc_dens(C, km, k + 1, tlo, slo);
                P3(f->ric, k, j, i) = P3(f->viv, k + 1, j, i) * od0 * grav * (rholo - rhoup) * g->odzt[k + 1];
                tup = P3(abT, k, j, i) - g->to[k + 1];
                sup = P3(abS, k, j, i) - g->so[k + 1];
                tlo = P3(abT, k + 1, j, i) - g->to[k + 1];
                slo = P3(abS, k + 1, j, i) - g->so[k + 1];
                rhoup = orc_dens(C, km, k + 1, tup, sup);
                rholo = orc_dens(C, km, k + 1, tlo, slo);
                P3(f->rict, k, j, i) = P3(f->vit, k + 1, j, i) * od0 * grav * (rholo - rhoup) * g->odzt[k + 1];
            }

    /* readyt_cu_2 (readyt.cpp:160-214): potential density, its weight, T and
     * S interpolated to the top interface */
#pragma omp parallel for
    for (int k = 0; k < km; ++k)
        for (int j = 0; j < jmt; ++j)
            for (int i = 0; i < imt; ++i) {
                if (P3(f->vit, k, j, i) > 0.0) {
                    const double tq = P3(atT, k, j, i) - g->to[k];
                    const double sq = P3(atS, k, j, i) - g->so[k];
                    P3(f->pdensity, k, j, i) = 1.0e+3 + g->po[k] + orc_dens(C, km, k, tq, sq);
                } else {
                    P3(f->pdensity, k, j, i) = 0.0;
                }
                P3(f->gg, k, j, i) = -od0 * grav * P3(f->pdensity, k, j, i) * P3(f->vit, k, j, i);
                if (k >= 1) {
                    const double w = g->dzp[k - 1], s = g->dzp[k - 1] + g->dzp[k];
                    P3(f->ppb, k, j, i) =
                        P3(f->vit, k, j, i) *
                        (P3(atT, k - 1, j, i) - (P3(atT, k - 1, j, i) - P3(atT, k, j, i)) * w / s);
                    P3(f->ppc, k, j, i) =
                        P3(f->vit, k, j, i) *
                        (P3(atS, k - 1, j, i) - (P3(atS, k - 1, j, i) - P3(atS, k, j, i)) * w / s);
                }
            }

    /* readyt_cu_2_2 (readyt.cpp:216-256): hydrostatic integrals */
#pragma omp parallel for
    for (int j = 0; j < jmt; ++j)
        for (int i = 0; i < imt; ++i) {
            P3(f->pp, 0, j, i) = 0.5e0 * g->dzp[0] * P3(f->gg, 0, j, i) * P3(f->vit, 0, j, i);
            P3(f->ppa, 0, j, i) = P2(f->psa, j, i) * P3(f->vit, 0, j, i);
            P3(f->ppb, 0, j, i) = P3(atT, 0, j, i) * P3(f->vit, 0, j, i);
            P3(f->ppc, 0, j, i) = P3(atS, 0, j, i) * P3(f->vit, 0, j, i);
            for (int k = 1; k < km; ++k) {
                P3(f->pp, k, j, i) =
                    P3(f->vit, k, j, i) *
                    (P3(f->pp, k - 1, j, i) +
                     0.5e0 * (P3(f->gg, k, j, i) * g->dzp[k] + P3(f->gg, k - 1, j, i) * g->dzp[k - 1]));
                P3(f->ppa, k, j, i) =
                    P3(f->vit, k, j, i) * (P3(f->ppa, k - 1, j, i) + P3(f->gg, k - 1, j, i) * g->dzp[k - 1]);
            }
        }

    /* thermal_cu (readyt.cpp:258-297) */
#pragma omp parallel for
    for (int k = 0; k < km; ++k)
        for (int j = 0; j < jmt; ++j)
            for (int i = 1; i < imt - 1; ++i)
                orc_thermal(P3(f->ppb, k, j, i), P3(f->ppc, k, j, i), P3(f->ppa, k, j, i),
                            P3(f->vit, k, j, i), od0, &P3(f->alpha, k, j, i), &P3(f->beta, k, j, i));

    /* readyt_cu_3 (readyt.cpp:299-385) */
    const double epsln = 1.0e-25;
    for (int j = 0; j < jmt; ++j)
        for (int i = 0; i < imt; ++i) {
            P2(f->buoytur, j, i) = P3(f->vit, 0, j, i) * P2(f->nswv, j, i) * P3(f->alpha, 0, j, i) * grav * g->od0cp;
            P2(f->buoysol, j, i) = P3(f->vit, 0, j, i) * P2(f->swv, j, i) * P3(f->alpha, 0, j, i) * grav * g->od0cp;
        }
#pragma omp parallel for
    for (int k = 0; k < km; ++k)
        for (int j = 0; j < jmt; ++j)
            for (int i = 0; i < imt; ++i) {
                if (k < kmm1 && 1 <= i && i < imt - 1) {
                    const double dT = P3(atT, k, j, i) - P3(atT, k + 1, j, i);
                    const double dS = P3(atS, k, j, i) - P3(atS, k + 1, j, i);
                    P3(f->ricdttms, k, j, i) =
                        P3(f->vit, k + 1, j, i) * grav *
                        (dT * P3(f->alpha, k + 1, j, i) + 1000.0e0 * dS * P3(f->beta, k + 1, j, i)) *
                        g->odzt[k + 1];
                    P3(f->ricdt, k, j, i) = P3(f->vit, k + 1, j, i) /
                                            ((P3(atT, k, j, i) - P3(atT, k + 1, j, i) + epsln) *
                                             P3(f->alpha, k + 1, j, i)) *
                                            1000.0e0 * dS * P3(f->beta, k + 1, j, i);
                }
                P3(f->gg, k, j, i) =
                    -od0 * grav * (P3(f->pdensity, k, j, i) - g->po[k] - 1000.0e0) * P3(f->vit, k, j, i);
                P3(f->dlu, k, j, i) = 0.0;
                P3(f->dlv, k, j, i) = 0.0;
                if (1 <= i && j < jmt - 1 && k <= P2(g->kmu, j, i) - 1) {
                    P3(f->dlu, k, j, i) = P2(g->dxur, j, i) * p5 *
                                          (P3(f->pp, k, j + 1, i) - P3(f->pp, k, j, i - 1) -
                                           P3(f->pp, k, j + 1, i - 1) + P3(f->pp, k, j, i));
                    P3(f->dlv, k, j, i) = P2(g->dyur, j, i) * p5 *
                                          (P3(f->pp, k, j + 1, i) - P3(f->pp, k, j, i - 1) +
                                           P3(f->pp, k, j + 1, i - 1) - P3(f->pp, k, j, i));
                }
            }

    /* readyt_cu_3_2 (readyt.cpp:387-427): barotropic pressure-gradient
     * integrals, then dlu's planes 0 and 1 reused for the column integrals of
     * gg */
#pragma omp parallel for
    for (int j = 0; j < jmt; ++j)
        for (int i = 0; i < imt; ++i) {
            double *px = &P2(f->pxb, j, i), *py = &P2(f->pyb, j, i);
            *px = 0.0;
            *py = 0.0;
            for (int k = 0; k < km; ++k) {
                *px += g->dzp[k] * P2(g->ohbu, j, i) * P3(f->dlu, k, j, i) * P3(f->viv, k, j, i);
                *py += g->dzp[k] * P2(g->ohbu, j, i) * P3(f->dlv, k, j, i) * P3(f->viv, k, j, i);
            }
            P3(f->dlu, 0, j, i) = 0.0e0;
            P3(f->dlu, 1, j, i) = 0.0e0;
            for (int k = 0; k < km; ++k) {
                const double abcd = P3(f->gg, k, j, i) * P2(g->ohbt, j, i) * g->dzp[k];
                P3(f->dlu, 0, j, i) += abcd;
                P3(f->dlu, 1, j, i) += abcd * g->zkt[k];
            }
        }

    /* readyt_cu_4 (readyt.cpp:429-448) */
    for (int j = 0; j < jmt; ++j)
        for (int i = 0; i < imt; ++i) {
            P3(f->dlv, 0, j, i) = (P3(f->dlu, 0, j, i) + P3(f->dlu, 1, j, i) * P2(g->ohbt, j, i)) / grav;
            P3(f->dlv, 1, j, i) = P3(f->dlu, 1, j, i) * P2(g->ohbt, j, i) * P2(g->ohbt, j, i);
        }

    /* readyt_cu_5 (readyt.cpp:450-509) */
    for (int j = 0; j < jmt; ++j)
        for (int i = 0; i < imt; ++i) {
            if (j < jmt - 1 && 1 <= i) {
                P2(f->wgp, j, i) = P2(g->au0, j, i) * P3(f->dlv, 0, j, i) + P2(g->aus, j, i) * P3(f->dlv, 0, j + 1, i) +
                                   P2(g->auw, j, i) * P3(f->dlv, 0, j, i - 1) +
                                   P2(g->ausw, j, i) * P3(f->dlv, 0, j + 1, i - 1);
                P2(f->work, j, i) = P2(g->au0, j, i) * P3(f->dlv, 1, j, i) + P2(g->aus, j, i) * P3(f->dlv, 1, j + 1, i) +
                                    P2(g->auw, j, i) * P3(f->dlv, 1, j, i - 1) +
                                    P2(g->ausw, j, i) * P3(f->dlv, 1, j + 1, i - 1);
            } else {
                P2(f->wgp, j, i) = 0.0e0;
                P2(f->work, j, i) = 0.0e0;
            }
            P2(f->whx, j, i) = P2(g->hbx, j, i) * P2(f->work, j, i) * P3(f->viv, 0, j, i);
            P2(f->why, j, i) = P2(g->hby, j, i) * P2(f->work, j, i) * P3(f->viv, 0, j, i);
            P2(f->work1, j, i) = 0.0;
            P2(f->work2, j, i) = 0.0;
            if (1 <= i && j < jmt - 1 && 0 <= P2(g->kmu, j, i) - 1) {
                P2(f->work1, j, i) = P2(g->dxur, j, i) * p5 *
                                     (P2(f->psa, j + 1, i) - P2(f->psa, j, i - 1) - P2(f->psa, j + 1, i - 1) +
                                      P2(f->psa, j, i));
                P2(f->work2, j, i) = P2(g->dyur, j, i) * p5 *
                                     (P2(f->psa, j + 1, i) - P2(f->psa, j, i - 1) + P2(f->psa, j + 1, i - 1) -
                                      P2(f->psa, j, i));
            }
            P2(f->pay, j, i) = -od0 * P2(f->work2, j, i);
            P2(f->pax, j, i) = -od0 * P2(f->work1, j, i);
        }

    /* ax, ay, az start the diagnostic period at zero (readyt.cpp:591-595) */
    if (ist == 0) {
        if (f->ax) memset(f->ax, 0, 2 * n3 * sizeof(double));
        if (f->ay) memset(f->ay, 0, 2 * n3 * sizeof(double));
        if (f->az) memset(f->az, 0, 2 * n3 * sizeof(double));
    }
}
