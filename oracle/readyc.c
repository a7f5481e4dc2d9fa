/* readyc.c -- CPU oracle of readyc_'s vertical-mixing preparation, test
 * infrastructure only.  One loop nest per reference kernel, in launch order,
 * with the reference's operation order (compiled with -ffp-contract=off).
 * See readyc.h for the layout. */
#include "readyc.h"

#include <string.h>

#define P2(a, j, i) (a)[(size_t)(j) * imt + (i)]
#define P3(a, k, j, i) (a)[((size_t)(k) * jmt + (j)) * imt + (i)]

void orc_readyc_shear(const orc_readyc_grid *g, const orc_readyc_fields *f) {
    const int imt = g->imt, jmt = g->jmt, km = g->km, kmm1 = km - 1;
    const size_t n3 = (size_t)imt * jmt * km;
    /* clears (readyc.cpp:1636-1643) */
    memset(f->s2t, 0, n3 * sizeof(double));
    memset(f->riu, 0, n3 * sizeof(double));
    memset(f->wp12, 0, n3 * sizeof(double));
    memset(f->wp13, 0, n3 * sizeof(double));
    memset(f->riv1, 0, n3 * sizeof(double));
    memset(f->riv2, 0, n3 * sizeof(double));

    /* readyc1 (readyc.cpp:44-107) */
    for (int j = 0; j < jmt; ++j)
        for (int i = 0; i < imt; ++i) {
            P2(f->h0bl, j, i) = P2(f->h0bf, j, i);
            P2(f->h0bf, j, i) = P2(f->h0, j, i);
        }
#pragma omp parallel for
    for (int k = 0; k < km; ++k)
        for (int j = 1; j < jmt; ++j)
            for (int i = 0; i < imt - 1; ++i) {
                const double den = P3(f->viv, k, j, i) * P2(g->at0, j, i) +
                                   P3(f->viv, k, j - 1, i) * P2(g->atn, j, i) +
                                   P3(f->viv, k, j, i + 1) * P2(g->ate, j, i) +
                                   P3(f->viv, k, j - 1, i + 1) * P2(g->atne, j, i) + 1.0e-25;
                P3(f->wp12, k, j, i) =
                    P3(f->vit, k, j, i) *
                    (P2(g->at0, j, i) * P3(f->up, k, j, i) * P3(f->viv, k, j, i) +
                     P2(g->atn, j, i) * P3(f->up, k, j - 1, i) * P3(f->viv, k, j - 1, i) +
                     P2(g->ate, j, i) * P3(f->up, k, j, i + 1) * P3(f->viv, k, j, i + 1) +
                     P2(g->atne, j, i) * P3(f->up, k, j - 1, i + 1) * P3(f->viv, k, j - 1, i + 1)) /
                    den;
                P3(f->wp13, k, j, i) =
                    P3(f->vit, k, j, i) *
                    (P2(g->at0, j, i) * P3(f->vp, k, j, i) * P3(f->viv, k, j, i) +
                     P2(g->atn, j, i) * P3(f->vp, k, j - 1, i) * P3(f->viv, k, j - 1, i) +
                     P2(g->ate, j, i) * P3(f->vp, k, j, i + 1) * P3(f->viv, k, j, i + 1) +
                     P2(g->atne, j, i) * P3(f->vp, k, j - 1, i + 1) * P3(f->viv, k, j - 1, i + 1)) /
                    den;
            }

    /* readyc1_2 (readyc.cpp:109-170): riv1/riv2 hold the T-grid differences
     * while s2t and rit are formed, then the U-grid ones */
    const double epsln = 1.0e-25;
#pragma omp parallel for
    for (int k = 0; k < kmm1; ++k)
        for (int j = 0; j < jmt; ++j)
            for (int i = 0; i < imt; ++i) {
                const double oz = g->odzt[k + 1];
                if (1 <= j && i < imt - 1) {
                    P3(f->riv1, k, j, i) = P3(f->wp12, k, j, i) * P3(f->vit, k, j, i) -
                                           P3(f->wp12, k + 1, j, i) * P3(f->vit, k + 1, j, i);
                    P3(f->riv2, k, j, i) = P3(f->wp13, k, j, i) * P3(f->vit, k, j, i) -
                                           P3(f->wp13, k + 1, j, i) * P3(f->vit, k + 1, j, i);
                    P3(f->s2t, k, j, i) = P3(f->vit, k + 1, j, i) *
                                          (P3(f->riv1, k, j, i) * P3(f->riv1, k, j, i) +
                                           P3(f->riv2, k, j, i) * P3(f->riv2, k, j, i)) *
                                          oz * oz;
                    P3(f->rit, k, j, i) =
                        P3(f->vit, k + 1, j, i) * P3(f->rict, k, j, i) / (P3(f->s2t, k, j, i) + epsln);
                }
                P3(f->riv1, k, j, i) = P3(f->up, k, j, i) - P3(f->up, k + 1, j, i);
                P3(f->riv2, k, j, i) = P3(f->vp, k, j, i) - P3(f->vp, k + 1, j, i);
                P3(f->s2u, k, j, i) = P3(f->viv, k + 1, j, i) *
                                      (P3(f->riv1, k, j, i) * P3(f->riv1, k, j, i) +
                                       P3(f->riv2, k, j, i) * P3(f->riv2, k, j, i)) *
                                      oz * oz;
                P3(f->riu, k + 1, j, i) =
                    P3(f->viv, k + 1, j, i) * P3(f->ric, k, j, i) / (P3(f->s2u, k, j, i) + epsln);
            }
}

void orc_readyc_akmu(const orc_readyc_grid *g, const double *akmt, const double *viv, double *akmu) {
    const int imt = g->imt, jmt = g->jmt, km = g->km;
    memset(akmu, 0, (size_t)imt * jmt * km * sizeof(double));
#pragma omp parallel for
    for (int k = 0; k < km - 1; ++k)
        for (int j = 0; j < jmt - 1; ++j)
            for (int i = 1; i < imt; ++i)
                P3(akmu, k, j, i) = (P2(g->au0, j, i) * P3(akmt, k, j, i) +
                                     P2(g->aus, j, i) * P3(akmt, k, j + 1, i) +
                                     P2(g->auw, j, i) * P3(akmt, k, j, i - 1) +
                                     P2(g->ausw, j, i) * P3(akmt, k, j + 1, i - 1)) *
                                    P3(viv, k + 1, j, i);
}
