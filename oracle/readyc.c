/* readyc.c -- CPU oracle of readyc_ (vertical-mixing preparation and the
 * momentum tendencies), test infrastructure only.  One loop nest per reference kernel, in launch order,
 * with the reference's operation order (compiled with -ffp-contract=off).
 * See readyc.h for the layout. */
#include "readyc.h"

#include <math.h>
#include <stdlib.h>
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

/* ------------------------------------------------------------------------
 * Momentum tendencies after readyc2 (readyc.cpp:1700-1793).  3-D scratch
 * arrays the reference keeps on the device (uk, vk, wka, work, u_wface,
 * v_sface, hduk2/hdvk2, am_factor, d2uk/d2vk) are local here. */

static double *zalloc(size_t n) {
    double *p = (double *)calloc(n ? n : 1, sizeof(double));
    if (!p) abort();
    return p;
}

/* upwell_cu_0 + upwell_cu (readyc.cpp:599-682): the barotropically scaled
 * velocities, their divergence at T points, its column integral and the
 * recursion of ws from the surface, ws[0] kept. */
void orc_rm_upwell(const orc_readyc_grid *g, const orc_readyc_mgrid *m, const orc_readyc_mfields *f) {
    const int imt = g->imt, jmt = g->jmt, km = g->km;
    const size_t n2 = (size_t)imt * jmt, n3 = n2 * km;
    double *uk = zalloc(n3), *vk = zalloc(n3), *wka = zalloc(n3), *work = zalloc(n2);
    for (int j = 0; j < jmt - 1; ++j)
        for (int i = 1; i < imt; ++i) {
            P2(work, j, i) = P2(g->au0, j, i) * P2(f->h0, j, i) + P2(g->aus, j, i) * P2(f->h0, j + 1, i) +
                             P2(g->auw, j, i) * P2(f->h0, j, i - 1) +
                             P2(g->ausw, j, i) * P2(f->h0, j + 1, i - 1);
            for (int k = 0; k < km; ++k) {
                P3(uk, k, j, i) = (1.0e0 + P2(work, j, i) * P2(m->ohbu, j, i)) * P3(f->u, k, j, i);
                P3(vk, k, j, i) = (1.0e0 + P2(work, j, i) * P2(m->ohbu, j, i)) * P3(f->v, k, j, i);
            }
        }
    for (int j = 0; j < jmt; ++j)
        for (int i = 0; i < imt; ++i) {
            if (i < imt - 1 && j >= 1)
                for (int k = 0; k < km; ++k) {
                    P3(wka, k, j, i) = 0.0e0;
                    if (k <= P2(m->kmt, j, i) - 1)
                        P3(wka, k, j, i) =
                            0.5 *
                            ((P3(uk, k, j, i + 1) + P3(uk, k, j - 1, i + 1)) * P2(m->htw, j, i + 1) -
                             (P3(uk, k, j, i) + P3(uk, k, j - 1, i)) * P2(m->htw, j, i) +
                             (P3(vk, k, j, i + 1) + P3(vk, k, j, i)) * P2(m->hts, j, i) -
                             (P3(vk, k, j - 1, i + 1) + P3(vk, k, j - 1, i)) * P2(m->hts, j - 1, i)) *
                            P2(m->tarea_r, j, i);
                }
            P2(work, j, i) = 0.0e0;
        }
    for (int j = 1; j < jmt - 1; ++j)
        for (int i = 1; i < imt - 1; ++i) {
            for (int k = 0; k < km; ++k)
                P2(work, j, i) -= m->dzp[k] * P3(wka, k, j, i) * P3(f->vit, k, j, i);
            for (int k = 1; k < km; ++k)
                P3(f->ws, k, j, i) =
                    P3(f->vit, k, j, i) *
                    (P3(f->ws, k - 1, j, i) +
                     m->dzp[k - 1] * (P2(work, j, i) * P2(m->ohbt, j, i) + P3(wka, k - 1, j, i)));
            P2(work, j, i) = 1.0e0 / (1.0e0 + P2(f->h0, j, i) * P2(m->ohbt, j, i));
            for (int k = 1; k < km; ++k) P3(f->ws, k, j, i) *= P2(work, j, i);
        }
    free(uk);
    free(vk);
    free(wka);
    free(work);
}

/* the clears of dlu, dlv, u_wface, v_sface, wka (readyc.cpp:1712-1716),
 * readyc3 (:248-271) and advection_momentum_cu (:684-719) */
void orc_rm_advect(const orc_readyc_grid *g, const orc_readyc_mgrid *m, const orc_readyc_mfields *f) {
    const int imt = g->imt, jmt = g->jmt, km = g->km;
    const size_t n3 = (size_t)imt * jmt * km;
    const double p25 = 0.25, p5 = 0.5;
    double *wka = zalloc(n3), *uw = zalloc(n3), *vs = zalloc(n3);
    memset(f->dlu, 0, n3 * sizeof(double));
    memset(f->dlv, 0, n3 * sizeof(double));
#pragma omp parallel for
    for (int k = 0; k < km; ++k)
        for (int j = 0; j < jmt - 1; ++j)
            for (int i = 1; i < imt; ++i) {
                P3(wka, k, j, i) = P2(g->au0, j, i) * (P3(f->ws, k, j, i)) +
                                   P2(g->aus, j, i) * (P3(f->ws, k, j + 1, i)) +
                                   P2(g->auw, j, i) * (P3(f->ws, k, j, i - 1)) +
                                   P2(g->ausw, j, i) * (P3(f->ws, k, j + 1, i - 1));
                P3(uw, k, j, i) = (P3(f->u, k, j, i - 1) + P3(f->u, k, j, i)) * p25 * P2(m->hue, j, i - 1);
                P3(vs, k, j, i) = (P3(f->v, k, j, i) + P3(f->v, k, j + 1, i)) * p25 * P2(m->hun, j + 1, i);
            }
#pragma omp parallel for
    for (int k = 0; k < km; ++k)
        for (int j = 1; j < jmt - 1; ++j)
            for (int i = 1; i < imt - 1; ++i) {
                const double *u = f->u, *v = f->v;
                double z1 = 0.0e0, z2 = 0.0e0, z3 = 0.0e0, z4 = 0.0e0;
                P3(f->dlu, k, j, i) = (-P3(uw, k, j, i) * (P3(u, k, j, i) - P3(u, k, j, i - 1)) -
                                       P3(uw, k, j, i + 1) * (P3(u, k, j, i + 1) - P3(u, k, j, i)) -
                                       P3(vs, k, j, i) * (P3(u, k, j + 1, i) - P3(u, k, j, i)) -
                                       P3(vs, k, j - 1, i) * (P3(u, k, j, i) - P3(u, k, j - 1, i))) *
                                      P2(m->uarea_r, j, i);
                P3(f->dlv, k, j, i) = (-P3(uw, k, j, i) * (P3(v, k, j, i) - P3(v, k, j, i - 1)) -
                                       P3(uw, k, j, i + 1) * (P3(v, k, j, i + 1) - P3(v, k, j, i)) -
                                       P3(vs, k, j, i) * (P3(v, k, j + 1, i) - P3(v, k, j, i)) -
                                       P3(vs, k, j - 1, i) * (P3(v, k, j, i) - P3(v, k, j - 1, i))) *
                                      P2(m->uarea_r, j, i);
                if (k > 0) {
                    z1 = P3(wka, k, j, i) * (P3(u, k - 1, j, i) - P3(u, k, j, i));
                    z3 = P3(wka, k, j, i) * (P3(v, k - 1, j, i) - P3(v, k, j, i));
                }
                if (k < km - 1) {
                    z2 = P3(wka, k + 1, j, i) * (P3(u, k, j, i) - P3(u, k + 1, j, i));
                    z4 = P3(wka, k + 1, j, i) * (P3(v, k, j, i) - P3(v, k + 1, j, i));
                }
                P3(f->dlu, k, j, i) += -p5 * m->odzp[k] * (z1 + z2);
                P3(f->dlv, k, j, i) += -p5 * m->odzp[k] * (z3 + z4);
            }
    free(wka);
    free(uw);
    free(vs);
}

/* the 5-point operator of hdiffu_* on (a, b) at (k, j, i): the du
 * coefficients on a, the dm coefficients on b (readyc.cpp:289-300) */
static double lap_du(const orc_readyc_grid *g, const orc_readyc_mgrid *m, const double *a, int k, int j,
                     int i) {
    const int imt = g->imt, jmt = g->jmt;
    return (P2(m->duc, j, i) + P2(m->dum, j, i)) * P3(a, k, j, i) + P2(m->dun, j, i) * P3(a, k, j - 1, i) +
           P2(m->dus, j, i) * P3(a, k, j + 1, i) + P2(m->due, j, i) * P3(a, k, j, i + 1) +
           P2(m->duw, j, i) * P3(a, k, j, i - 1);
}
static double lap_dm(const orc_readyc_grid *g, const orc_readyc_mgrid *m, const double *b, int k, int j,
                     int i) {
    const int imt = g->imt, jmt = g->jmt;
    return P2(m->dmc, j, i) * P3(b, k, j, i) + P2(m->dmn, j, i) * P3(b, k, j - 1, i) +
           P2(m->dms, j, i) * P3(b, k, j + 1, i) + P2(m->dme, j, i) * P3(b, k, j, i + 1) +
           P2(m->dmw, j, i) * P3(b, k, j, i - 1);
}

/* hdel4_grad (readyc.cpp:405-441): am_factor += |grad w|^2 at U points */
static void del4_grad(const orc_readyc_grid *g, const orc_readyc_mgrid *m, const double *w, double *amfac) {
    const int imt = g->imt, jmt = g->jmt, km = g->km;
    const double p5 = 0.5;
#pragma omp parallel for
    for (int k = 0; k < km; ++k)
        for (int j = 0; j < jmt; ++j)
            for (int i = 0; i < imt; ++i) {
                double gx = 0.0, gy = 0.0;
                if (1 <= i && j < jmt - 1 && k <= P2(m->kmu, j, i) - 1) {
                    gx = P2(m->dxur, j, i) * p5 *
                         (P3(w, k, j + 1, i) - P3(w, k, j, i - 1) - P3(w, k, j + 1, i - 1) + P3(w, k, j, i));
                    gy = P2(m->dyur, j, i) * p5 *
                         (P3(w, k, j + 1, i) - P3(w, k, j, i - 1) + P3(w, k, j + 1, i - 1) - P3(w, k, j, i));
                }
                P3(amfac, k, j, i) += gx * gx + gy * gy;
            }
}

/* horizontal viscosity: hdiffu_del2_cu (readyc.cpp:273-341), or with del4
 * hdel4_div, am_factor clear, hdel4_grad, hdel4_zcurl, hdel4_grad,
 * hdiffu_del4_0, hdiffu_del4_1 (:343-576, launched :1754-1771); both end
 * with wka = c0f |u| */
void orc_rm_hdiff(const orc_readyc_grid *g, const orc_readyc_mgrid *m, const orc_readyc_mfields *f) {
    const int imt = g->imt, jmt = g->jmt, km = g->km;
    const int jb = m->jb, je = m->je, ib = m->ib, ie = m->ie;
    const size_t n3 = (size_t)imt * jmt * km;
    const double *up = f->up, *vp = f->vp;
    double *d2u = NULL, *d2v = NULL;
    if (m->del4) {
        double *w = zalloc(n3), *amfac = zalloc(n3);
        d2u = zalloc(n3);
        d2v = zalloc(n3);
        /* hdel4_div (:372-402) */
#pragma omp parallel for
        for (int k = 0; k < km; ++k)
            for (int j = 0; j < jmt; ++j)
                for (int i = 0; i < imt; ++i) {
                    P3(w, k, j, i) = 0.0e0;
                    if (i < imt - 1 && j >= 1 && k <= P2(m->kmt, j, i) - 1)
                        P3(w, k, j, i) =
                            0.5 *
                            ((P3(up, k, j, i + 1) + P3(up, k, j - 1, i + 1)) * P2(m->htw, j, i + 1) -
                             (P3(up, k, j, i) + P3(up, k, j - 1, i)) * P2(m->htw, j, i) +
                             (P3(vp, k, j, i + 1) + P3(vp, k, j, i)) * P2(m->hts, j, i) -
                             (P3(vp, k, j - 1, i + 1) + P3(vp, k, j - 1, i)) * P2(m->hts, j - 1, i)) *
                            P2(m->tarea_r, j, i);
                }
        del4_grad(g, m, w, amfac);
        /* hdel4_zcurl (:343-370) */
#pragma omp parallel for
        for (int k = 0; k < km; ++k)
            for (int j = 0; j < jmt; ++j)
                for (int i = 0; i < imt; ++i) {
                    P3(w, k, j, i) = 0.0e0;
                    if (i >= 1 && j >= 1 && k <= P2(m->kmt, j, i) - 1)
                        P3(w, k, j, i) =
                            0.5 * ((P3(vp, k, j, i) * P2(m->dyu, j, i) +
                                    P3(vp, k, j - 1, i) * P2(m->dyu, j - 1, i) -
                                    P3(vp, k, j, i - 1) * P2(m->dyu, j, i - 1) -
                                    P3(vp, k, j - 1, i - 1) * P2(m->dyu, j - 1, i - 1) -
                                    P3(up, k, j, i) * P2(m->dxu, j, i) -
                                    P3(up, k, j, i - 1) * P2(m->dxu, j, i - 1) +
                                    P3(up, k, j - 1, i) * P2(m->dxu, j - 1, i) +
                                    P3(up, k, j - 1, i - 1) * P2(m->dxu, j - 1, i - 1)));
                }
        del4_grad(g, m, w, amfac);
        /* hdiffu_del4_0 (:444-507) */
#pragma omp parallel for
        for (int k = 0; k < km; ++k)
            for (int j = 0; j < jmt; ++j)
                for (int i = 0; i < imt; ++i) {
                    double am_f = 1.0e0, hu = 0.0, hv = 0.0;
                    if (j >= jb - 2 && j <= je && i >= ib - 2 && i <= ie && k <= P2(m->kmu, j, i) - 1) {
                        const double t_area = sqrt(1.0e0 / P2(m->uarea_r, j, i));
                        const double dxdy = t_area * t_area * t_area * t_area * t_area * 45.0e0;
                        am_f = sqrt(P3(amfac, k, j, i)) * dxdy / fabs(m->am * P2(m->amf, j, i));
                        am_f = fmin(40.0e0, am_f);
                        am_f = fmax(1.0e0, am_f);
                        hu = lap_du(g, m, up, k, j, i) + lap_dm(g, m, vp, k, j, i);
                        hv = lap_du(g, m, vp, k, j, i) - lap_dm(g, m, up, k, j, i);
                    }
                    P3(d2u, k, j, i) = am_f * P2(m->amf, j, i) * hu;
                    P3(d2v, k, j, i) = am_f * P2(m->amf, j, i) * hv;
                }
        free(w);
        free(amfac);
    }
    /* hdiffu_del2_cu (:273-341) or hdiffu_del4_1 (:509-576) */
    const double *a = m->del4 ? d2u : up, *b = m->del4 ? d2v : vp;
    const int j0 = jb - 1, j1 = je, i0 = ib - 1, i1 = ie;
#pragma omp parallel for
    for (int k = 0; k < km; ++k)
        for (int j = 0; j < jmt; ++j)
            for (int i = 0; i < imt; ++i) {
                double hu = 0.0, hv = 0.0;
                if (j >= j0 && j < j1 && i >= i0 && i < i1) {
                    if (m->del4) {
                        hu = m->am * (lap_du(g, m, a, k, j, i) + lap_dm(g, m, b, k, j, i));
                        hv = m->am * (lap_du(g, m, b, k, j, i) - lap_dm(g, m, a, k, j, i));
                    } else {
                        hu = m->am * (lap_du(g, m, a, k, j, i) + lap_dm(g, m, b, k, j, i)) * P3(f->viv, k, j, i);
                        hv = m->am * (lap_du(g, m, b, k, j, i) - lap_dm(g, m, a, k, j, i)) * P3(f->viv, k, j, i);
                    }
                }
                if (k > P2(m->kmu, j, i) - 1) hu = hv = 0.0;
                if (1 <= j && j < jmt - 1 && 1 <= i && i < imt - 1) {
                    P3(f->dlv, k, j, i) = P3(f->dlv, k, j, i) + hv;
                    P3(f->dlu, k, j, i) = P3(f->dlu, k, j, i) + hu;
                }
                P3(f->wka, k, j, i) = m->c0f * sqrt(P3(up, k, j, i) * P3(up, k, j, i) +
                                                    P3(vp, k, j, i) * P3(vp, k, j, i));
            }
    free(d2u);
    free(d2v);
}

/* readyc4 (readyc.cpp:721-790): depth means of dlu, dlv, surface stress,
 * bottom drag at the deepest wet U level */
void orc_rm_stress(const orc_readyc_grid *g, const orc_readyc_mgrid *m, const orc_readyc_mfields *f) {
    const int imt = g->imt, jmt = g->jmt, km = g->km;
#pragma omp parallel for
    for (int j = 0; j < jmt; ++j)
        for (int i = 0; i < imt; ++i) {
            P2(f->dlub, j, i) = 0.0;
            P2(f->dlvb, j, i) = 0.0;
            for (int k = 0; k < km; ++k) {
                P2(f->dlub, j, i) += m->dzp[k] * P2(m->ohbu, j, i) * P3(f->dlu, k, j, i) * P3(f->viv, k, j, i);
                P2(f->dlvb, j, i) += m->dzp[k] * P2(m->ohbu, j, i) * P3(f->dlv, k, j, i) * P3(f->viv, k, j, i);
            }
            if (1 <= j && j < jmt - 1 && 1 <= i && i < imt - 1) {
                const int kmb = P2(m->kmu, j, i) - 1;
                if (kmb >= 0) {
                    const double ub = P3(f->up, kmb, j, i), vb = P3(f->vp, kmb, j, i);
                    P2(f->sbcx, j, i) = P2(f->su, j, i) * m->od0;
                    P2(f->sbcy, j, i) = P2(f->sv, j, i) * m->od0;
                    P2(f->bbcx, j, i) =
                        m->c0f * sqrt(ub * ub + vb * vb) * (ub * m->cag + P2(m->snlat, j, i) * vb * m->sag);
                    P2(f->bbcy, j, i) =
                        m->c0f * sqrt(ub * ub + vb * vb) * (-P2(m->snlat, j, i) * ub * m->sag + vb * m->cag);
                } else {
                    P2(f->sbcx, j, i) = 0.0e0;
                    P2(f->sbcy, j, i) = 0.0e0;
                    P2(f->bbcx, j, i) = 0.0e0;
                    P2(f->bbcy, j, i) = 0.0e0;
                }
                P2(f->dlub, j, i) = P2(f->dlub, j, i) + (P2(f->sbcx, j, i) - P2(f->bbcx, j, i)) * P2(m->ohbu, j, i);
                P2(f->dlvb, j, i) = P2(f->dlvb, j, i) + (P2(f->sbcy, j, i) - P2(f->bbcy, j, i)) * P2(m->ohbu, j, i);
            }
        }
}

/* readyc4_2 (readyc.cpp:793-877): the explicit half (1 - aidif) of the
 * vertical viscosity, surface stress at the top, bottom drag below the
 * deepest wet level */
void orc_rm_vdiff(const orc_readyc_grid *g, const orc_readyc_mgrid *m, const orc_readyc_mfields *f) {
    const int imt = g->imt, jmt = g->jmt, km = g->km;
    const double aidif = 0.5, sag = m->sag, cag = m->cag;
    const double *up = f->up, *vp = f->vp, *akmu = f->akmu, *viv = f->viv, *wka = f->wka;
#pragma omp parallel for
    for (int k = 0; k < km; ++k)
        for (int j = 1; j < jmt - 1; ++j)
            for (int i = 1; i < imt - 1; ++i) {
                const double sn = P2(m->snlat, j, i);
                double du1, du2, dv1, dv2;
                if (k == 0) {
                    dv1 = P2(f->sv, j, i) * m->od0 * (1 - aidif);
                    du1 = P2(f->su, j, i) * m->od0 * (1 - aidif);
                } else {
                    dv1 = P3(akmu, k - 1, j, i) * (1 - aidif) * (P3(vp, k - 1, j, i) - P3(vp, k, j, i)) *
                              g->odzt[k] * P3(viv, k, j, i) +
                          (1.0e0 - P3(viv, k, j, i)) * P3(wka, k - 1, j, i) * (1 - aidif) *
                              (-sn * P3(up, k - 1, j, i) * sag + P3(vp, k - 1, j, i) * cag);
                    du1 = P3(akmu, k - 1, j, i) * (1 - aidif) * (P3(up, k - 1, j, i) - P3(up, k, j, i)) *
                              g->odzt[k] * P3(viv, k, j, i) +
                          (1.0e0 - P3(viv, k, j, i)) * P3(wka, k - 1, j, i) * (1 - aidif) *
                              (P3(up, k - 1, j, i) * cag + sn * P3(vp, k - 1, j, i) * sag);
                }
                if (k == km - 1) {
                    dv2 = P3(wka, km - 1, j, i) * (-sn * P3(up, km - 1, j, i) * sag + P3(vp, km - 1, j, i) * cag) *
                          (1 - aidif);
                    du2 = P3(wka, km - 1, j, i) * (P3(up, km - 1, j, i) * cag + sn * P3(vp, km - 1, j, i) * sag) *
                          (1 - aidif);
                } else {
                    dv2 = P3(akmu, k, j, i) * (1 - aidif) * (P3(vp, k, j, i) - P3(vp, k + 1, j, i)) *
                              g->odzt[k + 1] * P3(viv, k + 1, j, i) +
                          (1.0e0 - P3(viv, k + 1, j, i)) * P3(wka, k, j, i) * (1 - aidif) *
                              (-sn * P3(up, k, j, i) * sag + P3(vp, k, j, i) * cag);
                    du2 = P3(akmu, k, j, i) * (1 - aidif) * (P3(up, k, j, i) - P3(up, k + 1, j, i)) *
                              g->odzt[k + 1] * P3(viv, k + 1, j, i) +
                          (1.0e0 - P3(viv, k + 1, j, i)) * P3(wka, k, j, i) * (1 - aidif) *
                              (P3(up, k, j, i) * cag + sn * P3(vp, k, j, i) * sag);
                }
                P3(f->dlv, k, j, i) = P3(f->dlv, k, j, i) + m->odzp[k] * (dv1 - dv2);
                P3(f->dlu, k, j, i) = P3(f->dlu, k, j, i) + m->odzp[k] * (du1 - du2);
            }
}

void orc_readyc_momentum(const orc_readyc_grid *g, const orc_readyc_mgrid *m, const orc_readyc_mfields *f) {
    orc_rm_upwell(g, m, f);
    orc_rm_advect(g, m, f);
    orc_rm_hdiff(g, m, f);
    orc_rm_stress(g, m, f);
    orc_rm_vdiff(g, m, f);
}
