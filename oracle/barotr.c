/* barotr.c -- CPU oracle of barotr_'s subcycle, test infrastructure only.
 * One loop nest per reference kernel, in launch order, with the reference's
 * operation order (compiled with -ffp-contract=off).  See barotr.h. */
#include "barotr.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define P2(a, j, i) (a)[(size_t)(j) * imt + (i)]
#define W(p, j, i) f->wka[((size_t)(p) * jmt + (j)) * imt + (i)]

static const double grav = 9.806e0;   /* constant_mod.h's g */

static double *zalloc(size_t n) {
    double *p = (double *)calloc(n ? n : 1, sizeof(double));
    if (!p) abort();
    return p;
}

/* the U-point average of h0 (barotr1_cu :40-49, barotr2_cu :207-215) */
static double h0_at_u(const orc_barotr_grid *g, const double *h0, int j, int i) {
    const int imt = g->imt, jmt = g->jmt;
    if (j == jmt - 1 || i == 0) return 0.0e0;
    return P2(g->au0, j, i) * P2(h0, j, i) + P2(g->aus, j, i) * P2(h0, j + 1, i) +
           P2(g->auw, j, i) * P2(h0, j, i - 1) + P2(g->ausw, j, i) * P2(h0, j + 1, i - 1);
}

/* the divergence of the transports in wka planes 0, 1 at a T point */
static double tdiv(const orc_barotr_grid *g, const orc_barotr_fields *f, int j, int i) {
    const int imt = g->imt, jmt = g->jmt;
    return 0.5 * ((W(0, j, i + 1) + W(0, j - 1, i + 1)) * P2(g->htw, j, i + 1) -
                  (W(0, j, i) + W(0, j - 1, i)) * P2(g->htw, j, i) +
                  (W(1, j, i + 1) + W(1, j, i)) * P2(g->hts, j, i) -
                  (W(1, j - 1, i + 1) + W(1, j - 1, i)) * P2(g->hts, j - 1, i));
}

static double lap_du(const orc_barotr_grid *g, const double *a, int j, int i) {
    const int imt = g->imt;
    return (P2(g->duc, j, i) + P2(g->dum, j, i)) * P2(a, j, i) + P2(g->dun, j, i) * P2(a, j - 1, i) +
           P2(g->dus, j, i) * P2(a, j + 1, i) + P2(g->due, j, i) * P2(a, j, i + 1) +
           P2(g->duw, j, i) * P2(a, j, i - 1);
}
static double lap_dm(const orc_barotr_grid *g, const double *b, int j, int i) {
    const int imt = g->imt;
    return P2(g->dmc, j, i) * P2(b, j, i) + P2(g->dmn, j, i) * P2(b, j - 1, i) + P2(g->dms, j, i) * P2(b, j + 1, i) +
           P2(g->dme, j, i) * P2(b, j, i + 1) + P2(g->dmw, j, i) * P2(b, j, i - 1);
}

/* bhdel4_grad (barotr.cpp:310-338) */
static void grad2(const orc_barotr_grid *g, const double *w, double *amfac) {
    const int imt = g->imt, jmt = g->jmt;
    for (int j = 0; j < jmt; ++j)
        for (int i = 0; i < imt; ++i) {
            double gx = 0.0, gy = 0.0;
            if (1 <= i && j < jmt - 1 && 0 <= P2(g->kmu, j, i) - 1) {
                gx = P2(g->dxur, j, i) * 0.5 * (P2(w, j + 1, i) - P2(w, j, i - 1) - P2(w, j + 1, i - 1) + P2(w, j, i));
                gy = P2(g->dyur, j, i) * 0.5 * (P2(w, j + 1, i) - P2(w, j, i - 1) + P2(w, j + 1, i - 1) - P2(w, j, i));
            }
            P2(amfac, j, i) += gx * gx + gy * gy;
        }
}

/* the masked 5-point operator of barotr3_cu1 / barotr3_tdel4_1 on x */
static double lap_t(const orc_barotr_grid *g, const double *x, int j, int i, double *cc_out) {
    const int imt = g->imt;
    double hdtk = 0, cc = 0;
    if (0 <= P2(g->kmt, j, i) - 1) {
        double c;
        if (0 <= P2(g->kmtn, j, i) - 1) {
            c = P2(g->dtn, j, i);
            hdtk += c * P2(x, j - 1, i);
            cc -= c;
        }
        if (0 <= P2(g->kmts, j, i) - 1) {
            c = P2(g->dts, j, i);
            hdtk += c * P2(x, j + 1, i);
            cc -= c;
        }
        if (0 <= P2(g->kmte, j, i) - 1) {
            c = P2(g->dte, j, i);
            hdtk += c * P2(x, j, i + 1);
            cc -= c;
        }
        if (0 <= P2(g->kmtw, j, i) - 1) {
            c = P2(g->dtw, j, i);
            hdtk += c * P2(x, j, i - 1);
            cc -= c;
        }
    }
    *cc_out = cc;
    return hdtk;
}

void orc_barotr_stage(const orc_barotr_grid *g, const orc_barotr_fields *f, int nc, int stage) {
    const int imt = g->imt, jmt = g->jmt;
    const int jb = g->jb, je = g->je, ib = g->ib, ie = g->ie;
    const size_t n2 = (size_t)imt * jmt;
    if (stage == 0) {
        /* the clears (barotr.cpp:903-904), barotr1_cu (:32-56), barotr1_cu1 (:58-83) */
        memset(f->wka, 0, 6 * n2 * sizeof(double));
        memset(f->work, 0, n2 * sizeof(double));
        for (int j = 0; j < jmt; ++j)
            for (int i = 0; i < imt; ++i) {
                P2(f->work, j, i) = h0_at_u(g, f->h0, j, i);
                if (i < imt - 1 && j > 0) {
                    W(0, j, i) = P2(f->ub, j, i) * (P2(f->dzph, j, i) + P2(f->work, j, i));
                    W(1, j, i) = P2(f->vb, j, i) * (P2(f->dzph, j, i) + P2(f->work, j, i));
                }
            }
        for (int j = 2; j < jmt - 2; ++j)
            for (int i = 2; i < imt - 2; ++i) {
                if (0 <= P2(g->kmt, j, i) - 1)
                    P2(f->work, j, i) = tdiv(g, f, j, i) * P2(g->tarea_r, j, i) * P2(f->vit, j, i) * -0.250;
                else
                    P2(f->work, j, i) = 0;
            }
    } else if (stage == 1) {
        /* barotr2_cu1 (:85-95) */
        for (size_t p = 0; p < n2; ++p) f->h0[p] = f->h0p[p] + f->work[p] * g->dtb;
        double *d2u = NULL, *d2v = NULL;
        if (g->del4) {
            /* bhdel4_div, bhdel4_grad, bhdel4_zcurl, bhdel4_grad (:272-338),
             * barotr2_udel4_0 (:342-407) */
            double *amfac = zalloc(n2);
            d2u = zalloc(n2);
            d2v = zalloc(n2);
            for (int j = 0; j < jmt; ++j)
                for (int i = 0; i < imt; ++i) {
                    W(0, j, i) = 0.0e0;
                    if (i < imt - 1 && j >= 1 && 0 <= P2(g->kmt, j, i) - 1)
                        W(0, j, i) = 0.5 *
                                     ((P2(f->ubp, j, i + 1) + P2(f->ubp, j - 1, i + 1)) * P2(g->htw, j, i + 1) -
                                      (P2(f->ubp, j, i) + P2(f->ubp, j - 1, i)) * P2(g->htw, j, i) +
                                      (P2(f->vbp, j, i + 1) + P2(f->vbp, j, i)) * P2(g->hts, j, i) -
                                      (P2(f->vbp, j - 1, i + 1) + P2(f->vbp, j - 1, i)) * P2(g->hts, j - 1, i)) *
                                     P2(g->tarea_r, j, i);
                }
            grad2(g, f->wka, amfac);
            for (int j = 0; j < jmt; ++j)
                for (int i = 0; i < imt; ++i) {
                    W(0, j, i) = 0.0e0;
                    if (i >= 1 && j >= 1 && 0 <= P2(g->kmt, j, i) - 1)
                        W(0, j, i) = 0.5 * ((P2(f->vbp, j, i) * P2(g->dyu, j, i) + P2(f->vbp, j - 1, i) * P2(g->dyu, j - 1, i) -
                                             P2(f->vbp, j, i - 1) * P2(g->dyu, j, i - 1) -
                                             P2(f->vbp, j - 1, i - 1) * P2(g->dyu, j - 1, i - 1) -
                                             P2(f->ubp, j, i) * P2(g->dxu, j, i) - P2(f->ubp, j, i - 1) * P2(g->dxu, j, i - 1) +
                                             P2(f->ubp, j - 1, i) * P2(g->dxu, j - 1, i) +
                                             P2(f->ubp, j - 1, i - 1) * P2(g->dxu, j - 1, i - 1)));
                }
            grad2(g, f->wka, amfac);
            for (int j = 0; j < jmt; ++j)
                for (int i = 0; i < imt; ++i) {
                    double hu = 0, hv = 0, am_f = 1.0e0;
                    if (j >= jb - 2 && j <= je && i >= ib - 2 && i <= ie && P2(g->kmu, j, i) - 1 >= 0) {
                        const double t_area = sqrt(1.0e0 / P2(g->uarea_r, j, i));
                        const double dxdy = t_area * t_area * t_area * t_area * t_area * 45.0e0;
                        am_f = sqrt(P2(amfac, j, i)) * dxdy / fabs(g->am * P2(g->amf, j, i));
                        am_f = fmin(40.0e0, am_f);
                        am_f = fmax(1.0e0, am_f);
                        hu = lap_du(g, f->ubp, j, i) + lap_dm(g, f->vbp, j, i);
                        hv = lap_du(g, f->vbp, j, i) - lap_dm(g, f->ubp, j, i);
                    }
                    P2(d2u, j, i) = am_f * P2(g->amf, j, i) * hu;
                    P2(d2v, j, i) = am_f * P2(g->amf, j, i) * hv;
                }
            free(amfac);
        }
        /* barotr2_cu (:98-253) or barotr2_udel4_1 (:410-488) */
        const double *a = g->del4 ? d2u : f->ubp, *b = g->del4 ? d2v : f->vbp;
        for (int j = 0; j < jmt; ++j)
            for (int i = 0; i < imt; ++i) {
                double hu = 0, hv = 0;
                if (P2(g->kmu, j, i) - 1 >= 0 && j >= jb - 1 && j < je && i >= ib - 1 && i < ie) {
                    if (g->del4) {
                        hu = g->am * (lap_du(g, a, j, i) + lap_dm(g, b, j, i));
                        hv = g->am * (lap_du(g, b, j, i) - lap_dm(g, a, j, i));
                    } else {
                        hu = g->am * (lap_du(g, a, j, i) + lap_dm(g, b, j, i)) * P2(f->viv, j, i);
                        hv = g->am * (lap_du(g, b, j, i) - lap_dm(g, a, j, i)) * P2(f->viv, j, i);
                    }
                }
                const int gr = g->del4 ? (i > 1 && i < imt - 2 && j > 1 && j < jmt - 2)
                                       : (i > 0 && i < imt - 1 && j > 1 && j < jmt);
                if (gr) {
                    double gx = 0.0e0, gy = 0.0e0;
                    if (P2(g->kmu, j, i) - 1 >= 0 && j < jmt - 1) {
                        gx = P2(g->dxur, j, i) * 0.5 *
                             (P2(f->h0, j + 1, i) - P2(f->h0, j, i - 1) - P2(f->h0, j + 1, i - 1) + P2(f->h0, j, i));
                        gy = P2(g->dyur, j, i) * 0.5 *
                             (P2(f->h0, j + 1, i) - P2(f->h0, j, i - 1) + P2(f->h0, j + 1, i - 1) - P2(f->h0, j, i));
                    }
                    const double gstar = (P2(f->wgp, j, i) - 1.0) * grav;
                    W(0, j, i) = hu + gstar * gx;
                    W(1, j, i) = hv + gstar * gy;
                    W(4, j, i) = hu;
                    W(5, j, i) = hv;
                }
                P2(f->work, j, i) = h0_at_u(g, f->h0, j, i);
                if (i > 1 && i < imt - 2 && j > 1 && j < jmt - 2) {
                    W(0, j, i) = P2(f->viv, j, i) * (W(0, j, i) + P2(f->dlub, j, i) - P2(g->fcor, j, i) * P2(f->vbp, j, i) +
                                                      P2(f->pax, j, i) + P2(f->pxb, j, i) -
                                                      P2(f->work, j, i) * P2(f->whx, j, i));
                    W(1, j, i) = P2(f->viv, j, i) * (W(1, j, i) + P2(f->dlvb, j, i) + P2(g->fcor, j, i) * P2(f->ubp, j, i) +
                                                      P2(f->pay, j, i) + P2(f->pyb, j, i) -
                                                      P2(f->work, j, i) * P2(f->why, j, i));
                    W(2, j, i) = P2(g->ebea, j, i) * W(0, j, i) - P2(g->ebeb, j, i) * W(1, j, i);
                    W(3, j, i) = P2(g->ebea, j, i) * W(1, j, i) + P2(g->ebeb, j, i) * W(0, j, i);
                }
            }
        free(d2u);
        free(d2v);
    } else if (stage == 2) {
        /* barotr3_cu (:614-634) */
        for (int j = 0; j < jmt; ++j)
            for (int i = 0; i < imt; ++i) {
                P2(f->ub, j, i) = P2(f->ubp, j, i) + W(2, j, i) * g->dtb;
                P2(f->vb, j, i) = P2(f->vbp, j, i) + W(3, j, i) * g->dtb;
                if (j >= 1 && i < imt - 1) {
                    W(0, j, i) = P2(f->ub, j, i) * (P2(f->dzph, j, i) + P2(f->work, j, i));
                    W(1, j, i) = P2(f->vb, j, i) * (P2(f->dzph, j, i) + P2(f->work, j, i));
                }
            }
        const int mix = (nc + 1) % 4 == 0;
        double *d2t = NULL;
        if (g->del4) {
            /* the clear of d2tk and barotr3_tdel4_1 (:735-797) */
            d2t = zalloc(n2);
            for (int j = 1; j < jmt; ++j)
                for (int i = 0; i < imt - 1; ++i) {
                    double hdtk = 0;
                    if (mix) {
                        if (j >= jb - 2 && j <= je && i >= ib - 2 && i <= ie) {
                            double cc;
                            hdtk = lap_t(g, f->h0p, j, i, &cc);
                            hdtk = P2(g->ahf, j, i) * (cc * P2(f->h0p, j, i) + hdtk);
                        } else {
                            hdtk = 0;
                        }
                    }
                    P2(d2t, j, i) = hdtk;
                }
        }
        /* barotr3_cu1 (:636-733) */
        const double *x = g->del4 ? d2t : f->h0p;
        for (int j = 1; j < jmt; ++j)
            for (int i = 0; i < imt - 1; ++i) {
                double div_out = 0, hdtk = 0;
                if (0 <= P2(g->kmt, j, i) - 1) div_out = tdiv(g, f, j, i) * P2(g->tarea_r, j, i);
                if (mix) {
                    if (j >= jb - 1 && j < je && i >= ib - 1 && i < ie) {
                        double cc;
                        hdtk = lap_t(g, x, j, i, &cc);
                        hdtk = g->ah * (cc * P2(x, j, i) + hdtk);
                    } else {
                        hdtk = 0;
                    }
                }
                P2(f->work, j, i) = (hdtk * 1.0e0 - div_out) * P2(f->vit, j, i);
            }
        free(d2t);
    } else if (stage == 3) {
        /* barotr4_cu (:858-882) */
        for (size_t p = 0; p < n2; ++p) {
            const double h0 = f->h0p[p] + (f->work[p] * g->dtb);
            f->h0p[p] = h0;
            f->ubp[p] = f->ub[p];
            f->vbp[p] = f->vb[p];
            f->h0f[p] += h0;
            f->h0bf[p] += h0;
            f->h0[p] = h0;
        }
    }
}
