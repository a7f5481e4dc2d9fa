/*
 * oracle.c -- CPU restatement of the LICOM3 tracer time step and convadj.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Every function follows the
 * reference statement cited above it, pass by pass, keeping the reference's
 * floating-point operation order so that a HIP implementation compiled with
 * -ffp-contract=off can be compared bit for bit.  Build with
 * -ffp-contract=off (the Makefile next to this file does).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define I2(j, i) ((size_t)(j) * (size_t)imt + (size_t)(i))
#define I3(k, j, i) (((size_t)(k) * (size_t)jmt + (size_t)(j)) * (size_t)imt + (size_t)(i))

static const double kRadius = 6371000.0;      /* constant_mod.F90:31 */

int orc_nthreads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_nthreads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------ halo */
/* POP_HaloUpdate for one block covering the whole domain with a cyclic
 * east-west boundary and closed north-south boundaries (domain.F90:162-163):
 * ghost columns copy the opposite interior columns, ghost rows of a closed
 * boundary take the fill value 0 (POP_HaloMod.F90:2942-2946). */
void orc_halo_fill_f64(double *a, int imt, int jmt, int nz) {
    for (int z = 0; z < nz; ++z) {
        double *p = a + (size_t)z * jmt * imt;
        for (int j = 2; j < jmt - 2; ++j) {
            p[I2(j, 0)] = p[I2(j, imt - 4)];
            p[I2(j, 1)] = p[I2(j, imt - 3)];
            p[I2(j, imt - 2)] = p[I2(j, 2)];
            p[I2(j, imt - 1)] = p[I2(j, 3)];
        }
        for (int i = 0; i < imt; ++i) {
            p[I2(0, i)] = 0.0;
            p[I2(1, i)] = 0.0;
            p[I2(jmt - 2, i)] = 0.0;
            p[I2(jmt - 1, i)] = 0.0;
        }
    }
}

void orc_halo_fill_i32(int32_t *p, int imt, int jmt) {
    for (int j = 2; j < jmt - 2; ++j) {
        p[I2(j, 0)] = p[I2(j, imt - 4)];
        p[I2(j, 1)] = p[I2(j, imt - 3)];
        p[I2(j, imt - 2)] = p[I2(j, 2)];
        p[I2(j, imt - 1)] = p[I2(j, 3)];
    }
    for (int i = 0; i < imt; ++i) {
        p[I2(0, i)] = 0;
        p[I2(1, i)] = 0;
        p[I2(jmt - 2, i)] = 0;
        p[I2(jmt - 1, i)] = 0;
    }
}

static int imin4(int a, int b, int c, int d) {
    int m = a;
    if (b < m) m = b;
    if (c < m) m = c;
    if (d < m) m = d;
    return m;
}

/* ---------------------------------------------------------- derived grid */
void orc_derive(orc_grid *g) {
    const int imt = g->imt, jmt = g->jmt, km = g->km;
    const size_t n2 = (size_t)imt * jmt;

    /* boundary.F90:29-41: KMT halo, then KMU = min of the four T columns
     * sharing the U point (SW corner of T(i,j); j grows southward). */
    orc_halo_fill_i32(g->kmt, imt, jmt);
    for (size_t q = 0; q < n2; ++q) g->kmu[q] = 0;
    for (int j = 0; j < jmt - 1; ++j)
        for (int i = 1; i < imt; ++i)
            g->kmu[I2(j, i)] = imin4(g->kmt[I2(j, i - 1)], g->kmt[I2(j, i)],
                                     g->kmt[I2(j + 1, i - 1)], g->kmt[I2(j + 1, i)]);
    orc_halo_fill_i32(g->kmu, imt, jmt);

    /* boundary.F90:43-52 */
    for (int k = 0; k < km; ++k)
        for (int j = 0; j < jmt; ++j)
            for (int i = 0; i < imt; ++i) {
                g->vit[I3(k, j, i)] = (k < g->kmt[I2(j, i)]) ? 1.0 : 0.0;
                g->viv[I3(k, j, i)] = (k < g->kmu[I2(j, i)]) ? 1.0 : 0.0;
            }

    /* grid.F90:1096-1106 */
    double *zkt = (double *)malloc(sizeof(double) * km);
    for (int k = 0; k < km; ++k) {
        g->dzp[k] = g->zkp[k] - g->zkp[k + 1];
        zkt[k] = (g->zkp[k] + g->zkp[k + 1]) * 0.5;
        g->odzp[k] = 1.0 / g->dzp[k];
    }
    g->odzt[0] = 2.0 * g->odzp[0];
    for (int k = 1; k < km; ++k) g->odzt[k] = 1.0 / (zkt[k - 1] - zkt[k]);
    free(zkt);

    /* grid.F90:466 */
    for (size_t q = 0; q < n2; ++q) g->tarea_r[q] = 1.0 / g->tarea[q];

    /* grid.F90:2029-2050: inverse column depths at T and U points */
    for (int j = 0; j < jmt; ++j)
        for (int i = 0; i < imt; ++i) {
            double s = 0.0;
            for (int k = 0; k < km; ++k) s = s + g->vit[I3(k, j, i)] * g->dzp[k];
            g->ohbt[I2(j, i)] = (s > 0.0) ? 1.0 / s : 0.0;
            s = 0.0;
            for (int k = 0; k < km; ++k) s = s + g->viv[I3(k, j, i)] * g->dzp[k];
            g->ohbu[I2(j, i)] = (s > 0.0) ? 1.0 / s : 0.0;
        }

    /* hmix_del4.F90:357-392 / hmix_del2.F90:346-385.  eoshift is end-off:
     * a neighbour outside the array contributes 0. */
    const double pi = 4.0 * atan(1.0);            /* constant_mod.F90:25 */
    const double dxg = 2.0 * kRadius * pi / (double)g->imt_global;
    double *w1 = (double *)malloc(sizeof(double) * n2);
    if (g->hmix == 4) {
        for (size_t q = 0; q < n2; ++q)
            g->ahf[q] = pow(g->tarea[q] / pow(dxg, 2.0), 1.5);
        for (size_t q = 0; q < n2; ++q) w1[q] = g->hts[q] / g->hue[q];
        for (int j = 0; j < jmt; ++j)
            for (int i = 0; i < imt; ++i) {
                g->dts[I2(j, i)] = w1[I2(j, i)] * g->tarea_r[I2(j, i)];
                g->dtn[I2(j, i)] = (j >= 1 ? w1[I2(j - 1, i)] : 0.0) * g->tarea_r[I2(j, i)];
            }
        for (size_t q = 0; q < n2; ++q) w1[q] = g->htw[q] / g->hun[q];
        for (int j = 0; j < jmt; ++j)
            for (int i = 0; i < imt; ++i) {
                g->dtw[I2(j, i)] = w1[I2(j, i)] * g->tarea_r[I2(j, i)];
                g->dte[I2(j, i)] = (i + 1 < imt ? w1[I2(j, i + 1)] : 0.0) * g->tarea_r[I2(j, i)];
            }
    } else {
        for (size_t q = 0; q < n2; ++q) g->ahf[q] = sqrt(g->tarea[q]) / dxg;
        for (int j = 0; j < jmt; ++j)
            for (int i = 0; i < imt; ++i) {
                double an = (j + 1 < jmt) ? g->ahf[I2(j + 1, i)] : 0.0;
                w1[I2(j, i)] = g->hts[I2(j, i)] / g->hue[I2(j, i)] * 0.5 * (g->ahf[I2(j, i)] + an);
            }
        for (int j = 0; j < jmt; ++j)
            for (int i = 0; i < imt; ++i) {
                g->dts[I2(j, i)] = w1[I2(j, i)] * g->tarea_r[I2(j, i)];
                g->dtn[I2(j, i)] = (j >= 1 ? w1[I2(j - 1, i)] : 0.0) * g->tarea_r[I2(j, i)];
            }
        for (int j = 0; j < jmt; ++j)
            for (int i = 0; i < imt; ++i) {
                double aw = (i >= 1) ? g->ahf[I2(j, i - 1)] : 0.0;
                w1[I2(j, i)] = g->htw[I2(j, i)] / g->hun[I2(j, i)] * 0.5 * (g->ahf[I2(j, i)] + aw);
            }
        for (int j = 0; j < jmt; ++j)
            for (int i = 0; i < imt; ++i) {
                g->dtw[I2(j, i)] = w1[I2(j, i)] * g->tarea_r[I2(j, i)];
                g->dte[I2(j, i)] = (i + 1 < imt ? w1[I2(j, i + 1)] : 0.0) * g->tarea_r[I2(j, i)];
            }
    }
    free(w1);

    /* grid.F90:651: ocean surface area over the physical (interior) cells */
    double a = 0.0;
    for (int j = 2; j < jmt - 2; ++j)
        for (int i = 2; i < imt - 2; ++i)
            if (g->kmt[I2(j, i)] >= 1) a = a + g->tarea[I2(j, i)];
    g->area_t = a;
}

/* ----------------------------------------------------------- utilities */
static double max6(double a, double b, double c, double d, double e, double f) {
    double m = a;                      /* tracer.cpp:47-82 (min1/max1)   */
    if (m < b) m = b;
    if (m < c) m = c;
    if (m < d) m = d;
    if (m < e) m = e;
    if (m < f) m = f;
    return m;
}
static double min6(double a, double b, double c, double d, double e, double f) {
    double m = a;
    if (m > b) m = b;
    if (m > c) m = c;
    if (m > d) m = d;
    if (m > e) m = e;
    if (m > f) m = f;
    return m;
}

double orc_dens(const double *c, int km, int k, double t, double s) {
    /* dens.F90:12-14 with C(k,m) at c[(m-1)*km + k] */
    return (c[k] + (c[3 * km + k] + c[6 * km + k] * s) * s +
            (c[2 * km + k] + c[7 * km + k] * s + c[5 * km + k] * t) * t) * t +
           (c[1 * km + k] + (c[4 * km + k] + c[8 * km + k] * s) * s) * s;
}

/* invtrit for one column (tracer.cpp:923-978, invtri.F90:5-65).  kz = kmt. */
void orc_invtri_column(int kz, const double *wkc, const double *odzt,
                       const double *odzp, double dts, double stf, double vit0,
                       double *x) {
    double a[kz > 0 ? kz : 1], b[kz > 0 ? kz : 1], c[kz > 0 ? kz : 1], dd[kz > 0 ? kz : 1];
    double e[kz + 1], f[kz + 1];
    if (kz <= 0) return;
    for (int k = 1; k < kz; ++k) {
        a[k] = wkc[k - 1] * odzt[k] * odzp[k] * dts * 0.5;
        dd[k] = x[k];
    }
    for (int k = 1; k < kz - 1; ++k) {
        c[k] = wkc[k] * odzt[k + 1] * odzp[k] * dts * 0.5;
        b[k] = 1.0 + a[k] + c[k];
    }
    a[0] = odzp[0] * dts * 0.5;
    c[0] = wkc[0] * odzt[1] * odzp[0] * dts * 0.5;
    b[0] = 1.0 + c[0];
    dd[0] = x[0];
    b[kz - 1] = 1.0 + a[kz - 1];
    c[kz - 1] = odzp[kz - 1] * dts * 0.5;
    e[kz] = 0.0;
    f[kz] = 0.0;
    for (int k = kz - 1; k >= 0; --k) {
        double g0 = 1.0 / (b[k] - c[k] * e[k + 1]);
        e[k] = a[k] * g0;
        f[k] = (dd[k] + c[k] * f[k + 1]) * g0;
    }
    x[0] = (e[0] * stf + f[0]) * vit0;
    for (int k = 1; k < kz; ++k) x[k] = (e[k] * x[k - 1] + f[k]); /* vit(k)=1 for k<kz */
}

/* ------------------------------------------------------------- tspas */
/* advection_tracer_2 predictor terms (tracer.cpp:311-414). */
typedef struct {
    double x0, y0, xy1, xy2, xy3, xy4, c1, c2, za, zb1, zb2, zc;
} pterms;

static void predictor_terms(const double *T, const double *U, const double *V,
                            const double *W, const orc_grid *g, double dts,
                            int k, int j, int i, pterms *t) {
    const int imt = g->imt, jmt = g->jmt, km = g->km;
    const double tr = g->tarea_r[I2(j, i)];
    const double *odzp = g->odzp, *odzt = g->odzt;
    const double Tc = T[I3(k, j, i)];
    const double Te = T[I3(k, j, i + 1)], Tw = T[I3(k, j, i - 1)];
    const double Ts = T[I3(k, j + 1, i)], Tn = T[I3(k, j - 1, i)];
    const double Ue = U[I3(k, j, i + 1)], Uw = U[I3(k, j, i)];
    const double Vs = V[I3(k, j, i)], Vn = V[I3(k, j - 1, i)];

    t->x0 = (Te + Tc) * Ue * tr - (Tc + Tw) * Uw * tr;
    t->y0 = ((Ts + Tc) * Vs - (Tc + Tn) * Vn) * tr;
    t->xy1 = -dts * (Te - Tc) * 2.0 * tr * Ue * Ue /
             (g->htw[I2(j, i + 1)] * g->hun[I2(j, i + 1)]);
    t->xy2 = dts * (Tc - Tw) * 2.0 * tr * Uw * Uw /
             (g->htw[I2(j, i)] * g->hun[I2(j, i)]);
    t->xy3 = -dts * (Ts - Tc) * 2.0 * tr * Vs * Vs /
             (g->hts[I2(j, i)] * g->hue[I2(j, i)]);
    t->xy4 = dts * (Tc - Tn) * 2.0 * tr * Vn * Vn /
             (g->hts[I2(j - 1, i)] * g->hue[I2(j - 1, i)]);
    t->c1 = -Tc * (Ue - Uw) * tr * 2.0;
    t->c2 = -Tc * (Vs - Vn) * tr * 2.0;
    if (k == 0) {
        const double W1 = W[I3(1, j, i)], T1 = T[I3(1, j, i)];
        t->za = -0.5 * odzp[0] * W1 * (T1 + Tc);
        t->zb1 = 0.0;
        t->zb2 = 0.5 * odzp[0] * W1 * W1 * odzt[1] * (Tc - T1) * dts;
        t->zc = odzp[0] * Tc * W1;
    } else if (k == km - 1) {
        const double Wk = W[I3(k, j, i)], Tu = T[I3(k - 1, j, i)];
        t->za = 0.5 * odzp[k] * Wk * (Tc + Tu);
        t->zb1 = -0.5 * odzp[k] * Wk * Wk * odzt[k] * (Tu - Tc) * dts;
        t->zb2 = 0.0;
        t->zc = -odzp[k] * Tc * Wk;
    } else {
        const double Wk = W[I3(k, j, i)], Wb = W[I3(k + 1, j, i)];
        const double Tu = T[I3(k - 1, j, i)], Td = T[I3(k + 1, j, i)];
        t->za = 0.5 * odzp[k] * Wk * (Tc + Tu) - 0.5 * odzp[k] * Wb * (Tc + Td);
        t->zb1 = -0.5 * odzp[k] * Wk * Wk * odzt[k] * (Tu - Tc) * dts;
        t->zb2 = 0.5 * odzp[k] * Wb * Wb * odzt[k + 1] * (Tc - Td) * dts;
        t->zc = -odzp[k] * Tc * (Wk - Wb);
    }
}

static void tspas(const orc_grid *g, double dts, const double *T,
                  const double *U, const double *V, const double *W,
                  double *at00, double *atmax, double *atmin, double *adv_tt,
                  double *ax, double *ay, double *az) {
    const int imt = g->imt, jmt = g->jmt, km = g->km;
    const double *vit = g->vit;

    /* advection_tracer_2: predictor and the 7-point bounds (tracer.cpp:291-547) */
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < km; ++k)
        for (int j = 1; j < jmt - 1; ++j)
            for (int i = 1; i < imt - 1; ++i) {
                pterms t;
                predictor_terms(T, U, V, W, g, dts, k, j, i, &t);
                double xx = -(t.x0 + t.xy1 + t.xy2 + t.c1);
                double yy = -(t.y0 + t.xy3 + t.xy4 + t.c2);
                double zz = -(t.zb1 + t.zb2 + t.za + t.zc);
                size_t c = I3(k, j, i);
                at00[c] = T[c] + (xx + yy + zz) * dts;

                const double lo = -1.0e10, hi = 1.0e10;
#define MV(q, w) (T[q] * vit[q] + (1.0 - vit[q]) * (w))
                size_t qn = I3(k, j - 1, i), qs = I3(k, j + 1, i);
                size_t qw = I3(k, j, i - 1), qe = I3(k, j, i + 1);
                if (k == 0) {
                    size_t qd = I3(1, j, i);
                    atmax[c] = max6(MV(c, lo), MV(qn, lo), MV(qs, lo), MV(qw, lo), MV(qe, lo), MV(qd, lo));
                    atmin[c] = min6(MV(c, hi), MV(qn, hi), MV(qs, hi), MV(qw, hi), MV(qe, hi), MV(qd, hi));
                } else if (k == km - 1) {
                    size_t qu = I3(k - 1, j, i);
                    atmax[c] = max6(MV(c, lo), MV(qn, lo), MV(qs, lo), MV(qw, lo), MV(qe, lo), MV(qu, lo));
                    atmin[c] = min6(MV(c, hi), MV(qn, hi), MV(qs, hi), MV(qw, hi), MV(qe, hi), MV(qu, hi));
                } else {
                    size_t qu = I3(k - 1, j, i), qd = I3(k + 1, j, i);
                    double m = max6(MV(c, lo), MV(qn, lo), MV(qs, lo), MV(qw, lo), MV(qe, lo), MV(qu, lo));
                    double v = MV(qd, lo);
                    if (m < v) m = v;
                    atmax[c] = m;
                    m = min6(MV(c, hi), MV(qn, hi), MV(qs, hi), MV(qw, hi), MV(qe, hi), MV(qu, hi));
                    v = MV(qd, hi);
                    if (m > v) m = v;
                    atmin[c] = m;
                }
#undef MV
            }

    /* advection_tracer_3: limiter by gather (tracer.cpp:549-921).  A cell is
     * "flagged" when its predictor leaves its local bounds; a face flux turns
     * first-order upwind when either adjacent cell is flagged.  The guard
     * terms reproduce the reference's parenthesisation at k=0 and k=km-1
     * (tracer.cpp:568-573) versus interior levels (:671-675). */
#define FL(q) (at00[q] > atmax[q] || at00[q] < atmin[q])
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < km; ++k)
        for (int j = 1; j < jmt - 1; ++j)
            for (int i = 1; i < imt - 1; ++i) {
                pterms t;
                predictor_terms(T, U, V, W, g, dts, k, j, i, &t);
                const double tr = g->tarea_r[I2(j, i)];
                const double *odzp = g->odzp;
                size_t c = I3(k, j, i);
                const double Tc = T[c];
                const int fc = FL(c);
                const int fe = FL(I3(k, j, i + 1)), fw = FL(I3(k, j, i - 1));
                const int fs = FL(I3(k, j + 1, i)), fn = FL(I3(k, j - 1, i));
                const double Ue = U[I3(k, j, i + 1)], Uw = U[I3(k, j, i)];
                const double Vs = V[I3(k, j, i)], Vn = V[I3(k, j - 1, i)];
                int up1, up2, up3, up4, upb1, upb2;
                if (k == 0 || k == km - 1) {
                    up1 = fc || (fe && i <= imt - 3);
                    up2 = fc || (fw && i >= 2);
                    up3 = fc || (fs && j <= jmt - 3);
                    up4 = fc || (fn && j >= 2);
                } else {
                    up1 = (fc || fe) && i <= imt - 3;
                    up2 = (fc || fw) && i >= 2;
                    up3 = (fc || fs) && j <= jmt - 3;
                    up4 = (fc || fn) && j >= 2;
                }
                double xy1 = up1 ? -(T[I3(k, j, i + 1)] - Tc) * fabs(Ue) * tr : t.xy1;
                double xy2 = up2 ? (Tc - T[I3(k, j, i - 1)]) * fabs(Uw) * tr : t.xy2;
                double xy3 = up3 ? -(T[I3(k, j + 1, i)] - Tc) * fabs(Vs) * tr : t.xy3;
                double xy4 = up4 ? (Tc - T[I3(k, j - 1, i)]) * fabs(Vn) * tr : t.xy4;
                double zb1, zb2;
                if (k == 0) {
                    zb1 = 0.0;
                    upb2 = fc || FL(I3(1, j, i));
                    zb2 = upb2 ? 0.5 * fabs(W[I3(1, j, i)]) * odzp[0] * (Tc - T[I3(1, j, i)]) : t.zb2;
                } else if (k == km - 1) {
                    upb1 = fc || FL(I3(k - 1, j, i));
                    zb1 = upb1 ? -0.5 * fabs(W[I3(k, j, i)]) * odzp[k] * (T[I3(k - 1, j, i)] - Tc) : t.zb1;
                    zb2 = 0.0;
                } else {
                    upb1 = fc || FL(I3(k - 1, j, i));
                    upb2 = fc || FL(I3(k + 1, j, i));
                    zb1 = upb1 ? -0.5 * fabs(W[I3(k, j, i)]) * odzp[k] * (T[I3(k - 1, j, i)] - Tc) : t.zb1;
                    zb2 = upb2 ? 0.5 * fabs(W[I3(k + 1, j, i)]) * odzp[k] * (Tc - T[I3(k + 1, j, i)]) : t.zb2;
                }
                double xx = -(t.x0 + xy1 + xy2 + t.c1);
                double yy = -(t.y0 + xy3 + xy4 + t.c2);
                double zz = -(t.za + zb1 + zb2 + t.zc);
                adv_tt[c] = xx + yy + zz;
                ax[c] = xx;
                ay[c] = yy;
                az[c] = zz;
            }
#undef FL
}

/* ---------------------------------------------------- horizontal mixing */
/* One 5-point pass of the tracer del2 operator (hmix_del2.F90:632-659,
 * tracer.cpp:1198-1234 and 1253-1290).  Returns hdtk (before the outer AH or
 * AHF factor) and cc through pointers. */
static void del2_sums(const orc_grid *g, const double *X, int k, int j, int i,
                      double *hdtk_out, double *cc_out) {
    const int imt = g->imt, jmt = g->jmt;
    double hdtk = 0.0, cc = 0.0;
    if (k <= g->kmt[I2(j, i)] - 1) {
        double c;
        if (k <= g->kmt[I2(j - 1, i)] - 1) {
            c = g->dtn[I2(j, i)];
            hdtk += c * X[I2(j - 1, i)];
            cc -= c;
        }
        if (k <= g->kmt[I2(j + 1, i)] - 1) {
            c = g->dts[I2(j, i)];
            hdtk += c * X[I2(j + 1, i)];
            cc -= c;
        }
        if (k <= g->kmt[I2(j, i + 1)] - 1) {
            c = g->dte[I2(j, i)];
            hdtk += c * X[I2(j, i + 1)];
            cc -= c;
        }
        if (k <= g->kmt[I2(j, i - 1)] - 1) {
            c = g->dtw[I2(j, i)];
            hdtk += c * X[I2(j, i - 1)];
            cc -= c;
        }
    }
    *hdtk_out = hdtk;
    *cc_out = cc;
}

void orc_advect(const orc_grid *g, double dts, const double *T, const double *uw,
                const double *vs, const double *ws, double *adv_tt) {
    const size_t n3 = (size_t)g->imt * g->jmt * g->km;
    double *b = (double *)calloc(6 * n3, sizeof(double));
    tspas(g, dts, T, uw, vs, ws, b, b + n3, b + 2 * n3, adv_tt, b + 3 * n3, b + 4 * n3, b + 5 * n3);
    free(b);
}

/* advection_tracer_1_0 (tracer.cpp:83-119, Fortran twin advection.F90:169-188):
 * face transports from the U-point velocities wkd (u), wkb (v): centered and
 * tspas (adv 0, 2) scale the sum of the face's two U corners by hts / htw,
 * flux (adv 1) weights each corner by its U-cell width dxu / dyu.  Faces
 * outside [1,imt-1) (v) and [1,jmt-1) (u) are left unwritten. */
void orc_faces(const orc_grid *g, int adv, const double *wkd, const double *wkb, double *uw,
               double *vs) {
    const int imt = g->imt, jmt = g->jmt, km = g->km;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < km; ++k)
        for (int j = 0; j < jmt; ++j)
            for (int i = 0; i < imt; ++i) {
                size_t c = I3(k, j, i);
                if (adv == 1) {
                    if (i >= 1 && i < imt - 1)
                        vs[c] = (wkb[c] * g->dxu[I2(j, i)] + wkb[c + 1] * g->dxu[I2(j, i + 1)]) * 0.25;
                    if (j >= 1 && j < jmt - 1)
                        uw[c] = (wkd[I3(k, j - 1, i)] * g->dyu[I2(j - 1, i)] + wkd[c] * g->dyu[I2(j, i)]) * 0.25;
                } else {
                    if (i >= 1 && i < imt - 1)
                        vs[c] = (wkb[c] + wkb[c + 1]) * g->hts[I2(j, i)] * 0.25;
                    if (j >= 1 && j < jmt - 1)
                        uw[c] = (wkd[I3(k, j - 1, i)] + wkd[c]) * g->htw[I2(j, i)] * 0.25;
                }
            }
}

/* advection_tracer_1 (tracer.cpp:120-289, Fortran twin advection.F90:190-247):
 * the centered (adv=0) and flux (adv=1) forms on the physical cells.
 *   centered: adv_tt = -(u.grad T) with the face transports, vertical terms
 *             -0.5*odzp*(w_k (T_{k-1}-T_k) + w_{k+1} (T_k-T_{k+1})); ax, ay,
 *             az accumulate the x, y and z parts divided by nss.
 *   flux:     adv_tt = (F_{i+1}-F_i + G_j-G_{j-1})*tarea_r with F = u_w (T+T),
 *             minus odzp*(Z_{k+1}-Z_k) with Z = w (T+T)/2; no diagnostics.
 * The reference's centered kernel runs its level-0 and level-(km-1)
 * statements in every k thread (tracer.cpp:137-160 and 210-256 carry no k
 * guard), so on the GPU those two levels' adv_tt and ax/ay/az increments are
 * applied a schedule-dependent number of times; this restatement applies
 * them once, as the Fortran twin does.  Level 0 and km-1 use the HIP
 * statements' single vertical term (the Fortran's 0 + z2 differs only in the
 * sign of a zero). */
void orc_advect_cf(const orc_grid *g, int adv, int nss, const double *T, const double *uw,
                   const double *vs, const double *ws, double *adv_tt, double *ax, double *ay,
                   double *az) {
    const int imt = g->imt, jmt = g->jmt, km = g->km;
    const double *odzp = g->odzp;
    const double p5 = 0.5;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < km; ++k)
        for (int j = 2; j < jmt - 2; ++j)
            for (int i = 2; i < imt - 2; ++i) {
                const size_t c = I3(k, j, i);
                const double tr = g->tarea_r[I2(j, i)];
                const double Tc = T[c], Tw = T[c - 1], Te = T[c + 1];
                const double Tn = T[I3(k, j - 1, i)], Ts = T[I3(k, j + 1, i)];
                const double Uw = uw[c], Ue = uw[c + 1];
                const double Vs = vs[c], Vn = vs[I3(k, j - 1, i)];
                if (adv == 0) {
                    double a = (-Uw * (Tc - Tw) - Ue * (Te - Tc) - Vs * (Ts - Tc) - Vn * (Tc - Tn)) * tr;
                    if (ax) ax[c] = ax[c] + (-Uw * (Tc - Tw) - Ue * (Te - Tc)) * tr / (double)nss;
                    if (ay) ay[c] = ay[c] + (-Vs * (Ts - Tc) - Vn * (Tc - Tn)) * tr / (double)nss;
                    double z;
                    if (k == 0) {
                        z = ws[I3(1, j, i)] * (Tc - T[I3(1, j, i)]);
                    } else if (k == km - 1) {
                        z = ws[c] * (T[I3(k - 1, j, i)] - Tc);
                    } else {
                        const double z1 = ws[c] * (T[I3(k - 1, j, i)] - Tc);
                        const double z2 = ws[I3(k + 1, j, i)] * (Tc - T[I3(k + 1, j, i)]);
                        z = z1 + z2;
                    }
                    a = a - p5 * odzp[k] * z;
                    if (az) az[c] = az[c] - p5 * odzp[k] * z / (double)nss;
                    adv_tt[c] = a;
                } else {
                    double a = (-Uw * (Tc + Tw) + Ue * (Te + Tc) - Vn * (Tc + Tn) + Vs * (Ts + Tc)) * tr;
                    const double z1 = (k == 0) ? 0.0 : ws[c] * (T[I3(k - 1, j, i)] + Tc) * p5;
                    const double z2 =
                        (k == km - 1) ? 0.0 : ws[I3(k + 1, j, i)] * (Tc + T[I3(k + 1, j, i)]) * p5;
                    adv_tt[c] = a - odzp[k] * (z2 - z1);
                }
            }
}

void orc_del2_level(const orc_grid *g, const double *X, int k, int use_ahf, double ah,
                    double *out) {
    const int imt = g->imt, jmt = g->jmt;
    for (int j = 2; j < jmt - 2; ++j)
        for (int i = 2; i < imt - 2; ++i) {
            double h, cc;
            del2_sums(g, X, k, j, i, &h, &cc);
            const double f = use_ahf ? g->ahf[I2(j, i)] : ah;
            out[I2(j, i)] = f * (cc * X[I2(j, i)] + h);
        }
}

/* ------------------------------------------------------------ the step */
void orc_tracer_step(const orc_grid *g, const orc_params *p, orc_fields *f,
                     orc_diag *d) {
    const int imt = g->imt, jmt = g->jmt, km = g->km;
    const size_t n2 = (size_t)imt * jmt, n3 = n2 * km;
    const double *vit = g->vit;
    const double aidif = 0.5;

    double *stf = (double *)calloc(n2, sizeof(double));
    double *work = (double *)calloc(n2, sizeof(double));
    double *temp11 = (double *)calloc(n2, sizeof(double));
    double *wkd = (double *)calloc(n3, sizeof(double));
    double *wkb = (double *)calloc(n3, sizeof(double));
    double *wka = (double *)calloc(n3, sizeof(double));
    double *uk = (double *)calloc(n3, sizeof(double));
    double *vk = (double *)calloc(n3, sizeof(double));
    double *uw = (double *)calloc(n3, sizeof(double));
    double *vs = (double *)calloc(n3, sizeof(double));
    double *at00 = (double *)calloc(n3, sizeof(double));
    double *atmax = (double *)calloc(n3, sizeof(double));
    double *atmin = (double *)calloc(n3, sizeof(double));
    double *adv = (double *)calloc(n3, sizeof(double));
    double *tf = (double *)calloc(n3, sizeof(double));
    double *wkc = (double *)calloc(n3, sizeof(double));
    double *d2tk = (double *)calloc(n3, sizeof(double));
    double *vtl = (double *)calloc(n3, sizeof(double));
    double *ori = (double *)calloc(n3, sizeof(double));

    /* tracer_1_0 (tracer.cpp:1081-1093) */
    for (size_t q = 0; q < n2; ++q) {
        f->h0f[q] = f->h0f[q] * p->onbc;
        stf[q] = 0.5 * f->h0f[q] + 0.5 * f->h0l[q];
    }
    /* tracer_1 (tracer.cpp:1096-1150) */
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < n3; ++q) {
        f->utf[q] *= p->oncc;
        f->vtf[q] *= p->oncc;
        wkd[q] = 0.5 * f->utf[q] + 0.5 * f->utl[q];
        wkb[q] = 0.5 * f->vtf[q] + 0.5 * f->vtl[q];
    }
    for (int j = 0; j < jmt - 1; ++j)
        for (int i = 1; i < imt; ++i)
            work[I2(j, i)] = g->au0[I2(j, i)] * stf[I2(j, i)] + g->aus[I2(j, i)] * stf[I2(j + 1, i)] +
                             g->auw[I2(j, i)] * stf[I2(j, i - 1)] +
                             g->ausw[I2(j, i)] * stf[I2(j + 1, i - 1)];
    /* upwell_2 (tracer.cpp:1008-1027) */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < km; ++k)
        for (int j = 0; j < jmt - 1; ++j)
            for (int i = 1; i < imt; ++i) {
                size_t c = I3(k, j, i);
                uk[c] = (1.0 + work[I2(j, i)] * g->ohbu[I2(j, i)]) * wkd[c];
                vk[c] = (1.0 + work[I2(j, i)] * g->ohbu[I2(j, i)]) * wkb[c];
            }
    /* operators_div (tracer.cpp:980-1006) */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < km; ++k)
        for (int j = 1; j < jmt; ++j)
            for (int i = 0; i < imt - 1; ++i) {
                if (k > g->kmt[I2(j, i)] - 1) continue;
                wka[I3(k, j, i)] =
                    0.5 *
                    ((uk[I3(k, j, i + 1)] + uk[I3(k, j - 1, i + 1)]) * g->htw[I2(j, i + 1)] -
                     (uk[I3(k, j, i)] + uk[I3(k, j - 1, i)]) * g->htw[I2(j, i)] +
                     (vk[I3(k, j, i + 1)] + vk[I3(k, j, i)]) * g->hts[I2(j, i)] -
                     (vk[I3(k, j - 1, i + 1)] + vk[I3(k, j - 1, i)]) * g->hts[I2(j - 1, i)]) *
                    g->tarea_r[I2(j, i)];
            }
    /* upwell_3 / upwell_4 (tracer.cpp:1029-1080) */
    for (size_t q = 0; q < n2; ++q) work[q] = 0.0;
#pragma omp parallel for schedule(static)
    for (int j = 1; j < jmt - 1; ++j)
        for (int i = 1; i < imt - 1; ++i) {
            double w = 0.0;
            for (int k = 0; k < km; ++k) w = w - g->dzp[k] * wka[I3(k, j, i)] * vit[I3(k, j, i)];
            for (int k = 1; k < km; ++k)
                f->ws[I3(k, j, i)] =
                    vit[I3(k, j, i)] * (f->ws[I3(k - 1, j, i)] +
                                        g->dzp[k - 1] * (w * g->ohbt[I2(j, i)] + wka[I3(k - 1, j, i)]));
            double s = 1.0 / (1.0 + stf[I2(j, i)] * g->ohbt[I2(j, i)]);
            work[I2(j, i)] = s;
            for (int k = 1; k < km; ++k) f->ws[I3(k, j, i)] = f->ws[I3(k, j, i)] * s;
        }

    double fw_norm2 = 0.0;
    for (int n = 0; n < 2; ++n) {
        double *T = f->at + (size_t)n * n3;
        double *TB = f->atb + (size_t)n * (n3 + n2); /* level k at plane k+1 */
        double *akt = f->akt + (size_t)n * n3;
        double *ax = d->ax + (size_t)n * n3, *ay = d->ay + (size_t)n * n3;
        double *az = d->az + (size_t)n * n3, *dz = d->dz + (size_t)n * n3;
        memset(tf, 0, n3 * sizeof(double));
        memset(adv, 0, n3 * sizeof(double));

        orc_faces(g, p->adv, wkd, wkb, uw, vs);
        if (p->adv == 2)
            tspas(g, p->dts, T, uw, vs, f->ws, at00, atmax, atmin, adv, ax, ay, az);
        else
            orc_advect_cf(g, p->adv, p->nss, T, uw, vs, f->ws, adv, ax, ay, az);

        /* horizontal mixing + AKT floor (tracer.cpp:1167-1362) */
        if (g->hmix == 4) {
            memset(d2tk, 0, n3 * sizeof(double));
#pragma omp parallel for schedule(static)
            for (int k = 0; k < km; ++k)
                for (int j = 1; j < jmt - 1; ++j)
                    for (int i = 1; i < imt - 1; ++i) {
                        double h, cc;
                        const double *X = TB + (size_t)(k + 1) * n2;
                        del2_sums(g, X, k, j, i, &h, &cc);
                        d2tk[I3(k, j, i)] = g->ahf[I2(j, i)] * (cc * X[I2(j, i)] + h);
                    }
        }
        if (n == 0 || g->hmix == 2)
            for (int j = 1; j < jmt - 1; ++j)
                for (int i = 1; i < imt - 1; ++i)
                    if (akt[I2(j, i)] < p->dwndmix) akt[I2(j, i)] = p->dwndmix;
#pragma omp parallel for schedule(static)
        for (int k = 0; k < km; ++k)
            for (int j = 1; j < jmt - 1; ++j)
                for (int i = 1; i < imt - 1; ++i) {
                    size_t c = I3(k, j, i);
                    int inner = (j >= 2 && j < jmt - 2 && i >= 2 && i < imt - 2);
                    if (inner) tf[c] = adv[c] * vit[c];
                    wkc[c] = akt[c];
                    double h, cc;
                    if (g->hmix == 4) {
                        const double *X = d2tk + (size_t)k * n2;
                        del2_sums(g, X, k, j, i, &h, &cc);
                        h = inner ? p->ah * (cc * X[I2(j, i)] + h) : 0.0;
                    } else {
                        const double *X = TB + (size_t)(k + 1) * n2;
                        del2_sums(g, X, k, j, i, &h, &cc);
                        h = inner ? p->ah * (cc * X[I2(j, i)] + h) : 0.0;
                    }
                    tf[c] += h;
                }

        /* tracer_5 (tracer.cpp:1395-1525) */
#pragma omp parallel for schedule(static)
        for (int j = 2; j < jmt - 2; ++j)
            for (int i = 2; i < imt - 2; ++i) {
                const size_t q = I2(j, i);
#define AB(k) TB[(size_t)((k) + 1) * n2 + q]
#define C3(k) ((size_t)(k) * n2 + q)
                if (n == 0) {
                    const double sw = f->swv[q];
                    for (int k = 1; k < km - 1; ++k) {
                        double wt1 = sw * g->pen[k - 1] * vit[C3(k)];
                        double wt2 = sw * g->pen[k] * vit[C3(k + 1)];
                        tf[C3(k)] = tf[C3(k)] + (wt1 - wt2) * g->odzp[k];
                        d->penetrate[C3(k)] = (wt1 - wt2) * g->odzp[k];
                    }
                    double wt1 = sw * g->pen[0] * vit[C3(1)];
                    double wt2 = sw * g->pen[km - 2] * vit[C3(km - 1)];
                    tf[C3(0)] = tf[C3(0)] - g->odzp[0] * wt1;
                    tf[C3(km - 1)] = tf[C3(km - 1)] + g->odzp[km - 1] * wt2;
                    d->penetrate[C3(0)] = -wt1 * g->odzp[0];
                    d->penetrate[C3(km - 1)] = wt2 * g->odzp[km - 1];
                }
                for (int k = 1; k < km - 1; ++k) {
                    double wt1 = wkc[C3(k - 1)] * (AB(k - 1) - AB(k)) * g->odzt[k] * vit[C3(k)];
                    double wt2 = wkc[C3(k)] * (AB(k) - AB(k + 1)) * g->odzt[k + 1] * vit[C3(k + 1)];
                    tf[C3(k)] = tf[C3(k)] + g->odzp[k] * (wt1 - wt2) * 0.5;
                    dz[C3(k)] = g->odzp[k] * (wt1 - wt2) * 0.5;
                }
                double wt1 = wkc[C3(0)] * (AB(0) - AB(1)) * g->odzt[1] * vit[C3(1)];
                double wt2 = wkc[C3(km - 2)] * (AB(km - 2) - AB(km - 1)) * g->odzt[km - 1] * vit[C3(km - 1)];
                tf[C3(0)] = tf[C3(0)] - g->odzp[0] * wt1 * 0.5;
                tf[C3(km - 1)] = tf[C3(km - 1)] + g->odzp[km - 1] * wt2 * (1.0 - aidif);
                dz[C3(0)] = -g->odzp[0] * wt1 * 0.5;
                dz[C3(km - 1)] = g->odzp[km - 1] * wt2 * 0.5;
            }
        /* surface fluxes: tracer_5 (n=1, FRC_CORE, :1479-1524) and
         * tracer_6 (n=0, FRC_CORE, :1543-1566); all columns with kmt>0 */
        for (int j = 0; j < jmt; ++j)
            for (int i = 0; i < imt; ++i) {
                const size_t q = I2(j, i);
                if (g->kmt[q] <= 0) continue;
                if (n == 1) {
                    const double sb = AB(0);
                    stf[q] = f->fresh[q] * 34.7 * p->od0 / 1000.0 +
                             p->gamma * (f->sss[q] - sb) * f->seaice[q] / g->odzp[0] +
                             p->gamma * 30.0 / 365.0 / 4.0 * 50.0 * (f->sss[q] - sb) / g->odzp[0] *
                                 (1.0 - f->seaice[q]);
                    tf[C3(0)] = tf[C3(0)] + stf[q] * (1.0 - aidif) * g->odzp[0];
                    d->net[n2 + q] = stf[q] * g->odzp[0];
                    temp11[q] = g->tarea[q] * d->net[n2 + q] * vit[C3(0)];
                }
            }
        double err_norm2 = 0.0;
        if (n == 1 && p->sssnorm) {
            /* tracer.cpp:1849-1868: plain sum over the physical cells */
            double s = 0.0;
            for (int j = 2; j < jmt - 2; ++j)
                for (int i = 2; i < imt - 2; ++i) s = s + temp11[I2(j, i)];
            err_norm2 = -s / g->area_t;
            fw_norm2 = err_norm2;
            for (int j = 0; j < jmt; ++j)
                for (int i = 0; i < imt; ++i)
                    tf[I2(j, i)] = tf[I2(j, i)] + err_norm2 * vit[I2(j, i)];
        }
        if (n == 0) {
            for (int j = 0; j < jmt; ++j)
                for (int i = 0; i < imt; ++i) {
                    const size_t q = I2(j, i);
                    if (g->kmt[q] <= 0) continue;
                    stf[q] = (f->swv[q] + f->nswv[q]) * p->od0cp +
                             f->seaice[q] * p->gamma * (f->sst[q] - AB(0)) / g->odzp[0];
                    tf[C3(0)] = tf[C3(0)] + stf[q] * g->odzp[0] * 0.5;
                    d->net[q] = stf[q] * g->odzp[0];
                }
        }
        /* tracer_7 (tracer.cpp:1571-1627) */
        if (p->boundary_restore == 1 && f->restore) {
            const double lamda = 1.0 / (15.0 * 86400.0);
            const double *R = f->restore + (size_t)n * n3;
            for (int k = 1; k < km; ++k)
                for (int j = 2; j < jmt - 2; ++j)
                    for (int i = 2; i < imt - 2; ++i) {
                        const size_t q = I2(j, i);
                        tf[C3(k)] = tf[C3(k)] + vit[C3(k)] * (R[C3(k)] - AB(k)) * lamda;
                    }
        }
        if (d->tf) memcpy(d->tf + (size_t)n * n3, tf, n3 * sizeof(double));
        /* tracer_8 + invtrit (tracer.cpp:1629-1653, 923-978) */
#pragma omp parallel for schedule(static)
        for (int j = 2; j < jmt - 2; ++j)
            for (int i = 2; i < imt - 2; ++i) {
                const size_t q = I2(j, i);
                double x[km > 0 ? km : 1], wc[km > 0 ? km : 1];
                for (int k = 0; k < km; ++k) {
                    vtl[C3(k)] = T[C3(k)] + p->dts * tf[C3(k)];
                    ori[C3(k)] = vtl[C3(k)];
                    x[k] = vtl[C3(k)];
                    wc[k] = wkc[C3(k)];
                }
                int kz = g->kmt[q];
                if (kz > 0) {
                    orc_invtri_column(kz, wc, g->odzt, g->odzp, p->dts, stf[q], vit[C3(0)], x);
                    for (int k = 0; k < kz; ++k) vtl[C3(k)] = x[k];
                }
            }
        orc_halo_fill_f64(vtl, imt, jmt, km);
        /* tracer_9 (tracer.cpp:1655-1699): diagnostics on the physical cells,
         * state on every cell. */
        double *tend = d->tend + (size_t)n * n3, *dtd = d->dt_diff + (size_t)n * n3;
#pragma omp parallel for schedule(static)
        for (int j = 2; j < jmt - 2; ++j)
            for (int i = 2; i < imt - 2; ++i) {
                const size_t q = I2(j, i);
                dtd[C3(0)] = (vtl[C3(0)] - ori[C3(0)]) / p->dts * vit[C3(0)] - stf[q] * 0.5 * g->odzp[0];
                for (int k = 1; k < km; ++k)
                    dtd[C3(k)] = (vtl[C3(k)] - ori[C3(k)]) / p->dts * vit[C3(k)];
                for (int k = 0; k < km; ++k)
                    tend[C3(k)] = (vtl[C3(k)] - T[C3(k)]) / p->dts * vit[C3(k)];
            }
        for (size_t q = 0; q < n3; ++q) {
            T[q] = vtl[q];
            TB[n2 + q] = T[q];
        }
#undef AB
#undef C3
    }
    d->fw_norm2 = fw_norm2;

    free(stf); free(work); free(temp11); free(wkd); free(wkb); free(wka);
    free(uk); free(vk); free(uw); free(vs); free(at00); free(atmax);
    free(atmin); free(adv); free(tf); free(wkc); free(d2tk); free(vtl); free(ori);

    if (p->convadj) orc_convadj(g, p, f, d);
}

/* --------------------------------------------------------------- convadj */
/* convadj_cu (convadj.cpp:29-288): column by column, mix statically
 * unstable neighbours (T and S together, dz-weighted) until stable, then
 * dt_conv = (at - atb)/dts, tend += dt_conv, atb = at.  Columns are all
 * (i,j) including ghosts; diagnostics are kept for the physical cells. */
void orc_convadj(const orc_grid *g, const orc_params *p, orc_fields *f,
                 orc_diag *d) {
    const int imt = g->imt, jmt = g->jmt, km = g->km;
    const size_t n2 = (size_t)imt * jmt, n3 = n2 * km;
    const double *C = g->eosc, *to = g->to, *so = g->so, *dzp = g->dzp;
#pragma omp parallel for collapse(2) schedule(dynamic, 16)
    for (int j = 0; j < jmt; ++j)
        for (int i = 0; i < imt; ++i) {
            const size_t q = I2(j, i);
            double *A[2] = {f->at + q, f->at + n3 + q};
#define AT(n, l) A[n][(size_t)(l) * n2]
            double rhoup[km], rholo[km];
            for (int l = 0; l < km; ++l) rhoup[l] = rholo[l] = 0.0;
            int kcon = g->kmt[q];
            if (kcon != 0) {
                int lcon = 0;
                for (int l = 0; l < km - 1; ++l) {
                    int l1 = l + 1;
                    rhoup[l1] = orc_dens(C, km, l1, AT(0, l1) - to[l1], AT(1, l1) - so[l1]);
                    rholo[l] = orc_dens(C, km, l1, AT(0, l) - to[l1], AT(1, l) - so[l1]);
                }
                for (int k = kcon - 1; k >= 1; --k)
                    if (rholo[k - 1] > rhoup[k]) lcon = k;
                if (lcon != 0) {
                    for (;;) { /* CONV_1 */
                        int lcona = lcon, lconb = lcon + 1;
                        double dztsum = dzp[lcona - 1] + dzp[lconb - 1];
                        double trasum[2];
                        for (int n = 0; n < 2; ++n) {
                            trasum[n] = AT(n, lcona - 1) * dzp[lcona - 1] + AT(n, lconb - 1) * dzp[lconb - 1];
                            double tramix = trasum[n] / dztsum;
                            AT(n, lcona - 1) = tramix;
                            AT(n, lconb - 1) = tramix;
                        }
                        for (;;) { /* CONV_2 */
                            if (lconb != kcon) {
                                int l1 = lconb + 1;
                                rholo[lconb - 1] = orc_dens(C, km, l1 - 1, AT(0, lconb - 1) - to[l1 - 1],
                                                            AT(1, lconb - 1) - so[l1 - 1]);
                                if (rholo[lconb - 1] > rhoup[l1 - 1]) {
                                    lconb = lconb + 1;
                                    dztsum = dztsum + dzp[lconb - 1];
                                    for (int n = 0; n < 2; ++n) {
                                        trasum[n] = trasum[n] + AT(n, lconb - 1) * dzp[lconb - 1];
                                        double tramix = trasum[n] / dztsum;
                                        for (int lm = lcona - 1; lm < lconb; ++lm) AT(n, lm) = tramix;
                                    }
                                    continue;
                                }
                            }
                            if (lcona > 1) {
                                int l1 = lcona - 1;
                                rholo[l1 - 1] = orc_dens(C, km, lcona - 1, AT(0, l1 - 1) - to[lcona - 1],
                                                         AT(1, l1 - 1) - so[lcona - 1]);
                                rhoup[lcona - 1] = orc_dens(C, km, lcona - 1, AT(0, lcona - 1) - to[lcona - 1],
                                                            AT(1, lcona - 1) - so[lcona - 1]);
                                if (rholo[lcona - 2] > rhoup[lcona - 1]) {
                                    lcona = lcona - 1;
                                    dztsum = dztsum + dzp[lcona - 1];
                                    for (int n = 0; n < 2; ++n) {
                                        trasum[n] = trasum[n] + AT(n, lcona - 1) * dzp[lcona - 1];
                                        double tramix = trasum[n] / dztsum;
                                        for (int lm = lcona; lm <= lconb; ++lm) AT(n, lm - 1) = tramix;
                                    }
                                    continue;
                                }
                            }
                            break;
                        }
                        if (lconb == kcon) break;
                        lcon = lconb;
                        int done = 0;
                        for (;;) { /* CONV_3 */
                            lcon = lcon + 1;
                            if (lcon == kcon) { done = 1; break; }
                            if (rholo[lcon - 1] <= rhoup[lcon]) continue;
                            break;
                        }
                        if (done) break;
                    }
                }
            }
            const int inner = (j >= 2 && j < jmt - 2 && i >= 2 && i < imt - 2);
            for (int n = 0; n < 2; ++n) {
                double *TB = f->atb + (size_t)n * (n3 + n2) + n2 + q;
                for (int k = 0; k < km; ++k) {
                    double dc = (AT(n, k) - TB[(size_t)k * n2]) / p->dts * g->vit[(size_t)k * n2 + q];
                    if (inner) {
                        d->dt_conv[(size_t)n * n3 + (size_t)k * n2 + q] = dc;
                        d->tend[(size_t)n * n3 + (size_t)k * n2 + q] += dc;
                    }
                    TB[(size_t)k * n2] = AT(n, k);
                }
            }
#undef AT
        }
}

/* --------------------------------------------------------------- icesnow */
/* icesnow_ (icesnow.cpp:70-85) with the active build's FRC_CORE reset of
 * licomqice (icesnow.cpp:76-78, def-undef.h:6) when reset is nonzero, then
 * icesnow_cu (icesnow.cpp:15-68) on every column, ghosts included:
 *   1. a wet column whose surface T is below TBICE freezes back to TBICE;
 *      licomqice takes tdiff*dzp(0)/dzp(0) and the surface S the rejected
 *      salt tdiff*(34.7-4.0)*CP/LATICE*0.001 (icesnow.cpp:33-44);
 *   2. stored ice melts against a surface above TBICE (icesnow.cpp:47-59);
 *   3. T is clamped to TBICE at every level (icesnow.cpp:61-67).
 * TBICE = -1.8 and CP = 3996 (constant_mod.F90:164,168), LATICE = 3.337e5
 * (shr_const_mod.h:51). */
void orc_icesnow(const orc_grid *g, double *at, double *licomqice, int reset) {
    const int imt = g->imt, jmt = g->jmt, km = g->km;
    const size_t n2 = (size_t)imt * jmt, n3 = n2 * km;
    const double tbice = -1.8, cp = 3996.0, latice = 3.337e5;
    const double sal_ocn = 34.7, sal_ice = 4.0;
    if (reset)
        for (size_t q = 0; q < n2; ++q) licomqice[q] = 0.0;
#pragma omp parallel for collapse(2) schedule(static)
    for (int j = 0; j < jmt; ++j)
        for (int i = 0; i < imt; ++i) {
            const size_t q = I2(j, i);
            double *T = at + q, *S = at + n3 + q;
            if (T[0] < tbice && g->kmt[q] > 0) {
                const double tdiff = tbice - T[0];
                licomqice[q] += tdiff * g->dzp[0] / g->dzp[0];
                S[0] += tdiff * (sal_ocn - sal_ice) * cp / latice * 0.001;
                T[0] = tbice;
            }
            if (licomqice[q] > 0 && T[0] > tbice) {
                const double tdiff = fmin(T[0] - tbice, licomqice[q]);
                licomqice[q] -= tdiff;
                T[0] -= tdiff;
                S[0] -= tdiff * (sal_ocn - sal_ice) * cp / latice * 0.001;
            }
            for (int k = 0; k < km; ++k)
                if (T[(size_t)k * n2] < tbice) T[(size_t)k * n2] = tbice;
        }
}
