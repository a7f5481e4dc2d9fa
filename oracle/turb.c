/* turb.c -- CPU oracle of the Canuto AKT producer (test infrastructure only;
 * see turb.h).  Plain C, -ffp-contract=off: every expression keeps the
 * reference's operation order, so the only differences from the GPU product
 * come from the elementary functions (atan, sin, cos, acosh, log), whose
 * last bits differ between host and device math libraries.
 *
 * Where the reference reads a variable before any assignment (an, slq2,
 * sm, sh, ss, amtaun2 and the background sm/sh/ss/slq2 of turb_device,
 * readyc.cpp:1126-1144), this restatement starts it at 0; where its table
 * search could step outside the table (a log-estimated index rounding past
 * MT, readyc.cpp:926-929) the index is kept inside it, and a NaN argument
 * leaves the interpolation's outputs alone instead of indexing with it.
 */
#include "turb.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MT ORC_TURB_MT
#define MT0 ORC_TURB_MT0
#define NTH ORC_TURB_NTH
#define W2 (2 * MT + 1)

/* canuto_mod.h constants used on this path */
static const double kEpson2 = 0.288;          /* epson2__ */
static const double kEplatMin = 7.0e-2;       /* eplatidependmin */
static const double kS2min = 1.0e-14;         /* s2min */
/* pconst_mod.h:242-247 (TIDEMIX) */
static const double kLocalFrac = 0.33, kMixEf = 0.20, kBackTide = 1.0e-6, kMaxTide = 1.0e-2;

/* Fortran SIGN(a, b) for doubles and integers (readyc.cpp:898-900) */
static double fsign(double a, double b) { return b >= 0.0 ? fabs(a) : -fabs(a); }
static int isign(int a, int b) { return b >= 0 ? abs(a) : -abs(a); }

double orc_turb_formld(int n, const double *z, const double *t) {
    for (int k = 0; k < n; ++k) {
        if (fabs(t[k] - t[0]) > 0.1) {
            const double tm = t[0] - fsign(0.1, t[0] - t[k]);
            return z[k] + (z[k - 1] - z[k]) * (tm - t[k]) / (t[k - 1] - t[k] + 1.0e-20);
        }
    }
    return z[n - 1];
}

double orc_turb_eplatidepend(double f, double an) {
    const double an0 = 5.24e-3;
    const double f30 = 4.0 * atan(1.0) / 43082.0;   /* Coriolis parameter at 30 degrees */
    const double den = f30 * acosh(an0 / f30);
    const double num = f * acosh(an / f);
    return num / den;
}

/* Cell (l0, l1) of x in an expabs table xt (uniform step delta for |x| below
 * xt[MT+MT0], ratio rat beyond): readyc.cpp:929-951 and, for the 2-D
 * table, :1011-1033 / :1045-1065.  Returns 0 when the found cell does not
 * bracket x (the reference then returns without touching its outputs). */
static int expabs_cell(const double *xt, double x, double delta, double rat, int *pl0, int *pl1) {
    const double ax = fabs(x), xs = xt[MT + MT0], xe = xt[2 * MT];
    int l0, l1;
    if (ax >= xs && ax >= xe) {
        l0 = l1 = (int)fsign((double)MT, x);
    } else {
        if (ax < xs) {
            l1 = (int)(x / delta) + (int)fsign(1.0, x);
        } else {
            const double idx = fsign((double)MT0 + ((log(ax) - log(xs)) / log(rat)), x);
            l1 = (int)idx + (int)fsign(1.0, x);
        }
        if (l1 > MT) l1 = MT;
        if (l1 < -MT) l1 = -MT;
        if (fabs(xt[MT + l1]) < ax)
            l1 += isign(1, l1);
        else if (fabs(xt[MT + l1 - isign(1, l1)]) > ax)
            l1 -= isign(1, l1);
        l0 = l1 - (int)fsign(1.0, x);
        if (ax <= 0.0) l1 = 1;
    }
    *pl0 = l0;
    *pl1 = l1;
    if ((x > 0.0 && (x < xt[MT + l0] || x > xt[MT + l1])) ||
        (x < 0.0 && (x > xt[MT + l0] || x < xt[MT + l1])))
        return 0;
    return 1;
}

int orc_turb_interp1d(const orc_turb_tables *t, double x, double out[4]) {
    const double *xt = t->and2on2a1;
    if (isnan(x)) return 0;
    if (x > xt[2 * MT])
        x = xt[2 * MT];
    else if (x < xt[0])
        x = xt[0];
    int l0, l1;
    if (!expabs_cell(xt, x, t->dand2on2, t->rnd2on2, &l0, &l1)) return 0;
    const double span = xt[MT + l1] - xt[MT + l0];
    const double dx = x - xt[MT + l0];
    const double *f[4] = {t->amtaun2a1, t->sma1, t->sha1, t->ssa1};
    for (int m = 0; m < 4; ++m) {
        const double slope = (l1 == l0) ? 0.0 : (f[m][MT + l1] - f[m][MT + l0]) / span;
        out[m] = f[m][MT + l0] + slope * dx;
    }
    return 1;
}

int orc_turb_interp2d(const orc_turb_tables *t, double *pri, double *prid, double out[4]) {
    const double *rb = t->rib;
    const double lo = rb[0], hi = rb[2 * MT];
    /* bring (ri, rid) back into the table along the ray through the origin
     * (readyc.cpp:984-1006); each statement uses the values already updated */
    if (*pri > hi) {
        if (fabs(*prid) <= *pri) {
            *prid = hi * (*prid / *pri);
            *pri = hi;
        } else if (*prid > *pri) {
            *pri = hi * (*pri / *prid);
            *prid = hi;
        } else if (*prid < -*pri) {
            *pri = lo * (*pri / *prid);
            *prid = lo;
        }
    } else if (*pri < lo) {
        if (fabs(*prid) <= -*pri) {
            *prid = lo * (*prid / *pri);
            *pri = lo;
        } else if (*prid > -*pri) {
            *pri = hi * (*pri / *prid);
            *prid = hi;
        } else if (*prid < *pri) {
            *pri = lo * (*pri / *prid);
            *prid = lo;
        }
    } else if (*prid > hi) {
        *pri = hi * (*pri / *prid);
        *prid = hi;
    } else if (*prid < lo) {
        *pri = lo * (*pri / *prid);
        *prid = lo;
    }
    const double ri = *pri, rid = *prid;
    if (isnan(ri) || isnan(rid)) return 0;
    int d0, d1, r0, r1;
    if (!expabs_cell(rb, rid, t->dri, t->rri, &d0, &d1)) return 0;
    const double rimax0 = rb[MT + t->irimax[MT + d0]], rimax1 = rb[MT + t->irimax[MT + d1]];
    if (ri > fmin(rimax0, rimax1)) {
        out[0] = out[1] = out[2] = out[3] = 0.0;
        return 2;
    }
    if (!expabs_cell(rb, ri, t->dri, t->rri, &r0, &r1)) return 0;
    const double span_d = rb[MT + d1] - rb[MT + d0], span_r = rb[MT + r1] - rb[MT + r0];
    const double dd = rid - rb[MT + d0], dr = ri - rb[MT + r0];
    const double *f[4] = {t->slq2b, t->smb, t->shb, t->ssb};
    const size_t c00 = (size_t)(MT + d0) * W2 + (size_t)(MT + r0);
    const size_t c01 = (size_t)(MT + d1) * W2 + (size_t)(MT + r0);   /* next Ri_d */
    const size_t c10 = (size_t)(MT + d0) * W2 + (size_t)(MT + r1);   /* next Ri */
    for (int m = 0; m < 4; ++m) {
        double s_d = 0.0, s_r = 0.0;
        if (d1 != d0) {
            s_d = (f[m][c01] - f[m][c00]) / span_d;
            s_r = (f[m][c10] - f[m][c00]) / span_r;
        }
        out[m] = f[m][c00] + s_r * dr + s_d * dd;
    }
    return 1;
}

static void bump(long *counts, int which) {
    if (counts) counts[which] += 1;
}

int orc_turb_column(const orc_turb_tables *t, int n, int nmax, const double *z, const double *tt,
                    const double *rh, const double *ri, const double *rid, double *s2,
                    const double *an2, double fricmx, double wndmix, double buoytur,
                    double buoysol, double coriol, double *amld, double *akm, double *akh,
                    double *aks, long *counts) {
    const double b1 = t->b1, dth = t->deltheta_r;
    const double quarter = atan(1.0);   /* pi/4 */
    const double buoytot = buoytur + buoysol;
    if (n > 0)
        *amld = orc_turb_formld(n, z, tt);
    else if (n == 0)
        *amld = z[0];
    else
        *amld = 0.0;
    const double al0 = 0.17 * *amld;

    /* carried from level to level, as in the reference */
    int below = 0;
    double an = 0.0, slq2 = 0.0, sm = 0.0, sh = 0.0, ss = 0.0, amtaun2 = 0.0;
    double sm_bk = 0.0, sh_bk = 0.0, ss_bk = 0.0, slq2_bk = 0.0, back_ri1 = 0.0;

    for (int k = 0; k < n; ++k) {
        int nofsmall = 0;
        if (an2[k] >= 0.0) an = sqrt(an2[k]);
        if (an < fabs(coriol)) nofsmall = 1;
        double ri1 = ri[k], rid1 = rid[k];
        const double and2 = rid[k] / (ri[k] + 1.0e-25) * an2[k];
        const double and2on2 = and2 / (an2[k] + 1.0e-25);
        const int pure = an2[k] < t->rib[0] * s2[k];
        double o[4];
        if (pure) {
            bump(counts, ORC_TB_PURE);
            if (orc_turb_interp1d(t, and2on2, o)) {
                amtaun2 = o[0];
                sm = o[1];
                sh = o[2];
                ss = o[3];
            }
            slq2 = (-amtaun2) / ((b1 * b1) * ri[k] + 1.0e-25);
        } else {
            bump(counts, ORC_TB_TABLE2D);
            if (orc_turb_interp2d(t, &ri1, &rid1, o)) {
                slq2 = o[0];
                sm = o[1];
                sh = o[2];
                ss = o[3];
            }
            if (slq2 < 0.0) return k;
            if (slq2 == 0.0) below = 1;
        }
        if (below) bump(counts, ORC_TB_BELOW);
        const int upper = (below == 0 || nofsmall == 1 || (ri1 < 0.0 && k <= 1)) && slq2 > 0.0;

        double epsy = 0.0;
        int use_epsy = 0;
        if (slq2 == 0.0)
            epsy = 0.0;
        else
            epsy = -buoytot / ((1.0 / ((b1 * b1) * slq2)) - 0.5 * sm);
        use_epsy = epsy >= 0.0 && upper;
        if (use_epsy) bump(counts, ORC_TB_EPSY);

        const double akz = 0.4 * z[k];
        const double al = akz * al0 / (al0 + akz);
        double al2 = al * al;
        if (!below && !use_epsy && ri1 > 0.0) {
            const double anlq2 = slq2 * ri1;
            if (anlq2 > 0.281) {   /* Deardorff limit (icondear = 0) */
                bump(counts, ORC_TB_DEARDORFF);
                al2 *= 0.281 / anlq2;
                slq2 = 0.281 / (ri1 + 1.0e-20);
            }
        }
        double epson2;
        if (an2[k] < 0.0) {
            epson2 = kEpson2;
        } else {
            double lat = nofsmall ? 0.0 : orc_turb_eplatidepend(fabs(coriol), an);
            lat = fmax(lat, kEplatMin);
            epson2 = kEpson2 * lat;
        }

        if (ri[k] > 0.0) {
            /* background mixing on the polar grid of (Ri_T, Ri_C) */
            const double rit = (ri[k] + rid[k]) / 2.0, ric = (ri[k] - rid[k]) / 2.0;
            const double ra_r = sqrt(rit * rit + ric * ric);
            double theta;
            if (fabs(rit) <= 0.0)
                theta = fabs(ric) <= 0.0 ? quarter : 2.0 * quarter;
            else
                theta = atan(ric / rit);
            if (fabs(theta) > 2.0 * quarter) return k;
            if (theta < -quarter) theta += 4.0 * quarter;
            const int j0 = (int)((theta + quarter) / dth);
            int i0 = j0 - NTH, i1 = i0 + 1;
            double th0 = i0 * dth, th1 = i1 * dth;
            if (th0 <= t->theta_rcrp && theta > t->theta_rcrp) {
                bump(counts, ORC_TB_THETA_P);
                theta = th1;
                th0 = th1;
                i0 = i1;
                ++i1;
                th1 += dth;
            } else if (th1 >= t->theta_rcrn && theta < t->theta_rcrn) {
                bump(counts, ORC_TB_THETA_N);
                theta = th0;
                th1 = th0;
                i1 = i0;
                --i0;
                th0 += -dth;
            }
            if (i1 > 3 * NTH || i0 < -NTH) return k;
            double w = theta - th0;
            const double *br = t->back_ra_r;
            double back_ra = br[NTH + i0] + w * ((br[NTH + i1] - br[NTH + i0]) / dth);
            int glt = 0;
            if ((theta <= t->theta_rcrp || theta >= t->theta_rcrn) && back_ra > ra_r) {
                glt = 1;
                back_ra = ra_r;
            }
            if (back_ra < 0.0) return k;
            const double brit = cos(theta) * back_ra, bric = sin(theta) * back_ra;
            back_ri1 = brit + bric;
            double back_rid1 = brit - bric;
            if (glt) {
                bump(counts, ORC_TB_RAFGLT);
                if (orc_turb_interp2d(t, &back_ri1, &back_rid1, o)) {
                    slq2_bk = o[0];
                    sm_bk = o[1];
                    sh_bk = o[2];
                    ss_bk = o[3];
                }
            } else {
                w = theta - i0 * dth;
                sm_bk = t->sm_r1[NTH + i0] + w * ((t->sm_r1[NTH + i1] - t->sm_r1[NTH + i0]) / dth);
                sh_bk = t->sh_r1[NTH + i0] + w * ((t->sh_r1[NTH + i1] - t->sh_r1[NTH + i0]) / dth);
                ss_bk = t->ss_r1[NTH + i0] + w * ((t->ss_r1[NTH + i1] - t->ss_r1[NTH + i0]) / dth);
                slq2_bk = t->slq2_r1[NTH + i0] +
                          w * ((t->slq2_r1[NTH + i1] - t->slq2_r1[NTH + i0]) / dth);
            }
            if (slq2_bk < 0.0) return k;
        }
        if (ri[k] <= 0.0) back_ri1 = 0.0;
        if (ri1 <= 0.0) sm_bk = sh_bk = ss_bk = 0.0;
        if (sm_bk < 0.0 || sh_bk < 0.0 || ss_bk < 0.0) return k;
        const double tb = 0.5 * b1 * b1 * back_ri1 * slq2_bk * epson2;
        const double v_bk = tb * sm_bk, t_bk = tb * sh_bk, s_bk = tb * ss_bk;
        if (v_bk < 0.0 || t_bk < 0.0 || s_bk < 0.0) return k;
        if (ri[k] > 0.0 && (fabs(v_bk) <= 0.0 || fabs(t_bk) <= 0.0 || fabs(s_bk) < 0.0)) return k;

        double tmp = 0.0;
        if (below && ri1 >= 0.0) {
            tmp = 0.5 * b1 * b1 * ri1 * slq2 * epson2;
        } else if (below && k >= 2) {
            /* length scale from the density curvature below the mixed layer */
            bump(counts, ORC_TB_DEEP);
            double dz, drh, d2rh;
            if (k == n - 1) {
                dz = z[k] - z[k - 1];
                drh = rh[k] - rh[k - 1];
                d2rh = rh[k] - 2.0 * rh[k - 1] + rh[k - 2];
            } else {
                dz = z[k + 1] - z[k - 1];
                drh = rh[k + 1] - rh[k - 1];
                d2rh = rh[k + 1] - 2.0 * rh[k] + rh[k - 1];
            }
            const double dzrh = drh / dz;
            const double d2zrh = 4.0 * d2rh / (dz * dz);
            const double ratio = dzrh / d2zrh;
            const double al0deep = 0.17 * fabs(ratio);
            const double akzd = 0.4 * z[k];
            const double aldeep = akzd * al0deep / (al0deep + akzd);
            al2 = aldeep * aldeep;
            if (!pure) {
                s2[k] = fmax(s2[k], kS2min);
                tmp = 0.5 * b1 * al2 * sqrt(s2[k] / (slq2 + 1.0e-40));
            }
        } else if (!pure) {
            s2[k] = fmax(s2[k], kS2min);
            if (use_epsy)
                tmp = 0.5 * epsy / (s2[k] + 1.0e-40);
            else
                tmp = 0.5 * b1 * al2 * sqrt(s2[k] / (slq2 + 1.0e-40));
        }
        if (pure) tmp = 0.5 * (b1 * b1) * al2 * sqrt(-an2[k] / amtaun2);
        akm[k] = fmin(tmp * sm + v_bk, fricmx);
        akh[k] = fmin(tmp * sh + t_bk, fricmx);
        aks[k] = fmin(tmp * ss + s_bk, fricmx);
    }
    for (int k = n > 0 ? n : 0; k < nmax; ++k) akm[k] = akh[k] = aks[k] = 0.0;
    if (n > 0) {
        if (akm[0] < wndmix) akm[0] = wndmix;
        if (akh[0] < wndmix) akh[0] = wndmix;
        if (aks[0] < wndmix) aks[0] = wndmix;
    }
    return -1;
}

void orc_turb_run(int imt, int jmt, int km, const int *kmt, const double *zkp, const double *dzp,
                  const orc_turb_tables *t, int ncc, double dfricmx, double dwndmix,
                  const orc_turb_fields *f, long *counts) {
    const size_t n2 = (size_t)imt * jmt;
    const double fricmx = dfricmx * 1.0e4, wndmix = dwndmix * 1.0e4;
    const double nccf = (double)(float)ncc;
    for (int k = 0; k < km; ++k) {
        memset(f->akmt + (size_t)k * n2, 0, n2 * sizeof(double));
        memset(f->ak_tide + (size_t)k * n2, 0, n2 * sizeof(double));
    }
    memset(f->amld, 0, n2 * sizeof(double));
    long total[ORC_TB_NCOUNT];
    memset(total, 0, sizeof(total));
#pragma omp parallel
    {
        long cnt[ORC_TB_NCOUNT];
        memset(cnt, 0, sizeof(cnt));
        double *col = (double *)calloc((size_t)10 * km, sizeof(double));
        double *z = col, *tt = col + km, *rh = col + 2 * km, *ri = col + 3 * km, *rid = col + 4 * km;
        double *s2 = col + 5 * km, *an2 = col + 6 * km, *akm = col + 7 * km, *akh = col + 8 * km,
               *aks = col + 9 * km;
#pragma omp for schedule(dynamic, 64)
        for (long q = 0; q < (long)n2; ++q) {
            const int i = (int)(q % imt), j = (int)(q / imt);
            if (!(i > 0 && i < imt - 1 && j > 0 && j < jmt - 1)) continue;
            if (!(f->vit[q] > 0.5)) continue;
            const int iwk = kmt[q] - 1;
            memset(col, 0, (size_t)10 * km * sizeof(double));
            for (int k = 0; k < iwk; ++k) {
                const size_t a = (size_t)k * n2 + q, b = a + n2;
                const double v = f->vit[b];
                z[k] = -v * zkp[k + 1] * 1.0e2;
                tt[k] = v * (f->at[a] - (f->at[a] - f->at[b]) * dzp[k] / (dzp[k] + dzp[k + 1]));
                ri[k] = v * f->rit[a];
                rid[k] = v * f->ricdttms[a];
                s2[k] = v * f->s2t[a];
                an2[k] = v * f->rict[a];
                rh[k] = v * (f->pdensity[a] + (f->pdensity[b] - f->pdensity[a]) * dzp[k] /
                                                  (dzp[k] + dzp[k + 1])) * 1.0e-3;
            }
            const double btur = f->vit[q] * f->buoytur[q] * 1.0e4;
            const double bsol = f->vit[q] * f->buoysol[q] * 1.0e4;
            double mld = 0.0;
            const int ex = orc_turb_column(t, iwk, km - 1, z, tt, rh, ri, rid, s2, an2, fricmx, wndmix,
                                           btur, bsol, f->fcort[q], &mld, akm, akh, aks, cnt);
            if (ex >= 0) cnt[ORC_TB_EXIT] += 1;
            f->amld[q] = mld * 1.0e-2;
            /* tidal mixing (TIDEMIX), then its minimum over the levels below */
            for (int k = 0; k < iwk; ++k) {
                const size_t a = (size_t)k * n2 + q;
                double v = kBackTide + kMixEf * kLocalFrac * f->wave_dis[q] * f->fz_tide[a] /
                                           (fmax(f->rict[a], 1.0e-8) * (rh[k] * 1000.0));
                if (v > kMaxTide) v = kMaxTide;
                f->ak_tide[a] = v;
                f->richardson[a] = f->rict[a];
                f->fztidal[a] = f->fz_tide[a];
                f->wp3_tidal[a] = rh[k];
            }
            for (int k = iwk - 2; k >= 0; --k) {
                const size_t a = (size_t)k * n2 + q;
                f->ak_tide[a] = fmin(f->ak_tide[a], f->ak_tide[a + n2]);
            }
            for (int k = 0; k < km - 1; ++k) {
                const size_t a = (size_t)k * n2 + q;
                const double tide = f->ak_tide[a];
                f->akmt[a] = akm[k] * 1.0e-4 + tide * 5.0;
                f->akt[a] = (akh[k] * 1.0e-4 + tide) / nccf;
                f->akt[(size_t)km * n2 + a] = (aks[k] * 1.0e-4 + tide) / nccf;
            }
        }
        free(col);
#pragma omp critical
        for (int m = 0; m < ORC_TB_NCOUNT; ++m) total[m] += cnt[m];
    }
    if (counts)
        for (int m = 0; m < ORC_TB_NCOUNT; ++m) counts[m] += total[m];
}
