/* turb.h -- CPU oracle of LICOM3-HIP's Canuto AKT producer (test
 * infrastructure only: imported by tests/, the smoke check and the bench's
 * CPU baseline leg, never by the product path).
 *
 * Restates turb_cu (readyc.cpp:1456-1611) with turb_device (:1116-1454),
 * interp1d_expabs / interp2d_expabs (:915-1115), formld (:902-913) and
 * eplatidepend_fun (:882-896), plus the clears that precede the launch
 * (readyc.cpp:1642-1671: akmt, ak_tide, amld and the wk columns start at 0).
 * Tables are the ones turbinit uploads (turb_init.cpp:37-88), in the
 * reference's layout: index -MT..MT stored at MT+i, the 2-D tables with the
 * Ri index fastest ((MT+irid)*(2MT+1) + MT+iri), the background tables at
 * NTH+itheta (itheta -NTH..3NTH).
 */
#ifndef ORC_TURB_H
#define ORC_TURB_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_TURB_MT 1462   /* canuto_mod.h: mt */
#define ORC_TURB_MT0 400   /* canuto_mod.h: mt0 */
#define ORC_TURB_NTH 75    /* canuto_mod.h: n_theta_r_oct */

typedef struct orc_turb_tables {
    const double *rib;                        /* 2MT+1: Ri (and Ri_d) table values */
    const int *irimax;                        /* 2MT+1: last realizable Ri index per Ri_d */
    const double *slq2b, *smb, *shb, *ssb;    /* (2MT+1)^2 */
    const double *and2on2a1, *amtaun2a1, *sma1, *sha1, *ssa1;   /* 2MT+1: zero-shear table */
    const double *back_ra_r, *sm_r1, *sh_r1, *ss_r1, *slq2_r1;  /* 4NTH+1: background */
    /* turb_param (turb_init.cpp:83) */
    double rri, rnd2on2, dri, deltheta_r, b1, theta_rcrp, theta_rcrn, dand2on2;
} orc_turb_tables;

/* Branch counters of the column routine (for the tests' coverage checks). */
enum {
    ORC_TB_PURE = 0,      /* zero-shear 1-D table */
    ORC_TB_TABLE2D,       /* 2-D (Ri, Ri_d) table */
    ORC_TB_BELOW,         /* levels with ifbelow set */
    ORC_TB_DEEP,          /* deep length scale (ifbelow, Ri < 0, k >= 2) */
    ORC_TB_EPSY,          /* lifepsy: dissipation from the surface buoyancy flux */
    ORC_TB_DEARDORFF,     /* slq2 * Ri > 0.281 clamp */
    ORC_TB_RAFGLT,        /* background radius clamped to ra_r: 2-D table */
    ORC_TB_THETA_P,       /* theta moved past theta_rcrp */
    ORC_TB_THETA_N,       /* theta moved before theta_rcrn */
    ORC_TB_EXIT,          /* columns left early (turb_device's returns) */
    ORC_TB_NCOUNT
};

/* formld (readyc.cpp:902-913): depth where t first departs from t[0] by more
 * than 0.1, linearly interpolated; z[n-1] if it never does. */
double orc_turb_formld(int n, const double *z, const double *t);

/* eplatidepend_fun (readyc.cpp:884-896). */
double orc_turb_eplatidepend(double f, double an);

/* interp1d_expabs (readyc.cpp:915-975) over the 1-D zero-shear table; out =
 * {amtaun2, sm, sh, ss}.  Returns 1 when out was set, 0 when the routine
 * leaves its outputs alone. */
int orc_turb_interp1d(const orc_turb_tables *t, double x, double out[4]);

/* interp2d_expabs (readyc.cpp:977-1115); ri, rid are updated in place (the
 * reference clamps them through its pointer arguments); out = {slq2, sm, sh,
 * ss}.  Returns 0 (outputs left alone), 1 (interpolated) or 2 (beyond the
 * realizable Ri: zeros). */
int orc_turb_interp2d(const orc_turb_tables *t, double *ri, double *rid, double out[4]);

/* turb_device (readyc.cpp:1116-1454) for one column of n levels (nmax
 * outputs).  s2 is modified as in the reference.  akm/akh/aks must hold
 * zeros on entry.  counts (may be NULL) accumulates ORC_TB_* counters.
 * Returns the level of an early exit, or -1. */
int orc_turb_column(const orc_turb_tables *t, int n, int nmax, const double *z, const double *tt,
                    const double *rh, const double *ri, const double *rid, double *s2,
                    const double *an2, double fricmx, double wndmix, double buoytur,
                    double buoysol, double coriol, double *amld, double *akm, double *akh,
                    double *aks, long *counts);

typedef struct orc_turb_fields {
    /* inputs, [k][j][i] with i fastest (3-D: km levels) */
    const double *at;          /* temperature (the first tracer of at) */
    const double *pdensity, *rit, *ricdttms, *s2t, *rict, *fz_tide, *vit;
    const double *buoytur, *buoysol, *fcort, *wave_dis;   /* 2-D */
    /* outputs */
    double *akmt;              /* km levels: cleared, then k < km-1 of wet columns */
    double *akt;               /* 2 x km levels: k < km-1 of wet interior columns */
    double *ak_tide;           /* km levels: cleared, then k < kmt-1 */
    double *amld;              /* 2-D: cleared, then the mixed-layer depth (m) */
    double *richardson, *fztidal, *wp3_tidal;   /* km levels: k < kmt-1 only */
} orc_turb_fields;

/* turb_cu over the block (imt x jmt x km, kmt[j][i], zkp km+1 levels, dzp km
 * levels), with the clears of readyc.cpp:1642-1671.  counts may be NULL. */
void orc_turb_run(int imt, int jmt, int km, const int *kmt, const double *zkp, const double *dzp,
                  const orc_turb_tables *t, int ncc, double dfricmx, double dwndmix,
                  const orc_turb_fields *f, long *counts);

#ifdef __cplusplus
}
#endif
#endif
