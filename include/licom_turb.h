/*
 * licom_turb.h -- C ABI of the Canuto vertical-mixing producer in
 * liblicom_tracer.so: the AKT (and AKMT) that the tracer step's implicit
 * vertical diffusion consumes.
 *
 * Drop-in boundary:
 *   licom_turb_create  replaces turbinit (turb_init.cpp:37-88): uploads the
 *                      Canuto lookup tables built by LICOM's Fortran turb_ini
 *                      (canuto_ini_2002.F90:192-970, canuto_mod arrays) and
 *                      turb_param / fcort;
 *   licom_turb_run     replaces the launch of turb_cu (readyc.cpp:1673-1678,
 *                      kernel readyc.cpp:1456-1611 with turb_device
 *                      :1116-1454) together with the clears that precede it
 *                      (readyc.cpp:1642, 1647, 1666-1671: amld, akmt,
 *                      ak_tide and the wk columns start at zero).  Built
 *                      with TIDEMIX and without BCKMEX (def-undef.h:23-24).
 *
 * Arrays are LICOM's block layout, i fastest: 3-D fields [k][j][i] over
 * imt x jmt (the block with its halo), akt [2][km][j][i].  All field
 * pointers are device pointers; the tables and grid arrays given to
 * licom_turb_create are host arrays.
 *
 * Written cells (as turb_cu): every (i, j) with 0 < i < imt-1, 0 < j < jmt-1
 * and vit(i,j,0) > 0.5; there akmt and akt get levels 0..km-2, ak_tide,
 * richardson, fztidal and wp3_tidal levels 0..kmt-2, amld the mixed-layer
 * depth.  akmt, ak_tide and amld are zero everywhere else; akt, richardson,
 * fztidal and wp3_tidal are left alone everywhere else.
 */
#ifndef LICOM_TURB_H
#define LICOM_TURB_H

#include "licom_tracer.h"

#ifdef __cplusplus
extern "C" {
#endif

#define LICOM_TURB_MT 1462   /* table half-width (canuto_mod.h: mt) */
#define LICOM_TURB_MT0 400   /* linear part of the tables (mt0) */
#define LICOM_TURB_NTH 75    /* background angles per octant (n_theta_r_oct) */

/* The canuto_mod arrays turbinit uploads (turb_init.cpp:58-80), host memory,
 * in the reference's layout: a table over index -MT..MT holds entry i at
 * [MT+i]; the 2-D tables hold (iri, irid) at [(MT+irid)*(2MT+1) + MT+iri];
 * the background tables hold itheta (-NTH..3NTH) at [NTH+itheta]. */
typedef struct licom_turb_tables {
    const double *rib;          /* 2MT+1: Ri and Ri_d table values */
    const int32_t *irimax;      /* 2MT+1: last realizable Ri index (-MT..MT) per Ri_d */
    const double *slq2b, *smb, *shb, *ssb;                      /* (2MT+1)^2 */
    const double *and2on2a1, *amtaun2a1, *sma1, *sha1, *ssa1;   /* 2MT+1 */
    const double *back_ra_r, *sm_r1, *sh_r1, *ss_r1, *slq2_r1;  /* 4NTH+1 */
    /* turb_param (turb_init.cpp:83) */
    double rri, rnd2on2, dri, deltheta_r, b1, theta_rcrp, theta_rcrn, dand2on2;
} licom_turb_tables;

typedef struct licom_turb licom_turb;

/* Uploads tables and grid constants for a block of imt x jmt x km (km <=
 * LICOM_MAX_KM).  kmt [jmt][imt], zkp km+1 levels (m, decreasing), dzp km
 * levels (m), fcort [jmt][imt] (Coriolis parameter at T points), ncc
 * (baroclinic steps per tracer step), dfricmx and dwndmix (m^2/s) are the
 * pconst_mod values.  Errors: LICOM_EINVAL (bad extents, NULL table,
 * rib not increasing, irimax out of range, ncc < 1), LICOM_ENODEV,
 * LICOM_ENOMEM. */
int licom_turb_create(int imt, int jmt, int km, const int32_t *kmt, const double *zkp,
                      const double *dzp, const double *fcort, const licom_turb_tables *tables,
                      int ncc, double dfricmx, double dwndmix, licom_turb **out);

/* Binds the producer to a hipStream_t (NULL = default stream). */
int licom_turb_set_stream(licom_turb *t, void *hip_stream);

/* Device pointers of one call (turb_cu's arguments, readyc.cpp:1456-1463). */
typedef struct licom_turb_fields {
    /* inputs */
    const double *at;           /* [km][j][i]: temperature (first tracer of at) */
    const double *pdensity;     /* [km][j][i] */
    const double *rit, *ricdttms, *s2t, *rict;   /* [km][j][i]: Ri, Ri_d, S^2, N^2 */
    const double *fz_tide;      /* [km][j][i] */
    const double *vit;          /* [km][j][i]: T-cell wet mask */
    const double *buoytur, *buoysol, *wave_dis;  /* [j][i] */
    /* outputs */
    double *akmt;               /* [km][j][i] */
    double *akt;                /* [2][km][j][i] */
    double *ak_tide;            /* [km][j][i] */
    double *amld;               /* [j][i], m */
    double *richardson, *fztidal, *wp3_tidal;    /* [km][j][i] */
} licom_turb_fields;

/* One turb_cu pass on the producer's stream; returns after launch. */
int licom_turb_run(licom_turb *t, const licom_turb_fields *f);

/* Waits for the producer's stream. */
int licom_turb_sync(licom_turb *t);

/* Per-kernel time of the last run in ms (HIP events on the producer's
 * stream; waits for it).  0 before the first run. */
int licom_turb_last_ms(licom_turb *t, double *ms);

void licom_turb_destroy(licom_turb *t);

#ifdef __cplusplus
}
#endif
#endif
