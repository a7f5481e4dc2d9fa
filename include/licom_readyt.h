/*
 * licom_readyt.h -- C ABI of the thermodynamic preparation that precedes the
 * momentum and tracer steps in liblicom_tracer.so: potential density, the
 * hydrostatic pressure integrals, thermal expansion and saline contraction,
 * the Richardson-number buoyancy terms the Canuto producer reads, and the
 * baroclinic / barotropic pressure-gradient forcing.
 *
 * Drop-in boundary: licom_readyt_run replaces readyt_ (readyt.cpp:511-598):
 * the clears of alpha, beta, rit, ric, rict, ricdttms, ricdt and akt
 * (:536-543), the kernels readyt_cu_1 (:106-158), dens_cu (:37-77),
 * readyt_cu_2 (:160-214), readyt_cu_2_2 (:216-256), thermal_cu (:258-297),
 * readyt_cu_3 (:299-385), readyt_cu_3_2 (:387-427), readyt_cu_4 (:429-448),
 * readyt_cu_5 (:450-509), and the clear of ax, ay, az when ist == 0
 * (:591-595).  licom_readyt_create replaces the uploads of the grid and
 * pconst arrays those kernels read (cuda_init.cpp:708-864).
 *
 * Layout: LICOM's block, i fastest: 3-D fields [k][j][i] over imt x jmt (the
 * block with its halo), two-tracer fields [n][k][j][i], atb
 * [n][km+1][j][i] with level k at plane k+1, ric / rict / ricdttms / ricdt /
 * rit km-1 planes.  Every cell of the block is computed, halos included, as
 * the reference's kernels do.  Field pointers are device pointers; the grid
 * given to licom_readyt_create is host memory.
 *
 * Final contents (the reference reuses dlu and dlv): dlu and dlv hold the
 * baroclinic pressure gradient at levels 2..km-1; dlu planes 0 and 1 hold
 * the column integrals sum(gg*ohbt*dzp) and sum(gg*ohbt*dzp*zkt); dlv
 * planes 0 and 1 the barotropic terms readyt_cu_4 derives from them.
 */
#ifndef LICOM_READYT_H
#define LICOM_READYT_H

#include "licom_tracer.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Static grid and pconst arrays (host pointers, copied). */
typedef struct licom_readyt_grid {
    int imt, jmt, km;                       /* block extents, 2 <= km <= LICOM_MAX_KM */
    const int32_t *kmu;                     /* [jmt][imt] levels per U column         */
    const double *au0, *aus, *auw, *ausw;   /* [jmt][imt] T->U weights                */
    const double *dxur, *dyur;              /* [jmt][imt] 1/dxu, 1/dyu                */
    const double *ohbu, *ohbt;              /* [jmt][imt] 1/depth at U and T points   */
    const double *hbx, *hby;                /* [jmt][imt] pconst hbx, hby             */
    const double *dzp, *odzt, *zkt;         /* [km] thickness, 1/dz at the top
                                               interface, T-level depth            */
    const double *to, *so, *po;             /* [km] EOS reference T, S, density
                                               offset (dncoef.h1)                   */
    const double *eosc;                     /* [9][km] = Fortran C(km,9), kg/m^3      */
    double od0, od0cp;                      /* 1/D0, 1/(D0*CP)                        */
} licom_readyt_grid;

typedef struct licom_readyt licom_readyt;

/* Copies the grid to the current HIP device.  Errors: LICOM_EINVAL (NULL
 * array, bad extents), LICOM_ENODEV, LICOM_ENOMEM. */
int licom_readyt_create(const licom_readyt_grid *grid, licom_readyt **out);

/* Binds the context to a hipStream_t (NULL = default stream). */
int licom_readyt_set_stream(licom_readyt *r, void *hip_stream);

/* Device arrays of one call (readyt_'s globals). */
typedef struct licom_readyt_fields {
    /* inputs */
    const double *vit, *viv;                /* [km][j][i] T and U wet masks (0/1)     */
    const double *at;                       /* [2][km][j][i]                          */
    const double *atb;                      /* [2][km+1][j][i]                        */
    const double *h0;                       /* [j][i] new surface height              */
    const double *u, *v;                    /* [km][j][i] new velocities              */
    const double *psa, *swv, *nswv;         /* [j][i] surface pressure, shortwave,
                                               non-shortwave heat flux              */
    /* time levels, shifted: h0l = h0f, h0f = h0; utl = utf, utf = u; vtl = vtf,
       vtf = v */
    double *h0l, *h0f;                      /* [j][i]                                 */
    double *utl, *vtl, *utf, *vtf;          /* [km][j][i]                             */
    /* outputs */
    double *dlu, *dlv;                      /* [km][j][i] (contents: see above)      */
    double *pdensity;                       /* [km][j][i] (a plane km is not touched) */
    double *gg;                             /* [km][j][i] -od0*g*(rho - po - 1000)    */
    double *pp, *ppa, *ppb, *ppc;           /* [km][j][i] pressure integrals, T and S
                                               at the top interface                */
    double *alpha, *beta;                   /* [km][j][i] thermal expansion, saline
                                               contraction                          */
    double *ric, *rict, *ricdttms, *ricdt;  /* [km-1][j][i]                           */
    double *rit;                            /* [km-1][j][i] cleared                   */
    double *akt;                            /* [2][km][j][i] cleared                  */
    double *buoytur, *buoysol;              /* [j][i]                                 */
    double *pxb, *pyb;                      /* [j][i]                                 */
    double *wgp, *work, *whx, *why;         /* [j][i]                                 */
    double *work1, *work2, *pax, *pay;      /* [j][i]                                 */
    double *ax, *ay, *az;                   /* [2][km][j][i], cleared when ist == 0;
                                               may be NULL                          */
} licom_readyt_fields;

/* One readyt_ on the context's stream; ist is pconst_mod ist (the step
 * counter: 0 clears ax, ay, az).  Returns after launch.  Errors:
 * LICOM_EINVAL (NULL context or a NULL array other than ax/ay/az). */
int licom_readyt_run(licom_readyt *r, const licom_readyt_fields *f, int ist);

/* Waits for the context's stream. */
int licom_readyt_sync(licom_readyt *r);

/* Kernel times of the last run in ms (HIP events on the context's stream;
 * waits for it): ms[0] the whole run, ms[1] the column pass, ms[2] the
 * pressure-gradient pass, ms[3] the surface pass.  Zeros before the first
 * run. */
int licom_readyt_last_ms(licom_readyt *r, double ms[4]);

void licom_readyt_destroy(licom_readyt *r);

#ifdef __cplusplus
}
#endif
#endif
