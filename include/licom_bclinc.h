/*
 * licom_bclinc.h -- C ABI of the baroclinic momentum update bclinc_ in
 * liblicom_tracer.so: pressure gradient and semi-implicit Coriolis, the
 * implicit vertical viscosity of u and v, the barotropic replacement of the
 * depth mean, the Robert-Asselin filter and the time sums utf/vtf.
 *
 * Drop-in boundary (bclinc.cpp:728-827).  The reference updates halos
 * between its launches through LICOM's MPI halo library, which stays the
 * caller's; bclinc_ becomes up to three calls with the caller's halo
 * exchanges between them:
 *   isc >= 1:  licom_bclinc_stage(b, f, isc, 0)  bclinc12, bclinc11, bclinc2,
 *                                                bclinc3 (:741-803)
 *              halo exchange of wka              (:805-807)
 *              licom_bclinc_stage(b, f, isc, 1)  bclinc4 (:809-814)
 *              halo exchange of wka              (:816-818)
 *              licom_bclinc_stage(b, f, isc, 2)  bclinc5 (:820-824)
 *   isc < 1:   licom_bclinc_stage(b, f, isc, 0)  bclinc12, bclinc11, bclinc6,
 *                                                bclinc71, bclinc72 (:741-773)
 *              halo exchange of u and v          (:766-779: the reference
 *                                                exchanges host copies and
 *                                                feeds them to bclinc81/82)
 *              licom_bclinc_stage(b, f, isc, 1)  bclinc81, bclinc82 (:779-788)
 * The caller advances isc after the last stage (:826).  The reference's
 * scratch arrays wka1, wka2 and a8..f8 are not materialised; the context
 * holds two 3-D scratch arrays of the block for the tridiagonal solves.
 *
 * Layout: LICOM's block, i fastest: 3-D fields [k][j][i] over imt x jmt x
 * km, 2-D fields [j][i].  Field pointers are device pointers; the grid given
 * to licom_bclinc_create is host memory.
 */
#ifndef LICOM_BCLINC_H
#define LICOM_BCLINC_H

#include "licom_tracer.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct licom_bclinc_grid {
    int imt, jmt, km;                        /* block extents, imt, jmt >= 5,
                                                2 <= km <= LICOM_MAX_KM         */
    double od0, c0f, cag, sag;               /* 1/rho0, drag, turning angle      */
    double onbb, dtc, dtc2, afc1, afc2;      /* 1/(nbb+1), dtc, 2 dtc, filter    */
    const int *kmu;                          /* [jmt][imt] 0..km                  */
    const double *snlat, *ohbt, *ohbu;       /* [jmt][imt]                        */
    const double *fcor, *dxur, *dyur;        /* [jmt][imt]                        */
    const double *epea, *epeb, *epla, *eplb; /* [jmt][imt] Coriolis factors       */
    const double *dzp, *odzt, *odzp, *zkt;   /* [km]                              */
} licom_bclinc_grid;

typedef struct licom_bclinc licom_bclinc;

/* Copies the grid to the current HIP device and allocates the solve
 * scratch.  Errors: LICOM_EINVAL (NULL array, bad extents, kmu outside
 * 0..km), LICOM_ENODEV, LICOM_ENOMEM. */
int licom_bclinc_create(const licom_bclinc_grid *grid, licom_bclinc **out);

/* Binds the context to a hipStream_t (NULL = default stream). */
int licom_bclinc_set_stream(licom_bclinc *b, void *hip_stream);

typedef struct licom_bclinc_fields {
    /* inputs */
    const double *vit, *viv;                 /* [km][j][i] T and U wet masks      */
    const double *gg, *akmu;                 /* [km][j][i] buoyancy, viscosity     */
    const double *su, *sv, *psa, *h0bl;      /* [j][i] wind stress, surface
                                                pressure, surface height         */
    const double *ub, *vb;                   /* [j][i] barotropic velocities      */
    /* inputs and outputs */
    double *h0bf;                            /* [j][i] scaled by onbb in stage 0  */
    double *u, *v, *up, *vp;                 /* [km][j][i] two time levels        */
    double *dlu, *dlv;                       /* [km][j][i] tendencies in (readyc_),
                                                accelerations out                */
    double *wka;                             /* [km][j][i] work array the halo
                                                exchanges carry between stages   */
    double *utf, *vtf;                       /* [km][j][i] time sums              */
    double *sbcx, *sbcy, *bbcx, *bbcy;       /* [j][i] written where kmu > 0 in
                                                1 <= j < jmt-1, 1 <= i < imt-1   */
} licom_bclinc_fields;

/* One stage of bclinc_ (above); returns after launch.  Errors: LICOM_EINVAL
 * (NULL context or field, stage outside 0..2, stage 2 with isc < 1). */
int licom_bclinc_stage(licom_bclinc *b, const licom_bclinc_fields *f, int isc, int stage);

/* Waits for the context's stream. */
int licom_bclinc_sync(licom_bclinc *b);

/* Kernel time in ms of the last licom_bclinc_stage (HIP events on the
 * context's stream; waits for it).  0 before the first call. */
int licom_bclinc_last_ms(licom_bclinc *b, double *ms);

void licom_bclinc_destroy(licom_bclinc *b);

#ifdef __cplusplus
}
#endif
#endif
