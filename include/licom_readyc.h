/*
 * licom_readyc.h -- C ABI of the vertical-mixing preparation of readyc_ in
 * liblicom_tracer.so: the velocity shear and the Richardson numbers that the
 * Canuto producer (licom_turb.h) reads, and the momentum diffusivity it
 * derives from the producer's AKMT.
 *
 * Drop-in boundary (readyc.cpp:1614-1678 and 1684-1685, the part of readyc_
 * before the momentum terms):
 *   licom_readyc_shear replaces the clears of s2t, riu, wp12, wp13, riv1,
 *                      riv2 (readyc.cpp:1636-1643) and the kernels readyc1
 *                      (:44-107: h0bl = h0bf, h0bf = h0; up, vp averaged to
 *                      T points with the wet U cells' weights) and readyc1_2
 *                      (:109-170: vertical shear S^2 and Ri on the T grid,
 *                      s2t and rit; on the U grid, s2u and riu);
 *   licom_turb_run     (licom_turb.h) is the turb_cu launch that follows;
 *   licom_readyc_akmu  replaces the clear of akmu (:1679) and readyc2
 *                      (:223-246, launched at :1684): AKMT averaged to U
 *                      points.
 *
 * Layout: LICOM's block, i fastest: 3-D fields [k][j][i] over imt x jmt;
 * rit and rict have km-1 planes, every other 3-D field km.  Every cell of
 * the block is computed as the reference's kernels compute it, halos
 * included.  Field pointers are device pointers; the grid given to
 * licom_readyc_create is host memory.
 */
#ifndef LICOM_READYC_H
#define LICOM_READYC_H

#include "licom_tracer.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct licom_readyc_grid {
    int imt, jmt, km;                        /* block extents, 2 <= km <= LICOM_MAX_KM */
    const double *at0, *atn, *ate, *atne;    /* [jmt][imt] U->T weights (grid.F90)    */
    const double *au0, *aus, *auw, *ausw;    /* [jmt][imt] T->U weights               */
    const double *odzt;                      /* [km] 1/dz at the top interface        */
} licom_readyc_grid;

typedef struct licom_readyc licom_readyc;

/* Copies the grid to the current HIP device.  Errors: LICOM_EINVAL (NULL
 * array, bad extents), LICOM_ENODEV, LICOM_ENOMEM. */
int licom_readyc_create(const licom_readyc_grid *grid, licom_readyc **out);

/* Binds the context to a hipStream_t (NULL = default stream). */
int licom_readyc_set_stream(licom_readyc *r, void *hip_stream);

typedef struct licom_readyc_fields {
    /* inputs */
    const double *vit, *viv;                 /* [km][j][i] T and U wet masks (0/1)    */
    const double *up, *vp;                   /* [km][j][i] velocities at U points     */
    const double *h0;                        /* [j][i]                                 */
    const double *rict;                      /* [km-1][j][i] N^2 on the T grid (readyt) */
    const double *ric;                       /* [km-1][j][i] N^2 on the U grid (readyt) */
    /* h0bl = h0bf, h0bf = h0 */
    double *h0bl, *h0bf;                     /* [j][i]                                 */
    /* outputs */
    double *wp12, *wp13;                     /* [km][j][i] up, vp at T points          */
    double *riv1, *riv2;                     /* [km][j][i] vertical differences of up, vp */
    double *s2t;                             /* [km][j][i] S^2 at T points             */
    double *rit;                             /* [km-1][j][i] rict / S^2 (written where
                                                1 <= j and i < imt-1; else untouched) */
    double *s2u;                             /* [km][j][i] S^2 at U points (levels
                                                0..km-2; level km-1 untouched)        */
    double *riu;                             /* [km][j][i] ric / S^2 at U points,
                                                level k+1 from interface k            */
} licom_readyc_fields;

/* readyc1 + readyc1_2 with the clears before them; returns after launch.
 * Errors: LICOM_EINVAL (NULL context or field). */
int licom_readyc_shear(licom_readyc *r, const licom_readyc_fields *f);

/* readyc2 with the clear of akmu: akmu [km][j][i] from akmt [km][j][i] (the
 * producer's output) and viv. */
int licom_readyc_akmu(licom_readyc *r, const double *akmt, const double *viv, double *akmu);

/* Waits for the context's stream. */
int licom_readyc_sync(licom_readyc *r);

/* Kernel time of the last licom_readyc_shear in ms (HIP events on the
 * context's stream; waits for it).  0 before the first call. */
int licom_readyc_last_ms(licom_readyc *r, double *ms);

void licom_readyc_destroy(licom_readyc *r);

#ifdef __cplusplus
}
#endif
#endif
