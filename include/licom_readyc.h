/*
 * licom_readyc.h -- C ABI of the vertical-mixing preparation of readyc_ in
 * liblicom_tracer.so: the velocity shear and the Richardson numbers that the
 * Canuto producer (licom_turb.h) reads, the momentum diffusivity it
 * derives from the producer's AKMT, and the momentum tendencies that follow
 * (advection, horizontal and vertical viscosity, surface stress and bottom
 * drag, their depth means).
 *
 * Drop-in boundary (readyc.cpp:1614-1793, readyc_ without the BCKMEX
 * background diffusivity readyc1_3):
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

/* ---------------------------------------------------------------------
 * The momentum tendencies of readyc_ after readyc2 (readyc.cpp:1700-1793):
 *   licom_readyc_momentum replaces the clears of uk, vk, wka, work, dlu,
 *                      dlv, u_wface, v_sface (:1700-1716) and the kernels
 *                      upwell_cu_0 + upwell_cu (:599-682, launched
 *                      :1705-1708: w at T points from u, v), readyc3
 *                      (:248-271, :1718), advection_momentum_cu (:684-719,
 *                      :1723), the horizontal viscosity (BIHAR: hdel4_div,
 *                      hdel4_grad, hdel4_zcurl, hdel4_grad, hdiffu_del4_0,
 *                      hdiffu_del4_1, :343-576 launched :1754-1771; else
 *                      hdiffu_del2_cu :273-341, :1778), readyc4 (:721-790,
 *                      :1784) and readyc4_2 (:793-877, :1789).
 * The scratch arrays of the reference (uk, vk, u_wface, v_sface, hduk2,
 * hdvk2, am_factor, work) are not materialised; with del4 the context holds
 * two 3-D scratch arrays of the block (the first Laplacians of up and vp,
 * scaled by the viscosity factor). */
typedef struct licom_readyc_mgrid {
    int ib, ie, jb, je;                      /* the block's first and last physical
                                                points, 1-based as in POP's blocks
                                                (this_block, readyc.cpp:1621-1624);
                                                del2 needs 2 <= jb, je <= jmt-1,
                                                del4 3 <= jb, je <= jmt-2 (same for i) */
    int del4;                                /* 1: biharmonic (the BIHAR build), 0: del2 */
    double am;                               /* viscosity (hmix_del2/del4 am)           */
    double c0f, od0, cag, sag;               /* drag coefficient, 1/rho0, cos/sin of the
                                                turning angle                           */
    const int *kmt, *kmu;                    /* [jmt][imt] levels of T and U columns,
                                                0..km                                   */
    const double *ohbu, *ohbt;               /* [jmt][imt] 1/H at U and T points        */
    const double *hts, *htw, *hue, *hun;     /* [jmt][imt] cell face lengths            */
    const double *tarea_r, *uarea_r;         /* [jmt][imt] 1/area of T and U cells      */
    const double *dxu, *dyu, *dxur, *dyur;   /* [jmt][imt] U cell sides and inverses    */
    const double *snlat;                     /* [jmt][imt] hemisphere sign              */
    const double *amf;                       /* [jmt][imt] del4 viscosity factor        */
    const double *duc, *dum, *dun, *dus, *due, *duw;   /* [jmt][imt] momentum Laplacian */
    const double *dmc, *dmn, *dms, *dme, *dmw;         /* [jmt][imt] its rotation terms */
    const double *dzp, *odzp;                /* [km] layer thickness and inverse        */
} licom_readyc_mgrid;

/* Copies the momentum grid to the context's device (and, with del4,
 * allocates the scratch).  Errors: LICOM_EINVAL (NULL, block bounds out of
 * the ranges above, kmt or kmu outside 0..km), LICOM_ENOMEM. */
int licom_readyc_momentum_grid(licom_readyc *r, const licom_readyc_mgrid *g);

typedef struct licom_readyc_mfields {
    /* inputs */
    const double *u, *v;                     /* [km][j][i] velocities (U points)        */
    const double *up, *vp;                   /* [km][j][i] the other time level, the
                                                viscosity's and the drag's              */
    const double *h0, *su, *sv;              /* [j][i] surface height, wind stress      */
    const double *vit, *viv;                 /* [km][j][i] T and U wet masks            */
    const double *akmu;                      /* [km][j][i] from licom_readyc_akmu       */
    double *ws;                              /* [km][j][i] vertical velocity at T points:
                                                level 0 read, levels 1..km-1 of columns
                                                1 <= j < jmt-1, 1 <= i < imt-1 written */
    /* outputs */
    double *dlu, *dlv;                       /* [km][j][i] momentum tendencies (zero
                                                outside 1 <= j < jmt-1, 1 <= i < imt-1) */
    double *wka;                             /* [km][j][i] c0f |(up, vp)|               */
    double *dlub, *dlvb;                     /* [j][i] depth means of the tendencies
                                                plus (surface - bottom stress) / H      */
    double *sbcx, *sbcy, *bbcx, *bbcy;       /* [j][i] surface stress and bottom drag
                                                (written inside the frame only)         */
} licom_readyc_mfields;

/* The momentum tendencies; returns after launch.  Errors: LICOM_EINVAL
 * (NULL context or field, no momentum grid). */
int licom_readyc_momentum(licom_readyc *r, const licom_readyc_mfields *f);

/* Waits for the context's stream. */
int licom_readyc_sync(licom_readyc *r);

/* Kernel time in ms of the last licom_readyc_shear or
 * licom_readyc_momentum, whichever ran last (HIP events on the context's
 * stream; waits for it).  0 before the first call. */
int licom_readyc_last_ms(licom_readyc *r, double *ms);

void licom_readyc_destroy(licom_readyc *r);

#ifdef __cplusplus
}
#endif
#endif
