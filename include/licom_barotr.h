/*
 * licom_barotr.h -- C ABI of the barotropic subcycle barotr_ in
 * liblicom_tracer.so: the free-surface predictor, the barotropic momentum
 * (viscosity, surface slope, forcing, semi-implicit Coriolis), the
 * corrector with the free-surface diffusion of every fourth subcycle, and
 * the time levels and sums.
 *
 * Drop-in boundary (barotr.cpp:885-1029).  barotr_ runs nbb subcycles, each
 * with three halo updates through LICOM's MPI halo library, which stays the
 * caller's; one subcycle nc becomes four calls with the caller's halo
 * exchanges between them:
 *   licom_barotr_stage(b, f, nc, 0)   the clears of wka and work, barotr1_cu,
 *                                     barotr1_cu1 (:903-913)
 *   halo exchange of work             (:915-917)
 *   licom_barotr_stage(b, f, nc, 1)   barotr2_cu1, then barotr2_cu (del2) or
 *                                     bhdel4_div/grad/zcurl/grad,
 *                                     barotr2_udel4_0/1 (del4) (:918-970)
 *   halo exchange of wka planes 2, 3  (:971-973)
 *   licom_barotr_stage(b, f, nc, 2)   barotr3_cu, (del4: barotr3_tdel4_1),
 *                                     barotr3_cu1 (:975-1011)
 *   halo exchange of work             (:1014-1016)
 *   licom_barotr_stage(b, f, nc, 3)   barotr4_cu (:1018-1020)
 * The caller advances isb after each subcycle (:1025).
 *
 * Layout: LICOM's block, i fastest, 2-D fields [j][i].  wka is the
 * reference's work array: planes 2 and 3 ([2..3][j][i]) carry the momentum
 * tendency from stage 1 to stage 2 and are written as the reference writes
 * them; planes 0, 1, 4, 5 are the reference's scratch, formed in registers
 * here and not written; other planes are untouched.  Field pointers are
 * device pointers; the grid given to licom_barotr_create is host memory.
 */
#ifndef LICOM_BAROTR_H
#define LICOM_BAROTR_H

#include "licom_tracer.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct licom_barotr_grid {
    int imt, jmt;                            /* block extents, >= 5               */
    int ib, ie, jb, je;                      /* POP's 1-based block bounds: del2
                                                2 <= jb, je <= jmt-1; del4
                                                3 <= jb, je <= jmt-2 (same for i) */
    int del4;                                /* 1: the BIHAR build, 0: del2        */
    double dtb, am, ah;                      /* barotropic step, viscosity, the
                                                free surface's diffusivity        */
    const int *kmt, *kmu;                    /* [jmt][imt]                        */
    const int *kmtn, *kmts, *kmte, *kmtw;    /* [jmt][imt] kmt of the neighbours  */
    const double *au0, *aus, *auw, *ausw;    /* [jmt][imt] T->U weights           */
    const double *htw, *hts, *tarea_r, *uarea_r;
    const double *dxu, *dyu, *dxur, *dyur;
    const double *fcor, *ebea, *ebeb;        /* Coriolis and its factors          */
    const double *amf, *ahf;                 /* del4 factors                      */
    const double *dtn, *dts, *dte, *dtw;     /* the T-grid Laplacian              */
    const double *duc, *dum, *dun, *dus, *due, *duw;   /* the momentum Laplacian  */
    const double *dmc, *dmn, *dms, *dme, *dmw;
} licom_barotr_grid;

typedef struct licom_barotr licom_barotr;

/* Copies the grid to the current HIP device (and, with del4, allocates two
 * 2-D scratch arrays).  Errors: LICOM_EINVAL (NULL array, bad extents or
 * bounds), LICOM_ENODEV, LICOM_ENOMEM. */
int licom_barotr_create(const licom_barotr_grid *grid, licom_barotr **out);

/* Binds the context to a hipStream_t (NULL = default stream). */
int licom_barotr_set_stream(licom_barotr *b, void *hip_stream);

typedef struct licom_barotr_fields {
    /* inputs */
    const double *vit, *viv;                 /* [j][i] level 0 of the wet masks   */
    const double *dzph;                      /* [j][i] column depth               */
    const double *dlub, *dlvb;               /* [j][i] baroclinic forcing         */
    const double *wgp, *pax, *pay, *pxb, *pyb, *whx, *why;   /* [j][i]            */
    /* inputs and outputs */
    double *h0, *h0p;                        /* [j][i] surface height, two levels */
    double *ub, *vb, *ubp, *vbp;             /* [j][i] barotropic velocities      */
    double *h0f, *h0bf;                      /* [j][i] sums over subcycles        */
    double *work;                            /* [j][i] carried between stages     */
    double *wka;                             /* [>= 4][j][i] planes 2, 3 carried  */
} licom_barotr_fields;

/* One stage of subcycle nc (above); returns after launch.  Errors:
 * LICOM_EINVAL (NULL context or field, stage outside 0..3). */
int licom_barotr_stage(licom_barotr *b, const licom_barotr_fields *f, int nc, int stage);

/* Waits for the context's stream. */
int licom_barotr_sync(licom_barotr *b);

/* Kernel time in ms of the last licom_barotr_stage (HIP events on the
 * context's stream; waits for it).  0 before the first call. */
int licom_barotr_last_ms(licom_barotr *b, double *ms);

void licom_barotr_destroy(licom_barotr *b);

#ifdef __cplusplus
}
#endif
#endif
