/*
 * oracle.h -- CPU restatement of the LICOM3 tracer time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker that the HIP library
 * is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product library never
 * links, loads or calls it.
 *
 * It restates, loop by loop and in the same floating-point operation order,
 * the active (HIP) semantics of the reference:
 *   tracer.cpp:1081-1699  (tracer_1_0 .. tracer_9, advection_tracer_*,
 *                          invtrit, operators_div, upwell_2..4)
 *   tracer.cpp:1701-1903  (host sequence of tracer_)
 *   convadj.cpp:29-288    (convadj_cu)
 *   icesnow.cpp:15-85     (icesnow_cu, icesnow_)
 * and the initialisation of the path's static inputs:
 *   boundary.F90:29-52, grid.F90:1096-1106, grid.F90:1343-1346,
 *   grid.F90:2029-2050, grid.F90:466, hmix_del2.F90:346-385,
 *   hmix_del4.F90:357-392.
 *
 * Arrays are C row-major with i fastest: 2-D [jmt][imt], 3-D [k][jmt][imt],
 * tracer arrays [n][k][jmt][imt], atb [n][km+1][jmt][imt] (level k at k+1).
 * imt = nx+4 and jmt = ny+4 (two ghost cells each side).
 */
#ifndef LICOM_ORACLE_H
#define LICOM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_grid {
    int imt, jmt, km;
    int imt_global;   /* zonal count used by the AHF normalisation          */
    int hmix;         /* 2 = Laplacian (hmix_del2), 4 = biharmonic (del4)    */
    /* primary inputs (caller fills) */
    int32_t *kmt;     /* [jmt][imt]; halos are (re)filled by orc_derive      */
    double *zkp;      /* [km+1] interface depths, zkp[0] = 0, negative down  */
    double *pen;      /* [km-1] shortwave fraction, already times OD0CP      */
    double *htw, *hts, *hun, *hue, *tarea;       /* [jmt][imt] metrics       */
    double *au0, *aus, *auw, *ausw;              /* [jmt][imt] T->U weights  */
    double *to, *so;  /* [km] EOS reference T, S (model units)               */
    double *eosc;     /* [9][km]: Fortran C(km,9) column-major, kg/m^3       */
    /* derived (caller allocates, orc_derive fills) */
    int32_t *kmu;     /* [jmt][imt]                                          */
    double *vit, *viv;/* [km][jmt][imt] 0/1 masks                            */
    double *tarea_r, *ohbt, *ohbu;               /* [jmt][imt]               */
    double *dtn, *dts, *dte, *dtw, *ahf;         /* [jmt][imt]               */
    double *dzp, *odzp, *odzt;                   /* [km]                     */
    double area_t;    /* ocean surface area over the interior               */
    /* U-cell widths (caller fills; read by the flux scheme only, else NULL) */
    double *dxu, *dyu;                           /* [jmt][imt]               */
} orc_grid;

typedef struct orc_params {
    double dts;       /* tracer time step (s)                               */
    double onbc, oncc;
    double dwndmix;   /* AKT surface floor                                  */
    double gamma;     /* restoring rate (1/s)                               */
    double od0, od0cp;
    double ah;        /* AH_HOR (del2) or AH_BIHAR (del4)                   */
    int boundary_restore;
    int sssnorm;      /* 1: global freshwater-flux normalisation of S       */
    int convadj;      /* 1: run convadj after the tracer step               */
    int adv;          /* 0 centered, 1 flux, 2 tspas (tracer.cpp:1712-1714)  */
    int nss;          /* steps per diagnostic period: the centered scheme
                         accumulates ax/ay/az increments divided by nss      */
} orc_params;

typedef struct orc_fields {
    double *at, *atb, *akt;               /* state / mixing                */
    double *utf, *vtf, *utl, *vtl;        /* [km][jmt][imt]                */
    double *ws;                           /* [km+1][jmt][imt]              */
    double *h0f, *h0l;                    /* [jmt][imt]                    */
    double *swv, *nswv, *sst, *seaice, *sss, *fresh;   /* [jmt][imt]       */
    double *restore;                      /* [2][km][jmt][imt] or NULL     */
} orc_fields;

typedef struct orc_diag {
    double *ax, *ay, *az, *dz;            /* [2][km][jmt][imt]             */
    double *tend, *dt_diff, *dt_conv;     /* [2][km][jmt][imt]             */
    double *penetrate;                    /* [km][jmt][imt]                */
    double *net;                          /* [2][jmt][imt]                 */
    double fw_norm2;                      /* SSSNORM correction applied    */
    double *tf;                           /* [2][km][jmt][imt] or NULL: the
                                             explicit tendency entering
                                             tracer_8 (for the pin tests)  */
} orc_diag;

void orc_halo_fill_f64(double *a, int imt, int jmt, int nz);
void orc_halo_fill_i32(int32_t *a, int imt, int jmt);
void orc_derive(orc_grid *g);
void orc_tracer_step(const orc_grid *g, const orc_params *p, orc_fields *f,
                     orc_diag *d);
void orc_convadj(const orc_grid *g, const orc_params *p, orc_fields *f,
                 orc_diag *d);
/* Stand-alone pieces used by the invariant tests. */
/* tspas advection only (advection_tracer_2/3) of one tracer T [km][jmt][imt]
 * with face velocities uw, vs and vertical velocity ws; writes adv_tt. */
void orc_advect(const orc_grid *g, double dts, const double *T, const double *uw,
                const double *vs, const double *ws, double *adv_tt);
/* advection_tracer_1_0 face transports uw, vs [km][jmt][imt] from the U-point
 * velocities wkd, wkb for scheme adv (0 centered, 1 flux, 2 tspas). */
void orc_faces(const orc_grid *g, int adv, const double *wkd, const double *wkb, double *uw,
               double *vs);
/* centered (adv=0) or flux (adv=1) advection of one tracer (advection_tracer_1,
 * tracer.cpp:120-289): writes adv_tt on the physical cells; the centered form
 * also adds its increments / nss to ax, ay, az (any of them may be NULL). */
void orc_advect_cf(const orc_grid *g, int adv, int nss, const double *T, const double *uw,
                   const double *vs, const double *ws, double *adv_tt, double *ax, double *ay,
                   double *az);
/* One 5-point del2 pass hdtk = fac*(cc*X + sum c*X_nbr) on the physical cells
 * of level k, fac = AH (del2 build) or AHF (first del4 pass). */
void orc_del2_level(const orc_grid *g, const double *X, int k, int use_ahf, double ah,
                    double *out);
void orc_invtri_column(int kz, const double *wkc_col, const double *odzt,
                       const double *odzp, double dts, double stf, double vit0,
                       double *x /* in: d, out: solution */);
double orc_dens(const double *eosc, int km, int k, double t, double s);
/* icesnow_ (icesnow.cpp:15-85): surface freezing adjustment and TBICE clamp of
 * at [2][km][jmt][imt] on every column; licomqice [jmt][imt] is zeroed first
 * when reset is nonzero (the FRC_CORE build). */
void orc_icesnow(const orc_grid *g, double *at, double *licomqice, int reset);
int orc_nthreads(void);
void orc_set_nthreads(int n);   /* OpenMP threads of the following calls */

#ifdef __cplusplus
}
#endif
#endif
