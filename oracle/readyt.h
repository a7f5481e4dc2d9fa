/* readyt.h -- CPU oracle of LICOM3-HIP's readyt_ (test infrastructure only:
 * imported by tests/, the smoke check and the bench's CPU baseline leg,
 * never by the product path).
 *
 * Restates readyt_ (readyt.cpp:511-598) kernel by kernel, in launch order,
 * over whole arrays: the clears (:531-543), readyt_cu_1 (:106-158), dens_cu
 * (:37-77) with dens_c_cu (:25-35), readyt_cu_2 (:160-214), readyt_cu_2_2
 * (:216-256), thermal_cu (:258-297), readyt_cu_3 (:299-385), readyt_cu_3_2
 * (:387-427), readyt_cu_4 (:429-448), readyt_cu_5 (:450-509) and the ist == 0
 * clear of ax/ay/az (:591-595).
 *
 * Layout: LICOM's block, i fastest; 3-D [k][j][i] over imt x jmt, two-tracer
 * arrays [n][k][j][i] (atb [n][km+1][j][i], level k at plane k+1),
 * ric/rict/ricdttms/ricdt/rit km-1 planes.
 */
#ifndef ORC_READYT_H
#define ORC_READYT_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_readyt_grid {
    int imt, jmt, km;
    const int *kmu;                          /* [j][i] */
    const double *au0, *aus, *auw, *ausw;    /* [j][i] T->U weights */
    const double *dxur, *dyur;               /* [j][i] 1/dxu, 1/dyu */
    const double *ohbu, *ohbt;               /* [j][i] */
    const double *hbx, *hby;                 /* [j][i] */
    const double *dzp, *odzt, *zkt;          /* [km] */
    const double *to, *so, *po;              /* [km] */
    const double *eosc;                      /* [9][km]: C(k,m) at (m-1)*km+k */
    double od0, od0cp;
} orc_readyt_grid;

typedef struct orc_readyt_fields {
    /* inputs */
    const double *vit, *viv;                 /* [km][j][i] */
    const double *at, *atb;                  /* [2][km][j][i], [2][km+1][j][i] */
    const double *h0, *u, *v;                /* [j][i], [km][j][i] x2 */
    const double *psa, *swv, *nswv;          /* [j][i] */
    /* time-level shift, in/out */
    double *h0l, *h0f;                       /* [j][i] */
    double *utl, *vtl, *utf, *vtf;           /* [km][j][i] */
    /* outputs */
    double *dlu, *dlv, *pdensity, *gg, *pp, *ppa, *ppb, *ppc, *alpha, *beta; /* [km][j][i] */
    double *ric, *rict, *ricdttms, *ricdt, *rit;                            /* [km-1][j][i] */
    double *akt;                                                            /* [2][km][j][i] */
    double *buoytur, *buoysol, *pxb, *pyb, *wgp, *work, *whx, *why, *work1, *work2, *pax,
        *pay;                                                               /* [j][i] */
    double *ax, *ay, *az;                     /* [2][km][j][i] or NULL */
} orc_readyt_fields;

/* LICOM's gravity (constant_mod.h:27) */
#define ORC_READYT_G 9.806

void orc_readyt(const orc_readyt_grid *g, const orc_readyt_fields *f, int ist);

/* thermal_cu's expression (readyt.cpp:272-294) at one cell: tt potential
 * temperature (C), ss salinity in model units (S-35)/1000, ppa the pressure
 * integral (Pa times -od0), vit the mask; returns alpha, beta. */
void orc_thermal(double tt, double ss, double ppa, double vit, double od0, double *alpha,
                 double *beta);

#ifdef __cplusplus
}
#endif
#endif
