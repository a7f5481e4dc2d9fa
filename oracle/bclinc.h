/* bclinc.h -- CPU oracle of LICOM3-HIP's baroclinic momentum update bclinc_
 * (test infrastructure only: imported by tests/, the smoke check and the
 * bench's CPU baseline leg, never by the product path).
 *
 * bclinc_ (bclinc.cpp:728-827) is split where the reference updates halos
 * between launches; the caller exchanges halos between the stages:
 *   isc >= 1 (leapfrog): stage 0 = bclinc12, bclinc11, bclinc2, bclinc3;
 *                        halo of wka; stage 1 = bclinc4; halo of wka;
 *                        stage 2 = bclinc5;
 *   isc < 1 (first step): stage 0 = bclinc12, bclinc11, bclinc6, bclinc71,
 *                        bclinc72; halo of u and v; stage 1 = bclinc81,
 *                        bclinc82 (stage 2 does nothing).
 * Layout: LICOM's block, i fastest, 3-D [k][j][i].  The reference's 3-D
 * scratch (wka1, wka2, a8..f8) is local to each call.
 */
#ifndef ORC_BCLINC_H
#define ORC_BCLINC_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_bclinc_grid {
    int imt, jmt, km;
    double od0, c0f, cag, sag, onbb, dtc, dtc2, afc1, afc2;
    const int *kmu;                                        /* [j][i] */
    const double *snlat, *ohbt, *ohbu, *fcor, *dxur, *dyur; /* [j][i] */
    const double *epea, *epeb, *epla, *eplb;               /* [j][i] */
    const double *dzp, *odzt, *odzp, *zkt;                 /* [km]   */
} orc_bclinc_grid;

typedef struct orc_bclinc_fields {
    const double *vit, *viv, *gg, *akmu;                   /* [km][j][i] */
    const double *su, *sv, *psa, *h0bl, *ub, *vb;          /* [j][i]     */
    double *h0bf;                                          /* [j][i] in/out (*= onbb) */
    double *u, *v, *up, *vp, *dlu, *dlv, *wka, *utf, *vtf; /* [km][j][i] in/out */
    double *sbcx, *sbcy, *bbcx, *bbcy;                     /* [j][i] out (interior, wet) */
} orc_bclinc_fields;

void orc_bclinc_stage(const orc_bclinc_grid *g, const orc_bclinc_fields *f, int isc, int stage);

#ifdef __cplusplus
}
#endif
#endif
