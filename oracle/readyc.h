/* readyc.h -- CPU oracle of the vertical-mixing preparation of LICOM3-HIP's
 * readyc_ (test infrastructure only: imported by tests/, the smoke check and
 * the bench's CPU baseline leg, never by the product path).
 *
 * Restates, in launch order over whole arrays: the clears of s2t, riu,
 * wp12, wp13, riv1, riv2 (readyc.cpp:1636-1643), readyc1 (:44-107),
 * readyc1_2 (:109-170), and after turb_cu the clear of akmu (:1679) and
 * readyc2 (:223-246).  Layout: LICOM's block, i fastest; 3-D [k][j][i];
 * rit and rict with km-1 planes.
 */
#ifndef ORC_READYC_H
#define ORC_READYC_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_readyc_grid {
    int imt, jmt, km;
    const double *at0, *atn, *ate, *atne;     /* [j][i] U->T weights */
    const double *au0, *aus, *auw, *ausw;     /* [j][i] T->U weights */
    const double *odzt;                       /* [km] */
} orc_readyc_grid;

typedef struct orc_readyc_fields {
    const double *vit, *viv, *up, *vp, *h0, *rict, *ric;
    double *h0bl, *h0bf;
    double *wp12, *wp13, *riv1, *riv2, *s2t, *rit, *s2u, *riu;
} orc_readyc_fields;

void orc_readyc_shear(const orc_readyc_grid *g, const orc_readyc_fields *f);
void orc_readyc_akmu(const orc_readyc_grid *g, const double *akmt, const double *viv, double *akmu);

#ifdef __cplusplus
}
#endif
#endif
