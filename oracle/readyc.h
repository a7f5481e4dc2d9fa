/* readyc.h -- CPU oracle of the vertical-mixing preparation of LICOM3-HIP's
 * readyc_ (test infrastructure only: imported by tests/, the smoke check and
 * the bench's CPU baseline leg, never by the product path).
 *
 * Restates, in launch order over whole arrays: the clears of s2t, riu,
 * wp12, wp13, riv1, riv2 (readyc.cpp:1636-1643), readyc1 (:44-107),
 * readyc1_2 (:109-170), and after turb_cu the clear of akmu (:1679) and
 * readyc2 (:223-246), then the momentum tendencies (below).  Layout:
 * LICOM's block, i fastest; 3-D [k][j][i]; rit and rict with km-1 planes.
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

/* The momentum tendencies of readyc_ after readyc2 (readyc.cpp:1700-1793):
 * the clears, upwell_cu_0 + upwell_cu (:599-682, w at T points from u, v),
 * readyc3 (:248-271, w and face velocities at U points),
 * advection_momentum_cu (:684-719), the horizontal viscosity (del2:
 * hdiffu_del2_cu :273-341; del4: hdel4_div, hdel4_grad, hdel4_zcurl,
 * hdel4_grad, hdiffu_del4_0, hdiffu_del4_1 :343-576), readyc4 (:721-790:
 * depth means, surface and bottom stress) and readyc4_2 (:793-877:
 * explicit half of the vertical viscosity). */
typedef struct orc_readyc_mgrid {
    int ib, ie, jb, je;                       /* block bounds, 1-based (POP blocks)      */
    int del4;                                 /* 1: BIHAR build, 0: Laplacian            */
    double am, c0f, od0, cag, sag;
    const int *kmt, *kmu;                     /* [j][i] */
    const double *ohbu, *ohbt, *hts, *htw, *hue, *hun, *tarea_r, *uarea_r;
    const double *dxu, *dyu, *dxur, *dyur, *snlat, *amf;
    const double *duc, *dum, *dun, *dus, *due, *duw, *dmc, *dmn, *dms, *dme, *dmw;
    const double *dzp, *odzp;                 /* [km] */
} orc_readyc_mgrid;

typedef struct orc_readyc_mfields {
    const double *u, *v, *up, *vp;            /* [km][j][i] */
    const double *h0, *su, *sv;               /* [j][i] */
    const double *vit, *viv, *akmu;           /* [km][j][i] */
    double *ws;                               /* [km][j][i] in/out */
    double *dlu, *dlv, *wka;                  /* [km][j][i] out */
    double *dlub, *dlvb, *sbcx, *sbcy, *bbcx, *bbcy;   /* [j][i] out */
} orc_readyc_mfields;

/* The pieces, in launch order; orc_readyc_momentum runs them all. */
void orc_rm_upwell(const orc_readyc_grid *g, const orc_readyc_mgrid *m, const orc_readyc_mfields *f);
void orc_rm_advect(const orc_readyc_grid *g, const orc_readyc_mgrid *m, const orc_readyc_mfields *f);
void orc_rm_hdiff(const orc_readyc_grid *g, const orc_readyc_mgrid *m, const orc_readyc_mfields *f);
void orc_rm_stress(const orc_readyc_grid *g, const orc_readyc_mgrid *m, const orc_readyc_mfields *f);
void orc_rm_vdiff(const orc_readyc_grid *g, const orc_readyc_mgrid *m, const orc_readyc_mfields *f);
void orc_readyc_momentum(const orc_readyc_grid *g, const orc_readyc_mgrid *m, const orc_readyc_mfields *f);

#ifdef __cplusplus
}
#endif
#endif
