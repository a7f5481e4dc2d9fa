/* barotr.h -- CPU oracle of LICOM3-HIP's barotropic subcycle barotr_ (test
 * infrastructure only: imported by tests/, the smoke check and the bench's
 * CPU baseline leg, never by the product path).
 *
 * One subcycle nc of barotr_ (barotr.cpp:901-1024) split where the reference
 * updates halos; the caller exchanges halos between the stages:
 *   stage 0: the clears of wka and work, barotr1_cu, barotr1_cu1;
 *            halo of work
 *   stage 1: barotr2_cu1 and the viscosity (del2: barotr2_cu; del4:
 *            bhdel4_div, bhdel4_grad, bhdel4_zcurl, bhdel4_grad,
 *            barotr2_udel4_0, barotr2_udel4_1); halo of wka planes 2, 3
 *   stage 2: barotr3_cu, (del4: barotr3_tdel4_1), barotr3_cu1; halo of work
 *   stage 3: barotr4_cu.
 * Layout: LICOM's block, i fastest, 2-D [j][i]; wka holds 6 planes
 * [6][j][i] (the planes of the reference's 3-D wka that barotr_ uses).
 * The reference's barotr2_cu reads h0 one row past the block at
 * j = jmt-1 for wka planes 0, 1 of that row; those cells are overwritten in
 * stage 2 before anything reads them, and here the gradient there is 0.
 */
#ifndef ORC_BAROTR_H
#define ORC_BAROTR_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_barotr_grid {
    int imt, jmt, ib, ie, jb, je, del4;
    double dtb, am, ah;
    const int *kmt, *kmu, *kmtn, *kmts, *kmte, *kmtw;
    const double *au0, *aus, *auw, *ausw, *htw, *hts, *tarea_r, *uarea_r;
    const double *dxu, *dyu, *dxur, *dyur, *fcor, *ebea, *ebeb, *amf, *ahf;
    const double *dtn, *dts, *dte, *dtw;
    const double *duc, *dum, *dun, *dus, *due, *duw, *dmc, *dmn, *dms, *dme, *dmw;
} orc_barotr_grid;

typedef struct orc_barotr_fields {
    const double *vit, *viv;                         /* level 0 of the masks */
    const double *dzph, *dlub, *dlvb, *wgp, *pax, *pay, *pxb, *pyb, *whx, *why;
    double *h0, *h0p, *ub, *vb, *ubp, *vbp, *h0f, *h0bf, *work;
    double *wka;                                     /* [6][j][i] */
} orc_barotr_fields;

void orc_barotr_stage(const orc_barotr_grid *g, const orc_barotr_fields *f, int nc, int stage);

#ifdef __cplusplus
}
#endif
#endif
