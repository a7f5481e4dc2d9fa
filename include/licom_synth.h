/*
 * licom_synth.h -- deterministic synthetic inputs for the tracer step.
 *
 * Host-only C ABI (no GPU).  Produces the grid, the static coefficient
 * inputs and an initial state in the reference's own layout (i fastest,
 * imt = nx+4, jmt = ny+4, halos filled periodic in i and closed in j), so
 * the same buffers feed the HIP library and its CPU oracle.
 *
 * The reference reads these fields from files (grid.F90:587 read_vert_grid,
 * grid.F90:780-990 read_horiz_grid, grid.F90:1115-1240 read_topography,
 * constant_mod.F90:331-364 dncoef.h1, inirun.F90:540 initial S units); none
 * of those files ship with it, so this generator stands in for them.
 *
 * Random numbers: SplitMix64 counter streams, stream key
 * seed ^ (field_id << 32); field ids 1 bathymetry, 2 T noise, 3 S noise,
 * 4 streamfunction, 5 AKT, 6 ensemble T perturbation.  Every value depends
 * only on (seed, field, interior cell index), never on thread count.
 */
#ifndef LICOM_SYNTH_H
#define LICOM_SYNTH_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct licom_synth_cfg {
    int nx, ny, km;          /* interior extents                           */
    int lat_mode;            /* 0 uniform rows, 1 Mercator rows            */
    double dlon;             /* zonal spacing (degrees)                    */
    double lat_north;        /* uniform mode: northern edge of first row   */
    double dlat;             /* uniform mode: row spacing (degrees)        */
    double lat_c;            /* Mercator mode: centre latitude (degrees)   */
    uint64_t seed;
    double depth_min, depth_max;  /* uniform random depth range (m)        */
    double sigma;            /* Gaussian smoothing of the depth (cells)    */
    double land_depth;       /* columns shallower than this are land       */
    double dz_top, dz_bot, depth_total;  /* tanh-stretched levels          */
    double umax;             /* max |u| at the surface (m/s)               */
    double div_frac;         /* divergent (potential) share of the flow    */
    double t_noise, s_noise; /* N(0,.) amplitudes of the initial T, S      */
    double t_pert;           /* >0: extra N(0,t_pert) on T (field 6)       */
    double akt_lo, akt_hi;   /* log-uniform AKT range (m^2/s)              */
} licom_synth_cfg;

typedef struct licom_synth_grid {
    /* all host buffers, allocated by the caller */
    int32_t *kmt;                                  /* [jmt][imt]           */
    double *zkp;                                   /* [km+1]               */
    double *pen;                                   /* [km-1]               */
    double *htw, *hts, *hun, *hue, *tarea, *dxu, *dyu;   /* [jmt][imt]     */
    double *au0, *aus, *auw, *ausw;                /* [jmt][imt]           */
    double *to, *so;                               /* [km]                 */
    double *eosc;                                  /* [9][km]              */
    double *lat_t;                                 /* [jmt] degrees        */
    int imt_global;                                /* out                  */
} licom_synth_grid;

typedef struct licom_synth_state {
    double *at;        /* [2][km][jmt][imt]                                */
    double *atb;       /* [2][km+1][jmt][imt]                              */
    double *akt;       /* [2][km][jmt][imt]                                */
    double *utf, *vtf, *utl, *vtl;   /* [km][jmt][imt]                     */
    double *ws;        /* [km+1][jmt][imt]                                 */
    double *h0f, *h0l; /* [jmt][imt]                                       */
} licom_synth_state;

/* Fill cfg with the defaults shared by every workload (depth 500-5000 m,
 * sigma 4, land < 200 m, dz 10->300 m over 5000 m, |u| <= 0.5 m/s,
 * T noise 0.05, S noise 0.02, AKT 1e-5..1e-2). */
void licom_synth_defaults(licom_synth_cfg *cfg);
int licom_synth_make_grid(const licom_synth_cfg *cfg, licom_synth_grid *out);
int licom_synth_make_state(const licom_synth_cfg *cfg, const int32_t *kmt,
                           const double *zkp, licom_synth_state *out);

#ifdef __cplusplus
}
#endif
#endif
