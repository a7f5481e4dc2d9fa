/*
 * licom_tracer.h -- C ABI of the MI355X (gfx950) LICOM3 tracer time step.
 *
 * Drop-in for the reference's tracer path:
 *   extern "C" void tracer_(int ii)      tracer.cpp:1701-1903
 *   extern "C" void convadj_()           convadj.cpp:290-298
 * called from stepon_ (stepon.cpp:69,77).  The reference reaches every array
 * through file-scope device globals (cuda_data.h:35-345) and Fortran module
 * symbols (pconst_mod_mp_dts_ ...); this ABI passes the same arrays
 * explicitly, in the same layout:
 *
 *   Fortran (imt,jmt,km[,ntra]), i fastest == C [ntra][km][jmt][imt],
 *   imt = nx + 4, jmt = ny + 4 (two ghost cells, param_mod.h:18-20),
 *   atb is (imt,jmt,0:km,ntra): level k lives at plane k+1 (tracer_mod.h:9),
 *   ws is (imt,jmt,km+1) (dyn_mod.h:29), 2-D fields are [jmt][imt].
 *
 * Grid and coefficient inputs are host pointers (read once by create);
 * state, forcing and diagnostics are device pointers.  Every function returns
 * LICOM_OK (0) or a negative error code and never exits the process (the
 * reference exits through CHECK/iCheck, cuda_data.h:4-11, common.h:6-16).
 * One context serves one stream; contexts are not thread safe.
 */
#ifndef LICOM_TRACER_H
#define LICOM_TRACER_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LICOM_MAX_KM 64

enum licom_status {
    LICOM_OK = 0,
    LICOM_EINVAL = -1,   /* bad argument or grid (kmt > km, imt < 8, ...) */
    LICOM_ENOMEM = -2,   /* device allocation failed                      */
    LICOM_EHIP = -3,     /* HIP runtime error (message: licom_last_error)  */
    LICOM_ENODEV = -4,   /* no usable gfx950 device                        */
    LICOM_ENOTSUP = -5   /* option not built (scheme, precision)           */
};

enum licom_hmix { LICOM_HMIX_DEL2 = 2, LICOM_HMIX_DEL4 = 4 };
/* advection schemes, numbered as the reference's Str code
 * (tracer.cpp:1712-1714): centered and flux are advection_tracer_1
 * (tracer.cpp:120-289), tspas advection_tracer_2/3 (:291-921) */
enum licom_adv { LICOM_ADV_CENTERED = 0, LICOM_ADV_FLUX = 1, LICOM_ADV_TSPAS = 2 };

/* licom_tracer_step flags */
#define LICOM_STEP_CONVADJ 0x1      /* run convadj after the tracer update    */
#define LICOM_STEP_SIDE_EFFECTS 0x2 /* reproduce the reference's in-place
                                       mutations: utf,vtf *= oncc; h0f *= onbc;
                                       AKT surface floor written back
                                       (tracer.cpp:1088,1107-1108,1187,1315) */
#define LICOM_STEP_VTL_SCRATCH 0x4  /* leave the new S in vtl, as the reference
                                       does with its scratch reuse
                                       (tracer.cpp:1638,1891): the S of
                                       tracer_, before convadj and icesnow
                                       (with CONVADJ, convadj then runs as a
                                       separate pass over every column)       */
#define LICOM_STEP_STAGED 0x8       /* run one kernel per reference stage
                                       instead of the fused passes (same bits;
                                       a cross-check, not a fast path)        */
#define LICOM_STEP_ICESNOW 0x10     /* apply icesnow_ between the tracer update
                                       and convadj, as stepon_ does
                                       (stepon.cpp:69-75), inside the column
                                       solve; licomqice goes to the array set
                                       by licom_tracer_set_icesnow            */

/* Static grid (host pointers, copied by licom_tracer_create).  Replaces the
 * grid/pconst module arrays read through cuda_init.cpp:708-864. */
typedef struct licom_grid {
    int nx, ny, km;           /* physical extents; imt = nx+4, jmt = ny+4      */
    int imt_global;           /* zonal count in the AHF normalisation
                                 (hmix_del2.F90:346, hmix_del4.F90:357)        */
    const int32_t *kmt;       /* [jmt][imt] levels per T column; ghost cells
                                 are refilled (cyclic i, closed j)             */
    const double *zkp;        /* [km+1] interface depths: 0, -dz1, ...         */
    const double *pen;        /* [km-1] shortwave fraction times OD0CP, or NULL*/
    const double *htw, *hts;  /* [jmt][imt] west/south face lengths of T cells */
    const double *hun, *hue;  /* [jmt][imt] north/east face lengths of U cells */
    const double *tarea;      /* [jmt][imt] T-cell area                        */
    const double *au0, *aus, *auw, *ausw; /* [jmt][imt] T->U weights, or NULL
                                 for the lat-lon 0.25 (grid.F90:1654-1657)     */
    const double *to, *so;    /* [km] EOS reference T and S                    */
    const double *eosc;       /* [9][km] = Fortran C(km,9), kg/m^3             */
    const double *dxu, *dyu;  /* [jmt][imt] U-cell widths: the flux scheme's
                                 face transports (tracer.cpp:106-117); may be
                                 NULL for centered and tspas                   */
} licom_grid;

typedef struct licom_params {
    double dts;               /* tracer step (s), pconst_mod dts              */
    double onbc, oncc;        /* 1/(NBB*NCC+1), 1/(NCC+1)                      */
    double dwndmix;           /* AKT surface floor (m^2/s)                    */
    double gamma;             /* restoring rate (1/s)                         */
    double od0, od0cp;        /* 1/D0, 1/(D0*CP)                              */
    double ah;                /* AH_HOR (del2) or AH_BIHAR (del4)             */
    int hmix;                 /* LICOM_HMIX_DEL2 / LICOM_HMIX_DEL4            */
    int adv;                  /* LICOM_ADV_CENTERED / _FLUX / _TSPAS          */
    int boundary_restore;     /* 1: restoring term of tracer_7                 */
    int sssnorm;              /* 1: global freshwater normalisation of S       */
    int nss;                  /* pconst_mod nss: the centered scheme adds its
                                 advective increments / nss to ax, ay, az
                                 (tracer.cpp:140-159); >= 1 for centered      */
} licom_params;

/* Device arrays of one step (the d_* globals of cuda_data.h). */
typedef struct licom_fields {
    double *at;               /* [2][km][jmt][imt]        in/out               */
    double *atb;              /* [2][km+1][jmt][imt]      in/out               */
    double *akt;              /* [2][km][jmt][imt]        in (k=0 floor out)   */
    double *utf, *vtf;        /* [km][jmt][imt]           in (scaled out)      */
    const double *utl;        /* [km][jmt][imt]           in                   */
    double *vtl;              /* [km][jmt][imt]           in (S out if flag)   */
    double *ws;               /* [km+1][jmt][imt]         out (plane 0 in)     */
    double *h0f;              /* [jmt][imt]               in (scaled out)      */
    const double *h0l;        /* [jmt][imt]               in                   */
    const double *swv, *nswv, *sst, *seaice, *sss, *fresh; /* [jmt][imt] or NULL */
    const double *restore;    /* [2][km][jmt][imt] or NULL                     */
} licom_fields;

/* Device diagnostics; any pointer may be NULL (not produced).  Values are
 * written on the physical cells [2,jmt-2) x [2,imt-2); ghost cells are left
 * untouched.  tspas overwrites ax, ay, az with the step's advective
 * tendencies (tracer.cpp:664,809,917); centered adds them divided by nss
 * (an accumulation over the caller's diagnostic period, tracer.cpp:140-205);
 * flux leaves them alone (tracer.cpp:258-288). */
typedef struct licom_diag {
    double *ax, *ay, *az;     /* [2][km][jmt][imt] advective tendencies        */
    double *dz;               /* [2][km][jmt][imt] explicit vertical diffusion */
    double *tend;             /* [2][km][jmt][imt] total tendency              */
    double *dt_diff;          /* [2][km][jmt][imt] implicit diffusion          */
    double *dt_conv;          /* [2][km][jmt][imt] convective adjustment       */
    double *penetrate;        /* [km][jmt][imt]    shortwave penetration       */
    double *net;              /* [2][jmt][imt]     surface flux (every column) */
    double *fw_norm2;         /* [1] SSSNORM correction (device scalar)        */
} licom_diag;

/* Derived static fields, for inspection (host buffers sized by the caller). */
typedef struct licom_grid_derived {
    int32_t *kmt, *kmu;                      /* [jmt][imt]                     */
    double *tarea_r, *ohbt, *ohbu;           /* [jmt][imt]                     */
    double *dtn, *dts, *dte, *dtw, *ahf;     /* [jmt][imt]                     */
    double *dzp, *odzp, *odzt;               /* [km]                           */
    double area_t;                           /* out                            */
} licom_grid_derived;

typedef struct licom_ctx licom_ctx;

/* Creates a context on the current HIP device: validates and copies the grid,
 * derives the masks and coefficients of boundary.F90:29-52,
 * grid.F90:1096-1106/1343-1346/2029-2050 and hmix_del{2,4}.F90 init_del*t,
 * and allocates scratch for one step.  Replaces cudainit()/turbinit()
 * (stepon.cpp:50-51, cuda_init.cpp:372-864). */
int licom_tracer_create(const licom_grid *grid, const licom_params *params,
                        licom_ctx **out);

/* Binds the context to a hipStream_t (NULL = default stream). */
int licom_tracer_set_stream(licom_ctx *ctx, void *hip_stream);

/* One tracer step for T and S: tracer_(ii) of tracer.cpp:1701-1903, followed
 * by convadj_() (convadj.cpp:290) when LICOM_STEP_CONVADJ is set.  Enqueued
 * on the context's stream; returns after launch. */
int licom_tracer_step(licom_ctx *ctx, const licom_fields *f,
                      const licom_diag *d, int flags);

/* The same step as licom_tracer_step, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops): it
 * calls no hipMalloc, hipFree or allocating routine, and keeps the step's
 * scratch in a static device array of 160 MiB per device instead of the
 * context's allocation.  Grids whose scratch, 32*imt*jmt*km + 44*imt*jmt
 * bytes, exceeds it (e.g. larger than 360 x 218 x 55) get LICOM_EINVAL;
 * LICOM_STEP_STAGED, layouts that need the per-stage kernels (odd imt,
 * arrays not 16-byte aligned) and ICESNOW with VTL_SCRATCH get
 * LICOM_ENOTSUP.  Every call on a device shares that device's array:
 * calls that overlap on one device race.  No counterpart in the reference
 * (tracer.cpp:1701 allocates its scratch once at start-up, cuda_init.cpp). */
int licom_tracer_step_noalloc(licom_ctx *ctx, const licom_fields *f,
                              const licom_diag *d, int flags);

/* Mixed-precision form of the step: the tracer state at/atb and every
 * 2-D field stay FP64; velocities, ws, akt and the 3-D diagnostics are FP32.
 * Advective fluxes, the del4/del2 stencils and the explicit vertical
 * diffusion run in FP32 on differences formed in FP64 (anomalies about the
 * cell's own value), and the increments are added to the FP64 state; the
 * implicit vertical solve, convadj and the SSSNORM sum run in FP64.  Results
 * differ from the FP64 form at roundoff level of the fluxes (and where that
 * flips a limiter decision).  No reference counterpart: a variant of
 * tracer_ (tracer.cpp:1701-1903) for callers that keep FP32 dynamics. */
typedef struct licom_fields_mixed {
    double *at;               /* [2][km][jmt][imt]        in/out (FP64)       */
    double *atb;              /* [2][km+1][jmt][imt]      in/out (FP64)       */
    float *akt;               /* [2][km][jmt][imt]        in (k=0 floor out)  */
    float *utf, *vtf;         /* [km][jmt][imt]           in (scaled out)     */
    const float *utl, *vtl;   /* [km][jmt][imt]           in                  */
    float *ws;                /* [km+1][jmt][imt]         out (plane 0 in)    */
    double *h0f;              /* [jmt][imt]               in (scaled out)     */
    const double *h0l;        /* [jmt][imt]               in                  */
    const double *swv, *nswv, *sst, *seaice, *sss, *fresh; /* [jmt][imt] or NULL */
    const double *restore;    /* [2][km][jmt][imt] or NULL                    */
} licom_fields_mixed;

typedef struct licom_diag_mixed {
    float *ax, *ay, *az, *dz, *tend, *dt_diff, *dt_conv; /* [2][km][jmt][imt]   */
    float *penetrate;         /* [km][jmt][imt]                               */
    double *net;              /* [2][jmt][imt]                                */
    double *fw_norm2;         /* [1]                                          */
} licom_diag_mixed;

/* The step in the mixed-precision form; flags as licom_tracer_step except
 * LICOM_STEP_VTL_SCRATCH and LICOM_STEP_STAGED (LICOM_ENOTSUP).  Needs an
 * even imt and 16-byte aligned at/atb; tspas only (other schemes:
 * LICOM_ENOTSUP, as for batches). */
int licom_tracer_step_mixed(licom_ctx *ctx, const licom_fields_mixed *f,
                            const licom_diag_mixed *d, int flags);

/* ---- Batched ensemble engine (mixed-precision form) --------------------
 * Advances many members, each a context of its own grid (any size), in one
 * launch per stage and step: every stage kernel walks a work list of
 * (member, block) items; the k-march runs persistent workgroups over
 * (member, tile, tracer) items.  A member's results are bit-identical
 * whether it runs alone (licom_tracer_step_mixed), in any batch, or under
 * any launch-shape choice.  No reference counterpart (the reference runs one
 * domain per rank). */
typedef struct licom_batch licom_batch;

/* Launch-shape choices of one member (mapping only, never the arithmetic). */
typedef struct licom_member_tuning {
    int prelude_shape;        /* faces/continuity workgroup: 0: 64x4 columns,
                                 1: 32x8, 2: 128x2                          */
    int solve_shape;          /* column solve workgroup: 0: 4 rows x 32
                                 columns, 1: 2 x 64, 2: 1 x 128             */
} licom_member_tuning;

/* members[m] with its fields[m] and diags[m] (diags may be NULL, or an entry
 * with NULL pointers); tuning may be NULL (shape 0 everywhere).  All members
 * must live on the current device.  The batch owns a stream of its own
 * (default priority) unless licom_batch_set_stream is called.  Each member's
 * licomqice array (licom_tracer_set_icesnow) is taken at creation. */
int licom_batch_create(licom_ctx *const *members, const licom_fields_mixed *fields,
                       const licom_diag_mixed *diags, const licom_member_tuning *tuning,
                       int n, licom_batch **out);
int licom_batch_set_stream(licom_batch *b, void *hip_stream);
/* Workgroups of the persistent k-march; 0 = one workgroup per item. */
int licom_batch_set_march_grid(licom_batch *b, int workgroups);
/* on != 0: each step is ONE launch (steps and captured graphs alike): a
 * grid of co-resident workgroups (occupancy x CUs) walks every stage's work
 * list, with a grid-wide barrier between dependent phases (surface |
 * SSSNORM partials | SSSNORM total + prelude | k-march | solve | side effects
 * + halo).  Same results bit for bit as the default of one launch per stage.
 * The barrier waits are bounded: if the workgroups cannot all be resident
 * (other work sharing the device) a wait times out after about a second and
 * licom_batch_sync returns LICOM_EHIP (the members' state is then
 * undefined), so run persistent batches on an otherwise idle device. */
int licom_batch_set_persistent(licom_batch *b, int on);
/* Workgroups of the persistent step (0 before licom_batch_set_persistent). */
int licom_batch_persistent_grid(const licom_batch *b);
/* Test hook: makes the next persistent step's first barrier report a
 * timeout (every workgroup leaves at once), as licom_batch_sync then shows. */
int licom_batch_fail_next_barrier(licom_batch *b);
/* Enqueues nsteps steps of every member (flags: CONVADJ, SIDE_EFFECTS,
 * ICESNOW when every member has a licomqice array). */
int licom_batch_step(licom_batch *b, int nsteps, int flags);
/* Captures nsteps steps into a HIP graph (replacing any previous one) and
 * launches nothing; licom_batch_replay launches the graph. */
int licom_batch_capture(licom_batch *b, int nsteps, int flags);
int licom_batch_replay(licom_batch *b);
int licom_batch_sync(licom_batch *b);
void licom_batch_destroy(licom_batch *b);

/* Autotuner: times every licom_member_tuning candidate on this member (reps
 * steps each, HIP events, a one-member batch per candidate) in three
 * interleaved rounds, keeps each candidate's fastest round and returns the
 * fastest candidate with its milliseconds per step -- the default shape
 * {0, 0} unless another is faster by more than 2% (run-to-run noise is about
 * 1%).  Advances the member's state. */
int licom_autotune_member(licom_ctx *ctx, const licom_fields_mixed *f, const licom_diag_mixed *d,
                          int reps, licom_member_tuning *best, double *best_ms);

/* Convective adjustment alone: convadj_() of convadj.cpp:290-298 over every
 * column, ghost columns included. */
int licom_convadj(licom_ctx *ctx, const licom_fields *f, const licom_diag *d);

/* Surface freezing adjustment alone: icesnow_() of icesnow.cpp:70-85
 * (icesnow_cu, icesnow.cpp:15-68) on at [2][km][jmt][imt] over every column,
 * ghost columns included: a wet column whose surface T is below TBICE
 * (-1.8 degC, constant_mod.F90:168) is reset to TBICE, its licomqice takes
 * the cooling and its surface S the rejected salt; stored ice melts against
 * a warmer surface; T is clamped to TBICE at every level.  licomqice is a
 * device [jmt][imt] array; reset_qice != 0 zeroes it first, as the
 * reference's FRC_CORE build does (icesnow.cpp:76-78, def-undef.h:6). */
int licom_icesnow(licom_ctx *ctx, double *at, double *licomqice, int reset_qice);

/* Output array (device [jmt][imt]) of the icesnow that steps run with
 * LICOM_STEP_ICESNOW (FRC_CORE semantics: rewritten on every column each
 * step); NULL detaches it.  Replaces the d_licomqice global
 * (cuda_data.h, icesnow.cpp:79). */
int licom_tracer_set_icesnow(licom_ctx *ctx, double *licomqice);

/* Waits for the context's stream. */
int licom_tracer_sync(licom_ctx *ctx);

/* ---- State I/O (host side; no context, no GPU) ----
 * The restart stream that inirun.F90:598-700 reads: Fortran sequential
 * unformatted records (4-byte length, payload, length) of the global grid
 * nx x ny in double precision, i fastest, j north to south, big-endian as
 * the reference's -convert big_endian build (Makefile:4) reads and writes
 * them (little_endian != 0: native order):
 * h0; u, v, T, S, ws level by level; su, sv, swv, sshf, lthf, fresh; then
 * (number_month, number_day).  It is the concatenation of the reference's
 * fort.22 .. fort.27 (ssave-cdf-instant.F90:100-830, COUP off). */
typedef struct licom_restart {
    double *h0;                     /* [ny][nx]                              */
    double *u, *v;                  /* [km][ny][nx]                          */
    double *at;                     /* [2][km][ny][nx]: T, then S as (S-35)/1000 */
    double *ws;                     /* [km][ny][nx]                          */
    double *su, *sv, *swv, *sshf, *lthf, *fresh; /* [ny][nx]                 */
    int32_t number_month, number_day;
    int32_t little_endian;          /* 0: big-endian records (the reference's
                                       build); 1: little-endian             */
} licom_restart;

/* Reads a restart into caller-allocated host arrays; a NULL array skips its
 * records.  A record of another length, a bad trailing length or a short
 * file gives LICOM_EINVAL with the record index in licom_last_error(). */
int licom_restart_read(const char *path, int nx, int ny, int km, licom_restart *r);

/* Writes the stream (every field required).  Land is written as 0, as the
 * reference's writer zeroes it in place (ssave-cdf-instant.F90:173-191): h0,
 * T, S and ws below kmt, u and v below kmu ([ny][nx] global masks); the
 * caller's arrays are not modified. */
int licom_restart_write(const char *path, int nx, int ny, int km, const int32_t *kmt,
                        const int32_t *kmu, const licom_restart *r);

/* Reads dncoef.h1 as constant_mod.F90:331-364 does (list-directed: a header
 * line, km values of to; the same for so and po; per level a header line and
 * C(k,1:9)) and scales C by 1000 (g/cm^3 -> kg/m^3): eosc is [9][km], the
 * layout of licom_grid.eosc.  po may be NULL. */
int licom_eos_read(const char *path, int km, double *to, double *so, double *po, double *eosc);

/* Copies the derived masks and coefficients to host buffers. */
int licom_grid_export(const licom_ctx *ctx, licom_grid_derived *out);

/* Per-kernel timing with HIP events on the context stream.  When enabled,
 * every launch of a step is bracketed by events; licom_kernel_times returns
 * accumulated milliseconds and launch counts per kernel name. */
int licom_set_profiling(licom_ctx *ctx, int enable);
int licom_kernel_times(licom_ctx *ctx, int max, const char **names,
                       double *ms, int *counts, int *n_out);

/* Bytes of device scratch held by the context. */
size_t licom_scratch_bytes(const licom_ctx *ctx);

/* Number of columns that convadj mixed in the last fused step (waits for the
 * stream).  Diagnostic with no reference counterpart (the reference counts
 * convective levels in icmon only under LOWRES, convadj.cpp:247-254). */
int licom_last_mix_count(licom_ctx *ctx, int *count);

/* Diagnostic output of the tspas limiter's decisions: every step of the
 * fused FP64 path and of the mixed form writes, for each physical cell of
 * each tracer, 1 where the cell's predictor left its local bounds (its faces
 * take the upwind flux, advection_tracer_3, tracer.cpp:549-921) and 0
 * elsewhere, into flags ([2][km][jmt][imt] bytes, device); NULL stops it.
 * Used to count the decisions the mixed form takes differently from FP64.
 * No reference counterpart (the reference keeps them in its at00/atmax/atmin
 * scratch). */
int licom_tracer_set_flag_output(licom_ctx *ctx, void *flags);

/* Number of tracer columns (one tracer of one column) whose new state held a
 * NaN or an infinity at a wet level in the last step (waits for the
 * stream).  A batch
 * member's count is in its own context, so a member that goes non-finite is
 * flagged without touching the others (each member has its own arrays).
 * The reference aborts the run instead (energy_c.c:109-112). */
int licom_last_nonfinite(licom_ctx *ctx, int *count);

/* Device-to-device streaming copy of `bytes` on `hip_stream`, used to
 * calibrate achievable HBM bandwidth and profiler byte counters.  No
 * reference counterpart. */
int licom_stream_copy(void *dst, const void *src, size_t bytes, void *hip_stream);

void licom_tracer_destroy(licom_ctx *ctx);
const char *licom_strerror(int status);
const char *licom_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
