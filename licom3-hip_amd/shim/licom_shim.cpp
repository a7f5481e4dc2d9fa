// licom_shim.cpp -- link-compatible replacements of the reference's
//   extern "C" void tracer_(int ii)   tracer.cpp:1701-1903
//   extern "C" void convadj_()        convadj.cpp:290-298
//   extern "C" void icesnow_()        icesnow.cpp:70-85
// on top of the C ABI (include/licom_tracer.h), for an in-place relink of an
// unmodified LICOM3-HIP build: compile this file with the build's block sizes
// and drop tracer.cpp, convadj.cpp and icesnow.cpp from the link.
//
// The shim reads what the reference's functions read: the Fortran module
// symbols (pconst_mod.h, grid.h, hmix_del*.h, pmix_mod.h) for the grid and
// the scalars, and the device arrays of cuda_data.h (allocated and filled by
// the reference's cudainit/steponinit).  It declares those symbols itself,
// with the reference's names and types, instead of including LICOM's headers,
// so it needs only the block sizes:
//   -DLICOM_NX_BLOCK=<nx_block> -DLICOM_NY_BLOCK=<ny_block> -DLICOM_KM=<km>
//   -DLICOM_IMT_GLOBAL=<imt_global>   (param_mod.h:19-25)
//   -DBIHAR                           (the biharmonic build, as the reference's)
// It keeps the reference's calling convention and its failure behaviour
// (message and exit(1), cuda_data.h:4-11).  One rank, one block per rank.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "licom_tracer.h"

#ifndef LICOM_NX_BLOCK
#define LICOM_NX_BLOCK 132
#endif
#ifndef LICOM_NY_BLOCK
#define LICOM_NY_BLOCK 100
#endif
#ifndef LICOM_KM
#define LICOM_KM 30
#endif
#ifndef LICOM_IMT_GLOBAL
#define LICOM_IMT_GLOBAL 720
#endif

namespace {
constexpr int kNx = LICOM_NX_BLOCK, kNy = LICOM_NY_BLOCK, kKm = LICOM_KM;
}

// Fortran module symbols (pconst_mod.h:25-203, grid.h:10-39,
// hmix_del4.h:14, hmix_del2.h:13, pmix_mod.h:7).
extern "C" {
extern double pconst_mod_mp_dts_, pconst_mod_mp_onbc_, pconst_mod_mp_oncc_;
extern double pconst_mod_mp_dwndmix_, pconst_mod_mp_gamma_, pconst_mod_mp_od0_, pconst_mod_mp_od0cp_;
extern int pconst_mod_mp_nss_, pconst_mod_mp_ist_, pconst_mod_mp_boundary_restore_;
extern char pconst_mod_mp_adv_tracer_[80];
extern double pconst_mod_mp_zkp_[kKm + 1];
extern double pconst_mod_mp_to_[kKm], pconst_mod_mp_so_[kKm];
extern double pconst_mod_mp_c_[9][kKm];
extern double pmix_mod_mp_pen_[kKm - 1];
extern int (*grid_mp_kmt_)[kNy][kNx];
extern double (*grid_mp_htw_)[kNy][kNx], (*grid_mp_hts_)[kNy][kNx];
extern double (*grid_mp_hun_)[kNy][kNx], (*grid_mp_hue_)[kNy][kNx];
extern double (*grid_mp_tarea_)[kNy][kNx];
extern double (*grid_mp_dxu_)[kNy][kNx], (*grid_mp_dyu_)[kNy][kNx];
#ifdef BIHAR
extern double hmix_del4_mp_ah_;
#else
extern double hmix_del2_mp_ah_;
#endif
}

// Device arrays of cuda_data.h:137-317 (C++ linkage, as the reference
// declares them).
extern double *d_at, *d_atb, *d_akt, *d_utf, *d_vtf, *d_utl, *d_vtl, *d_ws, *d_h0f, *d_h0l;
extern double *d_swv, *d_nswv, *d_sst, *d_seaice, *d_sss, *d_fresh, *d_restore;
extern double *d_ax, *d_ay, *d_az, *d_dz, *d_tend, *d_dt_diff, *d_dt_conv, *d_penetrate, *d_net;
extern double *d_licomqice;

namespace {

licom_ctx *g_ctx = nullptr;

[[noreturn]] void fail(const char *what, int rc) {
    std::fprintf(stderr, "licom %s: %s: %s\n", what, licom_strerror(rc), licom_last_error());
    std::exit(1);
}

// The scheme switch of tracer.cpp:1712-1714.
int scheme() {
    if (std::strncmp(pconst_mod_mp_adv_tracer_, "centered", 8) == 0) return LICOM_ADV_CENTERED;
    if (std::strncmp(pconst_mod_mp_adv_tracer_, "flux", 4) == 0) return LICOM_ADV_FLUX;
    return LICOM_ADV_TSPAS;
}

// First call: the grid and scalars the reference's kernels read
// (cuda_init.cpp:708-864), the coefficients of init_del4t/init_del2t and the
// masks are derived by licom_tracer_create.
void setup() {
    licom_grid g{};
    g.nx = kNx - 4;
    g.ny = kNy - 4;
    g.km = kKm;
    g.imt_global = LICOM_IMT_GLOBAL;
    g.kmt = &grid_mp_kmt_[0][0][0];
    g.zkp = pconst_mod_mp_zkp_;
    g.pen = pmix_mod_mp_pen_;
    g.htw = &grid_mp_htw_[0][0][0];
    g.hts = &grid_mp_hts_[0][0][0];
    g.hun = &grid_mp_hun_[0][0][0];
    g.hue = &grid_mp_hue_[0][0][0];
    g.tarea = &grid_mp_tarea_[0][0][0];
    g.dxu = &grid_mp_dxu_[0][0][0];
    g.dyu = &grid_mp_dyu_[0][0][0];
    g.to = pconst_mod_mp_to_;
    g.so = pconst_mod_mp_so_;
    g.eosc = &pconst_mod_mp_c_[0][0];
    licom_params p{};
    p.dts = pconst_mod_mp_dts_;
    p.onbc = pconst_mod_mp_onbc_;
    p.oncc = pconst_mod_mp_oncc_;
    p.dwndmix = pconst_mod_mp_dwndmix_;
    p.gamma = pconst_mod_mp_gamma_;
    p.od0 = pconst_mod_mp_od0_;
    p.od0cp = pconst_mod_mp_od0cp_;
#ifdef BIHAR
    p.ah = hmix_del4_mp_ah_;
    p.hmix = LICOM_HMIX_DEL4;
#else
    p.ah = hmix_del2_mp_ah_;
    p.hmix = LICOM_HMIX_DEL2;
#endif
    p.adv = scheme();
    p.boundary_restore = pconst_mod_mp_boundary_restore_;
    p.sssnorm = 1;                       // SSSNORM is defined in the active build
    p.nss = pconst_mod_mp_nss_;
    if (int rc = licom_tracer_create(&g, &p, &g_ctx)) fail("tracer_ setup", rc);
}

licom_fields fields() {
    licom_fields f{};
    f.at = d_at;
    f.atb = d_atb;
    f.akt = d_akt;
    f.utf = d_utf;
    f.vtf = d_vtf;
    f.utl = d_utl;
    f.vtl = d_vtl;
    f.ws = d_ws;
    f.h0f = d_h0f;
    f.h0l = d_h0l;
    f.swv = d_swv;
    f.nswv = d_nswv;
    f.sst = d_sst;
    f.seaice = d_seaice;
    f.sss = d_sss;
    f.fresh = d_fresh;
    f.restore = pconst_mod_mp_boundary_restore_ == 1 ? d_restore : nullptr;
    return f;
}

licom_diag diags() {
    licom_diag d{};
    d.ax = d_ax;
    d.ay = d_ay;
    d.az = d_az;
    d.dz = d_dz;
    d.tend = d_tend;
    d.dt_diff = d_dt_diff;
    d.dt_conv = d_dt_conv;
    d.penetrate = d_penetrate;
    d.net = d_net;
    return d;
}

}  // namespace

// tracer.cpp:1701: the in-place side effects of the reference (utf, vtf,
// h0f, the AKT floor) and its reuse of vtl, which is left holding S.
extern "C" void tracer_(int /*ii*/) {
    if (!g_ctx) setup();
    const licom_fields f = fields();
    const licom_diag d = diags();
    if (int rc = licom_tracer_step(g_ctx, &f, &d, LICOM_STEP_SIDE_EFFECTS | LICOM_STEP_VTL_SCRATCH))
        fail("tracer_", rc);
    pconst_mod_mp_ist_ += 1;             // tracer.cpp:1766
}

// convadj.cpp:290; returns after the work is done, as the reference's
// blocking copies that follow it imply.
extern "C" void convadj_() {
    if (!g_ctx) setup();
    const licom_fields f = fields();
    const licom_diag d = diags();
    if (int rc = licom_convadj(g_ctx, &f, &d)) fail("convadj_", rc);
    if (int rc = licom_tracer_sync(g_ctx)) fail("convadj_", rc);
}

// icesnow.cpp:70-85 with the FRC_CORE reset of licomqice (:76-78).
extern "C" void icesnow_() {
    if (!g_ctx) setup();
    if (int rc = licom_icesnow(g_ctx, d_at, d_licomqice, 1)) fail("icesnow_", rc);
}
