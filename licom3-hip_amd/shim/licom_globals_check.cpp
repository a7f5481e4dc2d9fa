// licom_globals_check.cpp -- the LICOM side of the shim's link interface,
// for checking the shim outside a LICOM build: definitions of the Fortran
// module symbols and cuda_data.h device pointers that licom_shim.cpp reads
// (the names and types of pconst_mod.h, grid.h, hmix_del2.h, pmix_mod.h and
// cuda_data.h).  Linked with the shim and liblicom_tracer.so into
// libshim_check.so, whose globals a test fills (ctypes) before calling
// tracer_ / convadj_ / icesnow_ as stepon_ would (stepon.cpp:69-77).
// Not part of a LICOM relink, where LICOM itself defines these symbols.
#ifndef LICOM_NX_BLOCK
#define LICOM_NX_BLOCK 132
#endif
#ifndef LICOM_NY_BLOCK
#define LICOM_NY_BLOCK 100
#endif
#ifndef LICOM_KM
#define LICOM_KM 30
#endif

namespace {
constexpr int kNx = LICOM_NX_BLOCK, kNy = LICOM_NY_BLOCK, kKm = LICOM_KM;
}

extern "C" {
double pconst_mod_mp_dts_, pconst_mod_mp_onbc_, pconst_mod_mp_oncc_;
double pconst_mod_mp_dwndmix_, pconst_mod_mp_gamma_, pconst_mod_mp_od0_, pconst_mod_mp_od0cp_;
int pconst_mod_mp_nss_ = 1, pconst_mod_mp_ist_, pconst_mod_mp_boundary_restore_;
char pconst_mod_mp_adv_tracer_[80] = "tspas";
double pconst_mod_mp_zkp_[kKm + 1];
double pconst_mod_mp_to_[kKm], pconst_mod_mp_so_[kKm];
double pconst_mod_mp_c_[9][kKm];
double pmix_mod_mp_pen_[kKm - 1];
int (*grid_mp_kmt_)[kNy][kNx];
double (*grid_mp_htw_)[kNy][kNx], (*grid_mp_hts_)[kNy][kNx];
double (*grid_mp_hun_)[kNy][kNx], (*grid_mp_hue_)[kNy][kNx];
double (*grid_mp_tarea_)[kNy][kNx];
double (*grid_mp_dxu_)[kNy][kNx], (*grid_mp_dyu_)[kNy][kNx];
double hmix_del2_mp_ah_, hmix_del4_mp_ah_;
}

double *d_at, *d_atb, *d_akt, *d_utf, *d_vtf, *d_utl, *d_vtl, *d_ws, *d_h0f, *d_h0l;
double *d_swv, *d_nswv, *d_sst, *d_seaice, *d_sss, *d_fresh, *d_restore;
double *d_ax, *d_ay, *d_az, *d_dz, *d_tend, *d_dt_diff, *d_dt_conv, *d_penetrate, *d_net;
double *d_licomqice;
