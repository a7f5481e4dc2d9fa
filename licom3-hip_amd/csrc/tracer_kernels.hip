// tracer_kernels.hip -- gfx950 kernels of the LICOM3 tracer step (FP64).
//
// Semantics follow the reference's active HIP path (tracer.cpp, convadj.cpp)
// expression by expression; the library is built with -ffp-contract=off so
// that every product and sum rounds exactly where the reference's does.
//
// Launch sequence of one step (both tracers advance in every launch):
//   k_surface   2-D   stf, T->U average, surface fluxes, SSSNORM integrand
//   k_sss_*     2-D   deterministic compensated global sum (SSSNORM)
//   k_faces     3-D   face velocities u_wface, v_sface
//   k_ws        col   divergence, column integral, continuity w
//   k_predict   3-D   tspas predictor and local bounds -> limiter flags
//   k_d2tk      3-D   first del2 pass of the biharmonic operator
//   k_correct   3-D   limited fluxes, mixing, explicit vertical terms
//   k_column    col   implicit vertical diffusion, tendencies, convadj
//   k_halo      ghost cells of at/atb (cyclic i, closed j)
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>

#include "tracer_device.h"
#include "tracer_internal.h"
#include "tracer_surface.h"

namespace licom {
namespace {

constexpr int KMAX = LICOM_MAX_KM;

__device__ __forceinline__ size_t i2(const KGrid &g, int j, int i) {
    return (size_t)j * g.imt + i;
}
__device__ __forceinline__ size_t i3(const KGrid &g, int k, int j, int i) {
    return ((size_t)k * g.jmt + j) * g.imt + i;
}
__device__ __forceinline__ double vitf(const KGrid &g, int k, int j, int i) {
    return (k < g.kmt[i2(g, j, i)]) ? 1.0 : 0.0;
}

// ---------------------------------------------------------------- surface
// (bodies in tracer_surface.h)
__global__ void k_surface(KGrid g, KPar p, SurfaceIO o) {
    surface_cell(g, p, o, blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y * blockDim.y + threadIdx.y);
}

__global__ __launch_bounds__(kRedThreads) void k_sss_partial(KGrid g, const double *__restrict__ temp11,
                                                             double *partials) {
    sss_partial_block(g, temp11, partials, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(kRedThreads) void k_sss_final(const double *__restrict__ partials, int nblk,
                                                           double area_t, int enabled, double *err_norm2,
                                                           double *out_copy) {
    sss_final_block(partials, nblk, area_t, enabled, err_norm2, out_copy);
}

// ------------------------------------------------------------ velocities
__device__ __forceinline__ double wkd_at(const double *utf, const double *utl, double oncc,
                                         size_t c) {
    return 0.5 * (utf[c] * oncc) + 0.5 * utl[c];   // tracer_1 (:1107-1119)
}

// advection_tracer_1_0 (tracer.cpp:83-119): centered and tspas faces from
// hts/htw, flux faces from the U-cell widths dxu/dyu
__global__ void k_faces(KGrid g, KPar p, const double *__restrict__ utf,
                        const double *__restrict__ utl, const double *__restrict__ vtf,
                        const double *__restrict__ vtl, double *uw, double *vs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= g.imt || j >= g.jmt) return;
    const size_t c = i3(g, k, j, i), q = i2(g, j, i);
    double v = 0.0, u = 0.0;
    if (p.adv == LICOM_ADV_FLUX) {
        if (i >= 1 && i < g.imt - 1)
            v = (wkd_at(vtf, vtl, p.oncc, c) * g.dxu[q] + wkd_at(vtf, vtl, p.oncc, c + 1) * g.dxu[q + 1]) * 0.25;
        if (j >= 1 && j < g.jmt - 1)
            u = (wkd_at(utf, utl, p.oncc, c - g.imt) * g.dyu[q - g.imt] + wkd_at(utf, utl, p.oncc, c) * g.dyu[q]) *
                0.25;
    } else {
        if (i >= 1 && i < g.imt - 1)
            v = (wkd_at(vtf, vtl, p.oncc, c) + wkd_at(vtf, vtl, p.oncc, c + 1)) * g.hts[q] * 0.25;
        if (j >= 1 && j < g.jmt - 1)
            u = (wkd_at(utf, utl, p.oncc, c - g.imt) + wkd_at(utf, utl, p.oncc, c)) * g.htw[q] * 0.25;
    }
    vs[c] = v;
    uw[c] = u;
}

// upwell_2, operators_div, upwell_3, upwell_4 (tracer.cpp:980-1080):
// barotropic-corrected continuity w, one column per lane.
__global__ void k_ws(KGrid g, KPar p, const double *__restrict__ utf,
                     const double *__restrict__ utl, const double *__restrict__ vtf,
                     const double *__restrict__ vtl, const double *__restrict__ work_u,
                     const double *__restrict__ stf_pre, double *ws) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i < 1 || i >= g.imt - 1 || j < 1 || j >= g.jmt - 1) return;
    const int kmt = g.kmt[i2(g, j, i)];
    const size_t qc = i2(g, j, i), qe = i2(g, j, i + 1), qn = i2(g, j - 1, i),
                 qne = i2(g, j - 1, i + 1);
    const double fc = 1.0 + work_u[qc] * g.ohbu[qc], fe = 1.0 + work_u[qe] * g.ohbu[qe];
    const double fn = 1.0 + work_u[qn] * g.ohbu[qn], fne = 1.0 + work_u[qne] * g.ohbu[qne];
    const double htw_c = g.htw[qc], htw_e = g.htw[qe], hts_c = g.hts[qc], hts_n = g.hts[qn];
    const double tr = g.tarea_r[qc];
    double wka[KMAX];
    double w = 0.0;
    for (int k = 0; k < g.km; ++k) {
        double d = 0.0;
        if (k <= kmt - 1) {
            const size_t o = (size_t)k * g.imt * g.jmt;
            const double ukc = fc * wkd_at(utf, utl, p.oncc, o + qc);
            const double uke = fe * wkd_at(utf, utl, p.oncc, o + qe);
            const double ukn = fn * wkd_at(utf, utl, p.oncc, o + qn);
            const double ukne = fne * wkd_at(utf, utl, p.oncc, o + qne);
            const double vkc = fc * wkd_at(vtf, vtl, p.oncc, o + qc);
            const double vke = fe * wkd_at(vtf, vtl, p.oncc, o + qe);
            const double vkn = fn * wkd_at(vtf, vtl, p.oncc, o + qn);
            const double vkne = fne * wkd_at(vtf, vtl, p.oncc, o + qne);
            d = 0.5 * ((uke + ukne) * htw_e - (ukc + ukn) * htw_c + (vke + vkc) * hts_c -
                       (vkne + vkn) * hts_n) * tr;
        }
        wka[k] = d;
        w = w - g.dzp[k] * d * (k < kmt ? 1.0 : 0.0);
    }
    const double ohbt = g.ohbt[qc];
    const double sc = 1.0 / (1.0 + stf_pre[qc] * ohbt);
    double prev = ws[qc];
    for (int k = 1; k < g.km; ++k) {
        const double cur = (k < kmt ? 1.0 : 0.0) * (prev + g.dzp[k - 1] * (w * ohbt + wka[k - 1]));
        ws[i3(g, k, j, i)] = cur * sc;
        prev = cur;
    }
}

// ------------------------------------------------------------------ tspas
using dev::PTerms;

// advection_tracer_2 predictor terms (tracer.cpp:311-414) read from global
// memory for cell (k,j,i).
__device__ __forceinline__ void predictor_terms(const KGrid &g, double dts, const double *T,
                                                const double *U, const double *V,
                                                const double *W, int k, int j, int i,
                                                PTerms &t) {
    const size_t c = i3(g, k, j, i), q = i2(g, j, i);
    const size_t sj = g.imt, sk = (size_t)g.imt * g.jmt;
    const double Tu = k > 0 ? T[c - sk] : 0.0;
    const double Td = k < g.km - 1 ? T[c + sk] : 0.0;
    const double Wk = W[c], Wb = k < g.km - 1 ? W[c + sk] : 0.0;
    dev::pterms(T[c], T[c + 1], T[c - 1], T[c + sj], T[c - sj], Tu, Td, U[c + 1], U[c], V[c],
                V[c - sj], Wk, Wb, g.tarea_r[q], g.htw[q + 1] * g.hun[q + 1], g.htw[q] * g.hun[q],
                g.hts[q] * g.hue[q], g.hts[q - sj] * g.hue[q - sj], k, g.km, g.odzp, g.odzt, dts, t);
}

// Value entering the 7-point bounds: at on ocean cells, the sentinel on land
// (at*vit + (1-vit)*wt of tracer.cpp:418-545, exact for vit in {0,1}).
__device__ __forceinline__ double bval(const KGrid &g, const double *T, int k, int j, int i,
                                       double sentinel) {
    return vitf(g, k, j, i) != 0.0 ? T[i3(g, k, j, i)] : sentinel;
}

// advection_tracer_2 (tracer.cpp:291-547): predictor and the limiter flag
// at00 > atmax || at00 < atmin, for both tracers.
__global__ void k_predict(KGrid g, KPar p, const double *__restrict__ at,
                          const double *__restrict__ uw, const double *__restrict__ vs,
                          const double *__restrict__ ws, uint8_t *flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= g.imt || j >= g.jmt) return;
    const size_t n3 = (size_t)g.imt * g.jmt * g.km, c = i3(g, k, j, i);
    const bool inside = i >= 1 && i < g.imt - 1 && j >= 1 && j < g.jmt - 1;
    for (int n = 0; n < 2; ++n) {
        uint8_t fl = 0;
        if (inside) {
            const double *T = at + n * n3;
            PTerms t;
            predictor_terms(g, p.dts, T, uw, vs, ws, k, j, i, t);
            const double a00 = dev::predictor(t, T[c], p.dts);
            const double lo = -1.0e10, hi = 1.0e10;
            double mx = bval(g, T, k, j, i, lo), mn = bval(g, T, k, j, i, hi);
            const int nb[4][2] = {{j - 1, i}, {j + 1, i}, {j, i - 1}, {j, i + 1}};
            for (int m = 0; m < 4; ++m) {
                mx = fmax(mx, bval(g, T, k, nb[m][0], nb[m][1], lo));
                mn = fmin(mn, bval(g, T, k, nb[m][0], nb[m][1], hi));
            }
            if (k > 0) {
                mx = fmax(mx, bval(g, T, k - 1, j, i, lo));
                mn = fmin(mn, bval(g, T, k - 1, j, i, hi));
            }
            if (k < g.km - 1) {
                mx = fmax(mx, bval(g, T, k + 1, j, i, lo));
                mn = fmin(mn, bval(g, T, k + 1, j, i, hi));
            }
            fl = (a00 > mx || a00 < mn) ? 1 : 0;
        }
        flags[n * n3 + c] = fl;
    }
}

// tracer_del4_1 (tracer.cpp:1238-1293): d2tk = AHF * del2(atb), on
// [1,jmt-1) x [1,imt-1); zero elsewhere.
__device__ __forceinline__ void del2_sums(const KGrid &g, const double *X, int k, int j, int i,
                                          double &hdtk, double &cc) {
    const size_t q = i2(g, j, i), sj = g.imt;
    hdtk = 0.0;
    cc = 0.0;
    if (k <= g.kmt[q] - 1) {
        double c;
        if (k <= g.kmt[q - sj] - 1) { c = g.dtn[q]; hdtk += c * X[q - sj]; cc -= c; }
        if (k <= g.kmt[q + sj] - 1) { c = g.dts[q]; hdtk += c * X[q + sj]; cc -= c; }
        if (k <= g.kmt[q + 1] - 1) { c = g.dte[q]; hdtk += c * X[q + 1]; cc -= c; }
        if (k <= g.kmt[q - 1] - 1) { c = g.dtw[q]; hdtk += c * X[q - 1]; cc -= c; }
    }
}

__global__ void k_d2tk(KGrid g, const double *__restrict__ atb, double *d2tk) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= g.imt || j >= g.jmt) return;
    const size_t n2 = (size_t)g.imt * g.jmt, n3 = n2 * g.km;
    const bool inside = i >= 1 && i < g.imt - 1 && j >= 1 && j < g.jmt - 1;
    for (int n = 0; n < 2; ++n) {
        double v = 0.0;
        if (inside) {
            const double *X = atb + n * (n3 + n2) + (size_t)(k + 1) * n2;
            double h, cc;
            del2_sums(g, X, k, j, i, h, cc);
            v = g.ahf[i2(g, j, i)] * (cc * X[i2(g, j, i)] + h);
        }
        d2tk[n * n3 + i3(g, k, j, i)] = v;
    }
}

__device__ __forceinline__ double akt_floored(const KPar &p, const double *akt, int n, int k,
                                              size_t c) {
    return dev::floor_akt(akt[c], k, n, p.hmix, p.dwndmix);
}

// advection_tracer_3 (tracer.cpp:549-921), tracer_del4_2 / tracer_4
// (:1167-1362), tracer_5 (:1395-1525), tracer_6 (:1527-1569),
// tracer_7 (:1571-1627), tracer_8 (:1629-1653) on the physical cells.
__global__ void k_correct(KGrid g, KPar p, const double *__restrict__ at,
                          const double *__restrict__ atb, const double *__restrict__ akt,
                          const double *__restrict__ uw, const double *__restrict__ vs,
                          const double *__restrict__ ws, const uint8_t *__restrict__ flags,
                          const double *__restrict__ d2tk, const double *__restrict__ stf_n,
                          const double *__restrict__ err_norm2, const double *swv,
                          const double *restore, double *ori, double *ax, double *ay, double *az,
                          double *dzd, double *pen_out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 2;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 2;
    const int k = blockIdx.z;
    if (i >= g.imt - 2 || j >= g.jmt - 2) return;
    const int km = g.km;
    const size_t n2 = (size_t)g.imt * g.jmt, n3 = n2 * km;
    const size_t c = i3(g, k, j, i), q = i2(g, j, i), sj = g.imt;
    const double *odzp = g.odzp, *odzt = g.odzt;
    const double vitc = vitf(g, k, j, i);
    const double aidif = 0.5;
    for (int n = 0; n < 2; ++n) {
        const double *T = at + n * n3;
        double adv;
        if (p.adv != LICOM_ADV_TSPAS) {
            // advection_tracer_1 (tracer.cpp:120-289)
            const double Tc = T[c];
            const double Tu = k > 0 ? T[c - n2] : 0.0, Td = k < km - 1 ? T[c + n2] : 0.0;
            const double Wb = k < km - 1 ? ws[c + n2] : 0.0;
            const dev::CFAdv r = dev::adv_cf(p.adv, Tc, T[c + 1], T[c - 1], T[c + sj], T[c - sj], Tu, Td,
                                             uw[c + 1], uw[c], vs[c], vs[c - sj], ws[c], Wb,
                                             g.tarea_r[q], k, km, odzp);
            if (p.adv == LICOM_ADV_CENTERED) {
                const double nss = (double)p.nss;
                if (ax) ax[n * n3 + c] = ax[n * n3 + c] + r.hx / nss;
                if (ay) ay[n * n3 + c] = ay[n * n3 + c] + r.hy / nss;
                if (az) az[n * n3 + c] = az[n * n3 + c] - r.vz / nss;
            }
            adv = r.adv;
        } else {
        const uint8_t *F = flags + n * n3;
        PTerms t;
        predictor_terms(g, p.dts, T, uw, vs, ws, k, j, i, t);
        const double tr = g.tarea_r[q];
        const double Tc = T[c];
        const bool fc = F[c], fe = F[c + 1], fw = F[c - 1], fs = F[c + sj], fn = F[c - sj];
        bool up1, up2, up3, up4;
        if (k == 0 || k == km - 1) {
            up1 = fc || (fe && i <= g.imt - 3);
            up2 = fc || (fw && i >= 2);
            up3 = fc || (fs && j <= g.jmt - 3);
            up4 = fc || (fn && j >= 2);
        } else {
            up1 = (fc || fe) && i <= g.imt - 3;
            up2 = (fc || fw) && i >= 2;
            up3 = (fc || fs) && j <= g.jmt - 3;
            up4 = (fc || fn) && j >= 2;
        }
        const double xy1 = up1 ? dev::up_xy1(T[c + 1], Tc, uw[c + 1], tr) : t.xy1;
        const double xy2 = up2 ? dev::up_xy2(Tc, T[c - 1], uw[c], tr) : t.xy2;
        const double xy3 = up3 ? dev::up_xy1(T[c + sj], Tc, vs[c], tr) : t.xy3;
        const double xy4 = up4 ? dev::up_xy2(Tc, T[c - sj], vs[c - sj], tr) : t.xy4;
        double zb1, zb2;
        if (k == 0) {
            zb1 = 0.0;
            zb2 = (fc || F[c + n2]) ? dev::up_zb2(Tc, T[c + n2], ws[c + n2], odzp[0]) : t.zb2;
        } else if (k == km - 1) {
            zb1 = (fc || F[c - n2]) ? dev::up_zb1(T[c - n2], Tc, ws[c], odzp[k]) : t.zb1;
            zb2 = 0.0;
        } else {
            zb1 = (fc || F[c - n2]) ? dev::up_zb1(T[c - n2], Tc, ws[c], odzp[k]) : t.zb1;
            zb2 = (fc || F[c + n2]) ? dev::up_zb2(Tc, T[c + n2], ws[c + n2], odzp[k]) : t.zb2;
        }
        const double xx = -(t.x0 + xy1 + xy2 + t.c1);
        const double yy = -(t.y0 + xy3 + xy4 + t.c2);
        const double zz = -(t.za + zb1 + zb2 + t.zc);
        adv = xx + yy + zz;
        if (ax) ax[n * n3 + c] = xx;
        if (ay) ay[n * n3 + c] = yy;
        if (az) az[n * n3 + c] = zz;
        }

        double tf = adv * vitc;
        // horizontal mixing
        {
            double h, cc;
            if (p.hmix == LICOM_HMIX_DEL4) {
                const double *X = d2tk + n * n3 + (size_t)k * n2;
                del2_sums(g, X, k, j, i, h, cc);
                h = p.ah * (cc * X[q] + h);
            } else {
                const double *X = atb + n * (n3 + n2) + (size_t)(k + 1) * n2;
                del2_sums(g, X, k, j, i, h, cc);
                h = p.ah * (cc * X[q] + h);
            }
            tf += h;
        }
        // shortwave penetration (T only) and the explicit half of vertical
        // diffusion; level 0 and km-1 receive their boundary terms exactly as
        // the reference's k==0 thread applies them (tracer.cpp:1421-1475).
        const double *B = atb + n * (n3 + n2) + n2;   // level k at B[k*n2]
        const double *A = akt + n * n3;
        if (n == 0) {
            const double sw = swv ? swv[q] : 0.0;
            double pv;
            if (k == 0) {
                const double wt1 = sw * g.pen[0] * vitf(g, 1, j, i);
                tf = tf - odzp[0] * wt1;
                pv = -wt1 * odzp[0];
            } else if (k == km - 1) {
                const double wt2 = sw * g.pen[km - 2] * vitf(g, km - 1, j, i);
                tf = tf + odzp[km - 1] * wt2;
                pv = wt2 * odzp[km - 1];
            } else {
                const double wt1 = sw * g.pen[k - 1] * vitc;
                const double wt2 = sw * g.pen[k] * vitf(g, k + 1, j, i);
                tf = tf + (wt1 - wt2) * odzp[k];
                pv = (wt1 - wt2) * odzp[k];
            }
            if (pen_out) pen_out[c] = pv;
        }
        double dzv;
        if (k == 0) {
            const double wt1 = akt_floored(p, A, n, 0, q) * (B[q] - B[n2 + q]) * odzt[1] *
                               vitf(g, 1, j, i);
            tf = tf - odzp[0] * wt1 * 0.5;
            dzv = -odzp[0] * wt1 * 0.5;
        } else if (k == km - 1) {
            const double wt2 = akt_floored(p, A, n, km - 2, (size_t)(km - 2) * n2 + q) *
                               (B[(size_t)(km - 2) * n2 + q] - B[(size_t)(km - 1) * n2 + q]) *
                               odzt[km - 1] * vitf(g, km - 1, j, i);
            tf = tf + odzp[km - 1] * wt2 * (1.0 - aidif);
            dzv = odzp[km - 1] * wt2 * 0.5;
        } else {
            const double wt1 = akt_floored(p, A, n, k - 1, c - n2) * (B[c - n2] - B[c]) *
                               odzt[k] * vitc;
            const double wt2 = akt_floored(p, A, n, k, c) * (B[c] - B[c + n2]) * odzt[k + 1] *
                               vitf(g, k + 1, j, i);
            tf = tf + odzp[k] * (wt1 - wt2) * 0.5;
            dzv = odzp[k] * (wt1 - wt2) * 0.5;
        }
        if (dzd) dzd[n * n3 + c] = dzv;
        if (k == 0) {
            if (n == 1) {
                if (g.kmt[q] > 0) tf = tf + stf_n[n2 + q] * (1.0 - aidif) * odzp[0];
                if (p.sssnorm) tf = tf + err_norm2[0] * vitc;
            } else if (g.kmt[q] > 0) {
                tf = tf + stf_n[q] * odzp[0] * 0.5;
            }
        }
        if (p.boundary_restore == 1 && restore && k >= 1) {
            const double lamda = 1.0 / (15.0 * 86400.0);
            tf = tf + vitc * (restore[n * n3 + c] - B[c]) * lamda;
        }
        ori[n * n3 + c] = T[c] + p.dts * tf;
    }
}

// --------------------------------------------------------------- columns
// invtrit (tracer.cpp:923-978) on one column held in x[] (in: explicit
// update, out: solution for k < kz).
__device__ void thomas(const KGrid &g, const KPar &p, const double *A, int n, size_t q,
                       size_t n2, int kz, double stf, double *x, double *e, double *f) {
    const double *odzt = g.odzt, *odzp = g.odzp;
    const double dts = p.dts;
    // bottom row (kz-1) and the top row, as the reference sets them last
    auto wkc = [&](int k) { return akt_floored(p, A, n, k, (size_t)k * n2 + q); };
    e[kz] = 0.0;
    f[kz] = 0.0;
    for (int k = kz - 1; k >= 0; --k) {
        double a, b, c;
        if (k == kz - 1) {
            a = (k >= 1) ? wkc(k - 1) * odzt[k] * odzp[k] * dts * 0.5 : odzp[0] * dts * 0.5;
            b = 1.0 + a;
            c = odzp[kz - 1] * dts * 0.5;
        } else if (k == 0) {
            a = odzp[0] * dts * 0.5;
            c = wkc(0) * odzt[1] * odzp[0] * dts * 0.5;
            b = 1.0 + c;
        } else {
            a = wkc(k - 1) * odzt[k] * odzp[k] * dts * 0.5;
            c = wkc(k) * odzt[k + 1] * odzp[k] * dts * 0.5;
            b = 1.0 + a + c;
        }
        const double g0 = 1.0 / (b - c * e[k + 1]);
        e[k] = a * g0;
        f[k] = (x[k] + c * f[k + 1]) * g0;
    }
    x[0] = (e[0] * stf + f[0]) * 1.0;
    for (int k = 1; k < kz; ++k) x[k] = (e[k] * x[k - 1] + f[k]) * 1.0;
}

using dev::dens;

__device__ __forceinline__ void convect(const KGrid &g, int kcon, double *T, double *S,
                                        double *rhoup, double *rholo) {
    dev::convect_column(g.eosc, g.to, g.so, g.dzp, g.km, kcon, T, S, rhoup, rholo);
}

// invtrit + halo-independent part of tracer_9 (tracer.cpp:1655-1699) +
// convadj (convadj.cpp:258-275) on the physical columns, plus the reference's
// in-place side effects on every column.
__global__ void k_column(KGrid g, KPar p, int flags, double *at, double *atb, double *akt,
                         const double *__restrict__ ori, const double *__restrict__ stf_n,
                         double *utf, double *vtf, double *h0f, double *tend, double *dt_diff,
                         double *dt_conv, int *counters) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= g.imt || j >= g.jmt) return;
    const int km = g.km;
    const size_t n2 = (size_t)g.imt * g.jmt, n3 = n2 * km, q = i2(g, j, i);
    if (flags & LICOM_STEP_SIDE_EFFECTS) {
        for (int k = 0; k < km; ++k) {
            utf[(size_t)k * n2 + q] *= p.oncc;
            vtf[(size_t)k * n2 + q] *= p.oncc;
        }
        h0f[q] = h0f[q] * p.onbc;
        if (i >= 1 && i < g.imt - 1 && j >= 1 && j < g.jmt - 1) {
            for (int n = 0; n < 2; ++n)
                if ((n == 0 || p.hmix == LICOM_HMIX_DEL2) && akt[n * n3 + q] < p.dwndmix)
                    akt[n * n3 + q] = p.dwndmix;
        }
    }
    if (i < 2 || i >= g.imt - 2 || j < 2 || j >= g.jmt - 2) return;
    const int kz = g.kmt[q];
    double x[2][KMAX], e[KMAX + 1], f[KMAX + 1];
    for (int n = 0; n < 2; ++n) {
        const double *O = ori + n * n3;
        for (int k = 0; k < km; ++k) x[n][k] = O[(size_t)k * n2 + q];
        const double stf = stf_n[n * n2 + q];
        if (kz > 0) {
            // wkc is akt with the floor applied to level 0 (tracer.cpp:1315/1187)
            thomas(g, p, akt + n * n3, n, q, n2, kz, stf, x[n], e, f);
        }
        const double *T = at + n * n3;
        for (int k = 0; k < km; ++k) {
            const size_t c = (size_t)k * n2 + q;
            const double vk = k < kz ? 1.0 : 0.0;
            if (dt_diff) {
                double v = (x[n][k] - O[c]) / p.dts * vk;
                if (k == 0) v = v - stf * 0.5 * g.odzp[0];
                dt_diff[n * n3 + c] = v;
            }
            if (tend) tend[n * n3 + c] = (x[n][k] - T[c]) / p.dts * vk;
        }
        bool nonfinite = false;   // counted as in the fused solve
        for (int k = 0; k < km; ++k) nonfinite |= !__builtin_isfinite(x[n][k]);
        if (nonfinite && counters) atomicAdd(counters + 1, 1);
    }
    if (flags & LICOM_STEP_CONVADJ) {
        // x[] holds at after tracer_9 (atb == at); convadj mixes a copy
        double y[2][KMAX];
        for (int n = 0; n < 2; ++n)
            for (int k = 0; k < km; ++k) y[n][k] = x[n][k];
        convect(g, kz, y[0], y[1], e, f);
        for (int n = 0; n < 2; ++n)
            for (int k = 0; k < km; ++k) {
                const size_t c = (size_t)k * n2 + q;
                const double vk = k < kz ? 1.0 : 0.0;
                const double dc = (y[n][k] - x[n][k]) / p.dts * vk;
                if (dt_conv) dt_conv[n * n3 + c] = dc;
                if (tend) tend[n * n3 + c] += dc;
                x[n][k] = y[n][k];
            }
    }
    for (int n = 0; n < 2; ++n)
        for (int k = 0; k < km; ++k) {
            const size_t c = (size_t)k * n2 + q;
            at[n * n3 + c] = x[n][k];
            atb[n * (n3 + n2) + n2 + c] = x[n][k];
        }
}

// Ghost cells of at and atb (levels 1..km of the 0:km array): cyclic in i,
// zero in the closed rows (POP_HaloUpdate with fillValue 0, as
// pop_haloupdate_tracer2.F90:10 followed by tracer_9's copy).
__global__ void k_halo(int imt, int jmt, int nplanes, double *at, double *atb, int km) {
    (void)nplanes;
    halo_cell(imt, jmt, km, blockIdx.y, blockIdx.x * blockDim.x + threadIdx.x, at, atb);
}

// convadj_cu over every column, ghost columns included (convadj.cpp:29-288).
__global__ void k_convadj(KGrid g, KPar p, double *at, double *atb, double *tend,
                          double *dt_conv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= g.imt || j >= g.jmt) return;
    const int km = g.km;
    const size_t n2 = (size_t)g.imt * g.jmt, n3 = n2 * km, q = i2(g, j, i);
    const int kz = g.kmt[q];
    double y[2][KMAX], ru[KMAX], rl[KMAX];
    for (int n = 0; n < 2; ++n)
        for (int k = 0; k < km; ++k) y[n][k] = at[n * n3 + (size_t)k * n2 + q];
    convect(g, kz, y[0], y[1], ru, rl);
    const bool inner = i >= 2 && i < g.imt - 2 && j >= 2 && j < g.jmt - 2;
    for (int n = 0; n < 2; ++n)
        for (int k = 0; k < km; ++k) {
            const size_t c = (size_t)k * n2 + q;
            double *B = atb + n * (n3 + n2) + n2;
            const double dc = (y[n][k] - B[c]) / p.dts * (k < kz ? 1.0 : 0.0);
            if (inner) {
                if (dt_conv) dt_conv[n * n3 + c] = dc;
                if (tend) tend[n * n3 + c] += dc;
            }
            at[n * n3 + c] = y[n][k];
            B[c] = y[n][k];
        }
}

// icesnow_cu (icesnow.cpp:15-68) on every column, ghosts included; `reset`
// folds in the FRC_CORE zeroing of licomqice (icesnow.cpp:76-78).  Cells the
// reference leaves alone are not written.
__global__ __launch_bounds__(256) void k_icesnow(KGrid g, double *__restrict__ at,
                                                 double *__restrict__ qice, int reset) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= g.imt || j >= g.jmt) return;
    const size_t n2 = (size_t)g.imt * g.jmt, n3 = n2 * g.km, q = i2(g, j, i);
    const double tb = dev::kTbice;
    double T0 = at[q], S0 = at[n3 + q], qi = reset ? 0.0 : qice[q];
    const double T0in = T0, S0in = S0;
    if (T0 < tb && g.kmt[q] > 0) {
        const double tdiff = tb - T0;
        qi += tdiff * g.dzp[0] / g.dzp[0];
        S0 = S0 + dev::ice_salt(tdiff);
        T0 = tb;
    }
    if (qi > 0 && T0 > tb) {
        const double tdiff = fmin(T0 - tb, qi);
        qi -= tdiff;
        T0 -= tdiff;
        S0 = S0 - dev::ice_salt(tdiff);
    }
    if (T0 < tb) T0 = tb;
    if (T0 != T0in) at[q] = T0;
    if (S0 != S0in) at[n3 + q] = S0;
    qice[q] = qi;
    for (int k = 1; k < g.km; ++k) {
        double *p = at + (size_t)k * n2 + q;
        if (*p < tb) *p = tb;
    }
}

// 16-byte vector copy used for bandwidth calibration: one element per lane,
// one pass over the grid (faster on gfx950 than a grid-stride loop).
__global__ __launch_bounds__(256) void k_stream_copy(const double2 *__restrict__ src,
                                                     double2 *__restrict__ dst, size_t n) {
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t < n) dst[t] = src[t];
}

__global__ void k_copy_plane(const double *src, double *dst, size_t n) {
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (size_t)gridDim.x * blockDim.x)
        dst[t] = src[t];
}

// ------------------------------------------------------------ launch glue
inline dim3 grid2(int imt, int jmt, dim3 b) {
    return dim3((imt + b.x - 1) / b.x, (jmt + b.y - 1) / b.y, 1);
}

}  // namespace

Timer::Timer(licom_ctx *c, const char *n) : ctx(c), name(n) {
    if (ctx->profiling) {
        (void)hipEventCreate(&a);
        (void)hipEventCreate(&b);
        (void)hipEventRecord(a, ctx->stream);
    }
}

Timer::~Timer() {
    if (ctx->profiling) {
        (void)hipEventRecord(b, ctx->stream);
        ctx->pending.push_back({name, a, b});
    }
}

int launch_surface(licom_ctx *ctx, const licom_fields *f, const licom_diag *d, const Scratch &s) {
    const KGrid &g = ctx->g;
    const KPar &p = ctx->kp;
    hipStream_t st = ctx->stream;
    const dim3 b2(64, 4);
    {
        Timer t(ctx, "k_surface");
        SurfaceIO o{f->h0f, f->h0l, f->atb, f->swv, f->nswv, f->sst, f->seaice, f->sss, f->fresh,
                    s.stf_pre, s.work_u, s.stf_n, s.temp11, d ? d->net : nullptr, s.mix_count};
        k_surface<<<grid2(g.imt, g.jmt, b2), b2, 0, st>>>(g, p, o);
    }
    {
        Timer t(ctx, "k_sss");
        k_sss_partial<<<s.nblk, kRedThreads, 0, st>>>(g, s.temp11, s.partials);
        k_sss_final<<<1, kRedThreads, 0, st>>>(s.partials, s.nblk, ctx->area_t, p.sssnorm,
                                               s.err_norm2, d ? d->fw_norm2 : nullptr);
    }
    return hip_status(hipGetLastError(), "surface launch");
}

int launch_halo(licom_ctx *ctx, const licom_fields *f, int flags) {
    const KGrid &g = ctx->g;
    hipStream_t st = ctx->stream;
    {
        Timer t(ctx, "k_halo");
        const int cells = 4 * (g.jmt - 4) + 4 * g.imt;
        k_halo<<<dim3((cells + 255) / 256, 4 * g.km), 256, 0, st>>>(g.imt, g.jmt, 4 * g.km, f->at,
                                                                     f->atb, g.km);
        // licomqice of the fused icesnow: ghost columns copy their interior
        // columns, closed rows hold 0, as icesnow_cu gives on the refilled
        // ghost cells of at and kmt
        if ((flags & LICOM_STEP_ICESNOW) && ctx->qice)
            k_halo<<<dim3((cells + 255) / 256, 1), 256, 0, st>>>(g.imt, g.jmt, 1, ctx->qice, nullptr, g.km);
    }
    if (flags & LICOM_STEP_VTL_SCRATCH) {
        Timer t(ctx, "k_copy_vtl");
        k_copy_plane<<<1024, 256, 0, st>>>(f->at + ctx->n3, f->vtl, ctx->n3);
    }
    return hip_status(hipGetLastError(), "halo launch");
}

int launch_step_simple(licom_ctx *ctx, const licom_fields *f, const licom_diag *d, int flags) {
    const KGrid &g = ctx->g;
    const KPar &p = ctx->kp;
    Scratch &s = ctx->s;
    hipStream_t st = ctx->stream;
    const dim3 b2(64, 4), b3(64, 4, 1);
    const dim3 g2 = grid2(g.imt, g.jmt, b2);
    const dim3 g3(g2.x, g2.y, g.km);
    const dim3 g3i((g.imt - 4 + 63) / 64, (g.jmt - 4 + 3) / 4, g.km);
    if (int rc = launch_surface(ctx, f, d, s)) return rc;
    {
        Timer t(ctx, "k_faces");
        k_faces<<<g3, b3, 0, st>>>(g, p, f->utf, f->utl, f->vtf, f->vtl, s.uw, s.vs);
    }
    {
        Timer t(ctx, "k_ws");
        k_ws<<<g2, b2, 0, st>>>(g, p, f->utf, f->utl, f->vtf, f->vtl, s.work_u, s.stf_pre, f->ws);
    }
    if (p.adv == LICOM_ADV_TSPAS) {
        Timer t(ctx, "k_predict");
        k_predict<<<g3, b3, 0, st>>>(g, p, f->at, s.uw, s.vs, f->ws, s.flags);
    }
    if (p.hmix == LICOM_HMIX_DEL4) {
        Timer t(ctx, "k_d2tk");
        k_d2tk<<<g3, b3, 0, st>>>(g, f->atb, s.d2tk);
    }
    {
        Timer t(ctx, "k_correct");
        k_correct<<<g3i, b3, 0, st>>>(g, p, f->at, f->atb, f->akt, s.uw, s.vs, f->ws, s.flags,
                                      s.d2tk, s.stf_n, s.err_norm2, f->swv, f->restore, s.ori,
                                      d ? d->ax : nullptr, d ? d->ay : nullptr,
                                      d ? d->az : nullptr, d ? d->dz : nullptr,
                                      d ? d->penetrate : nullptr);
    }
    // with icesnow, the reference's order tracer_ -> icesnow_ -> convadj_
    // (stepon.cpp:69-75): the column kernel without convadj, the halo, then
    // icesnow and convadj over every column
    const bool ice = flags & LICOM_STEP_ICESNOW;
    {
        Timer t(ctx, "k_column");
        k_column<<<g2, b2, 0, st>>>(g, p, ice ? flags & ~LICOM_STEP_CONVADJ : flags, f->at, f->atb, f->akt,
                                    s.ori, s.stf_n, f->utf, f->vtf, f->h0f, d ? d->tend : nullptr,
                                    d ? d->dt_diff : nullptr, d ? d->dt_conv : nullptr, s.mix_count);
    }
    if (int rc = launch_halo(ctx, f, flags & ~LICOM_STEP_ICESNOW)) return rc;
    if (!ice) return LICOM_OK;
    if (int rc = launch_icesnow(ctx, f->at, ctx->qice, 1)) return rc;
    return (flags & LICOM_STEP_CONVADJ) ? launch_convadj(ctx, f, d) : LICOM_OK;
}

int launch_stream_copy(void *dst, const void *src, size_t bytes, hipStream_t st) {
    const size_t n = bytes / 16;
    if (n == 0) return LICOM_OK;
    if ((n + 255) / 256 > 0x7fffffffull) {
        set_error("licom_stream_copy: %zu bytes exceed one launch", bytes);
        return LICOM_EINVAL;
    }
    k_stream_copy<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(static_cast<const double2 *>(src),
                                                               static_cast<double2 *>(dst), n);
    return hip_status(hipGetLastError(), "stream copy launch");
}

int launch_icesnow(licom_ctx *ctx, double *at, double *qice, int reset) {
    const dim3 b2(64, 4);
    Timer t(ctx, "k_icesnow");
    k_icesnow<<<grid2(ctx->imt, ctx->jmt, b2), b2, 0, ctx->stream>>>(ctx->g, at, qice, reset);
    return hip_status(hipGetLastError(), "icesnow launch");
}

int launch_convadj(licom_ctx *ctx, const licom_fields *f, const licom_diag *d) {
    const dim3 b2(64, 4);
    Timer t(ctx, "k_convadj");
    k_convadj<<<grid2(ctx->imt, ctx->jmt, b2), b2, 0, ctx->stream>>>(
        ctx->g, ctx->kp, f->at, f->atb, d ? d->tend : nullptr, d ? d->dt_conv : nullptr);
    return hip_status(hipGetLastError(), "convadj launch");
}

}  // namespace licom
