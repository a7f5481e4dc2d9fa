// barotr.hip -- the barotropic subcycle barotr_ on gfx950
// (include/licom_barotr.h).
//
// The reference runs up to eight 2-D launches per subcycle and passes
// transports, gradients and Laplacians between them through the planes of
// its 3-D work array.  Here each stage is one lane per cell (stage 1 in two
// launches, three with del4: the height update first, as the kernel after it
// rewrites the work array the update reads): the transports of the four U cells around a T
// point, the surface height at the four T points around a U point, the
// divergence and vorticity of the del4 factor and the T-grid Laplacians are
// formed where they are used, and only what the next stage or the caller's
// halo exchange reads is stored (work, wka planes 2 and 3, the velocities
// and heights).  Operation order follows the reference statements
// (-ffp-contract=off), so results are identical bit for bit.
#include <hip/hip_runtime.h>

#include <new>
#include <vector>

#include "licom_barotr.h"
#include "tracer_internal.h"

namespace licom {
namespace {

constexpr double kGrav = 9.806e0;   // constant_mod.h's g

struct TGrid {
    int imt, jmt, ib, ie, jb, je, del4;
    size_t n2;
    double dtb, am, ah;
    const int *kmt, *kmu, *kmtn, *kmts, *kmte, *kmtw;
    const double *au0, *aus, *auw, *ausw, *htw, *hts, *tarea_r, *uarea_r, *dxu, *dyu, *dxur, *dyur, *fcor,
        *ebea, *ebeb, *amf, *ahf, *dtn, *dts, *dte, *dtw, *duc, *dum, *dun, *dus, *due, *duw, *dmc, *dmn,
        *dms, *dme, *dmw;
    double *d2u, *d2v;   // del4 scratch [j][i]
};

__device__ __forceinline__ bool interior2(const TGrid &g, int j, int i) {
    return i > 1 && i < g.imt - 2 && j > 1 && j < g.jmt - 2;
}

// the U-point average of a T-point height (barotr1_cu :40-49, barotr2_cu
// :207-215); h(jj, ii) is a functor
template <class H>
__device__ __forceinline__ double h_at_u(const TGrid &g, int j, int i, H h) {
    if (j == g.jmt - 1 || i == 0) return 0.0e0;
    const size_t q = (size_t)j * g.imt + i;
    return g.au0[q] * h(j, i) + g.aus[q] * h(j + 1, i) + g.auw[q] * h(j, i - 1) + g.ausw[q] * h(j + 1, i - 1);
}

// the divergence of the transports (t0, t1 at the U cells (j, i+1),
// (j-1, i+1), (j, i), (j-1, i)) at the T point (j, i), before tarea_r
__device__ __forceinline__ double t_div(const TGrid &g, size_t q, double a0, double b0, double c0, double d0,
                                        double a1, double b1, double c1, double d1) {
    const int imt = g.imt;
    return 0.5 * ((a0 + b0) * g.htw[q + 1] - (c0 + d0) * g.htw[q] + (a1 + c1) * g.hts[q] - (b1 + d1) * g.hts[q - imt]);
}

struct Lap {
    double c, n, s, e, w, mc, mn, ms, me, mw;
};
__device__ __forceinline__ Lap load_lap(const TGrid &g, size_t q) {
    Lap L;
    L.c = g.duc[q] + g.dum[q];
    L.n = g.dun[q];
    L.s = g.dus[q];
    L.e = g.due[q];
    L.w = g.duw[q];
    L.mc = g.dmc[q];
    L.mn = g.dmn[q];
    L.ms = g.dms[q];
    L.me = g.dme[q];
    L.mw = g.dmw[q];
    return L;
}
__device__ __forceinline__ double lap_du(const Lap &L, const double *a, size_t q, int imt) {
    return L.c * a[q] + L.n * a[q - imt] + L.s * a[q + imt] + L.e * a[q + 1] + L.w * a[q - 1];
}
__device__ __forceinline__ double lap_dm(const Lap &L, const double *b, size_t q, int imt) {
    return L.mc * b[q] + L.mn * b[q - imt] + L.ms * b[q + imt] + L.me * b[q + 1] + L.mw * b[q - 1];
}

// the masked 5-point operator of barotr3_cu1 / barotr3_tdel4_1 on x (a
// functor of the cell), with its diagonal cc
template <class X>
__device__ __forceinline__ double lap_t(const TGrid &g, int j, int i, X x, double &cc) {
    const size_t q = (size_t)j * g.imt + i;
    double hdtk = 0;
    cc = 0;
    if (0 <= g.kmt[q] - 1) {
        double c;
        if (0 <= g.kmtn[q] - 1) {
            c = g.dtn[q];
            hdtk += c * x(j - 1, i);
            cc -= c;
        }
        if (0 <= g.kmts[q] - 1) {
            c = g.dts[q];
            hdtk += c * x(j + 1, i);
            cc -= c;
        }
        if (0 <= g.kmte[q] - 1) {
            c = g.dte[q];
            hdtk += c * x(j, i + 1);
            cc -= c;
        }
        if (0 <= g.kmtw[q] - 1) {
            c = g.dtw[q];
            hdtk += c * x(j, i - 1);
            cc -= c;
        }
    }
    return hdtk;
}

// stage 0: the clears, barotr1_cu and barotr1_cu1 (barotr.cpp:32-83)
__global__ __launch_bounds__(256) void k_bt0(TGrid g, licom_barotr_fields f) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
    const int imt = g.imt, jmt = g.jmt;
    if (i >= imt || j >= jmt) return;
    const size_t n2 = g.n2, q = (size_t)j * imt + i;
    auto h0 = [&](int jj, int ii) { return f.h0[(size_t)jj * imt + ii]; };
    double w;
    if (interior2(g, j, i)) {
        w = 0;
        if (0 <= g.kmt[q] - 1) {
            // transports of the U cells (j, i+1), (j-1, i+1), (j, i), (j-1, i)
            auto tr = [&](int jj, int ii, double &t0, double &t1) {
                const size_t p = (size_t)jj * imt + ii;
                const double hw = h_at_u(g, jj, ii, h0);
                t0 = f.ub[p] * (f.dzph[p] + hw);
                t1 = f.vb[p] * (f.dzph[p] + hw);
            };
            double a0, a1, b0, b1, c0, c1, d0, d1;
            tr(j, i + 1, a0, a1);
            tr(j - 1, i + 1, b0, b1);
            tr(j, i, c0, c1);
            tr(j - 1, i, d0, d1);
            w = t_div(g, q, a0, b0, c0, d0, a1, b1, c1, d1) * g.tarea_r[q] * f.vit[q] * -0.250;
        }
    } else {
        w = h_at_u(g, j, i, h0);
    }
    f.work[q] = w;
    f.wka[2 * n2 + q] = 0.0;
    f.wka[3 * n2 + q] = 0.0;
}

// bhdel4_div, bhdel4_grad, bhdel4_zcurl, bhdel4_grad and barotr2_udel4_0
// (barotr.cpp:272-407): the first Laplacian of (ubp, vbp) scaled by the
// viscosity factor
__global__ __launch_bounds__(256) void k_bt1_lap1(TGrid g, licom_barotr_fields f) {
    // divergence and vorticity once per T point of the block (rows j0..j0+4,
    // columns i0-1..i0+63, flat offsets from the block's corner; each T point
    // serves four U points), shared through LDS
    constexpr int TX = 65, TN = TX * 5;
    __shared__ double sDV[TN], sCZ[TN];
    const int tx = threadIdx.x, ty = threadIdx.y, tl = tx + 64 * ty;
    const int i0 = blockIdx.x * 64, j0 = blockIdx.y * 4;
    const int imt = g.imt, jmt = g.jmt;
    const double *u = f.ubp, *v = f.vbp;
    for (int e = tl; e < TN; e += 256) {
        const int tr = e / TX, ii = i0 - 1 + (e - tr * TX), jj = j0 + tr;
        double dv = 0.0e0, cz = 0.0e0;
        // a T point outside the array is used by no U point inside the
        // bounds (column -1 is the previous row's last cell, as the flat
        // offsets of the reference give it)
        if (jj >= 1 && jj < jmt && ii < imt) {
            const size_t p = (size_t)((long long)jj * imt + ii);
            const bool wet = 0 <= g.kmt[p] - 1;
            if (ii < imt - 1 && wet)
                dv = 0.5 *
                     ((u[p + 1] + u[p - imt + 1]) * g.htw[p + 1] - (u[p] + u[p - imt]) * g.htw[p] +
                      (v[p + 1] + v[p]) * g.hts[p] - (v[p - imt + 1] + v[p - imt]) * g.hts[p - imt]) *
                     g.tarea_r[p];
            if (ii >= 1 && wet)
                cz = 0.5 * ((v[p] * g.dyu[p] + v[p - imt] * g.dyu[p - imt] - v[p - 1] * g.dyu[p - 1] -
                             v[p - imt - 1] * g.dyu[p - imt - 1] - u[p] * g.dxu[p] - u[p - 1] * g.dxu[p - 1] +
                             u[p - imt] * g.dxu[p - imt] + u[p - imt - 1] * g.dxu[p - imt - 1]));
        }
        sDV[e] = dv;
        sCZ[e] = cz;
    }
    __syncthreads();
    const int i = i0 + tx, j = j0 + ty;
    if (i >= imt || j >= jmt) return;
    const size_t q = (size_t)j * imt + i;
    double am_f = 1.0e0, hu = 0, hv = 0;
    if (j >= g.jb - 2 && j <= g.je && i >= g.ib - 2 && i <= g.ie && g.kmu[q] - 1 >= 0) {
        // the T points (j, i), (j+1, i), (j, i-1), (j+1, i-1)
        const int ec = ty * TX + tx + 1, es = ec + TX, ew = ec - 1, esw = es - 1;
        const double gxf = g.dxur[q] * 0.5, gyf = g.dyur[q] * 0.5;
        double wc = sDV[ec], ws = sDV[es], ww = sDV[ew], wsw = sDV[esw];
        double gx = gxf * (ws - ww - wsw + wc), gy = gyf * (ws - ww + wsw - wc);
        double af = 0.0 + (gx * gx + gy * gy);
        wc = sCZ[ec];
        ws = sCZ[es];
        ww = sCZ[ew];
        wsw = sCZ[esw];
        gx = gxf * (ws - ww - wsw + wc);
        gy = gyf * (ws - ww + wsw - wc);
        af += gx * gx + gy * gy;
        const double t_area = sqrt(1.0e0 / g.uarea_r[q]);
        const double dxdy = t_area * t_area * t_area * t_area * t_area * 45.0e0;
        am_f = sqrt(af) * dxdy / fabs(g.am * g.amf[q]);
        am_f = fmin(40.0e0, am_f);
        am_f = fmax(1.0e0, am_f);
        const Lap L = load_lap(g, q);
        hu = lap_du(L, u, q, imt) + lap_dm(L, v, q, imt);
        hv = lap_du(L, v, q, imt) - lap_dm(L, u, q, imt);
    }
    g.d2u[q] = am_f * g.amf[q] * hu;
    g.d2v[q] = am_f * g.amf[q] * hv;
}

// stage 1: barotr2_cu1 (barotr.cpp:85-95)
__global__ __launch_bounds__(256) void k_bt1_h0(TGrid g, licom_barotr_fields f) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
    if (i >= g.imt || j >= g.jmt) return;
    const size_t q = (size_t)j * g.imt + i;
    f.h0[q] = f.h0p[q] + f.work[q] * g.dtb;
}

// stage 1: barotr2_cu (:98-253) or barotr2_udel4_1 (:410-488), on the new
// h0 (this kernel rewrites work, which k_bt1_h0 read)
template <bool DEL4>
__global__ __launch_bounds__(256) void k_bt1(TGrid g, licom_barotr_fields f) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
    const int imt = g.imt, jmt = g.jmt;
    if (i >= imt || j >= jmt) return;
    const size_t n2 = g.n2, q = (size_t)j * imt + i;
    // the new surface height (k_bt1_h0), at this and the neighbouring T points
    auto h0 = [&](int jj, int ii) { return f.h0[(size_t)jj * imt + ii]; };
    const double h0c = h0(j, i);
    double hu = 0, hv = 0;
    if (g.kmu[q] - 1 >= 0 && j >= g.jb - 1 && j < g.je && i >= g.ib - 1 && i < g.ie) {
        const Lap L = load_lap(g, q);
        if (DEL4) {
            hu = g.am * (lap_du(L, g.d2u, q, imt) + lap_dm(L, g.d2v, q, imt));
            hv = g.am * (lap_du(L, g.d2v, q, imt) - lap_dm(L, g.d2u, q, imt));
        } else {
            hu = g.am * (lap_du(L, f.ubp, q, imt) + lap_dm(L, f.vbp, q, imt)) * f.viv[q];
            hv = g.am * (lap_du(L, f.vbp, q, imt) - lap_dm(L, f.ubp, q, imt)) * f.viv[q];
        }
    }
    const bool gr = DEL4 ? interior2(g, j, i) : (i > 0 && i < imt - 1 && j > 1 && j < jmt);
    double t0 = 0, t1 = 0;
    if (gr) {
        double gx = 0.0e0, gy = 0.0e0;
        // the reference reads one row past the block at j = jmt-1 (del2);
        // those cells of planes 0, 1 are overwritten before any read
        if (g.kmu[q] - 1 >= 0 && j < jmt - 1) {
            const double hs = h0(j + 1, i), hw = h0(j, i - 1), hsw = h0(j + 1, i - 1);
            gx = g.dxur[q] * 0.5 * (hs - hw - hsw + h0c);
            gy = g.dyur[q] * 0.5 * (hs - hw + hsw - h0c);
        }
        const double gstar = (f.wgp[q] - 1.0) * kGrav;
        t0 = hu + gstar * gx;
        t1 = hv + gstar * gy;
    }
    const double work = h_at_u(g, j, i, h0);
    double w2 = 0.0, w3 = 0.0;
    if (interior2(g, j, i)) {
        const double vv = f.viv[q], fc = g.fcor[q];
        t0 = vv * (t0 + f.dlub[q] - fc * f.vbp[q] + f.pax[q] + f.pxb[q] - work * f.whx[q]);
        t1 = vv * (t1 + f.dlvb[q] + fc * f.ubp[q] + f.pay[q] + f.pyb[q] - work * f.why[q]);
        w2 = g.ebea[q] * t0 - g.ebeb[q] * t1;
        w3 = g.ebea[q] * t1 + g.ebeb[q] * t0;
    }
    f.work[q] = work;
    f.wka[2 * n2 + q] = w2;
    f.wka[3 * n2 + q] = w3;
}

// stage 2: barotr3_cu (:614-634), barotr3_tdel4_1 (:735-797),
// barotr3_cu1 (:636-733)
template <bool DEL4>
__global__ __launch_bounds__(256) void k_bt2(TGrid g, licom_barotr_fields f, int nc) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
    const int imt = g.imt, jmt = g.jmt;
    if (i >= imt || j >= jmt) return;
    const size_t n2 = g.n2, q = (size_t)j * imt + i;
    const double dtb = g.dtb;
    const double *w2 = f.wka + 2 * n2, *w3 = f.wka + 3 * n2;
    f.ub[q] = f.ubp[q] + w2[q] * dtb;
    f.vb[q] = f.vbp[q] + w3[q] * dtb;
    if (!(i < imt - 1 && j >= 1)) return;
    auto h0 = [&](int jj, int ii) { return f.h0[(size_t)jj * imt + ii]; };
    // barotr3_cu's transports at a U cell; planes 0, 1 as the earlier
    // kernels left them where barotr3_cu does not write (row 0: zero;
    // column imt-1: zero, or with del4 the vorticity bhdel4_zcurl left)
    auto tr = [&](int jj, int ii, double &t0, double &t1) {
        const size_t p = (size_t)jj * imt + ii;
        if (jj >= 1 && ii < imt - 1) {
            const double ub = f.ubp[p] + w2[p] * dtb, vb = f.vbp[p] + w3[p] * dtb;
            const double hw = h_at_u(g, jj, ii, h0);   // stage 1's work
            t0 = ub * (f.dzph[p] + hw);
            t1 = vb * (f.dzph[p] + hw);
            return;
        }
        t0 = 0.0;
        t1 = 0.0;
        if (DEL4 && jj >= 1 && ii >= 1 && 0 <= g.kmt[p] - 1) {
            const double *u = f.ubp, *v = f.vbp;
            t0 = 0.5 * ((v[p] * g.dyu[p] + v[p - imt] * g.dyu[p - imt] - v[p - 1] * g.dyu[p - 1] -
                         v[p - imt - 1] * g.dyu[p - imt - 1] - u[p] * g.dxu[p] - u[p - 1] * g.dxu[p - 1] +
                         u[p - imt] * g.dxu[p - imt] + u[p - imt - 1] * g.dxu[p - imt - 1]));
        }
    };
    double div_out = 0, hdtk = 0;
    if (0 <= g.kmt[q] - 1) {
        double a0, a1, b0, b1, c0, c1, d0, d1;
        tr(j, i + 1, a0, a1);
        tr(j - 1, i + 1, b0, b1);
        tr(j, i, c0, c1);
        tr(j - 1, i, d0, d1);
        div_out = t_div(g, q, a0, b0, c0, d0, a1, b1, c1, d1) * g.tarea_r[q];
    }
    if ((nc + 1) % 4 == 0) {
        if (j >= g.jb - 1 && j < g.je && i >= g.ib - 1 && i < g.ie) {
            auto h0p = [&](int jj, int ii) { return f.h0p[(size_t)jj * imt + ii]; };
            double cc;
            if (DEL4) {
                // barotr3_tdel4_1's d2tk at a cell of its region
                auto d2t = [&](int jj, int ii) {
                    if (!(ii < imt - 1 && jj >= 1)) return 0.0;
                    if (!(jj >= g.jb - 2 && jj <= g.je && ii >= g.ib - 2 && ii <= g.ie)) return 0.0;
                    double c2;
                    double h = lap_t(g, jj, ii, h0p, c2);
                    return g.ahf[(size_t)jj * imt + ii] * (c2 * h0p(jj, ii) + h);
                };
                hdtk = lap_t(g, j, i, d2t, cc);
                hdtk = g.ah * (cc * d2t(j, i) + hdtk);
            } else {
                hdtk = lap_t(g, j, i, h0p, cc);
                hdtk = g.ah * (cc * h0p(j, i) + hdtk);
            }
        } else {
            hdtk = 0;
        }
    }
    f.work[q] = (hdtk * 1.0e0 - div_out) * f.vit[q];
}

// stage 3: barotr4_cu (:858-882)
__global__ __launch_bounds__(256) void k_bt3(TGrid g, licom_barotr_fields f) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
    if (i >= g.imt || j >= g.jmt) return;
    const size_t q = (size_t)j * g.imt + i;
    const double h0 = f.h0p[q] + (f.work[q] * g.dtb);
    f.h0p[q] = h0;
    f.ubp[q] = f.ub[q];
    f.vbp[q] = f.vb[q];
    f.h0f[q] += h0;
    f.h0bf[q] += h0;
    f.h0[q] = h0;
}

}  // namespace
}  // namespace licom

using namespace licom;

struct licom_barotr {
    TGrid g{};
    hipStream_t stream = nullptr;
    void *mem = nullptr, *scratch = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool ran = false;
};

extern "C" {

int licom_barotr_create(const licom_barotr_grid *grid, licom_barotr **out) {
    if (!grid || !out) {
        set_error("licom_barotr_create: NULL argument");
        return LICOM_EINVAL;
    }
    *out = nullptr;
    const licom_barotr_grid &G = *grid;
    if (G.imt < 5 || G.jmt < 5) {
        set_error("licom_barotr_create: invalid extents imt=%d jmt=%d (need >= 5)", G.imt, G.jmt);
        return LICOM_EINVAL;
    }
    const int lo = G.del4 ? 3 : 2, hij = G.del4 ? G.jmt - 2 : G.jmt - 1, hii = G.del4 ? G.imt - 2 : G.imt - 1;
    if (G.jb < lo || G.je > hij || G.ib < lo || G.ie > hii) {
        set_error("licom_barotr_create: block bounds ib=%d ie=%d jb=%d je=%d outside %d..%d x %d..%d (%s)", G.ib,
                  G.ie, G.jb, G.je, lo, hii, lo, hij, G.del4 ? "del4" : "del2");
        return LICOM_EINVAL;
    }
    const double *src[] = {G.au0, G.aus, G.auw, G.ausw, G.htw, G.hts, G.tarea_r, G.uarea_r, G.dxu, G.dyu, G.dxur,
                           G.dyur, G.fcor, G.ebea, G.ebeb, G.amf, G.ahf, G.dtn, G.dts, G.dte, G.dtw, G.duc, G.dum,
                           G.dun, G.dus, G.due, G.duw, G.dmc, G.dmn, G.dms, G.dme, G.dmw};
    const int *isrc[] = {G.kmt, G.kmu, G.kmtn, G.kmts, G.kmte, G.kmtw};
    constexpr int nd = sizeof(src) / sizeof(src[0]), ni = sizeof(isrc) / sizeof(isrc[0]);
    for (const double *p : src)
        if (!p) {
            set_error("licom_barotr_create: a grid array is NULL");
            return LICOM_EINVAL;
        }
    for (const int *p : isrc)
        if (!p) {
            set_error("licom_barotr_create: a grid array is NULL");
            return LICOM_EINVAL;
        }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        set_error("no HIP device available");
        return LICOM_ENODEV;
    }
    licom_barotr *b = new (std::nothrow) licom_barotr();
    if (!b) return LICOM_ENOMEM;
    const size_t n2 = (size_t)G.imt * G.jmt;
    const size_t bytes = nd * n2 * sizeof(double) + ni * n2 * sizeof(int);
    std::vector<double> h(nd * n2);
    for (int a = 0; a < nd; ++a) std::copy(src[a], src[a] + n2, h.begin() + a * n2);
    int rc = hip_status(hipMalloc(&b->mem, bytes), "hipMalloc");
    char *base = static_cast<char *>(b->mem);
    if (!rc) rc = hip_status(hipMemcpy(base, h.data(), nd * n2 * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
    int *ik = reinterpret_cast<int *>(base + nd * n2 * sizeof(double));
    for (int a = 0; a < ni && !rc; ++a)
        rc = hip_status(hipMemcpy(ik + a * n2, isrc[a], n2 * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy");
    if (!rc && G.del4) rc = hip_status(hipMalloc(&b->scratch, 2 * n2 * sizeof(double)), "hipMalloc");
    if (!rc) rc = hip_status(hipEventCreate(&b->ev0), "hipEventCreate");
    if (!rc) rc = hip_status(hipEventCreate(&b->ev1), "hipEventCreate");
    if (rc) {
        licom_barotr_destroy(b);
        return rc;
    }
    const double *d = reinterpret_cast<const double *>(base);
    TGrid &g = b->g;
    g.imt = G.imt;
    g.jmt = G.jmt;
    g.ib = G.ib;
    g.ie = G.ie;
    g.jb = G.jb;
    g.je = G.je;
    g.del4 = G.del4 ? 1 : 0;
    g.n2 = n2;
    g.dtb = G.dtb;
    g.am = G.am;
    g.ah = G.ah;
    const double **dst[] = {&g.au0, &g.aus, &g.auw, &g.ausw, &g.htw, &g.hts, &g.tarea_r, &g.uarea_r,
                            &g.dxu, &g.dyu, &g.dxur, &g.dyur, &g.fcor, &g.ebea, &g.ebeb, &g.amf,
                            &g.ahf, &g.dtn, &g.dts, &g.dte, &g.dtw, &g.duc, &g.dum, &g.dun,
                            &g.dus, &g.due, &g.duw, &g.dmc, &g.dmn, &g.dms, &g.dme, &g.dmw};
    for (int a = 0; a < nd; ++a) *dst[a] = d + a * n2;
    const int **idst[] = {&g.kmt, &g.kmu, &g.kmtn, &g.kmts, &g.kmte, &g.kmtw};
    for (int a = 0; a < ni; ++a) *idst[a] = ik + a * n2;
    double *sc = static_cast<double *>(b->scratch);
    g.d2u = sc;
    g.d2v = sc ? sc + n2 : nullptr;
    *out = b;
    return LICOM_OK;
}

int licom_barotr_set_stream(licom_barotr *b, void *hip_stream) {
    if (!b) {
        set_error("licom_barotr_set_stream: NULL context");
        return LICOM_EINVAL;
    }
    b->stream = (hipStream_t)hip_stream;
    return LICOM_OK;
}

int licom_barotr_stage(licom_barotr *b, const licom_barotr_fields *f, int nc, int stage) {
    if (!b || !f) {
        set_error("licom_barotr_stage: NULL argument");
        return LICOM_EINVAL;
    }
    if (stage < 0 || stage > 3) {
        set_error("licom_barotr_stage: no stage %d (0..3)", stage);
        return LICOM_EINVAL;
    }
    const void *need[] = {f->vit, f->viv, f->dzph, f->dlub, f->dlvb, f->wgp, f->pax, f->pay, f->pxb,
                          f->pyb, f->whx, f->why, f->h0,   f->h0p,  f->ub,  f->vb,  f->ubp, f->vbp,
                          f->h0f, f->h0bf, f->work, f->wka};
    for (const void *p : need)
        if (!p) {
            set_error("licom_barotr_stage: a field pointer is NULL");
            return LICOM_EINVAL;
        }
    const TGrid &g = b->g;
    const dim3 blk(64, 4), grd((g.imt + 63) / 64, (g.jmt + 3) / 4);
    if (int rc = hip_status(hipEventRecord(b->ev0, b->stream), "hipEventRecord")) return rc;
    switch (stage) {
    case 0:
        k_bt0<<<grd, blk, 0, b->stream>>>(g, *f);
        break;
    case 1:
        k_bt1_h0<<<grd, blk, 0, b->stream>>>(g, *f);
        if (int rc = hip_status(hipGetLastError(), "k_bt1_h0 launch")) return rc;
        if (g.del4) {
            k_bt1_lap1<<<grd, blk, 0, b->stream>>>(g, *f);
            if (int rc = hip_status(hipGetLastError(), "k_bt1_lap1 launch")) return rc;
            k_bt1<true><<<grd, blk, 0, b->stream>>>(g, *f);
        } else {
            k_bt1<false><<<grd, blk, 0, b->stream>>>(g, *f);
        }
        break;
    case 2:
        if (g.del4)
            k_bt2<true><<<grd, blk, 0, b->stream>>>(g, *f, nc);
        else
            k_bt2<false><<<grd, blk, 0, b->stream>>>(g, *f, nc);
        break;
    default:
        k_bt3<<<grd, blk, 0, b->stream>>>(g, *f);
    }
    if (int rc = hip_status(hipGetLastError(), "barotr stage launch")) return rc;
    if (int rc = hip_status(hipEventRecord(b->ev1, b->stream), "hipEventRecord")) return rc;
    b->ran = true;
    return LICOM_OK;
}

int licom_barotr_sync(licom_barotr *b) {
    if (!b) {
        set_error("licom_barotr_sync: NULL context");
        return LICOM_EINVAL;
    }
    return hip_status(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
}

int licom_barotr_last_ms(licom_barotr *b, double *ms) {
    if (!b || !ms) {
        set_error("licom_barotr_last_ms: NULL argument");
        return LICOM_EINVAL;
    }
    *ms = 0.0;
    if (!b->ran) return LICOM_OK;
    if (int rc = hip_status(hipEventSynchronize(b->ev1), "hipEventSynchronize")) return rc;
    float v = 0.0f;
    if (int rc = hip_status(hipEventElapsedTime(&v, b->ev0, b->ev1), "hipEventElapsedTime")) return rc;
    *ms = v;
    return LICOM_OK;
}

void licom_barotr_destroy(licom_barotr *b) {
    if (!b) return;
    if (b->ran) (void)hipStreamSynchronize(b->stream);
    if (b->mem) (void)hipFree(b->mem);
    if (b->scratch) (void)hipFree(b->scratch);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    delete b;
}

}  // extern "C"
