// tracer_device.h -- device arithmetic shared by every kernel of the tracer
// step.  Each expression keeps the operation order of the reference statement
// it cites (the library is compiled with -ffp-contract=off), so that all
// kernel variants produce identical bits.
#pragma once
#include <hip/hip_runtime.h>

#include "licom_tracer.h"

namespace licom {
namespace dev {

// wkd / wkb of tracer_1 (tracer.cpp:1107-1119): half of the (rescaled)
// leading velocity plus half of the trailing one.
__device__ __forceinline__ double wk(double uf, double ul, double oncc) {
    return 0.5 * (uf * oncc) + 0.5 * ul;
}

// Terms of the tspas predictor at one cell (advection_tracer_2,
// tracer.cpp:311-414).  T* are tracer values (c centre, e/w = i+-1,
// s/n = j+-1, u/d = k-+1), Ue/Uw the west-face velocities at i+1 and i,
// Vs/Vn the south-face velocities at j and j-1, Wk/Wb the vertical velocity at
// the top and bottom of the cell, D1..D4 the face metric products
// htw*hun (i+1), htw*hun (i), hts*hue (j), hts*hue (j-1).
struct PTerms {
    double x0, y0, xy1, xy2, xy3, xy4, c1, c2, za, zb1, zb2, zc;
};

// Cheap terms (no division) of the predictor.
__device__ __forceinline__ void pterms_cheap(double Tc, double Te, double Tw, double Ts,
                                             double Tn, double Tu, double Td, double Ue,
                                             double Uw, double Vs, double Vn, double Wk,
                                             double Wb, double tr, int k, int km,
                                             const double *odzp, PTerms &t) {
    t.x0 = (Te + Tc) * Ue * tr - (Tc + Tw) * Uw * tr;
    t.y0 = ((Ts + Tc) * Vs - (Tc + Tn) * Vn) * tr;
    t.c1 = -Tc * (Ue - Uw) * tr * 2.0;
    t.c2 = -Tc * (Vs - Vn) * tr * 2.0;
    if (k == 0) {
        t.za = -0.5 * odzp[0] * Wb * (Td + Tc);
        t.zc = odzp[0] * Tc * Wb;
    } else if (k == km - 1) {
        t.za = 0.5 * odzp[k] * Wk * (Tc + Tu);
        t.zc = -odzp[k] * Tc * Wk;
    } else {
        t.za = 0.5 * odzp[k] * Wk * (Tc + Tu) - 0.5 * odzp[k] * Wb * (Tc + Td);
        t.zc = -odzp[k] * Tc * (Wk - Wb);
    }
}

// Lax-Wendroff terms (the limiter may replace them by upwind ones).
__device__ __forceinline__ void pterms_lw(double Tc, double Te, double Tw, double Ts, double Tn,
                                          double Tu, double Td, double Ue, double Uw, double Vs,
                                          double Vn, double Wk, double Wb, double tr, double D1,
                                          double D2, double D3, double D4, int k, int km,
                                          const double *odzp, const double *odzt, double dts,
                                          PTerms &t) {
    t.xy1 = -dts * (Te - Tc) * 2.0 * tr * Ue * Ue / D1;
    t.xy2 = dts * (Tc - Tw) * 2.0 * tr * Uw * Uw / D2;
    t.xy3 = -dts * (Ts - Tc) * 2.0 * tr * Vs * Vs / D3;
    t.xy4 = dts * (Tc - Tn) * 2.0 * tr * Vn * Vn / D4;
    if (k == 0) {
        t.zb1 = 0.0;
        t.zb2 = 0.5 * odzp[0] * Wb * Wb * odzt[1] * (Tc - Td) * dts;
    } else if (k == km - 1) {
        t.zb1 = -0.5 * odzp[k] * Wk * Wk * odzt[k] * (Tu - Tc) * dts;
        t.zb2 = 0.0;
    } else {
        t.zb1 = -0.5 * odzp[k] * Wk * Wk * odzt[k] * (Tu - Tc) * dts;
        t.zb2 = 0.5 * odzp[k] * Wb * Wb * odzt[k + 1] * (Tc - Td) * dts;
    }
}

__device__ __forceinline__ void pterms(double Tc, double Te, double Tw, double Ts, double Tn,
                                       double Tu, double Td, double Ue, double Uw, double Vs,
                                       double Vn, double Wk, double Wb, double tr, double D1,
                                       double D2, double D3, double D4, int k, int km,
                                       const double *odzp, const double *odzt, double dts,
                                       PTerms &t) {
    pterms_cheap(Tc, Te, Tw, Ts, Tn, Tu, Td, Ue, Uw, Vs, Vn, Wk, Wb, tr, k, km, odzp, t);
    pterms_lw(Tc, Te, Tw, Ts, Tn, Tu, Td, Ue, Uw, Vs, Vn, Wk, Wb, tr, D1, D2, D3, D4, k, km, odzp,
              odzt, dts, t);
}

// Predictor value at00 (tracer.cpp:410-414).
__device__ __forceinline__ double predictor(const PTerms &t, double Tc, double dts) {
    const double xx = -(t.x0 + t.xy1 + t.xy2 + t.c1);
    const double yy = -(t.y0 + t.xy3 + t.xy4 + t.c2);
    const double zz = -(t.zb1 + t.zb2 + t.za + t.zc);
    return Tc + (xx + yy + zz) * dts;
}

// advection_tracer_1 (tracer.cpp:120-289) at one cell: the centered
// (Str=0) or flux (Str=1) advective tendency, from the same neighbours as
// the tspas terms (Wk/Wb: w at the top and bottom of the cell).  For the
// centered form hx, hy, vz are the parts the reference adds to ax, ay and
// subtracts from az, each before its division by nss (tracer.cpp:140-159,
// 174-205, 224-255); level 0 and km-1 take one vertical term, as in the
// level-0 / level-(km-1) statements of tracer.cpp:152-159 and 245-255.
struct CFAdv {
    double adv, hx, hy, vz;
};
__device__ __forceinline__ CFAdv adv_cf(int scheme, double Tc, double Te, double Tw, double Ts,
                                        double Tn, double Tu, double Td, double Ue, double Uw,
                                        double Vs, double Vn, double Wk, double Wb, double tr, int k,
                                        int km, const double *odzp) {
    CFAdv r;
    if (scheme == LICOM_ADV_CENTERED) {
        const double a = (-Uw * (Tc - Tw) - Ue * (Te - Tc) - Vs * (Ts - Tc) - Vn * (Tc - Tn)) * tr;
        r.hx = (-Uw * (Tc - Tw) - Ue * (Te - Tc)) * tr;
        r.hy = (-Vs * (Ts - Tc) - Vn * (Tc - Tn)) * tr;
        double z;
        if (k == 0) {
            z = Wb * (Tc - Td);
        } else if (k == km - 1) {
            z = Wk * (Tu - Tc);
        } else {
            const double z1 = Wk * (Tu - Tc);
            const double z2 = Wb * (Tc - Td);
            z = z1 + z2;
        }
        r.vz = 0.5 * odzp[k] * z;
        r.adv = a - 0.5 * odzp[k] * z;
    } else {
        const double a = (-Uw * (Tc + Tw) + Ue * (Te + Tc) - Vn * (Tc + Tn) + Vs * (Ts + Tc)) * tr;
        const double z1 = (k == 0) ? 0.0 : Wk * (Tu + Tc) * 0.5;
        const double z2 = (k == km - 1) ? 0.0 : Wb * (Tc + Td) * 0.5;
        r.adv = a - odzp[k] * (z2 - z1);
        r.hx = r.hy = r.vz = 0.0;
    }
    return r;
}

// First-order upwind replacements of the limiter (tracer.cpp:574-658).
__device__ __forceinline__ double up_xy1(double Te, double Tc, double Ue, double tr) {
    return -(Te - Tc) * fabs(Ue) * tr;
}
__device__ __forceinline__ double up_xy2(double Tc, double Tw, double Uw, double tr) {
    return (Tc - Tw) * fabs(Uw) * tr;
}
__device__ __forceinline__ double up_zb1(double Tu, double Tc, double Wk, double odzpk) {
    return -0.5 * fabs(Wk) * odzpk * (Tu - Tc);
}
__device__ __forceinline__ double up_zb2(double Tc, double Td, double Wb, double odzpk) {
    return 0.5 * fabs(Wb) * odzpk * (Tc - Td);
}

// icesnow (icesnow.cpp:15-68): freezing point TBICE and CP
// (constant_mod.F90:168,164), the latent heat of fusion (shr_const_mod.h:51)
// and the reference salinities of sea water and sea ice (icesnow.cpp:27-28).
constexpr double kTbice = -1.8, kCpIce = 3996.0, kLatIce = 3.337e5, kSalOcn = 34.7, kSalIce = 4.0;
// Surface S change (model units) for tdiff K of freezing, in the reference's
// operation order (icesnow.cpp:39-40, 55-57).
__device__ __forceinline__ double ice_salt(double tdiff) {
    return tdiff * (kSalOcn - kSalIce) * kCpIce / kLatIce * 0.001;
}

// AKT surface floor: biharmonic build floors T only (tracer.cpp:1315), the
// Laplacian build every tracer (:1187); only interface 0 is affected.
__device__ __forceinline__ double floor_akt(double a, int k, int n, int hmix, double dwndmix) {
    if (k == 0 && (n == 0 || hmix == LICOM_HMIX_DEL2) && a < dwndmix) return dwndmix;
    return a;
}

// Row k of invtrit's tridiagonal system (tracer.cpp:936-963) for a column
// with kz wet levels; wkm1 = wkc(k-1), wkk = wkc(k) (after the AKT floor).
__device__ __forceinline__ void tri_row(int k, int kz, double wkm1, double wkk,
                                        const double *odzp, const double *odzt, double dts,
                                        double &a, double &b, double &c) {
    if (k == kz - 1) {
        a = (k >= 1) ? wkm1 * odzt[k] * odzp[k] * dts * 0.5 : odzp[0] * dts * 0.5;
        b = 1.0 + a;
        c = odzp[kz - 1] * dts * 0.5;
    } else if (k == 0) {
        a = odzp[0] * dts * 0.5;
        c = wkk * odzt[1] * odzp[0] * dts * 0.5;
        b = 1.0 + c;
    } else {
        a = wkm1 * odzt[k] * odzp[k] * dts * 0.5;
        c = wkk * odzt[k + 1] * odzp[k] * dts * 0.5;
        b = 1.0 + a + c;
    }
}

// dens.F90:12-14 with C(k,m) at C[(m-1)*km+k].
__device__ __forceinline__ double dens(const double *C, int km, int k, double t, double s) {
    return (C[k] + (C[3 * km + k] + C[6 * km + k] * s) * s +
            (C[2 * km + k] + C[7 * km + k] * s + C[5 * km + k] * t) * t) * t +
           (C[1 * km + k] + (C[4 * km + k] + C[8 * km + k] * s) * s) * s;
}

__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e) {
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// Column view with a stride: level k of a (imt,jmt,km) array at p[k*s].
struct Strided {
    double *p;
    size_t s;
    __device__ __forceinline__ double &operator[](int k) const { return p[(size_t)k * s]; }
};

// convadj_cu mixing (convadj.cpp:52-256) on the columns T[], S[] (length km,
// private arrays or Strided views of the state in global memory).
template <class Col>
__device__ __noinline__ void convect_column(const double *C, const double *to, const double *so,
                                            const double *dzp, int km, int kcon, Col T, Col S,
                                            double *rhoup, double *rholo) {
    if (kcon == 0) return;
    for (int l = 0; l < km; ++l) rhoup[l] = rholo[l] = 0.0;
    for (int l = 0; l < km - 1; ++l) {
        const int l1 = l + 1;
        rhoup[l1] = dens(C, km, l1, T[l1] - to[l1], S[l1] - so[l1]);
        rholo[l] = dens(C, km, l1, T[l] - to[l1], S[l] - so[l1]);
    }
    int lcon = 0;
    for (int k = kcon - 1; k >= 1; --k)
        if (rholo[k - 1] > rhoup[k]) lcon = k;
    if (lcon == 0) return;
    Col X[2] = {T, S};
    for (;;) {   // CONV_1
        int lcona = lcon, lconb = lcon + 1;
        double dztsum = dzp[lcona - 1] + dzp[lconb - 1];
        double trasum[2];
        for (int n = 0; n < 2; ++n) {
            trasum[n] = X[n][lcona - 1] * dzp[lcona - 1] + X[n][lconb - 1] * dzp[lconb - 1];
            const double tramix = trasum[n] / dztsum;
            X[n][lcona - 1] = tramix;
            X[n][lconb - 1] = tramix;
        }
        for (;;) {   // CONV_2
            if (lconb != kcon) {
                const int l1 = lconb + 1;
                rholo[lconb - 1] = dens(C, km, l1 - 1, T[lconb - 1] - to[l1 - 1], S[lconb - 1] - so[l1 - 1]);
                if (rholo[lconb - 1] > rhoup[l1 - 1]) {
                    lconb = lconb + 1;
                    dztsum = dztsum + dzp[lconb - 1];
                    for (int n = 0; n < 2; ++n) {
                        trasum[n] = trasum[n] + X[n][lconb - 1] * dzp[lconb - 1];
                        const double tramix = trasum[n] / dztsum;
                        for (int lm = lcona - 1; lm < lconb; ++lm) X[n][lm] = tramix;
                    }
                    continue;
                }
            }
            if (lcona > 1) {
                const int l1 = lcona - 1;
                rholo[l1 - 1] = dens(C, km, lcona - 1, T[l1 - 1] - to[lcona - 1], S[l1 - 1] - so[lcona - 1]);
                rhoup[lcona - 1] = dens(C, km, lcona - 1, T[lcona - 1] - to[lcona - 1],
                                        S[lcona - 1] - so[lcona - 1]);
                if (rholo[lcona - 2] > rhoup[lcona - 1]) {
                    lcona = lcona - 1;
                    dztsum = dztsum + dzp[lcona - 1];
                    for (int n = 0; n < 2; ++n) {
                        trasum[n] = trasum[n] + X[n][lcona - 1] * dzp[lcona - 1];
                        const double tramix = trasum[n] / dztsum;
                        for (int lm = lcona; lm <= lconb; ++lm) X[n][lm - 1] = tramix;
                    }
                    continue;
                }
            }
            break;
        }
        if (lconb == kcon) return;
        lcon = lconb;
        for (;;) {   // CONV_3
            lcon = lcon + 1;
            if (lcon == kcon) return;
            if (rholo[lcon - 1] <= rhoup[lcon]) continue;
            break;
        }
    }
}

}  // namespace dev
}  // namespace licom
