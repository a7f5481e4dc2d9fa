// turb.hip -- Canuto vertical-mixing producer (AKT, AKMT) and its C ABI
// (include/licom_turb.h).
//
// Replaces turb_cu (readyc.cpp:1456-1611) with turb_device (:1116-1454),
// interp1d_expabs / interp2d_expabs (:915-1115), formld (:902-913) and
// eplatidepend_fun (:882-896), plus the clears before its launch
// (readyc.cpp:1642-1671), and turbinit's table upload (turb_init.cpp:37-88).
//
// One lane per water column, consecutive lanes on consecutive i, so every
// per-level load and store of a wave is one coalesced 512-byte row piece.
// The reference stages each column in eight km-long scratch arrays in HBM
// (wp1/wp3/wp7/wp8 in [i][j][k] order, wk1-3) and passes them to a serial
// routine; here each lane walks its column three times, keeping only
// registers:
//   1. top-down from the surface until the mixed-layer criterion holds
//      (formld; usually a few levels);
//   2. bottom-up: tidal mixing and its running minimum over the levels
//      below (the reference's second loop), written once;
//   3. top-down: the closure (turb_device's level loop) and the final AKMT
//      and AKT of each level, which read back the level's tidal term.
// Level 0 is stored last: the surface floor (wndmix) applies only when no
// deeper level left the closure early.  The density profile and the depths
// of the neighbouring levels are loaded only when the deep length scale
// needs them (ifbelow set, Ri < 0, k >= 2).
//
// The four 2-D Canuto tables are interleaved into one array of 32-byte
// entries {slq2, sm, sh, ss}, so a lookup touches the three corners it
// interpolates between in two cache lines instead of twelve; the zero-shear
// table likewise, and the five background tables in 64-byte entries.
//
// The closure is bound by FP64 divisions and elementary functions, not by
// memory: the values of elementary functions that do not change from level
// to level (pi/4, the logs of the table constants, eplatidepend_fun's
// denominator) are evaluated once per lane.  Forcing a third wave per SIMD
// (launch bounds) spills registers and is slower.
//
// Compiled with -ffp-contract=off and written in the reference's operation
// order; atan, sincos, acosh and log come from the device math library.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "licom_turb.h"
#include "tracer_internal.h"

namespace licom {
namespace {

constexpr int MT = LICOM_TURB_MT, MT0 = LICOM_TURB_MT0, NTH = LICOM_TURB_NTH;
constexpr int W2 = 2 * MT + 1;      // entries of a 1-D table
constexpr int NB = 4 * NTH + 1;     // entries of a background table
constexpr int TPB = 256;

struct alignas(32) Q4 {   // {slq2, sm, sh, ss} or {amtaun2, sm, sh, ss}
    double v[4];
};
struct alignas(64) BK {   // background tables at one angle
    double ra, sm, sh, ss, slq2, pad[3];
};

struct TurbArgs {
    int imt, jmt, km;
    size_t n2;
    const int32_t *kmt;
    const double *zkp, *dzp, *fcort;
    const double *rib, *x1;   // 2-D table keys, zero-shear table keys
    const int32_t *irimax;
    const Q4 *tab2, *tab1;
    const BK *back;
    double rri, rnd2on2, dri, dth, b1, rcrp, rcrn, dand2on2;
    double fricmx, wndmix, nccf;
    licom_turb_fields f;
};

// Elementary-function values that do not change from level to level,
// evaluated once per lane with the same device functions.
struct Consts {
    double quarter;    // atan(1) = pi/4
    double lxs2, lrat2;   // log of the 2-D table's first geometric key and ratio
    double lxs1, lrat1;   // the same for the zero-shear table
    double eden;       // eplatidepend_fun's denominator: f30 * acosh(an0 / f30)
};

__device__ __forceinline__ double fsign(double a, double b) { return b >= 0.0 ? fabs(a) : -fabs(a); }
__device__ __forceinline__ int isgn(int a, int b) { return b >= 0 ? abs(a) : -abs(a); }

// Cell (l0, l1) of x in an expabs table: uniform step delta below xt[MT+MT0],
// ratio rat beyond.  False when the cell does not bracket x (the reference
// then leaves its outputs alone).  The log estimate is kept inside the table.
// lxs = log(xt[MT+MT0]) and lrat = log(rat), the same for every call.
__device__ bool table_cell(const double *__restrict__ xt, double x, double delta, double lxs,
                           double lrat, int &l0, int &l1) {
    const double ax = fabs(x), xs = xt[MT + MT0], xe = xt[2 * MT];
    if (ax >= xs && ax >= xe) {
        l0 = l1 = (int)fsign((double)MT, x);
    } else {
        int l = ax < xs ? (int)(x / delta) + (int)fsign(1.0, x)
                        : (int)fsign((double)MT0 + ((log(ax) - lxs) / lrat), x) +
                              (int)fsign(1.0, x);
        l = min(max(l, -MT), MT);
        if (fabs(xt[MT + l]) < ax)
            l += isgn(1, l);
        else if (fabs(xt[MT + l - isgn(1, l)]) > ax)
            l -= isgn(1, l);
        l0 = l - (int)fsign(1.0, x);
        l1 = ax <= 0.0 ? 1 : l;
    }
    return !((x > 0.0 && (x < xt[MT + l0] || x > xt[MT + l1])) ||
             (x < 0.0 && (x > xt[MT + l0] || x < xt[MT + l1])));
}

// interp1d_expabs over the zero-shear table (readyc.cpp:915-975)
__device__ bool interp_shear0(const TurbArgs &a, const Consts &c, double x, double o[4]) {
    if (__builtin_isnan(x)) return false;
    const double *xt = a.x1;
    x = x > xt[2 * MT] ? xt[2 * MT] : (x < xt[0] ? xt[0] : x);
    int l0, l1;
    if (!table_cell(xt, x, a.dand2on2, c.lxs1, c.lrat1, l0, l1)) return false;
    const double span = xt[MT + l1] - xt[MT + l0], dx = x - xt[MT + l0];
    const Q4 f0 = a.tab1[MT + l0], f1 = a.tab1[MT + l1];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const double s = l1 == l0 ? 0.0 : (f1.v[m] - f0.v[m]) / span;
        o[m] = f0.v[m] + s * dx;
    }
    return true;
}

// interp2d_expabs over the (Ri, Ri_d) table (readyc.cpp:977-1115); ri and
// rid are clamped in place as in the reference.  0: outputs left alone,
// 1: interpolated, 2: beyond the realizable Ri (zeros).
__device__ int interp_table2(const TurbArgs &a, const Consts &c, double &ri, double &rid,
                             double o[4]) {
    const double *rb = a.rib;
    const double lo = rb[0], hi = rb[2 * MT];
    if (ri > hi) {
        if (fabs(rid) <= ri) {
            rid = hi * (rid / ri);
            ri = hi;
        } else if (rid > ri) {
            ri = hi * (ri / rid);
            rid = hi;
        } else if (rid < -ri) {
            ri = lo * (ri / rid);
            rid = lo;
        }
    } else if (ri < lo) {
        if (fabs(rid) <= -ri) {
            rid = lo * (rid / ri);
            ri = lo;
        } else if (rid > -ri) {
            ri = hi * (ri / rid);
            rid = hi;
        } else if (rid < ri) {
            ri = lo * (ri / rid);
            rid = lo;
        }
    } else if (rid > hi) {
        ri = hi * (ri / rid);
        rid = hi;
    } else if (rid < lo) {
        ri = lo * (ri / rid);
        rid = lo;
    }
    if (__builtin_isnan(ri) || __builtin_isnan(rid)) return 0;
    int d0, d1, r0, r1;
    if (!table_cell(rb, rid, a.dri, c.lxs2, c.lrat2, d0, d1)) return 0;
    if (ri > fmin(rb[MT + a.irimax[MT + d0]], rb[MT + a.irimax[MT + d1]])) {
        o[0] = o[1] = o[2] = o[3] = 0.0;
        return 2;
    }
    if (!table_cell(rb, ri, a.dri, c.lxs2, c.lrat2, r0, r1)) return 0;
    const double span_d = rb[MT + d1] - rb[MT + d0], span_r = rb[MT + r1] - rb[MT + r0];
    const double dd = rid - rb[MT + d0], dr = ri - rb[MT + r0];
    const size_t row0 = (size_t)(MT + d0) * W2, row1 = (size_t)(MT + d1) * W2;
    const double *e00 = a.tab2[row0 + MT + r0].v, *e10 = a.tab2[row0 + MT + r1].v,
                 *e01 = a.tab2[row1 + MT + r0].v;
#pragma unroll 1
    for (int m = 0; m < 4; ++m) {
        double s_d = 0.0, s_r = 0.0;
        const double f00 = e00[m];
        if (d1 != d0) {
            s_d = (e01[m] - f00) / span_d;
            s_r = (e10[m] - f00) / span_r;
        }
        o[m] = f00 + s_r * dr + s_d * dd;
    }
    return 1;
}

__device__ double eplat_den() {
    const double an0 = 5.24e-3;
    const double f30 = 4.0 * atan(1.0) / 43082.0;   // f at 30 degrees
    return f30 * acosh(an0 / f30);
}

__global__ __launch_bounds__(TPB) void k_turb(TurbArgs a) {
    __shared__ double sZ[LICOM_MAX_KM + 1], sD[LICOM_MAX_KM];
    const int km = a.km;
    for (int r = threadIdx.x; r <= km; r += TPB) {
        sZ[r] = a.zkp[r];
        if (r < km) sD[r] = a.dzp[r];
    }
    __syncthreads();
    const size_t n2 = a.n2;
    const size_t q = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (q >= n2) return;
    const licom_turb_fields &F = a.f;
    const int i = (int)(q % (size_t)a.imt), j = (int)(q / (size_t)a.imt);
    const bool wet = i > 0 && i < a.imt - 1 && j > 0 && j < a.jmt - 1 && F.vit[q] > 0.5;
    if (!wet) {   // cleared by the reference before the launch
        for (int k = 0; k < km; ++k) {
            F.akmt[(size_t)k * n2 + q] = 0.0;
            F.ak_tide[(size_t)k * n2 + q] = 0.0;
        }
        F.amld[q] = 0.0;
        return;
    }
    const int n = a.kmt[q] - 1;   // levels of the closure (interfaces above the bottom)
    auto vit1 = [&](int k) { return F.vit[(size_t)(k + 1) * n2 + q]; };
    // column values of the reference's wp8 (depth, cm), wp1 (temperature
    // at the interface) and wp3 (density at the interface)
    auto depth = [&](int k) { return k < n ? -vit1(k) * sZ[k + 1] * 1.0e2 : 0.0; };
    auto temp = [&](int k) {
        const double t0 = F.at[(size_t)k * n2 + q], t1 = F.at[(size_t)(k + 1) * n2 + q];
        return vit1(k) * (t0 - (t0 - t1) * sD[k] / (sD[k] + sD[k + 1]));
    };
    auto dens = [&](int k) {
        if (k >= n) return 0.0;
        const double p0 = F.pdensity[(size_t)k * n2 + q], p1 = F.pdensity[(size_t)(k + 1) * n2 + q];
        return vit1(k) * (p0 + (p1 - p0) * sD[k] / (sD[k] + sD[k + 1])) * 1.0e-3;
    };

    // ---- 1. mixed-layer depth (formld)
    double amld = 0.0;
    if (n > 0) {
        const double t0 = temp(0);
        double tp = t0, zp = depth(0);
        amld = depth(n - 1);
        for (int k = 1; k < n; ++k) {
            const double tk = temp(k), zk = depth(k);
            if (fabs(tk - t0) > 0.1) {
                const double tm = t0 - fsign(0.1, t0 - tk);
                amld = zk + (zp - zk) * (tm - tk) / (tp - tk + 1.0e-20);
                break;
            }
            tp = tk;
            zp = zk;
        }
    }
    F.amld[q] = amld * 1.0e-2;

    // ---- 2. tidal mixing, minimum over the levels below, bottom-up
    {
        const double wd = F.wave_dis[q];
        double below = 0.0;
        for (int k = n - 1; k >= 0; --k) {
            const size_t o = (size_t)k * n2 + q;
            const double rc = F.rict[o], fz = F.fz_tide[o], rh = dens(k);
            double v = 1.0e-6 + 0.20 * 0.33 * wd * fz / (fmax(rc, 1.0e-8) * (rh * 1000.0));
            if (v > 1.0e-2) v = 1.0e-2;
            if (k < n - 1) v = fmin(v, below);
            below = v;
            F.ak_tide[o] = v;
            F.richardson[o] = rc;
            F.fztidal[o] = fz;
            F.wp3_tidal[o] = rh;
        }
        for (int k = n > 0 ? n : 0; k < km; ++k) F.ak_tide[(size_t)k * n2 + q] = 0.0;
    }

    // ---- 3. the closure, level by level from the surface
    const double b1 = a.b1, dth = a.dth, coriol = a.fcort[q];
    Consts cst;
    cst.quarter = atan(1.0);
    cst.lxs2 = log(a.rib[MT + MT0]);
    cst.lrat2 = log(a.rri);
    cst.lxs1 = log(a.x1[MT + MT0]);
    cst.lrat1 = log(a.rnd2on2);
    cst.eden = eplat_den();
    const double quarter = cst.quarter;
    const double buoytot = F.vit[q] * F.buoytur[q] * 1.0e4 + F.vit[q] * F.buoysol[q] * 1.0e4;
    const double al0 = 0.17 * amld;
    bool alive = true, below = false;
    double an = 0.0, slq2 = 0.0, sm = 0.0, sh = 0.0, ss = 0.0, amtaun2 = 0.0;
    double sm_bk = 0.0, sh_bk = 0.0, ss_bk = 0.0, slq2_bk = 0.0, back_ri1 = 0.0;
    double akm0 = 0.0, akh0 = 0.0, aks0 = 0.0;
    double *akt_s = F.akt + (size_t)km * n2;
    for (int k = 0; k < km - 1; ++k) {
        const size_t o = (size_t)k * n2 + q;
        double akm = 0.0, akh = 0.0, aks = 0.0;
        if (alive && k < n) {
            // the level's closure; `break` leaves it early (turb_device's returns)
            do {
                const double v = vit1(k);
                const double ri = v * F.rit[o], rid = v * F.ricdttms[o];
                double s2 = v * F.s2t[o];
                const double an2 = v * F.rict[o];
                const double zk = -v * sZ[k + 1] * 1.0e2;
                bool nofsmall = false;
                if (an2 >= 0.0) an = sqrt(an2);
                if (an < fabs(coriol)) nofsmall = true;
                double ri1 = ri, rid1 = rid;
                const double and2 = rid / (ri + 1.0e-25) * an2;
                const double and2on2 = and2 / (an2 + 1.0e-25);
                const bool pure = an2 < a.rib[0] * s2;
                double t4[4];
                if (pure) {
                    if (interp_shear0(a, cst, and2on2, t4)) {
                        amtaun2 = t4[0];
                        sm = t4[1];
                        sh = t4[2];
                        ss = t4[3];
                    }
                    slq2 = (-amtaun2) / ((b1 * b1) * ri + 1.0e-25);
                } else {
                    if (interp_table2(a, cst, ri1, rid1, t4)) {
                        slq2 = t4[0];
                        sm = t4[1];
                        sh = t4[2];
                        ss = t4[3];
                    }
                    if (slq2 < 0.0) {
                        alive = false;
                        break;
                    }
                    if (slq2 == 0.0) below = true;
                }
                const bool upper = (!below || nofsmall || (ri1 < 0.0 && k <= 1)) && slq2 > 0.0;
                const double epsy =
                    slq2 == 0.0 ? 0.0 : -buoytot / ((1.0 / ((b1 * b1) * slq2)) - 0.5 * sm);
                const bool use_epsy = epsy >= 0.0 && upper;
                const double akz = 0.4 * zk;
                const double al = akz * al0 / (al0 + akz);
                double al2 = al * al;
                if (!below && !use_epsy && ri1 > 0.0) {
                    const double anlq2 = slq2 * ri1;
                    if (anlq2 > 0.281) {
                        al2 *= 0.281 / anlq2;
                        slq2 = 0.281 / (ri1 + 1.0e-20);
                    }
                }
                double epson2;
                if (an2 < 0.0) {
                    epson2 = 0.288;
                } else {
                    const double fa = fabs(coriol);
                    const double lat = fmax(nofsmall ? 0.0 : fa * acosh(an / fa) / cst.eden, 7.0e-2);
                    epson2 = 0.288 * lat;
                }
                if (ri > 0.0) {
                    const double rit = (ri + rid) / 2.0, ric = (ri - rid) / 2.0;
                    const double ra_r = sqrt(rit * rit + ric * ric);
                    double theta = fabs(rit) <= 0.0 ? (fabs(ric) <= 0.0 ? quarter : 2.0 * quarter)
                                                    : atan(ric / rit);
                    if (fabs(theta) > 2.0 * quarter) {
                        alive = false;
                        break;
                    }
                    if (theta < -quarter) theta += 4.0 * quarter;
                    int i0 = (int)((theta + quarter) / dth) - NTH, i1 = i0 + 1;
                    double th0 = i0 * dth, th1 = i1 * dth;
                    if (th0 <= a.rcrp && theta > a.rcrp) {
                        theta = th1;
                        th0 = th1;
                        i0 = i1;
                        ++i1;
                        th1 += dth;
                    } else if (th1 >= a.rcrn && theta < a.rcrn) {
                        theta = th0;
                        th1 = th0;
                        i1 = i0;
                        --i0;
                        th0 += -dth;
                    }
                    if (i1 > 3 * NTH || i0 < -NTH) {
                        alive = false;
                        break;
                    }
                    const BK g0 = a.back[NTH + i0], g1 = a.back[NTH + i1];
                    double back_ra = g0.ra + (theta - th0) * ((g1.ra - g0.ra) / dth);
                    const bool glt = (theta <= a.rcrp || theta >= a.rcrn) && back_ra > ra_r;
                    if (glt) back_ra = ra_r;
                    if (back_ra < 0.0) {
                        alive = false;
                        break;
                    }
                    double st, ct;
                    sincos(theta, &st, &ct);
                    const double brit = ct * back_ra, bric = st * back_ra;
                    back_ri1 = brit + bric;
                    double back_rid1 = brit - bric;
                    if (glt) {
                        if (interp_table2(a, cst, back_ri1, back_rid1, t4)) {
                            slq2_bk = t4[0];
                            sm_bk = t4[1];
                            sh_bk = t4[2];
                            ss_bk = t4[3];
                        }
                    } else {
                        const double w = theta - i0 * dth;
                        sm_bk = g0.sm + w * ((g1.sm - g0.sm) / dth);
                        sh_bk = g0.sh + w * ((g1.sh - g0.sh) / dth);
                        ss_bk = g0.ss + w * ((g1.ss - g0.ss) / dth);
                        slq2_bk = g0.slq2 + w * ((g1.slq2 - g0.slq2) / dth);
                    }
                    if (slq2_bk < 0.0) {
                        alive = false;
                        break;
                    }
                }
                if (ri <= 0.0) back_ri1 = 0.0;
                if (ri1 <= 0.0) sm_bk = sh_bk = ss_bk = 0.0;
                if (sm_bk < 0.0 || sh_bk < 0.0 || ss_bk < 0.0) {
                    alive = false;
                    break;
                }
                const double tb = 0.5 * b1 * b1 * back_ri1 * slq2_bk * epson2;
                const double v_bk = tb * sm_bk, t_bk = tb * sh_bk, s_bk = tb * ss_bk;
                if (v_bk < 0.0 || t_bk < 0.0 || s_bk < 0.0 ||
                    (ri > 0.0 && (fabs(v_bk) <= 0.0 || fabs(t_bk) <= 0.0 || fabs(s_bk) < 0.0))) {
                    alive = false;
                    break;
                }
                double tmp = 0.0;
                if (below && ri1 >= 0.0) {
                    tmp = 0.5 * b1 * b1 * ri1 * slq2 * epson2;
                } else if (below && k >= 2) {
                    // deep length scale from the density curvature (loaded on demand)
                    double dz, drh, d2rh;
                    if (k == n - 1) {
                        const double rk = dens(k), rk1 = dens(k - 1);
                        dz = zk - depth(k - 1);
                        drh = rk - rk1;
                        d2rh = rk - 2.0 * rk1 + dens(k - 2);
                    } else {
                        const double rp = dens(k + 1), rm = dens(k - 1);
                        dz = depth(k + 1) - depth(k - 1);
                        drh = rp - rm;
                        d2rh = rp - 2.0 * dens(k) + rm;
                    }
                    const double dzrh = drh / dz;
                    const double d2zrh = 4.0 * d2rh / (dz * dz);
                    const double al0deep = 0.17 * fabs(dzrh / d2zrh);
                    const double akzd = 0.4 * zk;
                    const double aldeep = akzd * al0deep / (al0deep + akzd);
                    al2 = aldeep * aldeep;
                    if (!pure) {
                        s2 = fmax(s2, 1.0e-14);
                        tmp = 0.5 * b1 * al2 * sqrt(s2 / (slq2 + 1.0e-40));
                    }
                } else if (!pure) {
                    s2 = fmax(s2, 1.0e-14);
                    tmp = use_epsy ? 0.5 * epsy / (s2 + 1.0e-40)
                                   : 0.5 * b1 * al2 * sqrt(s2 / (slq2 + 1.0e-40));
                }
                if (pure) tmp = 0.5 * (b1 * b1) * al2 * sqrt(-an2 / amtaun2);
                akm = fmin(tmp * sm + v_bk, a.fricmx);
                akh = fmin(tmp * sh + t_bk, a.fricmx);
                aks = fmin(tmp * ss + s_bk, a.fricmx);
            } while (false);
        }
        if (k == 0) {   // stored after the loop (surface floor)
            akm0 = akm;
            akh0 = akh;
            aks0 = aks;
            continue;
        }
        const double tide = k < n ? F.ak_tide[o] : 0.0;
        F.akmt[o] = akm * 1.0e-4 + tide * 5.0;
        F.akt[o] = (akh * 1.0e-4 + tide) / a.nccf;
        akt_s[o] = (aks * 1.0e-4 + tide) / a.nccf;
    }
    if (alive && n > 0) {
        if (akm0 < a.wndmix) akm0 = a.wndmix;
        if (akh0 < a.wndmix) akh0 = a.wndmix;
        if (aks0 < a.wndmix) aks0 = a.wndmix;
    }
    if (km > 1) {
        const double tide = n > 0 ? F.ak_tide[q] : 0.0;
        F.akmt[q] = akm0 * 1.0e-4 + tide * 5.0;
        F.akt[q] = (akh0 * 1.0e-4 + tide) / a.nccf;
        akt_s[q] = (aks0 * 1.0e-4 + tide) / a.nccf;
    }
    F.akmt[(size_t)(km - 1) * n2 + q] = 0.0;
}

}  // namespace
}  // namespace licom

using namespace licom;

struct licom_turb {
    int imt = 0, jmt = 0, km = 0, device = 0;
    size_t n2 = 0;
    hipStream_t stream = nullptr;
    std::vector<void *> allocs;
    TurbArgs args{};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool ran = false;
};

namespace {

template <class T>
int upload(licom_turb *t, const T *src, size_t count, const T **dst) {
    void *p = nullptr;
    if (int rc = hip_status(hipMalloc(&p, count * sizeof(T)), "hipMalloc")) return rc;
    t->allocs.push_back(p);
    if (int rc = hip_status(hipMemcpy(p, src, count * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy"))
        return rc;
    *dst = (const T *)p;
    return LICOM_OK;
}

int check_tables(const licom_turb_tables *T) {
    const void *ptrs[] = {T->rib,       T->irimax,    T->slq2b, T->smb,   T->shb,   T->ssb,
                          T->and2on2a1, T->amtaun2a1, T->sma1,  T->sha1,  T->ssa1,  T->back_ra_r,
                          T->sm_r1,     T->sh_r1,     T->ss_r1, T->slq2_r1};
    for (const void *p : ptrs)
        if (!p) {
            set_error("licom_turb_create: NULL table");
            return LICOM_EINVAL;
        }
    for (int m = 1; m < W2; ++m)
        if (!(T->rib[m] > T->rib[m - 1]) || !(T->and2on2a1[m] > T->and2on2a1[m - 1])) {
            set_error("licom_turb_create: table keys must increase (entry %d)", m - MT);
            return LICOM_EINVAL;
        }
    for (int m = 0; m < W2; ++m)
        if (T->irimax[m] < -MT || T->irimax[m] > MT) {
            set_error("licom_turb_create: irimax out of range at %d: %d", m - MT, T->irimax[m]);
            return LICOM_EINVAL;
        }
    if (!(T->dri > 0.0) || !(T->dand2on2 > 0.0) || !(T->rri > 1.0) || !(T->rnd2on2 > 1.0) ||
        !(T->deltheta_r > 0.0)) {
        set_error("licom_turb_create: table steps must be positive (ratios above 1)");
        return LICOM_EINVAL;
    }
    return LICOM_OK;
}

}  // namespace

extern "C" {

int licom_turb_create(int imt, int jmt, int km, const int32_t *kmt, const double *zkp,
                      const double *dzp, const double *fcort, const licom_turb_tables *T,
                      int ncc, double dfricmx, double dwndmix, licom_turb **out) {
    if (!out || !kmt || !zkp || !dzp || !fcort || !T) {
        set_error("licom_turb_create: NULL argument");
        return LICOM_EINVAL;
    }
    *out = nullptr;
    if (imt < 3 || jmt < 3 || km < 2 || km > LICOM_MAX_KM) {
        set_error("licom_turb_create: invalid extents imt=%d jmt=%d km=%d (need imt, jmt >= 3, "
                  "2 <= km <= %d)", imt, jmt, km, LICOM_MAX_KM);
        return LICOM_EINVAL;
    }
    if (ncc < 1) {
        set_error("licom_turb_create: ncc must be >= 1 (got %d)", ncc);
        return LICOM_EINVAL;
    }
    const size_t n2 = (size_t)imt * jmt;
    for (size_t q = 0; q < n2; ++q)
        if (kmt[q] < 0 || kmt[q] > km) {
            set_error("licom_turb_create: kmt out of range at %zu: %d (km=%d)", q, kmt[q], km);
            return LICOM_EINVAL;
        }
    if (int rc = check_tables(T)) return rc;
    int ndev = 0, dev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        set_error("no HIP device available");
        return LICOM_ENODEV;
    }
    if (int rc = hip_status(hipGetDevice(&dev), "hipGetDevice")) return rc;
    licom_turb *t = new (std::nothrow) licom_turb();
    if (!t) return LICOM_ENOMEM;
    t->imt = imt;
    t->jmt = jmt;
    t->km = km;
    t->n2 = n2;
    t->device = dev;
    TurbArgs &a = t->args;
    a.imt = imt;
    a.jmt = jmt;
    a.km = km;
    a.n2 = n2;
    int rc = LICOM_OK;
    // interleaved tables
    std::vector<Q4> h2((size_t)W2 * W2), h1(W2);
    for (size_t e = 0; e < h2.size(); ++e)
        h2[e] = Q4{{T->slq2b[e], T->smb[e], T->shb[e], T->ssb[e]}};
    for (int e = 0; e < W2; ++e) h1[e] = Q4{{T->amtaun2a1[e], T->sma1[e], T->sha1[e], T->ssa1[e]}};
    std::vector<BK> hb(NB);
    for (int e = 0; e < NB; ++e)
        hb[e] = BK{T->back_ra_r[e], T->sm_r1[e], T->sh_r1[e], T->ss_r1[e], T->slq2_r1[e], {0, 0, 0}};
    if (!rc) rc = upload(t, kmt, n2, &a.kmt);
    if (!rc) rc = upload(t, zkp, (size_t)km + 1, &a.zkp);
    if (!rc) rc = upload(t, dzp, (size_t)km, &a.dzp);
    if (!rc) rc = upload(t, fcort, n2, &a.fcort);
    if (!rc) rc = upload(t, T->rib, (size_t)W2, &a.rib);
    if (!rc) rc = upload(t, T->and2on2a1, (size_t)W2, &a.x1);
    if (!rc) rc = upload(t, T->irimax, (size_t)W2, &a.irimax);
    if (!rc) rc = upload(t, h2.data(), h2.size(), &a.tab2);
    if (!rc) rc = upload(t, h1.data(), h1.size(), &a.tab1);
    if (!rc) rc = upload(t, hb.data(), hb.size(), &a.back);
    if (!rc) rc = hip_status(hipEventCreate(&t->ev0), "hipEventCreate");
    if (!rc) rc = hip_status(hipEventCreate(&t->ev1), "hipEventCreate");
    if (rc) {
        licom_turb_destroy(t);
        return rc;
    }
    a.rri = T->rri;
    a.rnd2on2 = T->rnd2on2;
    a.dri = T->dri;
    a.dth = T->deltheta_r;
    a.b1 = T->b1;
    a.rcrp = T->theta_rcrp;
    a.rcrn = T->theta_rcrn;
    a.dand2on2 = T->dand2on2;
    a.fricmx = dfricmx * 1.0e4;   // turb_cu passes them in cm^2/s (readyc.cpp:1510)
    a.wndmix = dwndmix * 1.0e4;
    a.nccf = (double)(float)ncc;  // AKT divides by (float) ncc (readyc.cpp:1570)
    *out = t;
    return LICOM_OK;
}

int licom_turb_set_stream(licom_turb *t, void *hip_stream) {
    if (!t) {
        set_error("licom_turb_set_stream: NULL producer");
        return LICOM_EINVAL;
    }
    t->stream = (hipStream_t)hip_stream;
    return LICOM_OK;
}

int licom_turb_run(licom_turb *t, const licom_turb_fields *f) {
    if (!t || !f) {
        set_error("licom_turb_run: NULL argument");
        return LICOM_EINVAL;
    }
    const void *need[] = {f->at,   f->pdensity, f->rit,     f->ricdttms,   f->s2t,     f->rict,
                          f->fz_tide, f->vit,   f->buoytur, f->buoysol,    f->wave_dis, f->akmt,
                          f->akt,  f->ak_tide,  f->amld,    f->richardson, f->fztidal, f->wp3_tidal};
    for (const void *p : need)
        if (!p) {
            set_error("licom_turb_run: a field pointer is NULL");
            return LICOM_EINVAL;
        }
    t->args.f = *f;
    const unsigned blocks = (unsigned)((t->n2 + TPB - 1) / TPB);
    if (int rc = hip_status(hipEventRecord(t->ev0, t->stream), "hipEventRecord")) return rc;
    hipLaunchKernelGGL(k_turb, dim3(blocks), dim3(TPB), 0, t->stream, t->args);
    if (int rc = hip_status(hipGetLastError(), "k_turb launch")) return rc;
    if (int rc = hip_status(hipEventRecord(t->ev1, t->stream), "hipEventRecord")) return rc;
    t->ran = true;
    return LICOM_OK;
}

int licom_turb_sync(licom_turb *t) {
    if (!t) {
        set_error("licom_turb_sync: NULL producer");
        return LICOM_EINVAL;
    }
    return hip_status(hipStreamSynchronize(t->stream), "hipStreamSynchronize");
}

int licom_turb_last_ms(licom_turb *t, double *ms) {
    if (!t || !ms) {
        set_error("licom_turb_last_ms: NULL argument");
        return LICOM_EINVAL;
    }
    *ms = 0.0;
    if (!t->ran) return LICOM_OK;
    if (int rc = hip_status(hipEventSynchronize(t->ev1), "hipEventSynchronize")) return rc;
    float v = 0.0f;
    if (int rc = hip_status(hipEventElapsedTime(&v, t->ev0, t->ev1), "hipEventElapsedTime")) return rc;
    *ms = v;
    return LICOM_OK;
}

void licom_turb_destroy(licom_turb *t) {
    if (!t) return;
    if (t->ran) (void)hipStreamSynchronize(t->stream);
    for (void *p : t->allocs) (void)hipFree(p);
    if (t->ev0) (void)hipEventDestroy(t->ev0);
    if (t->ev1) (void)hipEventDestroy(t->ev1);
    delete t;
}

}  // extern "C"
