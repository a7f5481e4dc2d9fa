// readyc.hip -- readyc_ on gfx950 (include/licom_readyc.h): velocity shear
// and Richardson numbers for the Canuto producer (readyc1, readyc1_2), AKMT
// averaged to U points after it (readyc2), then the momentum tendencies.
//
// The reference clears six 3-D arrays and runs two 3-D launches, the first
// writing up / vp averaged to T points (wp12, wp13) that the second reads
// back a level apart.  Here one lane per column (i-consecutive lanes)
// marches k once: the T-point average of level k is formed from the four U
// columns around the T point and kept in registers for the interface below,
// every output cell (including the clears) is written exactly once, and each
// level's inputs are loaded a level ahead through plane buffers, before the
// level's stores.  Operation order follows the reference statements
// (-ffp-contract=off), so results are identical bit for bit.
#include <hip/hip_runtime.h>

#include <new>
#include <vector>

#include "licom_readyc.h"
#include "tracer_internal.h"

namespace licom {
namespace {

struct CGrid {
    int imt, jmt, km;
    size_t n2, n3;
    const double *at0, *atn, *ate, *atne, *au0, *aus, *auw, *ausw, *odzt;
};

__global__ __launch_bounds__(256) void k_readyc_shear(CGrid g, licom_readyc_fields f) {
    __shared__ double sOZ[LICOM_MAX_KM];
    {
        const int t = threadIdx.x + 64 * threadIdx.y;
        for (int r = t; r < g.km; r += 256) sOZ[r] = g.odzt[r];
        __syncthreads();
    }
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
    const int imt = g.imt, jmt = g.jmt, km = g.km;
    if (i >= imt || j >= jmt) return;
    const size_t n2 = g.n2, q = (size_t)j * imt + i;
    const bool tp = j >= 1 && i < imt - 1;   // T points readyc1 / readyc1_2 average to
    double w0 = 0.0, wn = 0.0, we = 0.0, wne = 0.0;
    if (tp) {
        w0 = g.at0[q];
        wn = g.atn[q];
        we = g.ate[q];
        wne = g.atne[q];
    }
    {
        const double h = f.h0bf[q];
        f.h0bl[q] = h;
        f.h0bf[q] = f.h0[q];
    }
    const size_t pb = n2 * sizeof(double);
    const uint32_t oq = (uint32_t)q * 8u;
    const uint32_t on = tp ? (uint32_t)(q - imt) * 8u : kBufOff;
    const uint32_t oe = tp ? (uint32_t)(q + 1) * 8u : kBufOff;
    const uint32_t one = tp ? (uint32_t)(q - imt + 1) * 8u : kBufOff;
    struct LvIn {
        double vt, vv, u, v, un, ue, une, vn, ve, vne, vvn, vve, vvne, rict, ric;
    };
    auto load_level = [&](int k, LvIn &L) {
        const size_t o = (size_t)k * n2;
        const auto rvv = plane_rsrc(f.viv + o, pb), ru = plane_rsrc(f.up + o, pb),
                   rv = plane_rsrc(f.vp + o, pb);
        L.vt = buf_load(plane_rsrc(f.vit + o, pb), oq, 0.0);
        L.vv = buf_load(rvv, oq, 0.0);
        L.u = buf_load(ru, oq, 0.0);
        L.v = buf_load(rv, oq, 0.0);
        L.un = buf_load(ru, on, 0.0);
        L.ue = buf_load(ru, oe, 0.0);
        L.une = buf_load(ru, one, 0.0);
        L.vn = buf_load(rv, on, 0.0);
        L.ve = buf_load(rv, oe, 0.0);
        L.vne = buf_load(rv, one, 0.0);
        L.vvn = buf_load(rvv, on, 0.0);
        L.vve = buf_load(rvv, oe, 0.0);
        L.vvne = buf_load(rvv, one, 0.0);
        // N^2 of the interface below level k (km-1 planes)
        const bool ri = k < km - 1;
        L.rict = buf_load(plane_rsrc(f.rict + (ri ? o : 0), ri ? pb : 0), oq, 0.0);
        L.ric = buf_load(plane_rsrc(f.ric + (ri ? o : 0), ri ? pb : 0), oq, 0.0);
    };
    constexpr double epsln = 1.0e-25;
    double wp12p = 0.0, wp13p = 0.0, vtp = 0.0, up = 0.0, vpp = 0.0, rictp = 0.0, ricp = 0.0;
    LvIn cur, nxt;
    load_level(0, cur);
    f.riu[q] = 0.0;
    for (int k = 0; k < km; ++k) {
        const size_t c = (size_t)k * n2 + q;
        if (k + 1 < km) load_level(k + 1, nxt);
        // readyc1 (readyc.cpp:58-104): up, vp at the T point, wet U cells only
        double wp12 = 0.0, wp13 = 0.0;
        if (tp) {
            const double den = cur.vv * w0 + cur.vvn * wn + cur.vve * we + cur.vvne * wne + 1.0e-25;
            wp12 = cur.vt * (w0 * cur.u * cur.vv + wn * cur.un * cur.vvn + we * cur.ue * cur.vve +
                             wne * cur.une * cur.vvne) / den;
            wp13 = cur.vt * (w0 * cur.v * cur.vv + wn * cur.vn * cur.vvn + we * cur.ve * cur.vve +
                             wne * cur.vne * cur.vvne) / den;
        }
        f.wp12[c] = wp12;
        f.wp13[c] = wp13;
        if (k >= 1) {
            // readyc1_2 (readyc.cpp:118-167) at the interface between levels
            // k-1 and k (its index k-1)
            const size_t cm = c - n2;
            const double oz = sOZ[k];
            double s2t = 0.0;
            if (tp) {
                const double r1 = wp12p * vtp - wp12 * cur.vt;
                const double r2 = wp13p * vtp - wp13 * cur.vt;
                s2t = cur.vt * (r1 * r1 + r2 * r2) * oz * oz;
                f.rit[cm] = cur.vt * rictp / (s2t + epsln);
            }
            f.s2t[cm] = s2t;
            const double u1 = up - cur.u, u2 = vpp - cur.v;
            f.riv1[cm] = u1;
            f.riv2[cm] = u2;
            const double s2u = cur.vv * (u1 * u1 + u2 * u2) * oz * oz;
            f.s2u[cm] = s2u;
            f.riu[c] = cur.vv * ricp / (s2u + epsln);
        }
        wp12p = wp12;
        wp13p = wp13;
        vtp = cur.vt;
        up = cur.u;
        vpp = cur.v;
        rictp = cur.rict;
        ricp = cur.ric;
        cur = nxt;
    }
    // the cleared bottom planes readyc1_2 does not write
    const size_t cb = (size_t)(km - 1) * n2 + q;
    f.s2t[cb] = 0.0;
    f.riv1[cb] = 0.0;
    f.riv2[cb] = 0.0;
}

// readyc2 (readyc.cpp:223-246) with the clear of akmu before turb_cu
__global__ __launch_bounds__(256) void k_readyc_akmu(CGrid g, const double *__restrict__ akmt,
                                                     const double *__restrict__ viv,
                                                     double *__restrict__ akmu) {
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= g.n3) return;
    const int imt = g.imt, jmt = g.jmt;
    const size_t n2 = g.n2, q = e % n2;
    const int k = (int)(e / n2), j = (int)(q / imt), i = (int)(q % imt);
    double v = 0.0;
    if (i > 0 && j < jmt - 1 && k < g.km - 1)
        v = (g.au0[q] * akmt[e] + g.aus[q] * akmt[e + imt] + g.auw[q] * akmt[e - 1] +
             g.ausw[q] * akmt[e + imt - 1]) *
            viv[e + n2];
    akmu[e] = v;
}


// ---------------------------------------------------------------------------
// Momentum tendencies after readyc2.  The reference runs up to twelve 3-D
// launches over whole arrays, with eight 3-D scratch arrays between them.
// Here:
//   k_rmom_upwell   one lane per T column: the scaled velocities of the four
//                   U cells around it (staged in LDS per level, each U cell
//                   scaled once for the block), their divergence level by
//                   level (kept in registers), its column integral, then the
//                   recursion of ws (upwell_cu_0 + upwell_cu);
//   k_rmom_lap1     (del4) a 64x4 block of U columns marching k with an
//                   LDS window of (up, vp): divergence and vorticity once per
//                   T point (hdel4_div, hdel4_zcurl), then per U column the
//                   viscosity factor from those of its four T points
//                   (hdel4_grad twice) and the first Laplacian scaled by it
//                   (hdiffu_del4_0) into scratch;
//   k_rmom_main     one lane per U column marching k once: w and the face
//                   velocities at U points (readyc3) formed on the fly,
//                   advection, the horizontal viscosity (hdiffu_del2_cu or
//                   hdiffu_del4_1), the depth sums and the surface and bottom
//                   stress (readyc4) and the explicit vertical viscosity
//                   (readyc4_2), every output cell written once; the
//                   neighbours' u, v, ws and viscosity operand come from a
//                   per-level LDS window of the block.
// Each level's inputs are loaded a level ahead through plane buffers
// (out-of-range offsets return zero for the lanes that do not use them).
// Operation order follows the reference statements (-ffp-contract=off).
struct MGrid {
    int ib, ie, jb, je, del4;
    double am, c0f, od0, cag, sag;
    const int *kmt, *kmu;
    const double *ohbu, *ohbt, *hts, *htw, *hue, *hun, *tarea_r, *uarea_r, *dxu, *dyu, *dxur, *dyur,
        *snlat, *amf, *duc, *dum, *dun, *dus, *due, *duw, *dmc, *dmn, *dms, *dme, *dmw, *dzp, *odzp;
    double *d2u, *d2v;   // del4 scratch, [km][j][i]
};

__device__ __forceinline__ uint32_t boff(bool ok, size_t idx) { return ok ? (uint32_t)idx * 8u : kBufOff; }

// The U cells of a 64x4 block of T columns: rows j0-1..j0+3, columns
// i0..i0+64 (each serves up to four T columns); the lane that owns one holds
// its upwell_cu_0 factor and stages (u, v) times it in LDS per level.
constexpr int UWX = 65, UWN = UWX * 5;
struct UCell {
    double s;        // 1 + (h0 at the U point) / H there
    uint32_t off;    // plane offset (kBufOff: no U cell the block uses)
    int e;           // index among the block's U cells (UWN: the lane owns none)
};
__device__ __forceinline__ UCell ucell(const CGrid &g, const MGrid &m, const licom_readyc_mfields &f, int i0,
                                       int j0, int e) {
    UCell u;
    u.e = e < UWN ? e : UWN;
    u.s = 0.0;
    u.off = kBufOff;
    if (e >= UWN) return u;
    const int r = e / UWX, iu = i0 + (e - r * UWX), ju = j0 - 1 + r;
    // the T columns 1 <= j <= jmt-2, 1 <= i <= imt-2 use U cells
    // 0 <= ju <= jmt-2, 1 <= iu <= imt-1
    if (ju < 0 || ju > g.jmt - 2 || iu < 1 || iu > g.imt - 1) return u;
    const size_t p = (size_t)ju * g.imt + iu, imt = g.imt;
    // upwell_cu_0 (readyc.cpp:609-621)
    const double w = g.au0[p] * f.h0[p] + g.aus[p] * f.h0[p + imt] + g.auw[p] * f.h0[p - 1] +
                     g.ausw[p] * f.h0[p + imt - 1];
    u.s = 1.0e0 + w * m.ohbu[p];
    u.off = (uint32_t)p * 8u;
    return u;
}

__global__ __launch_bounds__(256) void k_rmom_upwell(CGrid g, MGrid m, licom_readyc_mfields f) {
    __shared__ double sDZ[LICOM_MAX_KM];
    __shared__ double sUK[2][UWN], sVK[2][UWN];
    const int tx = threadIdx.x, ty = threadIdx.y, tl = tx + 64 * ty;
    for (int r = tl; r < g.km; r += 256) sDZ[r] = m.dzp[r];
    const int i0 = blockIdx.x * 64, j0 = blockIdx.y * 4;
    const int i = i0 + tx, j = j0 + ty;
    const int imt = g.imt, jmt = g.jmt, km = g.km;
    // upwell_cu writes ws in 1 <= j < jmt-1, 1 <= i < imt-1 only
    const bool col = i >= 1 && i <= imt - 2 && j >= 1 && j <= jmt - 2;
    const size_t n2 = g.n2, q = col ? (size_t)j * imt + i : 0;
    const UCell C0 = ucell(g, m, f, i0, j0, tl), C1 = ucell(g, m, f, i0, j0, tl + 256);
    double htwe = 0, htwc = 0, htsc = 0, htsn = 0, tr = 0, ohbt = 0, h0 = 0;
    int kmt = 0;
    if (col) {
        htwe = m.htw[q + 1];
        htwc = m.htw[q];
        htsc = m.hts[q];
        htsn = m.hts[q - imt];
        tr = m.tarea_r[q];
        ohbt = m.ohbt[q];
        h0 = f.h0[q];
        kmt = m.kmt[q];
    }
    const size_t pb = n2 * sizeof(double);
    const uint32_t oc = col ? (uint32_t)q * 8u : kBufOff;
    double lu0, lv0, lu1, lv1, vt, vtn;
    auto load = [&](int k) {
        const size_t o = (size_t)k * n2;
        const auto ru = plane_rsrc(f.u + o, pb), rv = plane_rsrc(f.v + o, pb);
        lu0 = buf_load(ru, C0.off, 0.0);
        lv0 = buf_load(rv, C0.off, 0.0);
        lu1 = buf_load(ru, C1.off, 0.0);
        lv1 = buf_load(rv, C1.off, 0.0);
        vtn = buf_load(plane_rsrc(f.vit + o, pb), oc, 0.0);
    };
    // this T column's U cells (j, i+1), (j-1, i+1), (j, i), (j-1, i)
    const int ec = (ty + 1) * UWX + tx, ea = ec + 1, ed = ty * UWX + tx, eb = ed + 1;
    // upwell_cu (readyc.cpp:636-671): divergence per level, its integral
    double wk[LICOM_MAX_KM];
    double work = 0.0e0;
    uint64_t vone = 0;      // levels whose vit is exactly 1.0
    bool v01 = true;        // every vit of the column is +0.0 or 1.0
    load(0);
#pragma unroll
    for (int k = 0; k < LICOM_MAX_KM; ++k) {
        if (k < km) {
            const int b = k & 1;
            if (C0.e < UWN) {
                sUK[b][C0.e] = C0.s * lu0;
                sVK[b][C0.e] = C0.s * lv0;
            }
            if (C1.e < UWN) {
                sUK[b][C1.e] = C1.s * lu1;
                sVK[b][C1.e] = C1.s * lv1;
            }
            vt = vtn;
            if (k + 1 < km) load(k + 1);
            __syncthreads();
            double d = 0.0e0;
            if (k <= kmt - 1) {
                const double *UK = sUK[b], *VK = sVK[b];
                const double uka = UK[ea], ukb = UK[eb], ukc = UK[ec], ukd = UK[ed];
                const double vka = VK[ea], vkb = VK[eb], vkc = VK[ec], vkd = VK[ed];
                d = 0.5 * ((uka + ukb) * htwe - (ukc + ukd) * htwc + (vka + vkc) * htsc - (vkb + vkd) * htsn) *
                    tr;
            }
            wk[k] = d;
            work -= sDZ[k] * d * vt;
            const uint64_t bits = __builtin_bit_cast(uint64_t, vt);
            if (bits == 0x3FF0000000000000ull)
                vone |= 1ull << k;
            else if (bits != 0)
                v01 = false;
        }
    }
    if (!col) return;
    const double wo = work * ohbt;
    const double r = 1.0e0 / (1.0e0 + h0 * ohbt);
    double prev = f.ws[q];
#pragma unroll
    for (int k = 1; k < LICOM_MAX_KM; ++k) {
        if (k < km) {
            const size_t c = (size_t)k * n2 + q;
            const double vt = v01 ? (((vone >> k) & 1ull) ? 1.0 : 0.0) : f.vit[c];
            const double w = vt * (prev + sDZ[k - 1] * (wo + wk[k - 1]));
            f.ws[c] = w * r;
            prev = w;
        }
    }
}

// the 5-point operator of hdiffu_*: the du coefficients on a, dm on b
struct Lap {
    double c, n, s, e, w, mc, mn, ms, me, mw;
};
__device__ __forceinline__ double lap_du(const Lap &L, double a, double an, double as, double ae, double aw) {
    return L.c * a + L.n * an + L.s * as + L.e * ae + L.w * aw;
}
__device__ __forceinline__ double lap_dm(const Lap &L, double b, double bn, double bs, double be, double bw) {
    return L.mc * b + L.mn * bn + L.ms * bs + L.me * be + L.mw * bw;
}
__device__ __forceinline__ Lap load_lap(const MGrid &m, size_t q) {
    Lap L;
    L.c = m.duc[q] + m.dum[q];   // (duc + dum), summed first as in the reference
    L.n = m.dun[q];
    L.s = m.dus[q];
    L.e = m.due[q];
    L.w = m.duw[q];
    L.mc = m.dmc[q];
    L.mn = m.dmn[q];
    L.ms = m.dms[q];
    L.me = m.dme[q];
    L.mw = m.dmw[q];
    return L;
}

// hdel4_div and hdel4_zcurl at the T points of a 64x4 block of U columns,
// hdel4_grad of both (readyc.cpp:343-441) and hdiffu_del4_0 (:444-507).
// Per level the block stages up and vp over rows j0-1..j0+4 and columns
// i0-2..i0+64 (flat offsets from the block's corner, so a column outside
// 0..imt-1 reads the neighbouring row's cell, as the flat stencil offsets of
// the reference do; offsets outside the plane read 0) in LDS, computes the
// divergence and the vorticity once per T point (rows j0..j0+4, columns
// i0-1..i0+63; each is shared by four U columns), then every U column forms
// its viscosity factor and first Laplacian from LDS.  A T point whose
// stencil leaves the arrays is used by no U column inside the block bounds
// and gets 0.
constexpr int L1WX = 67, L1WN = L1WX * 6;   // up/vp window
constexpr int L1TX = 65, L1TN = L1TX * 5;   // T points
struct TPoint {
    double htwe, htwc, htsc, htsn, tr, dyc, dyn, dyw, dynw, dxc, dxn, dxw, dxnw;
    int kt, e;       // kmt (-1: no T point here), LDS index
    bool cu;         // i >= 1 (hdel4_zcurl)
};
__device__ __forceinline__ void tpoint_load(const CGrid &g, const MGrid &m, int i0, int j0, int e, TPoint &T) {
    T.kt = -1;
    T.e = e;
    if (e >= L1TN) return;
    const int tr = e / L1TX, tc = e - tr * L1TX;
    const int it = i0 - 1 + tc, jt = j0 + tr;
    const long long p = (long long)jt * g.imt + it;
    T.cu = it >= 1;
    if (p - g.imt - (T.cu ? 1 : 0) < 0 || p + 1 >= (long long)g.n2) return;
    T.htwe = m.htw[p + 1];
    T.htwc = m.htw[p];
    T.htsc = m.hts[p];
    T.htsn = m.hts[p - g.imt];
    T.tr = m.tarea_r[p];
    T.kt = m.kmt[p];
    if (T.cu) {
        T.dyc = m.dyu[p];
        T.dyn = m.dyu[p - g.imt];
        T.dyw = m.dyu[p - 1];
        T.dynw = m.dyu[p - g.imt - 1];
        T.dxc = m.dxu[p];
        T.dxn = m.dxu[p - g.imt];
        T.dxw = m.dxu[p - 1];
        T.dxnw = m.dxu[p - g.imt - 1];
    }
}
// divergence and vorticity of level k at one T point from the window; the
// operands are named as the U column (j, i) with T point (j, i) names them
__device__ __forceinline__ void tpoint_eval(const TPoint &T, int k, const double *U, const double *V,
                                            double *sDV, double *sCZ) {
    if (T.e >= L1TN) return;
    double dv = 0.0e0, cz = 0.0e0;
    if (k <= T.kt - 1) {
        const int tr = T.e / L1TX, tc = T.e - tr * L1TX;
        const int a = (tr + 1) * L1WX + tc + 1, b = tr * L1WX + tc + 1;   // rows j, j-1; column i
        dv = 0.5 * ((U[a + 1] + U[b + 1]) * T.htwe - (U[a] + U[b]) * T.htwc + (V[a + 1] + V[a]) * T.htsc -
                    (V[b + 1] + V[b]) * T.htsn) *
             T.tr;
        if (T.cu)
            cz = 0.5 * ((V[a] * T.dyc + V[b] * T.dyn - V[a - 1] * T.dyw - V[b - 1] * T.dynw - U[a] * T.dxc -
                         U[a - 1] * T.dxw + U[b] * T.dxn + U[b - 1] * T.dxnw));
    }
    sDV[T.e] = dv;
    sCZ[T.e] = cz;
}

__global__ __launch_bounds__(256) void k_rmom_lap1(CGrid g, MGrid m, licom_readyc_mfields f) {
    __shared__ double sU[2][L1WN], sV[2][L1WN];
    __shared__ double sDV[L1TN], sCZ[L1TN];
    const int tx = threadIdx.x, ty = threadIdx.y, t = tx + 64 * ty;
    const int i0 = blockIdx.x * 64, j0 = blockIdx.y * 4;
    const int i = i0 + tx, j = j0 + ty;
    const int imt = g.imt, jmt = g.jmt, km = g.km;
    const size_t n2 = g.n2;
    const bool in = i < imt && j < jmt;
    const size_t q = in ? (size_t)j * imt + i : 0;
    // the reference compares 0-based j with POP's 1-based bounds as given
    const bool reg = in && j >= m.jb - 2 && j <= m.je && i >= m.ib - 2 && i <= m.ie;
    const double amf = in ? m.amf[q] : 0.0;
    const int kmu = in ? m.kmu[q] : 0;
    Lap L{};
    double gxf = 0, gyf = 0, dxdy = 0, amd = 1.0;
    if (reg) {
        L = load_lap(m, q);
        gxf = m.dxur[q] * 0.5;
        gyf = m.dyur[q] * 0.5;
        const double t_area = sqrt(1.0e0 / m.uarea_r[q]);
        dxdy = t_area * t_area * t_area * t_area * t_area * 45.0e0;
        amd = fabs(m.am * amf);
    }
    TPoint T0, T1;
    tpoint_load(g, m, i0, j0, t, T0);
    tpoint_load(g, m, i0, j0, t + 256, T1);
    // this lane's window cells e = t, t + 256
    const size_t pb = n2 * sizeof(double);
    uint32_t wo[2];
    bool wv[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int e = t + 256 * u, r = e / L1WX, c = e - r * L1WX;
        const long long fl = (long long)(j0 - 1 + r) * imt + (i0 - 2 + c);
        wv[u] = e < L1WN;
        wo[u] = boff(wv[u] && fl >= 0 && fl < (long long)n2, (size_t)fl);
    }
    double lu[2], lv[2];
    auto load = [&](int k) {
        const size_t o = (size_t)k * n2;
        const auto ru = plane_rsrc(f.up + o, pb), rv = plane_rsrc(f.vp + o, pb);
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            lu[u] = buf_load(ru, wo[u], 0.0);
            lv[u] = buf_load(rv, wo[u], 0.0);
        }
    };
    constexpr int DR[4] = {0, 1, 0, 1}, DC[4] = {0, 0, -1, -1};
    load(0);
    for (int k = 0; k < km; ++k) {
        double *U = sU[k & 1], *V = sV[k & 1];
#pragma unroll
        for (int u = 0; u < 2; ++u)
            if (wv[u]) {
                U[t + 256 * u] = lu[u];
                V[t + 256 * u] = lv[u];
            }
        if (k + 1 < km) load(k + 1);
        __syncthreads();
        tpoint_eval(T0, k, U, V, sDV, sCZ);
        tpoint_eval(T1, k, U, V, sDV, sCZ);
        __syncthreads();
        double am_f = 1.0e0, hu = 0.0, hv = 0.0;
        if (reg && k <= kmu - 1) {
            double dv[4], cz[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int e = (ty + DR[s]) * L1TX + tx + 1 + DC[s];
                dv[s] = sDV[e];
                cz[s] = sCZ[e];
            }
            const double gx = gxf * (dv[1] - dv[2] - dv[3] + dv[0]);
            const double gy = gyf * (dv[1] - dv[2] + dv[3] - dv[0]);
            double af = 0.0 + (gx * gx + gy * gy);
            const double gx2 = gxf * (cz[1] - cz[2] - cz[3] + cz[0]);
            const double gy2 = gyf * (cz[1] - cz[2] + cz[3] - cz[0]);
            af += gx2 * gx2 + gy2 * gy2;
            am_f = sqrt(af) * dxdy / amd;
            am_f = fmin(40.0e0, am_f);
            am_f = fmax(1.0e0, am_f);
            // window cell of (j + r - 1, i + c - 2) is (ty + r) * L1WX + tx + c
            const int w = (ty + 1) * L1WX + tx + 2;
            hu = lap_du(L, U[w], U[w - L1WX], U[w + L1WX], U[w + 1], U[w - 1]) +
                 lap_dm(L, V[w], V[w - L1WX], V[w + L1WX], V[w + 1], V[w - 1]);
            hv = lap_du(L, V[w], V[w - L1WX], V[w + L1WX], V[w + 1], V[w - 1]) -
                 lap_dm(L, U[w], U[w - L1WX], U[w + L1WX], U[w + 1], U[w - 1]);
        }
        if (in) {
            const size_t c = (size_t)k * n2 + q;
            m.d2u[c] = am_f * amf * hu;
            m.d2v[c] = am_f * amf * hv;
        }
    }
}

// Window of the U points around a 64x4 block: rows j0-1..j0+4, columns
// i0-1..i0+64, at flat offsets from the block's corner (0 outside the plane).
constexpr int MWX = 66, MWN = MWX * 6;

template <bool DEL4>
__global__ __launch_bounds__(256, 2) void k_rmom_main(CGrid g, MGrid m, licom_readyc_mfields f) {
    __shared__ double sDZ[LICOM_MAX_KM], sOZP[LICOM_MAX_KM], sOZT[LICOM_MAX_KM];
    // per level: u, v and the viscosity operand of level k, ws of level k+1
    __shared__ double sWU[2][MWN], sWV[2][MWN], sWA[2][MWN], sWB[2][MWN], sWW[2][MWN];
    const int tx = threadIdx.x, ty = threadIdx.y, tl = tx + 64 * ty;
    for (int r = tl; r < g.km; r += 256) {
        sDZ[r] = m.dzp[r];
        sOZP[r] = m.odzp[r];
        sOZT[r] = g.odzt[r];
    }
    const int i0 = blockIdx.x * 64, j0 = blockIdx.y * 4;
    const int i = i0 + tx, j = j0 + ty;
    const int imt = g.imt, jmt = g.jmt, km = g.km;
    const bool act = i < imt && j < jmt;
    const size_t n2 = g.n2, q = act ? (size_t)j * imt + i : 0;
    const bool in = act && j >= 1 && j < jmt - 1 && i >= 1 && i < imt - 1;
    // hdiffu_del2_cu / hdiffu_del4_1: 0-based j against POP's 1-based bounds
    const bool hreg = act && j >= m.jb - 1 && j < m.je && i >= m.ib - 1 && i < m.ie;
    const double ohbu = act ? m.ohbu[q] : 0.0, c0f = m.c0f, cag = m.cag, sag = m.sag, am = m.am;
    const int kmu = act ? m.kmu[q] : 0;
    double a0 = 0, as = 0, aw = 0, asw = 0, huew = 0, huec = 0, huns = 0, hunc = 0, ua = 0, sn = 0, su = 0,
           sv = 0;
    Lap L{};
    if (in) {
        a0 = g.au0[q];
        as = g.aus[q];
        aw = g.auw[q];
        asw = g.ausw[q];
        huew = m.hue[q - 1];
        huec = m.hue[q];
        huns = m.hun[q + imt];
        hunc = m.hun[q];
        ua = m.uarea_r[q];
        sn = m.snlat[q];
        su = f.su[q];
        sv = f.sv[q];
    }
    if (hreg) L = load_lap(m, q);
    const size_t pb = n2 * sizeof(double);
    const uint32_t oc = act ? (uint32_t)q * 8u : kBufOff, ocin = in ? oc : kBufOff;
    const double *ha = DEL4 ? m.d2u : f.up, *hb = DEL4 ? m.d2v : f.vp;
    // this lane's window cells e = tl, tl + 256
    uint32_t wo[2];
    bool wv[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int e = tl + 256 * u, r = e / MWX, c = e - r * MWX;
        const long long fl = (long long)(j0 - 1 + r) * imt + (i0 - 1 + c);
        wv[u] = e < MWN;
        wo[u] = boff(wv[u] && fl >= 0 && fl < (long long)n2, (size_t)fl);
    }
    // window values of the next level to stage: u, v, a, b of level k, ws of k+1
    double lu[2], lv[2], la[2], lb[2], lw[2];
    auto load_win = [&](int k) {
        const size_t o = (size_t)k * n2;
        const auto ru = plane_rsrc(f.u + o, pb), rv = plane_rsrc(f.v + o, pb);
        const auto ra = plane_rsrc(ha + o, pb), rb = plane_rsrc(hb + o, pb);
        const auto rw = plane_rsrc(f.ws + (k + 1 < km ? o + n2 : o), pb);
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            lu[u] = buf_load(ru, wo[u], 0.0);
            lv[u] = buf_load(rv, wo[u], 0.0);
            la[u] = buf_load(ra, wo[u], 0.0);
            lb[u] = buf_load(rb, wo[u], 0.0);
            lw[u] = buf_load(rw, wo[u], 0.0);
        }
    };
    // the lane's own cell: u, v (for level k+1's vertical advection), up, vp, viv, akmu
    struct Lv {
        double u, v, up, vp, vv, ak;
    };
    auto load = [&](int k, Lv &V) {
        const size_t o = (size_t)k * n2;
        V.u = buf_load(plane_rsrc(f.u + o, pb), ocin, 0.0);
        V.v = buf_load(plane_rsrc(f.v + o, pb), ocin, 0.0);
        V.up = buf_load(plane_rsrc(f.up + o, pb), oc, 0.0);
        V.vp = buf_load(plane_rsrc(f.vp + o, pb), oc, 0.0);
        V.vv = buf_load(plane_rsrc(f.viv + o, pb), oc, 0.0);
        V.ak = buf_load(plane_rsrc(f.akmu + o, pb), ocin, 0.0);
    };
    // this U point and its neighbours in the window
    const int w0 = (ty + 1) * MWX + tx + 1, wW = w0 - 1, wE = w0 + 1, wN = w0 - MWX, wS = w0 + MWX,
              wSW = wS - 1;
    constexpr double p25 = 0.25, p5 = 0.5, aid = 1 - 0.5;   // (1 - aidif)
    Lv cur, nxt{};
    load(0, cur);
    load_win(0);
    // w at the U point (readyc3), level 0
    double wu = 0.0;
    {
        const auto rw = plane_rsrc(f.ws, pb);
        const uint32_t os = in ? (uint32_t)(q + imt) * 8u : kBufOff, ow = in ? (uint32_t)(q - 1) * 8u : kBufOff,
                       osw = in ? (uint32_t)(q + imt - 1) * 8u : kBufOff;
        wu = a0 * (buf_load(rw, ocin, 0.0)) + as * (buf_load(rw, os, 0.0)) + aw * (buf_load(rw, ow, 0.0)) +
             asw * (buf_load(rw, osw, 0.0));
    }
    double uprv = 0, vprv = 0, upprv = 0, vpprv = 0, akprv = 0, wkprv = 0;
    double dlub = 0.0, dlvb = 0.0, wkb = 0.0, ub = 0.0, vb = 0.0;
    for (int k = 0; k < km; ++k) {
        const bool more = k + 1 < km;
        const int bf = k & 1;
        double *U = sWU[bf], *V = sWV[bf], *A = sWA[bf], *B = sWB[bf], *W = sWW[bf];
#pragma unroll
        for (int u = 0; u < 2; ++u)
            if (wv[u]) {
                const int e = tl + 256 * u;
                U[e] = lu[u];
                V[e] = lv[u];
                A[e] = la[u];
                B[e] = lb[u];
                W[e] = lw[u];
            }
        if (more) {
            load(k + 1, nxt);
            load_win(k + 1);
        }
        __syncthreads();
        const size_t c = (size_t)k * n2 + q;
        const double wkc = c0f * sqrt(cur.up * cur.up + cur.vp * cur.vp);   // hdiffu_*'s wka
        double dlu = 0.0, dlv = 0.0;
        double wun = 0.0;
        if (in) {
            const double cu = U[w0], cuw = U[wW], cue = U[wE], cun = U[wN], cus = U[wS];
            const double cv = V[w0], cvw = V[wW], cve = V[wE], cvn = V[wN], cvs = V[wS];
            // readyc3 (readyc.cpp:258-269) and advection_momentum_cu (:694-716)
            if (more) wun = a0 * (W[w0]) + as * (W[wS]) + aw * (W[wW]) + asw * (W[wSW]);
            const double uwi = (cuw + cu) * p25 * huew, uwe = (cu + cue) * p25 * huec;
            const double vsj = (cv + cvs) * p25 * huns, vsn = (cvn + cv) * p25 * hunc;
            dlu = (-uwi * (cu - cuw) - uwe * (cue - cu) - vsj * (cus - cu) - vsn * (cu - cun)) * ua;
            dlv = (-uwi * (cv - cvw) - uwe * (cve - cv) - vsj * (cvs - cv) - vsn * (cv - cvn)) * ua;
            double z1 = 0.0e0, z2 = 0.0e0, z3 = 0.0e0, z4 = 0.0e0;
            if (k > 0) {
                z1 = wu * (uprv - cu);
                z3 = wu * (vprv - cv);
            }
            if (k < km - 1) {
                z2 = wun * (cu - nxt.u);
                z4 = wun * (cv - nxt.v);
            }
            dlu += -p5 * sOZP[k] * (z1 + z2);
            dlv += -p5 * sOZP[k] * (z3 + z4);
            // hdiffu_del2_cu (:285-335) / hdiffu_del4_1 (:521-559)
            double hu = 0.0, hv = 0.0;
            if (hreg) {
                const double Ha = DEL4 ? A[w0] : cur.up, Hb = DEL4 ? B[w0] : cur.vp;
                const double an = A[wN], as_ = A[wS], ae = A[wE], aw_ = A[wW];
                const double bn = B[wN], bs = B[wS], be = B[wE], bw = B[wW];
                if (DEL4) {
                    hu = am * (lap_du(L, Ha, an, as_, ae, aw_) + lap_dm(L, Hb, bn, bs, be, bw));
                    hv = am * (lap_du(L, Hb, bn, bs, be, bw) - lap_dm(L, Ha, an, as_, ae, aw_));
                } else {
                    hu = am * (lap_du(L, Ha, an, as_, ae, aw_) + lap_dm(L, Hb, bn, bs, be, bw)) * cur.vv;
                    hv = am * (lap_du(L, Hb, bn, bs, be, bw) - lap_dm(L, Ha, an, as_, ae, aw_)) * cur.vv;
                }
            }
            if (k > kmu - 1) hu = hv = 0.0;
            dlv = dlv + hv;
            dlu = dlu + hu;
            uprv = cu;
            vprv = cv;
        }
        // readyc4's depth sums (:730-743) see the tendencies before readyc4_2
        dlub += sDZ[k] * ohbu * dlu * cur.vv;
        dlvb += sDZ[k] * ohbu * dlv * cur.vv;
        if (k == kmu - 1) {
            wkb = wkc;
            ub = cur.up;
            vb = cur.vp;
        }
        if (in) {
            // readyc4_2 (:803-874)
            double du1, du2, dv1, dv2;
            if (k == 0) {
                dv1 = sv * m.od0 * aid;
                du1 = su * m.od0 * aid;
            } else {
                dv1 = akprv * aid * (vpprv - cur.vp) * sOZT[k] * cur.vv +
                      (1.0e0 - cur.vv) * wkprv * aid * (-sn * upprv * sag + vpprv * cag);
                du1 = akprv * aid * (upprv - cur.up) * sOZT[k] * cur.vv +
                      (1.0e0 - cur.vv) * wkprv * aid * (upprv * cag + sn * vpprv * sag);
            }
            if (k == km - 1) {
                dv2 = wkc * (-sn * cur.up * sag + cur.vp * cag) * aid;
                du2 = wkc * (cur.up * cag + sn * cur.vp * sag) * aid;
            } else {
                dv2 = cur.ak * aid * (cur.vp - nxt.vp) * sOZT[k + 1] * nxt.vv +
                      (1.0e0 - nxt.vv) * wkc * aid * (-sn * cur.up * sag + cur.vp * cag);
                du2 = cur.ak * aid * (cur.up - nxt.up) * sOZT[k + 1] * nxt.vv +
                      (1.0e0 - nxt.vv) * wkc * aid * (cur.up * cag + sn * cur.vp * sag);
            }
            dlv = dlv + sOZP[k] * (dv1 - dv2);
            dlu = dlu + sOZP[k] * (du1 - du2);
        }
        if (act) {
            f.dlu[c] = dlu;
            f.dlv[c] = dlv;
            f.wka[c] = wkc;
        }
        upprv = cur.up;
        vpprv = cur.vp;
        akprv = cur.ak;
        wkprv = wkc;
        wu = wun;
        cur = nxt;
    }
    if (!act) return;
    // readyc4 (:745-785): surface stress, bottom drag at the deepest wet level
    if (in) {
        double sbcx = 0.0e0, sbcy = 0.0e0, bbcx = 0.0e0, bbcy = 0.0e0;
        if (kmu - 1 >= 0) {
            sbcx = su * m.od0;
            sbcy = sv * m.od0;
            bbcx = wkb * (ub * cag + sn * vb * sag);
            bbcy = wkb * (-sn * ub * sag + vb * cag);
        }
        f.sbcx[q] = sbcx;
        f.sbcy[q] = sbcy;
        f.bbcx[q] = bbcx;
        f.bbcy[q] = bbcy;
        dlub = dlub + (sbcx - bbcx) * ohbu;
        dlvb = dlvb + (sbcy - bbcy) * ohbu;
    }
    f.dlub[q] = dlub;
    f.dlvb[q] = dlvb;
}

}  // namespace
}  // namespace licom

using namespace licom;

struct licom_readyc {
    CGrid g{};
    hipStream_t stream = nullptr;
    void *mem = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool ran = false;
    MGrid mg{};             // momentum grid (licom_readyc_momentum_grid)
    void *mmem = nullptr, *mscratch = nullptr;
    bool mom = false;
};

extern "C" {

int licom_readyc_create(const licom_readyc_grid *grid, licom_readyc **out) {
    if (!grid || !out) {
        set_error("licom_readyc_create: NULL argument");
        return LICOM_EINVAL;
    }
    *out = nullptr;
    const licom_readyc_grid &G = *grid;
    if (G.imt < 3 || G.jmt < 3 || G.km < 2 || G.km > LICOM_MAX_KM) {
        set_error("licom_readyc_create: invalid extents imt=%d jmt=%d km=%d (need imt, jmt >= 3, "
                  "2 <= km <= %d)", G.imt, G.jmt, G.km, LICOM_MAX_KM);
        return LICOM_EINVAL;
    }
    const double *src2[8] = {G.at0, G.atn, G.ate, G.atne, G.au0, G.aus, G.auw, G.ausw};
    for (const double *p : src2)
        if (!p) {
            set_error("licom_readyc_create: a grid array is NULL");
            return LICOM_EINVAL;
        }
    if (!G.odzt) {
        set_error("licom_readyc_create: a grid array is NULL");
        return LICOM_EINVAL;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        set_error("no HIP device available");
        return LICOM_ENODEV;
    }
    const size_t n2 = (size_t)G.imt * G.jmt, km = (size_t)G.km;
    const size_t bytes = (8 * n2 + km) * sizeof(double);
    licom_readyc *r = new (std::nothrow) licom_readyc();
    if (!r) return LICOM_ENOMEM;
    int rc = hip_status(hipMalloc(&r->mem, bytes), "hipMalloc");
    std::vector<double> h(8 * n2 + km);
    for (int a = 0; a < 8; ++a) std::copy(src2[a], src2[a] + n2, h.begin() + a * n2);
    std::copy(G.odzt, G.odzt + km, h.begin() + 8 * n2);
    if (!rc) rc = hip_status(hipMemcpy(r->mem, h.data(), bytes, hipMemcpyHostToDevice), "hipMemcpy");
    if (!rc) rc = hip_status(hipEventCreate(&r->ev0), "hipEventCreate");
    if (!rc) rc = hip_status(hipEventCreate(&r->ev1), "hipEventCreate");
    if (rc) {
        licom_readyc_destroy(r);
        return rc;
    }
    const double *d = reinterpret_cast<const double *>(r->mem);
    CGrid &g = r->g;
    g.imt = G.imt;
    g.jmt = G.jmt;
    g.km = G.km;
    g.n2 = n2;
    g.n3 = n2 * km;
    g.at0 = d;
    g.atn = d + n2;
    g.ate = d + 2 * n2;
    g.atne = d + 3 * n2;
    g.au0 = d + 4 * n2;
    g.aus = d + 5 * n2;
    g.auw = d + 6 * n2;
    g.ausw = d + 7 * n2;
    g.odzt = d + 8 * n2;
    *out = r;
    return LICOM_OK;
}

int licom_readyc_set_stream(licom_readyc *r, void *hip_stream) {
    if (!r) {
        set_error("licom_readyc_set_stream: NULL context");
        return LICOM_EINVAL;
    }
    r->stream = (hipStream_t)hip_stream;
    return LICOM_OK;
}

int licom_readyc_shear(licom_readyc *r, const licom_readyc_fields *f) {
    if (!r || !f) {
        set_error("licom_readyc_shear: NULL argument");
        return LICOM_EINVAL;
    }
    const void *need[] = {f->vit, f->viv, f->up,   f->vp,   f->h0,  f->rict, f->ric, f->h0bl, f->h0bf,
                          f->wp12, f->wp13, f->riv1, f->riv2, f->s2t, f->rit, f->s2u, f->riu};
    for (const void *p : need)
        if (!p) {
            set_error("licom_readyc_shear: a field pointer is NULL");
            return LICOM_EINVAL;
        }
    const CGrid &g = r->g;
    const dim3 blk(64, 4), grd((g.imt + 63) / 64, (g.jmt + 3) / 4);
    if (int rc = hip_status(hipEventRecord(r->ev0, r->stream), "hipEventRecord")) return rc;
    k_readyc_shear<<<grd, blk, 0, r->stream>>>(g, *f);
    if (int rc = hip_status(hipGetLastError(), "k_readyc_shear launch")) return rc;
    if (int rc = hip_status(hipEventRecord(r->ev1, r->stream), "hipEventRecord")) return rc;
    r->ran = true;
    return LICOM_OK;
}

int licom_readyc_akmu(licom_readyc *r, const double *akmt, const double *viv, double *akmu) {
    if (!r || !akmt || !viv || !akmu) {
        set_error("licom_readyc_akmu: NULL argument");
        return LICOM_EINVAL;
    }
    const CGrid &g = r->g;
    k_readyc_akmu<<<(unsigned)((g.n3 + 255) / 256), 256, 0, r->stream>>>(g, akmt, viv, akmu);
    return hip_status(hipGetLastError(), "k_readyc_akmu launch");
}

int licom_readyc_momentum_grid(licom_readyc *r, const licom_readyc_mgrid *mg) {
    if (!r || !mg) {
        set_error("licom_readyc_momentum_grid: NULL argument");
        return LICOM_EINVAL;
    }
    const licom_readyc_mgrid &G = *mg;
    const CGrid &g = r->g;
    const int imt = g.imt, jmt = g.jmt, km = g.km;
    // the reference's stencils stay inside the block for these bounds
    const int lo = G.del4 ? 3 : 2, hij = G.del4 ? jmt - 2 : jmt - 1, hii = G.del4 ? imt - 2 : imt - 1;
    if (G.jb < lo || G.je > hij || G.ib < lo || G.ie > hii) {
        set_error("licom_readyc_momentum_grid: block bounds ib=%d ie=%d jb=%d je=%d outside %d..%d x %d..%d "
                  "(%s)", G.ib, G.ie, G.jb, G.je, lo, hii, lo, hij, G.del4 ? "del4" : "del2");
        return LICOM_EINVAL;
    }
    const double *src[] = {G.ohbu, G.ohbt, G.hts, G.htw, G.hue, G.hun, G.tarea_r, G.uarea_r, G.dxu,
                           G.dyu,  G.dxur, G.dyur, G.snlat, G.amf, G.duc, G.dum, G.dun, G.dus,
                           G.due,  G.duw, G.dmc, G.dmn, G.dms, G.dme, G.dmw};
    constexpr int n2d = sizeof(src) / sizeof(src[0]);
    for (const double *p : src)
        if (!p) {
            set_error("licom_readyc_momentum_grid: a grid array is NULL");
            return LICOM_EINVAL;
        }
    if (!G.kmt || !G.kmu || !G.dzp || !G.odzp) {
        set_error("licom_readyc_momentum_grid: a grid array is NULL");
        return LICOM_EINVAL;
    }
    const size_t n2 = g.n2;
    for (size_t p = 0; p < n2; ++p)
        if (G.kmt[p] < 0 || G.kmt[p] > km || G.kmu[p] < 0 || G.kmu[p] > km) {
            set_error("licom_readyc_momentum_grid: kmt or kmu outside 0..%d at cell %zu", km, p);
            return LICOM_EINVAL;
        }
    if (r->mmem) (void)hipFree(r->mmem);
    if (r->mscratch) (void)hipFree(r->mscratch);
    r->mmem = r->mscratch = nullptr;
    r->mom = false;
    // doubles: the 2-D arrays, dzp, odzp; then the two int arrays
    const size_t nd = n2d * n2 + 2 * (size_t)km, bytes = nd * sizeof(double) + 2 * n2 * sizeof(int);
    std::vector<double> h(nd);
    for (int a = 0; a < n2d; ++a) std::copy(src[a], src[a] + n2, h.begin() + a * n2);
    std::copy(G.dzp, G.dzp + km, h.begin() + n2d * n2);
    std::copy(G.odzp, G.odzp + km, h.begin() + n2d * n2 + km);
    int rc = hip_status(hipMalloc(&r->mmem, bytes), "hipMalloc");
    char *base = static_cast<char *>(r->mmem);
    if (!rc) rc = hip_status(hipMemcpy(base, h.data(), nd * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
    int *ik = reinterpret_cast<int *>(base + nd * sizeof(double));
    if (!rc) rc = hip_status(hipMemcpy(ik, G.kmt, n2 * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy");
    if (!rc) rc = hip_status(hipMemcpy(ik + n2, G.kmu, n2 * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy");
    if (!rc && G.del4) rc = hip_status(hipMalloc(&r->mscratch, 2 * g.n3 * sizeof(double)), "hipMalloc");
    if (rc) {
        if (r->mmem) (void)hipFree(r->mmem);
        if (r->mscratch) (void)hipFree(r->mscratch);
        r->mmem = r->mscratch = nullptr;
        return rc;
    }
    const double *d = reinterpret_cast<const double *>(base);
    MGrid &m = r->mg;
    m.ib = G.ib;
    m.ie = G.ie;
    m.jb = G.jb;
    m.je = G.je;
    m.del4 = G.del4 ? 1 : 0;
    m.am = G.am;
    m.c0f = G.c0f;
    m.od0 = G.od0;
    m.cag = G.cag;
    m.sag = G.sag;
    const double **dst[] = {&m.ohbu, &m.ohbt, &m.hts, &m.htw, &m.hue, &m.hun, &m.tarea_r, &m.uarea_r, &m.dxu,
                            &m.dyu,  &m.dxur, &m.dyur, &m.snlat, &m.amf, &m.duc, &m.dum, &m.dun, &m.dus,
                            &m.due,  &m.duw, &m.dmc, &m.dmn, &m.dms, &m.dme, &m.dmw};
    for (int a = 0; a < n2d; ++a) *dst[a] = d + a * n2;
    m.dzp = d + n2d * n2;
    m.odzp = d + n2d * n2 + km;
    m.kmt = ik;
    m.kmu = ik + n2;
    double *sc = static_cast<double *>(r->mscratch);
    m.d2u = sc;
    m.d2v = sc ? sc + g.n3 : nullptr;
    r->mom = true;
    return LICOM_OK;
}

int licom_readyc_momentum(licom_readyc *r, const licom_readyc_mfields *f) {
    if (!r || !f) {
        set_error("licom_readyc_momentum: NULL argument");
        return LICOM_EINVAL;
    }
    if (!r->mom) {
        set_error("licom_readyc_momentum: no momentum grid (call licom_readyc_momentum_grid first)");
        return LICOM_EINVAL;
    }
    const void *need[] = {f->u,   f->v,   f->up,   f->vp,   f->h0,   f->su,   f->sv,   f->vit,  f->viv,
                          f->akmu, f->ws, f->dlu, f->dlv, f->wka, f->dlub, f->dlvb, f->sbcx, f->sbcy,
                          f->bbcx, f->bbcy};
    for (const void *p : need)
        if (!p) {
            set_error("licom_readyc_momentum: a field pointer is NULL");
            return LICOM_EINVAL;
        }
    const CGrid &g = r->g;
    const MGrid &m = r->mg;
    const dim3 blk(64, 4), grd((g.imt + 63) / 64, (g.jmt + 3) / 4);
    if (int rc = hip_status(hipEventRecord(r->ev0, r->stream), "hipEventRecord")) return rc;
    k_rmom_upwell<<<grd, blk, 0, r->stream>>>(g, m, *f);
    if (int rc = hip_status(hipGetLastError(), "k_rmom_upwell launch")) return rc;
    if (m.del4) {
        k_rmom_lap1<<<grd, blk, 0, r->stream>>>(g, m, *f);
        if (int rc = hip_status(hipGetLastError(), "k_rmom_lap1 launch")) return rc;
        k_rmom_main<true><<<grd, blk, 0, r->stream>>>(g, m, *f);
    } else {
        k_rmom_main<false><<<grd, blk, 0, r->stream>>>(g, m, *f);
    }
    if (int rc = hip_status(hipGetLastError(), "k_rmom_main launch")) return rc;
    if (int rc = hip_status(hipEventRecord(r->ev1, r->stream), "hipEventRecord")) return rc;
    r->ran = true;
    return LICOM_OK;
}

int licom_readyc_sync(licom_readyc *r) {
    if (!r) {
        set_error("licom_readyc_sync: NULL context");
        return LICOM_EINVAL;
    }
    return hip_status(hipStreamSynchronize(r->stream), "hipStreamSynchronize");
}

int licom_readyc_last_ms(licom_readyc *r, double *ms) {
    if (!r || !ms) {
        set_error("licom_readyc_last_ms: NULL argument");
        return LICOM_EINVAL;
    }
    *ms = 0.0;
    if (!r->ran) return LICOM_OK;
    if (int rc = hip_status(hipEventSynchronize(r->ev1), "hipEventSynchronize")) return rc;
    float v = 0.0f;
    if (int rc = hip_status(hipEventElapsedTime(&v, r->ev0, r->ev1), "hipEventElapsedTime")) return rc;
    *ms = v;
    return LICOM_OK;
}

void licom_readyc_destroy(licom_readyc *r) {
    if (!r) return;
    if (r->ran) (void)hipStreamSynchronize(r->stream);
    if (r->mem) (void)hipFree(r->mem);
    if (r->mmem) (void)hipFree(r->mmem);
    if (r->mscratch) (void)hipFree(r->mscratch);
    if (r->ev0) (void)hipEventDestroy(r->ev0);
    if (r->ev1) (void)hipEventDestroy(r->ev1);
    delete r;
}

}  // extern "C"
