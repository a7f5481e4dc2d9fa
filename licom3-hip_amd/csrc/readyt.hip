// readyt.hip -- readyt_ (readyt.cpp:511-598) on gfx950: the thermodynamic
// preparation of a LICOM step, behind licom_readyt_* (include/licom_readyt.h).
//
// The reference runs nine launches and two memsets over whole arrays, with
// dlu/dlv first holding T and S on the U grid and later the pressure
// gradient.  Here three passes, one lane per water column (i-consecutive
// lanes, so every level's loads and stores are coalesced row pieces):
//
//   k_readyt_col  marches k top to bottom: the time-level shift of h0, utf,
//                 vtf (readyt_cu_1), potential density and its weight
//                 (readyt_cu_2), the hydrostatic integrals pp / ppa as
//                 running sums in registers (readyt_cu_2_2), T and S at the
//                 top interface, alpha / beta (thermal_cu), the surface
//                 buoyancy terms and the Richardson-number terms of the
//                 interface above each level (dens_cu, readyt_cu_3; the U-grid
//                 T and S of dens_cu are formed per level from atb and never
//                 stored), the anomaly weight gg, the clears of akt, rit, and
//                 the column integrals of gg that end up in dlu planes 0, 1
//                 (readyt_cu_3_2);
//   k_readyt_pg   marches k again: the baroclinic pressure gradient from the
//                 four pp columns around each U point (readyt_cu_3), its
//                 vertical integrals pxb / pyb, and the barotropic terms of
//                 dlv planes 0, 1 (readyt_cu_4);
//   k_readyt_sfc  the 2-D surface terms of readyt_cu_5.
//
// Every expression keeps the reference's operation order (the library is
// compiled with -ffp-contract=off), so results equal the reference's
// kernels bit for bit.
#include <hip/hip_runtime.h>

#include <new>
#include <vector>

#include "licom_readyt.h"
#include "tracer_internal.h"

namespace licom {
namespace {

constexpr double kGrav = 9.806;   // constant_mod.h:27
constexpr int RK = LICOM_MAX_KM;

struct RGrid {
    int imt, jmt, km;
    size_t n2, n3;
    const int32_t *kmu;
    const double *au0, *aus, *auw, *ausw, *dxur, *dyur, *ohbu, *ohbt, *hbx, *hby;
    const double *dzp, *odzt, *zkt, *to, *so, *po, *eosc;
    double od0, od0cp;
};

// dens_c_cu (readyt.cpp:25-35): the 9-term polynomial of level k, C(k,m) at
// c[(m-1)*km + k]
__device__ __forceinline__ double dens(const double *c, int km, int k, double t, double s) {
    return (c[k] + (c[3 * km + k] + c[6 * km + k] * s) * s +
            (c[2 * km + k] + c[7 * km + k] * s + c[5 * km + k] * t) * t) * t +
           (c[1 * km + k] + (c[4 * km + k] + c[8 * km + k] * s) * s) * s;
}

// thermal_cu (readyt.cpp:272-294): McDougall (1987) thermal expansion and
// saline contraction from T, S (model units) and the pressure integral ppa
__device__ __forceinline__ void thermal(double tt1, double ppc, double ppa, double vit, double od0,
                                        double &alpha, double &beta) {
    const double tt2 = tt1 * tt1, tt3 = tt2 * tt1;
    const double ss1 = ppc * 1000.0e0, ss2 = ss1 * ss1;
    const double tm1 = -ppa / od0 / 10000.e0 * vit, tm2 = tm1 * tm1, tm3 = tm2 * tm1;
    beta = (0.785567e-3 - 0.301985e-5 * tt1 + 0.555579e-7 * tt2 - 0.415613e-9 * tt3 +
            ss1 * (-0.356603e-6 + 0.788212e-8 * tt1 + 0.408195e-10 * tm1 - 0.602281e-15 * tm2) +
            ss2 * (0.515032e-8) + tm1 * (-0.121555e-7 + 0.192867e-9 * tt1 - 0.2131127e-11 * tt2) +
            tm2 * (0.176621e-12 - 0.175379e-14 * tt1) + tm3 * (0.12155e-17)) *
           vit;
    alpha = (0.665157e-1 + 0.170907e-1 * tt1 - 0.203814e-3 * tt2 + 0.298357e-5 * tt3 -
             0.255019e-7 * tt3 * tt1 +
             ss1 * (0.378110e-2 - 0.846960e-4 * tt1 - 0.164759e-6 * tm1 - 0.251520e-11 * tm2) +
             ss2 * (-0.678662e-5) + tm1 * (0.380374e-4 - 0.933746e-6 * tt1 + 0.791325e-8 * tt2) +
             0.512857e-12 * tm2 * tt2 - 0.302285e-13 * tm3) *
            beta * vit;
}

// Per-level constants in LDS: to, so, po, dzp, odzt, zkt and the density
// coefficients (a global load inside the level loop would carry a vmcnt wait
// that also covers the level's stores).
struct RConst {
    double to[RK], so[RK], po[RK], dzp[RK], odzt[RK], zkt[RK], c[9 * RK];
};

__device__ __forceinline__ void load_const(const RGrid &g, RConst &s) {
    const int t = threadIdx.x + blockDim.x * threadIdx.y;
    const int nt = blockDim.x * blockDim.y;
    for (int r = t; r < g.km; r += nt) {
        s.to[r] = g.to[r];
        s.so[r] = g.so[r];
        s.po[r] = g.po[r];
        s.dzp[r] = g.dzp[r];
        s.odzt[r] = g.odzt[r];
        s.zkt[r] = g.zkt[r];
    }
    for (int r = t; r < 9 * g.km; r += nt) s.c[r] = g.eosc[r];
    __syncthreads();
}

__global__ __launch_bounds__(256) void k_readyt_col(RGrid g, licom_readyt_fields f) {
    __shared__ RConst S;
    load_const(g, S);
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
    const int imt = g.imt, jmt = g.jmt, km = g.km;
    if (i >= imt || j >= jmt) return;
    const size_t n2 = g.n2, n3 = g.n3, q = (size_t)j * imt + i;
    const double od0 = g.od0;
    const bool au_ok = j < jmt - 1 && i >= 1;        // U-grid cells (readyt_cu_1, dens_cu)
    const bool th_ok = i >= 1 && i < imt - 1;        // thermal_cu, ricdttms, ricdt
    double a0 = 0.0, as = 0.0, aw = 0.0, asw = 0.0;
    if (au_ok) {
        a0 = g.au0[q];
        as = g.aus[q];
        aw = g.auw[q];
        asw = g.ausw[q];
    }
    const double ohbt = g.ohbt[q], psa = f.psa[q];
    const double *atT = f.at, *atS = f.at + n3;
    const double *abT = f.atb + n2, *abS = f.atb + n3 + 2 * n2;   // level k at plane k+1
    {
        const double h = f.h0f[q];
        f.h0l[q] = h;
        f.h0f[q] = f.h0[q];
    }
    const double *C = S.c;
    // Each level's inputs are loaded through plane buffers one level ahead,
    // before the current level's stores: on gfx9 one counter (vmcnt) covers
    // loads and stores in issue order, so a load issued after the stores
    // would make the next level wait for all of them.  Masked neighbours (no
    // U point) read 0 from an out-of-range offset, without a branch.
    const size_t pbytes = n2 * sizeof(double);
    const uint32_t oq = (uint32_t)q * 8u;
    const uint32_t os = au_ok ? (uint32_t)(q + imt) * 8u : kBufOff;
    const uint32_t ow = au_ok ? (uint32_t)(q - 1) * 8u : kBufOff;
    const uint32_t osw = au_ok ? (uint32_t)(q + imt - 1) * 8u : kBufOff;
    struct LvIn {
        double uf, vf, un, vn, T, S, vt, vv, bT, bS, bTs, bTw, bTsw, bSs, bSw, bSsw;
    };
    auto load_level = [&](int k, LvIn &L) {
        const size_t o = (size_t)k * n2;
        L.uf = buf_load(plane_rsrc(f.utf + o, pbytes), oq, 0.0);
        L.vf = buf_load(plane_rsrc(f.vtf + o, pbytes), oq, 0.0);
        L.un = buf_load(plane_rsrc(f.u + o, pbytes), oq, 0.0);
        L.vn = buf_load(plane_rsrc(f.v + o, pbytes), oq, 0.0);
        L.T = buf_load(plane_rsrc(atT + o, pbytes), oq, 0.0);
        L.S = buf_load(plane_rsrc(atS + o, pbytes), oq, 0.0);
        L.vt = buf_load(plane_rsrc(f.vit + o, pbytes), oq, 0.0);
        L.vv = buf_load(plane_rsrc(f.viv + o, pbytes), oq, 0.0);
        const auto rT = plane_rsrc(abT + o, pbytes), rS = plane_rsrc(abS + o, pbytes);
        L.bT = buf_load(rT, oq, 0.0);
        L.bS = buf_load(rS, oq, 0.0);
        L.bTs = buf_load(rT, os, 0.0);
        L.bTw = buf_load(rT, ow, 0.0);
        L.bTsw = buf_load(rT, osw, 0.0);
        L.bSs = buf_load(rS, os, 0.0);
        L.bSw = buf_load(rS, ow, 0.0);
        L.bSsw = buf_load(rS, osw, 0.0);
    };
    double Tp = 0.0, Sp = 0.0, bTp = 0.0, bSp = 0.0, duTp = 0.0, duSp = 0.0;
    double ppp = 0.0, ppap = 0.0, ggp = 0.0;
    double s0 = 0.0, s1 = 0.0;
    LvIn cur, nxt;
    load_level(0, cur);
    for (int k = 0; k < km; ++k) {
        const size_t c = (size_t)k * n2 + q;
        if (k + 1 < km) load_level(k + 1, nxt);
        // readyt_cu_1: time levels
        f.utl[c] = cur.uf;
        f.vtl[c] = cur.vf;
        f.utf[c] = cur.un;
        f.vtf[c] = cur.vn;
        const double T = cur.T, Sl = cur.S, vt = cur.vt, vv = cur.vv;
        const double bT = cur.bT, bS = cur.bS;
        // T and S of atb on the U grid (readyt_cu_1's dlu / dlv, read by dens_cu)
        double duT = 0.0, duS = 0.0;
        if (au_ok) {
            duT = a0 * bT + as * cur.bTs + aw * cur.bTw + asw * cur.bTsw;
            duS = a0 * bS + as * cur.bSs + aw * cur.bSw + asw * cur.bSsw;
        }
        // readyt_cu_2: potential density and its weight in the pressure integral
        double pd = 0.0;
        if (vt > 0.0) pd = 1.0e+3 + S.po[k] + dens(C, km, k, T - S.to[k], Sl - S.so[k]);
        f.pdensity[c] = pd;
        const double gg0 = -od0 * kGrav * pd * vt;
        // T and S at the top interface (readyt_cu_2 for k >= 1, readyt_cu_2_2 at k = 0)
        double pb, pc;
        if (k == 0) {
            pb = T * vt;
            pc = Sl * vt;
        } else {
            const double w = S.dzp[k - 1], sm = S.dzp[k - 1] + S.dzp[k];
            pb = vt * (Tp - (Tp - T) * w / sm);
            pc = vt * (Sp - (Sp - Sl) * w / sm);
        }
        f.ppb[c] = pb;
        f.ppc[c] = pc;
        // readyt_cu_2_2: hydrostatic integrals
        double pp, ppa;
        if (k == 0) {
            pp = 0.5e0 * S.dzp[0] * gg0 * vt;
            ppa = psa * vt;
        } else {
            pp = vt * (ppp + 0.5e0 * (gg0 * S.dzp[k] + ggp * S.dzp[k - 1]));
            ppa = vt * (ppap + ggp * S.dzp[k - 1]);
        }
        f.pp[c] = pp;
        f.ppa[c] = ppa;
        // thermal_cu
        double al = 0.0, be = 0.0;
        if (th_ok) thermal(pb, pc, ppa, vt, od0, al, be);
        f.alpha[c] = al;
        f.beta[c] = be;
        // readyt_cu_3
        if (k == 0) {
            f.buoytur[q] = vt * f.nswv[q] * al * kGrav * g.od0cp;
            f.buoysol[q] = vt * f.swv[q] * al * kGrav * g.od0cp;
        } else {
            // the interface above level k: dens_cu and readyt_cu_3 at index k-1,
            // densities referred to level k
            const size_t cm = c - n2;
            double ric = 0.0, rict = 0.0;
            if (au_ok) {
                double rhoup = dens(C, km, k, duTp - S.to[k], duSp - S.so[k]);
                double rholo = dens(C, km, k, duT - S.to[k], duS - S.so[k]);
                ric = vv * od0 * kGrav * (rholo - rhoup) * S.odzt[k];
                rhoup = dens(C, km, k, bTp - S.to[k], bSp - S.so[k]);
                rholo = dens(C, km, k, bT - S.to[k], bS - S.so[k]);
                rict = vt * od0 * kGrav * (rholo - rhoup) * S.odzt[k];
            }
            f.ric[cm] = ric;
            f.rict[cm] = rict;
            double rt = 0.0, rd = 0.0;
            if (th_ok) {
                const double dT = Tp - T, dS = Sp - Sl;
                constexpr double epsln = 1.0e-25;
                rt = vt * kGrav * (dT * al + 1000.0e0 * dS * be) * S.odzt[k];
                rd = vt / ((Tp - T + epsln) * al) * 1000.0e0 * dS * be;
            }
            f.ricdttms[cm] = rt;
            f.ricdt[cm] = rd;
            f.rit[cm] = 0.0;
        }
        const double gg = -od0 * kGrav * (pd - S.po[k] - 1000.0e0) * vt;
        f.gg[c] = gg;
        // readyt_cu_3_2: column integrals of gg (dlu planes 0 and 1)
        const double abcd = gg * ohbt * S.dzp[k];
        s0 += abcd;
        s1 += abcd * S.zkt[k];
        f.akt[c] = 0.0;
        f.akt[n3 + c] = 0.0;
        Tp = T;
        Sp = Sl;
        bTp = bT;
        bSp = bS;
        duTp = duT;
        duSp = duS;
        ppp = pp;
        ppap = ppa;
        ggp = gg0;
        cur = nxt;
    }
    f.dlu[q] = s0;
    f.dlu[n2 + q] = s1;
}

// readyt_cu_3's pressure gradient at every level, readyt_cu_3_2's vertical
// integrals and readyt_cu_4's barotropic terms.
__global__ __launch_bounds__(256) void k_readyt_pg(RGrid g, licom_readyt_fields f) {
    __shared__ RConst S;
    load_const(g, S);
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
    const int imt = g.imt, jmt = g.jmt, km = g.km;
    if (i >= imt || j >= jmt) return;
    const size_t n2 = g.n2, q = (size_t)j * imt + i;
    const bool pg_ok = i >= 1 && j < jmt - 1;
    int kmu = 0;
    double dxr = 0.0, dyr = 0.0;
    if (pg_ok) {
        kmu = g.kmu[q];
        dxr = g.dxur[q];
        dyr = g.dyur[q];
    }
    const double ohbu = g.ohbu[q], ohbt = g.ohbt[q];
    const double p5 = 0.5;
    double px = 0.0, py = 0.0;
    for (int k = 0; k < km; ++k) {
        const size_t c = (size_t)k * n2 + q;
        double du = 0.0, dv = 0.0;
        if (pg_ok && k <= kmu - 1) {
            const double ps = f.pp[c + imt], pw = f.pp[c - 1], psw = f.pp[c + imt - 1], pc = f.pp[c];
            du = dxr * p5 * (ps - pw - psw + pc);
            dv = dyr * p5 * (ps - pw + psw - pc);
        }
        const double vv = f.viv[c];
        px += S.dzp[k] * ohbu * du * vv;
        py += S.dzp[k] * ohbu * dv * vv;
        if (k >= 2) {
            f.dlu[c] = du;
            f.dlv[c] = dv;
        }
    }
    f.pxb[q] = px;
    f.pyb[q] = py;
    const double s0 = f.dlu[q], s1 = f.dlu[n2 + q];
    f.dlv[q] = (s0 + s1 * ohbt) / kGrav;
    f.dlv[n2 + q] = s1 * ohbt * ohbt;
}

// readyt_cu_5 (readyt.cpp:450-509)
__global__ __launch_bounds__(256) void k_readyt_sfc(RGrid g, licom_readyt_fields f) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
    const int imt = g.imt, jmt = g.jmt;
    if (i >= imt || j >= jmt) return;
    const size_t n2 = g.n2, q = (size_t)j * imt + i;
    const double p5 = 0.5;
    const double *d0 = f.dlv, *d1 = f.dlv + n2;
    double wgp = 0.0, work = 0.0;
    const bool au_ok = j < jmt - 1 && 1 <= i;
    if (au_ok) {
        const double a0 = g.au0[q], as = g.aus[q], aw = g.auw[q], asw = g.ausw[q];
        wgp = a0 * d0[q] + as * d0[q + imt] + aw * d0[q - 1] + asw * d0[q + imt - 1];
        work = a0 * d1[q] + as * d1[q + imt] + aw * d1[q - 1] + asw * d1[q + imt - 1];
    }
    f.wgp[q] = wgp;
    f.work[q] = work;
    const double vv = f.viv[q];
    f.whx[q] = g.hbx[q] * work * vv;
    f.why[q] = g.hby[q] * work * vv;
    double w1 = 0.0, w2 = 0.0;
    if (au_ok && 0 <= g.kmu[q] - 1) {
        const double ps = f.psa[q + imt], pw = f.psa[q - 1], psw = f.psa[q + imt - 1], pc = f.psa[q];
        w1 = g.dxur[q] * p5 * (ps - pw - psw + pc);
        w2 = g.dyur[q] * p5 * (ps - pw + psw - pc);
    }
    f.work1[q] = w1;
    f.work2[q] = w2;
    f.pay[q] = -g.od0 * w2;
    f.pax[q] = -g.od0 * w1;
}

// Clear of a diagnostic array with a kernel (no memset node if the caller
// captures the run in a graph); 16-byte lanes.
__global__ __launch_bounds__(256) void k_fill_zero(double2 *p, size_t n2x) {
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t < n2x) p[t] = make_double2(0.0, 0.0);
}
__global__ __launch_bounds__(256) void k_fill_zero1(double *p, size_t n) {
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t < n) p[t] = 0.0;
}

}  // namespace
}  // namespace licom

using namespace licom;

struct licom_readyt {
    RGrid g{};
    int device = 0;
    hipStream_t stream = nullptr;
    void *mem = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    bool ran = false;
};

extern "C" {

int licom_readyt_create(const licom_readyt_grid *grid, licom_readyt **out) {
    if (!grid || !out) {
        set_error("licom_readyt_create: NULL argument");
        return LICOM_EINVAL;
    }
    *out = nullptr;
    const licom_readyt_grid &G = *grid;
    if (G.imt < 3 || G.jmt < 3 || G.km < 2 || G.km > LICOM_MAX_KM) {
        set_error("licom_readyt_create: invalid extents imt=%d jmt=%d km=%d (need imt, jmt >= 3, "
                  "2 <= km <= %d)", G.imt, G.jmt, G.km, LICOM_MAX_KM);
        return LICOM_EINVAL;
    }
    const void *need[] = {G.kmu, G.au0, G.aus,  G.auw, G.ausw, G.dxur, G.dyur, G.ohbu, G.ohbt, G.hbx,
                          G.hby, G.dzp, G.odzt, G.zkt, G.to,   G.so,   G.po,   G.eosc};
    for (const void *p : need)
        if (!p) {
            set_error("licom_readyt_create: a grid array is NULL");
            return LICOM_EINVAL;
        }
    const size_t n2 = (size_t)G.imt * G.jmt, km = (size_t)G.km;
    for (size_t q = 0; q < n2; ++q)
        if (G.kmu[q] < 0 || G.kmu[q] > G.km) {
            set_error("licom_readyt_create: kmu out of range at %zu: %d (km=%d)", q, G.kmu[q], G.km);
            return LICOM_EINVAL;
        }
    int ndev = 0, dev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        set_error("no HIP device available");
        return LICOM_ENODEV;
    }
    if (int rc = hip_status(hipGetDevice(&dev), "hipGetDevice")) return rc;
    // one allocation: ten 2-D double fields, kmu, seven 1-D fields (eosc 9 km)
    const size_t nd2 = 10 * n2, nd1 = 6 * km + 9 * km;
    const size_t bytes = (nd2 + nd1) * sizeof(double) + n2 * sizeof(int32_t);
    licom_readyt *r = new (std::nothrow) licom_readyt();
    if (!r) return LICOM_ENOMEM;
    r->device = dev;
    int rc = hip_status(hipMalloc(&r->mem, bytes), "hipMalloc");
    std::vector<char> host(bytes);
    double *h = reinterpret_cast<double *>(host.data());
    const double *src2[10] = {G.au0, G.aus, G.auw, G.ausw, G.dxur, G.dyur, G.ohbu, G.ohbt, G.hbx, G.hby};
    for (int a = 0; a < 10; ++a) std::copy(src2[a], src2[a] + n2, h + a * n2);
    const double *src1[6] = {G.dzp, G.odzt, G.zkt, G.to, G.so, G.po};
    for (int a = 0; a < 6; ++a) std::copy(src1[a], src1[a] + km, h + nd2 + a * km);
    std::copy(G.eosc, G.eosc + 9 * km, h + nd2 + 6 * km);
    int32_t *hk = reinterpret_cast<int32_t *>(h + nd2 + nd1);
    std::copy(G.kmu, G.kmu + n2, hk);
    if (!rc) rc = hip_status(hipMemcpy(r->mem, host.data(), bytes, hipMemcpyHostToDevice), "hipMemcpy");
    for (int e = 0; e < 4 && !rc; ++e) rc = hip_status(hipEventCreate(&r->ev[e]), "hipEventCreate");
    if (rc) {
        licom_readyt_destroy(r);
        return rc;
    }
    const double *d = reinterpret_cast<const double *>(r->mem);
    RGrid &g = r->g;
    g.imt = G.imt;
    g.jmt = G.jmt;
    g.km = G.km;
    g.n2 = n2;
    g.n3 = n2 * km;
    g.au0 = d;
    g.aus = d + n2;
    g.auw = d + 2 * n2;
    g.ausw = d + 3 * n2;
    g.dxur = d + 4 * n2;
    g.dyur = d + 5 * n2;
    g.ohbu = d + 6 * n2;
    g.ohbt = d + 7 * n2;
    g.hbx = d + 8 * n2;
    g.hby = d + 9 * n2;
    g.dzp = d + nd2;
    g.odzt = d + nd2 + km;
    g.zkt = d + nd2 + 2 * km;
    g.to = d + nd2 + 3 * km;
    g.so = d + nd2 + 4 * km;
    g.po = d + nd2 + 5 * km;
    g.eosc = d + nd2 + 6 * km;
    g.kmu = reinterpret_cast<const int32_t *>(d + nd2 + nd1);
    g.od0 = G.od0;
    g.od0cp = G.od0cp;
    *out = r;
    return LICOM_OK;
}

int licom_readyt_set_stream(licom_readyt *r, void *hip_stream) {
    if (!r) {
        set_error("licom_readyt_set_stream: NULL context");
        return LICOM_EINVAL;
    }
    r->stream = (hipStream_t)hip_stream;
    return LICOM_OK;
}

int licom_readyt_run(licom_readyt *r, const licom_readyt_fields *f, int ist) {
    if (!r || !f) {
        set_error("licom_readyt_run: NULL argument");
        return LICOM_EINVAL;
    }
    const void *need[] = {f->vit,   f->viv,   f->at,    f->atb,      f->h0,    f->u,       f->v,
                          f->psa,   f->swv,   f->nswv,  f->h0l,      f->h0f,   f->utl,     f->vtl,
                          f->utf,   f->vtf,   f->dlu,   f->dlv,      f->pdensity, f->gg,   f->pp,
                          f->ppa,   f->ppb,   f->ppc,   f->alpha,    f->beta,  f->ric,     f->rict,
                          f->ricdttms, f->ricdt, f->rit, f->akt,     f->buoytur, f->buoysol, f->pxb,
                          f->pyb,   f->wgp,   f->work,  f->whx,      f->why,   f->work1,   f->work2,
                          f->pax,   f->pay};
    for (const void *p : need)
        if (!p) {
            set_error("licom_readyt_run: a field pointer is NULL (only ax, ay, az may be)");
            return LICOM_EINVAL;
        }
    const RGrid &g = r->g;
    hipStream_t st = r->stream;
    const dim3 blk(64, 4), grd((g.imt + 63) / 64, (g.jmt + 3) / 4);
    if (int rc = hip_status(hipEventRecord(r->ev[0], st), "hipEventRecord")) return rc;
    k_readyt_col<<<grd, blk, 0, st>>>(g, *f);
    if (int rc = hip_status(hipGetLastError(), "k_readyt_col launch")) return rc;
    if (int rc = hip_status(hipEventRecord(r->ev[1], st), "hipEventRecord")) return rc;
    k_readyt_pg<<<grd, blk, 0, st>>>(g, *f);
    if (int rc = hip_status(hipGetLastError(), "k_readyt_pg launch")) return rc;
    if (int rc = hip_status(hipEventRecord(r->ev[2], st), "hipEventRecord")) return rc;
    k_readyt_sfc<<<grd, blk, 0, st>>>(g, *f);
    if (int rc = hip_status(hipGetLastError(), "k_readyt_sfc launch")) return rc;
    if (ist == 0) {
        // readyt.cpp:591-595: the diagnostic period starts at zero
        const size_t n = 2 * g.n3;
        for (double *p : {f->ax, f->ay, f->az}) {
            if (!p) continue;
            if ((reinterpret_cast<uintptr_t>(p) & 15) == 0 && n % 2 == 0)
                k_fill_zero<<<(unsigned)((n / 2 + 255) / 256), 256, 0, st>>>(reinterpret_cast<double2 *>(p), n / 2);
            else
                k_fill_zero1<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(p, n);
            if (int rc = hip_status(hipGetLastError(), "k_fill_zero launch")) return rc;
        }
    }
    if (int rc = hip_status(hipEventRecord(r->ev[3], st), "hipEventRecord")) return rc;
    r->ran = true;
    return LICOM_OK;
}

int licom_readyt_sync(licom_readyt *r) {
    if (!r) {
        set_error("licom_readyt_sync: NULL context");
        return LICOM_EINVAL;
    }
    return hip_status(hipStreamSynchronize(r->stream), "hipStreamSynchronize");
}

int licom_readyt_last_ms(licom_readyt *r, double ms[4]) {
    if (!r || !ms) {
        set_error("licom_readyt_last_ms: NULL argument");
        return LICOM_EINVAL;
    }
    for (int e = 0; e < 4; ++e) ms[e] = 0.0;
    if (!r->ran) return LICOM_OK;
    if (int rc = hip_status(hipEventSynchronize(r->ev[3]), "hipEventSynchronize")) return rc;
    float v = 0.0f;
    if (int rc = hip_status(hipEventElapsedTime(&v, r->ev[0], r->ev[3]), "hipEventElapsedTime")) return rc;
    ms[0] = v;
    for (int e = 0; e < 3; ++e) {
        if (int rc = hip_status(hipEventElapsedTime(&v, r->ev[e], r->ev[e + 1]), "hipEventElapsedTime"))
            return rc;
        ms[e + 1] = v;
    }
    return LICOM_OK;
}

void licom_readyt_destroy(licom_readyt *r) {
    if (!r) return;
    if (r->ran) (void)hipStreamSynchronize(r->stream);
    if (r->mem) (void)hipFree(r->mem);
    for (hipEvent_t e : r->ev)
        if (e) (void)hipEventDestroy(e);
    delete r;
}

}  // extern "C"
