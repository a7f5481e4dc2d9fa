// tracer_fused.hip -- production path of the FP64 tracer step on gfx950.
//
// Three HBM passes instead of the reference's ~21 launches per tracer:
//
//   k_prelude<KM>  one lane per column: face velocities u_wface/v_sface
//                  (advection_tracer_1_0) and the barotropic-corrected
//                  continuity w (upwell_2..4, operators_div); the column's
//                  divergence profile stays in registers between the
//                  column sum and the w recursion.
//   k_march        one workgroup per 32x16 tile of columns and tracer,
//                  marching k top to bottom.  Planes of at/atb (tile plus a
//                  2-cell halo) are staged in LDS; the tspas predictor of
//                  level k+1 (tile plus a 1-cell ring) feeds the limiter of
//                  level k through LDS flag planes; the del4 first pass is
//                  kept in LDS.  Writes the explicit update (tracer_8's vtl)
//                  and the advective / vertical-diffusion diagnostics.
//   k_solve<KM>    two lanes per column (T and S): the implicit vertical
//                  diffusion (invtrit) with checkpointed elimination
//                  coefficients in registers, tracer_9's tendencies, and a
//                  streaming stability check of convadj; statically unstable
//                  columns are then mixed in place exactly as convadj_cu does.
//
// Arithmetic is shared with the per-stage kernels through tracer_device.h,
// so both paths give identical bits.
#include <hip/hip_runtime.h>

#include "tracer_device.h"
#include "tracer_internal.h"

namespace licom {
namespace {

constexpr int KMAXF = LICOM_MAX_KM;

// ------------------------------------------------------------- prelude
template <int KM>
__global__ __launch_bounds__(256) void k_prelude(KGrid g, KPar p, const double *__restrict__ utf,
                                                 const double *__restrict__ utl,
                                                 const double *__restrict__ vtf,
                                                 const double *__restrict__ vtl,
                                                 const double *__restrict__ work_u,
                                                 const double *__restrict__ stf_pre,
                                                 double *__restrict__ ws, double *__restrict__ uw,
                                                 double *__restrict__ vs) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y;
    const int imt = g.imt, jmt = g.jmt, km = g.km;
    if (i >= imt || j >= jmt - 1) return;
    const size_t n2 = (size_t)imt * jmt, q = (size_t)j * imt + i;
    const bool do_u = j >= 1 && i >= 1;                // u_wface: j in [1,jmt-1), all i
    const bool do_v = i >= 1 && i < imt - 1;           // v_sface: i in [1,imt-1)
    const bool do_w = j >= 1 && i >= 1 && i < imt - 1; // w: [1,jmt-1) x [1,imt-1)
    const double oncc = p.oncc;
    const size_t qe = q + 1, qn = q - imt, qne = q - imt + 1;
    double fc = 0, fe = 0, fn = 0, fne = 0, htw_e = 0, hts_n = 0, tr = 0;
    const double htw_c = g.htw[q], hts_c = g.hts[q];
    int kmt = 0;
    if (do_w) {
        fc = 1.0 + work_u[q] * g.ohbu[q];
        fe = 1.0 + work_u[qe] * g.ohbu[qe];
        fn = 1.0 + work_u[qn] * g.ohbu[qn];
        fne = 1.0 + work_u[qne] * g.ohbu[qne];
        htw_e = g.htw[qe];
        hts_n = g.hts[qn];
        tr = g.tarea_r[q];
        kmt = g.kmt[q];
    }
    double wka[KM];
    double w = 0.0;
#pragma unroll
    for (int k = 0; k < KM; ++k) {
        wka[k] = 0.0;
        if (k < km) {
            const size_t o = (size_t)k * n2;
            const double dc = dev::wk(utf[o + q], utl[o + q], oncc);
            const double bc = dev::wk(vtf[o + q], vtl[o + q], oncc);
            if (do_u) {
                const double dn = dev::wk(utf[o + qn], utl[o + qn], oncc);
                uw[o + q] = (dn + dc) * htw_c * 0.25;
            }
            if (do_v) {
                const double be = dev::wk(vtf[o + qe], vtl[o + qe], oncc);
                vs[o + q] = (bc + be) * hts_c * 0.25;
            }
            if (do_w && k <= kmt - 1) {
                const double de = dev::wk(utf[o + qe], utl[o + qe], oncc);
                const double dn = dev::wk(utf[o + qn], utl[o + qn], oncc);
                const double dne = dev::wk(utf[o + qne], utl[o + qne], oncc);
                const double be = dev::wk(vtf[o + qe], vtl[o + qe], oncc);
                const double bn = dev::wk(vtf[o + qn], vtl[o + qn], oncc);
                const double bne = dev::wk(vtf[o + qne], vtl[o + qne], oncc);
                // operators_div (tracer.cpp:995-1003) on uk,vk of upwell_2
                const double d = 0.5 *
                                 ((fe * de + fne * dne) * htw_e - (fc * dc + fn * dn) * htw_c +
                                  (fe * be + fc * bc) * hts_c - (fne * bne + fn * bn) * hts_n) *
                                 tr;
                wka[k] = d;
                w = w - g.dzp[k] * d * 1.0;
            }
        }
    }
    if (!do_w) return;
    // upwell_3 / upwell_4 (tracer.cpp:1045-1079)
    const double ohbt = g.ohbt[q];
    const double sc = 1.0 / (1.0 + stf_pre[q] * ohbt);
    double prev = ws[q];
#pragma unroll
    for (int k = 1; k < KM; ++k) {
        if (k < km) {
            const double cur = (k < kmt ? 1.0 : 0.0) * (prev + g.dzp[k - 1] * (w * ohbt + wka[k - 1]));
            ws[(size_t)k * n2 + q] = cur * sc;
            prev = cur;
        }
    }
}

// --------------------------------------------------------------- march
constexpr int BX = 32, BY = 16;                 // interior tile (one lane per column)
constexpr int PX = BX + 2, PY = BY + 2;         // tile + 1-cell ring (predictor, d2tk)
constexpr int RX = BX + 4, RY = BY + 4;         // tile + 2-cell halo (staged planes)
constexpr int NT = BX * BY;                     // 512 lanes
constexpr int NP = PX * PY, NR = RX * RY;
constexpr int NRING = NP - BX * BY;             // 100 ring cells

struct MarchArgs {
    const double *at, *atb, *akt, *uw, *vs, *ws, *stf_n, *err_norm2, *swv, *restore;
    double *ori, *ax, *ay, *az, *dz, *pen;
    int ntx, ntiles, tiles_per_xcd;
};

// Ring cell r (0..NRING-1) -> P-plane coordinates.
__device__ __forceinline__ void ring_cell(int r, int &pj, int &pi) {
    if (r < PX) { pj = 0; pi = r; return; }
    r -= PX;
    if (r < PX) { pj = PY - 1; pi = r; return; }
    r -= PX;
    if (r < PY - 2) { pj = 1 + r; pi = 0; return; }
    r -= PY - 2;
    pj = 1 + r;
    pi = PX - 1;
}

// Own P cell: static metrics kept in registers for the whole march.
struct OwnCell {
    int rr;            // index in the staged R-planes
    bool ok;           // inside [1,jmt-1) x [1,imt-1): predicted and mixed
    double tr, D1, D2, D3, D4;
    double dtn, dts, dte, dtw, ahf;
};

// Lax-Wendroff terms of the own cell carried from the predictor of level k
// to the limiter of level k.
struct LW {
    double xy1, xy2, xy3, xy4, zb1, zb2;
};

// Velocities of one level at a P cell.
struct Vel {
    double Ue, Uw, Vs, Vn, Wk, Wb;
};

__device__ __forceinline__ void own_init(const KGrid &g, OwnCell &c, int gj, int gi, int rr) {
    c.rr = rr;
    c.ok = gj >= 1 && gj < g.jmt - 1 && gi >= 1 && gi < g.imt - 1;
    c.tr = c.D1 = c.D2 = c.D3 = c.D4 = 0.0;
    c.dtn = c.dts = c.dte = c.dtw = c.ahf = 0.0;
    if (!c.ok) return;
    const size_t q = (size_t)gj * g.imt + gi, sj = g.imt;
    c.tr = g.tarea_r[q];
    c.D1 = g.htw[q + 1] * g.hun[q + 1];
    c.D2 = g.htw[q] * g.hun[q];
    c.D3 = g.hts[q] * g.hue[q];
    c.D4 = g.hts[q - sj] * g.hue[q - sj];
    c.dtn = g.dtn[q];
    c.dts = g.dts[q];
    c.dte = g.dte[q];
    c.dtw = g.dtw[q];
    c.ahf = g.ahf[q];
}

__device__ __forceinline__ bool inside1(const KGrid &g, int gj, int gi) {
    return gj >= 1 && gj < g.jmt - 1 && gi >= 1 && gi < g.imt - 1;
}

// Velocities of level K at global cell (gj, gi) (cached global loads).
__device__ __forceinline__ Vel load_vel(const KGrid &g, const MarchArgs &a, int gj, int gi, int K) {
    const size_t n2 = (size_t)g.imt * g.jmt;
    const size_t o = (size_t)K * n2 + (size_t)gj * g.imt + gi;
    Vel v;
    v.Uw = a.uw[o];
    v.Ue = a.uw[o + 1];
    v.Vs = a.vs[o];
    v.Vn = a.vs[o - g.imt];
    v.Wk = a.ws[o];
    v.Wb = (K < g.km - 1) ? a.ws[o + n2] : 0.0;
    return v;
}

// Predictor and limiter flag of level K at R-plane index r from the staged
// planes Pk (level K), Pu (K-1), Pd (K+1); returns the LW terms in lw.
__device__ __forceinline__ bool pred_flag(const KGrid &g, int r, const double *Pk,
                                          const double *Pu, const double *Pd, const int *sK,
                                          int K, double dts, const Vel &v, double tr, double D1,
                                          double D2, double D3, double D4, LW &lw) {
    const double Tc = Pk[r], Te = Pk[r + 1], Tw = Pk[r - 1], Ts = Pk[r + RX], Tn = Pk[r - RX];
    const double Tu = K > 0 ? Pu[r] : 0.0;
    const double Td = K < g.km - 1 ? Pd[r] : 0.0;
    dev::PTerms t;
    dev::pterms(Tc, Te, Tw, Ts, Tn, Tu, Td, v.Ue, v.Uw, v.Vs, v.Vn, v.Wk, v.Wb, tr, D1, D2, D3, D4,
                K, g.km, g.odzp, g.odzt, dts, t);
    lw.xy1 = t.xy1;
    lw.xy2 = t.xy2;
    lw.xy3 = t.xy3;
    lw.xy4 = t.xy4;
    lw.zb1 = t.zb1;
    lw.zb2 = t.zb2;
    const double a00 = dev::predictor(t, Tc, dts);
    // 7-point bounds with land masked to -+1e10 (tracer.cpp:416-545)
    const double lo = -1.0e10, hi = 1.0e10;
    double mx = sK[r] > K ? Tc : lo, mn = sK[r] > K ? Tc : hi;
    const int nb[4] = {r - RX, r + RX, r - 1, r + 1};
    const double tv[4] = {Tn, Ts, Tw, Te};
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const bool wet = sK[nb[m]] > K;
        mx = fmax(mx, wet ? tv[m] : lo);
        mn = fmin(mn, wet ? tv[m] : hi);
    }
    if (K > 0) {
        const bool wet = sK[r] > K - 1;
        mx = fmax(mx, wet ? Tu : lo);
        mn = fmin(mn, wet ? Tu : hi);
    }
    if (K < g.km - 1) {
        const bool wet = sK[r] > K + 1;
        mx = fmax(mx, wet ? Td : lo);
        mn = fmin(mn, wet ? Td : hi);
    }
    return a00 > mx || a00 < mn;
}

// Ring cell: metrics loaded on the fly (the ring is ~1/5 of the tile).
__device__ __forceinline__ bool ring_flag(const KGrid &g, int gj, int gi, int r, const double *Pk,
                                          const double *Pu, const double *Pd, const int *sK, int K,
                                          double dts, const Vel &v) {
    const size_t q = (size_t)gj * g.imt + gi, sj = g.imt;
    LW lw;
    return pred_flag(g, r, Pk, Pu, Pd, sK, K, dts, v, g.tarea_r[q], g.htw[q + 1] * g.hun[q + 1],
                     g.htw[q] * g.hun[q], g.hts[q] * g.hue[q], g.hts[q - sj] * g.hue[q - sj], lw);
}

// del4 first pass d2tk = AHF * del2(atb) at R-plane index r (tracer.cpp:1249-1286).
__device__ __forceinline__ double d2tk_at(int r, const double *B, const int *sK, int k, double dtn,
                                          double dts, double dte, double dtw, double ahf) {
    double hdtk = 0.0, cc = 0.0;
    if (k <= sK[r] - 1) {
        double cf;
        if (k <= sK[r - RX] - 1) { cf = dtn; hdtk += cf * B[r - RX]; cc -= cf; }
        if (k <= sK[r + RX] - 1) { cf = dts; hdtk += cf * B[r + RX]; cc -= cf; }
        if (k <= sK[r + 1] - 1) { cf = dte; hdtk += cf * B[r + 1]; cc -= cf; }
        if (k <= sK[r - 1] - 1) { cf = dtw; hdtk += cf * B[r - 1]; cc -= cf; }
    }
    return ahf * (cc * B[r] + hdtk);
}

__device__ __forceinline__ void load_plane(double *dst, const double *src, int gj0, int gi0,
                                           int imt, int jmt) {
    for (int r = threadIdx.x; r < NR; r += NT) {
        const int gj = gj0 + r / RX, gi = gi0 + r % RX;
        dst[r] = (gj >= 0 && gj < jmt && gi >= 0 && gi < imt) ? src[(size_t)gj * imt + gi] : 0.0;
    }
}

__global__ __launch_bounds__(NT) void k_march(KGrid g, KPar p, MarchArgs a) {
    __shared__ double sAT[4][NR];
    __shared__ double sB[3][NR];
    __shared__ double sD2[NP];
    __shared__ unsigned char sFL[2][NP];
    __shared__ int sK[NR];

    // XCD-aware placement: consecutive blocks of one XCD take the T and S
    // halves of one tile, then the next tile of the same row of tiles.
    const int L = blockIdx.x, xcd = L & 7, slot = L >> 3;
    const int tile = xcd * a.tiles_per_xcd + (slot >> 1);
    const int n = slot & 1;
    if (tile >= a.ntiles) return;
    const int imt = g.imt, jmt = g.jmt, km = g.km;
    const size_t n2 = (size_t)imt * jmt, n3 = n2 * km;
    const int gi0 = 2 + (tile % a.ntx) * BX, gj0 = 2 + (tile / a.ntx) * BY;

    const int t = threadIdx.x, li = t % BX, lj = t / BX;
    const int gi = gi0 + li, gj = gj0 + lj;
    const bool act = gi < imt - 2 && gj < jmt - 2;
    const size_t q = (size_t)gj * imt + gi;

    const double *T = a.at + n * n3;
    const double *TB = a.atb + n * (n3 + n2) + n2;   // level k at plane k
    const double *A = a.akt + n * n3;
    const double dts = p.dts;

    OwnCell own;
    own_init(g, own, gj, gi, (lj + 2) * RX + (li + 2));
    const int pOwn = (lj + 1) * PX + (li + 1);
    // ring slot of this lane (lanes >= NRING have none)
    int rpj = 0, rpi = 0;
    if (t < NRING) ring_cell(t, rpj, rpi);
    const int rgj = gj0 - 1 + rpj, rgi = gi0 - 1 + rpi;
    const bool rok = t < NRING && inside1(g, rgj, rgi);
    const int rR = (rpj + 1) * RX + (rpi + 1), pR = rpj * PX + rpi;
    double r_dtn = 0, r_dts = 0, r_dte = 0, r_dtw = 0, r_ahf = 0;
    if (rok && p.hmix == LICOM_HMIX_DEL4) {
        const size_t rq = (size_t)rgj * imt + rgi;
        r_dtn = g.dtn[rq];
        r_dts = g.dts[rq];
        r_dte = g.dte[rq];
        r_dtw = g.dtw[rq];
        r_ahf = g.ahf[rq];
    }

    for (int r = t; r < NR; r += NT) {
        const int rj = gj0 - 2 + r / RX, ri = gi0 - 2 + r % RX;
        sK[r] = (rj >= 0 && rj < jmt && ri >= 0 && ri < imt) ? g.kmt[(size_t)rj * imt + ri] : 0;
    }
    const double sw = (n == 0 && act && a.swv) ? a.swv[q] : 0.0;
    const int kmt_own = act ? g.kmt[q] : 0;
    const double stf_n = act ? a.stf_n[n * n2 + q] : 0.0;
    const double enorm = (n == 1 && p.sssnorm && act) ? a.err_norm2[0] : 0.0;

    load_plane(sAT[0], T, gj0 - 2, gi0 - 2, imt, jmt);
    load_plane(sAT[1], T + n2, gj0 - 2, gi0 - 2, imt, jmt);
    load_plane(sB[0], TB, gj0 - 2, gi0 - 2, imt, jmt);
    Vel vc = own.ok ? load_vel(g, a, gj, gi, 0) : Vel{0, 0, 0, 0, 0, 0};
    Vel vr = rok ? load_vel(g, a, rgj, rgi, 0) : Vel{0, 0, 0, 0, 0, 0};
    __syncthreads();

    LW lc{}, ln{};
    {
        bool f0 = false;
        if (own.ok)
            f0 = pred_flag(g, own.rr, sAT[0], sAT[0], sAT[1], sK, 0, dts, vc, own.tr, own.D1,
                           own.D2, own.D3, own.D4, lc);
        sFL[0][pOwn] = f0;
        if (rok) sFL[0][pR] = ring_flag(g, rgj, rgi, rR, sAT[0], sAT[0], sAT[1], sK, 0, dts, vr);
    }
    double Tprev = 0.0, Bprev = 0.0, Aprev = 0.0;
    bool fprev = false;
    const double lamda = 1.0 / (15.0 * 86400.0);
    const double aidif = 0.5;

    for (int k = 0; k < km; ++k) {
        // ---- phase L: stage the planes of levels k+2 (at) and k+1 (atb)
        if (k + 2 < km) load_plane(sAT[(k + 2) & 3], T + (size_t)(k + 2) * n2, gj0 - 2, gi0 - 2, imt, jmt);
        if (k + 1 < km) load_plane(sB[(k + 1) % 3], TB + (size_t)(k + 1) * n2, gj0 - 2, gi0 - 2, imt, jmt);
        const double Acur = act ? dev::floor_akt(A[(size_t)k * n2 + q], k, n, p.hmix, p.dwndmix) : 0.0;
        const double Rk = (p.boundary_restore == 1 && a.restore && act && k >= 1)
                              ? a.restore[n * n3 + (size_t)k * n2 + q] : 0.0;
        Vel vn{0, 0, 0, 0, 0, 0}, vrn{0, 0, 0, 0, 0, 0};
        if (k + 1 < km) {
            if (own.ok) vn = load_vel(g, a, gj, gi, k + 1);
            if (rok) vrn = load_vel(g, a, rgj, rgi, k + 1);
        }
        __syncthreads();

        // ---- phase P: predictor of level k+1 (own + ring), d2tk of level k
        const double *Pk = sAT[k & 3], *Pk1 = sAT[(k + 1) & 3], *Pk2 = sAT[(k + 2) & 3];
        const double *Bk = sB[k % 3], *Bk1 = sB[(k + 1) % 3];
        if (k + 1 < km) {
            bool f = false;
            if (own.ok)
                f = pred_flag(g, own.rr, Pk1, Pk, Pk2, sK, k + 1, dts, vn, own.tr, own.D1, own.D2,
                              own.D3, own.D4, ln);
            sFL[(k + 1) & 1][pOwn] = f;
            if (rok) sFL[(k + 1) & 1][pR] = ring_flag(g, rgj, rgi, rR, Pk1, Pk, Pk2, sK, k + 1, dts, vrn);
        }
        if (p.hmix == LICOM_HMIX_DEL4) {
            sD2[pOwn] = own.ok ? d2tk_at(own.rr, Bk, sK, k, own.dtn, own.dts, own.dte, own.dtw, own.ahf)
                               : 0.0;
            if (rok) sD2[pR] = d2tk_at(rR, Bk, sK, k, r_dtn, r_dts, r_dte, r_dtw, r_ahf);
        }
        __syncthreads();

        // ---- phase C: limited fluxes and the explicit update of level k
        if (act) {
            const unsigned char *F = sFL[k & 1];
            const int r = own.rr, pp = pOwn;
            const double Tc = Pk[r], Te = Pk[r + 1], Tw = Pk[r - 1], Ts = Pk[r + RX], Tn = Pk[r - RX];
            const double Tu = Tprev, Td = (k + 1 < km) ? Pk1[r] : 0.0;
            const bool fc = F[pp], fe = F[pp + 1], fw = F[pp - 1], fs = F[pp + PX], fn = F[pp - PX];
            const bool fd = (k + 1 < km) ? sFL[(k + 1) & 1][pp] : false;
            bool up1, up2, up3, up4;
            if (k == 0 || k == km - 1) {
                up1 = fc || (fe && gi <= imt - 3);
                up2 = fc || (fw && gi >= 2);
                up3 = fc || (fs && gj <= jmt - 3);
                up4 = fc || (fn && gj >= 2);
            } else {
                up1 = (fc || fe) && gi <= imt - 3;
                up2 = (fc || fw) && gi >= 2;
                up3 = (fc || fs) && gj <= jmt - 3;
                up4 = (fc || fn) && gj >= 2;
            }
            const double tr = own.tr;
            dev::PTerms tc;
            dev::pterms_cheap(Tc, Te, Tw, Ts, Tn, Tu, Td, vc.Ue, vc.Uw, vc.Vs, vc.Vn, vc.Wk, vc.Wb, tr,
                              k, km, g.odzp, tc);
            const double xy1 = up1 ? dev::up_xy1(Te, Tc, vc.Ue, tr) : lc.xy1;
            const double xy2 = up2 ? dev::up_xy2(Tc, Tw, vc.Uw, tr) : lc.xy2;
            const double xy3 = up3 ? dev::up_xy1(Ts, Tc, vc.Vs, tr) : lc.xy3;
            const double xy4 = up4 ? dev::up_xy2(Tc, Tn, vc.Vn, tr) : lc.xy4;
            double zb1, zb2;
            if (k == 0) {
                zb1 = 0.0;
                zb2 = (fc || fd) ? dev::up_zb2(Tc, Td, vc.Wb, g.odzp[0]) : lc.zb2;
            } else if (k == km - 1) {
                zb1 = (fc || fprev) ? dev::up_zb1(Tu, Tc, vc.Wk, g.odzp[k]) : lc.zb1;
                zb2 = 0.0;
            } else {
                zb1 = (fc || fprev) ? dev::up_zb1(Tu, Tc, vc.Wk, g.odzp[k]) : lc.zb1;
                zb2 = (fc || fd) ? dev::up_zb2(Tc, Td, vc.Wb, g.odzp[k]) : lc.zb2;
            }
            const double xx = -(tc.x0 + xy1 + xy2 + tc.c1);
            const double yy = -(tc.y0 + xy3 + xy4 + tc.c2);
            const double zz = -(tc.za + zb1 + zb2 + tc.zc);
            const size_t c = (size_t)k * n2 + q;
            if (a.ax) a.ax[n * n3 + c] = xx;
            if (a.ay) a.ay[n * n3 + c] = yy;
            if (a.az) a.az[n * n3 + c] = zz;
            const double vk = k < kmt_own ? 1.0 : 0.0;
            double tf = (xx + yy + zz) * vk;
            // horizontal mixing (tracer_del4_2 / tracer_4)
            {
                const bool d4 = p.hmix == LICOM_HMIX_DEL4;
                const double *X = d4 ? sD2 : Bk;
                const int xs = d4 ? pp : r, xw = d4 ? PX : RX;
                double hdtk = 0.0, cc = 0.0;
                if (k <= sK[r] - 1) {
                    double cf;
                    if (k <= sK[r - RX] - 1) { cf = own.dtn; hdtk += cf * X[xs - xw]; cc -= cf; }
                    if (k <= sK[r + RX] - 1) { cf = own.dts; hdtk += cf * X[xs + xw]; cc -= cf; }
                    if (k <= sK[r + 1] - 1) { cf = own.dte; hdtk += cf * X[xs + 1]; cc -= cf; }
                    if (k <= sK[r - 1] - 1) { cf = own.dtw; hdtk += cf * X[xs - 1]; cc -= cf; }
                }
                tf += p.ah * (cc * X[xs] + hdtk);
            }
            // shortwave penetration (T) and explicit vertical diffusion
            // (tracer_5, boundary levels as its k==0 thread applies them)
            const double *odzp = g.odzp, *odzt = g.odzt;
            const double v1 = (k + 1 < km && k + 1 < kmt_own) ? 1.0 : 0.0;
            if (n == 0) {
                double pv;
                if (k == 0) {
                    const double wt1 = sw * g.pen[0] * (1 < kmt_own ? 1.0 : 0.0);
                    tf = tf - odzp[0] * wt1;
                    pv = -wt1 * odzp[0];
                } else if (k == km - 1) {
                    const double wt2 = sw * g.pen[km - 2] * vk;
                    tf = tf + odzp[km - 1] * wt2;
                    pv = wt2 * odzp[km - 1];
                } else {
                    const double wt1 = sw * g.pen[k - 1] * vk;
                    const double wt2 = sw * g.pen[k] * v1;
                    tf = tf + (wt1 - wt2) * odzp[k];
                    pv = (wt1 - wt2) * odzp[k];
                }
                if (a.pen) a.pen[c] = pv;
            }
            const double Bc = Bk[r];
            double dzv;
            if (k == 0) {
                const double wt1 = Acur * (Bc - Bk1[r]) * odzt[1] * v1;
                tf = tf - odzp[0] * wt1 * 0.5;
                dzv = -odzp[0] * wt1 * 0.5;
            } else if (k == km - 1) {
                const double wt2 = Aprev * (Bprev - Bc) * odzt[km - 1] * vk;
                tf = tf + odzp[km - 1] * wt2 * (1.0 - aidif);
                dzv = odzp[km - 1] * wt2 * 0.5;
            } else {
                const double wt1 = Aprev * (Bprev - Bc) * odzt[k] * vk;
                const double wt2 = Acur * (Bc - Bk1[r]) * odzt[k + 1] * v1;
                tf = tf + odzp[k] * (wt1 - wt2) * 0.5;
                dzv = odzp[k] * (wt1 - wt2) * 0.5;
            }
            if (a.dz) a.dz[n * n3 + c] = dzv;
            // surface fluxes (tracer_5 n=1, tracer_6) and restoring (tracer_7)
            if (k == 0) {
                if (n == 1) {
                    if (kmt_own > 0) tf = tf + stf_n * (1.0 - aidif) * odzp[0];
                    if (p.sssnorm) tf = tf + enorm * vk;
                } else if (kmt_own > 0) {
                    tf = tf + stf_n * odzp[0] * 0.5;
                }
            }
            if (p.boundary_restore == 1 && a.restore && k >= 1) tf = tf + vk * (Rk - Bc) * lamda;
            // tracer_8: the explicit update entering the implicit solve
            a.ori[n * n3 + c] = Tc + dts * tf;
            Tprev = Tc;
            Bprev = Bc;
            fprev = fc;
        }
        Aprev = Acur;
        lc = ln;
        vc = vn;
    }
}

// ---------------------------------------------------------------- solve
// Two lanes per column: lane parity = tracer; 32 columns along i per wave.
//
// invtrit (tracer.cpp:923-978) eliminates from the bottom and substitutes
// from the top.  Holding e(k), f(k) for the whole column would take 2*km
// doubles per lane; instead the elimination keeps them only at every SEG-th
// level, and the substitution recomputes each SEG-level segment from its
// checkpoint with the same operations (identical bits), re-reading the
// segment's inputs while they are still in cache.
constexpr int SEG = 8;

template <int KM>
__global__ __launch_bounds__(256) void k_solve(KGrid g, KPar p, int flags, double *__restrict__ at,
                                               double *__restrict__ atb,
                                               const double *__restrict__ akt,
                                               const double *__restrict__ ori,
                                               const double *__restrict__ stf_n,
                                               double *__restrict__ tend,
                                               double *__restrict__ dt_diff,
                                               double *__restrict__ dt_conv, int *mix_count) {
    constexpr int NSEG = (KM + SEG - 1) / SEG;
    const int t = threadIdx.x, n = t & 1, col = t >> 1;
    const int gi = 2 + blockIdx.x * 32 + (col & 31);
    const int gj = 2 + blockIdx.y * 4 + (col >> 5);
    const int imt = g.imt, jmt = g.jmt, km = g.km;
    const bool act = gi < imt - 2 && gj < jmt - 2;
    const size_t n2 = (size_t)imt * jmt, n3 = n2 * km;
    const size_t q = act ? (size_t)gj * imt + gi : 0;
    const int kz = act ? g.kmt[q] : 0;
    const double *O = ori + n * n3;
    const double *A = akt + n * n3;
    const double *odzp = g.odzp, *odzt = g.odzt;
    const double dts = p.dts;
    const double stf = act ? stf_n[n * n2 + q] : 0.0;
    auto wkc = [&](int k) { return dev::floor_akt(A[(size_t)k * n2 + q], k, n, p.hmix, p.dwndmix); };

    // elimination from the bottom; checkpoints e,f at levels SEG*m, m >= 1
    double cke[NSEG + 1], ckf[NSEG + 1];
#pragma unroll
    for (int m = 0; m <= NSEG; ++m) cke[m] = ckf[m] = 0.0;
    {
        double en = 0.0, fn = 0.0, wk_k = 0.0;
#pragma unroll
        for (int k = KM - 1; k >= 0; --k) {
            if (k < kz) {
                const double d = O[(size_t)k * n2 + q];
                const double wkm1 = k >= 1 ? wkc(k - 1) : 0.0;
                double aa, bb, cc;
                dev::tri_row(k, kz, wkm1, wk_k, odzp, odzt, dts, aa, bb, cc);
                const double g0 = 1.0 / (bb - cc * en);
                en = aa * g0;
                fn = (d + cc * fn) * g0;
                wk_k = wkm1;
            }
            if (k > 0 && k % SEG == 0) {
                cke[k / SEG] = en;
                ckf[k / SEG] = fn;
            }
        }
    }

    // substitution from the top, tracer_9, and the convadj stability scan
    const bool conv = flags & LICOM_STEP_CONVADJ;
    const double *C = g.eosc, *to = g.to, *so = g.so;
    double xprev = 0.0, Tp = 0.0, Sp = 0.0;
    bool unstable = false;
    double *TT = at + n * n3, *BB = atb + n * (n3 + n2) + n2;
#pragma unroll
    for (int m = 0; m < NSEG; ++m) {
        const int top = SEG * m;
        double se[SEG], sf[SEG], sd[SEG];
        {
            double ee = cke[m + 1], ff = ckf[m + 1], wk_k = 0.0;
#pragma unroll
            for (int r = SEG - 1; r >= 0; --r) {
                const int k = top + r;
                se[r] = sf[r] = sd[r] = 0.0;
                if (k < KM && k < kz) {
                    if (r == SEG - 1 && k < kz - 1) wk_k = wkc(k);
                    const double d = O[(size_t)k * n2 + q];
                    const double wkm1 = k >= 1 ? wkc(k - 1) : 0.0;
                    double aa, bb, cc;
                    dev::tri_row(k, kz, wkm1, wk_k, odzp, odzt, dts, aa, bb, cc);
                    const double g0 = 1.0 / (bb - cc * ee);
                    ee = aa * g0;
                    ff = (d + cc * ff) * g0;
                    se[r] = ee;
                    sf[r] = ff;
                    sd[r] = d;
                    wk_k = wkm1;
                }
            }
        }
#pragma unroll
        for (int r = 0; r < SEG; ++r) {
            const int k = top + r;
            if (k < KM && k < km) {
                const size_t c = (size_t)k * n2 + q;
                const bool wet = k < kz;
                const double o = wet ? sd[r] : (act ? O[c] : 0.0);
                double x = o;
                if (wet) x = (k == 0) ? (se[0] * stf + sf[0]) * 1.0 : (se[r] * xprev + sf[r]) * 1.0;
                xprev = x;
                const double vk = wet ? 1.0 : 0.0;
                if (act) {
                    const double told = TT[c];
                    if (dt_diff) {
                        double v = (x - o) / dts * vk;
                        if (k == 0) v = v - stf * 0.5 * odzp[0];
                        dt_diff[n * n3 + c] = v;
                    }
                    if (tend) tend[n * n3 + c] = (x - told) / dts * vk;
                    if (conv && dt_conv) dt_conv[n * n3 + c] = 0.0;
                    TT[c] = x;
                    BB[c] = x;
                }
                if (conv) {
                    const double xo = __shfl_xor(x, 1);
                    const double Tk = n == 0 ? x : xo, Sk = n == 0 ? xo : x;
                    if (k >= 1 && k <= kz - 1) {
                        const double up = dev::dens(C, km, k, Tk - to[k], Sk - so[k]);
                        const double lo = dev::dens(C, km, k, Tp - to[k], Sp - so[k]);
                        if (lo > up) unstable = true;
                    }
                    Tp = Tk;
                    Sp = Sk;
                }
            }
        }
    }

    // convadj_cu (convadj.cpp:52-275) for statically unstable columns, on
    // the state just written (at == atb == tracer_9's values, cache-resident)
    if (conv && act && n == 0 && unstable) {
        atomicAdd(mix_count, 1);
        double rhoup[KMAXF], rholo[KMAXF];
        const dev::Strided Tg{at + q, n2}, Sg{at + n3 + q, n2};
        dev::convect_column(C, to, so, g.dzp, km, kz, Tg, Sg, rhoup, rholo);
        for (int nn = 0; nn < 2; ++nn) {
            double *Tn = at + nn * n3, *Bn = atb + nn * (n3 + n2) + n2;
            for (int k = 0; k < kz; ++k) {
                const size_t c = (size_t)k * n2 + q;
                const double y = Tn[c], b = Bn[c];
                if (y != b) {
                    const double dc = (y - b) / dts * 1.0;
                    if (dt_conv) dt_conv[nn * n3 + c] = dc;
                    if (tend) tend[nn * n3 + c] += dc;
                    Bn[c] = y;
                }
            }
        }
    }
}

// In-place side effects of the reference on its inputs
// (tracer.cpp:1088, 1107-1108, 1187, 1315).
__global__ void k_side(KGrid g, KPar p, double *utf, double *vtf, double *h0f, double *akt) {
    const size_t n2 = (size_t)g.imt * g.jmt, n3 = n2 * g.km;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n3;
         t += (size_t)gridDim.x * blockDim.x) {
        utf[t] *= p.oncc;
        vtf[t] *= p.oncc;
        if (t < n2) {
            h0f[t] = h0f[t] * p.onbc;
            const int j = (int)(t / g.imt), i = (int)(t % g.imt);
            if (i >= 1 && i < g.imt - 1 && j >= 1 && j < g.jmt - 1)
                for (int n = 0; n < 2; ++n)
                    akt[n * n3 + t] = dev::floor_akt(akt[n * n3 + t], 0, n, p.hmix, p.dwndmix);
        }
    }
}

template <int KM>
void launch_prelude_solve(licom_ctx *ctx, const licom_fields *f, const licom_diag *d, int flags,
                          bool prelude) {
    const KGrid &g = ctx->g;
    const KPar &p = ctx->kp;
    Scratch &s = ctx->s;
    hipStream_t st = ctx->stream;
    if (prelude) {
        Timer t(ctx, "k_prelude");
        const dim3 b(64, 4), gr((g.imt + 63) / 64, (g.jmt + 3) / 4);
        k_prelude<KM><<<gr, b, 0, st>>>(g, p, f->utf, f->utl, f->vtf, f->vtl, s.work_u, s.stf_pre,
                                       f->ws, s.uw, s.vs);
    } else {
        Timer t(ctx, "k_solve");
        const dim3 gr((g.imt - 4 + 31) / 32, (g.jmt - 4 + 3) / 4);
        k_solve<KM><<<gr, 256, 0, st>>>(g, p, flags, f->at, f->atb, f->akt, s.ori, s.stf_n,
                                       d ? d->tend : nullptr, d ? d->dt_diff : nullptr,
                                       d ? d->dt_conv : nullptr, s.mix_count);
    }
}

void launch_column(licom_ctx *ctx, const licom_fields *f, const licom_diag *d, int flags,
                   bool prelude) {
    switch (ctx->km) {
        case 30: launch_prelude_solve<30>(ctx, f, d, flags, prelude); break;
        case 55: launch_prelude_solve<55>(ctx, f, d, flags, prelude); break;
        default: launch_prelude_solve<KMAXF>(ctx, f, d, flags, prelude); break;
    }
}

}  // namespace

int launch_step_fused(licom_ctx *ctx, const licom_fields *f, const licom_diag *d, int flags) {
    const KGrid &g = ctx->g;
    Scratch &s = ctx->s;
    hipStream_t st = ctx->stream;
    if (int rc = launch_surface(ctx, f, d)) return rc;
    launch_column(ctx, f, d, flags, true);
    {
        Timer t(ctx, "k_march");
        MarchArgs a{};
        a.at = f->at;
        a.atb = f->atb;
        a.akt = f->akt;
        a.uw = s.uw;
        a.vs = s.vs;
        a.ws = f->ws;
        a.stf_n = s.stf_n;
        a.err_norm2 = s.err_norm2;
        a.swv = f->swv;
        a.restore = f->restore;
        a.ori = s.ori;
        a.ax = d ? d->ax : nullptr;
        a.ay = d ? d->ay : nullptr;
        a.az = d ? d->az : nullptr;
        a.dz = d ? d->dz : nullptr;
        a.pen = d ? d->penetrate : nullptr;
        a.ntx = (g.imt - 4 + BX - 1) / BX;
        const int nty = (g.jmt - 4 + BY - 1) / BY;
        a.ntiles = a.ntx * nty;
        a.tiles_per_xcd = (a.ntiles + 7) / 8;
        k_march<<<16 * a.tiles_per_xcd, NT, 0, st>>>(g, ctx->kp, a);
    }
    launch_column(ctx, f, d, flags, false);
    if (flags & LICOM_STEP_SIDE_EFFECTS) {
        Timer t(ctx, "k_side");
        k_side<<<4096, 256, 0, st>>>(g, ctx->kp, f->utf, f->vtf, f->h0f, f->akt);
    }
    return launch_halo(ctx, f, flags);
}

}  // namespace licom
