// tracer_fused.hip -- production path of the FP64 tracer step on gfx950.
//
// Three HBM passes instead of the reference's ~21 launches per tracer:
//
//   k_prelude<KM>  one lane per column: face velocities u_wface/v_sface
//                  (advection_tracer_1_0) and the barotropic-corrected
//                  continuity w (upwell_2..4, operators_div); the column's
//                  divergence profile stays in registers between the
//                  column sum and the w recursion.
//   k_march        one workgroup per 16x16 tile of columns and tracer,
//                  marching k top to bottom.  Planes of at, atb and the
//                  velocities (tile plus a 2-cell halo) are staged in LDS,
//                  fetched a level ahead into registers; the tspas predictor of
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
constexpr int BX = 16, BY = 16;                 // interior tile (one lane per column)
constexpr int PX = BX + 2, PY = BY + 2;         // tile + 1-cell ring (predictor, d2tk)
constexpr int RX = BX + 4, RY = BY + 4;         // tile + 2-cell halo (staged planes)
constexpr int NT = BX * BY;                     // 256 lanes
constexpr int NP = PX * PY, NR = RX * RY;
constexpr int NRING = NP - BX * BY;             // 68 ring cells
constexpr int NPF = (NR + NT - 1) / NT;         // plane cells staged per lane

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

// Static metrics of a predicted cell: tarea_r, the four face products
// htw*hun (i+1, i), hts*hue (j, j-1) and the del2 stencil (DTN..DTW, AHF).
struct Metrics {
    double tr, D1, D2, D3, D4, dtn, dts, dte, dtw, ahf;
};

__device__ __forceinline__ Metrics load_metrics(const KGrid &g, int gj, int gi) {
    const size_t q = (size_t)gj * g.imt + gi, sj = g.imt;
    Metrics m;
    m.tr = g.tarea_r[q];
    m.D1 = g.htw[q + 1] * g.hun[q + 1];
    m.D2 = g.htw[q] * g.hun[q];
    m.D3 = g.hts[q] * g.hue[q];
    m.D4 = g.hts[q - sj] * g.hue[q - sj];
    m.dtn = g.dtn[q];
    m.dts = g.dts[q];
    m.dte = g.dte[q];
    m.dtw = g.dtw[q];
    m.ahf = g.ahf[q];
    return m;
}

// LW terms carried from the predictor of level k to the limiter of level k.
struct LW {
    double xy1, xy2, xy3, xy4, zb1, zb2;
};

// Staged planes of the levels around the one being predicted.
struct Lv {
    const double *Pk, *Pu, *Pd;     // at of levels K, K-1, K+1
    const double *U, *V, *Wk, *Wb;  // u_wface, v_sface of level K; ws at K and K+1
};

// Predictor and limiter flag of level K at R-plane index r.
__device__ __forceinline__ bool pred_flag(const KGrid &g, int r, const Lv &L, const int *sK, int K,
                                          double dts, double tr, double D1, double D2, double D3,
                                          double D4, LW &lw) {
    const double Tc = L.Pk[r], Te = L.Pk[r + 1], Tw = L.Pk[r - 1], Ts = L.Pk[r + RX],
                 Tn = L.Pk[r - RX];
    const double Tu = K > 0 ? L.Pu[r] : 0.0;
    const double Td = K < g.km - 1 ? L.Pd[r] : 0.0;
    const double Ue = L.U[r + 1], Uw = L.U[r], Vs = L.V[r], Vn = L.V[r - RX];
    const double Wk = L.Wk[r], Wb = K < g.km - 1 ? L.Wb[r] : 0.0;
    dev::PTerms t;
    dev::pterms(Tc, Te, Tw, Ts, Tn, Tu, Td, Ue, Uw, Vs, Vn, Wk, Wb, tr, D1, D2, D3, D4, K, g.km,
                g.odzp, g.odzt, dts, t);
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

// One lane's share of a staged plane: the R-plane cells t, t+NT, ...
struct PlaneSlots {
    size_t off[NPF];   // offset inside one (jmt,imt) plane
    bool in[NPF];      // cell inside the array (else staged as 0)
};

__device__ __forceinline__ void fetch(const PlaneSlots &s, const double *src, double (&v)[NPF]) {
#pragma unroll
    for (int u = 0; u < NPF; ++u) v[u] = s.in[u] ? src[s.off[u]] : 0.0;
}

__device__ __forceinline__ void stage(const double (&v)[NPF], double *dst) {
#pragma unroll
    for (int u = 0; u < NPF; ++u) {
        const int r = threadIdx.x + u * NT;
        if (r < NR) dst[r] = v[u];
    }
}

// k-march over one 16x16 tile of columns and one tracer.  Every level's
// inputs (at, atb, u_wface, v_sface, ws) are staged as LDS planes covering
// the tile plus a 2-cell halo; each lane fetches its share of the next level
// into registers while the current level is computed.
__global__ __launch_bounds__(NT) void k_march(KGrid g, KPar p, MarchArgs a) {
    __shared__ double sAT[3][NR];
    __shared__ double sB[2][NR];
    __shared__ double sU[2][NR];
    __shared__ double sV[2][NR];
    __shared__ double sW[3][NR];
    __shared__ double sD2[NP];
    __shared__ double sRM[10][NRING];
    __shared__ unsigned char sFL[2][NP];
    __shared__ int sK[NR];

    // XCD-aware placement: consecutive blocks of one XCD take the T and S
    // halves of one tile, then the next tile of the same row of tiles.
    const int Lb = blockIdx.x, xcd = Lb & 7, slot = Lb >> 3;
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

    PlaneSlots ps;
#pragma unroll
    for (int u = 0; u < NPF; ++u) {
        const int r = t + u * NT;
        const int rj = gj0 - 2 + r / RX, ri = gi0 - 2 + r % RX;
        ps.in[u] = r < NR && rj >= 0 && rj < jmt && ri >= 0 && ri < imt;
        ps.off[u] = ps.in[u] ? (size_t)rj * imt + ri : 0;
    }
    for (int r = t; r < NR; r += NT) {
        const int rj = gj0 - 2 + r / RX, ri = gi0 - 2 + r % RX;
        sK[r] = (rj >= 0 && rj < jmt && ri >= 0 && ri < imt) ? g.kmt[(size_t)rj * imt + ri] : 0;
    }
    // own cell and ring cell
    const int rOwn = (lj + 2) * RX + (li + 2), pOwn = (lj + 1) * PX + (li + 1);
    const bool own_ok = gj >= 1 && gj < jmt - 1 && gi >= 1 && gi < imt - 1;
    Metrics M{};
    if (own_ok) M = load_metrics(g, gj, gi);
    int rpj = 0, rpi = 0;
    if (t < NRING) ring_cell(t, rpj, rpi);
    const int rgj = gj0 - 1 + rpj, rgi = gi0 - 1 + rpi;
    const bool rok = t < NRING && rgj >= 1 && rgj < jmt - 1 && rgi >= 1 && rgi < imt - 1;
    const int rR = (rpj + 1) * RX + (rpi + 1), pR = rpj * PX + rpi;
    if (rok) {
        const Metrics m = load_metrics(g, rgj, rgi);
        sRM[0][t] = m.tr; sRM[1][t] = m.D1; sRM[2][t] = m.D2; sRM[3][t] = m.D3; sRM[4][t] = m.D4;
        sRM[5][t] = m.dtn; sRM[6][t] = m.dts; sRM[7][t] = m.dte; sRM[8][t] = m.dtw; sRM[9][t] = m.ahf;
    }
    const double sw = (n == 0 && act && a.swv) ? a.swv[q] : 0.0;
    const int kmt_own = act ? g.kmt[q] : 0;
    const double stf_n = act ? a.stf_n[n * n2 + q] : 0.0;
    const double enorm = (n == 1 && p.sssnorm && act) ? a.err_norm2[0] : 0.0;
    const bool rest = p.boundary_restore == 1 && a.restore != nullptr;

    // prologue: levels 0 and 1 of at and ws, level 0 of atb, u_wface, v_sface
    double v0[NPF], v1[NPF];
    fetch(ps, T, v0);
    fetch(ps, T + n2, v1);
    stage(v0, sAT[0]);
    stage(v1, sAT[1]);
    fetch(ps, TB, v0);
    stage(v0, sB[0]);
    fetch(ps, a.uw, v0);
    fetch(ps, a.vs, v1);
    stage(v0, sU[0]);
    stage(v1, sV[0]);
    fetch(ps, a.ws, v0);
    fetch(ps, a.ws + n2, v1);
    stage(v0, sW[0]);
    stage(v1, sW[1]);
    // registers holding the next level's share of every plane
    double fAT[NPF], fB[NPF], fU[NPF], fV[NPF], fW[NPF];
    if (km > 2) fetch(ps, T + 2 * n2, fAT);
    fetch(ps, TB + n2, fB);
    fetch(ps, a.uw + n2, fU);
    fetch(ps, a.vs + n2, fV);
    if (km > 2) fetch(ps, a.ws + 2 * n2, fW);
    double fA = act ? A[q] : 0.0;
    double fR = 0.0;
    __syncthreads();

    LW lc{}, ln{};
    {
        const Lv L0{sAT[0], sAT[0], sAT[1], sU[0], sV[0], sW[0], sW[1]};
        bool f0 = false;
        if (own_ok) f0 = pred_flag(g, rOwn, L0, sK, 0, dts, M.tr, M.D1, M.D2, M.D3, M.D4, lc);
        sFL[0][pOwn] = f0;
        if (rok) {
            LW tmp;
            sFL[0][pR] = pred_flag(g, rR, L0, sK, 0, dts, sRM[0][t], sRM[1][t], sRM[2][t],
                                   sRM[3][t], sRM[4][t], tmp);
        }
    }
    double Tprev = 0.0, Bprev = 0.0, Aprev = 0.0;
    bool fprev = false;
    const double lamda = 1.0 / (15.0 * 86400.0);
    const double aidif = 0.5;

    for (int k = 0; k < km; ++k) {
        // ---- phase L: stage the prefetched planes, prefetch the next level
        if (k + 2 < km) stage(fAT, sAT[(k + 2) % 3]);
        if (k + 1 < km) {
            stage(fB, sB[(k + 1) & 1]);
            stage(fU, sU[(k + 1) & 1]);
            stage(fV, sV[(k + 1) & 1]);
        }
        if (k + 2 < km) stage(fW, sW[(k + 2) % 3]);
        const double Acur = act ? dev::floor_akt(fA, k, n, p.hmix, p.dwndmix) : 0.0;
        const double Rk = fR;
        if (k + 3 < km) fetch(ps, T + (size_t)(k + 3) * n2, fAT);
        if (k + 2 < km) {
            fetch(ps, TB + (size_t)(k + 2) * n2, fB);
            fetch(ps, a.uw + (size_t)(k + 2) * n2, fU);
            fetch(ps, a.vs + (size_t)(k + 2) * n2, fV);
        }
        if (k + 3 < km) fetch(ps, a.ws + (size_t)(k + 3) * n2, fW);
        if (k + 1 < km) {
            fA = act ? A[(size_t)(k + 1) * n2 + q] : 0.0;
            fR = (rest && act) ? a.restore[n * n3 + (size_t)(k + 1) * n2 + q] : 0.0;
        }
        __syncthreads();

        // ---- phase P: predictor of level k+1 (own + ring), d2tk of level k
        const double *Pk = sAT[k % 3], *Pk1 = sAT[(k + 1) % 3];
        const double *Bk = sB[k & 1], *Bk1 = sB[(k + 1) & 1];
        if (k + 1 < km) {
            const Lv L1{Pk1, Pk, sAT[(k + 2) % 3], sU[(k + 1) & 1], sV[(k + 1) & 1],
                        sW[(k + 1) % 3], sW[(k + 2) % 3]};
            bool f = false;
            if (own_ok) f = pred_flag(g, rOwn, L1, sK, k + 1, dts, M.tr, M.D1, M.D2, M.D3, M.D4, ln);
            sFL[(k + 1) & 1][pOwn] = f;
            if (rok) {
                LW tmp;
                sFL[(k + 1) & 1][pR] = pred_flag(g, rR, L1, sK, k + 1, dts, sRM[0][t], sRM[1][t],
                                                 sRM[2][t], sRM[3][t], sRM[4][t], tmp);
            }
        }
        if (p.hmix == LICOM_HMIX_DEL4) {
            sD2[pOwn] = own_ok ? d2tk_at(rOwn, Bk, sK, k, M.dtn, M.dts, M.dte, M.dtw, M.ahf) : 0.0;
            if (rok)
                sD2[pR] = d2tk_at(rR, Bk, sK, k, sRM[5][t], sRM[6][t], sRM[7][t], sRM[8][t], sRM[9][t]);
        }
        __syncthreads();

        // ---- phase C: limited fluxes and the explicit update of level k
        if (act) {
            const unsigned char *F = sFL[k & 1];
            const int r = rOwn, pp = pOwn;
            const double Tc = Pk[r], Te = Pk[r + 1], Tw = Pk[r - 1], Ts = Pk[r + RX], Tn = Pk[r - RX];
            const double Tu = Tprev, Td = (k + 1 < km) ? Pk1[r] : 0.0;
            const double *U = sU[k & 1], *V = sV[k & 1];
            const double Ue = U[r + 1], Uw = U[r], Vs = V[r], Vn = V[r - RX];
            const double Wk = sW[k % 3][r], Wb = (k + 1 < km) ? sW[(k + 1) % 3][r] : 0.0;
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
            const double tr = M.tr;
            dev::PTerms tc;
            dev::pterms_cheap(Tc, Te, Tw, Ts, Tn, Tu, Td, Ue, Uw, Vs, Vn, Wk, Wb, tr, k, km, g.odzp, tc);
            const double xy1 = up1 ? dev::up_xy1(Te, Tc, Ue, tr) : lc.xy1;
            const double xy2 = up2 ? dev::up_xy2(Tc, Tw, Uw, tr) : lc.xy2;
            const double xy3 = up3 ? dev::up_xy1(Ts, Tc, Vs, tr) : lc.xy3;
            const double xy4 = up4 ? dev::up_xy2(Tc, Tn, Vn, tr) : lc.xy4;
            double zb1, zb2;
            if (k == 0) {
                zb1 = 0.0;
                zb2 = (fc || fd) ? dev::up_zb2(Tc, Td, Wb, g.odzp[0]) : lc.zb2;
            } else if (k == km - 1) {
                zb1 = (fc || fprev) ? dev::up_zb1(Tu, Tc, Wk, g.odzp[k]) : lc.zb1;
                zb2 = 0.0;
            } else {
                zb1 = (fc || fprev) ? dev::up_zb1(Tu, Tc, Wk, g.odzp[k]) : lc.zb1;
                zb2 = (fc || fd) ? dev::up_zb2(Tc, Td, Wb, g.odzp[k]) : lc.zb2;
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
                    if (k <= sK[r - RX] - 1) { cf = M.dtn; hdtk += cf * X[xs - xw]; cc -= cf; }
                    if (k <= sK[r + RX] - 1) { cf = M.dts; hdtk += cf * X[xs + xw]; cc -= cf; }
                    if (k <= sK[r + 1] - 1) { cf = M.dte; hdtk += cf * X[xs + 1]; cc -= cf; }
                    if (k <= sK[r - 1] - 1) { cf = M.dtw; hdtk += cf * X[xs - 1]; cc -= cf; }
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
            if (rest && k >= 1) tf = tf + vk * (Rk - Bc) * lamda;
            // tracer_8: the explicit update entering the implicit solve
            a.ori[n * n3 + c] = Tc + dts * tf;
            Tprev = Tc;
            Bprev = Bc;
            fprev = fc;
        }
        Aprev = Acur;
        lc = ln;
        __syncthreads();
    }
}

// ---------------------------------------------------------------- solve
// Two lanes per column: lane parity = tracer; 32 columns along i per wave.
//
// invtrit (tracer.cpp:923-978) eliminates from the bottom and substitutes
// from the top.  Holding e(k), f(k) for the whole column would take 2*km
// doubles per lane; instead the elimination keeps them only at every SEG-th
// level, and the substitution recomputes each SEG-level segment from its
// checkpoint with the same operations (identical bits).  Both sweeps load
// the next segment's inputs while computing the current one.
constexpr int SEG = 8;

// convadj_cu (convadj.cpp:52-256) for one column, driven by the mask of
// statically unstable interfaces found with the unmixed densities (bit k:
// rholo(k-1) > rhoup(k)).  Every density the reference reads after mixing
// starts is either recomputed by it or belongs to a level that mixing has not
// touched, so the mask plus on-demand recomputation reproduces its decisions.
// The two lanes of a column run this together: each holds its own tracer's
// column X (the state just written, stride n2) and they swap values with
// __shfl_xor.  Returns the mixed level range in lo/hi.
__device__ __noinline__ void convect_sparse(const double *C, const double *to, const double *so,
                                            const double *dzp, int km, int kcon, uint64_t mask,
                                            int n, double *X, size_t n2, int &lo, int &hi) {
    lo = km;
    hi = -1;
    int lcon = __ffsll((unsigned long long)mask) - 1;
    for (;;) {   // CONV_1
        int lcona = lcon, lconb = lcon + 1;
        double dztsum = dzp[lcona - 1] + dzp[lconb - 1];
        double trasum = X[(size_t)(lcona - 1) * n2] * dzp[lcona - 1] + X[(size_t)(lconb - 1) * n2] * dzp[lconb - 1];
        double tramix = trasum / dztsum;
        for (;;) {   // CONV_2
            if (lconb != kcon) {
                const double v = X[(size_t)lconb * n2];
                const double vo = __shfl_xor(v, 1), mo = __shfl_xor(tramix, 1);
                const double Tm = n == 0 ? tramix : mo, Sm = n == 0 ? mo : tramix;
                const double Tv = n == 0 ? v : vo, Sv = n == 0 ? vo : v;
                const double rl = dev::dens(C, km, lconb, Tm - to[lconb], Sm - so[lconb]);
                const double ru = dev::dens(C, km, lconb, Tv - to[lconb], Sv - so[lconb]);
                if (rl > ru) {
                    lconb = lconb + 1;
                    dztsum = dztsum + dzp[lconb - 1];
                    trasum = trasum + v * dzp[lconb - 1];
                    tramix = trasum / dztsum;
                    continue;
                }
            }
            if (lcona > 1) {
                const double v = X[(size_t)(lcona - 2) * n2];
                const double vo = __shfl_xor(v, 1), mo = __shfl_xor(tramix, 1);
                const double Tm = n == 0 ? tramix : mo, Sm = n == 0 ? mo : tramix;
                const double Tv = n == 0 ? v : vo, Sv = n == 0 ? vo : v;
                const double rl = dev::dens(C, km, lcona - 1, Tv - to[lcona - 1], Sv - so[lcona - 1]);
                const double ru = dev::dens(C, km, lcona - 1, Tm - to[lcona - 1], Sm - so[lcona - 1]);
                if (rl > ru) {
                    lcona = lcona - 1;
                    dztsum = dztsum + dzp[lcona - 1];
                    trasum = trasum + v * dzp[lcona - 1];
                    tramix = trasum / dztsum;
                    continue;
                }
            }
            break;
        }
        for (int l = lcona - 1; l < lconb; ++l) X[(size_t)l * n2] = tramix;
        lo = min(lo, lcona - 1);
        hi = max(hi, lconb - 1);
        if (lconb == kcon) return;
        // CONV_3: next unstable interface below the mixed region
        const uint64_t below = lconb >= 63 ? 0ull : (mask & ~((2ull << lconb) - 1ull));
        if (below == 0) return;
        lcon = __ffsll((unsigned long long)below) - 1;
    }
}

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
    const double *O = ori + n * n3 + q;
    const double *A = akt + n * n3 + q;
    double *TT = at + n * n3 + q, *BB = atb + n * (n3 + n2) + n2 + q;
    const double *odzp = g.odzp, *odzt = g.odzt;
    const double dts = p.dts;
    const double stf = act ? stf_n[n * n2 + q] : 0.0;
    const int hmix = p.hmix;
    const double dwm = p.dwndmix;

    // ---- elimination from the bottom; checkpoints e,f at levels SEG*m (m >= 1)
    double cke[NSEG + 1], ckf[NSEG + 1];
#pragma unroll
    for (int m = 0; m <= NSEG; ++m) cke[m] = ckf[m] = 0.0;
    {
        double pd[SEG], pa[SEG];
#pragma unroll
        for (int r = 0; r < SEG; ++r) {
            const int k = SEG * (NSEG - 1) + r;
            pd[r] = pa[r] = 0.0;
            if (k < KM && k < kz) {
                pd[r] = O[(size_t)k * n2];
                pa[r] = k >= 1 ? A[(size_t)(k - 1) * n2] : 0.0;
            }
        }
        double en = 0.0, fn = 0.0, wk_k = 0.0;
#pragma unroll
        for (int m = NSEG - 1; m >= 0; --m) {
            double cd[SEG], ca[SEG];
#pragma unroll
            for (int r = 0; r < SEG; ++r) {
                cd[r] = pd[r];
                ca[r] = pa[r];
            }
            if (m > 0) {
#pragma unroll
                for (int r = 0; r < SEG; ++r) {
                    const int k = SEG * (m - 1) + r;
                    pd[r] = pa[r] = 0.0;
                    if (k < kz) {
                        pd[r] = O[(size_t)k * n2];
                        pa[r] = k >= 1 ? A[(size_t)(k - 1) * n2] : 0.0;
                    }
                }
            }
#pragma unroll
            for (int r = SEG - 1; r >= 0; --r) {
                const int k = SEG * m + r;
                if (k < KM && k < kz) {
                    const double wkm1 = k >= 1 ? dev::floor_akt(ca[r], k - 1, n, hmix, dwm) : 0.0;
                    double aa, bb, cc;
                    dev::tri_row(k, kz, wkm1, wk_k, odzp, odzt, dts, aa, bb, cc);
                    const double g0 = 1.0 / (bb - cc * en);
                    en = aa * g0;
                    fn = (cd[r] + cc * fn) * g0;
                    wk_k = wkm1;
                }
            }
            if (m > 0) {
                cke[m] = en;
                ckf[m] = fn;
            }
        }
    }

    // ---- substitution from the top, tracer_9, convadj stability scan
    const bool conv = flags & LICOM_STEP_CONVADJ;
    const double *C = g.eosc, *to = g.to, *so = g.so;
    double xprev = 0.0, Tp = 0.0, Sp = 0.0;
    uint64_t mask = 0;
    double qd[SEG], qa[SEG], qt[SEG], qw = 0.0;
    auto load_seg = [&](int m) {
#pragma unroll
        for (int r = 0; r < SEG; ++r) {
            const int k = SEG * m + r;
            qd[r] = qa[r] = qt[r] = 0.0;
            if (k < KM && k < km && act) {
                qd[r] = O[(size_t)k * n2];
                qt[r] = TT[(size_t)k * n2];
                if (k >= 1 && k < kz) qa[r] = A[(size_t)(k - 1) * n2];
            }
        }
        const int kb = SEG * m + SEG - 1;
        qw = (kb < KM && kb < kz - 1) ? A[(size_t)kb * n2] : 0.0;
    };
    load_seg(0);
#pragma unroll
    for (int m = 0; m < NSEG; ++m) {
        double cd[SEG], ca[SEG], ct[SEG];
#pragma unroll
        for (int r = 0; r < SEG; ++r) {
            cd[r] = qd[r];
            ca[r] = qa[r];
            ct[r] = qt[r];
        }
        const double cw = qw;
        if (m + 1 < NSEG) load_seg(m + 1);
        const int top = SEG * m;
        double se[SEG], sf[SEG];
        {
            double ee = cke[m + 1], ff = ckf[m + 1];
            double wk_k = dev::floor_akt(cw, top + SEG - 1, n, hmix, dwm);
#pragma unroll
            for (int r = SEG - 1; r >= 0; --r) {
                const int k = top + r;
                se[r] = sf[r] = 0.0;
                if (k < KM && k < kz) {
                    const double wkm1 = k >= 1 ? dev::floor_akt(ca[r], k - 1, n, hmix, dwm) : 0.0;
                    double aa, bb, cc;
                    dev::tri_row(k, kz, wkm1, wk_k, odzp, odzt, dts, aa, bb, cc);
                    const double g0 = 1.0 / (bb - cc * ee);
                    ee = aa * g0;
                    ff = (cd[r] + cc * ff) * g0;
                    se[r] = ee;
                    sf[r] = ff;
                    wk_k = wkm1;
                }
            }
        }
#pragma unroll
        for (int r = 0; r < SEG; ++r) {
            const int k = top + r;
            if (k < KM && k < km) {
                const size_t c = (size_t)k * n2;
                const bool wet = k < kz;
                const double o = cd[r];
                double x = o;
                if (wet) x = (k == 0) ? (se[0] * stf + sf[0]) * 1.0 : (se[r] * xprev + sf[r]) * 1.0;
                xprev = x;
                const double vk = wet ? 1.0 : 0.0;
                if (act) {
                    if (dt_diff) {
                        double v = (x - o) / dts * vk;
                        if (k == 0) v = v - stf * 0.5 * odzp[0];
                        dt_diff[n * n3 + q + c] = v;
                    }
                    if (tend) tend[n * n3 + q + c] = (x - ct[r]) / dts * vk;
                    if (conv && dt_conv) dt_conv[n * n3 + q + c] = 0.0;
                    TT[c] = x;
                    BB[c] = x;
                }
                if (conv) {
                    const double xo = __shfl_xor(x, 1);
                    const double Tk = n == 0 ? x : xo, Sk = n == 0 ? xo : x;
                    if (k >= 1 && k <= kz - 1) {
                        const double up = dev::dens(C, km, k, Tk - to[k], Sk - so[k]);
                        const double lo = dev::dens(C, km, k, Tp - to[k], Sp - so[k]);
                        if (lo > up) mask |= 1ull << k;
                    }
                    Tp = Tk;
                    Sp = Sk;
                }
            }
        }
    }

    // ---- convadj for statically unstable columns (both lanes of the pair)
    if (conv && act && mask != 0) {
        if (n == 0) atomicAdd(mix_count, 1);
        int lo, hi;
        convect_sparse(C, to, so, g.dzp, km, kz, mask, n, TT, n2, lo, hi);
        for (int l = lo; l <= hi; ++l) {
            const size_t c = (size_t)l * n2;
            const double y = TT[c], b = BB[c];
            if (y != b) {
                const double dc = (y - b) / dts * 1.0;
                if (dt_conv) dt_conv[n * n3 + q + c] = dc;
                if (tend) tend[n * n3 + q + c] += dc;
                BB[c] = y;
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
