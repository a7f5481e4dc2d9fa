// tracer_mixed.hip -- mixed-precision form of the tracer step on gfx950.
//
// The tracer state (at, atb), the surface fields, the implicit vertical solve,
// convadj and the SSSNORM sum stay FP64; velocities, ws, akt, the face
// velocities, the divergence and the 3-D diagnostics are FP32.  The k-march
// forms every difference of the tracer (neighbour minus the cell's own value)
// in FP64 and runs the tspas fluxes, the limiter test, the del4/del2 stencils
// and the explicit vertical diffusion in FP32 on those anomalies; its FP32
// tendency is added to the FP64 state (ori = at + dts * tf).  Written in this
// anomaly form, the sums of tspas (x0 + c1, y0 + c2, za + zc of
// tracer.cpp:311-414) reduce to flux differences with no cancellation of the
// O(10 K) tracer values, so FP32 keeps ~1e-7 relative accuracy on the
// increments.
//
//   k_prelude_mx<KM>              face velocities, divergence, continuity w
//   k_march_mx                    four-wave k-march (as the FP64 k_march)
//   k_solve<KM, float, float>     FP64 invtrit + convadj (tracer_solve.h)
//   k_side3<float4> / k_side2_mx  in-place side effects
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "tracer_device.h"
#include "tracer_internal.h"
#include "tracer_solve.h"
#include "tracer_surface.h"

namespace licom {
namespace {

constexpr int KMAXM = LICOM_MAX_KM;

__device__ __forceinline__ float wkf(float uf, float ul, float oncc) {
    return 0.5f * (uf * oncc) + 0.5f * ul;
}

// ------------------------------------------------------------- prelude
// Launch shapes of the prelude: a 256-lane workgroup covers FX x (256/FX)
// columns; shape 0: 64 x 4, 1: 32 x 8, 2: 128 x 2.
__host__ __device__ constexpr int fd_cols(int shape) { return shape == 0 ? 64 : (shape == 1 ? 32 : 128); }
__host__ __device__ constexpr int fd_rows(int shape) { return 256 / fd_cols(shape); }

// One-pass prelude of the mixed form (see k_prelude_col of the FP64 form):
// a block of FX x FY columns (fd_cols / fd_rows) marches k with the
// (FX+1) x (FY+1) window of wkd / wkb staged in LDS a level ahead, metrics
// loaded once per column, the divergence profile in registers, the column
// sum and the w recursion in FP64.
template <int KM>
__device__ __forceinline__ void prelude_col_mx_body(const KGrid &g, float oncc,
                                                    const float *__restrict__ utf,
                                                    const float *__restrict__ utl,
                                                    const float *__restrict__ vtf,
                                                    const float *__restrict__ vtl,
                                                    const double *__restrict__ work_u,
                                                    const double *__restrict__ stf_pre,
                                                    float *__restrict__ ws, float *__restrict__ uw,
                                                    float *__restrict__ vs, int bx, int by, int shape) {
    __shared__ float sD[2][400], sBv[2][400];
    // dzp from LDS: a global load inside the level loop would carry a vmcnt
    // wait that also waits for the level's stores and the next level's loads
    __shared__ double sDZ[KMAXM];
    const int FX = fd_cols(shape), FY = fd_rows(shape), WX = FX + 1, WN = WX * (FY + 1);
    const int t = threadIdx.x, tx = t % FX, ty = t / FX;
    const int i0 = bx * FX, j0 = by * FY;
    const int i = i0 + tx, j = j0 + ty;
    const int imt = g.imt, jmt = g.jmt, km = g.km;
    for (int r = t; r < km; r += 256) sDZ[r] = g.dzp[r];
    const size_t n2 = (size_t)imt * jmt;
    size_t wq[2];
    bool wv[2], wok[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int e = t + 256 * u, wr = e / WX, wc = e - wr * WX;
        const int jj = j0 - 1 + wr, ii = i0 + wc;
        wv[u] = e < WN;
        wok[u] = wv[u] && jj >= 0 && jj < jmt && ii < imt;
        wq[u] = wok[u] ? (size_t)jj * imt + ii : 0;
    }
    // window loads and the face-velocity stores go through plane buffers
    // (masked lanes read 0 / drop their store without a branch)
    uint32_t wo[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) wo[u] = buf_off(wok[u], wq[u], 4);
    const size_t pb = n2 * sizeof(float);
    float f_ut[2], f_ul[2], f_vt[2], f_vl[2];
    auto fetch = [&](int k) {
        const size_t o = (size_t)k * n2;
        const auto r_ut = plane_rsrc(utf + o, pb), r_ul = plane_rsrc(utl + o, pb);
        const auto r_vt = plane_rsrc(vtf + o, pb), r_vl = plane_rsrc(vtl + o, pb);
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            f_ut[u] = buf_load(r_ut, wo[u], 0.0f);
            f_ul[u] = buf_load(r_ul, wo[u], 0.0f);
            f_vt[u] = buf_load(r_vt, wo[u], 0.0f);
            f_vl[u] = buf_load(r_vl, wo[u], 0.0f);
        }
    };
    fetch(0);
    const bool in = i < imt && j < jmt;
    const size_t q = in ? (size_t)j * imt + i : 0;
    const bool do_u = in && j >= 1 && j < jmt - 1;
    const bool do_v = in && i >= 1 && i < imt - 1;
    const bool do_w = do_u && do_v;
    const float htw_c = in ? (float)g.htw[q] : 0.0f, hts_c = in ? (float)g.hts[q] : 0.0f;
    float fc = 0, fe = 0, fn = 0, fne = 0, htw_e = 0, hts_n = 0, tr = 0;
    double ohbt = 0, stfp = 0, prev = 0;
    int kmt = 0;
    if (do_w) {
        ohbt = g.ohbt[q];
        stfp = stf_pre[q];
        prev = (double)ws[q];
        const size_t qe = q + 1, qn = q - imt, qne = q - imt + 1;
        fc = (float)(1.0 + work_u[q] * g.ohbu[q]);
        fe = (float)(1.0 + work_u[qe] * g.ohbu[qe]);
        fn = (float)(1.0 + work_u[qn] * g.ohbu[qn]);
        fne = (float)(1.0 + work_u[qne] * g.ohbu[qne]);
        htw_e = (float)g.htw[qe];
        hts_n = (float)g.hts[qn];
        tr = (float)g.tarea_r[q];
        kmt = g.kmt[q];
    }
    const uint32_t ou = buf_off(do_u, q, 4), ov = buf_off(do_v, q, 4);
    // everything loaded so far completes here, so that no wait for it stays
    // inside the level loop
    wait_all_vmem();
    const int w0 = (ty + 1) * WX + tx;
    float wka[KM];
    double w = 0.0;
#pragma unroll
    for (int k = 0; k < KM; ++k) {
        wka[k] = 0.0f;
        if (k < km) {
            const int b = k & 1;
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                if (wv[u]) {
                    sD[b][t + 256 * u] = wkf(f_ut[u], f_ul[u], oncc);
                    sBv[b][t + 256 * u] = wkf(f_vt[u], f_vl[u], oncc);
                }
            }
            if (k + 1 < km) fetch(k + 1);
            __syncthreads();
            const float dc = sD[b][w0], bc = sBv[b][w0];
            buf_store(plane_rsrc(uw + (size_t)k * n2, pb), ou, (sD[b][w0 - WX] + dc) * htw_c * 0.25f);
            buf_store(plane_rsrc(vs + (size_t)k * n2, pb), ov, (bc + sBv[b][w0 + 1]) * hts_c * 0.25f);
            if (do_w && k <= kmt - 1) {
                const float de = sD[b][w0 + 1], dn = sD[b][w0 - WX], dne = sD[b][w0 - WX + 1];
                const float be = sBv[b][w0 + 1], bn = sBv[b][w0 - WX], bne = sBv[b][w0 - WX + 1];
                const float d = 0.5f *
                                ((fe * de + fne * dne) * htw_e - (fc * dc + fn * dn) * htw_c +
                                 (fe * be + fc * bc) * hts_c - (fne * bne + fn * bn) * hts_n) *
                                tr;
                wka[k] = d;
                w = w - sDZ[k] * (double)d;
            }
        }
    }
    if (!do_w) return;
    const double sc = 1.0 / (1.0 + stfp * ohbt);
#pragma unroll
    for (int k = 1; k < KM; ++k) {
        if (k < km) {
            const double cur = (k < kmt ? 1.0 : 0.0) * (prev + sDZ[k - 1] * (w * ohbt + (double)wka[k - 1]));
            ws[(size_t)k * n2 + q] = (float)(cur * sc);
            prev = cur;
        }
    }
}

template <int KM>
__global__ __launch_bounds__(256) void k_prelude_mx(KGrid g, float oncc, const float *__restrict__ utf,
                                                    const float *__restrict__ utl,
                                                    const float *__restrict__ vtf,
                                                    const float *__restrict__ vtl,
                                                    const double *__restrict__ work_u,
                                                    const double *__restrict__ stf_pre,
                                                    float *__restrict__ ws, float *__restrict__ uw,
                                                    float *__restrict__ vs, int shape, XcdGrid xg) {
    int bx, by;
    if (!xcd_block(xg, bx, by)) return;
    prelude_col_mx_body<KM>(g, oncc, utf, utl, vtf, vtl, work_u, stf_pre, ws, uw, vs, bx, by, shape);
}

// --------------------------------------------------------------- march
// A 32x6 tile: each of the three tile waves owns two whole 32-cell rows, so
// that the level's FP32 stores leave in 128-byte row pieces (the march is
// bound by its scattered stores, and the 64-byte pieces of a 16-wide tile
// cost a third more time). Wave 3 is the ring/staging wave: its lane l
// takes ring cell l (the north and south rows) and, for l < 12, ring cell
// 64 + l (the west and east columns). Workgroups of four waves: with five
// or six, residency fell to under two workgroups per CU.
constexpr int BX = 32, BY = 6;
constexpr int PX = BX + 2, PY = BY + 2;
constexpr int RX = BX + 4, RY = BY + 4;
constexpr int NOWN = BX * BY, NRING = 2 * (BX + BY), NT = NOWN + 64;
static_assert(NOWN == 192 && 2 * BX == 64 && NRING > 64 && NRING <= 128,
              "three tile waves; one ring wave with a second ring cell on its first lanes");
constexpr int NP = PX * PY, NR = RX * RY;
constexpr int NCH = NR / 2, NCI = (NCH + 63) / 64;   // 16-byte chunks of an FP64 window
constexpr int NCF = (NR + 63) / 64;                  // dword loads of an FP32 window
constexpr int NCA = NOWN / 64;                       // dword loads of the tile's akt
constexpr int MARCH_BLOCKS = 3;                      // resident workgroups per CU (LDS)
constexpr float BIG = 1.0e10f;
// Look-ahead of the staged planes (as in the FP64 k_march): at and ws are
// issued for level k+LA+2 at the start of P(k); faces, atb, akt for k+LA+1.
constexpr int LA = 2;
constexpr int NAT = LA + 3, NW = LA + 3, NUV = LA + 2, NB = LA + 2;
constexpr int DMA_ATW = NCI + NCF;                  // at (16-byte), ws (dword)
constexpr int DMA_UVBA = 2 * NCF + NCI + NCA;       // u, v (dword), atb (16-byte), akt (dword)

struct MarchArgsMx {
    const double *at, *atb, *stf_n, *err_norm2, *swv, *restore;
    const float *akt, *uw, *vs, *ws;
    double *ori;
    float *ax, *ay, *az, *dz, *pen;
    uint8_t *flags_out = nullptr;   // the limiter flag of every physical cell, or NULL
    int ntx, ntiles, tiles_per_xcd;
};

// Ring cell r (0..NRING-1) -> P-plane coordinates (pi, pj): north row, south
// row, west column, east column, corners excluded.
__device__ __forceinline__ int2 ring_cell(int r) {
    const bool ns = r < 2 * BX;
    const int rr = ns ? r : r - 2 * BX;
    const int pi = ns ? 1 + (rr % BX) : (rr < BY ? 0 : PX - 1);
    const int pj = ns ? (rr < BX ? 0 : PY - 1) : 1 + (rr % BY);
    return make_int2(pi, pj);
}

// FP32 metrics of a cell: tarea_r, the LW coefficients 2 dts tarea_r / D
// (D = htw*hun at i+1, i; hts*hue at j, j-1) and the del2 stencil.
struct MetricsF {
    float tr, cx1, cx2, cx3, cx4, dtn, dts, dte, dtw, ahf;
};

__device__ __forceinline__ MetricsF load_metrics_f(const KGrid &g, int gj, int gi, double dts) {
    const size_t q = (size_t)gj * g.imt + gi, sj = g.imt;
    const double tr = g.tarea_r[q];
    MetricsF m;
    m.tr = (float)tr;
    m.cx1 = (float)(2.0 * dts * tr / (g.htw[q + 1] * g.hun[q + 1]));
    m.cx2 = (float)(2.0 * dts * tr / (g.htw[q] * g.hun[q]));
    m.cx3 = (float)(2.0 * dts * tr / (g.hts[q] * g.hue[q]));
    m.cx4 = (float)(2.0 * dts * tr / (g.hts[q - sj] * g.hue[q - sj]));
    m.dtn = (float)g.dtn[q];
    m.dts = (float)g.dts[q];
    m.dte = (float)g.dte[q];
    m.dtw = (float)g.dtw[q];
    m.ahf = (float)g.ahf[q];
    return m;
}

struct LWf {
    float xy1, xy2, xy3, xy4, zb1, zb2;
};

struct LvM {
    const double *Pk, *Pu, *Pd;
    const float *U, *V, *Wk, *Wb;
};

// Anomalies of the six neighbours about the cell value (FP64 differences).
struct Anom {
    float e, w, s, n, u, d;
};

// The advective terms of advection_tracer_2 in anomaly form: xc = x0 + c1,
// yc = y0 + c2, zc = za + zc (tracer.cpp:311-414).
__device__ __forceinline__ void cheap_mx(const Anom &a, float Ue, float Uw, float Vs, float Vn,
                                         float Wk, float Wb, float tr, float odzpk, float &xc,
                                         float &yc, float &zc) {
    xc = (a.e * Ue - a.w * Uw) * tr;
    yc = (a.s * Vs - a.n * Vn) * tr;
    zc = 0.5f * odzpk * (Wk * a.u - Wb * a.d);
}

// Predictor and limiter flag of level K at R-plane index r (mixed form).
__device__ __forceinline__ bool pred_flag_mx(const KGrid &g, int r, const LvM &L, const int *sK,
                                             int K, const float *odzp, const float *odzt, float dts,
                                             const MetricsF &M, LWf &lw) {
    const int km = g.km;
    const double Tc = L.Pk[r];
    Anom a;
    a.e = (float)(L.Pk[r + 1] - Tc);
    a.w = (float)(L.Pk[r - 1] - Tc);
    a.s = (float)(L.Pk[r + RX] - Tc);
    a.n = (float)(L.Pk[r - RX] - Tc);
    a.u = K > 0 ? (float)(L.Pu[r] - Tc) : 0.0f;
    a.d = K < km - 1 ? (float)(L.Pd[r] - Tc) : 0.0f;
    const float Ue = L.U[r + 1], Uw = L.U[r], Vs = L.V[r], Vn = L.V[r - RX];
    const float Wk = K > 0 ? L.Wk[r] : 0.0f, Wb = K < km - 1 ? L.Wb[r] : 0.0f;
    const float odzpk = odzp[K];
    float xc, yc, zc;
    cheap_mx(a, Ue, Uw, Vs, Vn, Wk, Wb, M.tr, odzpk, xc, yc, zc);
    lw.xy1 = -(a.e * Ue * Ue) * M.cx1;
    lw.xy2 = -(a.w * Uw * Uw) * M.cx2;
    lw.xy3 = -(a.s * Vs * Vs) * M.cx3;
    lw.xy4 = -(a.n * Vn * Vn) * M.cx4;
    lw.zb1 = K > 0 ? -0.5f * odzpk * Wk * Wk * odzt[K] * a.u * dts : 0.0f;
    lw.zb2 = K < km - 1 ? -0.5f * odzpk * Wb * Wb * odzt[K + 1] * a.d * dts : 0.0f;
    const float xx = -(xc + lw.xy1 + lw.xy2);
    const float yy = -(yc + lw.xy3 + lw.xy4);
    const float zz = -(lw.zb1 + lw.zb2 + zc);
    const float d00 = (xx + yy + zz) * dts;
    // 7-point bounds about the cell value, land at -+1e10 (tracer.cpp:416-545)
    float mx = sK[r] > K ? 0.0f : -BIG, mn = sK[r] > K ? 0.0f : BIG;
    const int nb[4] = {r - RX, r + RX, r - 1, r + 1};
    const float tv[4] = {a.n, a.s, a.w, a.e};
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const bool wet = sK[nb[m]] > K;
        mx = fmaxf(mx, wet ? tv[m] : -BIG);
        mn = fminf(mn, wet ? tv[m] : BIG);
    }
    if (K > 0) {
        const bool wet = sK[r] > K - 1;
        mx = fmaxf(mx, wet ? a.u : -BIG);
        mn = fminf(mn, wet ? a.u : BIG);
    }
    if (K < km - 1) {
        const bool wet = sK[r] > K + 1;
        mx = fmaxf(mx, wet ? a.d : -BIG);
        mn = fminf(mn, wet ? a.d : BIG);
    }
    return d00 > mx || d00 < mn;
}

// del4 first pass, anomaly form: AHF * sum over wet neighbours of DT*(B_nbr - B).
__device__ __forceinline__ float d2tk_mx(int r, const double *B, const int *sK, int k,
                                         const MetricsF &M) {
    float h = 0.0f;
    if (k <= sK[r] - 1) {
        const double Bc = B[r];
        if (k <= sK[r - RX] - 1) h += M.dtn * (float)(B[r - RX] - Bc);
        if (k <= sK[r + RX] - 1) h += M.dts * (float)(B[r + RX] - Bc);
        if (k <= sK[r + 1] - 1) h += M.dte * (float)(B[r + 1] - Bc);
        if (k <= sK[r - 1] - 1) h += M.dtw * (float)(B[r - 1] - Bc);
    }
    return M.ahf * h;
}

#define LDS_PTR(p) ((__attribute__((address_space(3))) void *)(p))

// Direct global->LDS loads issued through inline assembly: the compiler
// then neither reorders them nor inserts vector-memory waits before this
// wave's later LDS reads (which touch other slots); the staging wave waits
// for them itself (wait_vmem) before the barrier that publishes the planes.
// M0 holds the wave's LDS destination; nothing else in these kernels uses M0.
__device__ __forceinline__ uint32_t lds_addr(const void *p) {
    return __builtin_amdgcn_readfirstlane((uint32_t)(uintptr_t)LDS_PTR(p));
}
__device__ __forceinline__ void dma16(const void *g, uint32_t lds) {
    asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off" ::"s"(lds), "v"(g)
                 : "memory");
}
__device__ __forceinline__ void dma4(const void *g, uint32_t lds) {
    asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dword %1, off" ::"s"(lds), "v"(g)
                 : "memory");
}

__device__ __forceinline__ void stage_d(const double *src, double *dst, const int (&off)[NCI],
                                        int lane) {
#pragma unroll
    for (int u = 0; u < NCI; ++u)
        if (u * 64 + lane < NCH)
            dma16(src + off[u], lds_addr(dst + 128 * u));
}

__device__ __forceinline__ void stage_f(const float *src, float *dst, const int (&off)[NCF],
                                        int lane) {
#pragma unroll
    for (int u = 0; u < NCF; ++u)
        if (u * 64 + lane < NR)
            dma4(src + off[u], lds_addr(dst + 64 * u));
}

// akt of the tile (read only at the tile's own cells): NOWN floats in NCA
// dword instructions; element c of the slot is tile lane c.
__device__ __forceinline__ void stage_tile_f(const float *src, float *dst, const int (&off)[NCA],
                                             int lane) {
#pragma unroll
    for (int u = 0; u < NCA; ++u) dma4(src + off[u], lds_addr(dst + 64 * u));
}

__device__ __forceinline__ void wait_vmem() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }
template <int N>
__device__ __forceinline__ void wait_vmem_n() {
    static_assert(N >= 0 && N <= 63, "vmcnt range");
    asm volatile("s_waitcnt vmcnt(%0)" ::"I"(N) : "memory");
}
// Waits for every staged load except the `issued` issued at the start of
// the current level (direct-to-LDS loads complete in order): DMA_ATW +
// DMA_UVBA, DMA_UVBA or none.
__device__ __forceinline__ void wait_vmem_keep(int issued) {
    if (issued == DMA_ATW + DMA_UVBA) wait_vmem_n<DMA_ATW + DMA_UVBA>();
    else if (issued == DMA_UVBA) wait_vmem_n<DMA_UVBA>();
    else wait_vmem_n<0>();
}
__device__ __forceinline__ void bar_lds() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

__device__ __forceinline__ void march_body(const KGrid &g, const KPar &p, const MarchArgsMx &a, int tile,
                                           int n) {
    __shared__ __align__(16) double sAT[NAT][NR];
    __shared__ __align__(16) double sB[NB][NR];
    __shared__ __align__(16) float sU[NUV][NR];
    __shared__ __align__(16) float sV[NUV][NR];
    __shared__ __align__(16) float sW[NW][NR];
    __shared__ float sD2[NP];
    __shared__ unsigned char sFL[2][NP];
    __shared__ int sK[NR];
    // per-level constants and the tile's akt come from LDS, so that no
    // vector-memory load in the level loop carries a vmcnt wait (on gfx9
    // vmcnt also covers the staged planes and the level's stores)
    __shared__ float sZP[KMAXM], sZT[KMAXM], sPN[KMAXM];
    __shared__ __align__(16) float sAK[NB][NOWN];

    const int imt = g.imt, jmt = g.jmt, km = g.km;
    const size_t n2 = (size_t)imt * jmt, n3 = n2 * km;
    const int gi0 = 2 + (tile % a.ntx) * BX, gj0 = 2 + (tile / a.ntx) * BY;
    const long e0 = (long)(gj0 - 2) * imt + (gi0 - 2);

    const int t = threadIdx.x, lane = t & 63;
    // wave 3 (lanes NOWN..NT-1) is the ring/staging wave; the test goes
    // through a scalar register so that the two roles compile to two
    // separate loops behind a uniform branch (with a per-lane test they are
    // merged into one exec-masked loop, and the tile waves' loads and stores
    // then carry vmcnt waits into the staging wave's code)
    const bool ringw = __builtin_amdgcn_readfirstlane(t >> 6) == NOWN / 64;
    const int2 rc = ring_cell(ringw ? t - NOWN : 0);   // ring cell `lane` (< 64 < NRING)
    const int pi = ringw ? rc.x : 1 + t % BX, pj = ringw ? rc.y : 1 + t / BX;
    const int gi = gi0 - 1 + pi, gj = gj0 - 1 + pj;
    const int rC = (pj + 1) * RX + (pi + 1), pC = pj * PX + pi;
    const bool cell_ok = gj >= 1 && gj < jmt - 1 && gi >= 1 && gi < imt - 1;
    const bool act = !ringw && gi < imt - 2 && gj < jmt - 2;
    const size_t q = cell_ok ? (size_t)gj * imt + gi : 0;

    const double *T = a.at + n * n3;
    const double *TB = a.atb + n * (n3 + n2) + n2;
    const float *A = a.akt + n * n3;
    const double dts = p.dts;
    const float dtsf = (float)dts;

    for (int r = t; r < NR; r += NT) {
        const int rj = gj0 - 2 + r / RX, ri = gi0 - 2 + r % RX;
        sK[r] = (rj < jmt && ri < imt) ? g.kmt[(size_t)rj * imt + ri] : 0;
    }
    for (int r = t; r < km; r += NT) {
        sZP[r] = (float)g.odzp[r];
        sZT[r] = (float)g.odzt[r];
        sPN[r] = r < km - 1 ? (float)g.pen[r] : 0.0f;
    }
    MetricsF M{};
    if (cell_ok) M = load_metrics_f(g, gj, gi, dts);

    if (ringw) {
        int offd[NCI], offf[NCF];
#pragma unroll
        for (int u = 0; u < NCI; ++u) {
            const int c = min(u * 64 + lane, NCH - 1), row = c / (RX / 2), cp = c - row * (RX / 2);
            offd[u] = (int)min(e0 + (long)row * imt + 2 * cp, (long)n2 - 2);
        }
#pragma unroll
        for (int u = 0; u < NCF; ++u) {
            const int c = min(u * 64 + lane, NR - 1), row = c / RX, col = c - row * RX;
            offf[u] = (int)min(e0 + (long)row * imt + col, (long)n2 - 1);
        }
        int offa[NCA];
#pragma unroll
        for (int u = 0; u < NCA; ++u) {
            const int c = u * 64 + lane, row = c / BX, col = c - row * BX;
            offa[u] = (int)min((long)(gj0 + row) * imt + gi0 + col, (long)n2 - 1);
        }
        for (int l = 0; l < NAT - 1 && l < km; ++l) {
            stage_d(T + (size_t)l * n2, sAT[l], offd, lane);
            stage_f(a.ws + (size_t)l * n2, sW[l], offf, lane);
        }
        for (int l = 0; l < NUV - 1 && l < km; ++l) {
            stage_f(a.uw + (size_t)l * n2, sU[l], offf, lane);
            stage_f(a.vs + (size_t)l * n2, sV[l], offf, lane);
            stage_d(TB + (size_t)l * n2, sB[l], offd, lane);
            stage_tile_f(A + (size_t)l * n2, sAK[l], offa, lane);
        }
        // the second ring cell of lanes 0..NRING-65 (west and east columns)
        const bool hasB = lane < NRING - 64;
        const int2 rb = ring_cell(hasB ? 64 + lane : 64);
        const int giB = gi0 - 1 + rb.x, gjB = gj0 - 1 + rb.y;
        const int rCB = (rb.y + 1) * RX + (rb.x + 1), pCB = rb.y * PX + rb.x;
        const bool okB = hasB && gjB >= 1 && gjB < jmt - 1 && giB >= 1 && giB < imt - 1;
        // their metrics live in LDS: in registers they push the persistent
        // ensemble march over its register budget
        __shared__ MetricsF sMB[NRING - 64];
        const MetricsF &MB = sMB[min(lane, NRING - 65)];
        if (okB) sMB[lane] = load_metrics_f(g, gjB, giB, dts);
        wait_vmem();
        bar_lds();
        {
            LWf tmp;
            const LvM L0{sAT[0], sAT[0], sAT[1], sU[0], sV[0], sW[0], sW[1]};
            const bool f0 = cell_ok ? pred_flag_mx(g, rC, L0, sK, 0, sZP, sZT, dtsf, M, tmp) : false;
            sFL[0][pC] = f0;
            if (hasB) sFL[0][pCB] = okB ? pred_flag_mx(g, rCB, L0, sK, 0, sZP, sZT, dtsf, MB, tmp) : false;
        }
        bar_lds();
        for (int k = 0; k < km; ++k) {
            const int s0 = k % NAT, s1 = (k + 1) % NAT, s2 = (k + 2) % NAT, s3 = (k + LA + 2) % NAT;
            const int u0 = k % NUV, u1 = (k + 1) % NUV, u2 = (k + LA + 1) % NUV;
            int issued = 0;
            if (k + LA + 2 < km) {
                stage_d(T + (size_t)(k + LA + 2) * n2, sAT[s3], offd, lane);
                stage_f(a.ws + (size_t)(k + LA + 2) * n2, sW[s3], offf, lane);
                issued += DMA_ATW;
            }
            if (k + LA + 1 < km) {
                stage_f(a.uw + (size_t)(k + LA + 1) * n2, sU[u2], offf, lane);
                stage_f(a.vs + (size_t)(k + LA + 1) * n2, sV[u2], offf, lane);
                stage_d(TB + (size_t)(k + LA + 1) * n2, sB[u2], offd, lane);
                stage_tile_f(A + (size_t)(k + LA + 1) * n2, sAK[u2], offa, lane);
                issued += DMA_UVBA;
            }
            const LvM L1{sAT[s1], sAT[s0], sAT[s2], sU[u1], sV[u1], sW[s1], sW[s2]};
            if (k + 1 < km) {
                LWf tmp;
                sFL[(k + 1) & 1][pC] = cell_ok ? pred_flag_mx(g, rC, L1, sK, k + 1, sZP, sZT, dtsf, M, tmp) : false;
            }
            if (p.hmix == LICOM_HMIX_DEL4) {
                sD2[pC] = cell_ok ? d2tk_mx(rC, sB[u0], sK, k, M) : 0.0f;
                if (hasB) sD2[pCB] = okB ? d2tk_mx(rCB, sB[u0], sK, k, MB) : 0.0f;
            }
            bar_lds();
            // C(k): the column cells' flags of level k+1, first read in
            // C(k+1); the planes they read (at k..k+2, faces and ws of k+1)
            // are complete and not restaged before the end of level k
            if (k + 1 < km && hasB) {
                LWf tmp;
                sFL[(k + 1) & 1][pCB] = okB ? pred_flag_mx(g, rCB, L1, sK, k + 1, sZP, sZT, dtsf, MB, tmp) : false;
            }
            wait_vmem_keep(issued);
            bar_lds();
        }
        return;
    }

    const double sw = (n == 0 && act && a.swv) ? a.swv[q] : 0.0;
    const int kmt_own = act ? g.kmt[q] : 0;
    const double stf_n = act ? a.stf_n[n * n2 + q] : 0.0;
    const double enorm = (n == 1 && p.sssnorm && act) ? a.err_norm2[0] : 0.0;
    const bool rest = p.boundary_restore == 1 && a.restore != nullptr;
    const double odzp0 = g.odzp[0];
    // the per-cell loads above complete here, before the level loop, so
    // that no wait for them (which would also wait for the level's stores)
    // stays inside it
    __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0)
    bar_lds();
    LWf lc{}, ln{};
    {
        const LvM L0{sAT[0], sAT[0], sAT[1], sU[0], sV[0], sW[0], sW[1]};
        sFL[0][pC] = cell_ok ? pred_flag_mx(g, rC, L0, sK, 0, sZP, sZT, dtsf, M, lc) : false;
    }
    bar_lds();

    double Tprev = 0.0, Bprev = 0.0;
    float Aprev = 0.0f;
    bool fprev = false;
    const double lamda = 1.0 / (15.0 * 86400.0);
    const float ah = (float)p.ah;

    for (int k = 0; k < km; ++k) {
        const int s0 = k % NAT, s1 = (k + 1) % NAT, s2 = (k + 2) % NAT;
        const int u0 = k % NUV, u1 = (k + 1) % NUV;
        const float Acur = act ? (float)dev::floor_akt((double)sAK[u0][t], k, n, p.hmix, p.dwndmix) : 0.0f;
        if (k + 1 < km) {
            const LvM L1{sAT[s1], sAT[s0], sAT[s2], sU[u1], sV[u1], sW[s1], sW[s2]};
            sFL[(k + 1) & 1][pC] = cell_ok ? pred_flag_mx(g, rC, L1, sK, k + 1, sZP, sZT, dtsf, M, ln) : false;
        }
        if (p.hmix == LICOM_HMIX_DEL4) sD2[pC] = cell_ok ? d2tk_mx(rC, sB[u0], sK, k, M) : 0.0f;
        bar_lds();

        if (act) {
            const double *Pk = sAT[s0], *Pk1 = sAT[s1];
            const double *Bk = sB[u0], *Bk1 = sB[u1];
            const unsigned char *F = sFL[k & 1];
            const int r = rC, pp = pC;
            const double Tc = Pk[r];
            Anom an;
            an.e = (float)(Pk[r + 1] - Tc);
            an.w = (float)(Pk[r - 1] - Tc);
            an.s = (float)(Pk[r + RX] - Tc);
            an.n = (float)(Pk[r - RX] - Tc);
            an.u = k > 0 ? (float)(Tprev - Tc) : 0.0f;
            an.d = k + 1 < km ? (float)(Pk1[r] - Tc) : 0.0f;
            const float *U = sU[u0], *V = sV[u0];
            const float Ue = U[r + 1], Uw = U[r], Vs = V[r], Vn = V[r - RX];
            const float Wk = k > 0 ? sW[s0][r] : 0.0f, Wb = (k + 1 < km) ? sW[s1][r] : 0.0f;
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
            const float tr = M.tr;
            const float odzpk = sZP[k];
            float xc, yc, zc;
            cheap_mx(an, Ue, Uw, Vs, Vn, Wk, Wb, tr, odzpk, xc, yc, zc);
            const float xy1 = up1 ? -an.e * fabsf(Ue) * tr : lc.xy1;
            const float xy2 = up2 ? -an.w * fabsf(Uw) * tr : lc.xy2;
            const float xy3 = up3 ? -an.s * fabsf(Vs) * tr : lc.xy3;
            const float xy4 = up4 ? -an.n * fabsf(Vn) * tr : lc.xy4;
            float zb1 = 0.0f, zb2 = 0.0f;
            if (k > 0) zb1 = (fc || fprev) ? -0.5f * fabsf(Wk) * odzpk * an.u : lc.zb1;
            if (k + 1 < km) zb2 = (fc || fd) ? -0.5f * fabsf(Wb) * odzpk * an.d : lc.zb2;
            const float xx = -(xc + xy1 + xy2);
            const float yy = -(yc + xy3 + xy4);
            const float zz = -(zc + zb1 + zb2);
            const size_t c = (size_t)k * n2 + q;
            if (a.ax) a.ax[n * n3 + c] = xx;
            if (a.ay) a.ay[n * n3 + c] = yy;
            if (a.az) a.az[n * n3 + c] = zz;
            const float vk = k < kmt_own ? 1.0f : 0.0f;
            float tf = (xx + yy + zz) * vk;
            // horizontal mixing (tracer_del4_2 / tracer_4)
            {
                float h = 0.0f;
                if (p.hmix == LICOM_HMIX_DEL4) {
                    if (k <= sK[r] - 1) {
                        float cc = 0.0f;
                        if (k <= sK[r - RX] - 1) { h += M.dtn * sD2[pp - PX]; cc -= M.dtn; }
                        if (k <= sK[r + RX] - 1) { h += M.dts * sD2[pp + PX]; cc -= M.dts; }
                        if (k <= sK[r + 1] - 1) { h += M.dte * sD2[pp + 1]; cc -= M.dte; }
                        if (k <= sK[r - 1] - 1) { h += M.dtw * sD2[pp - 1]; cc -= M.dtw; }
                        h = cc * sD2[pp] + h;
                    }
                } else if (k <= sK[r] - 1) {
                    const double Bc0 = Bk[r];
                    if (k <= sK[r - RX] - 1) h += M.dtn * (float)(Bk[r - RX] - Bc0);
                    if (k <= sK[r + RX] - 1) h += M.dts * (float)(Bk[r + RX] - Bc0);
                    if (k <= sK[r + 1] - 1) h += M.dte * (float)(Bk[r + 1] - Bc0);
                    if (k <= sK[r - 1] - 1) h += M.dtw * (float)(Bk[r - 1] - Bc0);
                }
                tf += ah * h;
            }
            // shortwave penetration (T) and explicit vertical diffusion
            const float v1 = (k + 1 < km && k + 1 < kmt_own) ? 1.0f : 0.0f;
            if (n == 0) {
                float pv;
                const float swf = (float)sw;
                if (k == 0) {
                    pv = -swf * sPN[0] * (1 < kmt_own ? 1.0f : 0.0f) * odzpk;
                } else if (k == km - 1) {
                    pv = swf * sPN[km - 2] * vk * odzpk;
                } else {
                    pv = (swf * sPN[k - 1] * vk - swf * sPN[k] * v1) * odzpk;
                }
                tf += pv;
                if (a.pen) a.pen[c] = pv;
            }
            const double Bc = Bk[r];
            float dzv;
            if (k == 0) {
                dzv = -odzpk * Acur * (float)(Bc - Bk1[r]) * sZT[1] * v1 * 0.5f;
            } else if (k == km - 1) {
                dzv = odzpk * Aprev * (float)(Bprev - Bc) * sZT[km - 1] * vk * 0.5f;
            } else {
                const float wt1 = Aprev * (float)(Bprev - Bc) * sZT[k] * vk;
                const float wt2 = Acur * (float)(Bc - Bk1[r]) * sZT[k + 1] * v1;
                dzv = odzpk * (wt1 - wt2) * 0.5f;
            }
            tf += dzv;
            if (a.dz) a.dz[n * n3 + c] = dzv;
            if (a.flags_out) a.flags_out[n * n3 + c] = fc;
            if (k == 0 && kmt_own > 0) {
                tf += (float)(stf_n * 0.5 * odzp0);
                if (n == 1 && p.sssnorm) tf += (float)enorm * vk;
            } else if (k == 0 && n == 1 && p.sssnorm) {
                tf += (float)enorm * vk;
            }
            if (rest && k >= 1) {
                const double Rk = a.restore[n * n3 + c];
                tf += vk * (float)((Rk - Bc) * lamda);
            }
            a.ori[n * n3 + c] = Tc + dts * (double)tf;
            Tprev = Tc;
            Bprev = Bc;
            fprev = fc;
        }
        Aprev = Acur;
        lc = ln;
        bar_lds();
    }
}

__global__ __launch_bounds__(NT, MARCH_BLOCKS) void k_march_mx(KGrid g, KPar p, MarchArgsMx a) {
    // XCD-aware placement: consecutive blocks of one XCD take the T and S
    // halves of one tile, then the next tile of the same row of tiles.
    const int Lb = blockIdx.x, xcd = Lb & 7, slot = Lb >> 3;
    const int tile = xcd * a.tiles_per_xcd + (slot >> 1);
    if (tile >= a.ntiles) return;
    march_body(g, p, a, tile, slot & 1);
}

// --------------------------------------------------------------- side
__global__ __launch_bounds__(256) void k_side3_mx(float4 *__restrict__ utf, float4 *__restrict__ vtf,
                                                  size_t n, float oncc) {
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    float4 a = utf[t], b = vtf[t];
    a.x *= oncc; a.y *= oncc; a.z *= oncc; a.w *= oncc;
    b.x *= oncc; b.y *= oncc; b.z *= oncc; b.w *= oncc;
    utf[t] = a;
    vtf[t] = b;
}

__global__ __launch_bounds__(256) void k_side1_mx(float *__restrict__ utf, float *__restrict__ vtf,
                                                  size_t n, float oncc) {
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    utf[t] *= oncc;
    vtf[t] *= oncc;
}

__global__ __launch_bounds__(256) void k_side2_mx(KGrid g, KPar p, double *h0f, float *akt) {
    const size_t n2 = (size_t)g.imt * g.jmt, n3 = n2 * g.km;
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= n2) return;
    h0f[t] = h0f[t] * p.onbc;
    const int j = (int)(t / g.imt), i = (int)(t % g.imt);
    if (i >= 1 && i < g.imt - 1 && j >= 1 && j < g.jmt - 1)
        for (int n = 0; n < 2; ++n)
            akt[n * n3 + t] = (float)dev::floor_akt((double)akt[n * n3 + t], 0, n, p.hmix, p.dwndmix);
}

template <int KM>
void launch_prelude_solve_mx(licom_ctx *ctx, const licom_fields_mixed *f, const licom_diag_mixed *d,
                             int flags, bool prelude) {
    const KGrid &g = ctx->g;
    const KPar &p = ctx->kp;
    Scratch &s = ctx->s;
    hipStream_t st = ctx->stream;
    float *uw = reinterpret_cast<float *>(s.uw), *vs = reinterpret_cast<float *>(s.vs);
    if (prelude) {
        Timer t(ctx, "k_prelude");
        const XcdGrid xg = xcd_grid((g.imt + 63) / 64, (g.jmt + 3) / 4);
        k_prelude_mx<KM><<<xcd_blocks(xg), 256, 0, st>>>(g, (float)p.oncc, f->utf, f->utl, f->vtf, f->vtl,
                                                         s.work_u, s.stf_pre, f->ws, uw, vs, 0, xg);
    } else {
        Timer t(ctx, "k_solve");
        const dim3 gr = solve_grid(g.imt, g.jmt, ctx->solve_shape);
        if (flags & LICOM_STEP_ICESNOW)
            k_solve_ice<KM, float, float><<<gr, 256, 0, st>>>(g, p, flags, f->at, f->atb, f->akt, s.ori, s.stf_n,
                                                  d ? d->tend : nullptr, d ? d->dt_diff : nullptr,
                                                  d ? d->dt_conv : nullptr, s.mix_count, ctx->solve_shape,
                                                  ctx->qice);
        else
            k_solve<KM, float, float><<<gr, 256, 0, st>>>(g, p, flags, f->at, f->atb, f->akt, s.ori, s.stf_n,
                                              d ? d->tend : nullptr, d ? d->dt_diff : nullptr,
                                              d ? d->dt_conv : nullptr, s.mix_count, ctx->solve_shape);
    }
}

void launch_column_mx(licom_ctx *ctx, const licom_fields_mixed *f, const licom_diag_mixed *d, int flags,
                      bool prelude) {
    switch (ctx->km) {
        case 30: launch_prelude_solve_mx<30>(ctx, f, d, flags, prelude); break;
        case 55: launch_prelude_solve_mx<55>(ctx, f, d, flags, prelude); break;
        default: launch_prelude_solve_mx<KMAXM>(ctx, f, d, flags, prelude); break;
    }
}

}  // namespace

int launch_step_mixed(licom_ctx *ctx, const licom_fields_mixed *f, const licom_diag_mixed *d, int flags) {
    const KGrid &g = ctx->g;
    Scratch &s = ctx->s;
    hipStream_t st = ctx->stream;
    // the FP64 surface, SSSNORM and halo kernels only touch FP64 fields
    licom_fields f64{};
    f64.at = f->at;
    f64.atb = f->atb;
    f64.h0f = f->h0f;
    f64.h0l = f->h0l;
    f64.swv = f->swv;
    f64.nswv = f->nswv;
    f64.sst = f->sst;
    f64.seaice = f->seaice;
    f64.sss = f->sss;
    f64.fresh = f->fresh;
    f64.restore = f->restore;
    licom_diag d64{};
    if (d) {
        d64.net = d->net;
        d64.fw_norm2 = d->fw_norm2;
    }
    if (int rc = launch_surface(ctx, &f64, &d64, ctx->s)) return rc;
    launch_column_mx(ctx, f, d, flags, true);
    {
        Timer t(ctx, "k_march");
        MarchArgsMx a{};
        a.at = f->at;
        a.atb = f->atb;
        a.akt = f->akt;
        a.uw = reinterpret_cast<const float *>(s.uw);
        a.vs = reinterpret_cast<const float *>(s.vs);
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
        a.flags_out = ctx->flags_out;
        a.ntx = (g.imt - 4 + BX - 1) / BX;
        const int nty = (g.jmt - 4 + BY - 1) / BY;
        a.ntiles = a.ntx * nty;
        a.tiles_per_xcd = (a.ntiles + 7) / 8;
        k_march_mx<<<16 * a.tiles_per_xcd, NT, 0, st>>>(g, ctx->kp, a);
    }
    launch_column_mx(ctx, f, d, flags, false);
    if (flags & LICOM_STEP_SIDE_EFFECTS) {
        Timer t(ctx, "k_side");
        const size_t n3 = ctx->n3;
        const float oncc = (float)ctx->kp.oncc;
        const bool vec = ((reinterpret_cast<uintptr_t>(f->utf) | reinterpret_cast<uintptr_t>(f->vtf)) & 15) == 0 &&
                         n3 % 4 == 0;
        if (vec) {
            const size_t n = n3 / 4;
            k_side3_mx<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(
                reinterpret_cast<float4 *>(f->utf), reinterpret_cast<float4 *>(f->vtf), n, oncc);
        } else {
            k_side1_mx<<<(unsigned)((n3 + 255) / 256), 256, 0, st>>>(f->utf, f->vtf, n3, oncc);
        }
        k_side2_mx<<<(unsigned)((ctx->n2 + 255) / 256), 256, 0, st>>>(g, ctx->kp, f->h0f, f->akt);
    }
    return launch_halo(ctx, &f64, flags & ~LICOM_STEP_VTL_SCRATCH);
}

// ================================================================ ensemble
namespace {

// One member of a batch: its grid, parameters, fields, diagnostics, scratch
// and launch decomposition, read by every batched kernel.
struct MemberDev {
    KGrid g;
    KPar p;
    double *at, *atb, *h0f;
    const double *h0l, *swv, *nswv, *sst, *seaice, *sss, *fresh, *restore;
    float *akt, *utf, *vtf;
    const float *utl, *vtl;
    float *ws;
    float *ax, *ay, *az, *dz, *tend, *dt_diff, *dt_conv, *pen;
    double *net, *fw_norm2;
    double *stf_pre, *work_u, *stf_n, *temp11, *partials, *err_norm2, *ori;
    float *uw, *vs;
    int *mix_count;
    double *qice;                     // licomqice of the fused icesnow, or NULL
    double area_t;
    int nblk;
    int sf_bx, sf_by;                 // 64 x 4 blocks of the 2-D kernels
    int fd_shape, fd_bx, fd_by;       // prelude column blocks
    int sv_shape, sv_bx, sv_by;       // column-solve blocks
    int ntx, ntiles;                  // BX x BY k-march tiles
    int side_vec, side_blocks, side2_blocks;
    int halo_bx;                      // halo blocks per plane (4 km planes + licomqice)
};

enum { ST_SURF, ST_SSSP, ST_SSSF, ST_FD, ST_WC /* unused */, ST_MARCH, ST_SOLVE, ST_SIDE3, ST_SIDE2, ST_HALO,
       NSTAGE };

// Member owning global block b of a stage (starts: nm+1 prefix sums).
__device__ __forceinline__ int find_member(const int *st, int nm, int b) {
    int lo = 0, hi = nm - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (st[mid] <= b) lo = mid; else hi = mid - 1;
    }
    return lo;
}

#define LICOM_MEMBER(stage)                                        \
    const int *st_ = starts + (stage) * (nm + 1);                  \
    const int m_ = find_member(st_, nm, (int)blockIdx.x);          \
    const MemberDev &d = M[m_];                                    \
    const int l = (int)blockIdx.x - st_[m_];

__global__ __launch_bounds__(256) void kb_surface(const MemberDev *M, const int *starts, int nm) {
    LICOM_MEMBER(ST_SURF)
    const SurfaceIO o{d.h0f, d.h0l, d.atb, d.swv, d.nswv, d.sst, d.seaice, d.sss, d.fresh,
                      d.stf_pre, d.work_u, d.stf_n, d.temp11, d.net, d.mix_count};
    const int bx = l % d.sf_bx, by = l / d.sf_bx;
    surface_cell(d.g, d.p, o, bx * 64 + (int)threadIdx.x % 64, by * 4 + (int)threadIdx.x / 64);
}

__global__ __launch_bounds__(kRedThreads) void kb_sss_partial(const MemberDev *M, const int *starts, int nm) {
    LICOM_MEMBER(ST_SSSP)
    sss_partial_block(d.g, d.temp11, d.partials, l, d.nblk);
}

__global__ __launch_bounds__(kRedThreads) void kb_sss_final(const MemberDev *M, const int *starts, int nm) {
    LICOM_MEMBER(ST_SSSF)
    (void)l;
    sss_final_block(d.partials, d.nblk, d.area_t, d.p.sssnorm, d.err_norm2, d.fw_norm2);
}

template <int KM>
__global__ __launch_bounds__(256) void kb_prelude(const MemberDev *M, const int *starts, int nm) {
    LICOM_MEMBER(ST_FD)
    prelude_col_mx_body<KM>(d.g, (float)d.p.oncc, d.utf, d.utl, d.vtf, d.vtl, d.work_u, d.stf_pre, d.ws,
                            d.uw, d.vs, l % d.fd_bx, l / d.fd_bx, d.fd_shape);
}

// Persistent k-march: each workgroup walks the (member, tile, tracer) list.
__global__ __launch_bounds__(NT, MARCH_BLOCKS) void kb_march(const MemberDev *M, const int *starts, int nm,
                                                  int total) {
    const int *st_ = starts + ST_MARCH * (nm + 1);
    for (int it = blockIdx.x; it < total; it += gridDim.x) {
        const int m = find_member(st_, nm, it);
        const MemberDev &d = M[m];
        const int l = it - st_[m];
        MarchArgsMx a;
        a.at = d.at;
        a.atb = d.atb;
        a.stf_n = d.stf_n;
        a.err_norm2 = d.err_norm2;
        a.swv = d.swv;
        a.restore = d.restore;
        a.akt = d.akt;
        a.uw = d.uw;
        a.vs = d.vs;
        a.ws = d.ws;
        a.ori = d.ori;
        a.ax = d.ax;
        a.ay = d.ay;
        a.az = d.az;
        a.dz = d.dz;
        a.pen = d.pen;
        a.ntx = d.ntx;
        a.ntiles = d.ntiles;
        a.tiles_per_xcd = 0;
        march_body(d.g, d.p, a, l >> 1, l & 1);
    }
}

template <int KM>
__global__ __launch_bounds__(256) void kb_solve(const MemberDev *M, const int *starts, int nm, int flags) {
    LICOM_MEMBER(ST_SOLVE)
    solve_body<KM, float, float>(d.g, d.p, flags, d.at, d.atb, d.akt, d.ori, d.stf_n, d.tend,
                                 d.dt_diff, d.dt_conv, d.mix_count, l % d.sv_bx, l / d.sv_bx,
                                 d.sv_shape, nullptr);
}

// The same with icesnow between the update and convadj (LICOM_STEP_ICESNOW).
template <int KM>
__global__ __launch_bounds__(256, 2) void kb_solve_ice(const MemberDev *M, const int *starts, int nm,
                                                    int flags) {
    LICOM_MEMBER(ST_SOLVE)
    solve_body<KM, float, float, true>(d.g, d.p, flags, d.at, d.atb, d.akt, d.ori, d.stf_n, d.tend,
                                       d.dt_diff, d.dt_conv, d.mix_count, l % d.sv_bx, l / d.sv_bx,
                                       d.sv_shape, d.qice);
}

__global__ __launch_bounds__(256) void kb_side3(const MemberDev *M, const int *starts, int nm) {
    LICOM_MEMBER(ST_SIDE3)
    const size_t n3 = (size_t)d.g.imt * d.g.jmt * d.g.km;
    const size_t t = (size_t)l * 256 + threadIdx.x;
    const float oncc = (float)d.p.oncc;
    if (d.side_vec) {
        if (t >= n3 / 4) return;
        float4 *u = reinterpret_cast<float4 *>(d.utf), *v = reinterpret_cast<float4 *>(d.vtf);
        float4 a = u[t], b = v[t];
        a.x *= oncc; a.y *= oncc; a.z *= oncc; a.w *= oncc;
        b.x *= oncc; b.y *= oncc; b.z *= oncc; b.w *= oncc;
        u[t] = a;
        v[t] = b;
    } else {
        if (t >= n3) return;
        d.utf[t] *= oncc;
        d.vtf[t] *= oncc;
    }
}

__global__ __launch_bounds__(256) void kb_side2(const MemberDev *M, const int *starts, int nm) {
    LICOM_MEMBER(ST_SIDE2)
    const KGrid &g = d.g;
    const size_t n2 = (size_t)g.imt * g.jmt, n3 = n2 * g.km;
    const size_t t = (size_t)l * 256 + threadIdx.x;
    if (t >= n2) return;
    d.h0f[t] = d.h0f[t] * d.p.onbc;
    const int j = (int)(t / g.imt), i = (int)(t % g.imt);
    if (i >= 1 && i < g.imt - 1 && j >= 1 && j < g.jmt - 1)
        for (int n = 0; n < 2; ++n)
            d.akt[n * n3 + t] = (float)dev::floor_akt((double)d.akt[n * n3 + t], 0, n, d.p.hmix, d.p.dwndmix);
}

// Ghost cells of at/atb; plane 4 km is licomqice when icesnow ran.
__global__ __launch_bounds__(256) void kb_halo(const MemberDev *M, const int *starts, int nm, int ice) {
    LICOM_MEMBER(ST_HALO)
    const int plane = l / d.halo_bx, t = (l % d.halo_bx) * 256 + (int)threadIdx.x;
    if (plane < 4 * d.g.km)
        halo_cell(d.g.imt, d.g.jmt, d.g.km, plane, t, d.at, d.atb);
    else if (ice)
        halo_cell(d.g.imt, d.g.jmt, d.g.km, 0, t, d.qice, nullptr);
}

#undef LICOM_MEMBER

// ------------------------------------------------- persistent batch step
// One launch per step: a grid of co-resident workgroups walks every stage's
// work list (the same items and bodies as the kb_* launches, so the same
// bits), with a grid-wide barrier between dependent phases:
//   surface | SSSNORM partials | SSSNORM total + column prelude | k-march |
//   column solve | side effects + halo.
// The barrier waits are bounded: a workgroup that waits longer than
// kBarSpins polls (workgroups that are not co-resident, for example with
// other work on the device) sets the error flag and leaves, every later
// barrier then lets all workgroups leave at once, and licom_batch_sync
// reports the error.  The host can set the flag itself to check that path.
struct GridBar {
    unsigned count;   // arrivals at the current barrier
    unsigned gen;     // barrier generation
    int err;          // set by a timed-out wait (or by the host)
    int pad;
};
constexpr unsigned long long kBarSpins = 1ull << 24;   // ~1 s of s_sleep 2

struct StageTotals {
    int t[NSTAGE];
};

__device__ __forceinline__ bool grid_sync(GridBar *b, unsigned nblk) {
    __shared__ int s_ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        if (__hip_atomic_load(&b->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            ok = 0;
        } else {
            // the generation is read before arriving: the last arrival may
            // advance it at any time after its own atomicAdd
            const unsigned gen = __hip_atomic_load(&b->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
            __threadfence();   // this workgroup's results, before it arrives
            const unsigned arrived =
                __hip_atomic_fetch_add(&b->count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) + 1u;
            if (arrived == nblk) {
                __hip_atomic_store(&b->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_fetch_add(&b->gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                unsigned long long spins = 0;
                while (__hip_atomic_load(&b->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                    if (__hip_atomic_load(&b->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
                        ++spins > kBarSpins) {
                        __hip_atomic_store(&b->err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        ok = 0;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                }
            }
            __threadfence();   // the other workgroups' results, after the barrier
        }
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

template <int KM, bool ICE>
__global__ __launch_bounds__(256, 2) void kb_step(const MemberDev *M, const int *starts, int nm,
                                               StageTotals T, int flags, GridBar *bar) {
    const int G = (int)gridDim.x;
    const unsigned nblk = gridDim.x;
    auto member = [&](int stage, int it, int &l) -> const MemberDev & {
        const int *st_ = starts + stage * (nm + 1);
        const int m = find_member(st_, nm, it);
        l = it - st_[m];
        return M[m];
    };
    // surface fluxes, T->U average, SSSNORM integrand
    for (int it = blockIdx.x; it < T.t[ST_SURF]; it += G) {
        int l;
        const MemberDev &d = member(ST_SURF, it, l);
        const SurfaceIO o{d.h0f, d.h0l, d.atb, d.swv, d.nswv, d.sst, d.seaice, d.sss, d.fresh,
                          d.stf_pre, d.work_u, d.stf_n, d.temp11, d.net, d.mix_count};
        const int bx = l % d.sf_bx, by = l / d.sf_bx;
        surface_cell(d.g, d.p, o, bx * 64 + (int)threadIdx.x % 64, by * 4 + (int)threadIdx.x / 64);
    }
    if (!grid_sync(bar, nblk)) return;
    for (int it = blockIdx.x; it < T.t[ST_SSSP]; it += G) {
        int l;
        const MemberDev &d = member(ST_SSSP, it, l);
        sss_partial_block(d.g, d.temp11, d.partials, l, d.nblk);
        __syncthreads();
    }
    if (!grid_sync(bar, nblk)) return;
    // SSSNORM total and the column prelude (both need only the surface)
    for (int it = blockIdx.x; it < T.t[ST_SSSF] + T.t[ST_FD]; it += G) {
        int l;
        if (it < T.t[ST_SSSF]) {
            const MemberDev &d = member(ST_SSSF, it, l);
            sss_final_block(d.partials, d.nblk, d.area_t, d.p.sssnorm, d.err_norm2, d.fw_norm2);
        } else {
            const MemberDev &d = member(ST_FD, it - T.t[ST_SSSF], l);
            prelude_col_mx_body<KM>(d.g, (float)d.p.oncc, d.utf, d.utl, d.vtf, d.vtl, d.work_u, d.stf_pre,
                                    d.ws, d.uw, d.vs, l % d.fd_bx, l / d.fd_bx, d.fd_shape);
        }
        __syncthreads();
    }
    if (!grid_sync(bar, nblk)) return;
    for (int it = blockIdx.x; it < T.t[ST_MARCH]; it += G) {
        int l;
        const MemberDev &d = member(ST_MARCH, it, l);
        MarchArgsMx a;
        a.at = d.at;
        a.atb = d.atb;
        a.stf_n = d.stf_n;
        a.err_norm2 = d.err_norm2;
        a.swv = d.swv;
        a.restore = d.restore;
        a.akt = d.akt;
        a.uw = d.uw;
        a.vs = d.vs;
        a.ws = d.ws;
        a.ori = d.ori;
        a.ax = d.ax;
        a.ay = d.ay;
        a.az = d.az;
        a.dz = d.dz;
        a.pen = d.pen;
        a.ntx = d.ntx;
        a.ntiles = d.ntiles;
        a.tiles_per_xcd = 0;
        march_body(d.g, d.p, a, l >> 1, l & 1);
        __syncthreads();
    }
    if (!grid_sync(bar, nblk)) return;
    for (int it = blockIdx.x; it < T.t[ST_SOLVE]; it += G) {
        int l;
        const MemberDev &d = member(ST_SOLVE, it, l);
        solve_body<KM, float, float, ICE>(d.g, d.p, flags, d.at, d.atb, d.akt, d.ori, d.stf_n, d.tend,
                                          d.dt_diff, d.dt_conv, d.mix_count, l % d.sv_bx, l / d.sv_bx,
                                          d.sv_shape, ICE ? d.qice : nullptr);
        __syncthreads();
    }
    if (!grid_sync(bar, nblk)) return;
    // side effects and ghost cells (independent of each other)
    const int nside = (flags & LICOM_STEP_SIDE_EFFECTS) ? T.t[ST_SIDE3] + T.t[ST_SIDE2] : 0;
    for (int it = blockIdx.x; it < nside + T.t[ST_HALO]; it += G) {
        int l;
        const unsigned tid = threadIdx.x;
        if (it < T.t[ST_SIDE3] && it < nside) {
            const MemberDev &d = member(ST_SIDE3, it, l);
            const size_t n3 = (size_t)d.g.imt * d.g.jmt * d.g.km;
            const size_t t = (size_t)l * 256 + tid;
            const float oncc = (float)d.p.oncc;
            if (d.side_vec) {
                if (t < n3 / 4) {
                    float4 *u = reinterpret_cast<float4 *>(d.utf), *v = reinterpret_cast<float4 *>(d.vtf);
                    float4 a = u[t], b = v[t];
                    a.x *= oncc; a.y *= oncc; a.z *= oncc; a.w *= oncc;
                    b.x *= oncc; b.y *= oncc; b.z *= oncc; b.w *= oncc;
                    u[t] = a;
                    v[t] = b;
                }
            } else if (t < n3) {
                d.utf[t] *= oncc;
                d.vtf[t] *= oncc;
            }
        } else if (it < nside) {
            const MemberDev &d = member(ST_SIDE2, it - T.t[ST_SIDE3], l);
            const KGrid &g = d.g;
            const size_t n2 = (size_t)g.imt * g.jmt, n3 = n2 * g.km;
            const size_t t = (size_t)l * 256 + tid;
            if (t < n2) {
                d.h0f[t] = d.h0f[t] * d.p.onbc;
                const int j = (int)(t / g.imt), i = (int)(t % g.imt);
                if (i >= 1 && i < g.imt - 1 && j >= 1 && j < g.jmt - 1)
                    for (int n = 0; n < 2; ++n)
                        d.akt[n * n3 + t] =
                            (float)dev::floor_akt((double)d.akt[n * n3 + t], 0, n, d.p.hmix, d.p.dwndmix);
            }
        } else {
            const MemberDev &d = member(ST_HALO, it - nside, l);
            const int plane = l / d.halo_bx, t = (l % d.halo_bx) * 256 + (int)tid;
            if (plane < 4 * d.g.km)
                halo_cell(d.g.imt, d.g.jmt, d.g.km, plane, t, d.at, d.atb);
            else if (ICE)
                halo_cell(d.g.imt, d.g.jmt, d.g.km, 0, t, d.qice, nullptr);
        }
    }
}

}  // namespace
}  // namespace licom

using licom::MemberDev;

struct licom_batch {
    int n = 0;
    int kmax = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    bool stream_set = false;          // licom_batch_set_stream was called
    MemberDev *d_members = nullptr;
    int *d_starts = nullptr;
    int totals[licom::NSTAGE] = {};
    int march_grid = 0;
    bool all_qice = false;            // every member has a licomqice array
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // one persistent launch per step (licom_batch_set_persistent)
    bool persistent = false;
    int step_grid = 0;                // co-resident workgroups of kb_step
    licom::GridBar *d_bar = nullptr;
};

namespace {

// The batch's own stream, created on first use unless one was set.
int ensure_stream(licom_batch *b) {
    if (b->stream || b->stream_set) return LICOM_OK;
    const int rc = licom::hip_status(hipStreamCreate(&b->stream), "hipStreamCreate");
    b->own_stream = rc == LICOM_OK;
    return rc;
}

int batch_launch_persistent(licom_batch *b, int flags) {
    using namespace licom;
    StageTotals T;
    for (int s = 0; s < NSTAGE; ++s) T.t[s] = b->totals[s];
    const bool ice = flags & LICOM_STEP_ICESNOW;
    const dim3 gr((unsigned)b->step_grid);
    const MemberDev *M = b->d_members;
    const int *S = b->d_starts;
    hipStream_t st = b->stream;
    if (b->kmax <= 30) {
        if (ice) kb_step<30, true><<<gr, 256, 0, st>>>(M, S, b->n, T, flags, b->d_bar);
        else kb_step<30, false><<<gr, 256, 0, st>>>(M, S, b->n, T, flags, b->d_bar);
    } else if (b->kmax <= 55) {
        if (ice) kb_step<55, true><<<gr, 256, 0, st>>>(M, S, b->n, T, flags, b->d_bar);
        else kb_step<55, false><<<gr, 256, 0, st>>>(M, S, b->n, T, flags, b->d_bar);
    } else {
        if (ice) kb_step<LICOM_MAX_KM, true><<<gr, 256, 0, st>>>(M, S, b->n, T, flags, b->d_bar);
        else kb_step<LICOM_MAX_KM, false><<<gr, 256, 0, st>>>(M, S, b->n, T, flags, b->d_bar);
    }
    return hip_status(hipGetLastError(), "persistent batch launch");
}

// Workgroups of kb_step that fit on the device at once (the occupancy of the
// instantiation the batch runs, times the CU count).
int persistent_grid(licom_batch *b, int &grid) {
    using namespace licom;
    int dev = 0, cus = 0, per = 0;
    if (int rc = hip_status(hipGetDevice(&dev), "hipGetDevice")) return rc;
    if (int rc = hip_status(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev),
                            "hipDeviceGetAttribute"))
        return rc;
    const void *fn = b->kmax <= 30   ? (const void *)kb_step<30, false>
                     : b->kmax <= 55 ? (const void *)kb_step<55, false>
                                     : (const void *)kb_step<LICOM_MAX_KM, false>;
    const void *fi = b->kmax <= 30   ? (const void *)kb_step<30, true>
                     : b->kmax <= 55 ? (const void *)kb_step<55, true>
                                     : (const void *)kb_step<LICOM_MAX_KM, true>;
    int per_i = 0;
    if (int rc = hip_status(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, 256, 0), "occupancy"))
        return rc;
    if (int rc = hip_status(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_i, fi, 256, 0), "occupancy"))
        return rc;
    per = std::min(per, per_i);
    if (per < 1 || cus < 1) {
        set_error("persistent batch step: no workgroup of kb_step fits on a CU");
        return LICOM_ENOTSUP;
    }
    grid = per * cus;
    return LICOM_OK;
}

int batch_launch_step(licom_batch *b, int flags) {
    using namespace licom;
    if (b->persistent) return batch_launch_persistent(b, flags);
    const int nm = b->n;
    hipStream_t st = b->stream;
    const MemberDev *M = b->d_members;
    const int *S = b->d_starts;
    auto grid = [&](int stage) { return (unsigned)b->totals[stage]; };
    kb_surface<<<grid(ST_SURF), 256, 0, st>>>(M, S, nm);
    kb_sss_partial<<<grid(ST_SSSP), kRedThreads, 0, st>>>(M, S, nm);
    kb_sss_final<<<grid(ST_SSSF), kRedThreads, 0, st>>>(M, S, nm);
    if (b->kmax <= 30) kb_prelude<30><<<grid(ST_FD), 256, 0, st>>>(M, S, nm);
    else if (b->kmax <= 55) kb_prelude<55><<<grid(ST_FD), 256, 0, st>>>(M, S, nm);
    else kb_prelude<LICOM_MAX_KM><<<grid(ST_FD), 256, 0, st>>>(M, S, nm);
    const int items = b->totals[ST_MARCH];
    const int mg = b->march_grid > 0 ? std::min(b->march_grid, items) : items;
    kb_march<<<(unsigned)mg, NT, 0, st>>>(M, S, nm, items);
    if (flags & LICOM_STEP_ICESNOW) {
        if (b->kmax <= 30) kb_solve_ice<30><<<grid(ST_SOLVE), 256, 0, st>>>(M, S, nm, flags);
        else if (b->kmax <= 55) kb_solve_ice<55><<<grid(ST_SOLVE), 256, 0, st>>>(M, S, nm, flags);
        else kb_solve_ice<LICOM_MAX_KM><<<grid(ST_SOLVE), 256, 0, st>>>(M, S, nm, flags);
    } else {
        if (b->kmax <= 30) kb_solve<30><<<grid(ST_SOLVE), 256, 0, st>>>(M, S, nm, flags);
        else if (b->kmax <= 55) kb_solve<55><<<grid(ST_SOLVE), 256, 0, st>>>(M, S, nm, flags);
        else kb_solve<LICOM_MAX_KM><<<grid(ST_SOLVE), 256, 0, st>>>(M, S, nm, flags);
    }
    if (flags & LICOM_STEP_SIDE_EFFECTS) {
        kb_side3<<<grid(ST_SIDE3), 256, 0, st>>>(M, S, nm);
        kb_side2<<<grid(ST_SIDE2), 256, 0, st>>>(M, S, nm);
    }
    kb_halo<<<grid(ST_HALO), 256, 0, st>>>(M, S, nm, (flags & LICOM_STEP_ICESNOW) ? 1 : 0);
    return hip_status(hipGetLastError(), "batch launch");
}

int check_batch_flags(const licom_batch *b, int flags) {
    if (flags & (LICOM_STEP_VTL_SCRATCH | LICOM_STEP_STAGED)) {
        licom::set_error("batch: VTL_SCRATCH and STAGED are FP64-only");
        return LICOM_ENOTSUP;
    }
    if ((flags & LICOM_STEP_ICESNOW) && !b->all_qice) {
        licom::set_error("batch: LICOM_STEP_ICESNOW needs licom_tracer_set_icesnow on every member");
        return LICOM_EINVAL;
    }
    return LICOM_OK;
}

}  // namespace

extern "C" int licom_batch_create(licom_ctx *const *members, const licom_fields_mixed *fields,
                                  const licom_diag_mixed *diags, const licom_member_tuning *tuning,
                                  int n, licom_batch **out) {
    using namespace licom;
    if (!members || !fields || !out || n <= 0) {
        set_error("licom_batch_create: NULL argument or n <= 0");
        return LICOM_EINVAL;
    }
    *out = nullptr;
    std::vector<MemberDev> hm(n);
    std::vector<int> starts((size_t)NSTAGE * (n + 1), 0);
    int kmax = 0;
    for (int m = 0; m < n; ++m) {
        licom_ctx *c = members[m];
        const licom_fields_mixed &f = fields[m];
        if (!c || !f.at || !f.atb || !f.akt || !f.utf || !f.vtf || !f.utl || !f.vtl || !f.ws ||
            !f.h0f || !f.h0l) {
            set_error("licom_batch_create: member %d has a NULL context or field", m);
            return LICOM_EINVAL;
        }
        const uintptr_t al = reinterpret_cast<uintptr_t>(f.at) | reinterpret_cast<uintptr_t>(f.atb);
        if (c->imt % 2 != 0 || (al & 15) != 0) {
            set_error("licom_batch_create: member %d needs an even imt and 16-byte aligned at/atb", m);
            return LICOM_ENOTSUP;
        }
        if (c->p.boundary_restore == 1 && !f.restore) {
            set_error("licom_batch_create: member %d has boundary_restore=1 but no restore field", m);
            return LICOM_EINVAL;
        }
        if (c->kp.adv != LICOM_ADV_TSPAS) {
            set_error("licom_batch_create: member %d: batches run the mixed tspas form only (adv=%d)", m,
                      c->kp.adv);
            return LICOM_ENOTSUP;
        }
        const licom_member_tuning tu = tuning ? tuning[m] : licom_member_tuning{0, 0};
        if (tu.prelude_shape < 0 || tu.prelude_shape > 2 || tu.solve_shape < 0 || tu.solve_shape > 2) {
            set_error("licom_batch_create: member %d has an unknown launch shape", m);
            return LICOM_EINVAL;
        }
        MemberDev &d = hm[m];
        d = MemberDev{};
        d.g = c->g;
        d.p = c->kp;
        d.at = f.at;
        d.atb = f.atb;
        d.h0f = f.h0f;
        d.h0l = f.h0l;
        d.swv = f.swv;
        d.nswv = f.nswv;
        d.sst = f.sst;
        d.seaice = f.seaice;
        d.sss = f.sss;
        d.fresh = f.fresh;
        d.restore = f.restore;
        d.akt = f.akt;
        d.utf = f.utf;
        d.vtf = f.vtf;
        d.utl = f.utl;
        d.vtl = f.vtl;
        d.ws = f.ws;
        if (diags) {
            const licom_diag_mixed &dg = diags[m];
            d.ax = dg.ax;
            d.ay = dg.ay;
            d.az = dg.az;
            d.dz = dg.dz;
            d.tend = dg.tend;
            d.dt_diff = dg.dt_diff;
            d.dt_conv = dg.dt_conv;
            d.pen = dg.penetrate;
            d.net = dg.net;
            d.fw_norm2 = dg.fw_norm2;
        }
        Scratch &s = c->s;
        d.stf_pre = s.stf_pre;
        d.work_u = s.work_u;
        d.stf_n = s.stf_n;
        d.temp11 = s.temp11;
        d.partials = s.partials;
        d.err_norm2 = s.err_norm2;
        d.ori = s.ori;
        d.uw = reinterpret_cast<float *>(s.uw);
        d.vs = reinterpret_cast<float *>(s.vs);
        d.mix_count = s.mix_count;
        d.qice = c->qice;
        d.area_t = c->area_t;
        d.nblk = s.nblk;
        const int imt = c->imt, jmt = c->jmt, km = c->km;
        kmax = std::max(kmax, km);
        d.sf_bx = (imt + 63) / 64;
        d.sf_by = (jmt + 3) / 4;
        d.fd_shape = tu.prelude_shape;
        d.fd_bx = (imt + fd_cols(d.fd_shape) - 1) / fd_cols(d.fd_shape);
        d.fd_by = (jmt + fd_rows(d.fd_shape) - 1) / fd_rows(d.fd_shape);
        d.sv_shape = tu.solve_shape;
        const dim3 sg = solve_grid(imt, jmt, d.sv_shape);
        d.sv_bx = (int)sg.x;
        d.sv_by = (int)sg.y;
        d.ntx = (imt - 4 + BX - 1) / BX;
        d.ntiles = d.ntx * ((jmt - 4 + BY - 1) / BY);
        const size_t n3 = c->n3, n2 = c->n2;
        d.side_vec = ((reinterpret_cast<uintptr_t>(f.utf) | reinterpret_cast<uintptr_t>(f.vtf)) & 15) == 0 &&
                     n3 % 4 == 0;
        d.side_blocks = (int)(((d.side_vec ? n3 / 4 : n3) + 255) / 256);
        d.side2_blocks = (int)((n2 + 255) / 256);
        d.halo_bx = (4 * (jmt - 4) + 4 * imt + 255) / 256;
        const long cnt[NSTAGE] = {(long)d.sf_bx * d.sf_by,
                                  d.nblk,
                                  1,
                                  (long)d.fd_bx * d.fd_by,
                                  0,
                                  2L * d.ntiles,
                                  (long)d.sv_bx * d.sv_by,
                                  d.side_blocks,
                                  d.side2_blocks,
                                  (long)d.halo_bx * (4 * km + 1)};
        for (int st = 0; st < NSTAGE; ++st) {
            const long next = starts[(size_t)st * (n + 1) + m] + cnt[st];
            if (next > 0x7fffffffL) {
                set_error("licom_batch_create: too many workgroups in stage %d", st);
                return LICOM_EINVAL;
            }
            starts[(size_t)st * (n + 1) + m + 1] = (int)next;
        }
    }
    if (kmax > LICOM_MAX_KM) {
        set_error("licom_batch_create: km %d exceeds %d", kmax, LICOM_MAX_KM);
        return LICOM_EINVAL;
    }
    licom_batch *b = new (std::nothrow) licom_batch;
    if (!b) return LICOM_ENOMEM;
    b->n = n;
    b->kmax = kmax;
    b->all_qice = std::all_of(hm.begin(), hm.end(), [](const MemberDev &d) { return d.qice != nullptr; });
    for (int st = 0; st < NSTAGE; ++st) b->totals[st] = starts[(size_t)st * (n + 1) + n];
    int rc = hip_status(hipMalloc(&b->d_members, sizeof(MemberDev) * n), "hipMalloc batch");
    if (!rc) rc = hip_status(hipMalloc(&b->d_starts, sizeof(int) * starts.size()), "hipMalloc batch");
    if (!rc) rc = hip_status(hipMemcpy(b->d_members, hm.data(), sizeof(MemberDev) * n, hipMemcpyHostToDevice),
                             "hipMemcpy batch");
    if (!rc) rc = hip_status(hipMemcpy(b->d_starts, starts.data(), sizeof(int) * starts.size(),
                                       hipMemcpyHostToDevice), "hipMemcpy batch");
    // the batch's own stream is created on first use unless the caller sets
    // one (the request loop rebuilds batches often and always sets one)
    if (rc) {
        licom_batch_destroy(b);
        return rc;
    }
    *out = b;
    return LICOM_OK;
}

extern "C" int licom_batch_set_stream(licom_batch *b, void *hip_stream) {
    if (!b) return LICOM_EINVAL;
    if (b->own_stream) (void)hipStreamDestroy(b->stream);
    b->own_stream = false;
    b->stream_set = true;
    b->stream = static_cast<hipStream_t>(hip_stream);
    return LICOM_OK;
}

extern "C" int licom_batch_set_march_grid(licom_batch *b, int workgroups) {
    if (!b || workgroups < 0) return LICOM_EINVAL;
    b->march_grid = workgroups;
    return LICOM_OK;
}

extern "C" int licom_batch_step(licom_batch *b, int nsteps, int flags) {
    if (!b || nsteps < 0) return LICOM_EINVAL;
    if (int rc = check_batch_flags(b, flags)) return rc;
    if (int rc = ensure_stream(b)) return rc;
    for (int s = 0; s < nsteps; ++s)
        if (int rc = batch_launch_step(b, flags)) return rc;
    return LICOM_OK;
}

extern "C" int licom_batch_capture(licom_batch *b, int nsteps, int flags) {
    using namespace licom;
    if (!b || nsteps <= 0) return LICOM_EINVAL;
    if (int rc = check_batch_flags(b, flags)) return rc;
    if (int rc = ensure_stream(b)) return rc;
    if (!b->stream) {
        set_error("licom_batch_capture: graph capture needs a non-default stream");
        return LICOM_EINVAL;
    }
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    int rc = hip_status(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    if (rc) return rc;
    int lrc = LICOM_OK;
    for (int s = 0; s < nsteps && !lrc; ++s) lrc = batch_launch_step(b, flags);
    hipGraph_t g = nullptr;
    rc = hip_status(hipStreamEndCapture(b->stream, &g), "hipStreamEndCapture");
    if (lrc || rc) {
        if (g) (void)hipGraphDestroy(g);
        return lrc ? lrc : rc;
    }
    b->graph = g;
    return hip_status(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0), "hipGraphInstantiate");
}

extern "C" int licom_batch_replay(licom_batch *b) {
    using namespace licom;
    if (!b || !b->exec) {
        set_error("licom_batch_replay: nothing captured");
        return LICOM_EINVAL;
    }
    return hip_status(hipGraphLaunch(b->exec, b->stream), "hipGraphLaunch");
}

extern "C" int licom_batch_sync(licom_batch *b) {
    using namespace licom;
    if (!b) return LICOM_EINVAL;
    if (int rc = hip_status(hipStreamSynchronize(b->stream), "hipStreamSynchronize")) return rc;
    if (b->d_bar) {
        GridBar h{};
        if (int rc = hip_status(hipMemcpy(&h, b->d_bar, sizeof h, hipMemcpyDeviceToHost), "hipMemcpy barrier"))
            return rc;
        if (h.err) {
            // the steps since the last sync are incomplete; the barrier starts afresh
            (void)hipMemset(b->d_bar, 0, sizeof(GridBar));
            set_error("persistent batch step: a grid barrier wait timed out (workgroups not co-resident, "
                      "e.g. other work on the device); the members' state is undefined");
            return LICOM_EHIP;
        }
    }
    return LICOM_OK;
}

extern "C" int licom_batch_set_persistent(licom_batch *b, int on) {
    using namespace licom;
    if (!b) return LICOM_EINVAL;
    if (on && !b->d_bar) {
        int grid = 0;
        if (int rc = persistent_grid(b, grid)) return rc;
        if (int rc = hip_status(hipMalloc(&b->d_bar, sizeof(GridBar)), "hipMalloc barrier")) return rc;
        if (int rc = hip_status(hipMemset(b->d_bar, 0, sizeof(GridBar)), "hipMemset barrier")) return rc;
        b->step_grid = grid;
    }
    b->persistent = on != 0;
    return LICOM_OK;
}

extern "C" int licom_batch_fail_next_barrier(licom_batch *b) {
    using namespace licom;
    if (!b || !b->d_bar) return LICOM_EINVAL;
    const int one = 1;
    return hip_status(hipMemcpy(&b->d_bar->err, &one, sizeof one, hipMemcpyHostToDevice), "hipMemcpy barrier");
}

extern "C" int licom_batch_persistent_grid(const licom_batch *b) { return b ? b->step_grid : 0; }

extern "C" void licom_batch_destroy(licom_batch *b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->d_members) (void)hipFree(b->d_members);
    if (b->d_starts) (void)hipFree(b->d_starts);
    if (b->d_bar) (void)hipFree(b->d_bar);
    if (b->own_stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

extern "C" int licom_autotune_member(licom_ctx *ctx, const licom_fields_mixed *f,
                                     const licom_diag_mixed *d, int reps, licom_member_tuning *best,
                                     double *best_ms) {
    using namespace licom;
    if (!ctx || !f || !best || reps <= 0) {
        set_error("licom_autotune_member: NULL argument or reps <= 0");
        return LICOM_EINVAL;
    }
    const int flags = LICOM_STEP_CONVADJ;
    // Every candidate is timed in kRounds interleaved rounds (so that a
    // drift of the clock or of the memory system hits all of them alike) and
    // keeps its fastest round; a candidate replaces the default shape only
    // when it is faster by more than kMargin, so that run-to-run noise does
    // not pick the shape.
    constexpr int kCand = 9, kRounds = 3;
    constexpr double kMargin = 0.02;
    licom_batch *bs[kCand] = {};
    double tmin[kCand];
    for (int c = 0; c < kCand; ++c) tmin[c] = 1e300;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = hip_status(hipEventCreate(&e0), "hipEventCreate");
    if (!rc) rc = hip_status(hipEventCreate(&e1), "hipEventCreate");
    for (int c = 0; c < kCand && !rc; ++c) {
        const licom_member_tuning tu{c / 3, c % 3};
        rc = licom_batch_create(&ctx, f, d, &tu, 1, &bs[c]);
        if (!rc) rc = licom_batch_step(bs[c], 1, flags);                 // warm-up
        if (!rc) rc = licom_batch_sync(bs[c]);
    }
    for (int r = 0; r < kRounds && !rc; ++r)
        for (int c = 0; c < kCand && !rc; ++c) {
            rc = hip_status(hipEventRecord(e0, bs[c]->stream), "hipEventRecord");
            if (!rc) rc = licom_batch_step(bs[c], reps, flags);
            if (!rc) rc = hip_status(hipEventRecord(e1, bs[c]->stream), "hipEventRecord");
            if (!rc) rc = hip_status(hipEventSynchronize(e1), "hipEventSynchronize");
            float ms = 0.0f;
            if (!rc) rc = hip_status(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
            if (!rc) tmin[c] = std::min(tmin[c], (double)ms / reps);
        }
    for (int c = 0; c < kCand; ++c)
        if (bs[c]) licom_batch_destroy(bs[c]);
    int bc = 0;
    for (int c = 1; c < kCand; ++c)
        if (tmin[c] < tmin[bc]) bc = c;
    if (tmin[bc] > (1.0 - kMargin) * tmin[0]) bc = 0;
    const licom_member_tuning best_tu{bc / 3, bc % 3};
    const double best_t = tmin[bc];
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (rc) return rc;
    *best = best_tu;
    if (best_ms) *best_ms = best_t;
    return LICOM_OK;
}
