// tracer_fused.hip -- production path of the FP64 tracer step on gfx950.
//
// Four 3-D passes per step instead of the reference's ~21 launches per
// tracer (plus the 2-D surface, SSSNORM and halo kernels that the staged
// path in tracer_kernels.hip shares):
//
//   k_prelude_col<KM>  a 64x4 block of columns marches k: face velocities
//                  u_wface/v_sface (advection_tracer_1_0), the divergence of
//                  the barotropic-corrected velocity (upwell_2,
//                  operators_div) and the continuity w (upwell_3/4), the
//                  divergence profile held in registers.
//   k_march        one four-wave workgroup per 16x12 tile of columns and
//                  tracer, marching k top to bottom.  Waves 0-2 own the tile,
//                  wave 3 the 56-cell ring the limiter and del4 read around
//                  it; wave 3 also streams the planes of at, atb and the
//                  velocities (tile plus a 2-cell halo) straight into LDS a
//                  level ahead.  The tspas predictor of level k+1 feeds the
//                  limiter of level k through LDS flag planes; the del4 first
//                  pass stays in LDS.  Writes the explicit update (tracer_8's
//                  vtl) and the advective / vertical-diffusion diagnostics.
//   k_solve<KM>    two lanes per column (T and S): the implicit vertical
//                  diffusion (invtrit) with checkpointed elimination
//                  coefficients in registers, tracer_9's tendencies, and a
//                  streaming stability check of convadj; statically unstable
//                  columns are then mixed in place exactly as convadj_cu does.
//   k_side3/k_side2  the in-place side effects on utf, vtf, h0f and akt.
//
// Arithmetic is shared with the per-stage kernels through tracer_device.h,
// so both paths give identical bits.
#include <hip/hip_runtime.h>

#include "tracer_device.h"
#include "tracer_internal.h"
#include "tracer_solve.h"

namespace licom {
namespace {

constexpr int KMAXF = LICOM_MAX_KM;

// ------------------------------------------------------------- prelude
// One-pass prelude: a 64x4 block of columns marches k.  Each level's wkd /
// wkb window (the block plus its east column and north row, 65 x 5) is
// staged in LDS from registers loaded a level ahead, so every velocity value
// is read from HBM once; the metrics and the barotropic factors
// (1 + work_u*ohbu, upwell_2) are loaded once per column; the divergence
// profile stays in registers for the column sum and the continuity recursion
// of upwell_3/4 (tracer.cpp:1029-1080).
template <int KM, bool FLUX>
__global__ __launch_bounds__(256) void k_prelude_col(KGrid g, KPar p, const double *__restrict__ utf,
                                                     const double *__restrict__ utl,
                                                     const double *__restrict__ vtf,
                                                     const double *__restrict__ vtl,
                                                     const double *__restrict__ work_u,
                                                     const double *__restrict__ stf_pre,
                                                     double *__restrict__ ws, double *__restrict__ uw,
                                                     double *__restrict__ vs, XcdGrid xg) {
    constexpr int WX = 65, WN = 5 * WX;
    __shared__ double sD[2][WN], sBv[2][WN];
    // dzp from LDS: a global load inside the level loop would carry a vmcnt
    // wait that also waits for the level's stores and the next level's loads
    __shared__ double sDZ[KMAXF];
    int bx, by;
    if (!xcd_block(xg, bx, by)) return;
    const int t = threadIdx.x, tx = t % 64, ty = t / 64;
    const int i0 = bx * 64, j0 = by * 4;
    const int i = i0 + tx, j = j0 + ty;
    const int imt = g.imt, jmt = g.jmt, km = g.km;
    for (int r = t; r < km; r += 256) sDZ[r] = g.dzp[r];
    const size_t n2 = (size_t)imt * jmt;
    const double oncc = p.oncc;
    // this lane's window elements: e = t and e = t + 256
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
    for (int u = 0; u < 2; ++u) wo[u] = buf_off(wok[u], wq[u], 8);
    const size_t pb = n2 * sizeof(double);
    double f_ut[2], f_ul[2], f_vt[2], f_vl[2];
    auto fetch = [&](int k) {
        const size_t o = (size_t)k * n2;
        const auto r_ut = plane_rsrc(utf + o, pb), r_ul = plane_rsrc(utl + o, pb);
        const auto r_vt = plane_rsrc(vtf + o, pb), r_vl = plane_rsrc(vtl + o, pb);
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            f_ut[u] = buf_load(r_ut, wo[u], 0.0);
            f_ul[u] = buf_load(r_ul, wo[u], 0.0);
            f_vt[u] = buf_load(r_vt, wo[u], 0.0);
            f_vl[u] = buf_load(r_vl, wo[u], 0.0);
        }
    };
    fetch(0);
    const bool in = i < imt && j < jmt;
    const size_t q = in ? (size_t)j * imt + i : 0;
    const bool do_u = in && j >= 1 && j < jmt - 1;
    const bool do_v = in && i >= 1 && i < imt - 1;
    const bool do_w = do_u && do_v;
    const double htw_c = in ? g.htw[q] : 0.0, hts_c = in ? g.hts[q] : 0.0;
    // flux faces (tracer.cpp:106-117): U-cell widths at (j-1,i), (j,i) and
    // (j,i), (j,i+1)
    double dyu_n = 0.0, dyu_c = 0.0, dxu_c = 0.0, dxu_e = 0.0;
    if constexpr (FLUX) {
        if (do_u) {
            dyu_n = g.dyu[q - imt];
            dyu_c = g.dyu[q];
        }
        if (do_v) {
            dxu_c = g.dxu[q];
            dxu_e = g.dxu[q + 1];
        }
    }
    double fc = 0, fe = 0, fn = 0, fne = 0, htw_e = 0, hts_n = 0, tr = 0;
    double ohbt = 0, stfp = 0, prev = 0;
    int kmt = 0;
    if (do_w) {
        ohbt = g.ohbt[q];
        stfp = stf_pre[q];
        prev = ws[q];
        const size_t qe = q + 1, qn = q - imt, qne = q - imt + 1;
        fc = 1.0 + work_u[q] * g.ohbu[q];
        fe = 1.0 + work_u[qe] * g.ohbu[qe];
        fn = 1.0 + work_u[qn] * g.ohbu[qn];
        fne = 1.0 + work_u[qne] * g.ohbu[qne];
        htw_e = g.htw[qe];
        hts_n = g.hts[qn];
        tr = g.tarea_r[q];
        kmt = g.kmt[q];
    }
    const uint32_t ou = buf_off(do_u, q, 8), ov = buf_off(do_v, q, 8);
    // everything loaded so far completes here, so that no wait for it stays
    // inside the level loop
    wait_all_vmem();
    const int w0 = (ty + 1) * WX + tx;
    double wka[KM];
    double w = 0.0;
#pragma unroll
    for (int k = 0; k < KM; ++k) {
        wka[k] = 0.0;
        if (k < km) {
            const int b = k & 1;
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                if (wv[u]) {
                    sD[b][t + 256 * u] = dev::wk(f_ut[u], f_ul[u], oncc);
                    sBv[b][t + 256 * u] = dev::wk(f_vt[u], f_vl[u], oncc);
                }
            }
            if (k + 1 < km) fetch(k + 1);
            __syncthreads();
            const double dc = sD[b][w0], bc = sBv[b][w0];
            if constexpr (FLUX) {
                buf_store(plane_rsrc(uw + (size_t)k * n2, pb), ou, (sD[b][w0 - WX] * dyu_n + dc * dyu_c) * 0.25);
                buf_store(plane_rsrc(vs + (size_t)k * n2, pb), ov, (bc * dxu_c + sBv[b][w0 + 1] * dxu_e) * 0.25);
            } else {
                buf_store(plane_rsrc(uw + (size_t)k * n2, pb), ou, (sD[b][w0 - WX] + dc) * htw_c * 0.25);
                buf_store(plane_rsrc(vs + (size_t)k * n2, pb), ov, (bc + sBv[b][w0 + 1]) * hts_c * 0.25);
            }
            if (do_w && k <= kmt - 1) {
                const double de = sD[b][w0 + 1], dn = sD[b][w0 - WX], dne = sD[b][w0 - WX + 1];
                const double be = sBv[b][w0 + 1], bn = sBv[b][w0 - WX], bne = sBv[b][w0 - WX + 1];
                // operators_div (tracer.cpp:995-1003) on uk,vk of upwell_2
                const double d = 0.5 *
                                 ((fe * de + fne * dne) * htw_e - (fc * dc + fn * dn) * htw_c +
                                  (fe * be + fc * bc) * hts_c - (fne * bne + fn * bn) * hts_n) *
                                 tr;
                wka[k] = d;
                w = w - sDZ[k] * d * 1.0;
            }
        }
    }
    if (!do_w) return;
    // upwell_3 / upwell_4 (tracer.cpp:1045-1079)
    const double sc = 1.0 / (1.0 + stfp * ohbt);
#pragma unroll
    for (int k = 1; k < KM; ++k) {
        if (k < km) {
            const double cur = (k < kmt ? 1.0 : 0.0) * (prev + sDZ[k - 1] * (w * ohbt + wka[k - 1]));
            ws[(size_t)k * n2 + q] = cur * sc;
            prev = cur;
        }
    }
}

// --------------------------------------------------------------- march
constexpr int BX = 16, BY = 12;                 // interior tile (one lane per column)
constexpr int PX = BX + 2, PY = BY + 2;         // tile + 1-cell ring (predictor, d2tk)
constexpr int RX = BX + 4, RY = BY + 4;         // tile + 2-cell halo (staged planes)
constexpr int NOWN = BX * BY;                   // 192 tile lanes: waves 0-2
constexpr int NRING = 2 * (BX + BY);            // 56 ring cells: wave 3 (the ring's
                                                // corners are never read)
constexpr int NT = 256;                         // four waves
static_assert(NOWN == 192 && NRING <= 64, "three tile waves and one ring wave");
constexpr int NP = PX * PY, NR = RX * RY;
// Staged planes are loaded LA levels ahead of the phase that first reads
// them: at and ws are issued for level k+LA+2 at the start of P(k), the face
// velocities, atb and akt for level k+LA+1.  LDS slots per array follow.
constexpr int LA = 2;
constexpr int NAT = LA + 3, NW = LA + 3, NUV = LA + 2, NB = LA + 2;
// direct-to-LDS instructions the staging wave issues per level
constexpr int DMA_ATW = 2 * ((NR / 2 + 63) / 64);          // at, ws
constexpr int DMA_UVBA = 3 * ((NR / 2 + 63) / 64) + 2;     // u, v, atb + akt

struct MarchArgs {
    const double *at, *atb, *akt, *uw, *vs, *ws, *stf_n, *err_norm2, *swv, *restore;
    double *ori, *ax, *ay, *az, *dz, *pen;
    uint8_t *flags_out;   // the limiter flag of every physical cell, or NULL
    int ntx, ntiles, tiles_per_xcd;
};

// Ring cell r (0..NRING-1) -> P-plane coordinates (pi, pj): north row, south
// row, west column, east column, corners excluded.
__device__ __forceinline__ int2 ring_cell(int r) {
    const bool ns = r < 2 * BX;                       // north / south rows
    const int rr = ns ? r : r - 2 * BX;               // else west / east columns
    const int pi = ns ? 1 + (rr % BX) : (rr < BY ? 0 : PX - 1);
    const int pj = ns ? (rr < BX ? 0 : PY - 1) : 1 + (rr % BY);
    return make_int2(pi, pj);
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
                                          const double *odzp, const double *odzt, double dts,
                                          double tr, double D1, double D2, double D3, double D4,
                                          LW &lw) {
    const double Tc = L.Pk[r], Te = L.Pk[r + 1], Tw = L.Pk[r - 1], Ts = L.Pk[r + RX],
                 Tn = L.Pk[r - RX];
    const double Tu = K > 0 ? L.Pu[r] : 0.0;
    const double Td = K < g.km - 1 ? L.Pd[r] : 0.0;
    const double Ue = L.U[r + 1], Uw = L.U[r], Vs = L.V[r], Vn = L.V[r - RX];
    const double Wk = L.Wk[r], Wb = K < g.km - 1 ? L.Wb[r] : 0.0;
    dev::PTerms t;
    dev::pterms(Tc, Te, Tw, Ts, Tn, Tu, Td, Ue, Uw, Vs, Vn, Wk, Wb, tr, D1, D2, D3, D4, K, g.km,
                odzp, odzt, dts, t);
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

// One wave copies the R-window (tile plus 2-cell halo) of one plane straight
// from global memory into LDS (global_load_lds: lane l of an instruction
// lands at LDS base + 16 l), in 16-byte chunks, which needs an even imt and
// 16-byte aligned arrays (the launcher checks).  off[u] is this lane's chunk
// offset for instruction u inside the plane, clamped into the plane, so
// window cells outside the array receive other cells' values, which only
// feed cells whose results are discarded.  The loads complete asynchronously
// (vmcnt).
constexpr int NCH = NR / 2;                 // 16-byte chunks per window
constexpr int NCI = (NCH + 63) / 64;        // instructions per window
__device__ __forceinline__ void stage_plane(const double *src, double *dst, const int (&off)[NCI],
                                            int lane) {
#pragma unroll
    for (int u = 0; u < NCI; ++u)
        if (u * 64 + lane < NCH)
            dma16(src + off[u], lds_addr(dst + 128 * u));
}

// The 16x12 tile of one plane (akt: read only at the tile's own cells):
// 12 rows of 16 doubles, 8 chunks per row, 96 chunks in two instructions;
// chunk c lands at element 2c of the slot, which is the tile lane index.
constexpr int NCA = BX * BY / 2;
__device__ __forceinline__ void stage_tile(const double *src, double *dst, const int (&off)[2],
                                           int lane) {
#pragma unroll
    for (int u = 0; u < 2; ++u)
        if (u * 64 + lane < NCA)
            dma16(src + off[u], lds_addr(dst + 128 * u));
}

__device__ __forceinline__ void wait_vmem() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }
// Waits for every staged load except the `issued` issued at the start of
// the current level (direct-to-LDS loads complete in order).
__device__ __forceinline__ void wait_vmem_keep(int issued) {
    static_assert(DMA_ATW + DMA_UVBA == 17 && DMA_UVBA == 11, "wait counts below");
    if (LA < 2 || issued == 0) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    else if (issued >= 17) asm volatile("s_waitcnt vmcnt(17)" ::: "memory");
    else asm volatile("s_waitcnt vmcnt(11)" ::: "memory");
}

// Workgroup barrier that waits only for this wave's LDS traffic.  The
// staging wave's direct-to-LDS loads are waited for explicitly (wait_vmem)
// before the barrier that publishes them, so the tile waves never wait on
// vector-memory counters at a barrier.
__device__ __forceinline__ void bar_lds() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// k-march over one 16x12 tile of columns and one tracer, four waves.
// Waves 0-2 own the tile's cells; wave 3 owns the 56 ring cells the limiter
// and the del4 second pass read around the tile, and stages every plane
// (a four-wave workgroup keeps three workgroups resident per CU).
// Per level k two phases, each closed by a barrier:
//   P(k): wave 3 first issues the direct-to-LDS loads of the planes first
//         read LA levels later; then the tspas predictor of level k+1 and
//         the del4 first pass of level k at every cell (tile and ring, one
//         cell per lane).
//   C(k): limiter, fluxes and explicit update of level k on the tile (waves
//         0-2); wave 3 waits for the loads P(k+1) and C(k+1) read.
// LDS slots rotate: at and ws x(LA+3), u_wface/v_sface, atb and akt x(LA+2).
// ADV: the advection scheme (LICOM_ADV_*); centered and flux need no
// predictor, so their P(k) is the del4 first pass alone.
template <int ADV>
__global__ __launch_bounds__(NT, (LA >= 2 ? 2 : 3)) void k_march(KGrid g, KPar p, MarchArgs a) {
    constexpr bool tspas = ADV == LICOM_ADV_TSPAS;
    __shared__ __align__(16) double sAT[NAT][NR];
    __shared__ __align__(16) double sU[NUV][NR];
    __shared__ __align__(16) double sV[NUV][NR];
    __shared__ __align__(16) double sW[NW][NR];
    __shared__ __align__(16) double sB[NB][NR];
    __shared__ double sD2[NP];
    __shared__ unsigned char sFL[2][NP];
    __shared__ int sK[NR];
    // odzp, odzt and pen: per-level constants read from LDS, so that no
    // vector-memory load inside the level loop carries a vmcnt wait (on gfx9
    // vmcnt also covers the staged planes, the akt prefetch and the stores)
    __shared__ double sZP[KMAXF], sZT[KMAXF], sPN[KMAXF];
    __shared__ __align__(16) double sAK[NB][NOWN];    // akt of the tile, staged like atb

    // XCD-aware placement: consecutive blocks of one XCD take the T and S
    // halves of one tile, then the next tile of the same row of tiles.
    const int Lb = blockIdx.x, xcd = Lb & 7, slot = Lb >> 3;
    const int tile = xcd * a.tiles_per_xcd + (slot >> 1);
    const int n = slot & 1;
    if (tile >= a.ntiles) return;
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
    const int rl = t - NOWN;                     // ring lane (wave 3)
    const bool has_cell = !ringw || rl < NRING;
    const int2 rc = ring_cell(has_cell && ringw ? rl : 0);
    const int pi = ringw ? rc.x : 1 + t % BX, pj = ringw ? rc.y : 1 + t / BX;
    const int gi = gi0 - 1 + pi, gj = gj0 - 1 + pj;
    const int rC = (pj + 1) * RX + (pi + 1), pC = pj * PX + pi;
    const bool cell_ok = has_cell && gj >= 1 && gj < jmt - 1 && gi >= 1 && gi < imt - 1;
    const bool act = !ringw && gi < imt - 2 && gj < jmt - 2;
    const size_t q = cell_ok ? (size_t)gj * imt + gi : 0;

    const double *T = a.at + n * n3;
    const double *TB = a.atb + n * (n3 + n2) + n2;   // level k at plane k
    const double *A = a.akt + n * n3;
    const double dts = p.dts;

    for (int r = t; r < NR; r += NT) {
        const int rj = gj0 - 2 + r / RX, ri = gi0 - 2 + r % RX;
        sK[r] = (rj < jmt && ri < imt) ? g.kmt[(size_t)rj * imt + ri] : 0;
    }
    for (int r = t; r < km; r += NT) {
        sZP[r] = g.odzp[r];
        sZT[r] = g.odzt[r];
        sPN[r] = r < km - 1 ? g.pen[r] : 0.0;
    }
    Metrics M{};
    if (cell_ok) M = load_metrics(g, gj, gi);
    const double sw = (n == 0 && act && a.swv) ? a.swv[q] : 0.0;
    const int kmt_own = act ? g.kmt[q] : 0;
    const double stf_n = act ? a.stf_n[n * n2 + q] : 0.0;
    const double enorm = (n == 1 && p.sssnorm && act) ? a.err_norm2[0] : 0.0;
    const bool rest = p.boundary_restore == 1 && a.restore != nullptr;

    if (ringw) {
        // ---- wave 3: ring cells and plane staging
        int off[NCI];
#pragma unroll
        for (int u = 0; u < NCI; ++u) {
            const int c = min(u * 64 + lane, NCH - 1), row = c / (RX / 2), cp = c - row * (RX / 2);
            off[u] = (int)min(e0 + (long)row * imt + 2 * cp, (long)n2 - 2);
        }
        int offa[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int c = min(u * 64 + lane, NCA - 1), row = c / (BX / 2), cp = c - row * (BX / 2);
            offa[u] = (int)min((long)(gj0 + row) * imt + gi0 + 2 * cp, (long)n2 - 2);
        }
        // prologue: at, ws 0..LA+1; faces, atb, akt 0..LA
        for (int l = 0; l < NAT - 1 && l < km; ++l) {
            stage_plane(T + (size_t)l * n2, sAT[l], off, lane);
            stage_plane(a.ws + (size_t)l * n2, sW[l], off, lane);
        }
        for (int l = 0; l < NUV - 1 && l < km; ++l) {
            stage_plane(a.uw + (size_t)l * n2, sU[l], off, lane);
            stage_plane(a.vs + (size_t)l * n2, sV[l], off, lane);
            stage_plane(TB + (size_t)l * n2, sB[l], off, lane);
            stage_tile(A + (size_t)l * n2, sAK[l], offa, lane);
        }
        wait_vmem();
        bar_lds();
        if constexpr (tspas) {
            LW tmp;
            const Lv L0{sAT[0], sAT[0], sAT[1], sU[0], sV[0], sW[0], sW[1]};
            const bool f0 = cell_ok ? pred_flag(g, rC, L0, sK, 0, sZP, sZT, dts, M.tr, M.D1, M.D2, M.D3,
                                                M.D4, tmp)
                                    : false;
            if (has_cell) sFL[0][pC] = f0;
        }
        bar_lds();
        for (int k = 0; k < km; ++k) {
            const int s0 = k % NAT, s1 = (k + 1) % NAT, s2 = (k + 2) % NAT, s3 = (k + LA + 2) % NAT;
            const int u0 = k % NUV, u1 = (k + 1) % NUV, u2 = (k + LA + 1) % NUV;
            // planes for P(k+LA) and C(k+LA), issued first so that they have
            // LA levels to land; their slots were last read in C(k-1)
            int issued = 0;
            if (k + LA + 2 < km) {
                stage_plane(T + (size_t)(k + LA + 2) * n2, sAT[s3], off, lane);
                stage_plane(a.ws + (size_t)(k + LA + 2) * n2, sW[s3], off, lane);
                issued += DMA_ATW;
            }
            if (k + LA + 1 < km) {
                stage_plane(a.uw + (size_t)(k + LA + 1) * n2, sU[u2], off, lane);
                stage_plane(a.vs + (size_t)(k + LA + 1) * n2, sV[u2], off, lane);
                stage_plane(TB + (size_t)(k + LA + 1) * n2, sB[u2], off, lane);
                stage_tile(A + (size_t)(k + LA + 1) * n2, sAK[u2], offa, lane);
                issued += DMA_UVBA;
            }
            if (tspas && k + 1 < km) {
                LW tmp;
                const Lv L1{sAT[s1], sAT[s0], sAT[s2], sU[u1], sV[u1], sW[s1], sW[s2]};
                const bool f =
                    cell_ok ? pred_flag(g, rC, L1, sK, k + 1, sZP, sZT, dts, M.tr, M.D1, M.D2, M.D3,
                                        M.D4, tmp)
                            : false;
                if (has_cell) sFL[(k + 1) & 1][pC] = f;
            }
            if (p.hmix == LICOM_HMIX_DEL4 && has_cell)
                sD2[pC] = cell_ok ? d2tk_at(rC, sB[u0], sK, k, M.dtn, M.dts, M.dte, M.dtw, M.ahf) : 0.0;
            bar_lds();                // end of P(k)
            wait_vmem_keep(issued);   // C(k): the planes P(k+1) and C(k+1) read land
            bar_lds();                // end of C(k)
        }
        return;
    }

    // ---- waves 0-2: the tile
    // the per-cell loads above complete here, before the level loop, so
    // that no wait for them (which would also wait for the level's stores)
    // stays inside it
    __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0)
    bar_lds();
    LW lc{}, ln{};
    if constexpr (tspas) {
        const Lv L0{sAT[0], sAT[0], sAT[1], sU[0], sV[0], sW[0], sW[1]};
        sFL[0][pC] = cell_ok ? pred_flag(g, rC, L0, sK, 0, sZP, sZT, dts, M.tr, M.D1, M.D2, M.D3, M.D4,
                                         lc)
                             : false;
    }
    bar_lds();

    double Tprev = 0.0, Bprev = 0.0, Aprev = 0.0;
    bool fprev = false;
    const double lamda = 1.0 / (15.0 * 86400.0);
    const double aidif = 0.5;

    for (int k = 0; k < km; ++k) {
        const int s0 = k % NAT, s1 = (k + 1) % NAT, s2 = (k + 2) % NAT;
        const int u0 = k % NUV, u1 = (k + 1) % NUV;
        // ---- phase P(k)
        const double Acur = act ? dev::floor_akt(sAK[u0][t], k, n, p.hmix, p.dwndmix) : 0.0;
        if (tspas && k + 1 < km) {
            const Lv L1{sAT[s1], sAT[s0], sAT[s2], sU[u1], sV[u1], sW[s1], sW[s2]};
            sFL[(k + 1) & 1][pC] =
                cell_ok ? pred_flag(g, rC, L1, sK, k + 1, sZP, sZT, dts, M.tr, M.D1, M.D2, M.D3, M.D4,
                                    ln)
                        : false;
        }
        if (p.hmix == LICOM_HMIX_DEL4)
            sD2[pC] = cell_ok ? d2tk_at(rC, sB[u0], sK, k, M.dtn, M.dts, M.dte, M.dtw, M.ahf) : 0.0;
        bar_lds();

        // ---- phase C(k): limited fluxes and the explicit update of level k
        if (act) {
            const double *Pk = sAT[s0], *Pk1 = sAT[s1];
            const double *Bk = sB[u0], *Bk1 = sB[u1];
            const unsigned char *F = sFL[k & 1];
            const int r = rC, pp = pC;
            const double Tc = Pk[r], Te = Pk[r + 1], Tw = Pk[r - 1], Ts = Pk[r + RX], Tn = Pk[r - RX];
            const double Tu = Tprev, Td = (k + 1 < km) ? Pk1[r] : 0.0;
            const double *U = sU[u0], *V = sV[u0];
            const double Ue = U[r + 1], Uw = U[r], Vs = V[r], Vn = V[r - RX];
            const double Wk = sW[s0][r], Wb = (k + 1 < km) ? sW[s1][r] : 0.0;
            const size_t c = (size_t)k * n2 + q;
            const double vk = k < kmt_own ? 1.0 : 0.0;
            double tf;
            bool fc = false;
            if constexpr (!tspas) {
                // advection_tracer_1 (tracer.cpp:120-289); the centered form
                // accumulates its diagnostics (the caller's ax/ay/az carry the
                // sum over the diagnostic period)
                const dev::CFAdv cf = dev::adv_cf(ADV, Tc, Te, Tw, Ts, Tn, Tu, Td, Ue, Uw, Vs, Vn, Wk, Wb, M.tr,
                                                  k, km, sZP);
                if constexpr (ADV == LICOM_ADV_CENTERED) {
                    const double nss = (double)p.nss;
                    if (a.ax) a.ax[n * n3 + c] = a.ax[n * n3 + c] + cf.hx / nss;
                    if (a.ay) a.ay[n * n3 + c] = a.ay[n * n3 + c] + cf.hy / nss;
                    if (a.az) a.az[n * n3 + c] = a.az[n * n3 + c] - cf.vz / nss;
                }
                tf = cf.adv * vk;
            } else {
                fc = F[pp];
                const bool fe = F[pp + 1], fw = F[pp - 1], fs = F[pp + PX], fn = F[pp - PX];
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
                dev::pterms_cheap(Tc, Te, Tw, Ts, Tn, Tu, Td, Ue, Uw, Vs, Vn, Wk, Wb, tr, k, km, sZP, tc);
                const double xy1 = up1 ? dev::up_xy1(Te, Tc, Ue, tr) : lc.xy1;
                const double xy2 = up2 ? dev::up_xy2(Tc, Tw, Uw, tr) : lc.xy2;
                const double xy3 = up3 ? dev::up_xy1(Ts, Tc, Vs, tr) : lc.xy3;
                const double xy4 = up4 ? dev::up_xy2(Tc, Tn, Vn, tr) : lc.xy4;
                double zb1, zb2;
                if (k == 0) {
                    zb1 = 0.0;
                    zb2 = (fc || fd) ? dev::up_zb2(Tc, Td, Wb, sZP[0]) : lc.zb2;
                } else if (k == km - 1) {
                    zb1 = (fc || fprev) ? dev::up_zb1(Tu, Tc, Wk, sZP[k]) : lc.zb1;
                    zb2 = 0.0;
                } else {
                    zb1 = (fc || fprev) ? dev::up_zb1(Tu, Tc, Wk, sZP[k]) : lc.zb1;
                    zb2 = (fc || fd) ? dev::up_zb2(Tc, Td, Wb, sZP[k]) : lc.zb2;
                }
                const double xx = -(tc.x0 + xy1 + xy2 + tc.c1);
                const double yy = -(tc.y0 + xy3 + xy4 + tc.c2);
                const double zz = -(tc.za + zb1 + zb2 + tc.zc);
                if (a.ax) a.ax[n * n3 + c] = xx;
                if (a.ay) a.ay[n * n3 + c] = yy;
                if (a.az) a.az[n * n3 + c] = zz;
                tf = (xx + yy + zz) * vk;
            }
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
            const double *odzp = sZP, *odzt = sZT;
            const double v1 = (k + 1 < km && k + 1 < kmt_own) ? 1.0 : 0.0;
            if (n == 0) {
                double pv;
                if (k == 0) {
                    const double wt1 = sw * sPN[0] * (1 < kmt_own ? 1.0 : 0.0);
                    tf = tf - odzp[0] * wt1;
                    pv = -wt1 * odzp[0];
                } else if (k == km - 1) {
                    const double wt2 = sw * sPN[km - 2] * vk;
                    tf = tf + odzp[km - 1] * wt2;
                    pv = wt2 * odzp[km - 1];
                } else {
                    const double wt1 = sw * sPN[k - 1] * vk;
                    const double wt2 = sw * sPN[k] * v1;
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
            if (a.flags_out) a.flags_out[n * n3 + c] = fc;
            // surface fluxes (tracer_5 n=1, tracer_6) and restoring (tracer_7)
            if (k == 0) {
                if (n == 1) {
                    if (kmt_own > 0) tf = tf + stf_n * (1.0 - aidif) * odzp[0];
                    if (p.sssnorm) tf = tf + enorm * vk;
                } else if (kmt_own > 0) {
                    tf = tf + stf_n * odzp[0] * 0.5;
                }
            }
            if (rest && k >= 1) {
                const double Rk = a.restore[n * n3 + c];
                tf = tf + vk * (Rk - Bc) * lamda;
            }
            // tracer_8: the explicit update entering the implicit solve
            a.ori[n * n3 + c] = Tc + dts * tf;
            Tprev = Tc;
            Bprev = Bc;
            fprev = fc;
        }
        Aprev = Acur;
        lc = ln;
        bar_lds();
    }
}

// Side effects of tracer_1_0 / tracer_1 / tracer_del4_2 (tracer.cpp:1088,
// 1107-1108, 1187, 1315): utf, vtf *= oncc over the whole 3-D arrays (one
// or two elements per lane, a pure streaming pass), then h0f *= onbc and the
// AKT surface floor written back on the 2-D plane.
template <typename V>
__global__ __launch_bounds__(256) void k_side3(V *__restrict__ utf, V *__restrict__ vtf, size_t n,
                                               double oncc) {
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    V a = utf[t], b = vtf[t];
    if constexpr (sizeof(V) == 16) {
        a.x *= oncc; a.y *= oncc;
        b.x *= oncc; b.y *= oncc;
    } else {
        a *= oncc;
        b *= oncc;
    }
    utf[t] = a;
    vtf[t] = b;
}

__global__ __launch_bounds__(256) void k_side2(KGrid g, KPar p, double *h0f, double *akt) {
    const size_t n2 = (size_t)g.imt * g.jmt, n3 = n2 * g.km;
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= n2) return;
    h0f[t] = h0f[t] * p.onbc;
    const int j = (int)(t / g.imt), i = (int)(t % g.imt);
    if (i >= 1 && i < g.imt - 1 && j >= 1 && j < g.jmt - 1)
        for (int n = 0; n < 2; ++n)
            akt[n * n3 + t] = dev::floor_akt(akt[n * n3 + t], 0, n, p.hmix, p.dwndmix);
}

void launch_side(licom_ctx *ctx, const licom_fields *f) {
    const KGrid &g = ctx->g;
    hipStream_t st = ctx->stream;
    Timer t(ctx, "k_side");
    const size_t n3 = ctx->n3;
    const bool vec = ((reinterpret_cast<uintptr_t>(f->utf) | reinterpret_cast<uintptr_t>(f->vtf)) & 15) == 0 &&
                     n3 % 2 == 0;
    if (vec) {
        const size_t n = n3 / 2;
        k_side3<double2><<<(unsigned)((n + 255) / 256), 256, 0, st>>>(
            reinterpret_cast<double2 *>(f->utf), reinterpret_cast<double2 *>(f->vtf), n, ctx->kp.oncc);
    } else {
        k_side3<double><<<(unsigned)((n3 + 255) / 256), 256, 0, st>>>(f->utf, f->vtf, n3, ctx->kp.oncc);
    }
    k_side2<<<(unsigned)((ctx->n2 + 255) / 256), 256, 0, st>>>(g, ctx->kp, f->h0f, f->akt);
}

template <int KM>
void launch_prelude_solve(licom_ctx *ctx, const licom_fields *f, const licom_diag *d, int flags,
                          bool prelude, const Scratch &s) {
    const KGrid &g = ctx->g;
    const KPar &p = ctx->kp;
    hipStream_t st = ctx->stream;
    if (prelude) {
        Timer t(ctx, "k_prelude");
        const XcdGrid xg = xcd_grid((g.imt + 63) / 64, (g.jmt + 3) / 4);
        if (p.adv == LICOM_ADV_FLUX)
            k_prelude_col<KM, true><<<xcd_blocks(xg), 256, 0, st>>>(g, p, f->utf, f->utl, f->vtf, f->vtl,
                                                                    s.work_u, s.stf_pre, f->ws, s.uw, s.vs, xg);
        else
            k_prelude_col<KM, false><<<xcd_blocks(xg), 256, 0, st>>>(g, p, f->utf, f->utl, f->vtf, f->vtl,
                                                                     s.work_u, s.stf_pre, f->ws, s.uw, s.vs, xg);
    } else {
        Timer t(ctx, "k_solve");
        const dim3 gr = solve_grid(g.imt, g.jmt, ctx->solve_shape);
        if (flags & LICOM_STEP_ICESNOW)
            k_solve_ice<KM, double, double><<<gr, 256, 0, st>>>(g, p, flags, f->at, f->atb, f->akt, s.ori, s.stf_n,
                                                  d ? d->tend : nullptr, d ? d->dt_diff : nullptr,
                                                  d ? d->dt_conv : nullptr, s.mix_count, ctx->solve_shape,
                                                  ctx->qice);
        else
            k_solve<KM, double, double><<<gr, 256, 0, st>>>(g, p, flags, f->at, f->atb, f->akt, s.ori, s.stf_n,
                                              d ? d->tend : nullptr, d ? d->dt_diff : nullptr,
                                              d ? d->dt_conv : nullptr, s.mix_count, ctx->solve_shape);
    }
}

void launch_column(licom_ctx *ctx, const licom_fields *f, const licom_diag *d, int flags,
                   bool prelude, const Scratch &s) {
    switch (ctx->km) {
        case 30: launch_prelude_solve<30>(ctx, f, d, flags, prelude, s); break;
        case 55: launch_prelude_solve<55>(ctx, f, d, flags, prelude, s); break;
        default: launch_prelude_solve<KMAXF>(ctx, f, d, flags, prelude, s); break;
    }
}

}  // namespace

int launch_step_fused(licom_ctx *ctx, const licom_fields *f, const licom_diag *d, int flags,
                      const Scratch &s) {
    const KGrid &g = ctx->g;
    hipStream_t st = ctx->stream;
    if (int rc = launch_surface(ctx, f, d, s)) return rc;
    launch_column(ctx, f, d, flags, true, s);
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
        a.flags_out = ctx->flags_out;
        a.ntx = (g.imt - 4 + BX - 1) / BX;
        const int nty = (g.jmt - 4 + BY - 1) / BY;
        a.ntiles = a.ntx * nty;
        a.tiles_per_xcd = (a.ntiles + 7) / 8;
        const unsigned nb = 16 * a.tiles_per_xcd;
        switch (ctx->kp.adv) {
            case LICOM_ADV_CENTERED: k_march<LICOM_ADV_CENTERED><<<nb, NT, 0, st>>>(g, ctx->kp, a); break;
            case LICOM_ADV_FLUX: k_march<LICOM_ADV_FLUX><<<nb, NT, 0, st>>>(g, ctx->kp, a); break;
            default: k_march<LICOM_ADV_TSPAS><<<nb, NT, 0, st>>>(g, ctx->kp, a); break;
        }
    }
    launch_column(ctx, f, d, flags, false, s);
    if (flags & LICOM_STEP_SIDE_EFFECTS) launch_side(ctx, f);
    return launch_halo(ctx, f, flags);
}

}  // namespace licom
