// tracer_solve.h -- column solve of the tracer step (included by the FP64
// and mixed-precision kernel files; each gets its own instantiations).
#pragma once
#include <hip/hip_runtime.h>

#include "tracer_device.h"
#include "tracer_internal.h"

namespace licom {
namespace {

// ---------------------------------------------------------------- solve
// Two lanes per column, 32 columns along i per wave: lanes 0-31 hold T and
// lanes 32-63 S of the same 32 columns, so each quarter of a wave reads or
// writes 16 consecutive values of one tracer (a whole 128-byte line of an
// FP64 array) rather than 8 of each.
//
// invtrit (tracer.cpp:923-978) eliminates from the bottom and substitutes
// from the top.  Holding e(k), f(k) for the whole column would take 2*km
// doubles per lane; instead the elimination keeps them only at every SEG-th
// level, and the substitution recomputes each SEG-level segment from its
// checkpoint with the same operations (identical bits).  Both sweeps load
// the next segment's inputs while computing the current one.
constexpr int SEG = 4;

// convadj_cu (convadj.cpp:52-256) for one column, driven by the mask of
// statically unstable interfaces found with the unmixed densities (bit k:
// rholo(k-1) > rhoup(k)).  Every density the reference reads after mixing
// starts is either recomputed by it or belongs to a level that mixing has not
// touched, so the mask plus on-demand recomputation reproduces its decisions.
// The two lanes of a column run this together: each holds its own tracer's
// column X (the state just written, stride n2) and they swap values with
// __shfl_xor over 32 lanes.  Returns the mixed level range in lo/hi.
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
                const double vo = __shfl_xor(v, 32), mo = __shfl_xor(tramix, 32);
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
                const double vo = __shfl_xor(v, 32), mo = __shfl_xor(tramix, 32);
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

// AK: element type of akt; DG: element type of the diagnostics (double for
// the FP64 drop-in, float for the mixed form; the solve itself is FP64).
// Launch shapes of the solve: a 256-lane workgroup covers R rows of C
// columns (R*C = 128 columns, two lanes each); shape 0: 4 x 32, 1: 2 x 64,
// 2: 1 x 128.  Only the mapping changes, never the arithmetic.
__host__ __device__ constexpr int solve_rows(int shape) { return shape == 0 ? 4 : (shape == 1 ? 2 : 1); }
__host__ __device__ constexpr int solve_cols(int shape) { return 128 / solve_rows(shape); }
inline dim3 solve_grid(int imt, int jmt, int shape) {
    // up to 31 extra leading columns per row from the line alignment
    return dim3((imt - 4 + 31 + solve_cols(shape) - 1) / solve_cols(shape),
                (jmt - 4 + solve_rows(shape) - 1) / solve_rows(shape));
}

// AK: element type of akt; DG: element type of the diagnostics (double for
// the FP64 drop-in, float for the mixed form; the solve itself is FP64).
// ICE: icesnow between the update and convadj, licomqice to qice.
template <int KM, typename AK, typename DG, bool ICE = false>
__device__ __forceinline__ void solve_body(const KGrid &g, const KPar &p, int flags,
                                           double *__restrict__ at, double *__restrict__ atb,
                                           const AK *__restrict__ akt,
                                           const double *__restrict__ ori,
                                           const double *__restrict__ stf_n,
                                           DG *__restrict__ tend, DG *__restrict__ dt_diff,
                                           DG *__restrict__ dt_conv, int *mix_count, int bx,
                                           int by, int shape, double *__restrict__ qice) {
    constexpr int NSEG = (KM + SEG - 1) / SEG;
    const int t = threadIdx.x, n = (t >> 5) & 1, col = (t >> 6) * 32 + (t & 31);
    const int imt = g.imt, jmt = g.jmt, km = g.km;
    // The per-level constants (odzp, odzt, the density coefficients, to, so)
    // are read from LDS: as global loads after the first stores they cannot
    // be scalar loads, and every vector load's wait would also wait for the
    // stores issued before it (one vmcnt covers both on gfx9).
    __shared__ double sZP[LICOM_MAX_KM], sZT[LICOM_MAX_KM], sTO[LICOM_MAX_KM], sSO[LICOM_MAX_KM];
    __shared__ double sC[9 * LICOM_MAX_KM];
    for (int r = t; r < km; r += 256) {
        sZP[r] = g.odzp[r];
        sZT[r] = g.odzt[r];
        sTO[r] = g.to[r];
        sSO[r] = g.so[r];
    }
    for (int r = t; r < 9 * km; r += 256) sC[r] = g.eosc[r];
    __syncthreads();
    const int R = solve_rows(shape), NC = solve_cols(shape);
    const int gj = 2 + by * R + col / NC;
    // Each wave takes 32 consecutive columns of one row, starting at an
    // element index that is a multiple of 32, so the T half's 256-byte
    // (FP64) and 128-byte (FP32) segments of every level are whole cache
    // lines; partial lines cost the L2 a fill before their write-back.
    const long rowbase = (long)gj * imt;
    const int gi = (int)(((rowbase + 2) & ~31L) + (long)NC * bx + col % NC - rowbase);
    const bool act = gi >= 2 && gi < imt - 2 && gj < jmt - 2;
    const size_t n2 = (size_t)imt * jmt, n3 = n2 * km;
    const size_t q = act ? (size_t)gj * imt + gi : 0;
    const int kz = act ? g.kmt[q] : 0;
    double *TT = at + n * n3 + q, *BB = atb + n * (n3 + n2) + n2 + q;
    const double *odzp = sZP, *odzt = sZT;
    const double dts = p.dts;
    const double stf = act ? stf_n[n * n2 + q] : 0.0;
    // icesnow between the update and convadj: dzp(0) for the ice volume,
    // levels whose state icesnow changed (bit k)
    constexpr bool ice = ICE;
    const double dzp0 = ice ? g.dzp[0] : 1.0;
    uint64_t icemask = 0;
    double qv = 0.0;
    const int hmix = p.hmix;
    const double dwm = p.dwndmix;
    // With FP32 diagnostics (kBuf) every 3-D access of the substitution goes
    // through a buffer over one plane of this tracer: masked lanes (land,
    // outside the tile, beyond km) get an out-of-range offset and load 0 /
    // drop their store with no branch, so the waits for a prefetched segment
    // never include the stores that followed it; a NULL diagnostic gets an
    // empty buffer.  The FP64 form moves enough bytes to be bound by HBM
    // bandwidth and runs faster with plain conditional accesses.
    constexpr bool kBuf = sizeof(DG) == 4;
    const double *Ob = ori + n * n3;
    const AK *Ab = akt + n * n3;
    double *Tb = at + n * n3, *Bb = atb + n * (n3 + n2) + n2;
    const size_t pb8 = n2 * sizeof(double), pba = n2 * sizeof(AK), pbg = n2 * sizeof(DG);
    auto plane = [&](const void *base, size_t esz, int k, size_t bytes) {
        const int kk = k < km ? k : km - 1;
        return plane_rsrc(base ? (const char *)base + (size_t)kk * n2 * esz : nullptr, base ? bytes : 0);
    };
    const uint32_t q8 = (uint32_t)q * 8u, qak = (uint32_t)q * (uint32_t)sizeof(AK),
                   qg = (uint32_t)q * (uint32_t)sizeof(DG);
    auto ld_o = [&](int k, bool ok) { return buf_load(plane(Ob, 8, k, pb8), ok ? q8 : kBufOff, 0.0); };
    auto ld_t = [&](int k, bool ok) { return buf_load(plane(Tb, 8, k, pb8), ok ? q8 : kBufOff, 0.0); };
    auto ld_a = [&](int k, bool ok) {
        return (double)buf_load(plane(Ab, sizeof(AK), k, pba), ok ? qak : kBufOff, AK(0));
    };

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
                pd[r] = Ob[(size_t)k * n2 + q];
                pa[r] = k >= 1 ? (double)Ab[(size_t)(k - 1) * n2 + q] : 0.0;
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
                        pd[r] = Ob[(size_t)k * n2 + q];
                        pa[r] = k >= 1 ? (double)Ab[(size_t)(k - 1) * n2 + q] : 0.0;
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
    const double *C = sC, *to = sTO, *so = sSO;
    double xprev = 0.0, Tp = 0.0, Sp = 0.0;
    uint64_t mask = 0;
    double qd[SEG], qa[SEG], qt[SEG], qw = 0.0;
    auto load_seg = [&](int m) {
#pragma unroll
        for (int r = 0; r < SEG; ++r) {
            const int k = SEG * m + r;
            if constexpr (kBuf) {
                const bool ok = k < km && act;
                qd[r] = k < KM ? ld_o(k, ok) : 0.0;
                qt[r] = k < KM ? ld_t(k, ok) : 0.0;
                qa[r] = k < KM && k >= 1 ? ld_a(k - 1, k < kz) : 0.0;
            } else {
                qd[r] = qa[r] = qt[r] = 0.0;
                if (k < KM && k < km && act) {
                    qd[r] = Ob[(size_t)k * n2 + q];
                    qt[r] = Tb[(size_t)k * n2 + q];
                    if (k >= 1 && k < kz) qa[r] = (double)Ab[(size_t)(k - 1) * n2 + q];
                }
            }
        }
        const int kb = SEG * m + SEG - 1;
        if constexpr (kBuf) qw = kb < KM ? ld_a(kb, kb < kz - 1) : 0.0;
        else qw = (kb < KM && kb < kz - 1) ? (double)Ab[(size_t)kb * n2 + q] : 0.0;
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
            // levels k >= km (KM above km) store nothing and never count as wet
            if (k < KM) {
                const bool wet = k < kz;
                const double o = cd[r];
                double x = o;
                if (wet) x = (k == 0) ? (se[0] * stf + sf[0]) * 1.0 : (se[r] * xprev + sf[r]) * 1.0;
                xprev = x;
                const double vk = wet ? 1.0 : 0.0;
                double v = (x - o) / dts * vk;
                if (k == 0) v = v - stf * 0.5 * odzp[0];
                const double tn = (x - ct[r]) / dts * vk;
                // icesnow_cu (icesnow.cpp:33-67) on the updated column: the
                // surface freezing of a wet column (T lane) with its rejected
                // salt (S lane, through the lane pair), then the TBICE clamp
                // of T at every level; atb keeps the tracer_ state x, as the
                // reference's icesnow leaves atb alone.  Its melt branch
                // (icesnow.cpp:47-59) needs licomqice > 0 on entry, which the
                // FRC_CORE reset rules out.
                double xs = x;
                if constexpr (ice) {
                    if (k == 0) {
                        double td = 0.0;
                        if (n == 0 && kz > 0 && x < dev::kTbice) {
                            td = dev::kTbice - x;
                            xs = dev::kTbice;
                        }
                        const double tdo = __shfl_xor(td, 32);
                        if (n == 1 && tdo > 0.0) xs = x + dev::ice_salt(tdo);
                        qv = td > 0.0 ? 0.0 + td * dzp0 / dzp0 : 0.0;
                    }
                    if (n == 0 && xs < dev::kTbice) xs = dev::kTbice;
                    if (xs != x && k < km) icemask |= 1ull << k;
                    // keep the mask materialised level by level (sunk to its
                    // use in the tail, it keeps every level's values live)
                    asm volatile("" : "+v"(icemask));
                }
                if constexpr (kBuf) {
                    const bool st = act && k < km;
                    const uint32_t og = st ? qg : kBufOff, o8 = st ? q8 : kBufOff;
                    buf_store(plane(dt_diff ? dt_diff + n * n3 : nullptr, sizeof(DG), k, pbg), og, (DG)v);
                    buf_store(plane(tend ? tend + n * n3 : nullptr, sizeof(DG), k, pbg), og, (DG)tn);
                    buf_store(plane(conv && dt_conv ? dt_conv + n * n3 : nullptr, sizeof(DG), k, pbg), og,
                              (DG)0.0);
                    buf_store(plane(Tb, 8, k, pb8), o8, xs);
                    buf_store(plane(Bb, 8, k, pb8), o8, x);
                } else if (act && k < km) {
                    // non-temporal: the L2 otherwise fills each written line
                    // from HBM before merging the partial writes (the FP64
                    // solve fetched ~32 GB more than it reads at 0.1 deg);
                    // 8% faster here, no change for the FP32 buffer stores
                    const size_t c = n * n3 + (size_t)k * n2 + q;
                    if (dt_diff) __builtin_nontemporal_store((DG)v, &dt_diff[c]);
                    if (tend) __builtin_nontemporal_store((DG)tn, &tend[c]);
                    if (conv && dt_conv) __builtin_nontemporal_store((DG)0.0, &dt_conv[c]);
                    __builtin_nontemporal_store(xs, &Tb[(size_t)k * n2 + q]);
                    __builtin_nontemporal_store(x, &Bb[(size_t)k * n2 + q]);
                }
                if (conv) {
                    const double xo = __shfl_xor(xs, 32);
                    const double Tk = n == 0 ? xs : xo, Sk = n == 0 ? xo : xs;
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

    // A member whose state went non-finite is flagged (its counter), not
    // aborted: other members of a batch never read its arrays.  The
    // substitution carries a NaN or an infinity at any wet level down to the
    // bottom wet level (x(k) = e(k) x(k-1) + f(k); 0 * inf is NaN), so the
    // lane reads back that one value instead of testing every level, which
    // kept a lane mask live through the unrolled sweep and spilled registers.
    if (act && kz > 0 && !__builtin_isfinite(Bb[(size_t)(kz - 1) * n2 + q]))
        atomicAdd(mix_count + 1, 1);

    // ---- convadj for statically unstable columns (both lanes of the pair)
    if constexpr (!ice) {
        if (conv && act && mask != 0) {
            if (n == 0) atomicAdd(mix_count, 1);
            int lo, hi;
            convect_sparse(g.eosc, g.to, g.so, g.dzp, km, kz, mask, n, TT, n2, lo, hi);
            for (int l = lo; l <= hi; ++l) {
                const size_t c = (size_t)l * n2;
                const double y = TT[c], b = BB[c];
                if (y != b) {
                    const double dc = (y - b) / dts * 1.0;
                    if (dt_conv) dt_conv[n * n3 + q + c] = (DG)dc;
                    if (tend) tend[n * n3 + q + c] = (DG)((double)tend[n * n3 + q + c] + dc);
                    BB[c] = y;
                }
            }
        }
        return;
    }
    if (n == 0 && act) qice[q] = qv;
    // with icesnow, convadj's dt_conv = (at - atb)/dts, tend += dt_conv,
    // atb = at (convadj.cpp:259-273) also covers the levels icesnow changed
    // (every other level has at == atb here)
    if (conv && act && (mask != 0 || icemask != 0)) {
        int lo = km, hi = -1;
        if (mask != 0) {   // the same in both lanes of the pair
            if (n == 0) atomicAdd(mix_count, 1);
            convect_sparse(g.eosc, g.to, g.so, g.dzp, km, kz, mask, n, TT, n2, lo, hi);
        }
        for (int l = 0; l < km; ++l) {
            if ((l < lo || l > hi) && !((icemask >> l) & 1ull)) continue;
            const size_t c = (size_t)l * n2;
            const double y = TT[c], b = BB[c];
            if (y != b) {
                const double dc = (y - b) / dts * (l < kz ? 1.0 : 0.0);
                if (dt_conv) dt_conv[n * n3 + q + c] = (DG)dc;
                if (tend) tend[n * n3 + q + c] = (DG)((double)tend[n * n3 + q + c] + dc);
                BB[c] = y;
            }
        }
    }
}

// The single-model solve.  Its workgroups keep the hardware's round-robin
// order over the XCDs (neighbouring row chunks on different XCDs): handing
// each XCD a contiguous run of chunks instead made the 0.1 deg solve 45%
// slower (25.7 vs 17.7 ms).
template <int KM, typename AK, typename DG>
__global__ __launch_bounds__(256, 3) void k_solve(KGrid g, KPar p, int flags, double *__restrict__ at,
                                               double *__restrict__ atb,
                                               const AK *__restrict__ akt,
                                               const double *__restrict__ ori,
                                               const double *__restrict__ stf_n,
                                               DG *__restrict__ tend,
                                               DG *__restrict__ dt_diff,
                                               DG *__restrict__ dt_conv, int *mix_count, int shape) {
    solve_body<KM, AK, DG, false>(g, p, flags, at, atb, akt, ori, stf_n, tend, dt_diff, dt_conv, mix_count,
                                  blockIdx.x, blockIdx.y, shape, nullptr);
}

// The same with icesnow between the update and convadj (LICOM_STEP_ICESNOW);
// its extra live values exceed the three-waves-per-SIMD register budget.
template <int KM, typename AK, typename DG>
__global__ __launch_bounds__(256, 2) void k_solve_ice(KGrid g, KPar p, int flags, double *__restrict__ at,
                                                   double *__restrict__ atb,
                                                   const AK *__restrict__ akt,
                                                   const double *__restrict__ ori,
                                                   const double *__restrict__ stf_n,
                                                   DG *__restrict__ tend,
                                                   DG *__restrict__ dt_diff,
                                                   DG *__restrict__ dt_conv, int *mix_count, int shape,
                                                   double *__restrict__ qice) {
    solve_body<KM, AK, DG, true>(g, p, flags, at, atb, akt, ori, stf_n, tend, dt_diff, dt_conv, mix_count,
                                 blockIdx.x, blockIdx.y, shape, qice);
}

}  // namespace
}  // namespace licom
