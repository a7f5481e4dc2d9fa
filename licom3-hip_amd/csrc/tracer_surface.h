// tracer_surface.h -- 2-D surface stage, SSSNORM reduction and the halo
// fill of the tracer step, written as per-cell / per-block device bodies so
// that the single-domain kernels and the batched ensemble kernels share them.
#pragma once
#include <hip/hip_runtime.h>

#include "tracer_device.h"
#include "tracer_internal.h"

namespace licom {
namespace {

// tracer_1_0 (tracer.cpp:1081-1093), T->U average of tracer_1
// (:1131-1149), surface fluxes of tracer_5 n=1 (:1479-1524, FRC_CORE) and
// tracer_6 n=0 (:1543-1566, FRC_CORE), SSSNORM integrand (:1517-1522).
__device__ __forceinline__ double stf0(const double *h0f, const double *h0l, double onbc, size_t q) {
    return 0.5 * (h0f[q] * onbc) + 0.5 * h0l[q];
}

struct SurfaceIO {
    const double *h0f, *h0l, *atb, *swv, *nswv, *sst, *seaice, *sss, *fresh;
    double *stf_pre, *work_u, *stf_n, *temp11, *net;
    int *mix_count;
};

__device__ __forceinline__ void surface_cell(const KGrid &g, const KPar &p, const SurfaceIO &o, int i,
                                             int j) {
    if (i == 0 && j == 0 && o.mix_count) {
        o.mix_count[0] = 0;   // columns convadj mixes in this step
        o.mix_count[1] = 0;   // tracer columns whose new state is not finite
    }
    if (i >= g.imt || j >= g.jmt) return;
    const size_t q = (size_t)j * g.imt + i, n2 = (size_t)g.imt * g.jmt;
    const double *h0f = o.h0f, *h0l = o.h0l;
    const double s0 = stf0(h0f, h0l, p.onbc, q);
    o.stf_pre[q] = s0;
    double wu = 0.0;
    if (j < g.jmt - 1 && i >= 1) {
        wu = g.au0[q] * s0 + g.aus[q] * stf0(h0f, h0l, p.onbc, q + g.imt) +
             g.auw[q] * stf0(h0f, h0l, p.onbc, q - 1) +
             g.ausw[q] * stf0(h0f, h0l, p.onbc, q + g.imt - 1);
    }
    o.work_u[q] = wu;
    double sT = s0, sS = s0, t11 = 0.0;
    if (g.kmt[q] > 0) {
        const double odzp0 = g.odzp[0];
        const double tb = o.atb[n2 + q];                          // atb(:,:,1,1)
        const double sb = o.atb[(size_t)(g.km + 1) * n2 + n2 + q];  // atb(:,:,1,2)
        const double fw = o.fresh ? o.fresh[q] : 0.0, si = o.seaice ? o.seaice[q] : 0.0;
        const double ss = o.sss ? o.sss[q] : 0.0, st = o.sst ? o.sst[q] : 0.0;
        const double sw = o.swv ? o.swv[q] : 0.0, nsw = o.nswv ? o.nswv[q] : 0.0;
        sS = fw * 34.7 * p.od0 / 1000.0 + p.gamma * (ss - sb) * si / odzp0 +
             p.gamma * 30.0 / 365.0 / 4.0 * 50.0 * (ss - sb) / odzp0 * (1.0 - si);
        sT = (sw + nsw) * p.od0cp + si * p.gamma * (st - tb) / odzp0;
        const double n1 = sS * odzp0;
        t11 = g.tarea[q] * n1 * 1.0;
        if (o.net) {
            o.net[q] = sT * odzp0;
            o.net[n2 + q] = n1;
        }
    }
    o.stf_n[q] = sT;
    o.stf_n[n2 + q] = sS;
    o.temp11[q] = t11;
}

// Compensated (TwoSum) accumulation of the SSSNORM integrand over the
// physical cells; the reference sums on the host in j-major order and then
// MPI_ALLREDUCEs (tracer.cpp:1849-1868, mpi3.F90:26).  The order here is
// fixed by the block count nblk, so repeated runs, and a member run alone or
// inside a batch, give identical bits.
constexpr int kRedThreads = 256;

__device__ __forceinline__ void sss_partial_block(const KGrid &g, const double *__restrict__ temp11,
                                                  double *partials, int b, int nblk) {
    __shared__ double sh[kRedThreads], sl[kRedThreads];
    const int nx = g.imt - 4, ny = g.jmt - 4;
    const size_t total = (size_t)nx * ny;
    double hi = 0.0, lo = 0.0;
    for (size_t t = (size_t)b * kRedThreads + threadIdx.x; t < total; t += (size_t)nblk * kRedThreads) {
        const int jj = (int)(t / nx), ii = (int)(t % nx);
        double s, e;
        dev::two_sum(hi, temp11[(size_t)(jj + 2) * g.imt + ii + 2], s, e);
        hi = s;
        lo += e;
    }
    sh[threadIdx.x] = hi;
    sl[threadIdx.x] = lo;
    __syncthreads();
    for (int w = kRedThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            double s, e;
            dev::two_sum(sh[threadIdx.x], sh[threadIdx.x + w], s, e);
            sh[threadIdx.x] = s;
            sl[threadIdx.x] = sl[threadIdx.x] + sl[threadIdx.x + w] + e;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[2 * b] = sh[0];
        partials[2 * b + 1] = sl[0];
    }
}

__device__ __forceinline__ void sss_final_block(const double *__restrict__ partials, int nblk,
                                                double area_t, int enabled, double *err_norm2,
                                                double *out_copy) {
    __shared__ double sh[kRedThreads], sl[kRedThreads];
    double hi = 0.0, lo = 0.0;
    for (int b = threadIdx.x; b < nblk; b += kRedThreads) {
        double s, e;
        dev::two_sum(hi, partials[2 * b], s, e);
        hi = s;
        lo += e + partials[2 * b + 1];
    }
    sh[threadIdx.x] = hi;
    sl[threadIdx.x] = lo;
    __syncthreads();
    for (int w = kRedThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            double s, e;
            dev::two_sum(sh[threadIdx.x], sh[threadIdx.x + w], s, e);
            sh[threadIdx.x] = s;
            sl[threadIdx.x] = sl[threadIdx.x] + sl[threadIdx.x + w] + e;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double total = sh[0] + sl[0];
        const double v = enabled ? -total / area_t : 0.0;
        *err_norm2 = v;
        if (out_copy) *out_copy = v;
    }
}

// Ghost cells of at/atb after the step (tracer.cpp:1891-1898 with
// POP_HaloMod's cyclic-i / closed-j fill): lane t of plane `plane`
// (0..2km-1 at, 2km..4km-1 atb).
__device__ __forceinline__ void halo_cell(int imt, int jmt, int km, int plane, int t, double *at,
                                          double *atb) {
    const int ncol = 4 * (jmt - 4), nrow = 4 * imt;
    if (t >= ncol + nrow || plane >= 4 * km) return;
    const size_t n2 = (size_t)imt * jmt, n3 = n2 * km;
    double *P;
    if (plane < 2 * km) {
        P = at + (size_t)plane * n2;
    } else {
        const int pl = plane - 2 * km, n = pl / km, k = pl % km;
        P = atb + n * (n3 + n2) + n2 + (size_t)k * n2;
    }
    if (t < ncol) {
        const int j = 2 + t / 4, w = t % 4;
        const int dst = (w == 0) ? 0 : (w == 1) ? 1 : (w == 2) ? imt - 2 : imt - 1;
        const int src = (w == 0) ? imt - 4 : (w == 1) ? imt - 3 : (w == 2) ? 2 : 3;
        P[(size_t)j * imt + dst] = P[(size_t)j * imt + src];
    } else {
        const int r = t - ncol, w = r / imt, i = r % imt;
        const int j = (w == 0) ? 0 : (w == 1) ? 1 : (w == 2) ? jmt - 2 : jmt - 1;
        P[(size_t)j * imt + i] = 0.0;
    }
}

}  // namespace
}  // namespace licom
