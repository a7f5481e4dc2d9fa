// readyc.hip -- the vertical-mixing preparation of readyc_ on gfx950
// (include/licom_readyc.h): velocity shear and Richardson numbers for the
// Canuto producer (readyc1, readyc1_2), and AKMT averaged to U points after
// it (readyc2).
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

}  // namespace
}  // namespace licom

using namespace licom;

struct licom_readyc {
    CGrid g{};
    hipStream_t stream = nullptr;
    void *mem = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool ran = false;
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
    if (r->ev0) (void)hipEventDestroy(r->ev0);
    if (r->ev1) (void)hipEventDestroy(r->ev1);
    delete r;
}

}  // extern "C"
