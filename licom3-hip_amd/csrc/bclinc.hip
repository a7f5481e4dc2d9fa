// bclinc.hip -- the baroclinic momentum update bclinc_ on gfx950
// (include/licom_bclinc.h).
//
// The reference runs up to five launches per stage over whole arrays with
// eight 3-D scratch arrays (wka1, wka2, a8..f8).  Here every kernel is one
// lane per column (i-consecutive lanes) marching k:
//   k_bcl_prep     bclinc12 (stresses) and bclinc11 (h0bf, the free-surface
//                  term into wka where the next kernel does not overwrite it);
//   k_bcl_a        the pressure gradient and Coriolis of bclinc2 / bclinc6,
//                  with the hydrostatic pressure (bclinc11's wka1) and the
//                  free-surface term (wka2) of the four T columns around the
//                  U point formed on the fly from running column sums, then
//                  the explicit update and the implicit vertical viscosity
//                  of v (bclinc2 + bclinc3) or of u and v (bclinc71/72);
//   k_bcl_b        bclinc4 (v's barotropic replacement and filter, then u's
//                  explicit update and solve) or bclinc81/82;
//   k_bcl_c        bclinc5.
// The tridiagonal solves eliminate bottom-up and substitute top-down as the
// reference does; the elimination's two coefficient columns go through a
// scratch array pair (the lane's own cells, read back in reverse order).
// Operation order follows the reference statements (-ffp-contract=off), so
// results are identical bit for bit.
#include <hip/hip_runtime.h>

#include <new>
#include <vector>

#include "licom_bclinc.h"
#include "tracer_internal.h"

namespace licom {
namespace {

struct BGrid {
    int imt, jmt, km;
    size_t n2, n3;
    double od0, c0f, cag, sag, onbb, dtc, dtc2, afc1, afc2;
    const int *kmu;
    const double *snlat, *ohbt, *ohbu, *fcor, *dxur, *dyur, *epea, *epeb, *epla, *eplb;
    const double *dzp, *odzt, *odzp, *zkt;
    double *e8, *f8;   // solve scratch [km][j][i]
};

__device__ __forceinline__ bool interior2(const BGrid &g, int j, int i) {
    return i >= 2 && i < g.imt - 2 && j >= 2 && j < g.jmt - 2;
}

// The implicit vertical viscosity of one U column (bclinc3_cu :197-259,
// bclinc4_cu :301-374, bclinc71_cu :498-576, bclinc72_cu :589-667): x holds
// the explicit update at levels 0..kz-1 and receives the solution there.
__device__ void column_solve(const BGrid &g, const double *__restrict__ akmu, const double *__restrict__ viv,
                             double *__restrict__ x, size_t q, int kz, double dt, double sbc, double bbc) {
    const size_t n2 = g.n2;
    const double *odzt = g.odzt, *odzp = g.odzp;
    // elimination from the bottom: e8(k), f8(k) from level k+1's
    double en = 0.0, fn = 0.0;   // e8(kz) = f8(kz) = 0
    for (int k = kz - 1; k >= 0; --k) {
        const size_t c = (size_t)k * n2 + q;
        double a8, b8, c8, d8;
        if (k == kz - 1) {
            // bottom condition; a8 as the level above it set it
            a8 = k >= 1 ? akmu[c - n2] * odzt[k] * odzp[k] * dt * 0.5 : odzp[0] * dt * 0.5;
            b8 = 1.0 + a8;
            c8 = odzp[k] * dt * 0.5;
            d8 = x[c] - bbc * odzp[k] * dt * 0.5;
        } else if (k == 0) {
            a8 = odzp[0] * dt * 0.5;
            c8 = akmu[c] * odzt[1] * odzp[0] * dt * 0.5;
            b8 = 1.0 + c8;
            d8 = x[c];
        } else {
            a8 = akmu[c - n2] * odzt[k] * odzp[k] * dt * 0.5;
            c8 = akmu[c] * odzt[k + 1] * odzp[k] * dt * 0.5;
            b8 = 1.0 + a8 + c8;
            d8 = x[c];
        }
        const double g0 = 1.0 / (b8 - c8 * en);
        const double e = a8 * g0;
        const double fv = (d8 + c8 * fn) * g0;
        g.e8[c] = e;
        g.f8[c] = fv;
        en = e;
        fn = fv;
    }
    // substitution from the surface
    double prev = sbc;
    for (int k = 0; k < kz; ++k) {
        const size_t c = (size_t)k * n2 + q;
        const double v = (g.e8[c] * prev + g.f8[c]) * viv[c];
        x[c] = v;
        prev = v;
    }
}

// bclinc12_cu (bclinc.cpp:28-62) and bclinc11_cu (:64-104)
__global__ __launch_bounds__(256) void k_bcl_prep(BGrid g, licom_bclinc_fields f, int isc) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
    const int imt = g.imt, jmt = g.jmt, km = g.km;
    if (i >= imt || j >= jmt) return;
    const size_t n2 = g.n2, q = (size_t)j * imt + i;
    if (i >= 1 && i < imt - 1 && j >= 1 && j < jmt - 1) {
        const int kmb = g.kmu[q] - 1;
        if (kmb >= 0) {
            const size_t c = (size_t)kmb * n2 + q;
            const double ub = f.up[c], vb = f.vp[c], sn = g.snlat[q];
            const double temp = g.c0f * sqrt(ub * ub + vb * vb);
            f.sbcx[q] = f.su[q] * g.od0;
            f.sbcy[q] = f.sv[q] * g.od0;
            f.bbcx[q] = temp * (ub * g.cag + sn * vb * g.sag);
            f.bbcy[q] = temp * (vb * g.cag - sn * ub * g.sag);
        }
    }
    double aa = 0;
    if (isc != 0) aa = 0.5;
    const double h0bf = f.h0bf[q] * g.onbb;
    f.h0bf[q] = h0bf;
    // stage 0 of the leapfrog overwrites wka in the interior right after
    if (isc >= 1 && interior2(g, j, i)) return;
    const double work = aa * h0bf + (1.0 - aa) * f.h0bl[q], ohbt = g.ohbt[q];
    for (int k = 0; k < km; ++k) {
        const size_t c = (size_t)k * n2 + q;
        f.wka[c] = (1.0 + ohbt * g.zkt[k]) * work * f.vit[c];
    }
}

// bclinc11's T-column state for one T point of a 64x4 block of U columns
// (rows j0..j0+4, columns i0-1..i0+63; each T point serves four U columns),
// carried down the column by the lane that owns the T point, with its vit and
// gg of the next level loaded a level ahead
constexpr int BTX = 65, BTN = BTX * 5;
struct TCol {
    double wkk, work, ohbt, vt, gk;
    size_t q;
    int e;          // index among the block's T points (BTN: the lane owns none)
    bool ok;        // inside the array
};

__device__ __forceinline__ TCol tcol(const BGrid &g, const licom_bclinc_fields &f, int i0, int j0, int e,
                                     double aa) {
    TCol t;
    t.e = e < BTN ? e : BTN;
    t.ok = false;
    t.q = 0;
    t.wkk = t.work = t.ohbt = t.vt = t.gk = 0.0;
    if (e >= BTN) return t;
    const int tr = e / BTX, it = i0 - 1 + (e - tr * BTX), jt = j0 + tr;
    if (it < 0 || it >= g.imt || jt >= g.jmt) return t;
    const size_t q = (size_t)jt * g.imt + it;
    t.ok = true;
    t.q = q;
    t.work = aa * f.h0bf[q] + (1.0 - aa) * f.h0bl[q];   // h0bf already scaled by k_bcl_prep
    t.wkk = (f.psa[q] * g.od0 + t.work * 9.806e0) * f.vit[q];
    t.ohbt = g.ohbt[q];
    t.vt = f.vit[q];
    t.gk = f.gg[q];
    return t;
}

// level k of the recursion (bclinc11_cu) and the wka1/wka2 values there
__device__ __forceinline__ void tcol_level(const BGrid &g, const licom_bclinc_fields &f, TCol &t, int k,
                                           double dzk, double zk, double *W1, double *W2, double *G) {
    if (t.e >= BTN) return;
    double w1 = 0.0, w2 = 0.0, gk = 0.0;
    if (t.ok) {
        const double vt = t.vt;
        gk = t.gk;
        if (k + 1 < g.km) {
            const size_t cn = (size_t)(k + 1) * g.n2 + t.q;
            t.vt = f.vit[cn];
            t.gk = f.gg[cn];
        }
        const double nxt = t.wkk - gk * dzk * vt;
        w1 = 0.5 * (t.wkk + nxt);
        w2 = (1.0 + t.ohbt * zk) * t.work * vt;
        t.wkk = nxt;
    }
    W1[t.e] = w1;
    W2[t.e] = w2;
    G[t.e] = gk;
}

// bclinc2_cu / bclinc6_cu (bclinc.cpp:106-185, 418-486), then bclinc3_cu or
// bclinc71_cu + bclinc72_cu.  Per level the block's T points publish their
// wka1, wka2 and gg in LDS (double-buffered, one barrier per level); every
// U column reads its four.
template <bool EULER>
__global__ __launch_bounds__(256) void k_bcl_a(BGrid g, licom_bclinc_fields f, int isc) {
    __shared__ double sW1[2][BTN], sW2[2][BTN], sG[2][BTN];
    const int tx = threadIdx.x, ty = threadIdx.y, tl = tx + 64 * ty;
    const int i0 = blockIdx.x * 64, j0 = blockIdx.y * 4;
    const int i = i0 + tx, j = j0 + ty;
    const int imt = g.imt, jmt = g.jmt, km = g.km;
    const size_t n2 = g.n2;
    const bool grd = i >= 1 && i < imt && j < jmt - 1;
    const size_t q = grd ? (size_t)j * imt + i : 0;
    const bool in2 = grd && interior2(g, j, i);
    double aa = 0;
    if (isc != 0) aa = 0.5;
    TCol T0 = tcol(g, f, i0, j0, tl, aa), T1 = tcol(g, f, i0, j0, tl + 256, aa);
    int kmb = -1;
    double fc = 0, dxur = 0, dyur = 0, ea = 0, eb = 0;
    if (grd) {
        kmb = g.kmu[q] - 1;
        fc = g.fcor[q];
        dxur = g.dxur[q];
        dyur = g.dyur[q];
        ea = isc == 0 ? g.epea[q] : g.epla[q];
        eb = isc == 0 ? g.epeb[q] : g.eplb[q];
    }
    const double dt = EULER ? g.dtc : g.dtc2;
    // this U column's T points (j, i), (j, i-1), (j+1, i), (j+1, i-1)
    const int e0 = ty * BTX + tx + 1, e1 = e0 - 1, e2 = e0 + BTX, e3 = e2 - 1;
    for (int k = 0; k < km; ++k) {
        const int b = k & 1;
        const double dzk = g.dzp[k], zk = g.zkt[k];
        tcol_level(g, f, T0, k, dzk, zk, sW1[b], sW2[b], sG[b]);
        tcol_level(g, f, T1, k, dzk, zk, sW1[b], sW2[b], sG[b]);
        __syncthreads();
        if (!grd) continue;
        const size_t c = (size_t)k * n2 + q;
        const double *W1 = sW1[b], *W2 = sW2[b], *G = sG[b];
        double gradx = fc * f.vp[c];
        double grady = -fc * f.up[c];
        if (kmb >= k) {
            const double ggu = 0.25 * (G[e0] + G[e1] + G[e2] + G[e3]);
            gradx += dxur * 0.5 * (W1[e2] - W1[e1] - W1[e3] + W1[e0]);
            grady += dyur * 0.5 * (W1[e2] - W1[e1] + W1[e3] - W1[e0]);
            gradx -= ggu * dxur * 0.5 * (W2[e2] - W2[e1] - W2[e3] + W2[e0]);
            grady -= ggu * dyur * 0.5 * (W2[e2] - W2[e1] + W2[e3] - W2[e0]);
        }
        grady -= f.dlv[c];
        gradx -= f.dlu[c];
        const double wk1 = -(ea * grady + eb * gradx);
        const double wk2 = -(ea * gradx - eb * grady);
        const double vv = f.viv[c];
        const double dlv = wk1 * vv, dlu = wk2 * vv;
        f.dlv[c] = dlv;
        f.dlu[c] = dlu;
        if (in2) {
            // the explicit update, kept where the solve does not reach
            if (EULER) {
                f.u[c] = f.up[c] + dlu * dt;
                f.v[c] = f.vp[c] + dlv * dt;
            } else {
                f.wka[c] = f.vp[c] + dlv * dt;
            }
        }
    }
    if (!in2) return;
    const int kz = kmb + 1;
    if (kz <= 0) return;
    if (EULER) {
        column_solve(g, f.akmu, f.viv, f.u, q, kz, dt, f.sbcx[q], f.bbcx[q]);
        column_solve(g, f.akmu, f.viv, f.v, q, kz, dt, f.sbcy[q], f.bbcy[q]);
    } else {
        column_solve(g, f.akmu, f.viv, f.wka, q, kz, dt, f.sbcy[q], f.bbcy[q]);
    }
}

// A column of x in registers with its depth sum in bclinc4/5's product order
// (dzp ohbu x viv), x and viv read once, not again for the update; viv as a
// bit mask when every level's is +0.0 or 1.0.
struct Column {
    double x[LICOM_MAX_KM];
    uint64_t vone;
    bool v01;
    double work;
};

__device__ __forceinline__ void load_column(const BGrid &g, const double *__restrict__ x,
                                            const double *__restrict__ viv, size_t q, double ohbu, Column &C) {
    const size_t n2 = g.n2;
    const int km = g.km;
    C.vone = 0;
    C.v01 = true;
    C.work = 0.0;
#pragma unroll
    for (int k = 0; k < LICOM_MAX_KM; ++k) {
        if (k < km) {
            const size_t c = (size_t)k * n2 + q;
            const double xv = x[c], vv = viv[c];
            C.x[k] = xv;
            C.work += g.dzp[k] * ohbu * xv * vv;
            const uint64_t bits = __builtin_bit_cast(uint64_t, vv);
            if (bits == 0x3FF0000000000000ull)
                C.vone |= 1ull << k;
            else if (bits != 0)
                C.v01 = false;
        }
    }
}

__device__ __forceinline__ double viv_of(const Column &C, const double *__restrict__ viv, size_t c, int k) {
    return C.v01 ? (((C.vone >> k) & 1ull) ? 1.0 : 0.0) : viv[c];
}

// bclinc4_cu (bclinc.cpp:266-379) or bclinc81_cu + bclinc82_cu (:670-726)
template <bool EULER>
__global__ __launch_bounds__(256) void k_bcl_b(BGrid g, licom_bclinc_fields f) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
    const int imt = g.imt, jmt = g.jmt, km = g.km;
    if (i >= imt || j >= jmt) return;
    const size_t n2 = g.n2, q = (size_t)j * imt + i;
    const double ohbu = g.ohbu[q];
    if (EULER) {
        // u and v as the caller's halo exchange left them; two plain passes
        // (the register column of both velocities would take 256 VGPRs)
        for (int pass = 0; pass < 2; ++pass) {
            double *x = pass ? f.v : f.u, *tf = pass ? f.vtf : f.utf;
            const double b = pass ? f.vb[q] : f.ub[q];
            double work = 0;
            for (int k = 0; k < km; ++k) {
                const size_t c = (size_t)k * n2 + q;
                work += g.dzp[k] * ohbu * f.viv[c] * x[c];
            }
            for (int k = 0; k < km; ++k) {
                const size_t c = (size_t)k * n2 + q;
                const double v = (x[c] - work + b) * f.viv[c];
                x[c] = v;
                tf[c] += v;
            }
        }
        return;
    }
    Column C;
    load_column(g, f.wka, f.viv, q, ohbu, C);
    const bool in2 = interior2(g, j, i);
    const double vb = f.vb[q];
#pragma unroll
    for (int k = 0; k < LICOM_MAX_KM; ++k) {
        if (k < km) {
            const size_t c = (size_t)k * n2 + q;
            const double w = (C.x[k] - C.work + vb) * viv_of(C, f.viv, c, k);
            f.vp[c] = g.afc2 * f.v[c] + g.afc1 * (f.vp[c] + w);
            f.v[c] = w;
            f.wka[c] = in2 ? f.up[c] + f.dlu[c] * g.dtc2 : w;
        }
    }
    if (!in2) return;
    const int kz = g.kmu[q];
    if (kz > 0) column_solve(g, f.akmu, f.viv, f.wka, q, kz, g.dtc2, f.sbcx[q], f.bbcx[q]);
}

// bclinc5_cu (bclinc.cpp:381-416)
__global__ __launch_bounds__(256) void k_bcl_c(BGrid g, licom_bclinc_fields f) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
    const int imt = g.imt, jmt = g.jmt, km = g.km;
    if (i >= imt || j >= jmt) return;
    const size_t n2 = g.n2, q = (size_t)j * imt + i;
    const double ohbu = g.ohbu[q], ub = f.ub[q];
    Column C;
    load_column(g, f.wka, f.viv, q, ohbu, C);
#pragma unroll
    for (int k = 0; k < LICOM_MAX_KM; ++k) {
        if (k < km) {
            const size_t c = (size_t)k * n2 + q;
            const double w = (C.x[k] - C.work + ub) * viv_of(C, f.viv, c, k);
            f.up[c] = g.afc2 * f.u[c] + g.afc1 * (f.up[c] + w);
            f.u[c] = w;
            f.utf[c] += w;
            f.vtf[c] += f.v[c];
        }
    }
}

}  // namespace
}  // namespace licom

using namespace licom;

struct licom_bclinc {
    BGrid g{};
    hipStream_t stream = nullptr;
    void *mem = nullptr, *scratch = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool ran = false;
};

extern "C" {

int licom_bclinc_create(const licom_bclinc_grid *grid, licom_bclinc **out) {
    if (!grid || !out) {
        set_error("licom_bclinc_create: NULL argument");
        return LICOM_EINVAL;
    }
    *out = nullptr;
    const licom_bclinc_grid &G = *grid;
    if (G.imt < 5 || G.jmt < 5 || G.km < 2 || G.km > LICOM_MAX_KM) {
        set_error("licom_bclinc_create: invalid extents imt=%d jmt=%d km=%d (need imt, jmt >= 5, "
                  "2 <= km <= %d)", G.imt, G.jmt, G.km, LICOM_MAX_KM);
        return LICOM_EINVAL;
    }
    const double *src2[] = {G.snlat, G.ohbt, G.ohbu, G.fcor, G.dxur, G.dyur, G.epea, G.epeb, G.epla, G.eplb};
    const double *src1[] = {G.dzp, G.odzt, G.odzp, G.zkt};
    constexpr int n2d = sizeof(src2) / sizeof(src2[0]), n1d = sizeof(src1) / sizeof(src1[0]);
    for (const double *p : src2)
        if (!p) {
            set_error("licom_bclinc_create: a grid array is NULL");
            return LICOM_EINVAL;
        }
    for (const double *p : src1)
        if (!p) {
            set_error("licom_bclinc_create: a grid array is NULL");
            return LICOM_EINVAL;
        }
    if (!G.kmu) {
        set_error("licom_bclinc_create: a grid array is NULL");
        return LICOM_EINVAL;
    }
    const size_t n2 = (size_t)G.imt * G.jmt, km = (size_t)G.km;
    for (size_t p = 0; p < n2; ++p)
        if (G.kmu[p] < 0 || G.kmu[p] > G.km) {
            set_error("licom_bclinc_create: kmu outside 0..%d at cell %zu", G.km, p);
            return LICOM_EINVAL;
        }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        set_error("no HIP device available");
        return LICOM_ENODEV;
    }
    licom_bclinc *b = new (std::nothrow) licom_bclinc();
    if (!b) return LICOM_ENOMEM;
    const size_t nd = n2d * n2 + n1d * km, bytes = nd * sizeof(double) + n2 * sizeof(int);
    std::vector<double> h(nd);
    for (int a = 0; a < n2d; ++a) std::copy(src2[a], src2[a] + n2, h.begin() + a * n2);
    for (int a = 0; a < n1d; ++a) std::copy(src1[a], src1[a] + km, h.begin() + n2d * n2 + a * km);
    int rc = hip_status(hipMalloc(&b->mem, bytes), "hipMalloc");
    char *base = static_cast<char *>(b->mem);
    if (!rc) rc = hip_status(hipMemcpy(base, h.data(), nd * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
    int *kmu = reinterpret_cast<int *>(base + nd * sizeof(double));
    if (!rc) rc = hip_status(hipMemcpy(kmu, G.kmu, n2 * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy");
    if (!rc) rc = hip_status(hipMalloc(&b->scratch, 2 * n2 * km * sizeof(double)), "hipMalloc");
    if (!rc) rc = hip_status(hipEventCreate(&b->ev0), "hipEventCreate");
    if (!rc) rc = hip_status(hipEventCreate(&b->ev1), "hipEventCreate");
    if (rc) {
        licom_bclinc_destroy(b);
        return rc;
    }
    const double *d = reinterpret_cast<const double *>(base);
    BGrid &g = b->g;
    g.imt = G.imt;
    g.jmt = G.jmt;
    g.km = G.km;
    g.n2 = n2;
    g.n3 = n2 * km;
    g.od0 = G.od0;
    g.c0f = G.c0f;
    g.cag = G.cag;
    g.sag = G.sag;
    g.onbb = G.onbb;
    g.dtc = G.dtc;
    g.dtc2 = G.dtc2;
    g.afc1 = G.afc1;
    g.afc2 = G.afc2;
    const double **dst2[] = {&g.snlat, &g.ohbt, &g.ohbu, &g.fcor, &g.dxur, &g.dyur, &g.epea, &g.epeb, &g.epla, &g.eplb};
    const double **dst1[] = {&g.dzp, &g.odzt, &g.odzp, &g.zkt};
    for (int a = 0; a < n2d; ++a) *dst2[a] = d + a * n2;
    for (int a = 0; a < n1d; ++a) *dst1[a] = d + n2d * n2 + a * km;
    g.kmu = kmu;
    g.e8 = static_cast<double *>(b->scratch);
    g.f8 = g.e8 + g.n3;
    *out = b;
    return LICOM_OK;
}

int licom_bclinc_set_stream(licom_bclinc *b, void *hip_stream) {
    if (!b) {
        set_error("licom_bclinc_set_stream: NULL context");
        return LICOM_EINVAL;
    }
    b->stream = (hipStream_t)hip_stream;
    return LICOM_OK;
}

int licom_bclinc_stage(licom_bclinc *b, const licom_bclinc_fields *f, int isc, int stage) {
    if (!b || !f) {
        set_error("licom_bclinc_stage: NULL argument");
        return LICOM_EINVAL;
    }
    if (stage < 0 || stage > 2 || (stage == 2 && isc < 1)) {
        set_error("licom_bclinc_stage: no stage %d at isc=%d (0..2 with isc >= 1, 0..1 with isc < 1)", stage,
                  isc);
        return LICOM_EINVAL;
    }
    const void *need[] = {f->vit, f->viv, f->gg, f->akmu, f->su, f->sv,  f->psa,  f->h0bl, f->ub,
                          f->vb,  f->h0bf, f->u, f->v,    f->up, f->vp, f->dlu, f->dlv, f->wka,
                          f->utf, f->vtf, f->sbcx, f->sbcy, f->bbcx, f->bbcy};
    for (const void *p : need)
        if (!p) {
            set_error("licom_bclinc_stage: a field pointer is NULL");
            return LICOM_EINVAL;
        }
    const BGrid &g = b->g;
    const dim3 blk(64, 4), grd((g.imt + 63) / 64, (g.jmt + 3) / 4);
    const bool euler = isc < 1;
    if (int rc = hip_status(hipEventRecord(b->ev0, b->stream), "hipEventRecord")) return rc;
    if (stage == 0) {
        k_bcl_prep<<<grd, blk, 0, b->stream>>>(g, *f, isc);
        if (int rc = hip_status(hipGetLastError(), "k_bcl_prep launch")) return rc;
        if (euler)
            k_bcl_a<true><<<grd, blk, 0, b->stream>>>(g, *f, isc);
        else
            k_bcl_a<false><<<grd, blk, 0, b->stream>>>(g, *f, isc);
        if (int rc = hip_status(hipGetLastError(), "k_bcl_a launch")) return rc;
    } else if (stage == 1) {
        if (euler)
            k_bcl_b<true><<<grd, blk, 0, b->stream>>>(g, *f);
        else
            k_bcl_b<false><<<grd, blk, 0, b->stream>>>(g, *f);
        if (int rc = hip_status(hipGetLastError(), "k_bcl_b launch")) return rc;
    } else {
        k_bcl_c<<<grd, blk, 0, b->stream>>>(g, *f);
        if (int rc = hip_status(hipGetLastError(), "k_bcl_c launch")) return rc;
    }
    if (int rc = hip_status(hipEventRecord(b->ev1, b->stream), "hipEventRecord")) return rc;
    b->ran = true;
    return LICOM_OK;
}

int licom_bclinc_sync(licom_bclinc *b) {
    if (!b) {
        set_error("licom_bclinc_sync: NULL context");
        return LICOM_EINVAL;
    }
    return hip_status(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
}

int licom_bclinc_last_ms(licom_bclinc *b, double *ms) {
    if (!b || !ms) {
        set_error("licom_bclinc_last_ms: NULL argument");
        return LICOM_EINVAL;
    }
    *ms = 0.0;
    if (!b->ran) return LICOM_OK;
    if (int rc = hip_status(hipEventSynchronize(b->ev1), "hipEventSynchronize")) return rc;
    float v = 0.0f;
    if (int rc = hip_status(hipEventElapsedTime(&v, b->ev0, b->ev1), "hipEventElapsedTime")) return rc;
    *ms = v;
    return LICOM_OK;
}

void licom_bclinc_destroy(licom_bclinc *b) {
    if (!b) return;
    if (b->ran) (void)hipStreamSynchronize(b->stream);
    if (b->mem) (void)hipFree(b->mem);
    if (b->scratch) (void)hipFree(b->scratch);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    delete b;
}

}  // extern "C"
