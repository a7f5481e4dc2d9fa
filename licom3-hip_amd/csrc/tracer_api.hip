// tracer_api.hip -- C ABI of liblicom_tracer.so: context set-up, the static
// coefficient derivation, entry points and error reporting.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "tracer_internal.h"

namespace licom {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int hip_status(hipError_t e, const char *what) {
    if (e == hipSuccess) return LICOM_OK;
    set_error("%s: %s", what, hipGetErrorString(e));
    return (e == hipErrorOutOfMemory) ? LICOM_ENOMEM : LICOM_EHIP;
}

namespace {

constexpr double kRadius = 6371000.0;   // constant_mod.F90:31

template <class T>
void halo_fill(std::vector<T> &a, int imt, int jmt) {
    for (int j = 2; j < jmt - 2; ++j) {
        T *r = a.data() + (size_t)j * imt;
        r[0] = r[imt - 4];
        r[1] = r[imt - 3];
        r[imt - 2] = r[2];
        r[imt - 1] = r[3];
    }
    for (int i = 0; i < imt; ++i) {
        a[i] = a[imt + i] = T(0);
        a[(size_t)(jmt - 2) * imt + i] = a[(size_t)(jmt - 1) * imt + i] = T(0);
    }
}

template <class T>
int scratch(licom_ctx *c, T **dst, size_t n) {
    void *p = nullptr;
    int rc = hip_status(hipMalloc(&p, n * sizeof(T)), "hipMalloc scratch");
    if (rc) return rc;
    c->allocs.push_back(p);
    c->scratch_bytes += n * sizeof(T);
    *dst = static_cast<T *>(p);
    return hip_status(hipMemset(p, 0, n * sizeof(T)), "hipMemset scratch");
}

// Derived static fields.  Formulas and their operation order follow
// boundary.F90:29-52 (KMU, VIT), grid.F90:1096-1106 (DZP, ODZP, ODZT),
// grid.F90:466 (TAREA_R), grid.F90:2029-2050 (OHBT, OHBU),
// hmix_del4.F90:357-392 / hmix_del2.F90:346-385 (AHF, DTN/S/E/W; eoshift is
// end-off) and grid.F90:651 (area_t over the physical ocean cells).
int derive(licom_ctx *c, const licom_grid *G) {
    const int imt = c->imt, jmt = c->jmt, km = c->km;
    const size_t n2 = c->n2;
    c->h_kmt.assign(G->kmt, G->kmt + n2);
    halo_fill(c->h_kmt, imt, jmt);
    for (size_t q = 0; q < n2; ++q)
        if (c->h_kmt[q] < 0 || c->h_kmt[q] > km) {
            set_error("kmt out of range at %zu: %d (km=%d)", q, c->h_kmt[q], km);
            return LICOM_EINVAL;
        }
    c->h_kmu.assign(n2, 0);
    for (int j = 0; j < jmt - 1; ++j)
        for (int i = 1; i < imt; ++i) {
            const int32_t *k = c->h_kmt.data();
            int m = k[(size_t)j * imt + i - 1];
            m = std::min(m, k[(size_t)j * imt + i]);
            m = std::min(m, k[(size_t)(j + 1) * imt + i - 1]);
            m = std::min(m, k[(size_t)(j + 1) * imt + i]);
            c->h_kmu[(size_t)j * imt + i] = m;
        }
    halo_fill(c->h_kmu, imt, jmt);

    c->h_dzp.resize(km);
    c->h_odzp.resize(km);
    c->h_odzt.resize(km);
    std::vector<double> zkt(km);
    for (int k = 0; k < km; ++k) {
        c->h_dzp[k] = G->zkp[k] - G->zkp[k + 1];
        zkt[k] = (G->zkp[k] + G->zkp[k + 1]) * 0.5;
        c->h_odzp[k] = 1.0 / c->h_dzp[k];
        if (!(c->h_dzp[k] > 0.0)) {
            set_error("zkp must decrease strictly (level %d)", k);
            return LICOM_EINVAL;
        }
    }
    c->h_odzt[0] = 2.0 * c->h_odzp[0];
    for (int k = 1; k < km; ++k) c->h_odzt[k] = 1.0 / (zkt[k - 1] - zkt[k]);

    c->h_tarea_r.resize(n2);
    for (size_t q = 0; q < n2; ++q) {
        if (!(G->tarea[q] > 0.0)) {
            set_error("tarea must be positive everywhere (cell %zu)", q);
            return LICOM_EINVAL;
        }
        c->h_tarea_r[q] = 1.0 / G->tarea[q];
    }
    c->h_ohbt.resize(n2);
    c->h_ohbu.resize(n2);
    for (size_t q = 0; q < n2; ++q) {
        double s = 0.0;
        for (int k = 0; k < km; ++k) s = s + (k < c->h_kmt[q] ? 1.0 : 0.0) * c->h_dzp[k];
        c->h_ohbt[q] = (s > 0.0) ? 1.0 / s : 0.0;
        s = 0.0;
        for (int k = 0; k < km; ++k) s = s + (k < c->h_kmu[q] ? 1.0 : 0.0) * c->h_dzp[k];
        c->h_ohbu[q] = (s > 0.0) ? 1.0 / s : 0.0;
    }

    const double pi = 4.0 * std::atan(1.0);
    const double dxg = 2.0 * kRadius * pi / (double)G->imt_global;
    c->h_ahf.resize(n2);
    c->h_dtn.resize(n2);
    c->h_dts.resize(n2);
    c->h_dte.resize(n2);
    c->h_dtw.resize(n2);
    std::vector<double> w1(n2);
    auto at = [imt](int j, int i) { return (size_t)j * imt + i; };
    if (c->p.hmix == LICOM_HMIX_DEL4) {
        for (size_t q = 0; q < n2; ++q) c->h_ahf[q] = std::pow(G->tarea[q] / std::pow(dxg, 2.0), 1.5);
        for (size_t q = 0; q < n2; ++q) w1[q] = G->hts[q] / G->hue[q];
    } else {
        for (size_t q = 0; q < n2; ++q) c->h_ahf[q] = std::sqrt(G->tarea[q]) / dxg;
        for (int j = 0; j < jmt; ++j)
            for (int i = 0; i < imt; ++i) {
                const double an = (j + 1 < jmt) ? c->h_ahf[at(j + 1, i)] : 0.0;
                w1[at(j, i)] = G->hts[at(j, i)] / G->hue[at(j, i)] * 0.5 * (c->h_ahf[at(j, i)] + an);
            }
    }
    for (int j = 0; j < jmt; ++j)
        for (int i = 0; i < imt; ++i) {
            c->h_dts[at(j, i)] = w1[at(j, i)] * c->h_tarea_r[at(j, i)];
            c->h_dtn[at(j, i)] = (j >= 1 ? w1[at(j - 1, i)] : 0.0) * c->h_tarea_r[at(j, i)];
        }
    if (c->p.hmix == LICOM_HMIX_DEL4) {
        for (size_t q = 0; q < n2; ++q) w1[q] = G->htw[q] / G->hun[q];
    } else {
        for (int j = 0; j < jmt; ++j)
            for (int i = 0; i < imt; ++i) {
                const double aw = (i >= 1) ? c->h_ahf[at(j, i - 1)] : 0.0;
                w1[at(j, i)] = G->htw[at(j, i)] / G->hun[at(j, i)] * 0.5 * (c->h_ahf[at(j, i)] + aw);
            }
    }
    for (int j = 0; j < jmt; ++j)
        for (int i = 0; i < imt; ++i) {
            c->h_dtw[at(j, i)] = w1[at(j, i)] * c->h_tarea_r[at(j, i)];
            c->h_dte[at(j, i)] = (i + 1 < imt ? w1[at(j, i + 1)] : 0.0) * c->h_tarea_r[at(j, i)];
        }
    double a = 0.0;
    for (int j = 2; j < jmt - 2; ++j)
        for (int i = 2; i < imt - 2; ++i)
            if (c->h_kmt[at(j, i)] >= 1) a = a + G->tarea[at(j, i)];
    c->area_t = a;
    return LICOM_OK;
}

int check_ptr(const void *p, const char *name) {
    if (!p) {
        set_error("required pointer %s is NULL", name);
        return LICOM_EINVAL;
    }
    return LICOM_OK;
}

int check_fields(const licom_fields *f) {
    const char *req[] = {"at", "atb", "akt", "utf", "vtf", "utl", "vtl", "ws", "h0f", "h0l"};
    const void *ptrs[] = {f->at, f->atb, f->akt, f->utf, f->vtf, f->utl, f->vtl, f->ws, f->h0f, f->h0l};
    for (int t = 0; t < 10; ++t)
        if (int rc = check_ptr(ptrs[t], req[t])) return rc;
    return LICOM_OK;
}

}  // namespace

// Checks the arguments of an FP64 step and settles its path: STAGED is added
// where the fused kernels cannot run the layout, and split is set when
// convadj must run as a separate pass after the step.
int plan_step(licom_ctx *c, const licom_fields *f, int &flags, bool &split, const char *who) {
    split = false;
    if (!c || !f) {
        set_error("%s: NULL argument", who);
        return LICOM_EINVAL;
    }
    if (int rc = check_fields(f)) return rc;
    if (c->p.boundary_restore == 1 && !f->restore) {
        set_error("boundary_restore=1 needs the restore field");
        return LICOM_EINVAL;
    }
    // The fused k-march stages planes in 16-byte chunks: it needs an even imt
    // and 16-byte aligned at, atb, ws and akt; other layouts take the
    // per-stage kernels (same results).
    const uintptr_t al = reinterpret_cast<uintptr_t>(f->at) | reinterpret_cast<uintptr_t>(f->atb) |
                         reinterpret_cast<uintptr_t>(f->ws) | reinterpret_cast<uintptr_t>(f->akt);
    if (c->imt % 2 != 0 || (al & 15) != 0) flags |= LICOM_STEP_STAGED;
    if (flags & LICOM_STEP_ICESNOW) {
        if (!c->qice) {
            set_error("LICOM_STEP_ICESNOW needs licom_tracer_set_icesnow first");
            return LICOM_EINVAL;
        }
        // the S left in vtl is the one before icesnow: the per-stage path
        // runs icesnow after that copy
        if (flags & LICOM_STEP_VTL_SCRATCH) flags |= LICOM_STEP_STAGED;
    } else if ((flags & LICOM_STEP_VTL_SCRATCH) && (flags & LICOM_STEP_CONVADJ)) {
        // tracer_ leaves the new S in vtl before convadj_ mixes at and atb
        // (tracer.cpp:1880-1893; convadj.cpp never touches vtl): the step
        // runs without convadj, so that the copy into vtl sees the unmixed
        // S, then convadj over every column as convadj_() does
        split = true;
    }
    return LICOM_OK;
}

}  // namespace licom

using namespace licom;

extern "C" {

int licom_tracer_create(const licom_grid *G, const licom_params *P, licom_ctx **out) {
    if (!G || !P || !out) {
        set_error("licom_tracer_create: NULL argument");
        return LICOM_EINVAL;
    }
    *out = nullptr;
    if (G->nx < 4 || G->ny < 1 || G->km < 2 || G->km > LICOM_MAX_KM || G->imt_global < 1) {
        set_error("invalid grid extents nx=%d ny=%d km=%d (need nx>=4, ny>=1, 2<=km<=%d)", G->nx,
                  G->ny, G->km, LICOM_MAX_KM);
        return LICOM_EINVAL;
    }
    const char *req[] = {"kmt", "zkp", "htw", "hts", "hun", "hue", "tarea", "to", "so", "eosc"};
    const void *ptrs[] = {G->kmt, G->zkp, G->htw, G->hts, G->hun, G->hue, G->tarea, G->to, G->so, G->eosc};
    for (int t = 0; t < 10; ++t)
        if (int rc = check_ptr(ptrs[t], req[t])) return rc;
    if (P->hmix != LICOM_HMIX_DEL2 && P->hmix != LICOM_HMIX_DEL4) {
        set_error("hmix must be 2 (del2) or 4 (del4), got %d", P->hmix);
        return LICOM_EINVAL;
    }
    if (P->adv != LICOM_ADV_TSPAS && P->adv != LICOM_ADV_CENTERED && P->adv != LICOM_ADV_FLUX) {
        set_error("adv must be 0 (centered), 1 (flux) or 2 (tspas), got %d", P->adv);
        return LICOM_EINVAL;
    }
    if (P->adv == LICOM_ADV_FLUX && (!G->dxu || !G->dyu)) {
        set_error("the flux scheme needs the U-cell widths dxu and dyu");
        return LICOM_EINVAL;
    }
    if (P->adv == LICOM_ADV_CENTERED && P->nss < 1) {
        set_error("the centered scheme divides its diagnostics by nss >= 1 (nss=%d)", P->nss);
        return LICOM_EINVAL;
    }
    if (!(P->dts > 0.0)) {
        set_error("dts must be positive");
        return LICOM_EINVAL;
    }
    int dev = 0, ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        set_error("no HIP device available");
        return LICOM_ENODEV;
    }
    if (int rc = hip_status(hipGetDevice(&dev), "hipGetDevice")) return rc;
    // the no-allocation step's static scratch: its device address, once per
    // device (a failure only disables licom_tracer_step_noalloc)
    (void)noalloc_init(dev);

    licom_ctx *c = new (std::nothrow) licom_ctx();
    if (!c) return LICOM_ENOMEM;
    c->nx = G->nx;
    c->ny = G->ny;
    c->km = G->km;
    c->imt = G->nx + 4;
    c->jmt = G->ny + 4;
    c->n2 = (size_t)c->imt * c->jmt;
    c->n3 = c->n2 * c->km;
    c->solve_shape = c->nx >= 512 ? 2 : 0;   // one 128-column row per workgroup on wide grids
    c->p = *P;
    c->device = dev;
    int rc = derive(c, G);
    if (rc) {
        delete c;
        return rc;
    }
    const size_t n2 = c->n2, n3 = c->n3;
    const int km = c->km;
    KGrid &g = c->g;
    g.imt = c->imt;
    g.jmt = c->jmt;
    g.km = km;
    std::vector<double> quarter(n2, 0.25), pen0(km > 1 ? km - 1 : 1, 0.0);
    // the U-cell widths only feed the flux scheme's faces
    const bool flux = P->adv == LICOM_ADV_FLUX;
    std::vector<double> zero2(flux ? 0 : n2, 0.0);
    // The static fields and the step scratch share one device allocation: a
    // context costs one hipMalloc and one clear instead of ~36 allocations
    // with a memset each (context creation is on the request loop's
    // admission path).  Parts are planned here and placed by place_parts.
    struct Part {
        void *slot;             // address of the pointer to set
        const void *src;        // host data to copy, or NULL: cleared scratch
        size_t bytes;
    };
    std::vector<Part> parts;
    auto up_d = [&](const double **d, const double *h, size_t n) {
        parts.push_back({(void *)d, h, n * sizeof(double)});
        return 0;
    };
    auto up_i = [&](const int32_t **d, const int32_t *h, size_t n) {
        parts.push_back({(void *)d, h, n * sizeof(int32_t)});
        return 0;
    };
    auto plan_scratch = [&](auto **d, size_t n) {
        const size_t b = n * sizeof(**d);
        parts.push_back({(void *)d, nullptr, b});
        c->scratch_bytes += b;
        return 0;
    };
    auto place_parts = [&]() -> int {
        auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
        size_t total = 0, zero_from = 0;
        for (const Part &pt : parts) total += al(pt.bytes);
        char *base = nullptr;
        if (int r = hip_status(hipMalloc((void **)&base, total), "hipMalloc context")) return r;
        c->allocs.push_back(base);
        size_t off = 0;
        for (const Part &pt : parts)   // copies first, then one clear of the scratch
            if (pt.src) {
                void *v = base + off;
                std::memcpy(pt.slot, &v, sizeof v);
                if (int r = hip_status(hipMemcpy(v, pt.src, pt.bytes, hipMemcpyHostToDevice), "hipMemcpy"))
                    return r;
                off += al(pt.bytes);
            }
        zero_from = off;
        for (const Part &pt : parts)
            if (!pt.src) {
                void *v = base + off;
                std::memcpy(pt.slot, &v, sizeof v);
                off += al(pt.bytes);
            }
        return off > zero_from ? hip_status(hipMemset(base + zero_from, 0, off - zero_from), "hipMemset scratch")
                               : (int)LICOM_OK;
    };
    rc = 0;
    rc = rc ? rc : up_i(&g.kmt, c->h_kmt.data(), n2);
    rc = rc ? rc : up_i(&g.kmu, c->h_kmu.data(), n2);
    rc = rc ? rc : up_d(&g.tarea, G->tarea, n2);
    rc = rc ? rc : up_d(&g.tarea_r, c->h_tarea_r.data(), n2);
    rc = rc ? rc : up_d(&g.htw, G->htw, n2);
    rc = rc ? rc : up_d(&g.hts, G->hts, n2);
    rc = rc ? rc : up_d(&g.hun, G->hun, n2);
    rc = rc ? rc : up_d(&g.hue, G->hue, n2);
    rc = rc ? rc : up_d(&g.ohbt, c->h_ohbt.data(), n2);
    rc = rc ? rc : up_d(&g.ohbu, c->h_ohbu.data(), n2);
    rc = rc ? rc : up_d(&g.dtn, c->h_dtn.data(), n2);
    rc = rc ? rc : up_d(&g.dts, c->h_dts.data(), n2);
    rc = rc ? rc : up_d(&g.dte, c->h_dte.data(), n2);
    rc = rc ? rc : up_d(&g.dtw, c->h_dtw.data(), n2);
    rc = rc ? rc : up_d(&g.ahf, c->h_ahf.data(), n2);
    rc = rc ? rc : up_d(&g.au0, G->au0 ? G->au0 : quarter.data(), n2);
    rc = rc ? rc : up_d(&g.aus, G->aus ? G->aus : quarter.data(), n2);
    rc = rc ? rc : up_d(&g.auw, G->auw ? G->auw : quarter.data(), n2);
    rc = rc ? rc : up_d(&g.ausw, G->ausw ? G->ausw : quarter.data(), n2);
    rc = rc ? rc : up_d(&g.dzp, c->h_dzp.data(), km);
    rc = rc ? rc : up_d(&g.odzp, c->h_odzp.data(), km);
    rc = rc ? rc : up_d(&g.odzt, c->h_odzt.data(), km);
    rc = rc ? rc : up_d(&g.pen, G->pen ? G->pen : pen0.data(), km - 1);
    rc = rc ? rc : up_d(&g.to, G->to, km);
    rc = rc ? rc : up_d(&g.so, G->so, km);
    rc = rc ? rc : up_d(&g.eosc, G->eosc, (size_t)9 * km);
    rc = rc ? rc : up_d(&g.dxu, flux ? G->dxu : zero2.data(), n2);
    rc = rc ? rc : up_d(&g.dyu, flux ? G->dyu : zero2.data(), n2);

    KPar &k = c->kp;
    k.dts = P->dts;
    k.onbc = P->onbc;
    k.oncc = P->oncc;
    k.dwndmix = P->dwndmix;
    k.gamma = P->gamma;
    k.od0 = P->od0;
    k.od0cp = P->od0cp;
    k.ah = P->ah;
    k.hmix = P->hmix;
    k.boundary_restore = P->boundary_restore;
    k.sssnorm = P->sssnorm;
    k.adv = P->adv;
    k.nss = P->nss >= 1 ? P->nss : 1;

    Scratch &s = c->s;
    // SSSNORM partial-sum blocks: about 8 physical cells per lane, at most 1024
    {
        const size_t cells = (size_t)(c->imt - 4) * (c->jmt - 4);
        s.nblk = (int)std::min<size_t>(1024, std::max<size_t>(1, (cells + 2047) / 2048));
    }
    rc = rc ? rc : plan_scratch(&s.stf_pre, n2);
    rc = rc ? rc : plan_scratch(&s.work_u, n2);
    rc = rc ? rc : plan_scratch(&s.stf_n, 2 * n2);
    rc = rc ? rc : plan_scratch(&s.temp11, n2);
    rc = rc ? rc : plan_scratch(&s.partials, (size_t)2 * s.nblk);
    rc = rc ? rc : plan_scratch(&s.err_norm2, 1);
    rc = rc ? rc : plan_scratch(&s.uw, n3);
    rc = rc ? rc : plan_scratch(&s.vs, n3);
    rc = rc ? rc : plan_scratch(&s.ori, 2 * n3);
    rc = rc ? rc : plan_scratch(&s.mix_count, 2);
    rc = rc ? rc : place_parts();
    if (rc) {
        licom_tracer_destroy(c);
        return rc;
    }
    *out = c;
    return LICOM_OK;
}

int licom_tracer_set_stream(licom_ctx *c, void *stream) {
    if (!c) return LICOM_EINVAL;
    c->stream = static_cast<hipStream_t>(stream);
    return LICOM_OK;
}

int licom_tracer_step(licom_ctx *c, const licom_fields *f, const licom_diag *d, int flags) {
    bool split = false;
    if (int rc = plan_step(c, f, flags, split, "licom_tracer_step")) return rc;
    const int run = split ? flags & ~LICOM_STEP_CONVADJ : flags;
    int rc;
    if (run & LICOM_STEP_STAGED) {
        // scratch of the per-stage kernels, allocated on first use
        if (!c->s.flags) {
            rc = scratch(c, &c->s.flags, 2 * c->n3);
            if (!rc && c->p.hmix == LICOM_HMIX_DEL4) rc = scratch(c, &c->s.d2tk, 2 * c->n3);
            if (rc) return rc;
        }
        rc = launch_step_simple(c, f, d, run);
    } else {
        rc = launch_step_fused(c, f, d, run, c->s);
    }
    return (rc || !split) ? rc : launch_convadj(c, f, d);
}

int licom_tracer_step_mixed(licom_ctx *c, const licom_fields_mixed *f, const licom_diag_mixed *d,
                            int flags) {
    if (!c || !f) {
        set_error("licom_tracer_step_mixed: NULL argument");
        return LICOM_EINVAL;
    }
    const char *req[] = {"at", "atb", "akt", "utf", "vtf", "utl", "vtl", "ws", "h0f", "h0l"};
    const void *ptrs[] = {f->at, f->atb, f->akt, f->utf, f->vtf, f->utl, f->vtl, f->ws, f->h0f, f->h0l};
    for (int t = 0; t < 10; ++t)
        if (int rc = check_ptr(ptrs[t], req[t])) return rc;
    if (c->p.boundary_restore == 1 && !f->restore) {
        set_error("boundary_restore=1 needs the restore field");
        return LICOM_EINVAL;
    }
    if (flags & (LICOM_STEP_VTL_SCRATCH | LICOM_STEP_STAGED)) {
        set_error("licom_tracer_step_mixed: VTL_SCRATCH and STAGED are FP64-only");
        return LICOM_ENOTSUP;
    }
    if (c->kp.adv != LICOM_ADV_TSPAS) {
        set_error("licom_tracer_step_mixed: the mixed form is built for tspas only (adv=%d)", c->kp.adv);
        return LICOM_ENOTSUP;
    }
    if ((flags & LICOM_STEP_ICESNOW) && !c->qice) {
        set_error("LICOM_STEP_ICESNOW needs licom_tracer_set_icesnow first");
        return LICOM_EINVAL;
    }
    const uintptr_t al = reinterpret_cast<uintptr_t>(f->at) | reinterpret_cast<uintptr_t>(f->atb);
    if (c->imt % 2 != 0 || (al & 15) != 0) {
        set_error("licom_tracer_step_mixed: needs an even imt and 16-byte aligned at/atb");
        return LICOM_ENOTSUP;
    }
    return launch_step_mixed(c, f, d, flags);
}

int licom_convadj(licom_ctx *c, const licom_fields *f, const licom_diag *d) {
    if (!c || !f || !f->at || !f->atb) {
        set_error("licom_convadj: NULL argument");
        return LICOM_EINVAL;
    }
    return launch_convadj(c, f, d);
}

int licom_icesnow(licom_ctx *c, double *at, double *licomqice, int reset_qice) {
    if (!c || !at || !licomqice) {
        set_error("licom_icesnow: NULL argument");
        return LICOM_EINVAL;
    }
    return launch_icesnow(c, at, licomqice, reset_qice);
}

int licom_tracer_set_icesnow(licom_ctx *c, double *licomqice) {
    if (!c) return LICOM_EINVAL;
    c->qice = licomqice;
    return LICOM_OK;
}

int licom_tracer_sync(licom_ctx *c) {
    if (!c) return LICOM_EINVAL;
    return hip_status(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
}

int licom_grid_export(const licom_ctx *c, licom_grid_derived *o) {
    if (!c || !o) return LICOM_EINVAL;
    auto cp = [](void *dst, const void *src, size_t bytes) {
        if (dst) std::memcpy(dst, src, bytes);
    };
    const size_t n2 = c->n2, km = c->km;
    cp(o->kmt, c->h_kmt.data(), n2 * 4);
    cp(o->kmu, c->h_kmu.data(), n2 * 4);
    cp(o->tarea_r, c->h_tarea_r.data(), n2 * 8);
    cp(o->ohbt, c->h_ohbt.data(), n2 * 8);
    cp(o->ohbu, c->h_ohbu.data(), n2 * 8);
    cp(o->dtn, c->h_dtn.data(), n2 * 8);
    cp(o->dts, c->h_dts.data(), n2 * 8);
    cp(o->dte, c->h_dte.data(), n2 * 8);
    cp(o->dtw, c->h_dtw.data(), n2 * 8);
    cp(o->ahf, c->h_ahf.data(), n2 * 8);
    cp(o->dzp, c->h_dzp.data(), km * 8);
    cp(o->odzp, c->h_odzp.data(), km * 8);
    cp(o->odzt, c->h_odzt.data(), km * 8);
    o->area_t = c->area_t;
    return LICOM_OK;
}

int licom_set_profiling(licom_ctx *c, int enable) {
    if (!c) return LICOM_EINVAL;
    c->profiling = enable != 0;
    return LICOM_OK;
}

int licom_kernel_times(licom_ctx *c, int max, const char **names, double *ms, int *counts,
                       int *n_out) {
    if (!c) return LICOM_EINVAL;
    if (int rc = hip_status(hipStreamSynchronize(c->stream), "sync")) return rc;
    for (auto &e : c->pending) {
        float t = 0.f;
        (void)hipEventElapsedTime(&t, e.start, e.stop);
        (void)hipEventDestroy(e.start);
        (void)hipEventDestroy(e.stop);
        size_t k = 0;
        while (k < c->prof_names.size() && c->prof_names[k] != e.name) ++k;
        if (k == c->prof_names.size()) {
            c->prof_names.push_back(e.name);
            c->prof_ms.push_back(0.0);
            c->prof_counts.push_back(0);
        }
        c->prof_ms[k] += t;
        c->prof_counts[k] += 1;
    }
    c->pending.clear();
    int n = (int)c->prof_names.size();
    if (n_out) *n_out = n;
    for (int k = 0; k < n && k < max; ++k) {
        if (names) names[k] = c->prof_names[k].c_str();
        if (ms) ms[k] = c->prof_ms[k];
        if (counts) counts[k] = c->prof_counts[k];
    }
    return LICOM_OK;
}

size_t licom_scratch_bytes(const licom_ctx *c) { return c ? c->scratch_bytes : 0; }

int licom_last_mix_count(licom_ctx *c, int *count) {
    if (!c || !count) return LICOM_EINVAL;
    if (int rc = hip_status(hipStreamSynchronize(c->stream), "sync")) return rc;
    return hip_status(hipMemcpy(count, c->s.mix_count, sizeof(int), hipMemcpyDeviceToHost),
                      "mix count copy");
}

int licom_tracer_set_flag_output(licom_ctx *c, void *flags) {
    if (!c) return LICOM_EINVAL;
    c->flags_out = static_cast<uint8_t *>(flags);
    return LICOM_OK;
}

int licom_last_nonfinite(licom_ctx *c, int *count) {
    if (!c || !count) return LICOM_EINVAL;
    if (int rc = hip_status(hipStreamSynchronize(c->stream), "sync")) return rc;
    return hip_status(hipMemcpy(count, c->s.mix_count + 1, sizeof(int), hipMemcpyDeviceToHost),
                      "non-finite count copy");
}

int licom_stream_copy(void *dst, const void *src, size_t bytes, void *stream) {
    if (!dst || !src || (bytes % 16) != 0) {
        set_error("licom_stream_copy: NULL pointer or size not a multiple of 16");
        return LICOM_EINVAL;
    }
    return launch_stream_copy(dst, src, bytes, static_cast<hipStream_t>(stream));
}

void licom_tracer_destroy(licom_ctx *c) {
    if (!c) return;
    if (!c->pending.empty()) (void)hipStreamSynchronize(c->stream);
    for (auto &e : c->pending) {
        (void)hipEventDestroy(e.start);
        (void)hipEventDestroy(e.stop);
    }
    for (void *p : c->allocs) (void)hipFree(p);
    delete c;
}

const char *licom_strerror(int s) {
    switch (s) {
        case LICOM_OK: return "ok";
        case LICOM_EINVAL: return "invalid argument";
        case LICOM_ENOMEM: return "out of device memory";
        case LICOM_EHIP: return "HIP runtime error";
        case LICOM_ENODEV: return "no usable device";
        case LICOM_ENOTSUP: return "not supported";
        default: return "unknown status";
    }
}

const char *licom_last_error(void) { return g_err; }

}  // extern "C"
