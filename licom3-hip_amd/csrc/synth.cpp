// synth.cpp -- deterministic synthetic grid and state for the tracer step.
//
// Host-only.  Stands in for the reference's input files: the vertical grid
// (grid.F90:1096-1106 consumes ZKP), the horizontal metrics
// (grid.F90:780-990), the topography (grid.F90:1115-1240 -> boundary.F90),
// the EOS table (constant_mod.F90:331-364, file dncoef.h1, not shipped) and
// the initial state (inirun.F90; salinity in model units (S-35)*1e-3,
// inirun.F90:540).  The shortwave profile follows grids.F90:51-60.
#include "licom_synth.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kRadius = 6371000.0;   // constant_mod.F90:31
constexpr double kD0 = 1026.0;          // constant_mod.F90:165
constexpr double kCp = 3996.0;          // constant_mod.F90:164
constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ULL;

inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
// n-th output of a SplitMix64 generator whose state starts at `key`.
inline uint64_t sm64(uint64_t key, uint64_t n) { return mix64(key + (n + 1) * kGolden); }
inline double uni(uint64_t key, uint64_t n) { return (double)(sm64(key, n) >> 11) * 0x1.0p-53; }
inline double gauss(uint64_t key, uint64_t idx) {
    double u1 = 1.0 - uni(key, 2 * idx);
    double u2 = uni(key, 2 * idx + 1);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
}
inline uint64_t stream(uint64_t seed, uint64_t field) { return seed ^ (field << 32); }

template <class T>
void halo_fill(T *p, int imt, int jmt, int nz) {
    for (int z = 0; z < nz; ++z) {
        T *a = p + (size_t)z * imt * jmt;
        for (int j = 2; j < jmt - 2; ++j) {
            T *r = a + (size_t)j * imt;
            r[0] = r[imt - 4];
            r[1] = r[imt - 3];
            r[imt - 2] = r[2];
            r[imt - 1] = r[3];
        }
        for (int i = 0; i < imt; ++i) {
            a[i] = T(0);
            a[imt + i] = T(0);
            a[(size_t)(jmt - 2) * imt + i] = T(0);
            a[(size_t)(jmt - 1) * imt + i] = T(0);
        }
    }
}

double merc(double lat_deg) { return std::log(std::tan(M_PI / 4.0 + lat_deg * M_PI / 360.0)); }
double imerc(double y) { return (2.0 * std::atan(std::exp(y)) - M_PI / 2.0) * 180.0 / M_PI; }

// Latitude (degrees) of the northern face of interior row r (r may be
// negative or >= ny for ghost rows).  Rows run north to south.
double face_lat(const licom_synth_cfg *c, double r) {
    if (c->lat_mode == 1) {
        double d = c->dlon * M_PI / 180.0;
        return imerc(merc(c->lat_c) + (0.5 * c->ny - r) * d);
    }
    return c->lat_north - r * c->dlat;
}

// tanh-stretched thicknesses from dz_top to dz_bot summing to depth_total.
std::vector<double> layer_dz(const licom_synth_cfg *c) {
    const int km = c->km;
    const double s = 4.0;
    auto prof = [&](double x0, std::vector<double> &dz) {
        double lo = std::tanh(-s * x0), hi = std::tanh(s * (1.0 - x0));
        double sum = 0.0;
        for (int k = 0; k < km; ++k) {
            double x = km > 1 ? (double)k / (km - 1) : 0.0;
            dz[k] = c->dz_top + (c->dz_bot - c->dz_top) * (std::tanh(s * (x - x0)) - lo) / (hi - lo);
            sum += dz[k];
        }
        return sum;
    };
    std::vector<double> dz(km);
    double a = -2.0, b = 3.0;   // total thickness decreases as x0 grows
    for (int it = 0; it < 200; ++it) {
        double m = 0.5 * (a + b);
        if (prof(m, dz) > c->depth_total) a = m; else b = m;
    }
    prof(0.5 * (a + b), dz);
    return dz;
}

int kmt_of_depth(const double *zkp, int km, double depth, double land) {
    if (depth < land) return 0;
    int n = 0;
    for (int k = 0; k < km; ++k)
        if (-zkp[k + 1] <= depth) n = k + 1;
    return n;
}

}  // namespace

extern "C" void licom_synth_defaults(licom_synth_cfg *c) {
    std::memset(c, 0, sizeof(*c));
    c->nx = 128; c->ny = 96; c->km = 30;
    c->lat_mode = 0; c->dlon = 0.5; c->dlat = 0.5; c->lat_north = 54.0; c->lat_c = 30.0;
    c->seed = 101;
    c->depth_min = 500.0; c->depth_max = 5000.0; c->sigma = 4.0; c->land_depth = 200.0;
    c->dz_top = 10.0; c->dz_bot = 300.0; c->depth_total = 5000.0;
    c->umax = 0.5; c->div_frac = 0.2;
    c->t_noise = 0.05; c->s_noise = 0.02; c->t_pert = 0.0;
    c->akt_lo = 1e-5; c->akt_hi = 1e-2;
}

extern "C" int licom_synth_make_grid(const licom_synth_cfg *c, licom_synth_grid *g) {
    if (!c || !g || c->nx < 4 || c->ny < 2 || c->km < 2) return -1;
    const int nx = c->nx, ny = c->ny, km = c->km, imt = nx + 4, jmt = ny + 4;
    const double rad = M_PI / 180.0;
    g->imt_global = (int)std::lround(360.0 / c->dlon);

    // vertical grid
    std::vector<double> dz = layer_dz(c);
    g->zkp[0] = 0.0;
    for (int k = 0; k < km; ++k) g->zkp[k + 1] = g->zkp[k] - dz[k];
    // grids.F90:51-60 (SOLAR two-band penetration), already times OD0CP
    {
        const double rpart = 0.58, rscl1 = 1.0 / 0.35, rscl2 = 1.0 / 23.0;
        const double od0cp = 1.0 / (kD0 * kCp);
        for (int k = 0; k < km - 1; ++k) {
            double a1 = std::max(g->zkp[k + 1] * rscl1, -70.0);
            double a2 = std::max(g->zkp[k + 1] * rscl2, -70.0);
            g->pen[k] = (rpart * std::exp(a1) + (1.0 - rpart) * std::exp(a2)) * od0cp;
        }
    }

    // horizontal metrics of a lat-lon B grid: T(i,j) centred, U(i,j) at the
    // south-west corner of T(i,j), j increasing southward.
    std::vector<double> latn(jmt + 1), latc(jmt + 1);
    for (int j = 0; j <= jmt; ++j) {
        latn[j] = face_lat(c, j - 2);
        latc[j] = 0.5 * (face_lat(c, j - 2) + face_lat(c, j - 1));
    }
    const double dl = c->dlon * rad;
    for (int j = 0; j < jmt; ++j) {
        const double ls = latn[j + 1];
        const double dyt = kRadius * (latn[j] - latn[j + 1]) * rad;
        const double dxt = kRadius * std::cos(latc[j] * rad) * dl;
        const double dyu = kRadius * (latc[j] - latc[j + 1]) * rad;
        const double dxu = kRadius * std::cos(ls * rad) * dl;
        g->lat_t[j] = latc[j];
        for (int i = 0; i < imt; ++i) {
            size_t q = (size_t)j * imt + i;
            g->tarea[q] = dxt * dyt;                   // grid.F90:465
            g->hts[q] = dxu;                           // south face of T
            g->htw[q] = dyt;                           // west face of T
            g->hun[q] = dxt;                           // lat_lon: hun = dxt (grid.F90:944)
            g->hue[q] = dyu;                           // lat_lon: hue = dyu
            g->dxu[q] = dxu;
            g->dyu[q] = dyu;
            g->au0[q] = g->aus[q] = g->auw[q] = g->ausw[q] = 0.25;  // grid.F90:1654-1657
        }
    }

    // bathymetry: U[depth_min, depth_max], Gaussian smoothed (periodic in i,
    // clamped in j), mapped to level counts; shallow columns are land.
    std::vector<double> d0((size_t)nx * ny), d1((size_t)nx * ny);
    const uint64_t kb = stream(c->seed, 1);
    for (size_t q = 0; q < (size_t)nx * ny; ++q)
        d0[q] = c->depth_min + uni(kb, q) * (c->depth_max - c->depth_min);
    if (c->sigma > 0.0) {
        int R = (int)std::ceil(4.0 * c->sigma);
        std::vector<double> w(2 * R + 1);
        double ws = 0.0;
        for (int d = -R; d <= R; ++d) ws += (w[d + R] = std::exp(-0.5 * d * d / (c->sigma * c->sigma)));
        for (auto &x : w) x /= ws;
#pragma omp parallel for schedule(static)
        for (int jj = 0; jj < ny; ++jj)
            for (int ii = 0; ii < nx; ++ii) {
                double s = 0.0;
                for (int d = -R; d <= R; ++d) s += w[d + R] * d0[(size_t)jj * nx + (((ii + d) % nx) + nx) % nx];
                d1[(size_t)jj * nx + ii] = s;
            }
#pragma omp parallel for schedule(static)
        for (int jj = 0; jj < ny; ++jj)
            for (int ii = 0; ii < nx; ++ii) {
                double s = 0.0;
                for (int d = -R; d <= R; ++d) {
                    int j2 = std::min(std::max(jj + d, 0), ny - 1);
                    s += w[d + R] * d1[(size_t)j2 * nx + ii];
                }
                d0[(size_t)jj * nx + ii] = s;
            }
    }
    std::memset(g->kmt, 0, sizeof(int32_t) * imt * jmt);
    for (int jj = 0; jj < ny; ++jj)
        for (int ii = 0; ii < nx; ++ii)
            g->kmt[(size_t)(jj + 2) * imt + ii + 2] =
                kmt_of_depth(g->zkp, km, d0[(size_t)jj * nx + ii], c->land_depth);
    halo_fill(g->kmt, imt, jmt, 1);

    // Synthetic 9-term EOS table in the dncoef.h1 layout C(km,9), already in
    // kg/m^3 (constant_mod.F90:364 multiplies by 1000).  Terms:
    // C1 t, C2 s, C3 t^2, C4 s t, C5 s^2, C6 t^3, C7 s^2 t, C8 s t^2, C9 s^3
    // (dens.F90:12-14), s in model units (psu-35)/1000.
    for (int k = 0; k < km; ++k) {
        double z = -0.5 * (g->zkp[k] + g->zkp[k + 1]);
        g->to[k] = 2.0 + 25.0 * std::exp(-z / 700.0) * std::cos(30.0 * rad);
        g->so[k] = (34.7 + 0.6 * std::exp(-z / 500.0) - 35.0) * 1.0e-3;
        double alpha = 1.0e-4 + 1.5e-4 * std::exp(-z / 1500.0);
        double cf[9] = {-kD0 * alpha, kD0 * 7.6e-4 * 1000.0, -kD0 * 5.0e-6, -kD0 * 1.0e-3,
                        0.0, kD0 * 2.0e-8, 0.0, 0.0, 0.0};
        for (int m = 0; m < 9; ++m) g->eosc[(size_t)m * km + k] = cf[m];
    }
    return 0;
}

extern "C" int licom_synth_make_state(const licom_synth_cfg *c, const int32_t *kmt,
                                      const double *zkp, licom_synth_state *s) {
    if (!c || !kmt || !zkp || !s) return -1;
    const int nx = c->nx, ny = c->ny, km = c->km, imt = nx + 4, jmt = ny + 4;
    const size_t n2 = (size_t)imt * jmt, n3 = n2 * km;
    const double rad = M_PI / 180.0;

    std::vector<double> zc(km), dzp(km);
    for (int k = 0; k < km; ++k) {
        zc[k] = -0.5 * (zkp[k] + zkp[k + 1]);
        dzp[k] = zkp[k] - zkp[k + 1];
    }
    std::vector<double> latc(jmt);
    for (int j = 0; j < jmt; ++j) latc[j] = 0.5 * (face_lat(c, j - 2) + face_lat(c, j - 1));

    // tracers (masked below the bottom and on land)
    const uint64_t kt = stream(c->seed, 2), ks = stream(c->seed, 3), kp = stream(c->seed, 6);
    const uint64_t ka = stream(c->seed, 5);
    const double lo = std::log(c->akt_lo), hi = std::log(c->akt_hi);
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < km; ++k)
        for (int jj = 0; jj < ny; ++jj) {
            const int j = jj + 2;
            const double clat = std::cos(latc[j] * rad);
            for (int ii = 0; ii < nx; ++ii) {
                const int i = ii + 2;
                const size_t q = (size_t)j * imt + i, c3 = (size_t)k * n2 + q;
                const uint64_t idx = ((uint64_t)k * ny + jj) * nx + ii;
                const bool wet = k < kmt[q];
                double T = 0.0, S = 0.0;
                if (wet) {
                    T = 2.0 + 25.0 * std::exp(-zc[k] / 700.0) * clat + c->t_noise * gauss(kt, idx);
                    if (c->t_pert > 0.0) T += c->t_pert * gauss(kp, idx);
                    double Sp = 34.7 + 0.6 * std::exp(-zc[k] / 500.0) + c->s_noise * gauss(ks, idx);
                    S = (Sp - 35.0) * 1.0e-3;
                }
                s->at[c3] = T;
                s->at[n3 + c3] = S;
                for (int n = 0; n < 2; ++n) {
                    const uint64_t ia = (((uint64_t)n * km + k) * ny + jj) * nx + ii;
                    double a = 0.0;
                    if (wet && k <= km - 2) a = std::exp(lo + uni(ka, ia) * (hi - lo));
                    s->akt[(size_t)n * n3 + c3] = a;
                }
            }
        }
    halo_fill(s->at, imt, jmt, 2 * km);
    halo_fill(s->akt, imt, jmt, 2 * km);
    for (int n = 0; n < 2; ++n) {
        double *b = s->atb + (size_t)n * (n3 + n2);
        std::memset(b, 0, n2 * sizeof(double));
        std::memcpy(b + n2, s->at + (size_t)n * n3, n3 * sizeof(double));
    }

    // velocity: U points (SW corners).  KMU as boundary.F90:29-41.
    std::vector<int32_t> kmu(n2, 0);
    for (int j = 0; j < jmt - 1; ++j)
        for (int i = 1; i < imt; ++i) {
            int a = kmt[(size_t)j * imt + i - 1], b = kmt[(size_t)j * imt + i];
            int e = kmt[(size_t)(j + 1) * imt + i - 1], f = kmt[(size_t)(j + 1) * imt + i];
            kmu[(size_t)j * imt + i] = std::min(std::min(a, b), std::min(e, f));
        }
    halo_fill(kmu.data(), imt, jmt, 1);
    // smooth random flow: 6 streamfunction modes + 2 potential modes
    const uint64_t kf = stream(c->seed, 4);
    struct Mode { double a, p, q, th, et; bool pot; } md[8];
    for (int m = 0; m < 8; ++m) {
        md[m].a = 0.5 + 0.5 * uni(kf, 8 * m);
        md[m].p = 1.0 + std::floor(4.0 * uni(kf, 8 * m + 1));
        md[m].q = 1.0 + std::floor(4.0 * uni(kf, 8 * m + 2));
        md[m].th = 2.0 * M_PI * uni(kf, 8 * m + 3);
        md[m].et = 2.0 * M_PI * uni(kf, 8 * m + 4);
        md[m].pot = m >= 6;
    }
    std::vector<double> u2(n2, 0.0), v2(n2, 0.0);
    double umax = 0.0;
#pragma omp parallel for reduction(max : umax) schedule(static)
    for (int jj = 0; jj < ny; ++jj)
        for (int ii = 0; ii < nx; ++ii) {
            const double X = (double)ii / nx, Y = (double)(jj + 1) / ny;
            double u = 0.0, v = 0.0;
            for (int m = 0; m < 8; ++m) {
                const double ax = 2.0 * M_PI * md[m].p * X + md[m].th, ay = M_PI * md[m].q * Y + md[m].et;
                const double fx = -md[m].a * 2.0 * M_PI * md[m].p * std::sin(ax) * std::sin(ay);
                const double fy = md[m].a * M_PI * md[m].q * std::cos(ax) * std::cos(ay);
                if (md[m].pot) { u += c->div_frac * fx; v += c->div_frac * fy; }
                else { u += -fy; v += fx; }
            }
            const size_t q = (size_t)(jj + 2) * imt + ii + 2;
            u2[q] = u; v2[q] = v;
            umax = std::max(umax, std::sqrt(u * u + v * v));
        }
    const double scale = umax > 0.0 ? c->umax / umax : 0.0;
    const double oncc = 0.5;   // utf = u / oncc with NCC = 1
    std::vector<double> e(km);
    for (int k = 0; k < km; ++k) e[k] = std::exp(-zc[k] / 1000.0);
#pragma omp parallel for schedule(static)
    for (int jj = 0; jj < ny; ++jj) {
        std::vector<double> gk(km);
        for (int ii = 0; ii < nx; ++ii) {
            const size_t q = (size_t)(jj + 2) * imt + ii + 2;
            const int m = kmu[q];
            // zero depth-integrated transport in every U column
            double num = 0.0, den = 0.0;
            for (int k = 0; k < m; ++k) { num += dzp[k] * e[k]; den += dzp[k]; }
            const double mean = den > 0.0 ? num / den : 0.0;
            for (int k = 0; k < km; ++k) gk[k] = k < m ? e[k] - mean : 0.0;
            for (int k = 0; k < km; ++k) {
                const size_t c3 = (size_t)k * n2 + q;
                const double u = u2[q] * scale * gk[k], v = v2[q] * scale * gk[k];
                s->utl[c3] = u;
                s->vtl[c3] = v;
                s->utf[c3] = u / oncc;
                s->vtf[c3] = v / oncc;
            }
        }
    }
    halo_fill(s->utl, imt, jmt, km);
    halo_fill(s->vtl, imt, jmt, km);
    halo_fill(s->utf, imt, jmt, km);
    halo_fill(s->vtf, imt, jmt, km);
    std::memset(s->ws, 0, (n3 + n2) * sizeof(double));
    std::memset(s->h0f, 0, n2 * sizeof(double));
    std::memset(s->h0l, 0, n2 * sizeof(double));
    return 0;
}
