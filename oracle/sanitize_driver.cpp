// sanitize_driver.cpp -- host AddressSanitizer / UndefinedBehaviorSanitizer
// run of the oracle and of the input generator (test infrastructure only:
// tests/test_sanitize_host.py builds it with `make -C oracle sanitize` and
// runs it).  It drives every oracle entry point the tests use over small
// seeded cases, with land, del2 and del4, forcing, restoring, icesnow and
// convadj, and a ragged grid whose extents are not multiples of any tile,
// so that an out-of-bounds index or an undefined operation in either
// component stops the run with a report.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "licom_synth.h"
#include "oracle.h"

namespace {

struct Case {
    int nx, ny, km, imt, jmt;
    size_t n2, n3;
    std::vector<int32_t> kmt, kmu;
    std::vector<double> zkp, pen, htw, hts, hun, hue, tarea, dxu, dyu, au0, aus, auw, ausw, to, so,
        eosc, lat;
    std::vector<double> vit, viv, tarea_r, ohbt, ohbu, dtn, dts, dte, dtw, ahf, dzp, odzp, odzt;
    std::vector<double> at, atb, akt, utf, vtf, utl, vtl, ws, h0f, h0l;
    std::vector<double> swv, nswv, sst, seaice, sss, fresh, restore;
    std::vector<double> ax, ay, az, dz, tend, dt_diff, dt_conv, penetrate, net, tf, qice;
    int imt_global = 0;
};

void make(Case &c, int nx, int ny, int km, uint64_t seed, double depth_min) {
    licom_synth_cfg cfg;
    licom_synth_defaults(&cfg);
    cfg.nx = nx;
    cfg.ny = ny;
    cfg.km = km;
    cfg.lat_mode = 0;
    cfg.dlon = 0.5;
    cfg.dlat = 0.5;
    cfg.lat_north = 54.0;
    cfg.seed = seed;
    cfg.depth_min = depth_min;
    cfg.sigma = 1.5;
    c.nx = nx;
    c.ny = ny;
    c.km = km;
    c.imt = nx + 4;
    c.jmt = ny + 4;
    c.n2 = (size_t)c.imt * c.jmt;
    c.n3 = c.n2 * km;
    const size_t n2 = c.n2, n3 = c.n3;
    c.kmt.assign(n2, 0);
    c.kmu.assign(n2, 0);
    for (auto *v : {&c.htw, &c.hts, &c.hun, &c.hue, &c.tarea, &c.dxu, &c.dyu, &c.au0, &c.aus, &c.auw,
                    &c.ausw, &c.tarea_r, &c.ohbt, &c.ohbu, &c.dtn, &c.dts, &c.dte, &c.dtw, &c.ahf,
                    &c.h0f, &c.h0l, &c.swv, &c.nswv, &c.sst, &c.seaice, &c.sss, &c.fresh, &c.qice})
        v->assign(n2, 0.0);
    c.zkp.assign(km + 1, 0.0);
    c.pen.assign(km - 1, 0.0);
    c.to.assign(km, 0.0);
    c.so.assign(km, 0.0);
    c.eosc.assign((size_t)9 * km, 0.0);
    c.lat.assign(c.jmt, 0.0);
    c.dzp.assign(km, 0.0);
    c.odzp.assign(km, 0.0);
    c.odzt.assign(km, 0.0);
    c.vit.assign(n3, 0.0);
    c.viv.assign(n3, 0.0);
    licom_synth_grid g{c.kmt.data(), c.zkp.data(), c.pen.data(), c.htw.data(), c.hts.data(),
                       c.hun.data(), c.hue.data(), c.tarea.data(), c.dxu.data(), c.dyu.data(),
                       c.au0.data(), c.aus.data(), c.auw.data(), c.ausw.data(), c.to.data(),
                       c.so.data(), c.eosc.data(), c.lat.data(), 0};
    if (licom_synth_make_grid(&cfg, &g) != 0) std::exit(3);
    c.imt_global = g.imt_global;
    c.at.assign(2 * n3, 0.0);
    c.atb.assign(2 * (n3 + n2), 0.0);
    c.akt.assign(2 * n3, 0.0);
    for (auto *v : {&c.utf, &c.vtf, &c.utl, &c.vtl}) v->assign(n3, 0.0);
    c.ws.assign(n3 + n2, 0.0);
    licom_synth_state s{c.at.data(), c.atb.data(), c.akt.data(), c.utf.data(), c.vtf.data(),
                        c.utl.data(), c.vtl.data(), c.ws.data(), c.h0f.data(), c.h0l.data()};
    if (licom_synth_make_state(&cfg, c.kmt.data(), c.zkp.data(), &s) != 0) std::exit(4);
    for (size_t q = 0; q < n2; ++q) {
        const double x = (double)(q % 97) / 97.0;
        c.swv[q] = 150.0 * x;
        c.nswv[q] = -60.0 + 40.0 * x;
        c.sst[q] = 5.0 + 20.0 * x;
        c.seaice[q] = (q % 5 == 0) ? 1.0 : 0.2 * x;
        c.sss[q] = -1e-3 + 2e-3 * x;
        c.fresh[q] = 1e-5 * (x - 0.5);
    }
    c.restore.assign(2 * n3, 0.0);
    for (size_t q = 0; q < 2 * n3; ++q) c.restore[q] = c.at[q] + 0.25;
    for (auto *v : {&c.ax, &c.ay, &c.az, &c.dz, &c.tend, &c.dt_diff, &c.dt_conv, &c.tf})
        v->assign(2 * n3, 0.0);
    c.penetrate.assign(n3, 0.0);
    c.net.assign(2 * n2, 0.0);
}

orc_grid grid_of(Case &c, int hmix) {
    orc_grid g{};
    g.imt = c.imt;
    g.jmt = c.jmt;
    g.km = c.km;
    g.imt_global = c.imt_global;
    g.hmix = hmix;
    g.kmt = c.kmt.data();
    g.zkp = c.zkp.data();
    g.pen = c.pen.data();
    g.htw = c.htw.data();
    g.hts = c.hts.data();
    g.hun = c.hun.data();
    g.hue = c.hue.data();
    g.tarea = c.tarea.data();
    g.au0 = c.au0.data();
    g.aus = c.aus.data();
    g.auw = c.auw.data();
    g.ausw = c.ausw.data();
    g.to = c.to.data();
    g.so = c.so.data();
    g.eosc = c.eosc.data();
    g.kmu = c.kmu.data();
    g.vit = c.vit.data();
    g.viv = c.viv.data();
    g.tarea_r = c.tarea_r.data();
    g.ohbt = c.ohbt.data();
    g.ohbu = c.ohbu.data();
    g.dtn = c.dtn.data();
    g.dts = c.dts.data();
    g.dte = c.dte.data();
    g.dtw = c.dtw.data();
    g.ahf = c.ahf.data();
    g.dzp = c.dzp.data();
    g.odzp = c.odzp.data();
    g.odzt = c.odzt.data();
    return g;
}

int run(int nx, int ny, int km, uint64_t seed, int hmix, double ah, int steps) {
    Case c;
    make(c, nx, ny, km, seed, -2000.0);
    orc_grid g = grid_of(c, hmix);
    orc_derive(&g);
    orc_params p{3600.0, 0.25, 0.5, 1e-3, 1.0 / (30.0 * 86400.0), 1.0 / 1026.0,
                 1.0 / (1026.0 * 3996.0), ah, 1, 1, 1};
    orc_fields f{c.at.data(), c.atb.data(), c.akt.data(), c.utf.data(), c.vtf.data(), c.utl.data(),
                 c.vtl.data(), c.ws.data(), c.h0f.data(), c.h0l.data(), c.swv.data(), c.nswv.data(),
                 c.sst.data(), c.seaice.data(), c.sss.data(), c.fresh.data(), c.restore.data()};
    orc_diag d{c.ax.data(), c.ay.data(), c.az.data(), c.dz.data(), c.tend.data(), c.dt_diff.data(),
               c.dt_conv.data(), c.penetrate.data(), c.net.data(), 0.0, c.tf.data()};
    for (int s = 0; s < steps; ++s) {
        orc_tracer_step(&g, &p, &f, &d);
        orc_icesnow(&g, c.at.data(), c.qice.data(), 1);
        orc_convadj(&g, &p, &f, &d);
    }
    std::vector<double> adv(c.n3, 0.0);
    orc_advect(&g, p.dts, c.at.data(), c.utf.data(), c.vtf.data(), c.ws.data(), adv.data());
    std::vector<double> out(c.n2, 0.0);
    orc_del2_level(&g, c.at.data(), km - 1, 1, 1.0, out.data());
    std::vector<double> col(km, 1.0);
    orc_invtri_column(km, c.akt.data(), c.odzt.data(), c.odzp.data(), p.dts, 1e-4, 1.0, col.data());
    double sum = 0.0;
    for (size_t q = 0; q < c.n3; ++q) sum += c.at[q] * c.vit[q];
    if (!std::isfinite(sum)) {
        std::fprintf(stderr, "non-finite state after %d steps (%dx%dx%d)\n", steps, nx, ny, km);
        return 1;
    }
    std::printf("ok %dx%dx%d hmix=%d steps=%d content=%.6e\n", nx, ny, km, hmix, steps, sum);
    return 0;
}

}  // namespace

int main() {
    int rc = 0;
    rc |= run(20, 13, 9, 7, 2, 1.0e3, 2);
    rc |= run(17, 11, 12, 8, 4, -1.0e10, 2);
    rc |= run(33, 6, 5, 9, 2, 1.0e3, 1);
    return rc;
}
