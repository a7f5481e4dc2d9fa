"""Pins of the oracle's readyt_ (oracle/readyt.c) against what the
mathematics and published values fix, so that a dropped term, a wrong level
index or a sign slip in the restatement fails here and not only in a GPU
comparison with the same restatement:

* thermal_cu (readyt.cpp:258-297): McDougall (1987)'s check values, beta =
  0.72088e-3 /psu and alpha/beta = 0.34763 at S = 40 psu, theta = 10 C,
  p = 4000 dbar; zero on land;
* potential density (readyt_cu_2): the 9-term polynomial of each level
  evaluated monomial by monomial with the level's own coefficients;
* hydrostatic integrals (readyt_cu_2_2): for a column of uniform density pp
  is gg times the depth of the T point and ppa is psa plus gg times the
  depth of the level's top interface;
* ppb / ppc: exact linear interpolation to the top interface on a linear
  profile;
* rict / ric (dens_cu): the buoyancy jump between two levels with both
  densities referred to the lower level's coefficients (a level-k
  evaluation differs); ric equals rict where atb is horizontally uniform;
* ricdttms / ricdt (readyt_cu_3): the identity ricdttms = g odzt alpha
  (dT + ricdt (dT + eps)) on wet cells;
* pressure gradient and its integrals (readyt_cu_3, _3_2): on pp linear in
  i and j the gradient is the slope times 1/dxu, 1/dyu; pxb equals the
  stencil of the column integrals of pp (integral of the gradient = gradient
  of the integral);
* dlu / dlv planes 0, 1 (readyt_cu_3_2, _4): for a uniform anomaly weight the
  column mean is gg and the first moment gg times the mean T-point depth;
* readyt_cu_5: T->U average of a constant, surface pressure gradient of a
  linear psa;
* the time-level shift, the clears (akt, rit, ax/ay/az only at ist = 0) and
  that every output cell is written.
"""
import numpy as np
import pytest

import licom_synth as ls
import orc

G = 9.806
OD0 = 1.0 / 1026.0


def base(nx=24, ny=12, km=8, seed=3, **kw):
    c = ls.make_case("C1", nx=nx, ny=ny, km=km, seed=seed, depth_min=-2000.0, sigma=1.5, **kw)
    g = ls.readyt_grid(c)
    f = ls.readyt_fields(c, g)
    return c, g, f


def run(c, g, f, ist=1):
    return orc.readyt(c.imt, c.jmt, c.km, g, f, ist=ist)


@pytest.fixture(scope="module")
def case():
    c, g, f = base()
    return c, g, f, run(c, g, f)


def test_thermal_mcdougall_check_values():
    # S = 40 psu -> ss = (40 - 35) / 1000; p = 4000 dbar -> ppa = -p * 1e4 * od0
    a, b = orc.thermal(10.0, 0.005, -4000.0 * 1.0e4 * OD0, 1.0, OD0)
    assert abs(b - 0.72088e-3) < 0.5e-8
    assert abs(a / b - 0.34763) < 0.5e-5
    assert orc.thermal(10.0, 0.005, -4.0e7 * OD0, 0.0, OD0) == (0.0, 0.0)
    # the surface value at 10 C, 35 psu: alpha ~ 1.67e-4 /C (sea-water tables)
    a0, b0 = orc.thermal(10.0, 0.0, 0.0, 1.0, OD0)
    assert 1.6e-4 < a0 < 1.75e-4 and 7.4e-4 < b0 < 7.7e-4


def test_potential_density_monomials(case):
    c, g, f, o = case
    C = g["eosc"]
    km = c.km
    vit = f["vit"]
    T = f["at"][0] - g["to"][:, None, None]
    S = f["at"][1] - g["so"][:, None, None]
    cc = [C[m][:, None, None] for m in range(9)]
    mono = (cc[0] * T + cc[1] * S + cc[2] * T ** 2 + cc[3] * S * T + cc[4] * S ** 2 +
            cc[5] * T ** 3 + cc[6] * S ** 2 * T + cc[7] * S * T ** 2 + cc[8] * S ** 3)
    ref = np.where(vit > 0, 1000.0 + g["po"][:, None, None] + mono, 0.0)
    assert np.max(np.abs(o["pdensity"] - ref)) <= 1e-12 * 1025.0
    # the anomaly weight: -od0 g (rho - po - 1000) vit
    gg = -g["od0"] * G * (ref - g["po"][:, None, None] - 1000.0) * vit
    assert np.max(np.abs(o["gg"] - gg)) <= 1e-12 * np.max(np.abs(gg))
    del km


def uniform_density_case():
    """Columns of uniform potential density: a linear EOS (C(k,1) only) with
    T = to + 1 and S = so at every level."""
    c, g, f = base(seed=5)
    km = c.km
    g["eosc"][:] = 0.0
    g["eosc"][0, :] = -0.2
    g["po"][:] = 3.0
    T = np.broadcast_to((g["to"] + 1.0)[:, None, None], f["at"][0].shape)
    S = np.broadcast_to(g["so"][:, None, None], f["at"][1].shape)
    f["at"][0] = T * f["vit"]
    f["at"][1] = S * f["vit"]
    del km
    return c, g, f


def test_hydrostatic_integrals_closed_form():
    c, g, f = uniform_density_case()
    o = run(c, g, f)
    kmt = c.grid["kmt"]
    rho = 1000.0 + 3.0 - 0.2
    gg0 = -g["od0"] * G * rho
    zkp = c.grid["zkp"]
    for (j, i) in [(5, 7), (9, 3), (2, 20)]:
        n = int(kmt[j, i])
        assert n >= 2
        for k in range(n):
            assert o["pp"][k, j, i] == pytest.approx(gg0 * (-g["zkt"][k]), rel=1e-13)
            assert o["ppa"][k, j, i] == pytest.approx(f["psa"][j, i] + gg0 * (-zkp[k]), rel=1e-12,
                                                      abs=1e-9)
        assert np.all(o["pp"][n:, j, i] == 0.0) and np.all(o["ppa"][n:, j, i] == 0.0)


def test_interface_interpolation_linear_profile():
    c, g, f = base(seed=6)
    a, b = 12.0, 3.0e-3
    T = a + b * g["zkt"]
    f["at"][0] = np.broadcast_to(T[:, None, None], f["at"][0].shape) * f["vit"]
    o = run(c, g, f)
    zkp = c.grid["zkp"]
    kmt = c.grid["kmt"]
    j, i = 6, 9
    n = int(kmt[j, i])
    assert o["ppb"][0, j, i] == T[0]
    for k in range(1, n):
        assert o["ppb"][k, j, i] == pytest.approx(a + b * zkp[k], rel=1e-13)


def test_rict_uses_the_lower_levels_coefficients():
    c, g, f = base(seed=7)
    km = c.km
    rng = np.random.default_rng(1)
    g["eosc"][:] = rng.normal(0.0, 1.0, g["eosc"].shape) * np.array(
        [0.2, 0.8, 5e-3, 3e-3, 1e-3, 2e-5, 1e-4, 1e-5, 1e-5])[:, None]
    o = run(c, g, f)
    C = g["eosc"]

    def rho(k, t, s):
        cc = C[:, k]
        return (cc[0] * t + cc[1] * s + cc[2] * t * t + cc[3] * s * t + cc[4] * s * s +
                cc[5] * t ** 3 + cc[6] * s * s * t + cc[7] * s * t * t + cc[8] * s ** 3)

    atb = f["atb"]
    j, i = 6, 9
    for k in range(km - 1):
        tu, su = atb[0, k + 1, j, i] - g["to"][k + 1], atb[1, k + 1, j, i] - g["so"][k + 1]
        tl, sl = atb[0, k + 2, j, i] - g["to"][k + 1], atb[1, k + 2, j, i] - g["so"][k + 1]
        want = f["vit"][k + 1, j, i] * g["od0"] * G * (rho(k + 1, tl, sl) - rho(k + 1, tu, su)) * g["odzt"][k + 1]
        assert o["rict"][k, j, i] == pytest.approx(want, rel=1e-10, abs=1e-18)
        if f["vit"][k + 1, j, i] > 0:
            wrong = f["vit"][k + 1, j, i] * g["od0"] * G * (rho(k, tl, sl) - rho(k, tu, su)) * g["odzt"][k + 1]
            assert abs(wrong - want) > 1e-6 * abs(want)


def test_ric_equals_rict_on_horizontally_uniform_atb():
    c, g, f = base(seed=8)
    prof = f["atb"][:, :, 6, 9].copy()
    f["atb"][:] = prof[:, :, None, None]
    o = run(c, g, f)
    jm, im = c.jmt, c.imt
    wet = (f["viv"][1:] > 0) & (f["vit"][1:] > 0)
    inner = np.zeros_like(wet)
    inner[:, :jm - 1, 1:] = True
    sel = wet & inner
    assert sel.sum() > 100
    assert np.allclose(o["ric"][sel], o["rict"][sel], rtol=1e-10, atol=1e-18)
    assert np.all(o["ric"][:, jm - 1, :] == 0.0) and np.all(o["ric"][:, :, 0] == 0.0)
    assert np.all(o["rict"][:, jm - 1, :] == 0.0) and np.all(o["rict"][:, :, 0] == 0.0)


def test_ricdttms_ricdt_identity(case):
    c, g, f, o = case
    T = f["at"][0]
    dT = T[:-1] - T[1:]
    al = o["alpha"][1:]
    wet = (f["vit"][1:] > 0)
    wet[:, :, 0] = False
    wet[:, :, -1] = False
    lhs = o["ricdttms"][wet]
    rhs = (G * g["odzt"][1:, None, None] * al * (dT + o["ricdt"] * (dT + 1.0e-25)))[wet]
    assert np.allclose(lhs, rhs, rtol=1e-9, atol=1e-16)
    assert wet.sum() > 100
    # the outer columns keep the clear's zero; on land (vit = 0, alpha = 0)
    # ricdt is 0/0, as in the reference
    assert np.all(o["ricdttms"][:, :, 0] == 0.0) and np.all(o["ricdt"][:, :, -1] == 0.0)
    land = ~(f["vit"][1:] > 0)
    land[:, :, 0] = False
    land[:, :, -1] = False
    assert land.sum() > 0 and np.all(np.isnan(o["ricdt"][land]))


def test_pressure_gradient_of_linear_pp_and_its_integral():
    c, g, f = uniform_density_case()
    # density linear in i and j: T = to + 1 + a i + b j (linear EOS)
    jj, ii = np.meshgrid(np.arange(c.jmt), np.arange(c.imt), indexing="ij")
    f["vit"][:] = 1.0          # every column wet to the bottom: pp linear in i, j
    f["viv"][:] = 1.0
    f["at"][0] = (g["to"] + 1.0)[:, None, None] + 0.01 * ii[None] - 0.02 * jj[None]
    f["at"][1] = np.broadcast_to(g["so"][:, None, None], f["at"][1].shape)
    g["kmu"][:] = c.km
    g["kmu"][-1, :] = 0
    g["kmu"][:, 0] = 0
    o = run(c, g, f)
    pp = o["pp"]
    for k in range(2, c.km):
        A = pp[k, 4, 6] - pp[k, 4, 5]
        B = pp[k, 5, 6] - pp[k, 4, 6]
        assert pp[k, 7, 11] - pp[k, 7, 10] == pytest.approx(A, rel=1e-9)
        sl = (slice(0, c.jmt - 1), slice(1, c.imt))
        assert np.allclose(o["dlu"][k][sl], A * g["dxur"][sl], rtol=1e-8, atol=0)
        assert np.allclose(o["dlv"][k][sl], B * g["dyur"][sl], rtol=1e-8, atol=0)
    # pxb: the stencil of the column integrals P = sum dzp pp
    P = np.tensordot(g["dzp"], pp, axes=1)
    ps = (slice(0, c.jmt - 1), slice(1, c.imt))
    sten_x = 0.5 * g["dxur"][ps] * (P[1:, 1:] - P[:-1, :-1] - P[1:, :-1] + P[:-1, 1:])
    sten_y = 0.5 * g["dyur"][ps] * (P[1:, 1:] - P[:-1, :-1] + P[1:, :-1] - P[:-1, 1:])
    assert np.allclose(o["pxb"][ps], g["ohbu"][ps] * sten_x, rtol=1e-9, atol=1e-20)
    assert np.allclose(o["pyb"][ps], g["ohbu"][ps] * sten_y, rtol=1e-9, atol=1e-20)


def test_column_integrals_of_a_uniform_anomaly():
    c, g, f = uniform_density_case()
    o = run(c, g, f)
    gg = -g["od0"] * G * (-0.2)
    kmt = c.grid["kmt"]
    j, i = 6, 9
    n = int(kmt[j, i])
    H = np.sum(g["dzp"][:n])
    s0 = o["dlu"][0, j, i]
    s1 = o["dlu"][1, j, i]
    assert s0 == pytest.approx(gg, rel=1e-13)
    assert s1 == pytest.approx(gg * np.sum(g["dzp"][:n] * g["zkt"][:n]) / H, rel=1e-12)
    oh = g["ohbt"][j, i]
    assert oh == pytest.approx(1.0 / H, rel=1e-15)
    assert o["dlv"][0, j, i] == pytest.approx((s0 + s1 * oh) / G, rel=1e-13)
    assert o["dlv"][1, j, i] == pytest.approx(s1 * oh * oh, rel=1e-13)


def test_surface_terms():
    c, g, f = base(seed=9)
    jj, ii = np.meshgrid(np.arange(c.jmt), np.arange(c.imt), indexing="ij")
    f["psa"][:] = 3.0 * ii - 5.0 * jj
    o = run(c, g, f)
    sl = (slice(0, c.jmt - 1), slice(1, c.imt))
    wet = g["kmu"][sl] > 0
    assert np.allclose(o["work1"][sl][wet], 3.0 * g["dxur"][sl][wet], rtol=1e-12)
    assert np.allclose(o["work2"][sl][wet], -5.0 * g["dyur"][sl][wet], rtol=1e-12)
    assert np.all(o["work1"][sl][~wet] == 0.0)
    assert np.allclose(o["pax"], -g["od0"] * o["work1"]) and np.allclose(o["pay"], -g["od0"] * o["work2"])
    # T->U average of a constant dlv plane, on a case where it is constant
    c2, g2, f2 = uniform_density_case()
    g2["ohbt"][:] = 0.5
    o2 = run(c2, g2, f2)
    d0 = o2["dlv"][0]
    j, i = 6, 9
    w = g2["au0"][j, i] * d0[j, i] + g2["aus"][j, i] * d0[j + 1, i] + g2["auw"][j, i] * d0[j, i - 1] + \
        g2["ausw"][j, i] * d0[j + 1, i - 1]
    assert o2["wgp"][j, i] == pytest.approx(w, rel=1e-14)
    # buoyancy fluxes share alpha at the surface: their ratio is nswv / swv
    a0 = o["alpha"][0]
    sel = (a0 != 0.0) & (f["vit"][0] > 0)
    assert np.allclose(o["buoytur"][sel] * f["swv"][sel], o["buoysol"][sel] * f["nswv"][sel], rtol=1e-12)


def test_shift_clears_and_coverage():
    c, g, f = base(seed=10)
    sh = ls.readyt_fields(c, g)
    f["ax"] = np.full((2, c.km, c.jmt, c.imt), 5.0)
    f["ay"] = f["ax"].copy()
    f["az"] = f["ax"].copy()
    o1 = run(c, g, f, ist=1)
    o0 = run(c, g, f, ist=0)
    for k in ("ax", "ay", "az"):
        assert np.all(o1[k] == 5.0) and np.all(o0[k] == 0.0)
    assert np.array_equal(o1["utl"], f["utf"]) and np.array_equal(o1["utf"], f["u"])
    assert np.array_equal(o1["vtl"], f["vtf"]) and np.array_equal(o1["vtf"], f["v"])
    assert np.array_equal(o1["h0l"], f["h0f"]) and np.array_equal(o1["h0f"], f["h0"])
    assert np.all(o1["akt"] == 0.0) and np.all(o1["rit"] == 0.0)
    for k in ("dlu", "dlv", "pdensity", "gg", "pp", "ppa", "ppb", "ppc", "alpha", "beta", "ric",
              "rict", "ricdttms", "ricdt", "buoytur", "buoysol", "pxb", "pyb", "wgp", "work",
              "whx", "why", "work1", "work2", "pax", "pay"):
        assert not np.any(o1[k] == 7.0), f"{k} has unwritten cells"
    del sh
