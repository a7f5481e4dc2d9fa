"""Pins of the oracle's Canuto producer (oracle/turb.c) against what the
mathematics fixes, so that a dropped term, a wrong index or a swapped table
in the restatement fails here and not only in a GPU comparison with itself:

* formld (readyc.cpp:902-913): on a linear temperature profile the
  interpolated depth is where T has moved 0.1 K from the surface value,
  warming or cooling, and the deepest depth when it never does;
* interp1d_expabs (:915-975): piecewise-linear interpolation on the expabs
  grid, equal to numpy.interp on a nonlinear table (a wrong cell would
  extrapolate), clamped at both ends;
* interp2d_expabs (:977-1115): on a separable table g(Ri) + h(Ri_d) the
  corner-and-slopes interpolation equals np.interp(Ri) + np.interp(Ri_d);
  beyond the realizable Ri of either bracketing Ri_d row the result is zero;
  points outside the table move back along their ray to the table edge;
* eplatidepend_fun (:882-896): its reference latitude is 30 degrees
  (f = Omega = 2 pi / sidereal day), where the factor is 1, and it vanishes
  at an = f;
* the column closure (turb_device): the zero-shear branch scales with
  sqrt(-N^2); the shear branch with sqrt(S^2) on top of an additive
  background; the background with the latitude factor; the surface floor;
* turb_cu's composition (:1487-1572): AKT divides by ncc, the tidal term is
  the running minimum of its column below, linear in wave_dis and fz_tide,
  clipped to [1e-6, 1e-2]; land and the outer ring are cleared.
"""
import numpy as np
import pytest

import licom_synth as ls
import orc

MT, MT0 = orc.TURB_MT, orc.TURB_MT0


@pytest.fixture(scope="module")
def tables():
    return ls.turb_tables()


@pytest.fixture(scope="module")
def oracle(tables):
    return orc.TurbOracle(tables)


# ---------------------------------------------------------------- formld
@pytest.mark.parametrize("sign", [-1.0, 1.0])
def test_formld_linear_profile(sign):
    z = np.arange(1, 31) * 1000.0            # cm
    g = 2.0e-5                               # K per cm
    t = 20.0 + sign * g * z
    d = orc.turb_formld(z, t)
    # T departs by 0.1 K from t[0] at z0 + 0.1/g
    assert d == pytest.approx(z[0] + 0.1 / g, rel=1e-12)


def test_formld_never_departs_returns_deepest():
    z = np.arange(1, 11) * 500.0
    t = 15.0 + 1e-3 * np.sin(np.arange(10))
    assert orc.turb_formld(z, t) == z[-1]


# ---------------------------------------------------------------- interp1d
def test_interp1d_equals_piecewise_linear_on_expabs_grid(tables):
    tb = dict(tables)
    x = tb["and2on2a1"]
    tb["amtaun2a1"] = x ** 2 + 1.0
    tb["sma1"] = np.sin(x)
    tb["sha1"] = np.abs(x) ** 1.5
    tb["ssa1"] = np.cos(3.0 * x)
    o = orc.TurbOracle(tb)
    rng = np.random.default_rng(1)
    pts = np.concatenate([rng.uniform(-4, 4, 200), rng.uniform(-57, 57, 200),
                          x[rng.integers(0, len(x), 50)], [0.0, 3.999, 4.0, 4.01, -4.0]])
    for p in pts:
        st, out = o.interp1d(p)
        assert st == 1
        for m, f in enumerate(("amtaun2a1", "sma1", "sha1", "ssa1")):
            ref = np.interp(p, x, tb[f])
            assert out[m] == pytest.approx(ref, rel=1e-12, abs=1e-14), (p, f)


def test_interp1d_clamps_at_table_ends(tables, oracle):
    for p, e in ((1e6, -1), (-1e6, 0), (tables["and2on2a1"][-1], -1)):
        st, out = oracle.interp1d(p)
        assert st == 1
        assert out[0] == tables["amtaun2a1"][e]


# ---------------------------------------------------------------- interp2d
def separable(tables):
    tb = dict(tables)
    r = tb["rib"]
    g = {"slq2b": r ** 2, "smb": np.sin(r), "shb": np.abs(r) ** 1.5, "ssb": np.cos(r)}
    h = {"slq2b": np.cos(0.3 * r), "smb": r, "shb": np.tanh(r), "ssb": r ** 3 * 1e-4}
    for k in g:
        tb[k] = np.ascontiguousarray(g[k][None, :] + h[k][:, None])
    tb["irimax"] = np.full(2 * MT + 1, MT, np.int32)     # everything realizable
    return tb, g, h


def test_interp2d_separable_table_is_sum_of_1d_interpolations(tables):
    tb, g, h = separable(tables)
    o = orc.TurbOracle(tb)
    r = tb["rib"]
    rng = np.random.default_rng(2)
    pts = [(rng.uniform(-4, 4), rng.uniform(-4, 4)) for _ in range(150)]
    pts += [(rng.uniform(-50, 50), rng.uniform(-50, 50)) for _ in range(150)]
    pts += [(0.0, 0.0), (r[MT + 7], r[MT - 3]), (4.0, -4.0), (30.0, 0.0)]
    for ri, rid in pts:
        st, out, ri2, rid2 = o.interp2d(ri, rid)
        assert st == 1 and (ri2, rid2) == (ri, rid)
        for m, k in enumerate(("slq2b", "smb", "shb", "ssb")):
            ref = np.interp(ri, r, g[k]) + np.interp(rid, r, h[k])
            assert out[m] == pytest.approx(ref, rel=1e-11, abs=1e-12), (ri, rid, k)


def test_interp2d_zero_beyond_realizable_ri(tables, oracle):
    r, irimax = tables["rib"], tables["irimax"]
    rng = np.random.default_rng(3)
    n_zero = n_set = 0
    for _ in range(400):
        ri, rid = rng.uniform(-2, 2), rng.uniform(-3, 3)
        # bracketing Ri_d rows, independently of the oracle's search
        j1 = int(np.searchsorted(r, rid))
        rows = (j1 - 1, j1) if rid > 0 else (j1, j1 - 1) if rid < 0 else (MT, MT + 1)
        lim = min(r[MT + irimax[rows[0]]], r[MT + irimax[rows[1]]])
        st, out, _, _ = oracle.interp2d(ri, rid)
        if ri > lim:
            assert st == 2 and not out.any()
            n_zero += 1
        else:
            assert st == 1 and out[0] > 0
            n_set += 1
    assert n_zero > 50 and n_set > 50


@pytest.mark.parametrize("ri,rid", [(80.0, 20.0), (80.0, 200.0), (80.0, -300.0), (-90.0, 10.0),
                                    (-90.0, 300.0), (-90.0, -300.0), (10.0, 100.0), (10.0, -100.0)])
def test_interp2d_moves_outside_points_along_their_ray(tables, ri, rid):
    tb, _, _ = separable(tables)
    o = orc.TurbOracle(tb)
    lo, hi = tb["rib"][0], tb["rib"][-1]
    st, out, ri2, rid2 = o.interp2d(ri, rid)
    assert lo <= ri2 <= hi and lo <= rid2 <= hi
    assert max(abs(ri2), abs(rid2)) == pytest.approx(hi, rel=1e-15)
    assert ri2 / rid2 == pytest.approx(ri / rid, rel=1e-13)


# ---------------------------------------------------------------- eplatidepend
def test_eplatidepend_reference_latitude_and_limit():
    omega = 2.0 * np.pi / 86164.0          # sidereal day: f at 30 degrees
    assert orc.turb_eplatidepend(omega, 5.24e-3) == pytest.approx(1.0, rel=2e-4)
    f30 = np.pi / 43082.0
    assert orc.turb_eplatidepend(f30, 5.24e-3) == pytest.approx(1.0, rel=1e-14)
    assert orc.turb_eplatidepend(1e-4, 1e-4) == 0.0
    # increasing in an, at fixed f
    a = [orc.turb_eplatidepend(5e-5, an) for an in (1e-4, 1e-3, 1e-2)]
    assert a[0] < a[1] < a[2]


# ---------------------------------------------------------------- column closure
def const_tables(tables, sm=0.11, sh=0.13, ss=0.12, amtaun2=1.7, slq2=2.0, back=None):
    tb = dict(tables)
    w = 2 * MT + 1
    for k, v in (("amtaun2a1", amtaun2), ("sma1", sm), ("sha1", sh), ("ssa1", ss)):
        tb[k] = np.full(w, v)
    for k, v in (("slq2b", slq2), ("smb", sm), ("shb", sh), ("ssb", ss)):
        tb[k] = np.full((w, w), v)
    tb["irimax"] = np.full(w, MT, np.int32)
    if back is not None:
        nb = len(tables["sm_r1"])
        for k, v in zip(("back_ra_r", "sm_r1", "sh_r1", "ss_r1", "slq2_r1"), back):
            tb[k] = np.full(nb, v)
    return tb


def one_level(o, an2, s2, ri, rid, z=3000.0, coriol=7e-5, nlev=2, buoy=-1e-3):
    """A column whose level 0 carries (an2, s2, ri, rid); deeper levels
    stable.  A destabilising surface buoyancy flux keeps the dissipation
    estimate from the surface flux (lifepsy) out of the way."""
    z_ = np.array([z, z + 2000.0] + [0.0] * (nlev - 2))
    t = np.array([20.0, 19.99] + [0.0] * (nlev - 2))
    rh = np.full(nlev, 1.025)
    an2_ = np.array([an2, 1e-5] + [0.0] * (nlev - 2))
    s2_ = np.array([s2, 1e-5] + [0.0] * (nlev - 2))
    ri_ = np.array([ri, 1.0] + [0.0] * (nlev - 2))
    rid_ = np.array([rid, 0.0] + [0.0] * (nlev - 2))
    return o.column(2, nlev, z_, t, rh, ri_, rid_, s2_, an2_, 1e9, 0.0, buoy, 0.0, coriol)


def test_zero_shear_branch_scales_with_sqrt_minus_n2(tables):
    o = orc.TurbOracle(const_tables(tables))
    r1 = one_level(o, -1e-6, 1e-12, -1e6, 0.0)
    r4 = one_level(o, -4e-6, 1e-12, -1e6, 0.0)
    assert r1["counts"]["pure"] == 1 and r1["exit"] == -1
    for k in ("akm", "akh", "aks"):
        assert r4[k][0] == pytest.approx(2.0 * r1[k][0], rel=1e-13)
    # the three diffusivities keep the table's ratios
    assert r1["akh"][0] / r1["akm"][0] == pytest.approx(0.13 / 0.11, rel=1e-13)


def test_shear_branch_sqrt_s2_over_an_additive_background(tables):
    o = orc.TurbOracle(const_tables(tables, back=(0.1, 0.05, 0.07, 0.06, 2.0)))
    ri, rid = 0.2, 0.05
    a = {m: one_level(o, 2e-6, m * 1e-5, ri, rid)["akm"][0] for m in (1, 4, 9)}
    v = 2.0 * a[1] - a[4]                      # background: akm(4 S^2) - v = 2 (akm(S^2) - v)
    assert v > 0
    assert a[9] - v == pytest.approx(3.0 * (a[1] - v), rel=1e-11)


def test_background_follows_the_latitude_factor(tables):
    o = orc.TurbOracle(const_tables(tables, back=(0.1, 0.05, 0.07, 0.06, 2.0)))
    ri, rid, an2 = 0.2, 0.05, 2e-5

    def background(f):
        a1 = one_level(o, an2, 1e-4, ri, rid, coriol=f)["akm"][0]
        a4 = one_level(o, an2, 4e-4, ri, rid, coriol=f)["akm"][0]
        return 2.0 * a1 - a4

    f1, f2 = 3e-5, 1.2e-4
    e1 = max(orc.turb_eplatidepend(f1, np.sqrt(an2)), 0.07)
    e2 = max(orc.turb_eplatidepend(f2, np.sqrt(an2)), 0.07)
    assert e1 != e2
    assert background(f1) / background(f2) == pytest.approx(e1 / e2, rel=1e-9)


def test_surface_floor_applies_to_level_zero_only(tables):
    o = orc.TurbOracle(const_tables(tables))
    z = np.array([3000.0, 5000.0, 8000.0])
    t = np.array([20.0, 19.99, 19.98])
    args = (np.full(3, 1.025), np.full(3, 50.0), np.zeros(3), np.full(3, 1e-8), np.full(3, 5e-7))
    r = o.column(2, 3, z, t, *args, 1e9, 1e6, 0.0, 0.0, 7e-5)
    assert r["akm"][0] == 1e6 and r["akh"][0] == 1e6
    assert r["akm"][1] < 1e6 and r["akm"][2] == 0.0      # level n is below the column


# ---------------------------------------------------------------- turb_cu
@pytest.fixture(scope="module")
def small_case(tables):
    c = ls.make_case("C1", nx=40, ny=30, km=20, seed=5)
    return ls.make_turb_case(c, seed=11, tables=tables)


def test_akt_divides_by_ncc(oracle, small_case):
    import copy
    a, _ = oracle.run(small_case)
    tc2 = copy.copy(small_case)
    tc2.params = dict(small_case.params, ncc=1)
    b, _ = oracle.run(tc2)
    ncc = small_case.params["ncc"]
    assert ncc == 2
    assert np.array_equal(b["akt"], a["akt"] * ncc)
    assert np.array_equal(b["akmt"], a["akmt"])


def test_tidal_term_is_running_minimum_and_linear(oracle, small_case):
    import copy
    out, _ = oracle.run(small_case)
    kmt = small_case.grid["kmt"]
    tide = out["ak_tide"]
    km = small_case.km
    jj, ii = np.nonzero(kmt > 2)
    jj, ii = jj[(jj > 0) & (jj < kmt.shape[0] - 1)], ii[(jj > 0) & (jj < kmt.shape[0] - 1)]
    for j, i in list(zip(jj, ii))[:300]:
        n = kmt[j, i] - 1
        col = tide[:n, j, i]
        if i == 0 or i == kmt.shape[1] - 1:
            continue
        assert np.all(col[:-1] <= col[1:])
        assert np.all((col >= 1e-6) & (col <= 1e-2))
        assert not tide[n:, j, i].any()
    # linear in wave_dis below the cap: (ak(2 wd) - 1e-6) = 2 (ak(wd) - 1e-6)
    tc2 = copy.copy(small_case)
    tc2.fields = dict(small_case.fields, wave_dis=small_case.fields["wave_dis"] * 2.0)
    out2, _ = oracle.run(tc2)
    a, b = tide - 1e-6, out2["ak_tide"] - 1e-6
    sel = (out2["ak_tide"] < 1e-2) & (tide > 0)
    assert sel.sum() > 1000
    assert np.allclose(b[sel], 2.0 * a[sel], rtol=1e-12, atol=1e-22)
    del km


def test_turb_cu_clears_land_and_ring_and_covers_every_branch(tables, oracle, small_case):
    km, jmt, imt = small_case.km, small_case.jmt, small_case.imt
    sentinel = dict(akmt=np.full((km, jmt, imt), 9.0), akt=np.full((2, km, jmt, imt), 9.0),
                    ak_tide=np.full((km, jmt, imt), 9.0), amld=np.full((jmt, imt), 9.0),
                    richardson=np.full((km, jmt, imt), 9.0), fztidal=np.full((km, jmt, imt), 9.0),
                    wp3_tidal=np.full((km, jmt, imt), 9.0))
    out, cnt = oracle.run(small_case, sentinel)
    kmt = small_case.grid["kmt"]
    wet = np.zeros((jmt, imt), bool)
    wet[1:-1, 1:-1] = kmt[1:-1, 1:-1] > 0
    assert not out["akmt"][:, ~wet].any() and not out["ak_tide"][:, ~wet].any()
    assert not out["amld"][~wet].any()
    assert np.all(out["akt"][:, :, ~wet] == 9.0)           # left alone
    assert np.all(out["akt"][:, km - 1] == 9.0)
    assert not out["akmt"][km - 1].any()
    assert np.all(out["richardson"][:, ~wet] == 9.0)
    wet_amld = out["amld"][wet]
    assert np.all(wet_amld >= 0)
    for k in ("pure", "table2d", "below", "deep", "epsy", "deardorff", "rafglt", "theta_p",
              "theta_n"):
        assert cnt[k] > 0, (k, cnt)


def test_early_exit_leaves_lower_levels_at_the_tidal_term(tables, small_case):
    """A background table with negative slq2 on part of the angle range makes
    turb_device return early (readyc.cpp:1334-1337); levels from there down
    keep zero closure diffusivity and level 0 loses its surface floor."""
    tb = dict(tables)
    th = (np.arange(len(tb["slq2_r1"])) - orc.TURB_NTH) * tb["deltheta_r"]
    tb["slq2_r1"] = np.where((th > 0.2) & (th < 0.6), -1.0, tb["slq2_r1"])
    o = orc.TurbOracle(tb)
    out, cnt = o.run(small_case)
    assert cnt["exit"] > 20
    ncc = small_case.params["ncc"]
    akh = (out["akt"][0] * ncc - out["ak_tide"])        # 1e-4 * akh
    assert (np.abs(akh) < 1e-18).sum() > 0
