"""Pins of the oracle's explicit tendency terms (tracer_5 .. tracer_7 and the
del4 / AKT-floor stages) against physical closed forms.

Each test switches every other term off (no velocity, AH = 0 or no
horizontal structure, AKT = 0 with a zero floor, no forcing) and reads the
explicit tendency `tf` that enters tracer_8 (tracer.cpp:1629-1653), which
the oracle exports for these tests.  The expected values are computed in a
different form from the reference's statements (interface fluxes and their
divergence, eigenvalues of the discrete Laplacian, relaxation rates), so an
index slip (pen(k) for pen(k-1)), a sign slip or a dropped factor fails:

* shortwave penetration  (tracer.cpp:1409-1433) -> divergence of the
  interface shortwave flux Q*pen, absorbed at the last wet level; zero
  column total
* explicit vertical diffusion (tracer.cpp:1436-1475) -> half the divergence
  of the interface diffusive flux kappa*dT/dz with no flux through the
  surface or the floor; zero on a linear profile, kappa*T'' on a quadratic
* FRC_CORE salt flux (tracer.cpp:1479-1524) and heat flux (:1543-1566) ->
  virtual salt flux of the freshwater input plus relaxation of the top
  layer towards sss / sst, at rate gamma under ice and gamma*50*30/(4*365)
  over open water (salt only); half in the explicit tendency, the whole
  flux in the step's content change
* restoring (tracer.cpp:1571-1627) -> relaxation at 1/(15 days) towards the
  restore field below the top level only
* del4 (tracer.cpp:1238-1362) -> AH*AHF*(13-point biharmonic) on a uniform
  grid; Fourier eigenvalue of the biharmonic
* AKT surface floor (tracer.cpp:1187 vs 1315) -> del2 floors both tracers'
  k=0 diffusivity, del4 only the temperature's, inside the one-cell ring
"""
import math

import numpy as np
import pytest

import licom_synth as ls
import orc

AIDIF = 0.5


def quiet_case(nx=12, ny=8, km=8, hmix=2, ah=0.0, dz=None, kmt=None, seed=1):
    """A case with no velocity, no horizontal mixing (AH = 0 unless given),
    AKT = 0 with a zero floor and no forcing, on uniform Cartesian metrics.
    dz: level thicknesses (default a stretched profile so that odzp/odzt
    differ per level); kmt: 2-D interior kmt (default km everywhere)."""
    if dz is None:
        dz = np.linspace(10.0, 80.0, km)
    dz = np.asarray(dz, np.float64)
    km = len(dz)
    c = ls.make_case("C1", nx=nx, ny=ny, km=km, hmix=hmix, ah=ah, seed=seed, depth_min=6000.0,
                     depth_max=6000.0)
    g = c.grid
    dx, dy = 40e3, 30e3
    g["htw"][:] = dy
    g["hts"][:] = dx
    g["hun"][:] = dx
    g["hue"][:] = dy
    g["tarea"][:] = dx * dy
    g["zkp"][:] = np.concatenate([[0.0], -np.cumsum(dz)])
    g["kmt"][:] = 0
    g["kmt"][2:-2, 2:-2] = km if kmt is None else kmt
    g["pen"][:] = 0.0
    st = c.state
    for f in ("utf", "vtf", "utl", "vtl", "ws", "h0f", "h0l", "akt"):
        st[f][:] = 0.0
    c.params.update(dwndmix=0.0, sssnorm=0, convadj=0, gamma=0.0)
    return c


def step_tf(c, st=None, **over):
    """One oracle step; returns (state after, diag with the explicit tendency)."""
    o = orc.Oracle(c)
    st = c.copy_state() if st is None else st
    d = orc.Oracle.new_diag(c)
    d["tf"] = np.zeros((2, c.km, c.jmt, c.imt))
    o.step(st, d, **over)
    return o, st, d


def vit_of(o):
    return o.d["vit"]


def set_profile(c, n, prof_at, prof_atb=None):
    """Horizontally uniform columns: at[n] = prof_at(k), atb[n] level k =
    prof_atb(k) (default: the same)."""
    st = c.state
    prof_atb = prof_at if prof_atb is None else prof_atb
    st["at"][n][:] = np.asarray(prof_at)[:, None, None]
    st["atb"][n][1:] = np.asarray(prof_atb)[:, None, None]
    st["atb"][n][0] = st["atb"][n][1]


# ------------------------------------------------------ shortwave penetration
def _shortwave_case(kmt=None):
    km = 9
    c = quiet_case(km=km, kmt=kmt)
    dz = -np.diff(c.grid["zkp"])
    zb = np.cumsum(dz)                                   # depth of each level's bottom
    frac = 0.58 * np.exp(-zb / 0.35) + 0.42 * np.exp(-zb / 23.0)   # two-band fractions
    c.grid["pen"][:] = frac[:-1] * c.params["od0cp"]
    rng = np.random.default_rng(5)
    c.forcing["swv"][:] = rng.uniform(50.0, 400.0, c.forcing["swv"].shape)
    set_profile(c, 0, np.linspace(20.0, 4.0, km))
    set_profile(c, 1, np.zeros(km))
    return c


def _penetration_flux_form(c, o):
    """Divergence of the shortwave flux through the level interfaces: the
    flux Q*pen(m-1) reaches interface m only when level m is wet; nothing
    enters through the surface (that part is in stf) or leaves through the
    floor."""
    km = c.km
    vit = vit_of(o)
    Q = c.forcing["swv"]
    flux = np.zeros((km + 1,) + Q.shape)
    for m in range(1, km):
        flux[m] = Q * c.grid["pen"][m - 1] * vit[m]
    return (flux[:-1] - flux[1:]) * o.d["odzp"][:, None, None]


@pytest.mark.parametrize("ragged", [False, True])
def test_shortwave_penetration_is_the_flux_divergence(ragged):
    kmt = None
    if ragged:
        rng = np.random.default_rng(2)
        kmt = rng.integers(1, 10, size=(8, 12)).astype(np.int32)
    c = _shortwave_case(kmt)
    o, st, d = step_tf(c)
    ref = _penetration_flux_form(c, o)[:, 2:-2, 2:-2]
    scale = np.abs(ref).max()
    assert scale > 0
    np.testing.assert_allclose(d["penetrate"][:, 2:-2, 2:-2], ref, rtol=1e-13, atol=1e-15 * scale)
    # the tendency adds the explicit half of the surface flux swv/(rho0 cp)
    # to the top level (tracer_6)
    surf = 0.5 * c.forcing["swv"] * c.params["od0cp"] / o.d["dzp"][0]
    ref_tf = ref.copy()
    ref_tf[0] += np.where(o.kmt > 0, surf, 0.0)[2:-2, 2:-2]
    got = d["tf"][0][:, 2:-2, 2:-2]
    np.testing.assert_allclose(got, ref_tf, rtol=1e-13, atol=1e-15 * scale)
    # salinity is never heated
    assert not np.any(d["tf"][1][:, 2:-2, 2:-2])
    # penetration only redistributes: every column's total is zero
    dz = o.d["dzp"][:, None, None]
    tot = (dz * d["penetrate"]).sum(0)[2:-2, 2:-2]
    assert np.abs(tot).max() <= 1e-13 * (dz * np.abs(ref).max()).sum()
    if ragged:
        # a one-level column keeps everything at the top: no penetration
        one = o.kmt[2:-2, 2:-2] == 1
        assert one.any() and not np.any(d["penetrate"][:, 2:-2, 2:-2][:, one])


def test_shortwave_heats_the_column_by_the_surface_flux():
    """With stf = swv*od0cp the column content gains exactly swv*od0cp*dts
    per unit area (half explicit, half implicit); penetration moves part of
    it below the top level but adds nothing."""
    c = _shortwave_case()
    o = orc.Oracle(c)
    st = c.copy_state()
    before = orc.content(c, o, st)[0]
    o.step(st)
    after = orc.content(c, o, st)[0]
    area = c.grid["tarea"][2:-2, 2:-2]
    q = (c.forcing["swv"] * c.params["od0cp"])[2:-2, 2:-2]
    assert after - before == pytest.approx(float((area * q).sum()) * c.params["dts"], rel=1e-11)
    # and the surface level alone gains less than the full flux
    top = (st["at"][0][0] - c.state["at"][0][0])[2:-2, 2:-2] * o.d["dzp"][0]
    assert np.all(top < q * c.params["dts"])


# ------------------------------------------------ explicit vertical diffusion
def _diffusive_flux_form(c, o, B, kappa):
    """Half the divergence of the downward diffusive flux
    D(m) = kappa(m-1) * (B(m-1) - B(m)) / (zt(m-1) - zt(m)) through interface
    m, with no flux through the surface or the sea floor; B and kappa are
    [km, jmt, imt]."""
    km = c.km
    vit = vit_of(o)
    zkp = c.grid["zkp"]
    zt = 0.5 * (zkp[:-1] + zkp[1:])
    D = np.zeros((km + 1,) + B.shape[1:])
    for m in range(1, km):
        D[m] = kappa[m - 1] * (B[m - 1] - B[m]) / (zt[m - 1] - zt[m]) * vit[m]
    dz = -np.diff(zkp)
    return 0.5 * (D[:-1] - D[1:]) / dz[:, None, None]


@pytest.mark.parametrize("ragged", [False, True])
def test_explicit_vertical_diffusion_is_half_the_flux_divergence(ragged):
    km = 10
    kmt = None
    if ragged:
        kmt = np.random.default_rng(7).integers(1, km + 1, size=(8, 12)).astype(np.int32)
    c = quiet_case(km=km, kmt=kmt)
    rng = np.random.default_rng(8)
    B = rng.normal(10.0, 3.0, (km,))
    set_profile(c, 0, 1.5 * B, B)            # at differs from atb: the term reads atb
    set_profile(c, 1, np.zeros(km), -0.5 * B)
    kappa = 10.0 ** rng.uniform(-5, -2, (2, km, c.jmt, c.imt))
    kappa[:, -1] = 0.0
    c.state["akt"][:] = kappa
    o, st, d = step_tf(c)
    for n, prof in ((0, B), (1, -0.5 * B)):
        Bn = np.broadcast_to(prof[:, None, None], (km, c.jmt, c.imt))
        ref = _diffusive_flux_form(c, o, Bn, kappa[n])[:, 2:-2, 2:-2]
        got = d["tf"][n][:, 2:-2, 2:-2]
        scale = np.abs(ref).max()
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-14 * scale)
        np.testing.assert_allclose(d["dz"][n][:, 2:-2, 2:-2], ref, rtol=1e-12, atol=1e-14 * scale)
        # no flux through the surface or the floor: the column total vanishes
        dz = o.d["dzp"][:, None, None]
        tot = (dz * d["tf"][n]).sum(0)[2:-2, 2:-2]
        assert np.abs(tot).max() <= 1e-12 * (dz * np.abs(ref).max()).sum()


def test_explicit_vertical_diffusion_linear_and_quadratic_profiles():
    km, h, kap = 12, 25.0, 3.0e-3
    c = quiet_case(km=km, dz=np.full(km, h))
    z = (np.arange(km) + 0.5) * h                      # depth of the level centres
    G, a = 0.02, 1.0e-4
    set_profile(c, 0, 15.0 - G * z)                  # linear
    set_profile(c, 1, a * z * z)                     # quadratic
    c.state["akt"][:] = kap
    c.state["akt"][:, -1] = 0.0
    o, st, d = step_tf(c)
    tT = d["tf"][0][:, 5, 6]
    tS = d["tf"][1][:, 5, 6]
    # linear: no interior divergence; the surface level loses and the bottom
    # level gains half the constant flux kappa*G over its thickness
    assert np.abs(tT[1:-1]).max() <= 1e-15
    assert tT[0] == pytest.approx(-0.5 * kap * G / h, rel=1e-12)
    assert tT[-1] == pytest.approx(0.5 * kap * G / h, rel=1e-12)
    # quadratic: half of kappa * T'' = kappa * a in the interior
    np.testing.assert_allclose(tS[1:-1], kap * a, rtol=1e-9)
    # top: only the flux below, kappa * 2a*z(1)-ish one-sided difference
    assert tS[0] == pytest.approx(0.5 * kap * (a * (z[1] ** 2 - z[0] ** 2) / h) / h, rel=1e-12)
    # bottom (the (1 - aidif) branch): only the flux from above
    assert tS[-1] == pytest.approx(-0.5 * kap * (a * (z[-1] ** 2 - z[-2] ** 2) / h) / h, rel=1e-12)


# ------------------------------------------------------------ surface fluxes
def _flux_case(seed=3):
    c = quiet_case(km=6, dz=[10.0, 20.0, 30.0, 40.0, 50.0, 60.0], seed=seed)
    rng = np.random.default_rng(seed)
    sh = c.forcing["swv"].shape
    set_profile(c, 0, np.linspace(18.0, 6.0, 6))
    set_profile(c, 1, np.linspace(-0.4e-3, 0.2e-3, 6))
    f = c.forcing
    f["sst"][:] = rng.uniform(-1.0, 28.0, sh)
    f["sss"][:] = rng.uniform(-2e-3, 1e-3, sh)
    f["fresh"][:] = rng.normal(0.0, 3e-5, sh)
    f["seaice"][:] = rng.choice([0.0, 1.0, 0.3], size=sh)
    c.params["gamma"] = 1.0 / (30.0 * 86400.0)
    return c


def _salt_flux(c):
    """Virtual salt flux of the freshwater input at S_ref = 34.7 psu in model
    units (/1000, divided by the reference density), plus relaxation of the
    top layer towards sss: the piston velocity is gamma*dz0 under ice and
    gamma*dz0*50*30/(4*365) over open water."""
    g = c.params["gamma"]
    dz0 = -c.grid["zkp"][1]
    f = c.forcing
    S0 = c.state["atb"][1][1]
    v_ice = g * dz0
    v_open = g * dz0 * 50.0 * 30.0 / (4.0 * 365.0)
    vfs = f["fresh"] * 34.7 / (1000.0 * ls.D0)
    return vfs + (f["sss"] - S0) * (v_ice * f["seaice"] + v_open * (1.0 - f["seaice"]))


def _heat_flux(c):
    """(swv + nswv)/(rho0 cp) plus relaxation of the top layer towards sst at
    rate gamma under ice only."""
    g = c.params["gamma"]
    dz0 = -c.grid["zkp"][1]
    f = c.forcing
    T0 = c.state["atb"][0][1]
    return (f["swv"] + f["nswv"]) * c.params["od0cp"] + g * dz0 * f["seaice"] * (f["sst"] - T0)


def test_frc_core_salt_flux_closed_form():
    c = _flux_case()
    o, st, d = step_tf(c)
    stf = _salt_flux(c)
    dz0 = -c.grid["zkp"][1]
    wet = o.kmt > 0
    np.testing.assert_allclose(d["net"][1][wet], (stf / dz0)[wet], rtol=1e-12, atol=1e-22)
    # half of the flux in the explicit tendency of the top level
    np.testing.assert_allclose(d["tf"][1][0][2:-2, 2:-2], ((1 - AIDIF) * stf / dz0)[2:-2, 2:-2],
                               rtol=1e-12, atol=1e-22)
    assert not np.any(d["tf"][1][1:, 2:-2, 2:-2])
    # the three regimes are all present and differ
    s = c.forcing["seaice"][2:-2, 2:-2]
    assert (s == 0).any() and (s == 1).any() and (s == 0.3).any()


def test_frc_core_salt_flux_changes_content_by_the_flux():
    c = _flux_case(4)
    o = orc.Oracle(c)
    st = c.copy_state()
    before = orc.content(c, o, st)[1]
    o.step(st)
    after = orc.content(c, o, st)[1]
    area = c.grid["tarea"][2:-2, 2:-2]
    expect = float((area * _salt_flux(c)[2:-2, 2:-2]).sum()) * c.params["dts"]
    assert after - before == pytest.approx(expect, rel=1e-10)


def test_frc_core_heat_flux_closed_form():
    c = _flux_case(5)
    c.forcing["nswv"][:] = np.random.default_rng(1).normal(0.0, 80.0, c.forcing["nswv"].shape)
    o, st, d = step_tf(c)
    stf = _heat_flux(c)
    dz0 = -c.grid["zkp"][1]
    wet = o.kmt > 0
    np.testing.assert_allclose(d["net"][0][wet], (stf / dz0)[wet], rtol=1e-12, atol=1e-20)
    np.testing.assert_allclose(d["tf"][0][0][2:-2, 2:-2], (0.5 * stf / dz0)[2:-2, 2:-2],
                               rtol=1e-12, atol=1e-20)
    # ice-free cells see no relaxation towards sst
    c2 = _flux_case(5)
    c2.forcing["nswv"][:] = c.forcing["nswv"]
    c2.forcing["sst"][:] += 5.0
    _, _, d2 = step_tf(c2)
    free = c.forcing["seaice"] == 0.0
    np.testing.assert_array_equal(d2["net"][0][free & wet], d["net"][0][free & wet])
    ice = (c.forcing["seaice"] == 1.0) & wet
    np.testing.assert_allclose((d2["net"][0] - d["net"][0])[ice], c.params["gamma"] * 5.0, rtol=1e-9)
    # content: the whole flux over the step
    o3 = orc.Oracle(c)
    st3 = c.copy_state()
    before = orc.content(c, o3, st3)[0]
    o3.step(st3)
    area = c.grid["tarea"][2:-2, 2:-2]
    expect = float((area * stf[2:-2, 2:-2]).sum()) * c.params["dts"]
    assert orc.content(c, o3, st3)[0] - before == pytest.approx(expect, rel=1e-10)


# ----------------------------------------------------------------- restoring
def test_restoring_relaxes_below_the_top_level_at_fifteen_days():
    km = 7
    c = quiet_case(km=km, kmt=np.random.default_rng(3).integers(0, km + 1, (8, 12)).astype(np.int32))
    B = np.arange(km, dtype=np.float64) + 3.0
    set_profile(c, 0, B)
    set_profile(c, 1, -B)
    off = np.array([0.5, -0.25])
    R = np.zeros((2, km, c.jmt, c.imt))
    R[0] = (B + off[0])[:, None, None]
    R[1] = (-B + off[1])[:, None, None]
    c.forcing["restore"] = R
    c.params["boundary_restore"] = 1
    o, st, d = step_tf(c)
    lam = 1.0 / (15.0 * 86400.0)
    vit = vit_of(o)
    for n in range(2):
        got = d["tf"][n][:, 2:-2, 2:-2]
        assert not np.any(got[0])                     # the top level is not restored
        np.testing.assert_allclose(got[1:], (off[n] * lam * vit[1:])[:, 2:-2, 2:-2], rtol=1e-14, atol=0)
    # switched off, nothing
    _, _, d0 = step_tf(c, boundary_restore=0)
    assert not np.any(d0["tf"][:, :, 2:-2, 2:-2])


# ---------------------------------------------------------------------- del4
def _lap_periodic(X, dx, dy):
    """5-point Laplacian, periodic in i (axis -1), rows (axis -2) as given."""
    return ((np.roll(X, -1, -1) + np.roll(X, 1, -1) - 2 * X) / dx ** 2
            + (np.roll(X, -1, -2) + np.roll(X, 1, -2) - 2 * X) / dy ** 2)


def _del4_case(nx=24, ny=16, km=3):
    c = quiet_case(nx=nx, ny=ny, km=km, hmix=4, ah=-1.0e10, dz=np.full(km, 50.0))
    return c


def test_del4_is_the_thirteen_point_biharmonic_on_a_uniform_grid():
    c = _del4_case()
    dx, dy = c.grid["hts"][0, 0], c.grid["htw"][0, 0]
    rng = np.random.default_rng(11)
    B = np.zeros((c.km, c.jmt, c.imt))
    B[:, 2:-2, 2:-2] = rng.normal(0.0, 1.0, (c.km, c.ny, c.nx))
    B = orc.halo_fill(B)
    c.state["atb"][0][1:] = B
    c.state["atb"][1][1:] = 2.0 * B
    c.state["at"][0] = B
    c.state["at"][1] = 2.0 * B
    o, st, d = step_tf(c)
    ahf = o.d["ahf"][5, 5]
    assert np.ptp(o.d["ahf"]) == 0.0
    # interior rows at least two away from the walls, every column (periodic)
    P = B[:, 2:-2, 2:-2]
    bih = _lap_periodic(_lap_periodic(P, dx, dy), dx, dy)
    rows = slice(2, -2)
    for n, f in ((0, 1.0), (1, 2.0)):
        ref = c.params["ah"] * ahf * f * bih[:, rows]
        got = d["tf"][n][:, 2:-2, 2:-2][:, rows]
        np.testing.assert_allclose(got, ref, rtol=1e-9, atol=1e-10 * np.abs(ref).max())


def test_del4_fourier_eigenvalue():
    """A zonal wave sin(kx) is an eigenfunction of the discrete biharmonic:
    L^2 sin = (2 - 2 cos(k dx))^2 / dx^4 * sin."""
    nx = 32
    c = _del4_case(nx=nx, ny=12)
    dx = c.grid["hts"][0, 0]
    i = np.arange(c.imt) - 2
    kdx = 2 * np.pi * 3 / nx
    w = np.sin(kdx * i)
    B = np.broadcast_to(w, (c.km, c.jmt, c.imt)).copy()
    B[:, :2] = B[:, -2:] = 0.0
    c.state["atb"][0][1:] = B
    c.state["at"][0] = B
    o, st, d = step_tf(c)
    lam = (2.0 - 2.0 * np.cos(kdx)) ** 2 / dx ** 4
    ref = c.params["ah"] * o.d["ahf"][5, 5] * lam * w[2:-2]
    got = d["tf"][0][1, 6, 2:-2]                 # a row far from the walls
    np.testing.assert_allclose(got, ref, rtol=1e-8, atol=1e-9 * np.abs(ref).max())


# --------------------------------------------------------------- AKT floor
@pytest.mark.parametrize("hmix", [2, 4])
def test_akt_surface_floor_depends_on_the_scheme(hmix):
    km = 5
    c = quiet_case(km=km, hmix=hmix, ah=0.0)
    c.params["dwndmix"] = 1.0e-3
    rng = np.random.default_rng(9)
    akt = 10.0 ** rng.uniform(-6, -4, (2, km, c.jmt, c.imt))      # all below the floor
    akt[:, -1] = 0.0
    c.state["akt"][:] = akt
    B = np.linspace(12.0, 2.0, km)
    set_profile(c, 0, B)
    set_profile(c, 1, 0.1 * B)
    o, st, d = step_tf(c)
    ring = (slice(1, -1), slice(1, -1))
    for n in range(2):
        floored = n == 0 or hmix == 2
        top = st["akt"][n][0]
        if floored:
            assert np.all(top[ring] == 1.0e-3)
        else:
            np.testing.assert_array_equal(top, akt[n][0])
        # the outermost ghost ring and the deeper levels are untouched
        np.testing.assert_array_equal(top[0], akt[n][0][0])
        np.testing.assert_array_equal(top[:, 0], akt[n][0][:, 0])
        np.testing.assert_array_equal(st["akt"][n][1:], akt[n][1:])
        # the surface level's explicit diffusion uses the floored value
        kap0 = 1.0e-3 if floored else akt[n][0][2:-2, 2:-2]
        dz = -np.diff(c.grid["zkp"])
        zt = 0.5 * (c.grid["zkp"][:-1] + c.grid["zkp"][1:])
        prof = B if n == 0 else 0.1 * B
        ref = -0.5 * kap0 * (prof[0] - prof[1]) / (zt[0] - zt[1]) / dz[0]
        np.testing.assert_allclose(d["tf"][n][0][2:-2, 2:-2], ref, rtol=1e-12)
