"""Pins of the CPU oracle against facts fixed independently of its own code.

The reference ships no golden vectors or tests for this path (it has no
tests at all), so each stage of the oracle is pinned by something that a
plausible mistake (dropped term, wrong sign or index, transposed operand)
would break:

* invtrit                -> LAPACK banded solve of the same system, and the
                            closed-form column budget sum(dzp*x) - sum(dzp*d)
                            = dts*stf/2
* dens                   -> the 9-term polynomial evaluated monomial by
                            monomial (pins the C(k,m) index map)
* masks / coefficients   -> integer closed forms, eoshift end-off rules, the
                            del2 flux symmetry DTE(i)*tarea(i) = DTW(i+1)*tarea(i+1)
* del2 / del4            -> 5-point textbook Laplacian on a uniform grid,
                            zero on a constant field, zero area-weighted sum
* tspas                  -> Lax-Wendroff closed forms (x and z) on smooth
                            fields, constant preservation, shape preservation
                            (no new extrema) on a step, telescoping flux sum
* convadj                -> stable column unchanged, two-layer inversion to
                            the dz-weighted mean, full mix of an unstable
                            column, per-column content conservation
* whole step             -> T/S content conserved with zero forcing; a
                            uniform surface heat flux changes T content by
                            exactly sum(tarea)*stf*dts; SSSNORM removes the
                            net freshwater flux
"""
import math

import numpy as np
import pytest
import scipy.linalg

import licom_synth as ls
import orc


def case_uniform(nx=24, ny=12, km=6, dx=50e3, dy=40e3, dz=100.0, hmix=2, ah=1e3):
    """Uniform Cartesian metrics (every face the same length) and no land."""
    c = ls.make_case("C1", nx=nx, ny=ny, km=km, hmix=hmix, ah=ah, depth_min=6000.0,
                     depth_max=6000.0, dz_top=dz, dz_bot=dz, depth_total=dz * km)
    g = c.grid
    g["htw"][:] = dy
    g["hts"][:] = dx
    g["hun"][:] = dx
    g["hue"][:] = dy
    g["tarea"][:] = dx * dy
    g["kmt"][:] = 0
    g["kmt"][2:-2, 2:-2] = km
    g["zkp"][:] = -dz * np.arange(km + 1)
    return c


# ------------------------------------------------------------------ invtri
@pytest.mark.parametrize("seed", range(6))
def test_invtri_matches_banded_solver(seed):
    rng = np.random.default_rng(seed)
    km = 30
    dz = rng.uniform(5.0, 300.0, km)
    zkp = np.concatenate([[0.0], -np.cumsum(dz)])
    zkt = 0.5 * (zkp[:-1] + zkp[1:])
    odzp = 1.0 / dz
    odzt = np.empty(km)
    odzt[0] = 2.0 * odzp[0]
    odzt[1:] = 1.0 / (zkt[:-1] - zkt[1:])
    dts = 3600.0
    for kz in (1, 2, 3, km // 2, km):
        wkc = 10.0 ** rng.uniform(-5, -2, km)
        d = rng.normal(10.0, 3.0, km)
        stf = rng.normal(0.0, 1e-4)
        x = orc.invtri_column(kz, wkc, odzt, odzp, dts, stf, 1.0, d)
        # the system of invtri.F90: -a x(k-1) + b x(k) - c x(k+1) = d(k),
        # x(-1) = stf, no flux through the bottom
        a = np.zeros(kz)
        c = np.zeros(kz)
        a[0] = odzp[0] * dts * 0.5
        for k in range(1, kz):
            a[k] = wkc[k - 1] * odzt[k] * odzp[k] * dts * 0.5
        for k in range(0, kz - 1):
            c[k] = wkc[k] * odzt[k + 1] * odzp[k] * dts * 0.5
        b = 1.0 + a + c
        b[0] = 1.0 + c[0]
        if kz == 1:
            b[0] = 1.0 + a[0]
        ab = np.zeros((3, kz))
        ab[0, 1:] = -c[:-1]
        ab[1] = b
        ab[2, :-1] = -a[1:]
        rhs = d[:kz].copy()
        rhs[0] += a[0] * stf
        ref = scipy.linalg.solve_banded((1, 1), ab, rhs)
        np.testing.assert_allclose(x[:kz], ref, rtol=1e-12, atol=0)
        if kz == 1:
            # tracer.cpp:958 sets b(kz-1) = 1 + a(kz-1) after b(0) = 1 + c(0),
            # so a one-level column solves (1 + a0) x0 = d0 + a0 stf
            assert x[0] == pytest.approx((d[0] + a[0] * stf) / (1.0 + a[0]), rel=1e-14)
            continue
        # column budget: the diffusion operator is in flux form
        budget = np.sum(dz[:kz] * x[:kz]) - np.sum(dz[:kz] * d[:kz])
        assert budget == pytest.approx(dts * stf * 0.5, rel=1e-9, abs=1e-9 * np.sum(dz))


# -------------------------------------------------------------------- dens
def test_dens_matches_monomials():
    rng = np.random.default_rng(3)
    km = 7
    C = rng.normal(size=(9, km))
    for _ in range(50):
        k = int(rng.integers(0, km))
        t, s = rng.normal(0, 5), rng.normal(0, 2e-3)
        c = C[:, k]
        mono = (c[0] * t + c[1] * s + c[2] * t * t + c[3] * s * t + c[4] * s * s
                + c[5] * t ** 3 + c[6] * s * s * t + c[7] * s * t * t + c[8] * s ** 3)
        assert orc.dens(C, k, t, s) == pytest.approx(mono, rel=1e-12, abs=1e-15)


# ------------------------------------------------------- masks and coefficients
def test_masks_and_coefficients_closed_forms():
    c = ls.make_case("C1", nx=40, ny=30, km=12, depth_min=-3000.0, sigma=1.0, seed=4)
    o = orc.Oracle(c)
    kmt, d = o.kmt, o.d
    km = c.km
    # halo: cyclic columns, closed (zero) rows
    assert np.array_equal(kmt[2:-2, 0], kmt[2:-2, -4]) and np.array_equal(kmt[2:-2, -1], kmt[2:-2, 3])
    assert not kmt[:2].any() and not kmt[-2:].any()
    # vit(k) = [k < kmt]
    assert np.array_equal(d["vit"], (np.arange(km)[:, None, None] < kmt[None]).astype(float))
    # kmu = min of the four T columns around the U point (SW corner of T(i,j))
    jj, ii = np.mgrid[2:c.jmt - 2, 2:c.imt - 2]
    kmu = np.minimum(np.minimum(kmt[jj, ii - 1], kmt[jj, ii]), np.minimum(kmt[jj + 1, ii - 1], kmt[jj + 1, ii]))
    assert np.array_equal(d["kmu"][2:-2, 2:-2], kmu)
    # vertical grid
    dz = -np.diff(c.grid["zkp"])
    np.testing.assert_allclose(d["dzp"], dz, rtol=0, atol=0)
    assert d["odzt"][0] == 2.0 * d["odzp"][0]
    zt = 0.5 * (c.grid["zkp"][:-1] + c.grid["zkp"][1:])
    np.testing.assert_allclose(d["odzt"][1:], 1.0 / (zt[:-1] - zt[1:]), rtol=1e-15)
    # inverse column depths
    depth = (d["vit"] * dz[:, None, None]).sum(0)
    ohbt = np.where(depth > 0, 1.0 / np.where(depth > 0, depth, 1.0), 0.0)
    np.testing.assert_allclose(d["ohbt"], ohbt, rtol=1e-14)
    # del2 coefficients are a conservative flux form: the east coefficient of
    # a cell times its area is the west coefficient of its neighbour times
    # that neighbour's area, and likewise north/south
    ta = c.grid["tarea"]
    np.testing.assert_allclose(d["dte"][:, :-1] * ta[:, :-1], d["dtw"][:, 1:] * ta[:, 1:], rtol=1e-14)
    np.testing.assert_allclose(d["dts"][:-1] * ta[:-1], d["dtn"][1:] * ta[1:], rtol=1e-14)
    assert not d["dte"][:, -1].any() and not d["dtn"][0].any()     # eoshift end-off


def test_del4_ahf_closed_form():
    c = ls.make_case("C1", nx=20, ny=10, km=5, hmix=4, ah=-1e10)
    o = orc.Oracle(c)
    dxg = 2 * 6371000.0 * math.pi / c.imt_global
    np.testing.assert_allclose(o.d["ahf"], (c.grid["tarea"] / dxg ** 2) ** 1.5, rtol=1e-14)


# ------------------------------------------------------------------- del2/4
def test_del2_is_the_textbook_laplacian_on_a_uniform_grid():
    dx, dy = 50e3, 40e3
    c = case_uniform(dx=dx, dy=dy)
    o = orc.Oracle(c)
    rng = np.random.default_rng(0)
    X = orc.halo_fill(rng.normal(size=(c.jmt, c.imt)))
    out = o.del2_level(X, 0, use_ahf=False, ah=1.0)
    # lap[j-1, i-1] is the Laplacian at X[j, i]
    lap = ((X[1:-1, 2:] + X[1:-1, :-2] - 2 * X[1:-1, 1:-1]) / dx ** 2
           + (X[2:, 1:-1] + X[:-2, 1:-1] - 2 * X[1:-1, 1:-1]) / dy ** 2)
    inner = out[2:-2, 2:-2]
    ref = o.d["ahf"][5, 5] * lap[1:-1, 1:-1]      # the coefficients carry AHF
    # rows away from the walls (a wall row loses its outer neighbour)
    np.testing.assert_allclose(inner[1:-1], ref[1:-1], rtol=1e-10, atol=1e-10 * np.abs(ref).max())
    # the southern wall row: no flux through the closed boundary
    row = c.jmt - 3
    wall = o.d["ahf"][5, 5] * ((X[row, 3:-1] + X[row, 1:-3] - 2 * X[row, 2:-2]) / dx ** 2
                               + (X[row - 1, 2:-2] - X[row, 2:-2]) / dy ** 2)
    np.testing.assert_allclose(out[row, 2:-2], wall, rtol=1e-10, atol=1e-10 * np.abs(ref).max())


def test_del2_constant_field_and_zero_net_flux():
    c = ls.make_case("C1", nx=40, ny=30, km=8, depth_min=-3000.0, sigma=1.0, seed=8)
    o = orc.Oracle(c)
    X = np.full((c.jmt, c.imt), 7.0)
    for k in range(c.km):
        out = o.del2_level(X, k, ah=1.0)
        assert np.max(np.abs(out)) <= 1e-13 * 7.0 * np.max(np.abs(o.d["dtn"]) + np.abs(o.d["dte"])) * 4
    rng = np.random.default_rng(1)
    Y = orc.halo_fill(rng.normal(size=(c.jmt, c.imt)))
    out = o.del2_level(Y, 3, ah=1.0)
    ta = c.grid["tarea"][2:-2, 2:-2]
    tot = np.sum(ta * out[2:-2, 2:-2])
    scale = np.sum(np.abs(ta * out[2:-2, 2:-2]))
    assert abs(tot) <= 1e-12 * scale


# ------------------------------------------------------------ upwell / div
def _faces(c):
    """tspas face velocities (advection_tracer_1_0) of the step's inputs."""
    s0, g, oncc = c.state, c.grid, c.params["oncc"]
    wkd = 0.5 * (s0["utf"] * oncc) + 0.5 * s0["utl"]
    wkb = 0.5 * (s0["vtf"] * oncc) + 0.5 * s0["vtl"]
    uw = np.zeros_like(wkd)
    vs = np.zeros_like(wkb)
    uw[:, 1:, :] = (wkd[:, :-1, :] + wkd[:, 1:, :]) * g["htw"][1:, :] * 0.25
    vs[:, :, :-1] = (wkb[:, :, :-1] + wkb[:, :, 1:]) * g["hts"][:, :-1] * 0.25
    return uw, vs


@pytest.mark.parametrize("net_transport", [False, True])
def test_w_closes_discrete_continuity_with_the_advective_faces(net_transport):
    """The w of upwell (and operators_div) is the one that makes the flux
    divergence of tspas's own face velocities vanish in 3-D, with the
    barotropic correction work*ohbt taking out any net column transport so
    that w reaches zero at the sea floor."""
    c = ls.make_case("C1", nx=32, ny=24, km=10, seed=3, depth_min=-2000.0, sigma=1.5)
    if net_transport:
        rng = np.random.default_rng(4)
        for f in ("utf", "utl", "vtf", "vtl"):
            c.state[f] += rng.normal(0.0, 0.05, c.state[f].shape)
    o = orc.Oracle(c)
    st = c.copy_state()
    o.step(st, convadj=0)
    ws = st["ws"]
    uw, vs = _faces(c)
    div = np.zeros_like(uw)
    div[:, :, :-1] = uw[:, :, 1:] - uw[:, :, :-1]
    div[:, 1:, :] += vs[:, 1:, :] - vs[:, :-1, :]
    tr = c.grid["tarea"].copy()
    tr[tr > 0] = 1.0 / tr[tr > 0]
    div *= 2.0 * tr
    dz, vit, kmt = o.d["dzp"], o.d["vit"], o.kmt
    work = -(dz[:, None, None] * div * vit).sum(0)
    corr = work * o.d["ohbt"]
    if not net_transport:
        assert np.abs(work[2:-2, 2:-2]).max() <= 1e-12 * np.abs(div).max() * dz.sum()
    scale = np.abs(ws).max()
    assert scale > 0
    for k in range(c.km - 1):
        wet = (vit[k + 1] > 0)[2:-2, 2:-2]          # interface k+1 inside the column
        got = (ws[k + 1] - ws[k])[2:-2, 2:-2][wet]
        ref = (dz[k] * (corr + div[k]))[2:-2, 2:-2][wet]
        if k == 0:
            got = ws[1][2:-2, 2:-2][wet]              # ws(0) is not part of the recursion
        assert np.abs(got - ref).max() <= 1e-11 * scale
    # sea floor: the unmasked continuation of the recursion below the last
    # wet level is zero
    jj, ii = np.nonzero(kmt[2:-2, 2:-2] > 1)
    jj, ii = jj + 2, ii + 2
    kb = kmt[jj, ii] - 1
    floor = ws[kb, jj, ii] + dz[kb] * (corr[jj, ii] + div[kb, jj, ii])
    assert np.abs(floor).max() <= 1e-11 * scale


# -------------------------------------------------------------------- tspas
def _adv_inputs(c, u=0.0, v=0.0, w=0.0):
    km, jmt, imt = c.km, c.jmt, c.imt
    dy, dx = c.grid["htw"][3, 3], c.grid["hts"][3, 3]
    uw = np.full((km, jmt, imt), u * dy / 2.0)
    vs = np.full((km, jmt, imt), v * dx / 2.0)
    ws = np.full((km + 1, jmt, imt), w)
    return uw, vs, ws


def test_tspas_constant_field_is_preserved():
    c = case_uniform()
    o = orc.Oracle(c)
    uw, vs, ws = _adv_inputs(c, u=0.3, v=-0.2, w=0.0)
    T = np.full((c.km, c.jmt, c.imt), 12.5)
    adv = o.advect(T, uw, vs, ws, 3600.0)
    assert not np.any(adv[:, 2:-2, 2:-2])


def test_tspas_smooth_zonal_field_is_lax_wendroff():
    dx, dts, u = 50e3, 3600.0, 0.5
    c = case_uniform(dx=dx, nx=32)
    o = orc.Oracle(c)
    uw, vs, ws = _adv_inputs(c, u=u)
    i = np.arange(c.imt) - 2
    T = np.broadcast_to(10.0 + 0.1 * np.sin(2 * np.pi * i / 32.0), (c.km, c.jmt, c.imt)).copy()
    adv = o.advect(T, uw, vs, ws, dts)
    Te, Tw, Tc = T[..., 3:-1], T[..., 1:-3], T[..., 2:-2]
    lw = -(Te - Tw) * u / (2 * dx) + dts * u * u * (Te - 2 * Tc + Tw) / (2 * dx * dx)
    np.testing.assert_allclose(adv[:, 2:-2, 2:-2], lw[:, 2:-2, :], rtol=1e-10, atol=1e-18)


def test_tspas_vertical_column_is_lax_wendroff():
    dz, dts, w = 100.0, 3600.0, 2e-3
    c = case_uniform(km=12, dz=dz)
    o = orc.Oracle(c)
    uw, vs, ws = _adv_inputs(c, w=w)
    z = np.arange(c.km)
    # monotone in depth: the limiter stays off away from extrema
    T = np.broadcast_to((5.0 + 2.0 * np.exp(-z / 6.0))[:, None, None], (c.km, c.jmt, c.imt)).copy()
    adv = o.advect(T, uw, vs, ws, dts)
    Tu, Td, Tc = T[:-2], T[2:], T[1:-1]
    lw = -w * (Tu - Td) / (2 * dz) + dts * w * w * (Tu - 2 * Tc + Td) / (2 * dz * dz)
    # levels 1..km-3: the bottom level's own one-sided terms feed level km-2
    np.testing.assert_allclose(adv[1:-2, 2:-2, 2:-2], lw[:-1, 2:-2, 2:-2], rtol=1e-9, atol=1e-18)


def test_tspas_step_is_shape_preserving_and_conservative():
    dx, dts, u = 50e3, 3600.0 * 6, 0.8      # CFL 0.35
    c = case_uniform(dx=dx, nx=40)
    o = orc.Oracle(c)
    uw, vs, ws = _adv_inputs(c, u=u)
    i = np.arange(c.imt) - 2
    prof = np.where((i % 40 >= 10) & (i % 40 < 20), 1.0, 0.0)
    T = np.broadcast_to(prof, (c.km, c.jmt, c.imt)).copy()
    adv = o.advect(T, uw, vs, ws, dts)
    new = (T + dts * adv)[:, 2:-2, 2:-2]
    assert new.max() <= 1.0 + 1e-12 and new.min() >= -1e-12        # no new extrema
    tv = lambda a: np.abs(np.diff(np.concatenate([a, a[:1]]))).sum()  # noqa: E731
    assert tv(new[0, 3]) <= tv(prof[2:-2]) + 1e-12
    # Lax-Wendroff alone overshoots this profile
    Te, Tw, Tc = T[..., 3:-1], T[..., 1:-3], T[..., 2:-2]
    lw = Tc + dts * (-(Te - Tw) * u / (2 * dx) + dts * u * u * (Te - 2 * Tc + Tw) / (2 * dx * dx))
    assert lw.max() > 1.0 + 1e-3
    # the periodic row sum of the flux divergence telescopes to zero
    assert abs(adv[0, 4, 2:-2].sum()) <= 1e-12 * np.abs(adv[0, 4, 2:-2]).sum()


# ------------------------------------------------------------------ convadj
def _column_case(T_col, S_col, dz=50.0, linear=True):
    km = len(T_col)
    c = ls.make_case("C1", nx=4, ny=2, km=km, depth_min=6000.0, depth_max=6000.0, dz_top=dz,
                     dz_bot=dz, depth_total=dz * km)
    c.grid["zkp"][:] = -dz * np.arange(km + 1)
    c.grid["kmt"][:] = 0
    c.grid["kmt"][2:-2, 2:-2] = km
    if linear:   # rho = -alpha t + beta s  (C1, C2 only)
        c.grid["eosc"][:] = 0.0
        c.grid["eosc"][0] = -0.2
        c.grid["eosc"][1] = 780.0
        c.grid["to"][:] = 0.0
        c.grid["so"][:] = 0.0
    st = c.state
    st["at"][:] = 0.0
    st["at"][0, :, 2:-2, 2:-2] = np.asarray(T_col)[:, None, None]
    st["at"][1, :, 2:-2, 2:-2] = np.asarray(S_col)[:, None, None]
    for n in range(2):
        st["atb"][n, 1:] = st["at"][n]
    return c


def _convadj(c):
    o = orc.Oracle(c)
    st = c.copy_state()
    d = o.convadj(st)
    return st, d


def test_convadj_stable_column_unchanged():
    T = np.linspace(20.0, 2.0, 8)
    c = _column_case(T, np.zeros(8))
    st, d = _convadj(c)
    assert np.array_equal(st["at"], c.state["at"])
    assert not np.any(d["dt_conv"])


def test_convadj_two_layer_inversion_mixes_to_the_mean():
    T = np.linspace(20.0, 2.0, 8)
    T[3], T[4] = 5.0, 15.0                    # cold (dense) over warm
    dz = 50.0
    c = _column_case(T, np.zeros(8), dz=dz)
    st, d = _convadj(c)
    col = st["at"][0, :, 3, 3]
    assert col[3] == col[4] == pytest.approx(10.0, rel=1e-15)
    np.testing.assert_array_equal(col[[0, 1, 2, 5, 6, 7]], T[[0, 1, 2, 5, 6, 7]])
    assert d["dt_conv"][0, 3, 3, 3] == pytest.approx((10.0 - 5.0) / c.params["dts"])


def test_convadj_unstable_column_mixes_completely_and_conserves():
    rng = np.random.default_rng(5)
    T = np.linspace(2.0, 20.0, 10)            # warm at depth everywhere
    S = rng.normal(0, 1e-4, 10)
    c = _column_case(T, S, dz=40.0)
    st, _ = _convadj(c)
    for n, ref in ((0, T), (1, S)):
        col = st["at"][n, :, 3, 3]
        assert np.ptp(col) <= 1e-12 * max(1.0, np.abs(col).max())
        assert col.sum() == pytest.approx(ref.sum(), rel=1e-13, abs=1e-16)


def test_convadj_conserves_column_content_on_random_columns():
    c = ls.make_case("C1", nx=24, ny=12, km=20, seed=9, t_noise=2.0, s_noise=0.5)
    o = orc.Oracle(c)
    st = c.copy_state()
    o.convadj(st)
    dz = o.d["dzp"][:, None, None]
    for n in range(2):
        before = (dz * c.state["at"][n]).sum(0)
        after = (dz * st["at"][n]).sum(0)
        np.testing.assert_allclose(after, before, rtol=1e-12, atol=1e-15)
    assert not np.array_equal(st["at"], c.state["at"])     # something mixed


# ------------------------------------------------------------------ the step
def test_step_conserves_content_without_forcing():
    c = ls.make_case("C1", depth_min=-2000.0, sigma=2.0, seed=12)
    o = orc.Oracle(c)
    st = c.copy_state()
    before = orc.content(c, o, st)
    for _ in range(3):
        o.step(st)
    after = orc.content(c, o, st)
    for b, a in zip(before, after):
        assert abs(a - b) <= 1e-13 * abs(b)


def test_uniform_heat_flux_changes_content_by_the_flux():
    c = ls.make_case("C1", nx=32, ny=24, km=12, seed=2)
    o = orc.Oracle(c)
    q = 200.0                                   # W/m^2 of non-solar heat
    c.forcing["nswv"][:] = q
    st = c.copy_state()
    before = orc.content(c, o, st)[0]
    o.step(st, convadj=0)
    after = orc.content(c, o, st)[0]
    stf = q * c.params["od0cp"]
    area = c.grid["tarea"][2:-2, 2:-2][o.kmt[2:-2, 2:-2] > 0].sum()
    assert after - before == pytest.approx(area * stf * c.params["dts"], rel=1e-9)


def test_sssnorm_removes_the_net_freshwater_flux():
    c = ls.make_case("C1", nx=32, ny=24, km=12, seed=6)
    o = orc.Oracle(c)
    rng = np.random.default_rng(6)
    c.forcing["fresh"][:] = rng.uniform(0.0, 2e-5, c.forcing["fresh"].shape)   # net input
    st = c.copy_state()
    before = orc.content(c, o, st)[1]
    d = o.step(st, convadj=0)
    after = orc.content(c, o, st)[1]
    assert d["fw_norm2"] < 0
    assert abs(after - before) <= 1e-11 * abs(before)
    # without the normalisation the same forcing changes the S content
    st2 = c.copy_state()
    o.step(st2, convadj=0, sssnorm=0)
    assert abs(orc.content(c, o, st2)[1] - before) > 1e3 * abs(after - before)


# ------------------------------------------------------------------ icesnow
# Salt rejected per kelvin of surface freezing, in model units ((S-35)/1000):
# (34.7 - 4.0) psu * CP / LATICE / 1000, constants of constant_mod.F90:164 and
# shr_const_mod.h:51, evaluated here in another order.
ICE_SALT_PER_K = 30.7 / 3.337e5 * 3996.0 * 1.0e-3


def _ice_case(seed=3):
    from helpers import plant_cold
    c = ls.make_case("C1", nx=24, ny=12, km=8, seed=seed, depth_min=-1500.0, sigma=2.0)
    st = plant_cold(c, c.copy_state(), seed)
    return c, orc.Oracle(c), st


def test_icesnow_clamps_and_freezes_wet_surface_cells():
    from helpers import TBICE
    c, o, st = _ice_case()
    T0, S0 = st["at"][0].copy(), st["at"][1].copy()
    qice = o.icesnow(st, np.full((c.jmt, c.imt), 7.0), reset=True)
    T, S = st["at"]
    kmt = o.d["kmt"] if "kmt" in o.d else c.grid["kmt"]
    wet = kmt > 0
    froze = (T0[0] < TBICE) & wet
    assert froze.sum() > 10 and ((T0[0] < TBICE) & ~wet).sum() > 10
    # every level: T = max(T0, TBICE); S below the surface untouched
    np.testing.assert_array_equal(T, np.maximum(T0, TBICE))
    np.testing.assert_array_equal(S[1:], S0[1:])
    # frozen wet cells: ice = the cooling below TBICE (x*dzp/dzp: 1 ulp)
    td = TBICE - T0[0][froze]
    np.testing.assert_allclose(qice[froze], td, rtol=2.3e-16, atol=0)
    np.testing.assert_allclose(S[0][froze] - S0[0][froze], td * ICE_SALT_PER_K, rtol=1e-9, atol=0)
    # no ice and no salt anywhere else (land cells only clamp)
    assert not np.any(qice[~froze])
    np.testing.assert_array_equal(S[0][~froze], S0[0][~froze])


def test_icesnow_melts_stored_ice_against_a_warm_surface():
    from helpers import TBICE
    c, o, st = _ice_case(4)
    st["at"][0] = np.maximum(st["at"][0], TBICE + 0.5)     # nothing freezes
    st["at"][0][0, 5, 6], st["at"][0][0, 7, 9] = -1.0, -1.0
    S0 = st["at"][1][0].copy()
    q = np.zeros((c.jmt, c.imt))
    q[5, 6], q[7, 9] = 0.5, 2.0
    o.icesnow(st, q, reset=False)
    T = st["at"][0][0]
    # melt min(T - TBICE, ice): 0.5 K of the 0.8 K, then all 0.8 K of 2.0
    assert T[5, 6] == pytest.approx(-1.5, abs=1e-15) and q[5, 6] == 0.0
    assert T[7, 9] == pytest.approx(TBICE, abs=1e-15) and q[7, 9] == pytest.approx(1.2, abs=1e-15)
    assert S0[5, 6] - st["at"][1][0][5, 6] == pytest.approx(0.5 * ICE_SALT_PER_K, rel=1e-9)
    assert S0[7, 9] - st["at"][1][0][7, 9] == pytest.approx(0.8 * ICE_SALT_PER_K, rel=1e-9)


def test_icesnow_is_idempotent_with_reset():
    c, o, st = _ice_case(5)
    q = o.icesnow(st, np.zeros((c.jmt, c.imt)), reset=True)
    assert np.any(q)
    at1 = st["at"].copy()
    q2 = o.icesnow(st, q, reset=True)
    np.testing.assert_array_equal(st["at"], at1)
    assert not np.any(q2)
