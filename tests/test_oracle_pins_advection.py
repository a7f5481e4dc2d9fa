"""Pins of the oracle's centered and flux advection (advection_tracer_1,
tracer.cpp:120-289, and the faces of advection_tracer_1_0, tracer.cpp:83-119)
against facts that do not come from its code:

* faces      -> impulse responses: a unit velocity at one U point reaches
                exactly the two faces that have it as a corner, scaled by
                hts/htw (centered, tspas) or by its own U-cell width dxu/dyu
                (flux)
* centered   -> zero on a constant field; on a linear field the discrete
                operator collapses to -(b (Uw+Ue) + c (Vs+Vn)) tarea_r and
                -0.5 odzp d (Wk+Wb) (one face term at the top and bottom
                levels); ax/ay/az accumulate the x, y, z parts / nss
* flux       -> telescoping: with periodic i, closed walls and no flux
                through the surface and the floor, the area-volume weighted
                sum of the tendency vanishes
* both       -> the algebraic identity flux + centered =
                T (2 tarea_r (Ue - Uw + Vs - Vn) + odzp (Wk' - Wb'))
                (Wk' = 0 at the top level, Wb' = 0 at the bottom one), which a
                sign, index or factor slip in either form breaks
* step       -> a flux-form step with zero forcing conserves T and S content
"""
import numpy as np
import pytest

import licom_synth as ls
import orc

CENTERED, FLUX = ls.ADV_CENTERED, ls.ADV_FLUX


def _case(nx=20, ny=14, km=7, seed=5, **kw):
    return ls.make_case("C1", nx=nx, ny=ny, km=km, seed=seed, depth_min=-2000.0, sigma=1.5, **kw)


def _random_inputs(c, seed):
    rng = np.random.default_rng(seed)
    km, jmt, imt = c.km, c.jmt, c.imt
    T = rng.normal(10.0, 3.0, (km, jmt, imt))
    uw = rng.normal(0.0, 1e4, (km, jmt, imt))
    vs = rng.normal(0.0, 1e4, (km, jmt, imt))
    ws = rng.normal(0.0, 1e-3, (km + 1, jmt, imt))
    return T, uw, vs, ws


# -------------------------------------------------------------------- faces
@pytest.mark.parametrize("adv", [CENTERED, FLUX, ls.ADV_TSPAS])
def test_face_impulse_response(adv):
    c = _case()
    o = orc.Oracle(c)
    g = c.grid
    km, jmt, imt = c.km, c.jmt, c.imt
    k0, j0, i0 = 3, 6, 9
    wkd = np.zeros((km, jmt, imt))
    wkb = np.zeros((km, jmt, imt))
    wkd[k0, j0, i0] = 1.0
    uw, vs = o.faces(adv, wkd, wkb)
    assert not vs.any()
    nz = sorted(zip(*np.nonzero(uw)))
    # u at U(i0,j0) is a corner of the west faces of T(i0,j0) and T(i0,j0+1)
    assert nz == [(k0, j0, i0), (k0, j0 + 1, i0)]
    if adv == FLUX:
        assert uw[k0, j0, i0] == g["dyu"][j0, i0] / 4
        assert uw[k0, j0 + 1, i0] == g["dyu"][j0, i0] / 4
    else:
        assert uw[k0, j0, i0] == g["htw"][j0, i0] / 4
        assert uw[k0, j0 + 1, i0] == g["htw"][j0 + 1, i0] / 4
    wkd[:] = 0.0
    wkb[k0, j0, i0] = 1.0
    uw, vs = o.faces(adv, wkd, wkb)
    assert not uw.any()
    nz = sorted(zip(*np.nonzero(vs)))
    # v at U(i0,j0) is a corner of the south faces of T(i0-1,j0) and T(i0,j0)
    assert nz == [(k0, j0, i0 - 1), (k0, j0, i0)]
    if adv == FLUX:
        assert vs[k0, j0, i0 - 1] == g["dxu"][j0, i0] / 4
        assert vs[k0, j0, i0] == g["dxu"][j0, i0] / 4
    else:
        assert vs[k0, j0, i0 - 1] == g["hts"][j0, i0 - 1] / 4
        assert vs[k0, j0, i0] == g["hts"][j0, i0] / 4


# ----------------------------------------------------------------- centered
def test_centered_zero_on_a_constant_field():
    c = _case()
    o = orc.Oracle(c)
    _, uw, vs, ws = _random_inputs(c, 1)
    T = np.full_like(uw, 7.25)
    ax, ay, az = (np.zeros_like(T) for _ in range(3))
    adv = o.advect_cf(CENTERED, T, uw, vs, ws, nss=3, ax=ax, ay=ay, az=az)
    assert not adv.any() and not ax.any() and not ay.any() and not az.any()


def test_centered_linear_field_closed_form_and_diagnostics():
    c = _case()
    o = orc.Oracle(c)
    _, uw, vs, ws = _random_inputs(c, 2)
    km, jmt, imt = c.km, c.jmt, c.imt
    b, cc, d = 0.37, -0.21, 0.8         # dT/di, dT/dj, dT/dk
    k, j, i = np.meshgrid(np.arange(km), np.arange(jmt), np.arange(imt), indexing="ij")
    T = 5.0 + b * i + cc * j + d * k
    nss = 4
    ax, ay, az = (np.zeros_like(T) for _ in range(3))
    adv = o.advect_cf(CENTERED, T, uw, vs, ws, nss=nss, ax=ax, ay=ay, az=az)
    o.advect_cf(CENTERED, T, uw, vs, ws, nss=nss, ax=ax, ay=ay, az=az)   # accumulates
    tr = 1.0 / c.grid["tarea"]
    Uw, Ue = uw, np.roll(uw, -1, axis=2)
    Vs, Vn = vs, np.roll(vs, 1, axis=1)
    hx = -b * (Uw + Ue) * tr
    hy = -cc * (Vs + Vn) * tr
    Wk, Wb = ws[:-1], ws[1:].copy()
    Wk = Wk.copy()
    Wk[0] = 0.0                          # no face above the top level
    Wb[-1] = 0.0                         # none below the bottom one
    odzp = o.d["odzp"][:, None, None]
    vz = -0.5 * odzp * d * (Wk + Wb)
    ref = hx + hy - vz
    sl = (slice(None), slice(2, -2), slice(2, -2))
    scale = np.abs(ref[sl]).max()
    assert np.abs(adv[sl] - ref[sl]).max() <= 1e-12 * scale
    assert np.abs(ax[sl] - 2 * hx[sl] / nss).max() <= 1e-12 * np.abs(hx[sl]).max()
    assert np.abs(ay[sl] - 2 * hy[sl] / nss).max() <= 1e-12 * np.abs(hy[sl]).max()
    assert np.abs(az[sl] + 2 * vz[sl] / nss).max() <= 1e-12 * np.abs(vz[sl]).max()
    # only the physical cells are written
    out = np.ones_like(adv, bool)
    out[sl] = False
    assert not ax[out].any() and not adv[out].any()


# --------------------------------------------------------------------- flux
def test_flux_form_telescopes():
    c = _case()
    o = orc.Oracle(c)
    T, uw, vs, ws = _random_inputs(c, 3)
    T = orc.halo_fill(T)
    uw = orc.halo_fill(uw)                # periodic faces in i
    vs[:, 1, :] = 0.0                     # closed northern and southern walls
    vs[:, -3, :] = 0.0
    adv = o.advect_cf(FLUX, T, uw, vs, ws)
    w = (o.d["dzp"][:, None, None] * c.grid["tarea"][None])
    sl = (slice(None), slice(2, -2), slice(2, -2))
    terms = (w * adv)[sl]
    parts = np.abs(w[sl] * uw[sl] * T[sl] / c.grid["tarea"][2:-2, 2:-2]).sum()
    assert abs(terms.sum()) <= 1e-12 * parts
    # an open wall is seen
    vs[:, -3, :] = 1e4
    adv2 = o.advect_cf(FLUX, T, uw, vs, ws)
    assert abs((w * adv2)[sl].sum()) > 1e-6 * parts


@pytest.mark.parametrize("seed", [4, 5])
def test_flux_plus_centered_identity(seed):
    c = _case(seed=seed)
    o = orc.Oracle(c)
    T, uw, vs, ws = _random_inputs(c, seed)
    f = o.advect_cf(FLUX, T, uw, vs, ws)
    ce = o.advect_cf(CENTERED, T, uw, vs, ws)
    tr = 1.0 / c.grid["tarea"]
    Uw, Ue = uw, np.roll(uw, -1, axis=2)
    Vs, Vn = vs, np.roll(vs, 1, axis=1)
    Wk, Wb = ws[:-1].copy(), ws[1:].copy()
    Wk[0] = 0.0
    Wb[-1] = 0.0
    odzp = o.d["odzp"][:, None, None]
    ref = T * (2.0 * tr * (Ue - Uw + Vs - Vn) + odzp * (Wk - Wb))
    sl = (slice(None), slice(2, -2), slice(2, -2))
    scale = np.abs(f[sl]).max() + np.abs(ce[sl]).max()
    assert np.abs((f + ce)[sl] - ref[sl]).max() <= 1e-12 * scale


# --------------------------------------------------------------------- step
def _content(o, c, st):
    return orc.content(c, o, st)


def test_flux_step_conserves_content():
    c = _case(nx=32, ny=20, km=10, seed=8)
    c.params["adv"] = FLUX
    o = orc.Oracle(c)
    st = c.copy_state()
    before = _content(o, c, st)
    for _ in range(3):
        o.step(st, convadj=0)
    after = _content(o, c, st)
    for n in range(2):
        scale = abs(before[n]) if n == 0 else max(abs(before[n]), 1e-3 * abs(before[0]))
        assert abs(after[n] - before[n]) <= 1e-12 * scale, (n, before[n], after[n])


def test_centered_step_keeps_a_constant_field():
    """A uniform tracer stays uniform under the centered scheme: every
    advective difference vanishes, the masked del2 and the vertical
    diffusion of a constant vanish, and the implicit solve returns the
    constant to roundoff."""
    c = _case(nx=24, ny=16, km=8, seed=9)
    c.params["adv"] = CENTERED
    c.params["nss"] = 2
    c.params["sssnorm"] = 0
    o = orc.Oracle(c)
    st = c.copy_state()
    vit = o.d["vit"]
    for n, val in ((0, 12.5), (1, -0.7e-3)):
        st["at"][n] = val * vit
        st["atb"][n][1:] = val * vit
    o.step(st, convadj=0)
    wet = vit[:, 2:-2, 2:-2] > 0
    for n, val in ((0, 12.5), (1, -0.7e-3)):
        got = st["at"][n][:, 2:-2, 2:-2][wet]
        assert np.abs(got - val).max() <= 1e-14 * abs(val), n
