"""Pins of the oracle's bclinc_ (oracle/bclinc.c) against what the
mathematics fixes, on a Cartesian block with unit spacing:

* Coriolis, semi-implicit: the new tendency d = dlu + i dlv satisfies
  (1 - i eps) d = -(gx + i gy), gx = f vp - dlu_in, gy = -f up - dlv_in, eps
  = f dtc / 2 on the first step (epea, epeb) and f dtc after (epla, eplb);
* pressure gradient: a surface pressure P x gives -P od0 at every wet level;
  a free surface H x with a uniform gg = G gives
  -(9.806 H - G (1 + ohbt zkt(k)) H); nothing below kmu;
* the implicit vertical viscosity: the solution satisfies the implicit
  system (x - x0) = dt/2 odzp (akmu (x_{k-1} - x_k) odzt_k
  - akmu (x_k - x_{k+1}) odzt_{k+1}) with the surface stress above level 0
  and the bottom drag below the deepest wet level, to rounding (a one-level
  column with the reference's extra dt/2 odzp on its diagonal);
* the barotropic replacement: the depth mean of the new velocity is ub;
* the Robert-Asselin filter: up' = u + afc1 (up - 2 u + u') with afc2 =
  1 - 2 afc1; utf and vtf accumulate the new velocities.
"""
import numpy as np
import pytest

import orc

NX, NY, KM = 12, 10, 6
DZ = 2.0


def grid(f=0.0, dtc=100.0, km=KM):
    s2 = (NY, NX)
    one = np.ones(s2)
    H = km * DZ
    eps_e, eps_l = 0.5 * f * dtc, f * dtc
    g = dict(od0=1.0e-3, c0f=2.0e-3, cag=1.0, sag=0.0, onbb=1.0, dtc=dtc, dtc2=2.0 * dtc, afc1=0.43,
             afc2=1.0 - 2.0 * 0.43, kmu=np.full(s2, km, np.int32), snlat=one.copy(), ohbt=one / H,
             ohbu=one / H, fcor=f * one, dxur=one.copy(), dyur=one.copy(),
             epea=one / (1.0 + eps_e * eps_e), epeb=eps_e * one / (1.0 + eps_e * eps_e),
             epla=one / (1.0 + eps_l * eps_l), eplb=eps_l * one / (1.0 + eps_l * eps_l),
             dzp=np.full(km, DZ), odzt=np.full(km, 1.0 / DZ), odzp=np.full(km, 1.0 / DZ),
             zkt=-DZ * (np.arange(km) + 0.5))
    return g


def fields(km=KM, **kw):
    s2, s3 = (NY, NX), (km, NY, NX)
    f = dict(vit=np.ones(s3), viv=np.ones(s3), gg=np.zeros(s3), akmu=np.zeros(s3), su=np.zeros(s2),
             sv=np.zeros(s2), psa=np.zeros(s2), h0bl=np.zeros(s2), ub=np.zeros(s2), vb=np.zeros(s2),
             h0bf=np.zeros(s2), u=np.zeros(s3), v=np.zeros(s3), up=np.zeros(s3), vp=np.zeros(s3),
             dlu=np.zeros(s3), dlv=np.zeros(s3), wka=np.zeros(s3), utf=np.zeros(s3), vtf=np.zeros(s3),
             sbcx=np.zeros(s2), sbcy=np.zeros(s2), bbcx=np.zeros(s2), bbcy=np.zeros(s2))
    f.update(kw)
    return f


def run(g, f, isc, stages):
    return orc.bclinc_stage(NX, NY, g["dzp"].size, g, f, isc, stages)


X = np.arange(NX, dtype=np.float64)[None, None, :] * np.ones((KM, NY, 1))
GRAD = (slice(None), slice(0, NY - 1), slice(1, NX))      # where the gradient is formed


@pytest.mark.parametrize("isc", [0, 3])
def test_coriolis_is_crank_nicolson(isc):
    f0, dtc = 1.0e-4, 1200.0
    g = grid(f=f0, dtc=dtc)
    rng = np.random.default_rng(isc)
    up, vp = rng.normal(0, 0.1, (KM, NY, NX)), rng.normal(0, 0.1, (KM, NY, NX))
    dlu, dlv = rng.normal(0, 1e-6, (KM, NY, NX)), rng.normal(0, 1e-6, (KM, NY, NX))
    o = run(g, fields(up=up, vp=vp, dlu=dlu, dlv=dlv), isc, [0])
    eps = (0.5 if isc == 0 else 1.0) * f0 * dtc
    gx, gy = f0 * vp - dlu, -f0 * up - dlv
    d = o["dlu"] + 1j * o["dlv"]
    lhs = (1.0 - 1j * eps) * d
    assert np.allclose(lhs[GRAD], -(gx + 1j * gy)[GRAD], rtol=1e-12, atol=1e-20)


def test_pressure_gradients():
    g = grid()
    P = 3.0
    psa = P * X[0]
    o = run(g, fields(psa=psa), 0, [0])
    assert np.allclose(o["dlu"][GRAD], -P * g["od0"], rtol=1e-12)
    assert np.allclose(o["dlv"][GRAD], 0.0, atol=1e-15)
    # free surface H x and a uniform gg = G: the hydrostatic and the
    # free-surface terms; h0bl enters on the first step (isc = 0)
    H, G = 0.2, 3.0e-3
    o = run(g, fields(h0bl=H * X[0], gg=np.full((KM, NY, NX), G)), 0, [0])
    zkt = g["zkt"][:, None, None]
    want = -(9.806 * H - G * (1.0 + g["ohbt"][0, 0] * zkt) * H)
    assert np.allclose(o["dlu"][GRAD], np.broadcast_to(want, o["dlu"].shape)[GRAD], rtol=1e-12)
    # after the first step h0bf and h0bl are averaged (aa = 0.5), and h0bf is
    # scaled by onbb first
    g2 = dict(g, onbb=0.5)
    o = run(g2, fields(h0bl=H * X[0], h0bf=4.0 * H * X[0]), 2, [0])
    work = 0.5 * (0.5 * 4.0 * H) + 0.5 * H          # per unit x
    assert np.allclose(o["dlu"][GRAD], -9.806 * work, rtol=1e-12)
    assert np.allclose(o["h0bf"], 0.5 * 4.0 * H * X[0])
    # nothing below kmu: only the Coriolis and dlu terms remain (zero here)
    kmu = np.full((NY, NX), KM, np.int32)
    kmu[4, 5] = 2
    o = run(dict(g, kmu=kmu), fields(psa=psa), 0, [0])
    assert np.all(o["dlu"][2:, 4, 5] == 0.0)


def implicit_residual(x, x0, akmu, odzt, odzp, dt, sbc, bbc, kz):
    """residual of the implicit vertical viscosity system of one column; a
    one-level column carries the reference's extra dt/2 odzp on its diagonal
    (its bottom condition adds the top's stress coefficient a8(0),
    bclinc.cpp:228-230 with kmb = 0)"""
    if kz == 1:
        return np.array([x[0] * (1.0 + 0.5 * dt * odzp[0]) - x0[0] - 0.5 * dt * odzp[0] * (sbc - bbc)])
    r = np.zeros(kz)
    for k in range(kz):
        flux_up = sbc if k == 0 else akmu[k - 1] * odzt[k] * (x[k - 1] - x[k])
        flux_dn = bbc if k == kz - 1 else akmu[k] * odzt[k + 1] * (x[k] - x[k + 1])
        r[k] = (x[k] - x0[k]) - 0.5 * dt * odzp[k] * (flux_up - flux_dn)
    return r


@pytest.mark.parametrize("isc", [0, 3])
def test_implicit_vertical_viscosity(isc):
    g = grid(dtc=600.0)
    rng = np.random.default_rng(7)
    kmu = np.full((NY, NX), KM, np.int32)
    kmu[5, 6] = 3
    kmu[6, 7] = 1
    g["kmu"] = kmu
    viv = (np.arange(KM)[:, None, None] < kmu[None]).astype(np.float64)
    akmu = rng.uniform(1.0e-3, 5.0e-2, (KM, NY, NX))
    up, vp = rng.normal(0, 0.1, (KM, NY, NX)) * viv, rng.normal(0, 0.1, (KM, NY, NX)) * viv
    su, sv = rng.normal(0, 0.1, (NY, NX)), rng.normal(0, 0.1, (NY, NX))
    f = fields(viv=viv, akmu=akmu, up=up, vp=vp, su=su, sv=sv)
    o = run(g, f, isc, [0])
    # the explicit part: Coriolis and pressure are zero here, so the first
    # guess is the previous velocity; bclinc12's stresses bound the column
    dt = g["dtc"] if isc == 0 else g["dtc2"]
    for (j, i) in ((3, 4), (5, 6), (6, 7)):
        kz = kmu[j, i]
        if isc == 0:
            x, x0, sbc, bbc = o["u"][:, j, i], up[:, j, i], o["sbcx"][j, i], o["bbcx"][j, i]
        else:
            x, x0, sbc, bbc = o["wka"][:, j, i], vp[:, j, i], o["sbcy"][j, i], o["bbcy"][j, i]
        r = implicit_residual(x, x0, akmu[:, j, i], g["odzt"], g["odzp"], dt, sbc, bbc, kz)
        assert np.all(np.abs(r) <= 1e-13 * (np.abs(x0[:kz]).max() + 1e-3)), (j, i, r)
    # bottom drag from the deepest wet level of up, vp (bclinc12)
    kb = kmu[5, 6] - 1
    ub_, vb_ = up[kb, 5, 6], vp[kb, 5, 6]
    assert o["bbcx"][5, 6] == pytest.approx(g["c0f"] * np.hypot(ub_, vb_) * ub_, rel=1e-14)
    assert o["sbcy"][5, 6] == pytest.approx(sv[5, 6] * g["od0"], rel=1e-15)


def test_barotropic_replacement_and_time_filter():
    g = grid()
    rng = np.random.default_rng(9)
    s3 = (KM, NY, NX)
    ub, vb = rng.normal(0, 0.1, (NY, NX)), rng.normal(0, 0.1, (NY, NX))
    # first step, stage 1: the halo-updated u, v get the barotropic mean ub, vb
    u1, v1 = rng.normal(0, 0.1, s3), rng.normal(0, 0.1, s3)
    utf0 = rng.normal(0, 0.1, s3)
    o = run(g, fields(u=u1, v=v1, ub=ub, vb=vb, utf=utf0.copy()), 0, [1])
    H = KM * DZ
    assert np.allclose((DZ * o["u"]).sum(0) / H, ub, rtol=1e-12, atol=1e-15)
    assert np.allclose((DZ * o["v"]).sum(0) / H, vb, rtol=1e-12, atol=1e-15)
    assert np.allclose(o["u"] - o["u"].mean(0), u1 - u1.mean(0), rtol=1e-12, atol=1e-15)
    assert np.allclose(o["utf"], utf0 + o["u"], rtol=0, atol=0)
    # leapfrog, stages 1 and 2: v from wka (stage 1), u from wka (stage 2),
    # both with the barotropic replacement and the Robert-Asselin filter
    u, v, up, vp = (rng.normal(0, 0.1, s3) for _ in range(4))
    wka = rng.normal(0, 0.1, s3)
    f = fields(u=u.copy(), v=v.copy(), up=up.copy(), vp=vp.copy(), wka=wka.copy(), ub=ub, vb=vb)
    o1 = run(g, f, 3, [1])
    vnew = o1["v"]
    assert np.allclose((DZ * vnew).sum(0) / H, vb, rtol=1e-12, atol=1e-15)
    assert np.allclose(o1["vp"], v + g["afc1"] * (vp - 2.0 * v + vnew), rtol=1e-12, atol=1e-15)
    f2 = dict(f, **{k: o1[k] for k in orc.BCLINC_INOUT})
    f2["wka"] = wka.copy()                       # as if the halo exchange returned this
    o2 = run(g, f2, 3, [2])
    unew = o2["u"]
    assert np.allclose((DZ * unew).sum(0) / H, ub, rtol=1e-12, atol=1e-15)
    assert np.allclose(o2["up"], u + g["afc1"] * (up - 2.0 * u + unew), rtol=1e-12, atol=1e-15)
    assert np.array_equal(o2["utf"], f2["utf"] + unew)
    assert np.array_equal(o2["vtf"], f2["vtf"] + o1["v"])
