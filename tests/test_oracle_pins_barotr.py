"""Pins of the oracle's barotr_ subcycle (oracle/barotr.c) against what the
mathematics fixes, on a Cartesian block with unit spacing and equal T<->U
weights:

* continuity: a uniform transport has no divergence; u = a x over a depth D
  gives the predictor work = -(1/4) a D (stage 0) and the corrector
  work = -a D (stage 2);
* Coriolis: with no other forcing the corrector turns (ub, vb) by
  (1 + i eps) / (1 - i eps), eps = f dtb / 2 -- a rotation that keeps the
  speed (Crank-Nicolson);
* the surface slope: h0 = H x gives the tendency (wgp - 1) 9.806 H;
* the free-surface diffusion of every fourth subcycle: h0p = x^2 gives
  work = 2 ah with del2 and nothing on the other subcycles;
* stage 3: h0p' = h0 = h0p + work dtb, ubp' = ub, h0f and h0bf accumulate.
"""
import numpy as np
import pytest

import orc

NX, NY = 14, 12
INT = (slice(2, NY - 2), slice(2, NX - 2))


def grid(del4=0, dtb=20.0, f=0.0, am=0.0, ah=0.0):
    s2 = (NY, NX)
    one = np.ones(s2)
    eps = 0.5 * f * dtb
    z = np.zeros(s2)
    kmt = np.full(s2, 5, np.int32)
    return dict(ib=3, ie=NX - 2, jb=3, je=NY - 2, del4=del4, dtb=dtb, am=am, ah=ah, kmt=kmt, kmu=kmt.copy(),
                kmtn=kmt.copy(), kmts=kmt.copy(), kmte=kmt.copy(), kmtw=kmt.copy(),
                au0=0.25 * one, aus=0.25 * one, auw=0.25 * one, ausw=0.25 * one, htw=one.copy(),
                hts=one.copy(), tarea_r=one.copy(), uarea_r=one.copy(), dxu=one.copy(), dyu=one.copy(),
                dxur=one.copy(), dyur=one.copy(), fcor=f * one, ebea=one / (1 + eps * eps),
                ebeb=eps * one / (1 + eps * eps), amf=one.copy(), ahf=one.copy(), dtn=one.copy(),
                dts=one.copy(), dte=one.copy(), dtw=one.copy(), duc=-4.0 * one, dum=z.copy(), dun=one.copy(),
                dus=one.copy(), due=one.copy(), duw=one.copy(), dmc=z.copy(), dmn=z.copy(), dms=z.copy(),
                dme=z.copy(), dmw=z.copy())


def fields(**kw):
    s2 = (NY, NX)
    f = {k: np.zeros(s2) for k in orc.BAROTR_IN + orc.BAROTR_INOUT if k != "wka"}
    f.update(vit=np.ones(s2), viv=np.ones(s2), wgp=np.ones(s2), wka=np.zeros((6, NY, NX)))
    f.update(kw)
    return f


X = np.arange(NX, dtype=np.float64)[None, :] * np.ones((NY, 1))


def run(g, f, nc, stages):
    return orc.barotr_stage(NX, NY, g, f, nc, stages)


def test_continuity_predictor_and_corrector():
    g = grid()
    D = 50.0
    o = run(g, fields(ub=np.full((NY, NX), 0.3), vb=np.full((NY, NX), -0.2), dzph=np.full((NY, NX), D)), 0, [0])
    assert np.all(o["work"][INT] == 0.0)
    a = 0.01
    o = run(g, fields(ub=a * X, dzph=np.full((NY, NX), D)), 0, [0])
    assert np.allclose(o["work"][INT], -0.25 * a * D, rtol=1e-13)
    # corrector: ub from ubp + wka2 dtb, the transport over dzph + (h0 at U)
    wka = np.zeros((6, NY, NX))
    o = run(g, fields(ubp=a * X, dzph=np.full((NY, NX), D), wka=wka), 0, [2])
    assert np.allclose(o["work"][2:, :-2], -a * D, rtol=1e-13)   # rows 0 and column imt-1 carry no transport


def test_coriolis_is_a_crank_nicolson_rotation():
    f0, dtb = 1.0e-4, 600.0
    g = grid(f=f0, dtb=dtb)
    U, V = 0.3, -0.1
    s2 = (NY, NX)
    o1 = run(g, fields(ubp=np.full(s2, U), vbp=np.full(s2, V)), 0, [1])
    o2 = run(g, fields(ubp=np.full(s2, U), vbp=np.full(s2, V), wka=o1["wka"]), 0, [2])
    eps = 0.5 * f0 * dtb
    want = (U + 1j * V) * (1 + 1j * eps) / (1 - 1j * eps)
    got = o2["ub"] + 1j * o2["vb"]
    assert np.allclose(got[INT], want, rtol=1e-13)
    assert np.allclose(np.abs(got[INT]), np.hypot(U, V), rtol=1e-13)


def test_surface_slope():
    g = grid()
    H = 0.02
    s2 = (NY, NX)
    wgp = np.full(s2, 0.95)
    o = run(g, fields(h0p=H * X, wgp=wgp), 0, [1])
    assert np.allclose(o["h0"], H * X)
    assert np.allclose(o["wka"][0][INT], (0.95 - 1.0) * 9.806 * H, rtol=1e-12)
    assert np.allclose(o["wka"][1][INT], 0.0, atol=1e-17)


@pytest.mark.parametrize("nc", [0, 3, 7])
def test_free_surface_diffusion_every_fourth_subcycle(nc):
    ah = 0.7
    g = grid(ah=ah)
    o = run(g, fields(h0p=X * X), nc, [2])
    reg = (slice(2, NY - 2), slice(2, NX - 2))
    want = 2.0 * ah if (nc + 1) % 4 == 0 else 0.0
    assert np.allclose(o["work"][reg], want, rtol=1e-13, atol=0)


def test_stage3_updates():
    g = grid(dtb=20.0)
    rng = np.random.default_rng(1)
    s2 = (NY, NX)
    f = fields(h0p=rng.normal(size=s2), work=rng.normal(size=s2), ub=rng.normal(size=s2), vb=rng.normal(size=s2),
               h0f=rng.normal(size=s2), h0bf=rng.normal(size=s2))
    o = run(g, f, 0, [3])
    h0 = f["h0p"] + f["work"] * 20.0
    assert np.array_equal(o["h0"], h0) and np.array_equal(o["h0p"], h0)
    assert np.array_equal(o["ubp"], f["ub"]) and np.array_equal(o["vbp"], f["vb"])
    assert np.array_equal(o["h0f"], f["h0f"] + h0) and np.array_equal(o["h0bf"], f["h0bf"] + h0)
