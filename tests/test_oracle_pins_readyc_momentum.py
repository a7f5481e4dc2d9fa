"""Pins of the oracle's readyc_ momentum tendencies (oracle/readyc.c,
orc_rm_*) against what the mathematics fixes, on a Cartesian block (unit
spacing unless stated, equal T<->U weights, the 5-point Laplacian as du*,
no metric or rotation terms unless stated):

* upwell (upwell_cu_0 + upwell_cu): a velocity of uniform divergence a gives
  ws(k) = (ws(0) + (1 - H ohbt) a (1 + h0 ohbu) z_k) / (1 + h0 ohbt), z_k the
  depth of interface k, in x and in y; zero below the bottom;
* advection (readyc3 + advection_momentum_cu): uniform flow has none; u = a x
  gives -u du/dx = -a^2 x, v = b y gives -b^2 y; a uniform w over u = g k
  gives -w du/dz with one-sided halves at the top and the bottom;
* del2 viscosity: am times the textbook 5-point Laplacian (u = x^2 -> 2 am),
  the rotation terms enter u with + and v with -, zero below kmu and outside
  the block's interior; wka = c0f |u| (a 3-4-5 triangle);
* del4 viscosity: u = x^4 / 12 has discrete biharmonic 2, so am_f = 1 gives
  2 am; the Smagorinsky-like factor am_f = sqrt(|grad div|^2 + |grad curl|^2)
  45 (1/uarea_r)^(5/2) / |am amf|, clamped to [1, 40], read off the jump of
  am_f across a change of cell area;
* readyc4: |bottom drag| = c0f |u_b|^2 turned by -snlat theta, taken at the
  deepest wet U level; the depth mean of a uniform tendency; land columns;
* readyc4_2: linear profiles have no interior viscous tendency, z^2 gives
  kappa; the surface stress and the bottom drag halves telescope in the
  column sum.
"""
import numpy as np
import pytest

import orc

NX, NY, KM = 20, 14, 6


def cart(nx=NX, ny=NY, km=KM, dz=2.0, del4=0, am=1.0):
    s2 = (ny, nx)
    one = np.ones(s2)
    q = 0.25 * one
    grid = dict(at0=q, atn=q, ate=q, atne=q, au0=q, aus=q, auw=q, ausw=q,
                odzt=np.full(km, 1.0 / dz))
    H = km * dz
    m = dict(ib=3, ie=nx - 2, jb=3, je=ny - 2, del4=del4, am=am, c0f=2.0e-3, od0=1.0e-3,
             cag=1.0, sag=0.0, kmt=np.full(s2, km, np.int32), kmu=np.full(s2, km, np.int32),
             ohbu=one / H, ohbt=one / H, hts=one.copy(), htw=one.copy(), hue=one.copy(),
             hun=one.copy(), tarea_r=one.copy(), uarea_r=one.copy(), dxu=one.copy(), dyu=one.copy(),
             dxur=one.copy(), dyur=one.copy(), snlat=one.copy(), amf=one.copy(),
             duc=-4.0 * one, dum=0.0 * one, dun=one.copy(), dus=one.copy(), due=one.copy(),
             duw=one.copy(), dmc=0.0 * one, dmn=0.0 * one, dms=0.0 * one, dme=0.0 * one,
             dmw=0.0 * one, dzp=np.full(km, dz), odzp=np.full(km, 1.0 / dz))
    return grid, m


def fields(nx=NX, ny=NY, km=KM, **kw):
    s2, s3 = (ny, nx), (km, ny, nx)
    f = dict(u=np.zeros(s3), v=np.zeros(s3), up=np.zeros(s3), vp=np.zeros(s3), h0=np.zeros(s2),
             su=np.zeros(s2), sv=np.zeros(s2), vit=np.ones(s3), viv=np.ones(s3), akmu=np.zeros(s3),
             ws=np.zeros(s3))
    f.update(kw)
    return f


def run(grid, m, f, pieces=None):
    km, ny, nx = f["u"].shape
    return orc.readyc_momentum(nx, ny, km, grid, m, f, pieces=pieces)


X = np.arange(NX, dtype=np.float64)[None, None, :] * np.ones((KM, NY, 1))
Y = np.arange(NY, dtype=np.float64)[None, :, None] * np.ones((KM, 1, NX))
INT = (slice(None), slice(1, NY - 1), slice(1, NX - 1))


def test_upwell_of_uniform_divergence():
    dz, H, a, c = 2.0, KM * 2.0, 3.0e-3, 4.0
    zk = dz * np.arange(KM)[:, None, None]           # depth of interface k
    for ohbt_h in (0.0, 0.5, 1.0):
        grid, m = cart(dz=dz)
        m["ohbt"] = np.full((NY, NX), ohbt_h / H)
        ws0 = np.zeros((KM, NY, NX))
        ws0[0] = 1.0e-4
        for vel in ("u", "v"):
            f = fields(h0=np.full((NY, NX), c), ws=ws0.copy(), **{vel: a * (X if vel == "u" else Y)})
            o = run(grid, m, f, ["upwell"])
            want = (1.0e-4 + (1.0 - ohbt_h) * a * (1.0 + c / H) * zk) / (1.0 + c * ohbt_h / H)
            want = np.broadcast_to(want, (KM, NY, NX))
            assert np.allclose(o["ws"][1:][INT[:1] + INT[1:]], want[1:][INT[:1] + INT[1:]],
                               rtol=1e-12, atol=1e-18), (ohbt_h, vel)
            # level 0 and the outer frame are left alone
            assert np.array_equal(o["ws"][0], ws0[0])
            assert np.all(o["ws"][1:, 0, :] == 0.0) and np.all(o["ws"][1:, :, -1] == 0.0)
    # below the bottom: vit = 0 -> ws = 0, and the divergence there is not counted
    grid, m = cart(dz=dz)
    m["ohbt"] = np.zeros((NY, NX))
    kmt = np.full((NY, NX), KM, np.int32)
    kmt[5, 7] = 3
    m["kmt"] = kmt
    vit = (np.arange(KM)[:, None, None] < kmt[None]).astype(np.float64)
    o = run(grid, m, fields(u=a * X, vit=vit), ["upwell"])
    assert np.all(o["ws"][3:, 5, 7] == 0.0)
    assert np.allclose(o["ws"][1:3, 5, 7], a * zk[1:3, 0, 0], rtol=1e-12)


def test_advection_closed_forms():
    grid, m = cart()
    ws = np.full((KM, NY, NX), 0.37)
    o = run(grid, m, fields(u=np.full((KM, NY, NX), 0.2), v=np.full((KM, NY, NX), -0.1), ws=ws),
            ["advect"])
    assert np.all(o["dlu"] == 0.0) and np.all(o["dlv"] == 0.0)
    a, b = 0.03, -0.02
    o = run(grid, m, fields(u=a * X, v=np.zeros((KM, NY, NX))), ["advect"])
    assert np.allclose(o["dlu"][INT], (-a * a * X)[INT], rtol=1e-13, atol=1e-18)
    assert np.all(o["dlv"] == 0.0)
    o = run(grid, m, fields(v=b * Y), ["advect"])
    assert np.allclose(o["dlv"][INT], (-b * b * Y)[INT], rtol=1e-13, atol=1e-18)
    # vertical: w uniform, u = g k -> -w du/dz, one-sided at top and bottom
    g, w0, dz = 0.05, 1.0e-3, 2.0
    K = np.arange(KM, dtype=np.float64)[:, None, None] * np.ones((1, NY, NX))
    o = run(grid, m, fields(u=g * K, ws=np.full((KM, NY, NX), w0)), ["advect"])
    want = np.full(KM, g * w0 / dz)
    want[0] = want[-1] = 0.5 * g * w0 / dz
    got = o["dlu"][:, 5, 7]
    assert np.allclose(got, want, rtol=1e-13), got
    # the frame outside 1 <= j < jmt-1, 1 <= i < imt-1 is cleared
    assert np.all(o["dlu"][:, 0] == 0.0) and np.all(o["dlu"][:, :, -1] == 0.0)


def test_del2_laplacian_and_rotation_terms():
    am = 3.0
    grid, m = cart(am=am)
    up = X * X
    vp = Y * Y
    o = run(grid, m, fields(up=up, vp=np.zeros_like(up)), ["hdiff"])
    # the reference's region: 0-based jb-1 <= j < je, ib-1 <= i < ie
    reg = (slice(None), slice(2, NY - 2), slice(2, NX - 2))
    assert np.allclose(o["dlu"][reg], 2.0 * am, rtol=1e-13)
    frame = np.ones((KM, NY, NX), bool)
    frame[reg] = False
    assert np.all(o["dlu"][frame] == 0.0)
    # rotation coupling: + dm on u's tendency, - dm on v's
    cm = 0.01
    m["dmc"] = np.full((NY, NX), cm)
    o = run(grid, m, fields(up=up, vp=vp), ["hdiff"])
    assert np.allclose(o["dlu"][reg], (am * (2.0 + cm * vp))[reg], rtol=1e-13)
    assert np.allclose(o["dlv"][reg], (am * (2.0 - cm * up))[reg], rtol=1e-13)
    # zero below kmu
    kmu = np.full((NY, NX), KM, np.int32)
    kmu[6, 8] = 2
    m["kmu"] = kmu
    o = run(grid, m, fields(up=up, vp=vp), ["hdiff"])
    assert np.all(o["dlu"][2:, 6, 8] == 0.0) and np.all(o["dlv"][2:, 6, 8] == 0.0)
    # wka = c0f |u|
    o = run(grid, m, fields(up=np.full((KM, NY, NX), 3.0), vp=np.full((KM, NY, NX), -4.0)), ["hdiff"])
    assert np.all(o["wka"] == 5.0 * m["c0f"])


def test_del4_biharmonic_and_factor_limits():
    am = -1.0e12                          # |am amf| huge: am_f = 1
    grid, m = cart(del4=1, am=am)
    up = X ** 4 / 12.0
    o = run(grid, m, fields(up=up, vp=np.zeros_like(up)), ["hdiff"])
    reg = (slice(None), slice(2, NY - 2), slice(2, NX - 2))
    assert np.allclose(o["dlu"][reg], 2.0 * am, rtol=1e-9)
    assert np.all(o["dlv"][reg] == 0.0)
    # am_f: with up = x^2, |grad div| = 2 and curl = 0; a cell area of 4
    # (t_area = 2) for i >= i0 multiplies am_f by 2^5 = 32
    i0 = 10
    # am_f = 2 * 45 * t_area^5 / |am|: 90 / |am| for i < i0, 2880 / |am| beyond
    for absam, lo, hi in ((80.0, 1.125, 36.0),        # both inside [1, 40]
                          (50.0, 1.8, 40.0),          # 57.6 -> 40
                          (200.0, 1.0, 14.4)):        # 0.45 -> 1
        grid, m = cart(del4=1, am=-absam)
        ua = np.ones((NY, NX))
        ua[:, i0:] = 0.25
        m["uarea_r"] = ua
        o = run(grid, m, fields(up=X * X, vp=np.zeros((KM, NY, NX))), ["hdiff"])
        # d2u = am_f * 2 on each side; am times its Laplacian across the jump
        jump = 2.0 * hi - 2.0 * lo
        d = o["dlu"][:, 5, :]
        assert np.allclose(d[:, i0 - 1], -absam * jump, rtol=1e-12), (absam, d[0, i0 - 1])
        assert np.allclose(d[:, i0], absam * jump, rtol=1e-12), (absam, d[0, i0])
        others = [i for i in range(2, NX - 2) if i not in (i0 - 1, i0)]
        assert np.allclose(d[:, others], 0.0, atol=1e-9 * absam)


def test_bottom_drag_and_depth_mean():
    th = np.deg2rad(25.0)
    grid, m = cart()
    m.update(cag=float(np.cos(th)), sag=float(np.sin(th)))
    sn = np.ones((NY, NX))
    sn[: NY // 2] = -1.0
    m["snlat"] = sn
    kmu = np.full((NY, NX), KM, np.int32)
    kmu[4, 6] = 3
    kmu[9, 9] = 0
    m["kmu"] = kmu
    rng = np.random.default_rng(1)
    up = rng.normal(0.0, 0.2, (KM, NY, NX))
    vp = rng.normal(0.0, 0.2, (KM, NY, NX))
    dlu = np.full((KM, NY, NX), 1.0e-6)
    viv = (np.arange(KM)[:, None, None] < kmu[None]).astype(np.float64)
    f = fields(up=up, vp=vp, su=np.full((NY, NX), 0.1), sv=np.full((NY, NX), -0.05), viv=viv,
               dlu=dlu, dlv=np.zeros_like(dlu))
    o = run(grid, m, f, ["stress"])
    for (j, i) in ((3, 5), (4, 6), (10, 12)):
        kb = kmu[j, i] - 1
        ub, vb = up[kb, j, i], vp[kb, j, i]
        bx, by = o["bbcx"][j, i], o["bbcy"][j, i]
        spd = np.hypot(ub, vb)
        assert np.hypot(bx, by) == pytest.approx(m["c0f"] * spd * spd, rel=1e-13)
        turn = np.angle(complex(bx, by) / complex(ub, vb))
        assert turn == pytest.approx(-sn[j, i] * th, abs=1e-12)
        # depth mean of a uniform 1e-6 over the wet levels plus the stresses
        H = 2.0 * KM
        wantx = 1.0e-6 * (2.0 * kmu[j, i]) / H + (0.1 * m["od0"] - bx) / H
        assert o["dlub"][j, i] == pytest.approx(wantx, rel=1e-12)
        wanty = (-0.05 * m["od0"] - by) / H
        assert o["dlvb"][j, i] == pytest.approx(wanty, rel=1e-12)
    # land column: no stress, no drag
    assert o["sbcx"][9, 9] == 0.0 and o["bbcx"][9, 9] == 0.0 and o["dlub"][9, 9] == 0.0
    # the frame keeps what it held
    assert np.all(o["sbcx"][0] == 0.0)


def test_vertical_viscosity_profiles_and_telescoping():
    dz, kap = 2.0, 0.3
    grid, m = cart(dz=dz)
    K = np.arange(KM, dtype=np.float64)[:, None, None] * np.ones((1, NY, NX))
    z = dz * K
    akmu = np.full((KM, NY, NX), kap)
    for prof, want in ((0.01 * z, 0.0), (z * z, kap)):
        o = run(grid, m, fields(up=prof, vp=-prof, akmu=akmu, wka=np.zeros((KM, NY, NX)),
                                dlu=np.zeros((KM, NY, NX)), dlv=np.zeros((KM, NY, NX))), ["vdiff"])
        inner = o["dlu"][1:KM - 1][(slice(None),) + INT[1:]]
        assert np.allclose(inner, want, rtol=1e-13, atol=1e-15)
        assert np.allclose(o["dlv"][1:KM - 1][(slice(None),) + INT[1:]], -want, rtol=1e-13, atol=1e-15)
    # column sums: dzp odzp = 1, so the interior terms telescope to
    # (surface stress - bottom drag) / 2, the drag at the deepest wet level
    th = np.deg2rad(20.0)
    m.update(cag=float(np.cos(th)), sag=float(np.sin(th)))
    kmu = np.full((NY, NX), KM, np.int32)
    kmu[5, 5] = 4
    viv = (np.arange(KM)[:, None, None] < kmu[None]).astype(np.float64)
    rng = np.random.default_rng(3)
    up = rng.normal(0.0, 0.2, (KM, NY, NX)) * viv
    vp = rng.normal(0.0, 0.2, (KM, NY, NX)) * viv
    wka = np.abs(rng.normal(0.0, 1.0e-3, (KM, NY, NX)))
    su = np.full((NY, NX), 0.08)
    o = run(grid, m, fields(up=up, vp=vp, akmu=rng.uniform(0, 1e-2, (KM, NY, NX)), viv=viv, wka=wka,
                            su=su, dlu=np.zeros((KM, NY, NX)), dlv=np.zeros((KM, NY, NX))), ["vdiff"])
    for (j, i) in ((5, 5), (7, 9)):
        kb = kmu[j, i] - 1
        drag = wka[kb, j, i] * (up[kb, j, i] * m["cag"] + vp[kb, j, i] * m["sag"])
        col = np.sum(dz * o["dlu"][: kb + 1, j, i])
        assert col == pytest.approx(0.5 * (su[j, i] * m["od0"] - drag), rel=1e-11, abs=1e-16)
