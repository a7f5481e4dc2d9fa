"""Pins of the oracle's readyc_ mixing preparation (oracle/readyc.c)
against what the mathematics fixes:

* readyc1: wp12 / wp13 are the wet-U-cell weighted means of up / vp around
  the T point (a dry neighbour's value does not enter, equal weights give the
  plain mean), zero on dry T cells and outside 1 <= j, i < imt-1;
* readyc1_2: for velocities linear in depth, u = a + b z and v = c + d z,
  the shear S^2 of every wet interface is b^2 + d^2, on the T grid (s2t) and
  on the U grid (s2u); rit = rict / S^2 and riu = ric / S^2, riu one level
  down (riu of level 0 is the clear's zero);
* readyc2: the T->U average of a uniform AKMT is AKMT times the sum of the
  weights on wet U interfaces, zero on the outer column and row and the
  bottom level;
* the clears and the cells each kernel leaves alone; the h0 time levels.
"""
import numpy as np
import pytest

import licom_synth as ls
import orc


def base(seed=3, nx=24, ny=12, km=8):
    c = ls.make_case("C1", nx=nx, ny=ny, km=km, seed=seed, depth_min=-2000.0, sigma=1.5)
    g = ls.readyt_grid(c)
    f = ls.readyt_fields(c, g)
    o = orc.readyt(c.imt, c.jmt, c.km, g, f)
    cg, cf = ls.readyc_inputs(c, g, o, seed=seed)
    return c, cg, cf


def shear(c, cg, cf):
    return orc.readyc_shear(c.imt, c.jmt, c.km, cg, cf)


def test_wp12_is_the_wet_weighted_mean():
    c, cg, cf = base()
    o = shear(c, cg, cf)
    up, viv, vit = cf["up"], cf["viv"], cf["vit"]
    w = [cg["at0"], cg["atn"], cg["ate"], cg["atne"]]
    sl = (slice(None), slice(1, None), slice(0, -1))
    nb = [up[:, 1:, :-1], up[:, :-1, :-1], up[:, 1:, 1:], up[:, :-1, 1:]]
    mb = [viv[:, 1:, :-1], viv[:, :-1, :-1], viv[:, 1:, 1:], viv[:, :-1, 1:]]
    ww = [x[1:, :-1][None] for x in w]
    num = sum(a * u * m for a, u, m in zip(ww, nb, mb))
    den = sum(a * m for a, m in zip(ww, mb)) + 1.0e-25
    ref = vit[sl] * num / den
    assert np.allclose(o["wp12"][sl], ref, rtol=1e-13, atol=1e-17)
    assert np.all(o["wp12"][:, 0, :] == 0.0) and np.all(o["wp12"][:, :, -1] == 0.0)
    assert np.all(o["wp12"][vit == 0.0] == 0.0)
    # a dry U neighbour's value does not enter
    dry = (viv == 0.0)
    cf2 = dict(cf)
    cf2["up"] = np.where(dry, 1.0e6, cf["up"])
    o2 = shear(c, cg, cf2)
    assert np.array_equal(o2["wp12"], o["wp12"])
    # equal weights, all four wet: the plain mean
    cg3 = dict(cg)
    for k in ("at0", "atn", "ate", "atne"):
        cg3[k] = np.full_like(cg[k], 0.25)
    o3 = shear(c, cg3, cf)
    k, j, i = 0, 6, 9
    assert viv[k, j, i] and viv[k, j - 1, i] and viv[k, j, i + 1] and viv[k, j - 1, i + 1]
    mean = (up[k, j, i] + up[k, j - 1, i] + up[k, j, i + 1] + up[k, j - 1, i + 1]) / 4.0
    assert o3["wp12"][k, j, i] == pytest.approx(mean, rel=1e-14)


def test_shear_of_linear_profiles():
    c, cg, cf = base(seed=4)
    zkt = 0.5 * (c.grid["zkp"][:-1] + c.grid["zkp"][1:])
    b, d = 2.0e-4, -3.0e-4
    prof_u = (0.1 + b * zkt)[:, None, None]
    prof_v = (-0.05 + d * zkt)[:, None, None]
    cf["up"] = np.broadcast_to(prof_u, cf["up"].shape) * cf["viv"]
    cf["vp"] = np.broadcast_to(prof_v, cf["vp"].shape) * cf["viv"]
    # odzt consistent with zkt (readyt_grid's construction)
    assert np.allclose(cg["odzt"][1:], 1.0 / (zkt[:-1] - zkt[1:]))
    o = shear(c, cg, cf)
    km = c.km
    s2 = b * b + d * d
    vit, viv = cf["vit"], cf["viv"]
    # U grid: every interface wet above and below
    wetu = (viv[1:] > 0) & (viv[:-1] > 0)
    assert np.allclose(o["s2u"][:km - 1][wetu], s2, rtol=1e-9)
    riu_ref = cf["ric"] / (o["s2u"][:km - 1] + 1.0e-25) * viv[1:]
    assert np.allclose(o["riu"][1:], riu_ref, rtol=1e-14, atol=0)
    assert np.all(o["riu"][0] == 0.0)
    # T grid: the mean of four identical wet U values is that value, so the
    # T-grid shear is the same where the T cell and its four U cells are wet
    wt = np.zeros_like(vit[1:], bool)
    inner = (viv[:, 1:, :-1] > 0) & (viv[:, :-1, :-1] > 0) & (viv[:, 1:, 1:] > 0) & (viv[:, :-1, 1:] > 0)
    both = inner[1:] & inner[:-1] & (vit[1:, 1:, :-1] > 0) & (vit[:-1, 1:, :-1] > 0)
    wt[:, 1:, :-1] = both
    assert wt.sum() > 50
    assert np.allclose(o["s2t"][:km - 1][wt], s2, rtol=1e-9)
    sel = np.zeros_like(vit[1:], bool)
    sel[:, 1:, :-1] = True
    rit_ref = vit[1:] * cf["rict"] / (o["s2t"][:km - 1] + 1.0e-25)
    assert np.allclose(o["rit"][sel], rit_ref[sel], rtol=1e-14, atol=0)


def test_clears_and_untouched_cells():
    c, cg, cf = base(seed=5)
    o = shear(c, cg, cf)
    km = c.km
    for k in ("s2t", "riv1", "riv2"):
        assert np.all(o[k][km - 1] == 0.0), k
    assert np.all(o["s2u"][km - 1] == 7.0)          # readyc1_2 never writes it
    assert np.all(o["rit"][:, 0, :] == 7.0) and np.all(o["rit"][:, :, -1] == 7.0)
    assert np.all(o["s2t"][:, 0, :] == 0.0) and np.all(o["s2t"][:, :, -1] == 0.0)
    assert np.array_equal(o["riv1"][:km - 1], cf["up"][:-1] - cf["up"][1:])
    assert np.array_equal(o["h0bl"], cf["h0bf"]) and np.array_equal(o["h0bf"], cf["h0"])
    for k in ("wp12", "wp13", "riv1", "riv2", "s2t", "riu"):
        assert not np.any(o[k] == 7.0), k


def test_akmu_of_uniform_akmt():
    c, cg, cf = base(seed=6)
    km, jmt, imt = c.km, c.jmt, c.imt
    akmt = np.full((km, jmt, imt), 3.0e-3)
    a = orc.readyc_akmu(imt, jmt, km, cg, akmt, cf["viv"])
    wsum = cg["au0"] + cg["aus"] + cg["auw"] + cg["ausw"]
    ref = 3.0e-3 * wsum[None, :-1, 1:] * cf["viv"][1:, :-1, 1:]
    assert np.allclose(a[:km - 1, :-1, 1:], ref, rtol=1e-15)
    assert np.all(a[:, :, 0] == 0.0) and np.all(a[:, -1, :] == 0.0) and np.all(a[km - 1] == 0.0)
