"""GPU parity of bclinc_ (licom_bclinc_stage) against the oracle, bit for
bit: the first step (isc = 0: Euler, bclinc6/71/72, then 81/82) and the
leapfrog (bclinc2/3, 4, 5) on a grid that tiles the launch and on a ragged
one at 55 levels, each stage compared; a stand-in halo exchange between the
stages (east-west periodic copy, the same on both sides); then readyc_'s
momentum tendencies feeding bclinc_ on the device, and the argument checks."""
import numpy as np
import pytest

import licom_bclinc as lb
import licom_readyc as lrc
import licom_synth as ls
import licom_tracer as lt
import orc

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[bclinc] {msg}", flush=True)


def differ(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return int(np.count_nonzero((a != b) & ~(np.isnan(a) & np.isnan(b))))


def periodic(a):
    """east-west halo exchange of a block with a two-cell halo"""
    a[..., :2] = a[..., -4:-2]
    a[..., -2:] = a[..., 2:4]


@pytest.mark.parametrize("isc", [0, 1, 5])
@pytest.mark.parametrize("shape", [dict(), dict(nx=100, ny=50, km=55, seed=31)])
def test_bclinc_stages_bitwise(shape, isc, need_gpu):
    c = ls.make_case("C1", **shape)
    announce(f"isc={isc} {c.imt}x{c.jmt}x{c.km}")
    rg = ls.readyt_grid(c)
    rf = ls.readyt_fields(c, rg)
    g = ls.bclinc_grid(c, rg)
    f = ls.bclinc_fields(c, rf)
    b = lb.Bclinc(c.imt, c.jmt, c.km, g, f)
    want = {k: f[k].copy() for k in lb.OUTPUTS}
    host_in = {k: f[k] for k in lb.FIELDS if k not in lb.OUTPUTS}
    stages = (0, 1) if isc < 1 else (0, 1, 2)
    for s in stages:
        want = orc.bclinc_stage(c.imt, c.jmt, c.km, g, dict(host_in, **want), isc, [s])
        b.stage(isc, s)
        got = b.host(lb.OUTPUTS)
        bad = {k: differ(got[k], want[k]) for k in lb.OUTPUTS}
        assert not any(bad.values()), (s, bad)
        # the caller's halo exchange between stages, on both sides
        names = ("u", "v") if isc < 1 else ("wka",)
        if s < stages[-1]:
            for k in names:
                periodic(want[k])
                t = b.fields[k]
                t[..., :2] = t[..., -4:-2].clone()
                t[..., -2:] = t[..., 2:4].clone()
    announce(f"  bit-exact over stages {stages}; last stage {b.last_ms():.3f} ms")
    b.close()


def test_readyc_momentum_then_bclinc_on_device(need_gpu):
    """dlu, dlv, sbcx..bbcy and akmu from licom_readyc_momentum feed
    licom_bclinc_stage in place; the oracle runs the same chain."""
    c = ls.make_case("C1", nx=100, ny=50, km=30, seed=8)
    km, jmt, imt = c.km, c.jmt, c.imt
    announce(f"readyc momentum -> bclinc (isc=2) on {imt}x{jmt}x{km}")
    rg = ls.readyt_grid(c)
    rf = ls.readyt_fields(c, rg)
    ot = orc.readyt(imt, jmt, km, rg, rf)
    cg, cf = ls.readyc_inputs(c, rg, ot)
    akmt = np.random.default_rng(4).uniform(0.0, 1.0e-2, (km, jmt, imt))
    akmu = orc.readyc_akmu(imt, jmt, km, cg, akmt, cf["viv"])
    mg = ls.readyc_mgrid(c, rg, del4=1)
    mf = ls.readyc_minputs(c, ot, cf, akmu)
    om = orc.readyc_momentum(imt, jmt, km, cg, mg, mf)
    rc = lrc.Readyc(imt, jmt, km, cg, cf)
    rc.momentum_grid(mg)
    dev = rc.momentum_fields(mf)
    rc.momentum(dev)
    g = ls.bclinc_grid(c, rg)
    f = ls.bclinc_fields(c, rf)
    f.update(up=mf["up"], vp=mf["vp"], u=mf["u"], v=mf["v"], akmu=akmu)
    hf = dict(f, dlu=om["dlu"], dlv=om["dlv"])
    want = orc.bclinc_stage(imt, jmt, km, g, hf, 2, [0, 1, 2])
    dfields = dict(f)
    dfields.update(up=dev["up"], vp=dev["vp"], u=dev["u"], v=dev["v"], akmu=dev["akmu"], dlu=dev["dlu"],
                   dlv=dev["dlv"])
    rc.sync()
    b = lb.Bclinc(imt, jmt, km, g, {k: (v.clone() if hasattr(v, "clone") else v) for k, v in dfields.items()})
    b.step(2)
    got = b.host(lb.OUTPUTS)
    bad = {k: differ(got[k], want[k]) for k in lb.OUTPUTS}
    assert not any(bad.values()), bad
    b.close()
    rc.close()


def test_bclinc_argument_checks(need_gpu):
    c = ls.make_case("C1", nx=16, ny=8, km=6, seed=2)
    rg = ls.readyt_grid(c)
    g = ls.bclinc_grid(c, rg)
    f = ls.bclinc_fields(c, ls.readyt_fields(c, rg))
    b = lb.Bclinc(c.imt, c.jmt, c.km, g, f)
    with pytest.raises(lt.LicomError, match="no stage 2"):
        b.stage(0, 2)
    with pytest.raises(lt.LicomError, match="no stage 3"):
        b.stage(1, 3)
    with pytest.raises(lt.LicomError, match="NULL"):
        lt.check(lb.lib().licom_bclinc_stage(b.handle, None, 1, 0), "licom_bclinc_stage")
    b.step(1)
    b.sync()
    b.close()
