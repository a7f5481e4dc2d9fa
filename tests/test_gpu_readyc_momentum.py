"""GPU parity of readyc_'s momentum tendencies (licom_readyc_momentum)
against the oracle, bit for bit, for the Laplacian and the biharmonic
viscosity, on a grid that tiles the launch exactly and on a ragged one at
55 levels; then the whole of readyc_ on the device from readyt_'s output
(shear -> Canuto producer's AKMT from the oracle -> akmu -> momentum), and
the argument checks."""
import numpy as np
import pytest

import licom_readyc as lrc
import licom_readyt as lr
import licom_synth as ls
import licom_tracer as lt
import orc

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[readyc-mom] {msg}", flush=True)


def differ(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return int(np.count_nonzero((a != b) & ~(np.isnan(a) & np.isnan(b))))


def inputs(c, del4, seed=12):
    g = ls.readyt_grid(c)
    f = ls.readyt_fields(c, g, seed=seed)
    ot = orc.readyt(c.imt, c.jmt, c.km, g, f)
    cg, cf = ls.readyc_inputs(c, g, ot)
    akmt = np.random.default_rng(seed + 5).uniform(0.0, 1.0e-2, (c.km, c.jmt, c.imt))
    akmu = orc.readyc_akmu(c.imt, c.jmt, c.km, cg, akmt, cf["viv"])
    mg = ls.readyc_mgrid(c, g, del4=del4)
    mf = ls.readyc_minputs(c, ot, cf, akmu)
    return cg, cf, mg, mf


@pytest.mark.parametrize("del4", [0, 1])
@pytest.mark.parametrize("shape", [dict(), dict(nx=100, ny=50, km=55, seed=31)])
def test_readyc_momentum_bitwise(shape, del4, need_gpu):
    c = ls.make_case("C1", **shape)
    announce(f"momentum {'del4' if del4 else 'del2'} {c.imt}x{c.jmt}x{c.km}")
    cg, cf, mg, mf = inputs(c, del4)
    want = orc.readyc_momentum(c.imt, c.jmt, c.km, cg, mg, mf)
    rc = lrc.Readyc(c.imt, c.jmt, c.km, cg, cf)
    rc.momentum_grid(mg)
    dev = rc.momentum_fields(mf)
    for rep in range(2):            # a second call from the same inputs gives the same bits
        rc.momentum(dev)
        rc.sync()
        got = {k: dev[k].cpu().numpy() for k in lrc.MOUTPUTS}
        bad = {k: differ(got[k], want[k]) for k in lrc.MOUTPUTS}
        assert not any(bad.values()), (rep, bad)
    # the cells each output leaves alone keep what they held
    assert np.all(got["sbcx"][0] == 7.0) and np.all(got["bbcy"][:, -1] == 7.0)
    assert np.all(got["ws"][1:, 0] == 7.0) and np.all(got["ws"][1:, :, -1] == 7.0)
    announce(f"  bit-exact; {rc.last_ms():.3f} ms")
    rc.close()


def test_readyc_whole_on_device(need_gpu):
    """readyt_ -> readyc_'s shear -> akmu -> momentum, each stage reading the
    previous one's device arrays in place (the AKMT between is the
    oracle's Canuto-independent random field uploaded once)."""
    c = ls.make_case("C1", nx=100, ny=50, km=30, seed=8)
    km, jmt, imt = c.km, c.jmt, c.imt
    announce(f"readyt -> readyc (shear, akmu, momentum del4) on {imt}x{jmt}x{km}")
    g = ls.readyt_grid(c)
    f = ls.readyt_fields(c, g, seed=12)
    ot = orc.readyt(imt, jmt, km, g, f)
    rt = lr.Readyt(imt, jmt, km, g, f)
    rt.readyt(1)
    cg, cf = ls.readyc_inputs(c, g, ot)
    fields = dict(cf)
    fields.update({k: rt.fields[k] for k in ("vit", "viv", "rict", "ric")})
    rc = lrc.Readyc(imt, jmt, km, cg, fields)
    rc.shear()
    akmt = np.random.default_rng(4).uniform(0.0, 1.0e-2, (km, jmt, imt))
    t_akmt = rc.torch.from_numpy(akmt).to(rc.device)
    akmu = rc.akmu(t_akmt)
    mg = ls.readyc_mgrid(c, g, del4=1)
    akmu_ref = orc.readyc_akmu(imt, jmt, km, cg, akmt, ot["viv"])
    mf = ls.readyc_minputs(c, ot, cf, akmu_ref)
    want = orc.readyc_momentum(imt, jmt, km, cg, mg, mf)
    rc.momentum_grid(mg)
    host = dict(mf)
    host.update(up=rc.fields["up"], vp=rc.fields["vp"], h0=rc.fields["h0"], vit=rt.fields["vit"],
                viv=rt.fields["viv"], akmu=akmu)
    dev = rc.momentum_fields(host)
    rc.momentum(dev)
    rc.sync()
    bad = {k: differ(dev[k].cpu().numpy(), want[k]) for k in lrc.MOUTPUTS}
    assert not any(bad.values()), bad
    rc.close()
    rt.close()


def test_readyc_momentum_argument_checks(need_gpu):
    c = ls.make_case("C1", nx=16, ny=8, km=6, seed=2)
    cg, cf, mg, mf = inputs(c, 1)
    rc = lrc.Readyc(c.imt, c.jmt, c.km, cg, cf)
    dev = rc.momentum_fields(mf)
    with pytest.raises(lt.LicomError, match="no momentum grid"):
        rc.momentum(dev)
    bad = dict(mg, jb=2)                 # del4 reads two cells beyond jb-1
    with pytest.raises(lt.LicomError, match="block bounds"):
        rc.momentum_grid(bad)
    bad = dict(mg, kmu=np.full((c.jmt, c.imt), c.km + 1, np.int32))
    with pytest.raises(lt.LicomError, match="kmt or kmu"):
        rc.momentum_grid(bad)
    rc.momentum_grid(dict(mg, del4=0, jb=2, ib=2, je=c.jmt - 1, ie=c.imt - 1))   # del2 allows these
    rc.momentum_grid(mg)
    with pytest.raises(lt.LicomError, match="NULL"):
        lt.check(lrc.lib().licom_readyc_momentum(rc.handle, None), "licom_readyc_momentum")
    rc.momentum(dev)
    rc.sync()
    rc.close()
