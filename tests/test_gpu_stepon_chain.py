"""stepon_'s momentum sequence on the device, each stage reading the previous
one's device arrays in place, against the oracle's chain on the same inputs,
bit for bit (stepon.cpp:59-66): readyt_ -> readyc_ (shear, AKMT at U points
from a given AKMT, momentum tendencies) -> barotr_ (four subcycles) ->
bclinc_ (a leapfrog step).  No halo exchange (a single block)."""
import numpy as np
import pytest

import licom_barotr as lbt
import licom_bclinc as lb
import licom_readyc as lrc
import licom_readyt as lr
import licom_synth as ls
import orc

pytestmark = pytest.mark.gpu

NBB = 4


def differ(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return int(np.count_nonzero((a != b) & ~(np.isnan(a) & np.isnan(b))))


def test_readyt_readyc_barotr_bclinc_chain(need_gpu):
    c = ls.make_case("C1", nx=100, ny=50, km=30, seed=8)
    km, jmt, imt = c.km, c.jmt, c.imt
    print(f"[stepon] readyt -> readyc -> barotr x{NBB} -> bclinc on {imt}x{jmt}x{km}", flush=True)
    # readyt_
    g = ls.readyt_grid(c)
    f = ls.readyt_fields(c, g, seed=12)
    ot = orc.readyt(imt, jmt, km, g, f)
    rt = lr.Readyt(imt, jmt, km, g, f)
    rt.readyt(1)
    # readyc_: shear, akmu, momentum (del4)
    cg, cf = ls.readyc_inputs(c, g, ot)
    rc_fields = dict(cf)
    rc_fields.update({k: rt.fields[k] for k in ("vit", "viv", "rict", "ric")})
    rc = lrc.Readyc(imt, jmt, km, cg, rc_fields)
    rc.shear()
    akmt = np.random.default_rng(4).uniform(0.0, 1.0e-2, (km, jmt, imt))
    akmu_o = orc.readyc_akmu(imt, jmt, km, cg, akmt, ot["viv"])
    akmu_d = rc.akmu(rc.torch.from_numpy(akmt).to(rc.device))
    mg = ls.readyc_mgrid(c, g, del4=1)
    mf = ls.readyc_minputs(c, ot, cf, akmu_o)
    om = orc.readyc_momentum(imt, jmt, km, cg, mg, mf)
    rc.momentum_grid(mg)
    dm = rc.momentum_fields(dict(mf, vit=rt.fields["vit"], viv=rt.fields["viv"], akmu=akmu_d,
                                 up=rc.fields["up"], vp=rc.fields["vp"], h0=rc.fields["h0"]))
    rc.momentum(dm)
    # barotr_: forcing from readyt_ and readyc_
    tg = ls.barotr_grid(c, g, del4=True)
    tf = ls.barotr_fields(c, f)
    pick2 = ("pax", "pay", "pxb", "pyb", "whx", "why", "wgp")
    tf_o = dict(tf, dlub=om["dlub"], dlvb=om["dlvb"], **{k: ot[k] for k in pick2},
                vit=np.ascontiguousarray(ot["vit"][0]), viv=np.ascontiguousarray(ot["viv"][0]))
    tf_d = dict(tf, dlub=dm["dlub"], dlvb=dm["dlvb"], **{k: rt.fields[k] for k in pick2},
                vit=rt.fields["vit"][0].contiguous(), viv=rt.fields["viv"][0].contiguous())
    bt = lbt.Barotr(imt, jmt, tg, tf_d)
    bo = {k: tf_o[k].copy() for k in lbt.OUTPUTS + ("wka",)}
    bin_o = {k: tf_o[k] for k in lbt.FIELDS if k not in lbt.OUTPUTS + ("wka",)}
    for nc in range(NBB):
        bo = orc.barotr_stage(imt, jmt, tg, dict(bin_o, **bo), nc, (0, 1, 2, 3))
    bt.run(NBB)
    got_b = bt.host(lbt.OUTPUTS)      # before bclinc_ scales h0bf in place
    # bclinc_: buoyancy from readyt_, tendencies from readyc_, barotropic
    # velocities and h0bf from barotr_
    bg = ls.bclinc_grid(c, g)
    bf = ls.bclinc_fields(c, f)
    bf_o = dict(bf, gg=ot["gg"], akmu=akmu_o, dlu=om["dlu"], dlv=om["dlv"], ub=bo["ub"], vb=bo["vb"],
                h0bf=bo["h0bf"], u=mf["u"], v=mf["v"], up=mf["up"], vp=mf["vp"], vit=ot["vit"], viv=ot["viv"])
    want = orc.bclinc_stage(imt, jmt, km, bg, bf_o, 1, (0, 1, 2))
    bf_d = dict(bf, gg=rt.fields["gg"], akmu=akmu_d, dlu=dm["dlu"], dlv=dm["dlv"], ub=bt.fields["ub"],
                vb=bt.fields["vb"], h0bf=bt.fields["h0bf"], u=dm["u"], v=dm["v"], up=dm["up"], vp=dm["vp"],
                vit=rt.fields["vit"], viv=rt.fields["viv"])
    rc.sync()
    bc = lb.Bclinc(imt, jmt, km, bg, bf_d)
    bc.step(1)
    got_c = bc.host(lb.OUTPUTS)
    bad = {("barotr", k): differ(got_b[k], bo[k]) for k in lbt.OUTPUTS}
    bad.update({("bclinc", k): differ(got_c[k], want[k]) for k in lb.OUTPUTS})
    assert not any(bad.values()), {k: v for k, v in bad.items() if v}
    for o in (bc, bt, rc, rt):
        o.close()
