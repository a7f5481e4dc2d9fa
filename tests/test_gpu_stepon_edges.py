"""GPU parity of readyc_'s momentum tendencies, bclinc_ and barotr_ against
the oracle, bit for bit, on level counts drawn at random per column (kmt and
kmu from 0 to km, masks that disagree with them included), so that every
branch of the column kernels runs: land columns, one-level columns (the
implicit solve's single-level condition), full columns, a wet level over a
dry one."""
import numpy as np
import pytest

import licom_barotr as lbt
import licom_bclinc as lb
import licom_readyc as lrc
import licom_synth as ls
import orc

pytestmark = pytest.mark.gpu


def differ(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return int(np.count_nonzero((a != b) & ~(np.isnan(a) & np.isnan(b))))


def random_levels(rng, km, shape):
    lv = rng.integers(0, km + 1, shape).astype(np.int32)
    lv[rng.random(shape) < 0.2] = 1          # plenty of one-level columns
    return lv


@pytest.fixture(scope="module")
def case():
    c = ls.make_case("C1", nx=70, ny=40, km=12, seed=41)
    g = ls.readyt_grid(c)
    f = ls.readyt_fields(c, g, seed=42)
    ot = orc.readyt(c.imt, c.jmt, c.km, g, f)
    return c, g, f, ot


def test_momentum_random_levels(case, need_gpu):
    c, g, f, ot = case
    rng = np.random.default_rng(1)
    cg, cf = ls.readyc_inputs(c, g, ot)
    akmu = orc.readyc_akmu(c.imt, c.jmt, c.km, cg, rng.uniform(0, 1e-2, (c.km, c.jmt, c.imt)), cf["viv"])
    for del4 in (0, 1):
        mg = ls.readyc_mgrid(c, g, del4=del4)
        mg.update(kmt=random_levels(rng, c.km, (c.jmt, c.imt)), kmu=random_levels(rng, c.km, (c.jmt, c.imt)))
        mf = ls.readyc_minputs(c, ot, cf, akmu)
        want = orc.readyc_momentum(c.imt, c.jmt, c.km, cg, mg, mf)
        rc = lrc.Readyc(c.imt, c.jmt, c.km, cg, cf)
        rc.momentum_grid(mg)
        dev = rc.momentum_fields(mf)
        rc.momentum(dev)
        rc.sync()
        bad = {k: differ(dev[k].cpu().numpy(), want[k]) for k in lrc.MOUTPUTS}
        assert not any(bad.values()), (del4, bad)
        rc.close()


@pytest.mark.parametrize("isc", [0, 2])
def test_bclinc_random_levels(case, isc, need_gpu):
    c, g, f, ot = case
    rng = np.random.default_rng(2 + isc)
    bg = ls.bclinc_grid(c, g)
    bg["kmu"] = random_levels(rng, c.km, (c.jmt, c.imt))
    bf = ls.bclinc_fields(c, f)
    want = orc.bclinc_stage(c.imt, c.jmt, c.km, bg, bf, isc, (0, 1) if isc < 1 else (0, 1, 2))
    b = lb.Bclinc(c.imt, c.jmt, c.km, bg, bf)
    b.step(isc)
    got = b.host(lb.OUTPUTS)
    bad = {k: differ(got[k], want[k]) for k in lb.OUTPUTS}
    assert not any(bad.values()), bad
    b.close()


@pytest.mark.parametrize("del4", [0, 1])
def test_barotr_random_levels(case, del4, need_gpu):
    c, g, f, ot = case
    rng = np.random.default_rng(5 + del4)
    tg = ls.barotr_grid(c, g, del4=bool(del4))
    s2 = (c.jmt, c.imt)
    for k in ("kmt", "kmu", "kmtn", "kmts", "kmte", "kmtw"):
        tg[k] = random_levels(rng, c.km, s2)
    tf = ls.barotr_fields(c, f)
    tin = {k: tf[k] for k in lbt.FIELDS if k not in lbt.OUTPUTS + ("wka",)}
    want = {k: tf[k].copy() for k in lbt.OUTPUTS + ("wka",)}
    for nc in range(4):
        want = orc.barotr_stage(c.imt, c.jmt, tg, dict(tin, **want), nc, (0, 1, 2, 3))
    b = lbt.Barotr(c.imt, c.jmt, tg, tf)
    b.run(4)
    got = b.host(lbt.OUTPUTS + ("wka",))
    bad = {k: differ(got[k], want[k]) for k in lbt.OUTPUTS}
    bad["wka[2:4]"] = differ(got["wka"][2:4], want["wka"][2:4])
    assert not any(bad.values()), bad
    b.close()
