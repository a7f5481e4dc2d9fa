"""GPU parity of readyc_'s mixing preparation (licom_readyc_shear,
licom_readyc_akmu) against the oracle, bit for bit, and the whole chain
that makes the tracer step's AKT from the model's own state: readyt_ ->
shear and Richardson numbers -> the Canuto producer -> the tracer step, as
stepon_ orders them (readyt_, readyc_, ..., tracer_)."""
import dataclasses

import numpy as np
import pytest

import licom_readyc as lrc
import licom_readyt as lr
import licom_synth as ls
import licom_tracer as lt
import licom_turb as ltu
import orc

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[readyc] {msg}", flush=True)


def differ(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return int(np.count_nonzero((a != b) & ~(np.isnan(a) & np.isnan(b))))


def readyt_both(c, seed=12):
    g = ls.readyt_grid(c)
    f = ls.readyt_fields(c, g, seed=seed)
    want = orc.readyt(c.imt, c.jmt, c.km, g, f)
    r = lr.Readyt(c.imt, c.jmt, c.km, g, f)
    r.readyt(1)
    return g, want, r


@pytest.mark.parametrize("shape", [dict(), dict(nx=100, ny=50, km=55, seed=31)])
def test_readyc_shear_and_akmu_bitwise(shape, need_gpu):
    c = ls.make_case("C1", **shape)
    announce(f"shear + akmu {c.imt}x{c.jmt}x{c.km}")
    g, ot, rt = readyt_both(c)
    cg, cf = ls.readyc_inputs(c, g, ot)
    want = orc.readyc_shear(c.imt, c.jmt, c.km, cg, cf)
    rc = lrc.Readyc(c.imt, c.jmt, c.km, cg, cf)
    rc.shear()
    got = rc.host()
    bad = {k: differ(got[k], want[k]) for k in lrc.OUTPUTS}
    assert not any(bad.values()), bad
    akmt = np.random.default_rng(2).uniform(0.0, 1.0e-2, (c.km, c.jmt, c.imt))
    akmu_ref = orc.readyc_akmu(c.imt, c.jmt, c.km, cg, akmt, cf["viv"])
    t_akmt = rc.torch.from_numpy(akmt).to(rc.device)
    akmu = rc.akmu(t_akmt).cpu().numpy()
    assert differ(akmu, akmu_ref) == 0
    announce(f"  bit-exact; shear pass {rc.last_ms():.3f} ms")
    rc.close()
    rt.close()


def test_readyt_readyc_turb_tracer_chain(need_gpu):
    """AKT from the model's own T, S and velocities, on the device end to end,
    then one FP64 tracer step; the oracle runs the same chain.  readyt_ and
    readyc_ agree bit for bit, AKT and the step within the path's 1e-12
    relative (the producer's elementary functions differ in their last bits
    between the device and host libraries)."""
    c = ls.make_case("C1")
    km, jmt, imt = c.km, c.jmt, c.imt
    announce(f"chain readyt -> readyc -> turb -> tracer step on {imt}x{jmt}x{km}")
    g, ot, rt = readyt_both(c)
    cg, cf = ls.readyc_inputs(c, g, ot)
    oc = orc.readyc_shear(imt, jmt, km, cg, cf)
    # the device chain reads readyt_'s outputs in place
    dev = {k: rt.fields[k] for k in ("vit", "viv", "rict", "ric")}
    fields = dict(cf)
    fields.update(dev)
    rc = lrc.Readyc(imt, jmt, km, cg, fields)
    rc.shear()
    torch = rc.torch

    def pad(a):            # km-1 planes -> the producer's km
        out = np.zeros((km, jmt, imt))
        out[:km - 1] = a
        return out

    tc = ls.make_turb_case(c, seed=5)
    tc.fields.update(at=np.ascontiguousarray(c.state["at"][0]), pdensity=ot["pdensity"],
                     rit=pad(oc["rit"]), ricdttms=pad(ot["ricdttms"]), s2t=oc["s2t"],
                     rict=pad(ot["rict"]), vit=ot["vit"], buoytur=ot["buoytur"],
                     buoysol=ot["buoysol"])
    tc.fields = {k: np.ascontiguousarray(v, np.float64) for k, v in tc.fields.items()}
    ref_turb, _ = orc.TurbOracle(tc.tables).run(tc)
    # the device producer reads the device chain's arrays in place
    m = lt.TracerModel(c)
    zeros = torch.zeros((1, jmt, imt), dtype=torch.float64, device=rc.device)
    dev_fields = dict(tc.fields)
    dev_fields.update(at=m.state["at"][0], pdensity=rt.fields["pdensity"],
                      rit=torch.cat([rc.fields["rit"], zeros]),
                      ricdttms=torch.cat([rt.fields["ricdttms"], zeros]),
                      rict=torch.cat([rt.fields["rict"], zeros]), s2t=rc.fields["s2t"],
                      vit=rt.fields["vit"], buoytur=rt.fields["buoytur"],
                      buoysol=rt.fields["buoysol"])
    p = ltu.TurbProducer(dataclasses.replace(tc, fields=dev_fields))
    p.run(akt=m.state["akt"])
    akmu = rc.akmu(p.outputs["akmt"])
    p.sync()
    got_akt = m.state["akt"].cpu().numpy()
    # the oracle's AKT where the producer writes it, the case's elsewhere
    akt_o = c.state["akt"].copy()
    wet = np.zeros((jmt, imt), bool)
    wet[1:-1, 1:-1] = c.grid["kmt"][1:-1, 1:-1] > 0
    akt_o[:, :km - 1] = np.where(wet[None, None], ref_turb["akt"][:, :km - 1], akt_o[:, :km - 1])
    d = np.abs(got_akt - akt_o)
    assert np.all(d <= 1e-12 * np.abs(akt_o)), float(np.max(d / np.maximum(np.abs(akt_o), 1e-300)))
    akmu_ref = orc.readyc_akmu(imt, jmt, km, cg, ref_turb["akmt"], ot["viv"])
    da = np.abs(akmu.cpu().numpy() - akmu_ref)
    assert np.all(da <= 1e-12 * np.abs(akmu_ref) + 1e-300), float(np.max(da))
    m.tracer(0)
    m.sync()
    st = c.copy_state()
    st["akt"] = akt_o
    o = orc.Oracle(c)
    diag = o.new_diag(c)
    o.step(st, diag, convadj=0)
    got = m.host_state()
    for k in ("at", "atb"):
        err = float(np.max(np.abs(got[k] - st[k])) / np.max(np.abs(st[k])))
        announce(f"  tracer {k}: rel err {err:.2e}")
        assert err <= 1e-12, (k, err)
    m.close()
    p.close()
    rc.close()
    rt.close()


def test_readyc_argument_checks(need_gpu):
    c = ls.make_case("C1", nx=16, ny=8, km=6, seed=2)
    g = ls.readyt_grid(c)
    f = ls.readyt_fields(c, g)
    o = orc.readyt(c.imt, c.jmt, c.km, g, f)
    cg, cf = ls.readyc_inputs(c, g, o)
    with pytest.raises(lt.LicomError, match="invalid extents"):
        lrc.Readyc(c.imt, c.jmt, 1, cg)
    rc = lrc.Readyc(c.imt, c.jmt, c.km, cg, cf)
    with pytest.raises(lt.LicomError, match="NULL"):
        lt.check(lrc.lib().licom_readyc_akmu(rc.handle, None, None, None), "licom_readyc_akmu")
    rc.close()
