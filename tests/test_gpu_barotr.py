"""GPU parity of barotr_'s subcycle (licom_barotr_stage) against the
oracle, bit for bit after every stage, with the Laplacian and the
biharmonic viscosity, over four subcycles (the fourth carries the
free-surface diffusion), on a grid that tiles the launch and on a ragged
one, with a stand-in halo exchange between the stages (east-west periodic
copy, the same on both sides); then the argument checks."""
import numpy as np
import pytest

import licom_barotr as lbt
import licom_synth as ls
import licom_tracer as lt
import orc

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[barotr] {msg}", flush=True)


def differ(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return int(np.count_nonzero((a != b) & ~(np.isnan(a) & np.isnan(b))))


def periodic_np(a):
    a[..., :2] = a[..., -4:-2]
    a[..., -2:] = a[..., 2:4]


def periodic_t(t):
    t[..., :2] = t[..., -4:-2].clone()
    t[..., -2:] = t[..., 2:4].clone()


@pytest.mark.parametrize("del4", [0, 1])
@pytest.mark.parametrize("shape", [dict(), dict(nx=100, ny=50, km=20, seed=31)])
def test_barotr_subcycles_bitwise(shape, del4, need_gpu):
    c = ls.make_case("C1", **shape)
    announce(f"{'del4' if del4 else 'del2'} {c.imt}x{c.jmt}, 4 subcycles")
    rg = ls.readyt_grid(c)
    rf = ls.readyt_fields(c, rg)
    g = ls.barotr_grid(c, rg, del4=bool(del4))
    f = ls.barotr_fields(c, rf)
    b = lbt.Barotr(c.imt, c.jmt, g, f)
    host_in = {k: f[k] for k in lbt.FIELDS if k not in lbt.OUTPUTS + ("wka",)}
    want = {k: f[k].copy() for k in lbt.OUTPUTS + ("wka",)}
    for nc in range(4):
        for s in range(4):
            want = orc.barotr_stage(c.imt, c.jmt, g, dict(host_in, **want), nc, [s])
            b.stage(nc, s)
            got = b.host(lbt.OUTPUTS + ("wka",))
            bad = {k: differ(got[k], want[k]) for k in lbt.OUTPUTS}
            # planes 0, 1, 4, 5 are the reference's scratch, formed in registers
            bad["wka[2:4]"] = differ(got["wka"][2:4], want["wka"][2:4])
            assert not any(bad.values()), (nc, s, bad)
            if s in lbt.HALO:
                if s == 1:
                    periodic_np(want["wka"][2:4])
                    periodic_t(b.fields["wka"][2:4])
                else:
                    periodic_np(want["work"])
                    periodic_t(b.fields["work"])
    announce(f"  bit-exact over 16 stages; last {b.last_ms():.3f} ms")
    b.close()


def test_barotr_argument_checks(need_gpu):
    c = ls.make_case("C1", nx=16, ny=8, km=6, seed=2)
    rg = ls.readyt_grid(c)
    g = ls.barotr_grid(c, rg, del4=True)
    f = ls.barotr_fields(c, ls.readyt_fields(c, rg))
    with pytest.raises(lt.LicomError, match="block bounds"):
        lbt.Barotr(c.imt, c.jmt, dict(g, jb=2), f)
    b = lbt.Barotr(c.imt, c.jmt, g, f)
    with pytest.raises(lt.LicomError, match="no stage 4"):
        b.stage(0, 4)
    with pytest.raises(lt.LicomError, match="NULL"):
        lt.check(lbt.lib().licom_barotr_stage(b.handle, None, 0, 0), "licom_barotr_stage")
    b.run(2)
    b.sync()
    b.close()
