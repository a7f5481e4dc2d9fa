"""GPU parity of readyt_ (licom_readyt_run) against the oracle: every field
the call writes, bit for bit (ricdt is 0/0 on land in the reference, so NaN
positions must agree), with cells the call leaves alone still holding the
sentinel they started with."""
import numpy as np
import pytest

import licom_readyt as lr
import licom_synth as ls
import licom_tracer as lt
import orc

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[readyt] {msg}", flush=True)


def differ(a, b):
    """Entries that differ, NaN equal to NaN."""
    a, b = np.asarray(a), np.asarray(b)
    both_nan = np.isnan(a) & np.isnan(b)
    return int(np.count_nonzero((a != b) & ~both_nan))


def check_case(case, ist, with_diag, seed=12):
    g = ls.readyt_grid(case)
    f = ls.readyt_fields(case, g, seed=seed)
    if with_diag:
        for k in ("ax", "ay", "az"):
            f[k] = np.full((2, case.km, case.jmt, case.imt), 5.0)
    announce(f"{case.imt}x{case.jmt}x{case.km} ist={ist} diag={with_diag}")
    want = orc.readyt(case.imt, case.jmt, case.km, g, f, ist=ist)
    r = lr.Readyt(case.imt, case.jmt, case.km, g, f)
    r.readyt(ist)
    got = r.host()
    ms = r.last_ms()
    r.close()
    bad = {k: differ(got[k], want[k]) for k in want}
    bad = {k: v for k, v in bad.items() if v}
    assert not bad, f"fields differing from the oracle: {bad}"
    nan = int(np.isnan(want["ricdt"]).sum())
    announce(f"  bit-exact, {nan} NaN cells of ricdt in both; run {ms[0]:.3f} ms")
    return want


def test_readyt_c1_bitwise(need_gpu):
    case = ls.make_case("C1")
    o = check_case(case, ist=1, with_diag=False)
    # the call wrote every output cell
    for k in lr.OUTPUTS:
        assert not np.any(o[k] == 7.0), k


def test_readyt_partial_blocks_55_levels_and_ist0(need_gpu):
    case = ls.make_case("C1", nx=100, ny=50, km=55, seed=31)
    o = check_case(case, ist=0, with_diag=True, seed=5)
    assert all(np.all(o[k] == 0.0) for k in ("ax", "ay", "az"))


def test_readyt_keeps_diagnostics_after_the_first_step(need_gpu):
    case = ls.make_case("C1", nx=64, ny=20, km=12, seed=8, depth_min=-2000.0, sigma=1.5)
    o = check_case(case, ist=3, with_diag=True, seed=6)
    assert all(np.all(o[k] == 5.0) for k in ("ax", "ay", "az"))


def test_readyt_full_width_band(need_gpu):
    """A band of the 0.1 deg grid: 3600 columns wide, 40 rows, 55 levels."""
    case = ls.make_case("C3", ny=40)
    check_case(case, ist=1, with_diag=False, seed=9)


def test_readyt_argument_checks(need_gpu):
    case = ls.make_case("C1", nx=16, ny=8, km=6, seed=2)
    g = ls.readyt_grid(case)
    bad = dict(g)
    bad["kmu"] = g["kmu"].copy()
    bad["kmu"][3, 3] = case.km + 1
    with pytest.raises(lt.LicomError, match="kmu out of range"):
        lr.Readyt(case.imt, case.jmt, case.km, bad)
    with pytest.raises(lt.LicomError, match="invalid extents"):
        lr.Readyt(case.imt, case.jmt, 1, g)
    f = ls.readyt_fields(case, g)
    r = lr.Readyt(case.imt, case.jmt, case.km, g, f)
    r.fields["pp"] = None
    with pytest.raises(lt.LicomError, match="NULL"):
        r.readyt(1)
    r.close()
