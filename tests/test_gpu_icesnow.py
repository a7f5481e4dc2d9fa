"""GPU parity of icesnow_ (icesnow.cpp:15-85): the stand-alone entry
licom_icesnow and the LICOM_STEP_ICESNOW form fused into the column solve,
against the oracle's orc_icesnow in the reference's order
tracer_ -> icesnow_ -> convadj_ (stepon.cpp:69-75).  FP64 results are
compared bit for bit (+0 == -0), licomqice included.

The initial states carry cells planted 0.01-1.2 K below the freezing point
(tests/helpers.py:plant_cold) on land, ghost and wet columns, at the surface
and at depth, so that every branch is taken.  Each test prints its case
before touching the GPU."""
import ctypes
import sys

import numpy as np
import pytest

import licom_synth as ls
import licom_tracer as lt
import orc
from helpers import TBICE, inner, mismatch, plant_cold, report
from test_gpu_parity import compare

pytestmark = pytest.mark.gpu

SIDE, CONV, ICE = lt.LICOM_STEP_SIDE_EFFECTS, lt.LICOM_STEP_CONVADJ, lt.LICOM_STEP_ICESNOW


def announce(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def cold_case(seed, **over):
    kw = dict(depth_min=-3000.0, depth_max=5000.0, sigma=1.5, seed=11)
    kw.update(over)
    case = ls.make_case("C1", **kw)
    plant_cold(case, case.state, seed)
    return case


def check_qice(model, ref, tag):
    got = model.licomqice.cpu().numpy()
    n = mismatch(got, ref)
    print(f"[{tag}] " + report("licomqice", got, ref) + f"; {int(np.count_nonzero(ref))} columns froze",
          flush=True)
    assert np.count_nonzero(ref) > 0, "the case exercises no freezing"
    assert n == 0, f"{tag}: licomqice differs in {n} columns"


def test_icesnow_alone_bitwise(need_gpu):
    announce("icesnow alone C1-land, reset and carried-over ice")
    case = cold_case(21)
    o = orc.Oracle(case)
    st = case.copy_state()
    q = o.icesnow(st, np.zeros((case.jmt, case.imt)), reset=True)
    m = lt.TracerModel(case)
    m.icesnow(reset=True)
    got = m.host_state()["at"]
    print("[icesnow] " + report("at", got, st["at"]), flush=True)
    assert mismatch(got, st["at"]) == 0
    assert (got[0] >= TBICE).all()
    check_qice(m, q, "icesnow reset")
    # melt branch: ice carried over (no reset) against surface T above TBICE
    rng = np.random.default_rng(4)
    q2 = np.where(rng.random(q.shape) < 0.3, rng.uniform(0.0, 1.0, q.shape), 0.0)
    st["at"][0][0] = np.maximum(st["at"][0][0], TBICE + rng.uniform(0.0, 0.6, q.shape))
    m.state["at"].copy_(m.torch.from_numpy(st["at"]))
    m.licomqice.copy_(m.torch.from_numpy(q2))
    o.icesnow(st, q2, reset=False)
    m.icesnow(reset=False)
    got = m.host_state()["at"]
    print("[icesnow melt] " + report("at", got, st["at"]), flush=True)
    assert mismatch(got, st["at"]) == 0
    got_q = m.licomqice.cpu().numpy()
    assert mismatch(got_q, q2) == 0
    m.close()


def oracle_step(o, st, od, q, convadj):
    o.step(st, od, convadj=0)
    o.icesnow(st, q, reset=True)
    if convadj:
        o.convadj(st, od)


@pytest.mark.parametrize("flags,tag", [
    (SIDE | CONV | ICE, "fused icesnow + convadj"),
    (SIDE | ICE, "fused icesnow"),
    (SIDE | CONV | ICE | lt.LICOM_STEP_STAGED, "staged icesnow + convadj"),
])
def test_step_with_icesnow_bitwise(flags, tag, need_gpu):
    announce(f"C1-land del4 x3, {tag}")
    case = cold_case(22, hmix=4, ah=-1.0e10)
    o = orc.Oracle(case)
    st = case.copy_state()
    od = orc.Oracle.new_diag(case)
    q = np.zeros((case.jmt, case.imt))
    m = lt.TracerModel(case)
    froze = 0
    for s in range(3):
        oracle_step(o, st, od, q, flags & CONV)
        m.tracer(s, flags)
        got = m.licomqice.cpu().numpy()
        assert mismatch(got, q) == 0, f"{tag}: licomqice differs after step {s}"
        froze += int(np.count_nonzero(q))
    print(f"[{tag}] {froze} column freezes over 3 steps", flush=True)
    assert froze > 0
    compare(m, st, od, tag, exact=True, tol=1e-10)
    m.close()


def test_step_with_icesnow_km55_bitwise(need_gpu):
    """The 55-level instantiation of the fused solve (the 0.1 deg grid's)."""
    announce("C1-land km55 x2, fused icesnow + convadj")
    case = cold_case(25, km=55, hmix=4, ah=-1.0e10)
    o = orc.Oracle(case)
    st = case.copy_state()
    od = orc.Oracle.new_diag(case)
    q = np.zeros((case.jmt, case.imt))
    m = lt.TracerModel(case)
    froze = 0
    for s in range(2):
        oracle_step(o, st, od, q, True)
        m.tracer(s, SIDE | CONV | ICE)
        assert mismatch(m.licomqice.cpu().numpy(), q) == 0, f"licomqice differs after step {s}"
        froze += int(np.count_nonzero(q))
    assert froze > 0
    compare(m, st, od, "C1 km55 icesnow + convadj", exact=True, tol=1e-10)
    m.close()


def test_mixed_fused_icesnow_equals_icesnow_after_the_step(need_gpu):
    """The mixed form's fused icesnow gives the same bits as its step
    followed by the stand-alone icesnow (same arithmetic, other order)."""
    announce("mixed C1-land km55 x2: fused icesnow vs step then licom_icesnow")
    case = cold_case(23, km=55)
    a = lt.TracerModel(case, precision="mixed")
    b = lt.TracerModel(case, precision="mixed")
    qb = b.torch.zeros_like(b.licomqice)
    L = lt.lib()
    froze = 0
    for s in range(2):
        a.tracer(s, SIDE | ICE)
        b.tracer(s, SIDE)
        lt.check(L.licom_icesnow(b.ctx, ctypes.c_void_p(b.state["at"].data_ptr()),
                                 ctypes.c_void_p(qb.data_ptr()), 1), "licom_icesnow")
        qa = a.licomqice.cpu().numpy()
        assert mismatch(qa, qb.cpu().numpy()) == 0, f"licomqice differs after step {s}"
        froze += int(np.count_nonzero(qa))
    assert froze > 0
    sa, sb = a.host_state(), b.host_state()
    da, db = a.host_diag(), b.host_diag()
    for k in ("at", "atb"):
        print("[mixed ice] " + report(k, sa[k], sb[k]), flush=True)
        assert mismatch(sa[k], sb[k]) == 0, k
    for k in ("tend", "dt_diff", "ax", "dz"):
        assert mismatch(inner(da[k]), inner(db[k])) == 0, k
    a.close()
    b.close()


def test_mixed_icesnow_convadj_against_oracle(need_gpu):
    announce("mixed C1-land x1 icesnow + convadj vs oracle")
    case = cold_case(24)
    o = orc.Oracle(case)
    st = case.copy_state()
    od = orc.Oracle.new_diag(case)
    q = np.zeros((case.jmt, case.imt))
    oracle_step(o, st, od, q, True)
    m = lt.TracerModel(case, precision="mixed")
    m.tracer(0, SIDE | CONV | ICE)
    gs = m.host_state()
    kmt = case.grid["kmt"]
    wet = (np.arange(case.km)[:, None, None] < kmt[None])[:, 2:-2, 2:-2]
    dT = np.abs(inner(gs["at"][0]) - inner(st["at"][0]))[wet].max()
    dS = np.abs(inner(gs["at"][1]) - inner(st["at"][1]))[wet].max() * 1000.0
    dq = np.abs(m.licomqice.cpu().numpy() - q).max()
    print(f"[mixed ice+conv] max|dT| {dT:.3e} K, max|dS| {dS:.3e} psu, max|dq| {dq:.3e}", flush=True)
    assert dT <= 1e-6 and dS <= 1e-6 and dq <= 1e-6
    m.close()


def test_icesnow_flag_needs_an_output_array(need_gpu):
    announce("LICOM_STEP_ICESNOW without licomqice")
    case = ls.make_case("C1")
    m = lt.TracerModel(case)
    lt.check(lt.lib().licom_tracer_set_icesnow(m.ctx, None), "detach")
    with pytest.raises(lt.LicomError):
        m.tracer(0, SIDE | ICE)
    m.close()


def test_full_size_fused_icesnow_equals_icesnow_after_the_step(need_gpu):
    """At the 0.1 deg workload size (3600x2302x55, del4, dt=600 s), with 2% of
    the surface planted below the freezing point: the step with
    LICOM_STEP_ICESNOW (no convadj) gives the same bits as the plain step
    followed by licom_icesnow (at, atb, licomqice), and T >= TBICE on every
    cell afterwards."""
    import torch
    announce("full size 3600x2302x55: fused icesnow vs step then licom_icesnow")
    case = ls.make_case("C3")
    rng = np.random.default_rng(26)
    T0 = case.state["at"][0][0]
    cold = rng.random(T0.shape) < 0.02
    T0[cold] = TBICE - rng.uniform(0.01, 1.2, int(cold.sum()))
    m = lt.TracerModel(case, diagnostics=False)
    case.state = None
    snap = {k: v.clone() for k, v in m.state.items()}
    m.tracer(0, SIDE | ICE)
    m.sync()
    a_at, a_atb, a_q = m.state["at"].clone(), m.state["atb"].clone(), m.licomqice.clone()
    for k, v in snap.items():
        m.state[k].copy_(v)
    del snap
    m.tracer(0, SIDE)
    m.icesnow(reset=True)
    m.sync()
    same = [torch.equal(a_at, m.state["at"]), torch.equal(a_atb, m.state["atb"]),
            torch.equal(a_q, m.licomqice)]
    frozen = int((a_q > 0).sum())
    tmin = float(a_at[0].min())
    print(f"[full size ice] same at/atb/licomqice {same}, {frozen} columns froze, min T {tmin:.6f}",
          flush=True)
    m.close()
    del a_at, a_atb, a_q, m
    torch.cuda.empty_cache()
    assert all(same) and frozen > 0 and tmin >= TBICE
