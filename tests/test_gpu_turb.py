"""GPU parity of the Canuto producer (licom_turb_run) against the oracle on
the same synthetic tables and fields.

Bit for bit: the mixed-layer depth, the tidal term and the tidal
diagnostics (only +, -, *, / and sqrt on both sides, -ffp-contract=off),
and everything the call clears or leaves alone.  AKMT and AKT go through
atan, sin, cos, acosh and log, whose last bits differ between the device
and host math libraries, so they are compared within the path's 1e-12
relative tolerance (and most entries are still identical).

Each test prints the case it runs before touching the GPU."""
import sys

import numpy as np
import pytest

import licom_synth as ls
import licom_tracer as lt
import licom_turb as ltu
import orc

pytestmark = pytest.mark.gpu

EXACT = ("ak_tide", "amld", "richardson", "fztidal", "wp3_tidal")
CLOSE = ("akmt", "akt")


def announce(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def sentinel_outputs(tc, value=9.0):
    km, jmt, imt = tc.km, tc.jmt, tc.imt
    return dict(akmt=np.full((km, jmt, imt), value), akt=np.full((2, km, jmt, imt), value),
                ak_tide=np.full((km, jmt, imt), value), amld=np.full((jmt, imt), value),
                richardson=np.full((km, jmt, imt), value), fztidal=np.full((km, jmt, imt), value),
                wp3_tidal=np.full((km, jmt, imt), value))


def run_both(tc, value=9.0):
    ref, counts = orc.TurbOracle(tc.tables).run(tc, sentinel_outputs(tc, value))
    p = ltu.TurbProducer(tc)
    for k, t in p.outputs.items():
        t.fill_(value)
    p.run()
    out = p.host_outputs()
    p.close()
    return out, ref, counts


def check(out, ref, tag, rtol=1e-12):
    msgs, bad = [], []
    for k in EXACT:
        n = int(np.count_nonzero(out[k] != ref[k]))
        msgs.append(f"{k}: {n} mismatches")
        if n:
            bad.append(k)
    for k in CLOSE:
        a, b = out[k], ref[k]
        d = np.abs(a - b)
        lim = rtol * np.maximum(np.abs(a), np.abs(b))
        n_over = int(np.count_nonzero(d > lim))
        n_diff = int(np.count_nonzero(a != b))
        msgs.append(f"{k}: {n_diff} differ, {n_over} beyond {rtol:g} relative, "
                    f"max rel {float(np.max(d / np.maximum(np.abs(b), 1e-300))):.2e}")
        if n_over:
            bad.append(k)
    print(f"[{tag}] " + "; ".join(msgs), flush=True)
    assert not bad, f"{tag}: {bad}: " + "; ".join(msgs)


def test_turb_c1_equals_oracle(need_gpu):
    announce("turb: C1 grid 132x100x30, synthetic tables, seed 7")
    c = ls.make_case("C1")
    tc = ls.make_turb_case(c, seed=7)
    out, ref, counts = run_both(tc)
    check(out, ref, "turb C1")
    for k in ("pure", "table2d", "below", "deep", "epsy", "deardorff", "rafglt", "theta_p",
              "theta_n"):
        assert counts[k] > 0, (k, counts)


def test_turb_km55_equals_oracle(need_gpu):
    announce("turb: 100x60x55 (the 0.1-degree column depth), seed 9")
    c = ls.make_case("C1", nx=100, ny=60, km=55, seed=9)
    tc = ls.make_turb_case(c, seed=9)
    out, ref, _ = run_both(tc, value=-3.0)
    check(out, ref, "turb km55")


def test_turb_early_exits_equal_oracle(need_gpu):
    announce("turb: background table with negative slq2 on an angle band (early exits)")
    tb = ls.turb_tables()
    th = (np.arange(len(tb["slq2_r1"])) - ltu.NTH) * tb["deltheta_r"]
    tb["slq2_r1"] = np.where((th > 0.2) & (th < 0.6), -1.0, tb["slq2_r1"])
    tb["sm_r1"] = np.where((th > 1.0) & (th < 1.2), 0.0, tb["sm_r1"])   # zero background
    c = ls.make_case("C1", nx=64, ny=48, km=30, seed=4)
    tc = ls.make_turb_case(c, seed=4, tables=tb)
    out, ref, counts = run_both(tc)
    assert counts["exit"] > 20, counts
    check(out, ref, "turb exits")


def test_turb_writes_akt_of_a_tracer_model_then_steps(need_gpu):
    """AKT produced on the device feeds the tracer step directly (the
    reference's readyc -> tracer order); the step matches the oracle's step
    on the oracle's AKT within the path's 1e-12 relative tolerance."""
    announce("turb -> tracer step: C1, AKT from the producer, one FP64 step vs oracle")
    c = ls.make_case("C1")
    tc = ls.make_turb_case(c, seed=3)
    ref, _ = orc.TurbOracle(tc.tables).run(tc, sentinel_outputs(tc, 0.0))
    # AKT at the bottom level is never written; both start from the case's AKT
    akt0 = c.state["akt"].copy()
    ref_akt = ref["akt"].copy()
    ref_akt[:, -1] = akt0[:, -1]
    wet = np.zeros((c.jmt, c.imt), bool)
    wet[1:-1, 1:-1] = c.grid["kmt"][1:-1, 1:-1] > 0
    akt_o = np.where(wet[None, None], ref_akt, akt0)
    m = lt.TracerModel(c)
    p = ltu.TurbProducer(tc)
    p.run(akt=m.state["akt"])
    p.sync()
    got_akt = m.state["akt"].cpu().numpy()
    d = np.abs(got_akt - akt_o)
    assert np.all(d <= 1e-12 * np.abs(akt_o)), float(np.max(d))
    m.tracer(0)
    m.sync()
    st = c.copy_state()
    st["akt"] = akt_o
    o = orc.Oracle(c)
    diag = o.new_diag(c)
    o.step(st, diag, convadj=0)
    got = m.host_state()
    for k in ("at", "atb"):
        a, b = got[k], st[k]
        err = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
        print(f"[turb->step] {k}: rel err {err:.2e}", flush=True)
        assert err <= 1e-12, (k, err)
    m.close()
    p.close()


def test_turb_rejects_bad_tables_and_extents(need_gpu):
    announce("turb: argument checks")
    c = ls.make_case("C1", nx=16, ny=12, km=10, seed=2)
    tc = ls.make_turb_case(c, seed=2)
    bad = dict(tc.tables)
    bad["rib"] = bad["rib"][::-1].copy()
    tc_bad = ls.TurbCase(imt=tc.imt, jmt=tc.jmt, km=tc.km, grid=tc.grid, tables=bad,
                         fields=tc.fields, params=tc.params)
    with pytest.raises(lt.LicomError, match="increase"):
        ltu.TurbProducer(tc_bad)
    tc_ncc = ls.TurbCase(imt=tc.imt, jmt=tc.jmt, km=tc.km, grid=tc.grid, tables=tc.tables,
                         fields=tc.fields, params=dict(tc.params, ncc=0))
    with pytest.raises(lt.LicomError, match="ncc"):
        ltu.TurbProducer(tc_ncc)
