"""GPU tests of the batched ensemble engine, the launch-shape autotuner and
the request loop (mixed-precision form).

A member's arithmetic does not depend on the batch it runs in, on the launch
shapes, or on graph capture: every comparison here is bit for bit."""
import sys

import numpy as np
import pytest

import licom_ensemble as le
import licom_synth as ls
import licom_tracer as lt

pytestmark = pytest.mark.gpu

SHAPES = ((48, 32, 12), (64, 40, 30), (96, 64, 30))


def announce(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def cases():
    return [le.member_case(4100 + m, nx, ny, km) for m, (nx, ny, km) in enumerate(SHAPES)]


def outputs(model):
    st, dg = model.host_state(), model.host_diag()
    return {"at": st["at"], "atb": st["atb"], "ws": st["ws"], "akt": st["akt"], "utf": st["utf"],
            "tend": dg["tend"], "ax": dg["ax"], "dt_conv": dg["dt_conv"]}


def assert_same(a, b, tag):
    for k in a:
        n = int(np.count_nonzero(a[k] != b[k]))
        assert n == 0, f"{tag}: {k} differs in {n} entries"


def alone(case, steps):
    m = lt.TracerModel(case, precision="mixed")
    for s in range(steps):
        m.tracer(s, le.MIXED_FLAGS)
    out = outputs(m)
    m.close()
    return out


def test_batch_equals_members_alone(need_gpu):
    announce("ensemble: 3 members batched vs alone, 5 steps")
    cs = cases()
    ref = [alone(c, 5) for c in cs]
    models = [lt.TracerModel(c, precision="mixed") for c in cs]
    ens = le.Ensemble(models)
    ens.step(5)
    ens.sync()
    for k, m in enumerate(models):
        assert_same(outputs(m), ref[k], f"member {k}")
        m.close()
    ens.close()


def test_launch_shapes_do_not_change_results(need_gpu):
    announce("ensemble: every launch shape gives the same bits")
    c = cases()[1]
    ref = alone(c, 3)
    for fd in range(3):
        for sv in range(3):
            m = lt.TracerModel(c, precision="mixed")
            ens = le.Ensemble([m], tuning=[(fd, sv)], march_grid=7)
            ens.step(3)
            ens.sync()
            assert_same(outputs(m), ref, f"shape ({fd},{sv})")
            ens.close()
            m.close()


def test_graph_replays_match_eager(need_gpu):
    announce("ensemble: 4-step graph replayed 3 times vs eager")
    cs = cases()
    g_models = [lt.TracerModel(c, precision="mixed") for c in cs]
    e_models = [lt.TracerModel(c, precision="mixed") for c in cs]
    g = le.Ensemble(g_models)
    e = le.Ensemble(e_models)
    g.capture(4)
    for r in range(3):
        g.replay()
        g.sync()
        e.step(4)
        e.sync()
        lt.lib()            # eager host work between replays
        for k in range(len(cs)):
            assert_same(outputs(g_models[k]), outputs(e_models[k]), f"replay {r} member {k}")
    for x in (g, e):
        x.close()
    for m in g_models + e_models:
        m.close()


def test_autotuner_picks_a_valid_shape_and_caches(tmp_path, need_gpu):
    announce("autotuner on a 64x40x30 member")
    path = tmp_path / "tune.json"
    t = le.Autotuner(str(path), reps=2)
    c = cases()[1]
    fd, sv = t.tune(c)
    assert fd in (0, 1, 2) and sv in (0, 1, 2)
    t2 = le.Autotuner(str(path))
    assert t2.tune(c) == (fd, sv)            # from the cache, no timing
    assert t2.table[le.Autotuner.key(c)]["ms_per_step"] > 0


@pytest.mark.parametrize("single,policy", [((), "classes"), ((0,), "classes"), ((), "srpt")],
                         ids=["together", "one-at-a-time", "srpt"])
def test_request_loop_results_same_alone_and_in_the_mix(single, policy, need_gpu):
    announce(f"request loop ({policy}): 5 requests, continuous batching vs alone, one_at_a_time={single}")
    reqs = []
    for k, (nx, ny, km) in enumerate(SHAPES + ((48, 32, 12), (64, 40, 30))):
        c = le.member_case(4200 + k, nx, ny, km)
        # all at once, and long enough (a few ms each) that later admissions
        # find earlier requests still running
        reqs.append(le.Request(k, 0.0, c, 40 + 10 * k))
    loop = le.RequestLoop(hbm_cap_bytes=2 * le.member_bytes(96, 64, 30), chunk_steps=4,
                          keep_results=True, one_at_a_time=single, policy=policy)
    stats = loop.run(reqs)
    print(f"[loop] {stats}", flush=True)
    assert stats["requests"] == len(reqs)
    assert stats["peak_active"] >= 2
    for r in reqs:
        ref = alone(r.case, r.n_steps)["at"]
        assert np.array_equal(loop.results[r.rid], ref), f"request {r.rid}"


def test_batch_with_icesnow_equals_members_alone(need_gpu):
    """LICOM_STEP_ICESNOW in a batch: each member's state, diagnostics and
    licomqice equal its own single-model steps with the flag."""
    from helpers import plant_cold
    announce("ensemble: 3 members with icesnow batched vs alone, 3 steps")
    flags = le.MIXED_FLAGS | lt.LICOM_STEP_ICESNOW
    cs = cases()
    for k, c in enumerate(cs):
        plant_cold(c, c.state, 30 + k)
    ref, qref = [], []
    for c in cs:
        m = lt.TracerModel(c, precision="mixed")
        for s in range(3):
            m.tracer(s, flags)
        ref.append(outputs(m))
        qref.append(m.licomqice.cpu().numpy())
        m.close()
    models = [lt.TracerModel(c, precision="mixed") for c in cs]
    ens = le.Ensemble(models)
    ens.step(3, flags)
    ens.sync()
    for k, m in enumerate(models):
        assert_same(outputs(m), ref[k], f"member {k}")
        q = m.licomqice.cpu().numpy()
        assert np.array_equal(q, qref[k]), f"member {k}: licomqice"
        m.close()
    ens.close()
