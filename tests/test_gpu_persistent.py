"""GPU tests of the persistent batch step (licom_batch_set_persistent): one
launch per step with grid-wide barriers between the stages gives the same
bits as each member alone (and so as the per-stage launches), eagerly, with
icesnow and replayed from a captured graph; a barrier timeout, forced from
the host through the test hook, is reported by licom_batch_sync and leaves
the batch usable."""
import numpy as np
import pytest

import licom_ensemble as le
import licom_tracer as lt
from helpers import plant_cold
from test_gpu_ensemble import alone, announce, assert_same, cases, outputs

pytestmark = pytest.mark.gpu


def test_persistent_steps_equal_members_alone(need_gpu):
    announce("persistent batch: 3 members, 5 eager steps vs alone")
    cs = cases()
    ref = [alone(c, 5) for c in cs]
    models = [lt.TracerModel(c, precision="mixed") for c in cs]
    ens = le.Ensemble(models, persistent=True)
    assert ens.persistent_grid() > 0
    ens.step(5)
    ens.sync()
    for k, m in enumerate(models):
        assert_same(outputs(m), ref[k], f"member {k}")
        m.close()
    ens.close()


def test_persistent_graph_replays_equal_members_alone(need_gpu):
    announce("persistent batch: 4-step graph replayed twice vs 8 steps alone")
    cs = cases()
    ref = [alone(c, 8) for c in cs]
    models = [lt.TracerModel(c, precision="mixed") for c in cs]
    ens = le.Ensemble(models, persistent=True)
    ens.capture(4)
    for _ in range(2):
        ens.replay()
        ens.sync()
        lt.lib()            # eager host work between replays
    for k, m in enumerate(models):
        assert_same(outputs(m), ref[k], f"member {k}")
        m.close()
    ens.close()


def test_persistent_with_icesnow_equals_members_alone(need_gpu):
    announce("persistent batch: 3 members with icesnow, 3 steps vs alone")
    flags = le.MIXED_FLAGS | lt.LICOM_STEP_ICESNOW
    cs = cases()
    for k, c in enumerate(cs):
        plant_cold(c, c.state, 40 + k)
    ref, qref = [], []
    for c in cs:
        m = lt.TracerModel(c, precision="mixed")
        for s in range(3):
            m.tracer(s, flags)
        ref.append(outputs(m))
        qref.append(m.licomqice.cpu().numpy())
        m.close()
    models = [lt.TracerModel(c, precision="mixed") for c in cs]
    ens = le.Ensemble(models, persistent=True)
    ens.step(3, flags)
    ens.sync()
    for k, m in enumerate(models):
        assert_same(outputs(m), ref[k], f"member {k}")
        assert np.array_equal(m.licomqice.cpu().numpy(), qref[k]), f"member {k}: licomqice"
        m.close()
    ens.close()


def test_persistent_barrier_timeout_is_reported(need_gpu):
    announce("persistent batch: a barrier timeout forced by the host is reported")
    c = cases()[0]
    m = lt.TracerModel(c, precision="mixed")
    ens = le.Ensemble([m], persistent=True)
    lt.check(le._lib().licom_batch_fail_next_barrier(ens.handle), "fail hook")
    ens.step(1)
    with pytest.raises(lt.LicomError, match="timed out"):
        ens.sync()
    ens.step(1)             # the barrier state was reset by the failed sync
    ens.sync()
    ens.close()
    m.close()
