"""A member whose state goes non-finite is flagged, not fatal: in a batch
(per-stage and persistent launches) the member with a planted NaN reports
non-finite tracer columns while the other members stay bit-identical to
their runs alone and report none; the single-model FP64 and staged paths
count the same way, and the request loop marks such a request as failed
while serving the rest."""
import numpy as np
import pytest

import licom_ensemble as le
import licom_synth as ls
import licom_tracer as lt
from test_gpu_ensemble import alone, announce, assert_same, cases, outputs

pytestmark = pytest.mark.gpu


def plant_nan(case, n=0):
    j, i = case.jmt // 2, case.imt // 2
    case.state["at"][n, 3, j, i] = np.nan
    return case


@pytest.mark.parametrize("persistent", [False, True], ids=["per-stage", "persistent"])
def test_batch_member_with_nan_is_flagged_and_isolated(persistent, need_gpu):
    announce(f"non-finite member in a batch (persistent={persistent})")
    cs = cases()
    plant_nan(cs[1])
    ref = {k: alone(c, 3) for k, c in enumerate(cs) if k != 1}
    models = [lt.TracerModel(c, precision="mixed") for c in cs]
    ens = le.Ensemble(models, persistent=persistent)
    ens.step(3)
    counts = ens.nonfinite()
    print(f"  non-finite tracer columns per member: {counts}", flush=True)
    assert counts[1] > 0 and counts[0] == 0 and counts[2] == 0
    for k, m in enumerate(models):
        if k != 1:
            assert_same(outputs(m), ref[k], f"member {k}")
        m.close()
    ens.close()


@pytest.mark.parametrize("flags", [lt.LICOM_STEP_CONVADJ, lt.LICOM_STEP_CONVADJ | lt.LICOM_STEP_STAGED],
                         ids=["fused", "staged"])
def test_single_model_counts_nonfinite_columns(flags, need_gpu):
    announce(f"non-finite count of the FP64 step (flags {flags})")
    c = ls.make_case("C1", nx=48, ny=24, km=12, seed=3)
    m = lt.TracerModel(c)
    m.tracer(0, flags)
    assert m.nonfinite() == 0
    m.close()
    plant_nan(c, n=1)
    m = lt.TracerModel(c)
    m.tracer(0, flags)
    assert m.nonfinite() > 0
    m.close()


def test_request_loop_marks_the_nonfinite_request(need_gpu):
    announce("request loop with one non-finite request")
    reqs = []
    for k, (nx, ny, km) in enumerate(((48, 32, 12), (64, 40, 30), (48, 32, 12))):
        c = le.member_case(4300 + k, nx, ny, km)
        if k == 1:
            plant_nan(c)
        reqs.append(le.Request(k, 0.0, c, 10 + 5 * k))
    loop = le.RequestLoop(hbm_cap_bytes=1e10, keep_results=True, policy="srpt")
    stats = loop.run(reqs)
    print(f"[loop] {stats}", flush=True)
    assert stats["requests"] == 3 and stats["nonfinite"] == 1
    assert [r.rid for r in loop.failed] == [1] and "non-finite" in reqs[1].error
    for k in (0, 2):
        assert reqs[k].error is None
        assert np.array_equal(loop.results[k], alone(reqs[k].case, reqs[k].n_steps)["at"])
