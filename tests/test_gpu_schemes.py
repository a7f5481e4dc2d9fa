"""GPU parity of the centered and flux advection schemes (advection_tracer_1,
tracer.cpp:120-289, selected by adv_tracer at tracer.cpp:1712-1714) against
the oracle, bit for bit: the fused k-march and the per-stage kernels, del2
and del4, with convadj and the side effects; the centered scheme's ax/ay/az
accumulate over the steps (divided by nss), the flux scheme leaves them
alone.  The mixed form and batches refuse both schemes."""
import numpy as np
import pytest

import licom_synth as ls
import licom_tracer as lt
import orc
from test_gpu_parity import announce, compare

pytestmark = pytest.mark.gpu

FLAGS = lt.LICOM_STEP_SIDE_EFFECTS | lt.LICOM_STEP_CONVADJ


def _run(case, steps, flags=FLAGS):
    o = orc.Oracle(case)
    st = case.copy_state()
    od = orc.Oracle.new_diag(case)
    m = lt.TracerModel(case)
    for s in range(steps):
        o.step(st, od, convadj=1)
        m.tracer(s, flags)
    compare(m, st, od, f"{case.name} x{steps}", exact=True, tol=1e-10)
    gd = m.host_diag()
    m.close()
    return st, od, gd


@pytest.mark.parametrize("adv", [ls.ADV_CENTERED, ls.ADV_FLUX], ids=["centered", "flux"])
@pytest.mark.parametrize("hmix", [2, 4], ids=["del2", "del4"])
@pytest.mark.parametrize("staged", [False, True], ids=["fused", "staged"])
def test_scheme_steps_bitwise(adv, hmix, staged, need_gpu):
    over = dict(depth_min=-3000.0, depth_max=5000.0, sigma=1.5, seed=11 + adv)
    if hmix == 4:
        over.update(hmix=4, ah=-1.0e10)
    case = ls.make_case("C1", **over)
    case.params["adv"] = adv
    case.params["nss"] = 3
    case.name = f"C1-land adv={adv} del{hmix} {'staged' if staged else 'fused'}"
    announce(f"scheme steps {case.name}")
    flags = FLAGS | (lt.LICOM_STEP_STAGED if staged else 0)
    _, od, gd = _run(case, 3, flags)
    ax = gd["ax"][..., 2:-2, 2:-2]
    if adv == ls.ADV_FLUX:
        assert not ax.any() and not gd["ay"].any() and not gd["az"].any()
    else:
        assert np.abs(ax).max() > 0


def test_centered_wide_grid_bitwise(need_gpu):
    """Tiles with partial edges and the 128-column solve rows."""
    case = ls.make_case("C1", nx=530, ny=26, km=12, seed=21, depth_min=-2500.0, sigma=2.0)
    case.params["adv"] = ls.ADV_CENTERED
    case.params["nss"] = 1
    case.name = "530x26x12 centered"
    announce(f"scheme steps {case.name}")
    _run(case, 2)


def test_flux_km55_bitwise(need_gpu):
    case = ls.make_case("C1", nx=100, ny=50, km=55, seed=23, hmix=4, ah=-1.0e10)
    case.params["adv"] = ls.ADV_FLUX
    case.name = "100x50x55 flux del4"
    announce(f"scheme steps {case.name}")
    _run(case, 2)


def test_mixed_form_and_batches_refuse_other_schemes(need_gpu):
    announce("mixed form with the centered scheme (refused)")
    case = ls.make_case("C1", nx=48, ny=24, km=12)
    case.params["adv"] = ls.ADV_CENTERED
    m = lt.TracerModel(case, precision="mixed")
    with pytest.raises(lt.LicomError, match="tspas only"):
        m.step()
    m.close()


def test_create_rejects_bad_scheme_arguments(need_gpu):
    announce("create with bad scheme arguments (refused)")
    case = ls.make_case("C1", nx=48, ny=24, km=12)
    case.params["adv"] = 7
    with pytest.raises(lt.LicomError, match="adv must be"):
        lt.TracerModel(case)
    case.params["adv"] = ls.ADV_CENTERED
    case.params["nss"] = 0
    with pytest.raises(lt.LicomError, match="nss"):
        lt.TracerModel(case)
