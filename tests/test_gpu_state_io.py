"""GPU check of the state I/O path into the engine (the restart stream of
inirun.F90:598-700 read back into the engine's layout).  Prints its case
before touching the GPU."""
import sys

import numpy as np
import pytest

import licom_synth as ls
import licom_tracer as lt
from helpers import mismatch, report

pytestmark = pytest.mark.gpu

SIDE, CONV = lt.LICOM_STEP_SIDE_EFFECTS, lt.LICOM_STEP_CONVADJ


def announce(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def test_state_restored_from_a_restart_steps_to_the_same_bits(need_gpu, tmp_path):
    """State I/O into the engine: T/S written as a restart stream, read back
    and put into the engine's layout give the same GPU step as the original
    state (land holds 0 in both)."""
    announce("C1-land: restart round trip, then one step")
    case = ls.make_case("C1", depth_min=-3000.0, depth_max=5000.0, sigma=1.5, seed=11)
    kmt = case.grid["kmt"]
    wet = np.arange(case.km)[:, None, None] < kmt[None]
    case.state["at"] = np.where(wet[None], case.state["at"], 0.0)
    at_g = np.ascontiguousarray(case.state["at"][:, :, 2:-2, 2:-2])
    f = {k: np.zeros((case.ny, case.nx)) for k in lt.RESTART_2D}
    z3 = np.zeros((case.km, case.ny, case.nx))
    f.update(u=z3, v=z3, ws=z3, at=at_g)
    kg = np.ascontiguousarray(kmt[2:-2, 2:-2])
    lt.write_restart(str(tmp_path / "restart"), f, kg, kg, 1, 1)
    r = lt.read_restart(str(tmp_path / "restart"), case.nx, case.ny, case.km)
    a = lt.TracerModel(case)
    restored = ls.make_case("C1", depth_min=-3000.0, depth_max=5000.0, sigma=1.5, seed=11)
    restored.state["at"] = lt.global_to_halo(r["at"])
    b = lt.TracerModel(restored)
    for m in (a, b):
        m.tracer(0, SIDE | CONV)
    sa, sb = a.host_state(), b.host_state()
    for k in ("at", "atb"):
        print("[restart] " + report(k, sa[k], sb[k]), flush=True)
        assert mismatch(sa[k], sb[k]) == 0, k
    a.close()
    b.close()
