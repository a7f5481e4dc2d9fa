"""GPU test of licom_tracer_step_noalloc, the FP64 step with its scratch in
a static device array (for callers that cannot allocate), against
licom_tracer_step on the same inputs: bit for bit, on two calls in a row
(the second on a smaller grid, so it reuses the array the first one left
behind), plus the size limit and the rejected flags.

Each test prints its case before touching the GPU."""
import ctypes
import sys

import numpy as np
import pytest

import licom_synth as ls
import licom_tracer as lt
from helpers import mismatch

pytestmark = pytest.mark.gpu

SIDE, CONV = lt.LICOM_STEP_SIDE_EFFECTS, lt.LICOM_STEP_CONVADJ


def announce(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def noalloc_step(m, flags):
    F, D = m._fields(), m._diag()
    lt.check(lt.lib().licom_tracer_step_noalloc(m.ctx, ctypes.byref(F), ctypes.byref(D), int(flags)),
             "licom_tracer_step_noalloc")


def outputs(m):
    st, dg = m.host_state(), m.host_diag()
    out = {k: st[k] for k in ("at", "atb", "ws", "utf", "vtf", "h0f", "akt", "vtl")}
    out.update({k: dg[k] for k in ("ax", "ay", "az", "dz", "tend", "dt_diff", "dt_conv", "penetrate",
                                   "net", "fw_norm2")})
    return out


def compare_pair(case, flags, steps, tag):
    ref = lt.TracerModel(case)
    got = lt.TracerModel(case)
    for s in range(steps):
        ref.tracer(s, flags)
        noalloc_step(got, flags)
    a, b = outputs(ref), outputs(got)
    ref.close()
    got.close()
    bad = {k: mismatch(a[k], b[k]) for k in a if mismatch(a[k], b[k])}
    print(f"[{tag}] differing arrays: {bad or 'none'}", flush=True)
    assert not bad, f"{tag}: {bad}"


def test_noalloc_equals_the_step_on_two_grids_in_a_row(need_gpu):
    announce("noalloc: C1 128x96x30 del2 x2, then 48x24x12 del4 x2 (smaller, reuses the array)")
    big = ls.make_case("C1", seed=41, depth_min=-2000.0, sigma=1.5)
    compare_pair(big, SIDE | CONV, 2, "noalloc 128x96x30")
    small = ls.make_case("C1", nx=48, ny=24, km=12, seed=42, hmix=4, ah=-1.0e10, depth_min=-2000.0,
                         sigma=1.5)
    compare_pair(small, SIDE | CONV, 2, "noalloc 48x24x12")


def test_noalloc_largest_member_and_vtl_split(need_gpu):
    announce("noalloc: 360x218x55 (largest ensemble member) with VTL_SCRATCH | CONVADJ")
    case = ls.make_case("C2", km=55, seed=43)
    compare_pair(case, SIDE | CONV | lt.LICOM_STEP_VTL_SCRATCH, 1, "noalloc 360x218x55 vtl")


def test_noalloc_rejects_large_grids_and_staged(need_gpu):
    announce("noalloc: size limit (400x260x55) and STAGED rejection")
    big = ls.make_case("C1", nx=400, ny=260, km=55, seed=44)
    m = lt.TracerModel(big, diagnostics=False)
    before = m.host_state()["at"]
    with pytest.raises(lt.LicomError, match="static buffer"):
        noalloc_step(m, SIDE | CONV)
    assert np.array_equal(m.host_state()["at"], before)      # nothing ran
    m.close()
    small = ls.make_case("C1", nx=16, ny=8, km=6, seed=45)
    m = lt.TracerModel(small)
    with pytest.raises(lt.LicomError, match="fused kernels only"):
        noalloc_step(m, SIDE | lt.LICOM_STEP_STAGED)
    m.close()
