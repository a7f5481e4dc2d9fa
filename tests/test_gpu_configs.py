"""GPU tests of the kernels and configurations the benchmark times, against
the oracle (FP64 bit for bit, the mixed form within its stated bounds) or,
for the batch engine and the request loop, against each member run alone.

* the 55-level FP64 column solve with convadj (the instantiation the
  0.1 deg benchmark runs), on the default and the wide-grid launch shape
* LICOM_STEP_VTL_SCRATCH on a convecting case: vtl holds tracer_'s S,
  before convadj (tracer.cpp:1880-1893, convadj.cpp)
* the mixed form on a grid that does not tile into the mixed march's 32x6
  tiles (partial tiles on both axes)
* the mixed form at the full 3600x2302x55 size: T/S content conservation and
  heat drift against the FP64 form
* the 64-member ensemble of the benchmark (domains and seeds of its
  generator, autotuned launch shapes, a 24-step graph replayed): every
  member bit-equal to its own single-model run; one member per shape
  against the oracle
* the request loop on Poisson arrivals with 24..240-step requests: every
  result bit-equal to the request run alone

Each test prints its case before touching the GPU."""
import sys

import numpy as np
import pytest

import licom_ensemble as le
import licom_synth as ls
import licom_tracer as lt
import orc
from helpers import inner, mismatch, report
from test_gpu_parity import compare, run_pair

pytestmark = pytest.mark.gpu

SIDE, CONV, VTL = lt.LICOM_STEP_SIDE_EFFECTS, lt.LICOM_STEP_CONVADJ, lt.LICOM_STEP_VTL_SCRATCH
PSU = 1000.0


def announce(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def wet(case):
    kmt = case.grid["kmt"]
    return (np.arange(case.km)[:, None, None] < kmt[None])[:, 2:-2, 2:-2]


def max_err(case, a, b):
    m = wet(case)
    dT = np.abs(inner(a[0]) - inner(b[0]))[m]
    dS = np.abs(inner(a[1]) - inner(b[1]))[m] * PSU
    return float(dT.max()), float(dS.max())


# ------------------------------------------------------------ FP64 km = 55
@pytest.mark.parametrize("nx,ny,tag", [(128, 96, "default launch"), (520, 24, "wide-grid launch")])
def test_km55_fp64_step_with_convadj_bitwise(nx, ny, tag, need_gpu):
    """k_prelude_col<55>, the k-march and k_solve<55, double, double> with
    convadj (no icesnow): the kernels of the 0.1 deg benchmark, with its
    del4 mixing and 600 s step."""
    announce(f"fp64 km55 del4 dt=600 x2, {nx}x{ny} ({tag})")
    case = ls.make_case("C1", nx=nx, ny=ny, km=55, seed=31, hmix=4, ah=-1.0e10, dts=600.0,
                        depth_min=-2000.0, sigma=1.5)
    case.name = f"km55-{nx}x{ny}"
    # the case must convect, or convadj is not exercised
    o = orc.Oracle(case)
    st = case.copy_state()
    o.step(st, convadj=0)
    st2 = {k: v.copy() for k, v in st.items()}
    o.convadj(st2)
    mixed_cols = int(np.count_nonzero((st2["at"] != st["at"]).any(axis=(0, 1))))
    print(f"[km55 {tag}] {mixed_cols} columns convect in the first step", flush=True)
    assert mixed_cols > 0
    run_pair(case, 2, tol=1e-10)


# ----------------------------------------------------------- vtl as scratch
@pytest.mark.parametrize("staged", [False, True], ids=["fused", "staged"])
def test_vtl_scratch_holds_the_unmixed_salinity(staged, need_gpu):
    announce(f"C1 VTL_SCRATCH | CONVADJ, convecting, {'staged' if staged else 'fused'}")
    case = ls.make_case("C1", nx=64, ny=48, km=30, seed=9, t_noise=1.0, s_noise=0.3)
    o = orc.Oracle(case)
    st = case.copy_state()
    od = orc.Oracle.new_diag(case)
    o.step(st, od, convadj=0)
    vtl_ref = st["at"][1].copy()        # tracer_'s S, halo filled (tracer.cpp:1891)
    o.convadj(st, od)
    changed = int(np.count_nonzero(st["at"][1] != vtl_ref))
    print(f"[vtl] convadj changes S at {changed} cells", flush=True)
    assert changed > 0
    flags = SIDE | CONV | VTL | (lt.LICOM_STEP_STAGED if staged else 0)
    m = lt.TracerModel(case)
    m.tracer(0, flags)
    compare(m, st, od, "vtl scratch", exact=True)
    got = m.host_state()["vtl"]
    print("[vtl] " + report("vtl", got, vtl_ref), flush=True)
    assert mismatch(got, vtl_ref) == 0
    m.close()


def test_vtl_scratch_without_convadj(need_gpu):
    announce("C1 VTL_SCRATCH without convadj")
    case = ls.make_case("C1", nx=64, ny=48, km=30, seed=10)
    o = orc.Oracle(case)
    st = case.copy_state()
    od = orc.Oracle.new_diag(case)
    o.step(st, od, convadj=0)
    m = lt.TracerModel(case)
    m.tracer(0, SIDE | VTL)
    compare(m, st, od, "vtl scratch, no convadj", exact=True)
    assert mismatch(m.host_state()["vtl"], st["at"][1]) == 0
    m.close()


# ---------------------------------------------------------- mixed, ragged
def test_mixed_partial_tiles_against_oracle(need_gpu):
    """100x50 columns: the mixed march's 32x6 tiles end in partial tiles on
    both axes (100 = 3*32 + 4, 50 = 8*6 + 2)."""
    announce("mixed 100x50x30 1 step vs oracle (partial tiles)")
    case = ls.make_case("C1", nx=100, ny=50, km=30, seed=14, depth_min=-1000.0, sigma=2.0)
    o = orc.Oracle(case)
    st = case.copy_state()
    od = orc.Oracle.new_diag(case)
    o.step(st, od, convadj=1)
    m = lt.TracerModel(case, precision="mixed")
    m.tracer(0, SIDE | CONV)
    gs, gd = m.host_state(), m.host_diag()
    m.close()
    eT, eS = max_err(case, gs["at"], st["at"])
    print(f"[mixed 100x50] max|dT| {eT:.3e} K, max|dS| {eS:.3e} psu", flush=True)
    assert eT <= 1e-6 and eS <= 1e-6
    # the tile edges: the last partial column and row of tiles
    for k in ("ax", "ay", "az", "tend"):
        a, b = inner(gd[k]), inner(od[k])
        scale = np.abs(b).max()
        err = np.abs(a - b).max() / scale
        edge = np.abs(a[..., :, 96:] - b[..., :, 96:]).max() / scale
        print(f"  {k}: rel {err:.3e} (last partial tile column {edge:.3e})", flush=True)
        assert err <= 1e-4, k


# --------------------------------------------------- mixed, full 0.1 deg
def test_full_size_mixed_conserves_content_and_tracks_fp64(need_gpu):
    """3600x2302x55, del4, dt = 600 s, two steps with convadj in both forms:
    the mixed form conserves T and S content to 1e-11 (compensated by the
    FP64 state and FP64 solve) and its heat content differs from the FP64
    form's by <= 1e-10 relative; max |dT| and max |dS| are reported."""
    import torch
    announce("full size 3600x2302x55: mixed vs fp64, 2 steps, content and heat")
    case = ls.make_case("C3")
    kmt = case.grid["kmt"]
    tarea = case.grid["tarea"]
    ref = lt.TracerModel(case, diagnostics=False)
    mix = lt.TracerModel(case, diagnostics=False, precision="mixed")
    case.state = None
    d = ref.derived()
    dev = ref.state["at"].device
    km = case.km
    vit = torch.arange(km, device=dev)[:, None, None] < torch.from_numpy(kmt).to(dev)[None]
    w = (torch.from_numpy(d["dzp"]).to(dev)[:, None, None] * torch.from_numpy(tarea).to(dev)[None]
         * vit)[:, 2:-2, 2:-2]
    del vit

    def content(m):
        m.sync()
        return [float((w * m.state["at"][n, :, 2:-2, 2:-2]).sum()) for n in range(2)]

    before = content(mix)
    for s in range(2):
        ref.tracer(s, SIDE | CONV)
        mix.tracer(s, SIDE | CONV)
    c_mix, c_ref = content(mix), content(ref)
    rel = [abs(a - b) / abs(b) for a, b in zip(c_mix, before)]
    drift = abs(c_mix[0] - c_ref[0]) / abs(c_ref[0])
    wetm = (torch.arange(km, device=dev)[:, None, None] < torch.from_numpy(kmt).to(dev)[None])
    dT = float(((mix.state["at"][0] - ref.state["at"][0]).abs() * wetm)[:, 2:-2, 2:-2].max())
    dS = float(((mix.state["at"][1] - ref.state["at"][1]).abs() * wetm)[:, 2:-2, 2:-2].max()) * PSU
    print(f"[full size mixed] content rel change T {rel[0]:.2e} S {rel[1]:.2e}; heat drift vs fp64 "
          f"{drift:.2e}; max|dT| {dT:.3e} K, max|dS| {dS:.3e} psu after 2 steps", flush=True)
    for m in (ref, mix):
        m.close()
    del w, wetm, ref, mix
    torch.cuda.empty_cache()
    assert rel[0] <= 1e-11 and rel[1] <= 1e-11
    assert drift <= 1e-10


# -------------------------------------------------- 64-member ensemble
def test_ensemble_of_64_graph_replay_equals_members_alone(tmp_path, need_gpu):
    """The benchmark's ensemble: 64 members from the workload generator
    (domains drawn with seed 4999, member m seeded 4000+m, T perturbed by
    N(0, 0.1)), launch shapes chosen by the autotuner, 24 steps captured in
    a HIP graph and replayed once.  Every member equals its own 24
    single-model steps bit for bit; one member of each shape is within the
    mixed form's one-step bound of the oracle."""
    announce("ensemble: 64 members, autotuned, 24-step graph vs alone")
    cases = le.ensemble_cases(64)
    tuner = le.Autotuner(str(tmp_path / "tune.json"), reps=2)
    tuning = [tuner.tune(c) for c in cases]
    models = [lt.TracerModel(c, precision="mixed") for c in cases]
    ens = le.Ensemble(models, tuning=tuning)
    ens.capture(24)
    ens.replay()
    ens.sync()
    got = [m.host_state()["at"] for m in models]
    for m in models:
        m.close()
    ens.close()
    del models
    shapes = {}
    for k, c in enumerate(cases):
        shapes.setdefault((c.nx, c.ny, c.km), k)
        m = lt.TracerModel(c, precision="mixed")
        for s in range(24):
            m.tracer(s, le.MIXED_FLAGS)
        ref = m.host_state()["at"]
        m.close()
        n = mismatch(got[k], ref)
        assert n == 0, f"member {k} ({c.nx}x{c.ny}x{c.km}): {n} entries differ from its run alone"
    print(f"[ensemble 64] every member bit-equal to its run alone; shapes {sorted(shapes)}",
          flush=True)
    for shape, k in sorted(shapes.items()):
        c = cases[k]
        o = orc.Oracle(c)
        st = c.copy_state()
        o.step(st, convadj=1)
        m = lt.TracerModel(c, precision="mixed")
        m.tracer(0, le.MIXED_FLAGS)
        eT, eS = max_err(c, m.host_state()["at"], st["at"])
        m.close()
        print(f"  member {k} {shape}: 1 step vs oracle max|dT| {eT:.3e} K, max|dS| {eS:.3e} psu",
              flush=True)
        assert eT <= 1e-6 and eS <= 1e-6, shape


# --------------------------------------------------------- request loop
@pytest.mark.parametrize("policy", ["classes", "srpt", "solo"])
def test_request_loop_poisson_requests_same_alone_and_in_the_mix(policy, need_gpu):
    """Requests of the serving workload (domains of the ensemble
    distribution, 24..240 steps, Poisson arrivals) served under a cap that
    holds only part of them at once: every result equals the request run
    alone, bit for bit."""
    reqs = le.poisson_requests(16, rate_per_s=40.0)
    cap = max(0.45 * sum(r.bytes for r in reqs), 1.2 * max(r.bytes for r in reqs))
    announce(f"request loop ({policy}): 16 Poisson requests, steps {min(r.n_steps for r in reqs)}.."
             f"{max(r.n_steps for r in reqs)}, cap {cap / 1e9:.2f} GB")
    loop = le.RequestLoop(hbm_cap_bytes=cap, keep_results=True, policy=policy)
    stats = loop.run(reqs)
    print(f"[loop poisson] {stats}", flush=True)
    assert stats["requests"] == len(reqs)
    assert stats["peak_active"] >= 2
    for r in reqs:
        m = lt.TracerModel(r.case, precision="mixed")
        for s in range(r.n_steps):
            m.tracer(s, le.MIXED_FLAGS)
        ref = m.host_state()["at"]
        m.close()
        n = mismatch(loop.results[r.rid], ref)
        assert n == 0, f"request {r.rid}: {n} entries differ from its run alone"
