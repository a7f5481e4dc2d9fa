#!/usr/bin/env python3
"""Benchmark: LICOM3 tracer step (T and S, tspas + del4 + implicit vertical
diffusion + convadj) on one MI355X at 3600 x 2302 x 55 (0.1 deg).

Prints ONE JSON line.  `value` is tracer cell-updates per second
(nx*ny*km*2 per step, land counted) with every input resident in HBM.
`roofline` is the dominant kernel's algorithmic HBM bytes per launch over its
average launch time (HIP events on the library's stream, timed region);
`step_roofline` is the same for the whole step with the path's 128 B per
tracer cell-update (FP64 drop-in, diagnostics on).  `cpu_baseline` times the
CPU restatement (oracle/, OpenMP) on a band of the same workload on this
host.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--variant fp64]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "licom3-hip_amd"))

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, vendor peak
STEP_BYTES_PER_UPDATE = {"fp64": 128.0, "mixed": 80.0}


def kernel_algorithmic_bytes(name, case, variant="fp64"):
    """Compulsory HBM bytes of one launch of `name` on `case`: every array
    the kernel must read or write, counted once, over the cells it covers.
    Per interior cell (nx*ny*km), both tracers; 2-D terms neglected."""
    cells = case.nx * case.ny * case.km
    del4 = case.params["hmix"] == 4
    if variant == "mixed":
        per_cell = {
            # utf vtf utl vtl (f32) -> u_wface v_sface ws (f32); the divergence
            # profile stays in registers
            "k_prelude": 4 * 4 + 4 * 3,
            # at atb (f64 x2) akt (f32 x2) faces ws (f32) -> ori (f64 x2), ax ay az dz (f32 x2), pen
            "k_march": 8 * 4 + 4 * 2 + 4 * 3 + 8 * 2 + 4 * 8 + 4,
            # ori at (f64 x2) akt (f32 x2) -> at atb (f64 x2) tend dt_diff dt_conv (f32 x2)
            "k_solve": 8 * 4 + 4 * 2 + 8 * 4 + 4 * 6,
            # utf vtf (f32) read and written
            "k_side": 4 * 4,
        }
        return per_cell.get(name, 0) * cells
    per_cell = {
        # fused path
        # utf vtf utl vtl -> u_wface v_sface ws; the divergence profile stays
        # in registers
        "k_prelude": 8 * 4 + 8 * 3,
        # at atb akt (x2) u_wface v_sface ws -> ori ax ay az dz (x2), penetrate
        "k_march": 8 * (6 + 3) + 8 * (10 + 1),
        # ori akt at (x2) -> at atb tend dt_diff dt_conv (x2)
        "k_solve": 8 * 6 + 8 * 10,
        # utf vtf read and written
        "k_side": 8 * 4,
        # per-stage path
        # at(2) atb(2) akt(2) uw vs ws flags(2x1B) [d2tk(2)] -> ori(2) ax ay az dz (x2) pen
        "k_correct": 8 * (2 + 2 + 2 + 2 + 1) + 2 + (16 if del4 else 0) + 8 * (2 + 8 + 1),
        # ori(2) akt(2) at(2) utf vtf -> at(2) atb(2) tend dt_diff dt_conv (x2) utf vtf
        "k_column": 8 * (2 + 2 + 2 + 2) + 8 * (2 + 2 + 6 + 2),
        "k_predict": 8 * (2 + 2 + 1) + 2,
        "k_faces": 8 * 4 + 8 * 2,
        "k_ws": 8 * 4 + 8,
        "k_d2tk": 8 * 2 + 8 * 2,
    }
    return per_cell.get(name, 0) * cells


COPY_BYTES = 4 << 30            # size of the calibration copy (read and written)


def traffic_file(workload, variant="fp64"):
    tag = "tracer_step" if variant == "fp64" else "tracer_step_mixed"
    return os.path.join(ROOT, "profiles", f"hbm_traffic_{tag}_{workload}.json")


def traffic_from_profiles(kernel, workload, variant="fp64"):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary of this workload, or None."""
    p = traffic_file(workload, variant)
    if not os.path.exists(p):
        return None
    with open(p) as f:
        k = json.load(f).get("kernels", {}).get(kernel)
    return None if k is None else k.get("hbm_bytes_per_launch")


def short_kernel_name(name):
    """'void licom::(anonymous namespace)::k_solve<55>(KGrid, ...)' -> 'k_solve'."""
    n = name.replace("(anonymous namespace)", "")
    n = n.split("(")[0]
    n = n.split("<")[0]
    return n.split("::")[-1].split()[-1]


def read_rocprof_csv(directory, suffix):
    """Rows of every rocprofv3 CSV under `directory` whose name ends in `suffix`."""
    import csv
    rows = []
    for dp, _, files in os.walk(directory):
        for fn in sorted(files):
            if fn.endswith(suffix):
                with open(os.path.join(dp, fn), newline="") as f:
                    rows.extend(csv.DictReader(f))
    if not rows:
        raise SystemExit(f"no *{suffix} under {directory}")
    return rows


def summarize_pmc(fetch_dir, write_dir, workload, commands, variant="fp64"):
    """Per-kernel HBM traffic from two rocprofv3 --pmc runs (FETCH_SIZE, then
    WRITE_SIZE; KiB per dispatch), each scaled by what the same counter reads
    for k_stream_copy moving COPY_BYTES each way.  Writes the summary json."""
    def per_kernel(d, counter):
        acc = {}
        for r in read_rocprof_csv(d, "counter_collection.csv"):
            if r.get("Counter_Name") != counter:
                continue
            k = short_kernel_name(r["Kernel_Name"])
            if k.endswith("_mx"):      # the mixed form's kernels, named as the FP64 ones
                k = k[:-3]
            acc.setdefault(k, {})
            acc[k].setdefault(r.get("Dispatch_Id", r.get("Correlation_Id")), 0.0)
            acc[k][r.get("Dispatch_Id", r.get("Correlation_Id"))] += float(r["Counter_Value"])
        return {k: sum(v.values()) / len(v) for k, v in acc.items()}
    fetch, write = per_kernel(fetch_dir, "FETCH_SIZE"), per_kernel(write_dir, "WRITE_SIZE")
    cf = COPY_BYTES / (fetch["k_stream_copy"] * 1024.0)
    cw = COPY_BYTES / (write["k_stream_copy"] * 1024.0)
    kernels = {}
    for k in sorted(set(fetch) & set(write)):
        b = fetch[k] * 1024.0 * cf + write[k] * 1024.0 * cw
        kernels[k] = {"fetch_kib": round(fetch[k], 1), "write_kib": round(write[k], 1),
                      "hbm_bytes_per_launch": round(b)}
    out = {"workload": workload, "commands": commands,
           "calibration": {"kernel": "k_stream_copy", "bytes_each_way": COPY_BYTES,
                           "fetch_factor": round(cf, 4), "write_factor": round(cw, 4)},
           "kernels": kernels}
    with open(traffic_file(workload, variant), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


def trace_file(workload, variant="fp64"):
    tag = "tracer_step" if variant == "fp64" else "tracer_step_mixed"
    return os.path.join(ROOT, "profiles", f"kernel_trace_{tag}_{workload}.md")


def summarize_trace(trace_dir, workload, command, variant="fp64"):
    """Markdown table of a rocprofv3 --kernel-trace --stats run."""
    rows = read_rocprof_csv(trace_dir, "kernel_stats.csv")
    form = "FP64 form" if variant == "fp64" else "mixed-precision form"
    lines = [f"# Kernel times, tracer step ({form}), {workload}", "", "Command (repo root):", "",
             "```", command, "```", "",
             "| kernel | calls | avg ms | min ms | max ms | share % |", "|---|---|---|---|---|---|"]
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"])):
        lines.append(f"| {short_kernel_name(r['Name'])} | {r['Calls']} | "
                     f"{float(r['AverageNs']) / 1e6:.3f} | {float(r['MinNs']) / 1e6:.3f} | "
                     f"{float(r['MaxNs']) / 1e6:.3f} | {float(r['Percentage']):.1f} |")
    with open(trace_file(workload, variant), "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def write_read_bandwidth(torch, nbytes, reps=5):
    """HBM GB/s of a write-only pass (fill) and of a read-only pass (sum) over
    `nbytes`: the tracer step writes ~95 GB per step at 0.1 deg, so the write
    rate bounds it as much as the copy rate does."""
    x = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda")
    x.fill_(1.0)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        x.fill_(2.0)
    e1.record()
    torch.cuda.synchronize()
    w = nbytes / (e0.elapsed_time(e1) / reps * 1e-3) / 1e9
    y = x.sum()                 # first call sets up the reduction
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        y = x.sum()
    e1.record()
    torch.cuda.synchronize()
    r = nbytes / (e0.elapsed_time(e1) / reps * 1e-3) / 1e9
    del x, y
    return w, r


def cpu_baseline(args):
    """CPU restatement (oracle/, C + OpenMP) on bands of the same workload
    (3600 x ny_band x 55 Mercator rows around the equator, del4, dt=600),
    median of timed steps after one warm-up step: on every OpenMP thread of
    this process (the host's CPU share), then on one thread over a narrower
    band."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import licom_synth as ls
    import orc

    def timed(band, seconds, nmin, nmax):
        case = ls.make_case("C3", ny=band)
        o = orc.Oracle(case)
        st = case.copy_state()
        dg = orc.Oracle.new_diag(case)
        o.step(st, dg)
        times = []
        t_end = time.time() + seconds
        while len(times) < nmin or (time.time() < t_end and len(times) < nmax):
            t0 = time.perf_counter()
            o.step(st, dg)
            times.append(time.perf_counter() - t0)
        times.sort()
        upd = case.nx * case.ny * case.km * 2
        return upd / times[len(times) // 2], len(times), case

    threads = orc.lib().orc_nthreads()
    v_all, n_all, case = timed(args.cpu_band, args.cpu_seconds, 5, 20)
    orc.lib().orc_set_nthreads(1)
    band1 = max(4, args.cpu_band // 8)
    v_one, n_one, case1 = timed(band1, args.cpu_seconds / 2, 3, 10)
    orc.lib().orc_set_nthreads(threads)
    # the Canuto producer's restatement on a band of the same workload
    tcase = ls.make_case("C3", ny=args.cpu_band)
    tc = ls.make_turb_case(tcase, seed=17)
    to = orc.TurbOracle(tc.tables)
    to.run(tc)
    tt = []
    t_end = time.time() + args.cpu_seconds / 2
    while len(tt) < 3 or (time.time() < t_end and len(tt) < 10):
        t0 = time.perf_counter()
        to.run(tc)
        tt.append(time.perf_counter() - t0)
    tt.sort()
    turb_cpu = {"value": tcase.nx * tcase.ny * tcase.km / tt[len(tt) // 2],
                "unit": "column cells/s", "cores": threads, "kind": "port",
                "sample": f"oracle turb_cu restatement on a {tcase.nx}x{tcase.ny}x{tcase.km} band, "
                          f"median of {len(tt)} runs"}
    return {"value": v_all, "unit": "tracer cell-updates/s", "cores": threads, "turb": turb_cpu,
            "kind": "port",
            "sample": f"oracle (C, OpenMP, {threads} threads) on a {case.nx}x{case.ny}x{case.km} band "
                      f"of the 0.1 deg workload, median of {n_all} steps after 1 warm-up",
            "one_core": {"value": v_one, "cores": 1,
                         "sample": f"the same on 1 thread, {case1.nx}x{case1.ny}x{case1.km} band, "
                                   f"median of {n_one} steps"},
            "host_cpus": os.cpu_count(),
            "threads_note": "OpenMP threads = OMP_NUM_THREADS (the host CPU share given to this job)"}


def run_icesnow(model, args, torch, lt, case, plain):
    """icesnow_ at the workload size: the stand-alone licom_icesnow (every
    column; algorithmic bytes: T at every level, kmt, surface S, licomqice
    written) and the step with LICOM_STEP_ICESNOW against the plain step
    (`plain`: the timed FP64 run)."""
    n2 = case.imt * case.jmt
    alg = n2 * (8 * case.km + 4 + 8 + 8)
    model.icesnow(reset=True)
    model.sync()
    model.kernel_times()
    model.set_profiling(True)
    reps = 10
    for _ in range(reps):
        model.icesnow(reset=True)
    kt = model.kernel_times()
    ms = kt["k_icesnow"][0] / kt["k_icesnow"][1]
    flags = lt.LICOM_STEP_SIDE_EFFECTS | lt.LICOM_STEP_CONVADJ | lt.LICOM_STEP_ICESNOW
    model.tracer(0, flags)
    model.sync()
    model.kernel_times()
    steps = max(2, args.steps // 2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        model.tracer(0, flags)
    model.sync()
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / steps
    kf = model.kernel_times()
    model.set_profiling(False)
    out = {"standalone_ms": round(ms, 4), "algorithmic_bytes": alg,
           "standalone_alg_gbs": round(alg / (ms * 1e-3) / 1e9, 1),
           "fused_step_ms": round(1000 * sec, 3), "plain_step_ms": round(plain["ms_per_step"], 3)}
    if "k_solve" in kf and "k_solve" in plain["kernel_ms"]:
        out["fused_k_solve_ms"] = round(kf["k_solve"][0] / kf["k_solve"][1], 4)
        out["plain_k_solve_ms"] = plain["kernel_ms"]["k_solve"]
    print(f"[bench] icesnow: {ms:.3f} ms alone, step {1000 * sec:.2f} ms fused", file=sys.stderr, flush=True)
    return out


def turb_algorithmic_bytes(kmt, km):
    """Bytes the Canuto producer must move once (each array read or written
    once): a wet interior column with n = kmt-1 closure levels reads at,
    pdensity and vit at n+1 levels and rit, ricdttms, s2t, rict, fz_tide at
    n, plus kmt, fcort, buoytur, buoysol, wave_dis; it writes akmt and
    ak_tide at km levels, akt at 2(km-1), richardson, fztidal, wp3_tidal at
    n and amld.  Any other column reads vit and kmt and writes akmt, ak_tide
    and amld."""
    import numpy as np
    wet = np.zeros(kmt.shape, bool)
    wet[1:-1, 1:-1] = kmt[1:-1, 1:-1] > 0
    n = (kmt[wet].astype(np.int64) - 1)
    per_wet = 8 * (3 * (n + 1) + 5 * n + 2 * km + 2 * (km - 1) + 3 * n + 1) + 4 + 8 * 4
    dry = int((~wet).sum())
    return int(per_wet.sum()) + dry * (8 * (2 * km + 1) + 8 + 4)


def run_turb(args, torch, lt, case, model):
    """The Canuto producer (turb_cu) at the workload size: alone, and ahead
    of the FP64 tracer step writing the model's AKT (readyc's turb_cu, then
    tracer_, as stepon orders them).  Fields are drawn on the device."""
    import licom_synth as ls
    import licom_turb as ltu
    dev = model.device
    t0 = time.time()
    tc = ls.make_turb_case_device(case, torch, dev, seed=17)
    p = ltu.TurbProducer(tc, device=dev, stream=model.stream)
    print(f"[bench] turb: inputs drawn in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    alg = turb_algorithmic_bytes(case.grid["kmt"], case.km)
    for _ in range(max(1, args.warmup)):
        p.run()
    p.sync()
    reps = max(3, args.steps)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(reps)]
    for a, b in ev:
        a.record(model.stream)
        p.run()
        b.record(model.stream)
    p.sync()
    ms = sorted(a.elapsed_time(b) for a, b in ev)
    avg = sum(ms) / len(ms)
    flags = lt.LICOM_STEP_SIDE_EFFECTS | lt.LICOM_STEP_CONVADJ
    p.run(akt=model.state["akt"])
    model.tracer(0, flags)
    model.sync()
    steps = max(2, args.steps // 2)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(steps):
        p.run(akt=model.state["akt"])
        model.tracer(0, flags)
    model.sync()
    torch.cuda.synchronize()
    fused = (time.perf_counter() - t1) / steps
    p.close()
    del p, tc
    torch.cuda.empty_cache()
    cells = case.nx * case.ny * case.km
    ach = alg / (avg * 1e-3) / 1e9
    print(f"[bench] turb: {avg:.3f} ms alone, turb + step {1000 * fused:.2f} ms", file=sys.stderr,
          flush=True)
    return {"kernel": "k_turb", "avg_launch_ms": round(avg, 4), "min_launch_ms": round(ms[0], 4),
            "launches": reps, "column_cells_per_s": cells / (avg * 1e-3),
            "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                         "traffic": traffic_from_profiles("k_turb", args.workload),
                         "algorithmic_bytes": alg},
            "turb_plus_fp64_step_ms": round(1000 * fused, 3),
            "tables": "synthetic Canuto tables on the reference grids (2925 x 2925 Ri x Ri_d)"}


def readyt_algorithmic_bytes(case):
    """Bytes readyt_ must move once per call, every cell of the block
    (halos included, as the reference computes them): reads vit, viv, at and
    atb of both tracers, u, v, utf, vtf (10 x 8 B); writes utl, vtl, utf,
    vtf, dlu, dlv, pdensity, gg, pp, ppa, ppb, ppc, alpha, beta, akt of
    both tracers (16 x 8 B) and ric, rict, ricdttms, ricdt, rit on km-1
    planes; the 2-D fields are under 1%."""
    n2 = case.imt * case.jmt
    return n2 * case.km * 8 * (10 + 16) + n2 * (case.km - 1) * 8 * 5


def run_readyc(args, torch, case, rgrid, rt_fields, st, model):
    """readyc_'s shear pass (readyc1 + readyc1_2) on the readyt_ outputs and
    the FP64 model's velocities, then its momentum tendencies with each
    viscosity (del4, del2): ms per call and their HBM rooflines.
    Algorithmic bytes: reads vit, viv, up, vp (km planes) and rict, ric (km-1
    planes); writes wp12, wp13, riv1, riv2, s2t, s2u, riu (km planes) and rit
    (km-1 planes), 8 bytes each."""
    import licom_readyc as lrc
    import licom_synth as ls
    km, jmt, imt = case.km, case.jmt, case.imt
    cg = ls.readyc_grid(case, rgrid)
    f64 = dict(dtype=torch.float64, device=model.device)
    fields = dict(rt_fields, up=st["utf"], vp=st["vtf"], h0=torch.zeros((jmt, imt), **f64),
                  h0bl=torch.zeros((jmt, imt), **f64), h0bf=torch.zeros((jmt, imt), **f64))
    rc = lrc.Readyc(imt, jmt, km, cg, fields, device=model.device, stream=model.stream)
    rc.shear()
    rc.sync()
    ms = []
    for _ in range(max(3, args.steps)):
        rc.shear()
        ms.append(rc.last_ms())
    avg = sum(ms) / len(ms)
    n2 = imt * jmt
    alg = n2 * km * 8 * (4 + 7) + n2 * (km - 1) * 8 * 3
    ach = alg / (avg * 1e-3) / 1e9
    print(f"[bench] readyc shear: {avg:.3f} ms per call ({ach:.0f} GB/s algorithmic)", file=sys.stderr,
          flush=True)
    out = {"kernel": "k_readyc_shear", "ms_per_call": round(avg, 4), "calls": len(ms),
           "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(ach / HBM_PEAK_GBS, 4), "algorithmic_bytes": alg}}
    # the momentum tendencies with the biharmonic viscosity, on the same
    # context: velocities of the FP64 model, a random AKMU, ws from level 0
    mom = {}
    for del4 in (1, 0):
        rc.momentum_grid(ls.readyc_mgrid(case, rgrid, del4=del4))
        gen = torch.Generator(device=model.device).manual_seed(3)
        mf = rc.momentum_fields(dict(u=st["utf"] * 1.01, v=st["vtf"] * 0.99, up=st["utf"], vp=st["vtf"],
                                     vit=rt_fields["vit"], viv=rt_fields["viv"],
                                     akmu=torch.rand((km, jmt, imt), generator=gen, **f64) * 1.0e-2,
                                     su=torch.full((jmt, imt), 0.1, **f64),
                                     sv=torch.full((jmt, imt), -0.05, **f64)))
        rc.momentum(mf)
        rc.sync()
        mms = []
        for _ in range(max(3, args.steps)):
            rc.momentum(mf)
            mms.append(rc.last_ms())
        del mf
        mavg = sum(mms) / len(mms)
        malg = readyc_momentum_algorithmic_bytes(case)
        mach = malg / (mavg * 1e-3) / 1e9
        name = "del4" if del4 else "del2"
        print(f"[bench] readyc momentum ({name}): {mavg:.3f} ms per call ({mach:.0f} GB/s algorithmic)",
              file=sys.stderr, flush=True)
        mom[name] = {"ms_per_call": round(mavg, 4), "calls": len(mms),
                     "roofline": {"bound": "hbm", "achieved": round(mach, 1), "peak": HBM_PEAK_GBS,
                                  "unit": "GB/s", "frac": round(mach / HBM_PEAK_GBS, 4),
                                  "algorithmic_bytes": malg}}
    out["momentum"] = mom
    rc.close()
    del rc, fields
    torch.cuda.empty_cache()
    return out


def bclinc_algorithmic_bytes(case):
    """bclinc_'s compulsory HBM bytes for a leapfrog step (its three stages):
    stage 0 reads vit, gg, up, vp, dlu, dlv, viv, akmu and writes dlu, dlv,
    wka; stage 1 reads wka, v, vp, viv, up, dlu, akmu and writes vp, v, wka;
    stage 2 reads wka, u, up, viv, utf, vtf, v and writes up, u, utf, vtf
    (8 bytes each); the 2-D fields."""
    n2, km = case.imt * case.jmt, case.km
    return n2 * km * 8 * (22 + 10) + n2 * 8 * (10 + 6)


def run_bclinc(args, torch, case, rgrid, rt_fields, st, model):
    """bclinc_ at the workload size: one leapfrog step (stages 0, 1, 2 with
    no halo exchange between, a single block) and one first step (stages 0,
    1), on the FP64 model's velocities and synthetic buoyancy, viscosity and
    forcing built on the device: ms per call and the HBM roofline."""
    import licom_bclinc as lb
    import licom_synth as ls
    km, jmt, imt = case.km, case.jmt, case.imt
    dev = model.device
    f64 = dict(dtype=torch.float64, device=dev)
    gen = torch.Generator(device=dev).manual_seed(5)
    rnd = lambda shape, scale: torch.randn(shape, generator=gen, **f64) * scale  # noqa: E731
    s2, s3 = (jmt, imt), (km, jmt, imt)
    viv, vit = rt_fields["viv"], rt_fields["vit"]
    fields = dict(vit=vit, viv=viv, gg=(rnd(s3, 2.0e-3) - 1.0e-2) * vit,
                  akmu=torch.rand(s3, generator=gen, **f64) * 5.0e-2 * viv,
                  su=rnd(s2, 0.1), sv=rnd(s2, 0.1), psa=rnd(s2, 5.0) + 1.0e3, h0bl=rnd(s2, 0.3),
                  ub=rnd(s2, 0.05), vb=rnd(s2, 0.05), h0bf=rnd(s2, 1.0), u=st["utf"] * viv,
                  v=st["vtf"] * viv, up=st["utf"] * viv, vp=st["vtf"] * viv, dlu=rnd(s3, 1.0e-6) * viv,
                  dlv=rnd(s3, 1.0e-6) * viv)
    b = lb.Bclinc(imt, jmt, km, ls.bclinc_grid(case, rgrid), fields, device=dev, stream=model.stream)
    b.step(2)
    b.sync()
    out = {}
    for name, isc, stages in (("leapfrog", 2, (0, 1, 2)), ("first_step", 0, (0, 1))):
        ms, per = [], [[] for _ in stages]
        for _ in range(max(3, args.steps)):
            tot = 0.0
            for n, s in enumerate(stages):
                b.stage(isc, s)
                t = b.last_ms()
                per[n].append(t)
                tot += t
            ms.append(tot)
        avg = sum(ms) / len(ms)
        entry = {"ms_per_call": round(avg, 4), "calls": len(ms),
                 "stage_ms": [round(sum(p) / len(p), 4) for p in per]}
        if name == "leapfrog":
            alg = bclinc_algorithmic_bytes(case)
            ach = alg / (avg * 1e-3) / 1e9
            entry["roofline"] = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                                 "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "algorithmic_bytes": alg}
        print(f"[bench] bclinc {name}: {avg:.3f} ms per call (stages {entry['stage_ms']})", file=sys.stderr,
              flush=True)
        out[name] = entry
    b.close()
    del b, fields
    torch.cuda.empty_cache()
    return out


def barotr_algorithmic_bytes(case):
    """One barotr_ subcycle's compulsory HBM bytes over its 2-D fields: 34
    reads and 16 writes of 8 bytes per cell (stage 0: h0, ub, vb, dzph, vit
    -> work, wka2, wka3; stage 1: h0p, work -> h0, then h0, ubp, vbp, viv and
    the nine forcing fields -> work, wka2, wka3; stage 2: ubp, vbp, wka2,
    wka3, h0, dzph, h0p, vit -> ub, vb, work; stage 3: h0p, work, ub, vb, h0f,
    h0bf -> h0p, ubp, vbp, h0f, h0bf, h0).  The grid metrics are not
    counted."""
    return case.imt * case.jmt * 8 * (34 + 16)


def run_barotr(args, torch, case, rgrid, rt_fields, model):
    """barotr_'s subcycle at the workload size with the biharmonic
    viscosity: stages 0-3 of subcycles 0..K-1 (every fourth with the surface
    diffusion), no halo exchange between (a single block), synthetic 2-D
    fields on the device: ms per subcycle and per stage."""
    import licom_barotr as lbt
    import licom_synth as ls
    jmt, imt = case.jmt, case.imt
    dev = model.device
    f64 = dict(dtype=torch.float64, device=dev)
    gen = torch.Generator(device=dev).manual_seed(7)
    rnd = lambda scale: torch.randn((jmt, imt), generator=gen, **f64) * scale  # noqa: E731
    vit, viv = rt_fields["vit"][0].contiguous(), rt_fields["viv"][0].contiguous()
    fields = dict(vit=vit, viv=viv, dzph=(3000.0 + rnd(100.0)) * viv, dlub=rnd(1e-6), dlvb=rnd(1e-6),
                  wgp=torch.full((jmt, imt), 0.95, **f64), pax=rnd(1e-7), pay=rnd(1e-7), pxb=rnd(1e-6),
                  pyb=rnd(1e-6), whx=rnd(1e-5), why=rnd(1e-5), h0=rnd(0.5) * vit, h0p=rnd(0.5) * vit,
                  ub=rnd(0.1) * viv, vb=rnd(0.1) * viv, ubp=rnd(0.1) * viv, vbp=rnd(0.1) * viv,
                  h0f=torch.zeros((jmt, imt), **f64), h0bf=torch.zeros((jmt, imt), **f64))
    b = lbt.Barotr(imt, jmt, ls.barotr_grid(case, rgrid, del4=True), fields, device=dev, stream=model.stream)
    b.run(1)
    b.sync()
    reps = max(4, args.steps)
    per = [[] for _ in range(4)]
    tot = []
    for nc in range(reps):
        t = 0.0
        for s in range(4):
            b.stage(nc, s)
            ms = b.last_ms()
            per[s].append(ms)
            t += ms
        tot.append(t)
    b.close()
    del b, fields
    torch.cuda.empty_cache()
    avg = sum(tot) / len(tot)
    alg = barotr_algorithmic_bytes(case)
    ach = alg / (avg * 1e-3) / 1e9
    stage_ms = [round(sum(p) / len(p), 4) for p in per]
    print(f"[bench] barotr subcycle: {avg:.3f} ms (stages {stage_ms})", file=sys.stderr, flush=True)
    return {"ms_per_subcycle": round(avg, 4), "subcycles": reps, "stage_ms": stage_ms,
            "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(ach / HBM_PEAK_GBS, 4), "algorithmic_bytes": alg}}


def readyc_momentum_algorithmic_bytes(case):
    """licom_readyc_momentum's compulsory HBM bytes: reads u, v, up, vp,
    vit, viv, akmu and ws level 0; writes ws (levels 1..km-1), dlu, dlv,
    wka (8 bytes each); the 2-D inputs and outputs."""
    n2, km = case.imt * case.jmt, case.km
    return n2 * km * 8 * (7 + 4) + n2 * 8 * (3 + 6)


def run_readyt(args, torch, lt, case, model):
    """readyt_ at the workload size on the FP64 model's at / atb, the other
    fields built on the device; its CPU baseline is the oracle on a band."""
    import licom_readyt as lr
    import licom_synth as ls
    dev = model.device
    km, jmt, imt = case.km, case.jmt, case.imt
    g = ls.readyt_grid(case)
    f64 = dict(dtype=torch.float64, device=dev)
    lev = torch.arange(km, device=dev)[:, None, None]
    kmt = torch.from_numpy(case.grid["kmt"]).to(dev)
    kmu = torch.from_numpy(g["kmu"]).to(dev)
    st = model.state
    f = dict(vit=(lev < kmt[None]).to(torch.float64), viv=(lev < kmu[None]).to(torch.float64),
             at=st["at"], atb=st["atb"], h0=torch.zeros((jmt, imt), **f64),
             u=st["utf"] * 1.01, v=st["vtf"] * 0.99,
             psa=torch.full((jmt, imt), 100.0, **f64), swv=torch.full((jmt, imt), 1.0e-5, **f64),
             nswv=torch.full((jmt, imt), -1.0e-5, **f64),
             h0l=torch.zeros((jmt, imt), **f64), h0f=torch.zeros((jmt, imt), **f64),
             utl=st["utl"].clone(), vtl=st["vtl"].clone(), utf=st["utf"].clone(), vtf=st["vtf"].clone())
    r = lr.Readyt(imt, jmt, km, g, f, device=dev, stream=model.stream)
    for _ in range(max(1, args.warmup)):
        r.readyt(1)
    r.sync()
    reps = max(3, args.steps)
    tot, parts = [], []
    for _ in range(reps):
        r.readyt(1)
        ms = r.last_ms()
        tot.append(ms[0])
        parts.append(ms[1:])
    keep = {k: r.fields[k] for k in ("vit", "viv", "rict", "ric")}
    r.close()
    del r, f
    torch.cuda.empty_cache()
    avg = sum(tot) / len(tot)
    part = [sum(p[i] for p in parts) / len(parts) for i in range(3)]
    alg = readyt_algorithmic_bytes(case)
    shear = run_readyc(args, torch, case, g, keep, st, model)
    bcl = run_bclinc(args, torch, case, g, keep, st, model)
    bcl["barotr"] = run_barotr(args, torch, case, g, keep, model)
    ach = alg / (avg * 1e-3) / 1e9
    print(f"[bench] readyt: {avg:.3f} ms per call ({ach:.0f} GB/s algorithmic)", file=sys.stderr,
          flush=True)
    out = {"ms_per_call": round(avg, 4), "calls": reps,
           "kernel_ms": {"k_readyt_col": round(part[0], 4), "k_readyt_pg": round(part[1], 4),
                         "k_readyt_sfc": round(part[2], 4)},
           "cells_per_s": imt * jmt * km / (avg * 1e-3),
           "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                        "traffic": traffic_from_profiles("k_readyt_col", args.workload),
                        "algorithmic_bytes": alg},
           "readyc_shear": shear, "bclinc": bcl}
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import orc
        band = ls.make_case(args.workload, ny=min(64, case.ny))
        gb = ls.readyt_grid(band)
        fb = ls.readyt_fields(band, gb)
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            ob = orc.readyt(band.imt, band.jmt, band.km, gb, fb, ist=1)
            times.append(time.perf_counter() - t0)
        cells = band.imt * band.jmt * band.km
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        shape = f"{band.imt}x{band.jmt}x{band.km}"
        out["cpu_baseline"] = {"value": cells / sorted(times)[1], "unit": "cells/s", "cores": cores,
                               "kind": "port",
                               "sample": f"oracle readyt_ restatement on a {shape} block, median of 3 calls"}
        out["cpu_baseline_rest"] = stepon_cpu_baselines(orc, ls, band, gb, fb, ob, cores)
    return out


def stepon_cpu_baselines(orc, ls, band, gb, fb, ob, cores):
    """The oracle's readyc_ (shear, momentum with del4), bclinc_ (a
    leapfrog step) and barotr_ (one subcycle, del4) on the same band as
    readyt_'s CPU baseline, median of 3 calls each, in cells (columns x
    levels; barotr_: columns) per second.  Their loops are partly OpenMP."""
    imt, jmt, km = band.imt, band.jmt, band.km
    cg, cf = ls.readyc_inputs(band, gb, ob)
    akmu = orc.readyc_akmu(imt, jmt, km, cg, np.full((km, jmt, imt), 1.0e-3), cf["viv"])
    mg = ls.readyc_mgrid(band, gb, del4=1)
    mf = ls.readyc_minputs(band, ob, cf, akmu)
    bg, bfl = ls.bclinc_grid(band, gb), ls.bclinc_fields(band, fb)
    tg, tf = ls.barotr_grid(band, gb, del4=True), ls.barotr_fields(band, fb)
    calls = {"readyc_shear": (lambda: orc.readyc_shear(imt, jmt, km, cg, cf), imt * jmt * km),
             "readyc_momentum_del4": (lambda: orc.readyc_momentum(imt, jmt, km, cg, mg, mf), imt * jmt * km),
             "bclinc_leapfrog": (lambda: orc.bclinc_stage(imt, jmt, km, bg, bfl, 2, (0, 1, 2)), imt * jmt * km),
             "barotr_subcycle": (lambda: orc.barotr_stage(imt, jmt, tg, tf, 0, (0, 1, 2, 3)), imt * jmt)}
    res = {}
    for name, (fn, cells) in calls.items():
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            fn()
            times.append(time.perf_counter() - t0)
        res[name] = {"value": cells / sorted(times)[1], "unit": "cells/s" if name != "barotr_subcycle"
                     else "columns/s", "ms": round(sorted(times)[1] * 1e3, 2)}
    return {"kind": "port", "cores": cores, "sample": f"the oracle on a {imt}x{jmt}x{km} band, median of 3",
            "calls": res}


def run_variant(model, args, torch, lt, case, updates):
    """Warm-up, then K timed steps of one precision form; per-kernel HIP-event
    times on the library's stream over K further steps."""
    flags = lt.LICOM_STEP_SIDE_EFFECTS | lt.LICOM_STEP_CONVADJ
    v = model.precision
    for _ in range(args.warmup):
        model.tracer(0, flags)
    model.sync()
    torch.cuda.synchronize()
    print(f"[bench] {v}: warm-up done, timing {args.steps} steps", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        model.tracer(0, flags)
    model.sync()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    kt = {}
    if not args.no_kernel_timing:
        # per-kernel HIP-event times over as many further steps; with them
        # on, the step runs every kernel in order on the library's stream
        model.kernel_times()       # drain
        model.set_profiling(True)
        for _ in range(args.steps):
            model.tracer(0, flags)
        model.sync()
        kt = model.kernel_times()
        model.set_profiling(False)
    sec = dt / args.steps
    roof = None
    if kt:
        dom = max(kt, key=lambda k: kt[k][0])
        avg_ms = kt[dom][0] / kt[dom][1]
        alg = kernel_algorithmic_bytes(dom, case, v)
        ach = alg / (avg_ms * 1e-3) / 1e9
        tr = traffic_from_profiles(dom, args.workload, v)
        roof = {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                "traffic": tr, "algorithmic_bytes": alg, "avg_launch_ms": round(avg_ms, 4)}
    sb = STEP_BYTES_PER_UPDATE[v]
    step_roof = {"bytes_per_update": sb, "achieved": round(sb * updates / sec / 1e9, 1),
                 "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": round(sb * updates / sec / 1e9 / HBM_PEAK_GBS, 4)}
    print(f"[bench] {v}: {1000 * sec:.2f} ms/step", file=sys.stderr, flush=True)
    # algorithmic GB/s of every kernel with a per-cell byte count
    kg = {}
    for k, t in kt.items():
        alg = kernel_algorithmic_bytes(k, case, v)
        if alg:
            kg[k] = round(alg / (t[0] / t[1] * 1e-3) / 1e9, 1)
    return {"value": updates / sec, "ms_per_step": 1000.0 * sec, "roofline": roof,
            "step_roofline": step_roof, "kernel_ms": {k: round(t[0] / t[1], 4) for k, t in kt.items()},
            "kernel_alg_gbs": kg}


def run_ensemble(args, torch, lt, le):
    """The 64-member ensemble workload (mixed form): members drawn from
    {96x64, 128x96, 256x192, 360x218} x km {30, 55}, launch shapes from the
    autotuner, 24 steps captured in a HIP graph and replayed."""
    t0 = time.time()
    cases = le.ensemble_cases(args.members)
    # launch shapes tuned afresh in every run (args.autotune_cache keeps them)
    tuner = le.Autotuner(args.autotune_cache, reps=5)
    tuning = [tuner.tune(c) for c in cases]
    print(f"[bench] ensemble: {len(cases)} members generated and tuned in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    models = [lt.TracerModel(c, diagnostics=True, precision="mixed") for c in cases]
    reps = max(1, args.ensemble_replays)

    def timed(persistent):
        # 24 steps captured in a graph: one persistent launch per step, or
        # one launch per stage and step
        ens = le.Ensemble(models, tuning=tuning, persistent=persistent)
        ups_ = ens.cell_updates_per_step()
        grid = ens.persistent_grid()
        ens.capture(24)
        ens.replay()               # warm-up
        ens.sync()
        t1 = time.perf_counter()
        for _ in range(reps):
            ens.replay()
        ens.sync()
        dt_ = time.perf_counter() - t1
        ens.close()
        return ups_, dt_, grid

    ups, dt, pgrid = timed(True)
    _, dt_stages, _ = timed(False)
    # isolated single-member step time per shape (for the serving slowdown)
    iso = {}
    for c in cases:
        key = (c.nx, c.ny, c.km)
        if key in iso:
            continue
        m = lt.TracerModel(c, diagnostics=True, precision="mixed")
        e1 = le.Ensemble([m], tuning=[tuner.tune(c)])
        e1.step(2)
        e1.sync()
        t2 = time.perf_counter()
        e1.step(10)
        e1.sync()
        iso[key] = (time.perf_counter() - t2) / 10
        e1.close()
        m.close()
    for m in models:
        m.close()
    torch.cuda.empty_cache()
    value = ups * 24 * reps / dt
    sb = STEP_BYTES_PER_UPDATE["mixed"]
    return {"members": len(cases), "value": value, "unit": "tracer cell-updates/s",
            "ms_per_step": 1000.0 * dt / (24 * reps), "graph_steps": 24, "replays": reps,
            "launches_per_step": 1, "persistent_workgroups": pgrid,
            "per_stage_launches": {"ms_per_step": 1000.0 * dt_stages / (24 * reps),
                                   "value": ups * 24 * reps / dt_stages, "launches_per_step": 9},
            "step_roofline": {"bytes_per_update": sb, "achieved": round(sb * value / 1e9, 1),
                              "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": round(sb * value / 1e9 / HBM_PEAK_GBS, 4)},
            "cell_updates_per_step": ups, "dtype": "f64/f32",
            "tuning": {k: v for k, v in tuner.table.items()},
            "isolated_step_ms": {f"{a}x{b}x{c}": round(1000 * t, 4) for (a, b, c), t in iso.items()}}, iso


def mixed_error_study(lt, ls, steps=100):
    """Error of the mixed form against the FP64 drop-in on the 1 deg grid
    (360x218x30, del2, dt=3600 s) after `steps` steps, beside the FP64 drop-in's
    own divergence from an initial T perturbed by N(0, 1e-8 K) on wet cells:
    the tspas limiter and convadj make discrete decisions at near-ties, so a
    perturbation of one mixed step's size flips some of them."""
    case = ls.make_case("C2")
    pcase = ls.make_case("C2")
    rng = np.random.default_rng(99)
    kmt = case.grid["kmt"]
    wet3 = np.arange(case.km)[:, None, None] < kmt[None]
    noise = 1e-8 * rng.standard_normal(pcase.state["at"][0].shape) * wet3
    pcase.state["at"][0] += noise
    pcase.state["atb"][0, 1:] += noise
    flags = lt.LICOM_STEP_SIDE_EFFECTS | lt.LICOM_STEP_CONVADJ
    models = [lt.TracerModel(case), lt.TracerModel(case, precision="mixed"), lt.TracerModel(pcase)]
    # tspas limiter decisions of every step: cells whose flag differs from the
    # FP64 run's at the same step (the two runs start from the same state;
    # from the second step on they also see slightly different states)
    import torch
    fl = [m.limiter_flags() for m in models]
    wet_t = torch.from_numpy(np.broadcast_to(wet3, (2,) + wet3.shape).copy()).to(fl[0].device)
    wet_t[:, :, :2] = False
    wet_t[:, :, -2:] = False
    wet_t[:, :, :, :2] = False
    wet_t[:, :, :, -2:] = False
    flips_mix, flips_ctl, flagged = [], [], []
    for st in range(steps):
        for m in models:
            m.tracer(st, flags)
        torch.cuda.synchronize()
        flips_mix.append(int(((fl[1] != fl[0]) & wet_t).sum()))
        flips_ctl.append(int(((fl[2] != fl[0]) & wet_t).sum()))
        flagged.append(int(((fl[0] != 0) & wet_t).sum()))
    nwet = int(wet_t.sum())
    a_ref, a_mix, a_ctl = [m.host_state()["at"] for m in models]
    for m in models:
        m.close()
    wet = wet3[:, 2:-2, 2:-2]

    def err(a):
        dT = np.abs(a[0] - a_ref[0])[:, 2:-2, 2:-2][wet]
        dS = np.abs(a[1] - a_ref[1])[:, 2:-2, 2:-2][wet] * 1000.0
        return {"max_dT_K": float(dT.max()), "rms_dT_K": float(np.sqrt(np.mean(dT * dT))),
                "max_dS_psu": float(dS.max()), "rms_dS_psu": float(np.sqrt(np.mean(dS * dS))),
                "wet_fraction_dT_above_1e-4": float(np.mean(dT > 1e-4))}

    dz = np.abs(np.diff(case.grid["zkp"]))[:case.km]
    w = dz[:, None, None] * case.grid["tarea"][None, 2:-2, 2:-2] * wet

    def heat(a):
        return float(np.sum(w * a[0][:, 2:-2, 2:-2]))

    h = heat(a_ref)

    def flips(f):
        return {"first_step": f[0], "mean_per_step": float(np.mean(f)), "max_per_step": int(max(f)),
                "total": int(sum(f))}

    em = err(a_mix)
    return {"grid": f"{case.nx}x{case.ny}x{case.km}", "steps": steps,
            "mixed_vs_fp64": dict(em, heat_drift_rel=abs(heat(a_mix) - h) / abs(h),
                                  limiter_flag_mismatches=flips(flips_mix)),
            "fp64_perturbed_1e-8K_vs_fp64": dict(err(a_ctl), heat_drift_rel=abs(heat(a_ctl) - h) / abs(h),
                                                 limiter_flag_mismatches=flips(flips_ctl)),
            "wet_tracer_cells": nwet, "fp64_flagged_cells_per_step": float(np.mean(flagged)),
            "target_max_dT_1e-4K_met": bool(em["max_dT_K"] <= 1e-4)}


def run_serving(args, torch, lt, le, offline_value, iso, load):
    """Request loop: Poisson arrivals at a fixed fraction `load` of the
    offline batched capacity, HBM admission under the cap, continuous
    batching.  Slowdown of a request = its latency / (its steps x the
    isolated step time of its member shape)."""
    mean_updates = np.mean([st * nx * ny * km * 2 for nx, ny, km, st in le.request_specs(args.requests)])
    rate = load * offline_value / mean_updates
    reqs = le.poisson_requests(args.requests, rate)
    print(f"[bench] serving: {len(reqs)} requests at {rate:.2f}/s (load {load})", file=sys.stderr,
          flush=True)
    loop = le.RequestLoop(hbm_cap_bytes=args.hbm_cap_gb * 1e9, chunk_steps=args.chunk_steps,
                          policy=args.serve_policy, lanes=args.serve_lanes)
    stats = loop.run(reqs)
    iso_t = lambda r: r.n_steps * iso.get((r.case.nx, r.case.ny, r.case.km), np.nan)  # noqa: E731
    served = [r for r in reqs if r.finished is not None]
    slow_of = {r.rid: (r.finished - r.arrival) / iso_t(r) for r in served}
    slow = np.array([v for v in slow_of.values() if np.isfinite(v)])
    # where the tail's time goes: the slowest requests' wait for admission
    # (context creation and upload) and their run in their class
    worst = sorted(served, key=lambda r: -slow_of[r.rid])[:4]
    stats["slowest"] = [{"shape": f"{r.case.nx}x{r.case.ny}x{r.case.km}", "steps": r.n_steps,
                         "isolated_s": round(iso_t(r), 5), "admission_s": round(r.admitted - r.arrival, 5),
                         "run_s": round(r.finished - r.admitted, 5), "slowdown": round(slow_of[r.rid], 2),
                         "admission_phases_s": {k: round(v, 5) for k, v in r.timing.items()}}
                        for r in worst]
    last_arrival = max(r.arrival for r in reqs)
    stats.update({"policy": args.serve_policy, "arrival_rate_per_s": rate, "load_fraction": load,
                  "hbm_cap_gb": args.hbm_cap_gb,
                  "drain_after_last_arrival_s": stats["makespan_s"] - last_arrival,
                  "vs_offline": stats["cell_updates_per_s"] / offline_value,
                  "p50_slowdown_vs_isolated": float(np.percentile(slow, 50)) if slow.size else None,
                  "p99_slowdown_vs_isolated": float(np.percentile(slow, 99)) if slow.size else None})
    return stats


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--variant", default="both", choices=["fp64", "mixed", "both"],
                    help="both: the FP64 drop-in is the headline value, the mixed form "
                         "and its error against FP64 are reported beside it")
    ap.add_argument("--workload", default="C3")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-band", type=int, default=128)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-kernel-timing", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-error-study", action="store_true")
    ap.add_argument("--no-icesnow", action="store_true")
    ap.add_argument("--no-turb", action="store_true")
    ap.add_argument("--no-readyt", action="store_true")
    ap.add_argument("--members", type=int, default=64)
    ap.add_argument("--autotune-cache", default=None,
                    help="JSON file to read and write the ensemble's tuned launch shapes "
                         "(default: none, shapes are tuned in every run)")
    ap.add_argument("--ensemble-replays", type=int, default=3)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--requests", type=int, default=64)
    ap.add_argument("--loads", default="0.25,0.8",
                    help="request arrival rates as fractions of the offline batched capacity")
    ap.add_argument("--chunk-steps", type=int, default=4)
    ap.add_argument("--serve-policy", default="solo,srpt,classes",
                    help="request loop scheduling, comma-separated: every request alone on one "
                         "of --serve-lanes streams, least remaining work first (solo), least "
                         "remaining work first over one batch (srpt), size classes advancing in "
                         "lock step (classes)")
    ap.add_argument("--serve-lanes", type=int, default=4,
                    help="streams of the solo policy")
    ap.add_argument("--hbm-cap-gb", type=float, default=200.0)
    ap.add_argument("--hw-queues", type=int, default=16,
                    help="GPU_MAX_HW_QUEUES for this process (0: leave the environment's): the "
                         "request loop's lanes and admission streams each want a queue of their own "
                         "(with HIP's default of 4 an admission's copies wait behind other streams' "
                         "kernels)")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="write the HBM traffic summary from two rocprofv3 --pmc runs")
    ap.add_argument("--summarize-trace", metavar="TRACE_DIR",
                    help="write the kernel-time table from a rocprofv3 --stats run")
    ap.add_argument("--command", action="append", default=[],
                    help="profiling command(s) recorded in the summary")
    args = ap.parse_args()
    if args.summarize_pmc:
        summarize_pmc(*args.summarize_pmc, args.workload, args.command,
                      "mixed" if args.variant == "mixed" else "fp64")
        return
    if args.summarize_trace:
        summarize_trace(args.summarize_trace, args.workload, " && ".join(args.command),
                        "mixed" if args.variant == "mixed" else "fp64")
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the tracer step would shard by 2-D horizontal tiles with 2-cell halos)",
              file=sys.stderr)
        sys.exit(2)

    if args.hw_queues:
        if not 1 <= args.hw_queues <= 32:
            print("bench.py: --hw-queues must be within 1..32", file=sys.stderr)
            sys.exit(2)
        os.environ["GPU_MAX_HW_QUEUES"] = str(args.hw_queues)
    import torch
    import licom_synth as ls
    import licom_tracer as lt

    copy_gbs = lt.stream_copy_bandwidth(COPY_BYTES)
    fill_gbs, read_gbs = write_read_bandwidth(torch, COPY_BYTES)
    t0 = time.time()
    case = ls.make_case(args.workload)
    print(f"[bench] generated {args.workload} {case.nx}x{case.ny}x{case.km} in "
          f"{time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    variants = ["fp64", "mixed"] if args.variant == "both" else [args.variant]
    models = {v: lt.TracerModel(case, diagnostics=True, precision=v) for v in variants}
    case.state = None          # free host copies
    updates = case.nx * case.ny * case.km * 2
    runs = {v: run_variant(models[v], args, torch, lt, case, updates) for v in variants}

    mixed_err = None
    if len(variants) == 2:
        # max |dT|, |dS| of the mixed form against the FP64 drop-in (bit-identical
        # to the oracle) after the same warmup + steps, over wet interior cells
        ref, mix = models["fp64"].state["at"], models["mixed"].state["at"]
        kmt = torch.from_numpy(case.grid["kmt"]).to(ref.device)
        wet = (torch.arange(case.km, device=ref.device)[:, None, None] < kmt[None])[:, 2:-2, 2:-2]
        err, rms, frac = [], [], []
        nwet = float(wet.sum())
        for n in range(2):
            d = (mix[n, :, 2:-2, 2:-2] - ref[n, :, 2:-2, 2:-2]).abs()
            d = torch.where(wet, d, torch.zeros_like(d))
            scale = 1.0 if n == 0 else 1000.0          # S is stored as (S - 35) / 1000
            err.append(float(d.max()) * scale)
            rms.append(float((d * d).sum() / nwet) ** 0.5 * scale)
            frac.append(float((d * scale > 1e-4).sum()) / nwet)
            del d
        mixed_err = {"max_dT_K": err[0], "max_dS_psu": err[1], "rms_dT_K": rms[0],
                     "rms_dS_psu": rms[1], "wet_fraction_dT_above_1e-4": frac[0],
                     "wet_fraction_dS_above_1e-4": frac[1],
                     "steps": args.warmup + args.steps, "against": "FP64 drop-in"}
    if "mixed" in models:
        models.pop("mixed").close()
        torch.cuda.empty_cache()
    ice = None
    if "fp64" in models and not args.no_icesnow and not args.no_kernel_timing:
        ice = run_icesnow(models["fp64"], args, torch, lt, case, runs["fp64"])
    turb = None
    if "fp64" in models and not args.no_turb:
        turb = run_turb(args, torch, lt, case, models["fp64"])
    readyt = None
    if "fp64" in models and not args.no_readyt:
        readyt = run_readyt(args, torch, lt, case, models["fp64"])
    for m in models.values():
        m.close()
    del models
    torch.cuda.empty_cache()

    err_study = None
    if "mixed" in variants and not args.no_error_study:
        print("[bench] mixed-form error study (360x218x30, 100 steps)", file=sys.stderr, flush=True)
        err_study = mixed_error_study(lt, ls)

    ens_out, serve_out = None, None
    if not args.no_ensemble:
        import licom_ensemble as le
        ens_out, iso = run_ensemble(args, torch, lt, le)
        if not args.no_serving:
            serve_out = []
            for pol in args.serve_policy.split(","):
                args_p = argparse.Namespace(**vars(args))
                args_p.serve_policy = pol
                serve_out += [run_serving(args_p, torch, lt, le, ens_out["value"], iso, float(x))
                              for x in args.loads.split(",")]

    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args)

    main_v = variants[0]
    r = runs[main_v]
    out = {
        "metric": "tracer cell-updates/s at 3600x2302x55, FP64 vs mixed; p50/p99 request latency",
        "value": r["value"], "unit": "tracer cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if main_v == "fp64" else "f64/f32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload} {case.nx}x{case.ny}x{case.km} tracer step "
                               f"(tspas, {'del4' if case.params['hmix'] == 4 else 'del2'}, "
                               f"implicit vertical diffusion, convadj), "
                               f"{'FP64 drop-in' if main_v == 'fp64' else 'mixed precision'}, "
                               f"diagnostics on",
                   "nx": case.nx, "ny": case.ny, "km": case.km, "ntra": 2,
                   "dts": case.params["dts"], "variant": main_v},
        "roofline": r["roofline"],
        "step_roofline": r["step_roofline"],
        "kernel_ms": r["kernel_ms"],
        "kernel_alg_gbs": r["kernel_alg_gbs"],
        "stream_copy_gbs": round(copy_gbs, 1),
        "write_only_gbs": round(fill_gbs, 1),
        "read_only_gbs": round(read_gbs, 1),
        "cpu_baseline": cpu,
    }
    if "mixed" in runs and main_v != "mixed":
        rm = runs["mixed"]
        out["mixed"] = {"value": rm["value"], "ms_per_step": rm["ms_per_step"], "dtype": "f64/f32",
                        "speedup_vs_fp64": rm["value"] / r["value"],
                        "roofline": rm["roofline"], "step_roofline": rm["step_roofline"],
                        "kernel_ms": rm["kernel_ms"], "kernel_alg_gbs": rm["kernel_alg_gbs"],
                        "error": mixed_err, "error_100_steps_1deg": err_study}
    if ice:
        out["icesnow"] = ice
    if turb:
        out["turb"] = turb
    if readyt:
        out["readyt"] = readyt
    if ens_out:
        out["ensemble"] = ens_out
    if serve_out:
        out["serving"] = serve_out
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
