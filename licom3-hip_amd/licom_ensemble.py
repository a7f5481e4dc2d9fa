"""Batched ensemble engine, launch-shape autotuner and request loop over the
C ABI (mixed-precision form).

* :class:`Ensemble` binds ``licom_batch_*``: many members, each a
  :class:`licom_tracer.TracerModel` of its own grid, advance together, one
  launch per stage and step (the k-march as persistent workgroups over
  (member, tile, tracer) items); N steps can be captured in a HIP graph and
  replayed.
* :class:`Autotuner` times every launch-shape candidate
  (``licom_autotune_member``) once per (imt, jmt, km) and caches the winner.
* :func:`ensemble_cases` draws the members of the ensemble workload: domains
  from {96x64, 128x96, 256x192, 360x218} x km in {30, 55}, member m seeded
  4000+m, domain draw seeded 4999, initial T perturbed by N(0, 0.1).
* :class:`RequestLoop` serves tracer-advance requests with continuous
  batching: requests that arrive while others run are admitted at the next
  chunk boundary when their HBM estimate fits under the cap; member state
  stays resident on the device until its request completes.

A member's results are bit-identical alone and in any batch (same kernels,
per-member reductions with per-member block counts).  There is no CPU
fallback.
"""
from __future__ import annotations

import ctypes
import json
import os
import time

import numpy as np

import licom_synth as ls
import licom_tracer as lt

MIXED_FLAGS = lt.LICOM_STEP_SIDE_EFFECTS | lt.LICOM_STEP_CONVADJ
DOMAINS = ((96, 64), (128, 96), (256, 192), (360, 218))
LEVELS = (30, 55)


class MemberTuning(ctypes.Structure):
    _fields_ = [("prelude_shape", ctypes.c_int), ("solve_shape", ctypes.c_int)]


def _lib():
    L = lt.lib()
    if not getattr(L, "_batch_bound", False):
        V = ctypes.c_void_p
        L.licom_batch_create.argtypes = [ctypes.POINTER(V), ctypes.POINTER(lt.Fields),
                                         ctypes.POINTER(lt.Diag), ctypes.POINTER(MemberTuning),
                                         ctypes.c_int, ctypes.POINTER(V)]
        L.licom_batch_set_stream.argtypes = [V, V]
        L.licom_batch_set_march_grid.argtypes = [V, ctypes.c_int]
        L.licom_batch_step.argtypes = [V, ctypes.c_int, ctypes.c_int]
        L.licom_batch_capture.argtypes = [V, ctypes.c_int, ctypes.c_int]
        L.licom_batch_replay.argtypes = [V]
        L.licom_batch_sync.argtypes = [V]
        L.licom_batch_destroy.argtypes = [V]
        L.licom_batch_set_persistent.argtypes = [V, ctypes.c_int]
        L.licom_batch_persistent_grid.argtypes = [V]
        L.licom_batch_fail_next_barrier.argtypes = [V]
        L.licom_autotune_member.argtypes = [V, ctypes.POINTER(lt.Fields), ctypes.POINTER(lt.Diag),
                                            ctypes.c_int, ctypes.POINTER(MemberTuning),
                                            ctypes.POINTER(ctypes.c_double)]
        L._batch_bound = True
    return L


class Ensemble:
    """Members (mixed-precision TracerModels) advanced as one batch.

    persistent=True: one launch per step (licom_batch_set_persistent), a grid
    of co-resident workgroups with grid-wide barriers between the stages;
    otherwise one launch per stage.  Same bits either way."""

    def __init__(self, members, tuning=None, march_grid=0, stream=None, persistent=False,
                 wait_members=True):
        if not members:
            raise lt.LicomError("Ensemble needs at least one member")
        for m in members:
            if m.precision != "mixed":
                raise lt.LicomError("ensemble members use the mixed-precision form")
        L = _lib()
        n = len(members)
        self.members = list(members)
        ctxs = (ctypes.c_void_p * n)(*[m.ctx.value for m in members])
        F = (lt.Fields * n)(*[m._fields() for m in members])
        D = (lt.Diag * n)(*[m._diag() for m in members])
        T = None
        if tuning is not None:
            T = (MemberTuning * n)(*[MemberTuning(int(a), int(b)) for a, b in tuning])
        self._keep = (ctxs, F, D, T)
        h = ctypes.c_void_p()
        lt.check(L.licom_batch_create(ctxs, F, D, T, n, ctypes.byref(h)), "licom_batch_create")
        self.handle = h
        # a stream of the ensemble's own (default priority): graph capture
        # needs one; it waits for the members' stream before every launch
        import torch
        self.torch = torch
        s = stream if stream is not None else torch.cuda.Stream(device=members[0].device)
        lt.check(L.licom_batch_set_stream(h, ctypes.c_void_p(s.cuda_stream)), "licom_batch_set_stream")
        self.stream = s
        # wait_members: every launch first waits for the members' stream (work
        # the caller queued there); the request loop, whose members' uploads
        # are complete when they join, turns this off so that a batch never
        # waits for the admission of other requests on that stream
        self.member_stream = members[0].stream if wait_members else None
        if march_grid:
            lt.check(L.licom_batch_set_march_grid(h, int(march_grid)), "licom_batch_set_march_grid")
        self.persistent = bool(persistent)
        if self.persistent:
            lt.check(L.licom_batch_set_persistent(h, 1), "licom_batch_set_persistent")
        self.captured = 0

    def set_stream(self, stream):
        """Moves the batch's launches to another stream; the caller orders it
        after the work already queued on the old one."""
        lt.check(_lib().licom_batch_set_stream(self.handle, ctypes.c_void_p(stream.cuda_stream)),
                 "licom_batch_set_stream")
        self.stream = stream

    def _order(self):
        if self.member_stream is not None:
            self.stream.wait_stream(self.member_stream)

    def step(self, nsteps=1, flags=MIXED_FLAGS):
        self._order()
        lt.check(_lib().licom_batch_step(self.handle, int(nsteps), int(flags)), "licom_batch_step")
        for m in self.members:
            m.ist += nsteps

    def capture(self, nsteps=24, flags=MIXED_FLAGS):
        self._order()
        self.sync()
        lt.check(_lib().licom_batch_capture(self.handle, int(nsteps), int(flags)), "licom_batch_capture")
        self.captured = int(nsteps)

    def replay(self):
        self._order()
        lt.check(_lib().licom_batch_replay(self.handle), "licom_batch_replay")
        for m in self.members:
            m.ist += self.captured

    def sync(self):
        """Waits for the batch's work; afterwards the members' host_state()
        reads the advanced state."""
        lt.check(_lib().licom_batch_sync(self.handle), "licom_batch_sync")

    def nonfinite(self):
        """Per member: tracer columns whose state went non-finite in the last
        step (a member is flagged; the others are not affected)."""
        self.sync()
        return [m.nonfinite() for m in self.members]

    def persistent_grid(self):
        """Workgroups of the persistent step (0 unless persistent)."""
        return int(_lib().licom_batch_persistent_grid(self.handle))

    def cell_updates_per_step(self):
        return sum(m.case.nx * m.case.ny * m.case.km * 2 for m in self.members)

    def close(self):
        if getattr(self, "handle", None):
            _lib().licom_batch_destroy(self.handle)
            self.handle = None


def member_case(seed, nx, ny, km, t_pert=0.1):
    """One ensemble member from the workload generator: the 1 deg global-like
    grid for 360x218, the 0.5 deg regional grid otherwise."""
    base = "C2" if (nx, ny) == (360, 218) else "C1"
    return ls.make_case(base, nx=nx, ny=ny, km=km, seed=seed, t_pert=t_pert)


def draw_domains(n, seed=4999):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nx, ny = DOMAINS[int(rng.integers(len(DOMAINS)))]
        out.append((nx, ny, LEVELS[int(rng.integers(len(LEVELS)))]))
    return out


def ensemble_cases(n=64, seed0=4000, shape_seed=4999):
    return [member_case(seed0 + m, *shape) for m, shape in enumerate(draw_domains(n, shape_seed))]


def member_bytes(nx, ny, km):
    """HBM held by one resident mixed-precision member: FP64 at/atb, FP32
    akt/velocities/ws and diagnostics, the context's scratch (ori and the face
    velocities in FP64-sized buffers) and the 2-D fields."""
    n2 = (nx + 4) * (ny + 4)
    n3 = n2 * km
    state = 8 * (2 * n3 + 2 * (n3 + n2)) + 4 * (2 * n3 + 4 * n3 + (n3 + n2))
    diag = 4 * (14 * n3 + n3) + 8 * 2 * n2
    scratch = 8 * (2 * n3 + 2 * n3) + 8 * 12 * n2
    return state + diag + scratch + 8 * 16 * n2


class Autotuner:
    """Chooses launch shapes per (imt, jmt, km) by timing every candidate
    on a throw-away member of that shape; results are cached in a JSON file."""

    def __init__(self, cache_path=None, reps=3):
        self.cache_path = cache_path
        self.reps = int(reps)
        self.table = {}
        if cache_path and os.path.exists(cache_path):
            with open(cache_path) as f:
                self.table = json.load(f)

    @staticmethod
    def key(case):
        return f"{case.imt}x{case.jmt}x{case.km}"

    def tune(self, case):
        k = self.key(case)
        if k not in self.table:
            m = lt.TracerModel(case, diagnostics=True, precision="mixed")
            m.torch.cuda.synchronize()
            best = MemberTuning()
            ms = ctypes.c_double()
            F, D = m._fields(), m._diag()
            lt.check(_lib().licom_autotune_member(m.ctx, ctypes.byref(F), ctypes.byref(D),
                                                  self.reps, ctypes.byref(best), ctypes.byref(ms)),
                     "licom_autotune_member")
            m.close()
            self.table[k] = {"prelude_shape": best.prelude_shape, "solve_shape": best.solve_shape,
                             "ms_per_step": ms.value}
            if self.cache_path:
                with open(self.cache_path, "w") as f:
                    json.dump(self.table, f, indent=1)
        e = self.table[k]
        return e["prelude_shape"], e["solve_shape"]


class Request:
    __slots__ = ("rid", "arrival", "case", "n_steps", "done_steps", "model", "admitted",
                 "finished", "bytes", "error", "ens", "lane", "queued", "timing")

    def __init__(self, rid, arrival, case, n_steps):
        self.rid, self.arrival, self.case, self.n_steps = rid, arrival, case, int(n_steps)
        self.done_steps = 0
        self.model = None
        self.admitted = None
        self.finished = None
        self.bytes = member_bytes(case.nx, case.ny, case.km)
        self.error = None     # reason the request was rejected, if it was
        self.ens = None       # solo policy: the request's own batch of one
        self.lane = None      # solo policy: the lane it is bound to once started
        self.queued = 0       # solo policy: steps launched and not yet retired
        self.timing = {}      # admission phases (s after arrival): handed to a worker,
                              # picked up, context + upload done, batch made, resident


class _Lane:
    """One size class of the request loop: its active members, their batch
    and a stream of its own, so that classes advance independently."""

    def __init__(self, stream):
        self.stream = stream
        self.active = []      # admitted members of this class
        self.running = []     # the members of the chunk in flight
        self.ens = None
        self.changed = False
        self.busy = False
        self.chunk = 0
        self.event = None
        self.single = False   # members one at a time (least remaining steps first)
        self.running_set = []


class RequestLoop:
    """Continuous batching of tracer-advance requests on one GPU.

    Requests are sorted into size classes by their cells per step
    (`class_cells` bounds); each class is a batch on a stream of its own, so a
    small member never waits for the steps of a large one, and the classes'
    kernels share the GPU.  The host loop polls: it admits every arrived
    request whose HBM estimate fits under the cap (arrival order) and hands it
    to an admission thread, which creates its context and uploads its state on
    a stream of its own (about 10 ms per member, which would otherwise stall
    every class's next chunk); members join their class when their state is
    resident.  For each class whose last chunk finished the loop retires finished
    members (results copied out), rebuilds the batch when the membership
    changed, and launches the next chunk of min(chunk_steps, fewest remaining
    steps).  Batches and member contexts of finished requests are freed when no
    class is running (a free synchronises the device) or when they hold more
    than a quarter of the cap.  Classes listed in `one_at_a_time` run one
    member at a time, least remaining steps first, instead of advancing their
    members together (shorter tails at low load, less throughput at high
    load).  Latency = completion - arrival.

    Admission needs both the HBM estimate under the cap and, on the device,
    that much free memory (hipMemGetInfo, less `free_margin_bytes`).  A
    request that can never be admitted -- its estimate exceeds the cap, or it
    exceeds the device's free memory while nothing else is resident -- is
    rejected: it is not run, its `error` says why, and it is listed in
    `rejected` (and counted in the statistics) instead of blocking the
    requests behind it.

    policy="solo": no lock step at all.  Every request runs alone (a batch of
    one, the launches of an isolated member) on one of `lanes` streams; a lane
    that can take a chunk runs the request with the least remaining work
    (remaining steps x cells per step) among those bound to it and those not
    started yet, so a short request that arrives while long ones run starts
    at the next chunk boundary of a lane and is not slowed to a batch's step
    time.  A request stays on the lane it started on (its state is only
    touched by that stream); chunks are queued two deep per lane.

    Streams share the process's hardware queues (HIP's default is four,
    GPU_MAX_HW_QUEUES): with more streams than queues an admission's copies
    and fills wait in order behind other streams' kernels (measured: 40-270
    ms instead of 2-20 ms per admission), so a serving process should set
    GPU_MAX_HW_QUEUES (up to 32) before HIP starts; bench.py uses 16."""

    def __init__(self, hbm_cap_bytes=200e9, chunk_steps=8, tuner=None, keep_results=False,
                 class_cells=(5.0e5, 2.0e6), one_at_a_time=(), one_at_a_time_upto=None,
                 free_margin_bytes=2.0e9, free_bytes="device", policy="classes",
                 fill_cells=4.0e6, admit_threads=2, lanes=4):
        if policy not in ("classes", "srpt", "solo"):
            raise ValueError(f"policy must be 'classes', 'srpt' or 'solo', not {policy!r}")
        if lanes < 1:
            raise ValueError("lanes must be at least 1")
        # "solo": streams the requests run on, one request per chunk
        self.lanes = int(lanes)
        # "srpt": one batch of the resident requests with the least remaining
        # work (remaining steps x cells per step), as many as fill_cells cells
        # per step (about what fills the GPU), re-chosen at every chunk
        # boundary (chunk_steps steps at most), so a short request that
        # arrives runs at once and leaves as soon as it is done; chunks are
        # queued two deep while no member of the batch finishes.
        self.policy = policy
        self.fill_cells = float(fill_cells)
        # threads creating the admitted requests' contexts and uploading their
        # state, each on a stream of its own
        self.admit_threads = int(admit_threads)
        self.cap = float(hbm_cap_bytes)
        # classes whose members run one at a time, least remaining steps first,
        # instead of advancing together, while at most `one_at_a_time_upto`
        # requests are active (None: always)
        self.one_at_a_time = frozenset(one_at_a_time)
        self.one_upto = one_at_a_time_upto
        self.chunk = int(chunk_steps)
        self.tuner = tuner
        self.keep_results = keep_results
        self.class_cells = tuple(class_cells)
        self.results = {}
        self.rejected = []
        self.failed = []      # served requests whose state went non-finite
        self.free_margin = float(free_margin_bytes)
        # device free-memory probe: "device" (hipMemGetInfo), a callable
        # returning bytes, or None for no probe
        self.free_bytes = free_bytes

    def _reject(self, r, why):
        r.error = why
        self.rejected.append(r)

    def _device_free(self):
        if self.free_bytes is None:
            return None
        if self.free_bytes == "device":
            import torch
            return float(torch.cuda.mem_get_info()[0])
        return float(self.free_bytes())

    def _class(self, r):
        cells = r.case.nx * r.case.ny * r.case.km
        for i, b in enumerate(self.class_cells):
            if cells <= b:
                return i
        return len(self.class_cells)

    def _check_finite(self, r, stream):
        # a request whose final state is not finite is reported as failed
        # (its result is still returned); the others are not affected.  The
        # check runs on the lane's stream, whose chunk has completed, so it
        # never waits for other work (admissions, other lanes)
        import torch
        if r.model is None:
            return
        with torch.cuda.stream(stream):
            ok = bool(torch.isfinite(r.model.state["at"]).all())
        if not ok:
            r.error = "the state became non-finite (NaN or infinity)"
            self.failed.append(r)

    @staticmethod
    def _remaining(r):
        return (r.n_steps - r.done_steps) * (r.case.nx * r.case.ny * r.case.km)

    def _srpt_pick(self, active):
        run, cells = [], 0.0
        for r in sorted(active, key=lambda r: (self._remaining(r), r.arrival)):
            c = r.case.nx * r.case.ny * r.case.km
            if run and cells + c > self.fill_cells:
                continue
            run.append(r)
            cells += c
        return run

    def _srpt_turn(self, ln, done, grave, clock, t0, torch):
        """One scheduling turn of the srpt policy on its single lane: retire
        the chunks that finished, then keep up to two chunks queued.
        Returns (launched, cell-updates retired, batch rebuilds)."""
        upd, rb, launched = 0, 0, False
        # ln.queue: chunks in flight, oldest first: (event, members, steps)
        q = ln.__dict__.setdefault("queue", [])
        while q and q[0][0].query():
            _, members, steps = q.pop(0)
            t_now = clock() - t0
            upd += steps * sum(r.case.nx * r.case.ny * r.case.km * 2 for r in members)
            finished = []
            for r in members:
                r.done_steps += steps
                if r.done_steps >= r.n_steps:
                    finished.append(r)
            for r in finished:
                r.finished = t_now
                done.append(r)
                self._check_finite(r, ln.stream)
                if self.keep_results:
                    self.results[r.rid] = r.model.host_state()["at"]
            if finished:
                grave.append((ln.ens, 0.0))
                ln.ens = None
                for r in finished:
                    grave.append((r.model, r.bytes))
                    r.model = None
                ln.active = [r for r in ln.active if r.finished is None]
                ln.changed = True
        # queue a chunk when the lane is idle, or a second one behind the
        # running chunk when no member of it finishes and nothing has joined
        while ln.active and len(q) < 2:
            if q:
                _, members, steps = q[-1]
                if ln.changed or any(r.done_steps + steps >= r.n_steps for r in members):
                    break
                run = members
                ahead = steps
            else:
                run = self._srpt_pick(ln.active)
                ahead = 0
            if ln.changed or ln.ens is None or run != ln.running_set:
                if ln.ens is not None:
                    grave.append((ln.ens, 0.0))
                tuning = [self.tuner.tune(r.case) for r in run] if self.tuner else None
                ln.ens = Ensemble([r.model for r in run], tuning=tuning, stream=ln.stream,
                                  wait_members=False)
                ln.changed = False
                ln.running_set = run
                rb += 1
            steps = min(self.chunk, min(r.n_steps - r.done_steps - ahead for r in run))
            ln.ens.step(steps)
            ev = torch.cuda.Event()
            ev.record(ln.stream)
            q.append((ev, run, steps))
            launched = True
        ln.busy = bool(q)
        ln.running = q[0][1] if q else []
        return launched, upd, rb

    @staticmethod
    def _solo_pick(active, li):
        """The request lane `li` runs next: least remaining work (steps not yet
        launched x cells per step) among the requests bound to the lane and
        those not started, earliest arrival on ties; None if there is none."""
        best, key = None, None
        for r in active:
            if r.lane not in (None, li):
                continue
            left = r.n_steps - r.done_steps - r.queued
            if left <= 0:
                continue
            k = (left * (r.case.nx * r.case.ny * r.case.km), r.arrival)
            if key is None or k < key:
                best, key = r, k
        return best

    def _solo_turn(self, lanes, active, done, grave, clock, t0, torch):
        """One scheduling turn of the solo policy: retire the lanes' finished
        chunks, then fill every lane's queue (two chunks deep), idle lanes
        first.  Returns (launched, cell-updates retired)."""
        upd, launched = 0, False
        for ln in lanes:
            q = ln.__dict__.setdefault("queue", [])
            while q and q[0][0].query():
                _, r, steps = q.pop(0)
                r.queued -= steps
                r.done_steps += steps
                upd += steps * r.case.nx * r.case.ny * r.case.km * 2
                if r.done_steps >= r.n_steps:
                    r.finished = clock() - t0
                    done.append(r)
                    self._check_finite(r, ln.stream)
                    if self.keep_results:
                        self.results[r.rid] = r.model.host_state()["at"]
                    grave.append((r.ens, 0.0))
                    grave.append((r.model, r.bytes))
                    r.ens = r.model = None
        active[:] = [r for r in active if r.finished is None]
        for depth in (0, 1):
            for li, ln in enumerate(lanes):
                if len(ln.queue) != depth:
                    continue
                r = self._solo_pick(active, li)
                if r is None:
                    continue
                if r.lane is None:
                    r.lane = li
                    r.ens.set_stream(ln.stream)
                steps = min(self.chunk, r.n_steps - r.done_steps - r.queued)
                r.ens.step(steps)
                ev = torch.cuda.Event()
                ev.record(ln.stream)
                ln.queue.append((ev, r, steps))
                r.queued += steps
                launched = True
        for ln in lanes:
            ln.busy = bool(ln.queue)
        return launched, upd

    def run(self, requests, clock=time.perf_counter, sleep=time.sleep):
        import queue
        import threading

        import torch
        pending = []
        for r in sorted(requests, key=lambda r: r.arrival):
            if r.bytes > self.cap:
                self._reject(r, f"needs {r.bytes / 1e9:.2f} GB, above the {self.cap / 1e9:.2f} GB cap")
            else:
                pending.append(r)
        done = []
        used = 0.0
        t0 = clock()
        updates = 0
        i = 0
        peak_active = 0
        rebuilds = 0
        lanes = []            # size classes, created with the first admission
        dev = None
        grave, grave_bytes = [], 0.0
        todo, ready, failures = queue.Queue(), queue.Queue(), []
        tune_lock = threading.Lock()
        solo_active = []      # solo policy: admitted requests not finished
        inflight = 0
        worker = None

        def start_lanes():
            nonlocal dev, worker
            dev = torch.cuda.current_device()
            nclass = {"srpt": 1, "solo": self.lanes}.get(self.policy, len(self.class_cells) + 1)
            for k in range(nclass):
                ln = _Lane(torch.cuda.Stream(device=dev))
                ln.single = k in self.one_at_a_time
                lanes.append(ln)
            worker = [threading.Thread(target=admit_worker, daemon=True) for _ in range(self.admit_threads)]
            for w in worker:
                w.start()

        def admit_worker():
            torch.cuda.set_device(dev)
            side = torch.cuda.Stream(device=dev)
            while True:
                r = todo.get()
                if r is None:
                    return
                try:
                    r.timing["picked"] = clock() - t0 - r.arrival
                    with torch.cuda.stream(side):
                        r.model = lt.TracerModel(r.case, diagnostics=True, precision="mixed", stream=side)
                        r.timing["context"] = clock() - t0 - r.arrival
                        if self.policy == "solo":
                            with tune_lock:
                                tuning = [self.tuner.tune(r.case)] if self.tuner else None
                            r.ens = Ensemble([r.model], tuning=tuning, stream=side, wait_members=False)
                            r.timing["batch"] = clock() - t0 - r.arrival
                    side.synchronize()     # state resident before a class stream reads it
                    r.timing["resident"] = clock() - t0 - r.arrival
                except Exception as e:     # handed to the loop, which raises it
                    failures.append(e)
                    ready.put(None)
                    return
                ready.put(r)

        def flush():
            nonlocal grave, grave_bytes, used
            for obj, nbytes in grave:
                obj.close()
                used -= nbytes
            grave, grave_bytes = [], 0.0

        while True:
            now = clock() - t0
            while i < len(pending) and pending[i].arrival <= now and used + pending[i].bytes <= self.cap:
                r = pending[i]
                free = self._device_free()
                if free is not None and r.bytes > free - self.free_margin:
                    if used == 0.0 and not grave and inflight == 0:
                        self._reject(r, f"needs {r.bytes / 1e9:.2f} GB, the device has "
                                        f"{free / 1e9:.2f} GB free")
                        i += 1
                        continue
                    break        # waits for resident members to finish
                if not lanes:
                    start_lanes()
                i += 1
                used += r.bytes
                r.timing["handed"] = clock() - t0 - r.arrival
                todo.put(r)
                inflight += 1
            while True:
                try:
                    r = ready.get_nowait()
                except queue.Empty:
                    break
                if r is None:
                    for _ in worker:
                        todo.put(None)
                    raise failures[0]
                inflight -= 1
                r.admitted = clock() - t0
                if self.policy == "solo":
                    solo_active.append(r)
                    continue
                ln = lanes[0 if self.policy == "srpt" else self._class(r)]
                ln.active.append(r)
                ln.changed = True
            launched = False
            if self.policy == "solo" and lanes:
                launched, upd = self._solo_turn(lanes, solo_active, done, grave, clock, t0, torch)
                updates += upd
                grave_bytes = sum(b for _, b in grave)
            if self.policy == "srpt" and lanes:
                ln = lanes[0]
                launched, upd, rb = self._srpt_turn(ln, done, grave, clock, t0, torch)
                updates += upd
                rebuilds += rb
                grave_bytes = sum(b for _, b in grave)
            for ln in (lanes if self.policy == "classes" else ()):
                if ln.busy:
                    if not ln.event.query():
                        continue
                    ln.busy = False
                    t_now = clock() - t0
                    updates += ln.chunk * sum(r.case.nx * r.case.ny * r.case.km * 2 for r in ln.running)
                    finished = []
                    for r in ln.running:
                        r.done_steps += ln.chunk
                        if r.done_steps >= r.n_steps:
                            finished.append(r)
                    for r in finished:
                        r.finished = t_now
                        done.append(r)
                        self._check_finite(r, ln.stream)
                        if self.keep_results:
                            self.results[r.rid] = r.model.host_state()["at"]
                    if finished:
                        # the batch holds the finished members' arrays: retire
                        # it with them
                        grave.append((ln.ens, 0.0))
                        ln.ens = None
                        for r in finished:
                            grave.append((r.model, r.bytes))
                            grave_bytes += r.bytes
                            r.model = None
                        ln.changed = True
                    ln.active = [r for r in ln.active if r.finished is None]
                    ln.running = []
                if not ln.busy and ln.active:
                    one = ln.single and (self.one_upto is None or
                                         sum(len(x.active) for x in lanes) <= self.one_upto)
                    run = ([min(ln.active, key=lambda r: (r.n_steps - r.done_steps, r.arrival))]
                           if one else list(ln.active))
                    if ln.changed or ln.ens is None or run != ln.running_set:
                        if ln.ens is not None:
                            grave.append((ln.ens, 0.0))
                        tuning = [self.tuner.tune(r.case) for r in run] if self.tuner else None
                        ln.ens = Ensemble([r.model for r in run], tuning=tuning, stream=ln.stream,
                                          wait_members=False)
                        ln.changed = False
                        ln.running_set = run
                        rebuilds += 1
                    ln.running = run
                    ln.chunk = min(self.chunk, min(r.n_steps - r.done_steps for r in ln.running))
                    ln.ens.step(ln.chunk)
                    ln.event = torch.cuda.Event()
                    ln.event.record(ln.stream)
                    ln.busy = True
                    launched = True
            peak_active = max(peak_active, sum(len(ln.active) for ln in lanes) + len(solo_active))
            idle = not any(ln.busy for ln in lanes)
            if grave and (idle or grave_bytes > 0.25 * self.cap):
                torch.cuda.synchronize()
                flush()
            if idle and not any(ln.active for ln in lanes) and not solo_active and inflight == 0:
                if i >= len(pending):
                    break
                wait = pending[i].arrival - (clock() - t0)
                # never spin: an arrived request that does not fit yet waits
                # for frees, a future one for its arrival
                sleep(min(wait, 0.005) if wait > 0 else 5e-4)
            elif not launched:
                sleep(5e-5)
        if worker is not None:
            for _ in worker:
                todo.put(None)
            for w in worker:
                w.join()
            torch.cuda.synchronize()
        flush()
        makespan = clock() - t0
        lat = np.array([r.finished - r.arrival for r in done])
        return {"requests": len(done), "rejected": len(self.rejected), "nonfinite": len(self.failed),
                "makespan_s": makespan,
                "cell_updates": updates,
                "cell_updates_per_s": updates / makespan if makespan > 0 else 0.0,
                "p50_latency_s": float(np.percentile(lat, 50)) if lat.size else None,
                "p99_latency_s": float(np.percentile(lat, 99)) if lat.size else None,
                "peak_active": peak_active, "batch_rebuilds": rebuilds,
                "chunk_steps": self.chunk, "class_cells": list(self.class_cells)}


def request_specs(n, seed_requests=5002):
    """(nx, ny, km, n_steps) of n requests: a member domain from the ensemble
    distribution and n_steps uniform in 24..240."""
    rr = np.random.default_rng(seed_requests)
    out = []
    for _ in range(n):
        nx, ny = DOMAINS[int(rr.integers(len(DOMAINS)))]
        km = LEVELS[int(rr.integers(len(LEVELS)))]
        out.append((nx, ny, km, int(rr.integers(24, 241))))
    return out


def poisson_requests(n, rate_per_s, seed_arrivals=5001, seed_requests=5002, t_pert=0.1):
    """n requests with exponential inter-arrival times (rate per second) and
    the shapes of request_specs; request k's member is seeded 6000+k."""
    ra = np.random.default_rng(seed_arrivals)
    t = 0.0
    out = []
    for k, (nx, ny, km, steps) in enumerate(request_specs(n, seed_requests)):
        t += float(ra.exponential(1.0 / rate_per_s))
        out.append(Request(k, t, member_case(6000 + k, nx, ny, km, t_pert), steps))
    return out
