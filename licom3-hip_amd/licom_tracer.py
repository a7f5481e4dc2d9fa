"""Host-side mirror of the reference's tracer interface over the C ABI.

``liblicom_tracer.so`` (include/licom_tracer.h) is the product: HIP kernels
for gfx950 behind a plain C ABI.  This module binds it with ctypes and keeps
the model state resident on the device (torch tensors are used only as
device allocations and for stream handles).

:class:`TracerModel` mirrors the reference's calling convention:

* ``tracer(ii)``  -> ``extern "C" void tracer_(int ii)``  (tracer.cpp:1701)
* ``convadj()``   -> ``extern "C" void convadj_()``        (convadj.cpp:290)
* ``icesnow()``   -> ``extern "C" void icesnow_()``        (icesnow.cpp:70)
* ``stepon(n)``   -> the tracer/icesnow/convadj part of ``stepon_``
  (stepon.cpp:57-77)

State I/O on the host (no GPU): ``read_restart`` / ``write_restart`` (the
restart stream of inirun.F90:598-700), ``read_eos`` (dncoef.h1,
constant_mod.F90:331-364) and ``global_to_halo``.

There is no CPU fallback for the step: if the library or a GPU is missing
every model entry point raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int32)
_VP = ctypes.c_void_p

LICOM_STEP_CONVADJ = 0x1
LICOM_STEP_SIDE_EFFECTS = 0x2
LICOM_STEP_VTL_SCRATCH = 0x4
LICOM_STEP_STAGED = 0x8
LICOM_STEP_ICESNOW = 0x10
HMIX_DEL2, HMIX_DEL4 = 2, 4
ADV_CENTERED, ADV_FLUX, ADV_TSPAS = 0, 1, 2   # the reference's Str codes (tracer.cpp:1712-1714)


class Grid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("km", ctypes.c_int),
                ("imt_global", ctypes.c_int),
                ("kmt", _IP), ("zkp", _DP), ("pen", _DP),
                ("htw", _DP), ("hts", _DP), ("hun", _DP), ("hue", _DP), ("tarea", _DP),
                ("au0", _DP), ("aus", _DP), ("auw", _DP), ("ausw", _DP),
                ("to", _DP), ("so", _DP), ("eosc", _DP), ("dxu", _DP), ("dyu", _DP)]


class Params(ctypes.Structure):
    _fields_ = [("dts", ctypes.c_double), ("onbc", ctypes.c_double), ("oncc", ctypes.c_double),
                ("dwndmix", ctypes.c_double), ("gamma", ctypes.c_double),
                ("od0", ctypes.c_double), ("od0cp", ctypes.c_double), ("ah", ctypes.c_double),
                ("hmix", ctypes.c_int), ("adv", ctypes.c_int),
                ("boundary_restore", ctypes.c_int), ("sssnorm", ctypes.c_int),
                ("nss", ctypes.c_int)]


class Fields(ctypes.Structure):
    _fields_ = [("at", _VP), ("atb", _VP), ("akt", _VP), ("utf", _VP), ("vtf", _VP),
                ("utl", _VP), ("vtl", _VP), ("ws", _VP), ("h0f", _VP), ("h0l", _VP),
                ("swv", _VP), ("nswv", _VP), ("sst", _VP), ("seaice", _VP), ("sss", _VP),
                ("fresh", _VP), ("restore", _VP)]


class Diag(ctypes.Structure):
    _fields_ = [("ax", _VP), ("ay", _VP), ("az", _VP), ("dz", _VP), ("tend", _VP),
                ("dt_diff", _VP), ("dt_conv", _VP), ("penetrate", _VP), ("net", _VP),
                ("fw_norm2", _VP)]


class GridDerived(ctypes.Structure):
    _fields_ = [("kmt", _IP), ("kmu", _IP), ("tarea_r", _DP), ("ohbt", _DP), ("ohbu", _DP),
                ("dtn", _DP), ("dts", _DP), ("dte", _DP), ("dtw", _DP), ("ahf", _DP),
                ("dzp", _DP), ("odzp", _DP), ("odzt", _DP), ("area_t", ctypes.c_double)]


EXPORTED = ["licom_tracer_create", "licom_tracer_set_stream", "licom_tracer_step",
            "licom_tracer_step_noalloc",
            "licom_tracer_step_mixed", "licom_batch_create", "licom_batch_set_stream",
            "licom_batch_set_march_grid", "licom_batch_set_persistent", "licom_batch_persistent_grid",
            "licom_batch_fail_next_barrier", "licom_batch_step", "licom_batch_capture",
            "licom_batch_replay", "licom_batch_sync", "licom_batch_destroy",
            "licom_autotune_member", "licom_convadj", "licom_icesnow", "licom_tracer_set_icesnow",
            "licom_tracer_sync", "licom_grid_export", "licom_set_profiling",
            "licom_kernel_times", "licom_scratch_bytes", "licom_last_mix_count", "licom_last_nonfinite",
            "licom_tracer_set_flag_output",
            "licom_stream_copy", "licom_tracer_destroy", "licom_strerror", "licom_last_error",
            "licom_restart_read", "licom_restart_write", "licom_eos_read"]


class Restart(ctypes.Structure):
    _fields_ = [(k, _DP) for k in ("h0", "u", "v", "at", "ws", "su", "sv", "swv", "sshf", "lthf",
                                   "fresh")] + [("number_month", ctypes.c_int32),
                                                ("number_day", ctypes.c_int32),
                                                ("little_endian", ctypes.c_int32)]


RESTART_2D = ("h0", "su", "sv", "swv", "sshf", "lthf", "fresh")


class LicomError(RuntimeError):
    pass


def lib_path():
    return os.path.join(_HERE, "liblicom_tracer.so")


def lib():
    """Load liblicom_tracer.so (raises if it was not built)."""
    global _LIB
    if _LIB is None:
        path = lib_path()
        if not os.path.exists(path):
            raise LicomError(f"{path} is missing: run __graft_entry__.build()")
        L = ctypes.CDLL(path)
        L.licom_tracer_create.argtypes = [ctypes.POINTER(Grid), ctypes.POINTER(Params),
                                          ctypes.POINTER(_VP)]
        L.licom_tracer_set_stream.argtypes = [_VP, _VP]
        L.licom_tracer_step.argtypes = [_VP, ctypes.POINTER(Fields), ctypes.POINTER(Diag),
                                        ctypes.c_int]
        L.licom_tracer_step_noalloc.argtypes = L.licom_tracer_step.argtypes
        # licom_fields_mixed / licom_diag_mixed have the same pointer layout
        L.licom_tracer_step_mixed.argtypes = [_VP, ctypes.POINTER(Fields), ctypes.POINTER(Diag),
                                              ctypes.c_int]
        L.licom_convadj.argtypes = [_VP, ctypes.POINTER(Fields), ctypes.POINTER(Diag)]
        L.licom_icesnow.argtypes = [_VP, _VP, _VP, ctypes.c_int]
        L.licom_tracer_set_icesnow.argtypes = [_VP, _VP]
        L.licom_tracer_sync.argtypes = [_VP]
        L.licom_grid_export.argtypes = [_VP, ctypes.POINTER(GridDerived)]
        L.licom_set_profiling.argtypes = [_VP, ctypes.c_int]
        L.licom_kernel_times.argtypes = [_VP, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p),
                                         _DP, _IP, _IP]
        L.licom_scratch_bytes.argtypes = [_VP]
        L.licom_scratch_bytes.restype = ctypes.c_size_t
        L.licom_last_mix_count.argtypes = [_VP, _IP]
        L.licom_last_nonfinite.argtypes = [_VP, _IP]
        L.licom_tracer_set_flag_output.argtypes = [_VP, _VP]
        L.licom_stream_copy.argtypes = [_VP, _VP, ctypes.c_size_t, _VP]
        L.licom_tracer_destroy.argtypes = [_VP]
        L.licom_strerror.argtypes = [ctypes.c_int]
        L.licom_strerror.restype = ctypes.c_char_p
        L.licom_last_error.restype = ctypes.c_char_p
        L.licom_restart_read.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(Restart)]
        L.licom_restart_write.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          _IP, _IP, ctypes.POINTER(Restart)]
        L.licom_eos_read.argtypes = [ctypes.c_char_p, ctypes.c_int, _DP, _DP, _DP, _DP]
        _LIB = L
    return _LIB


def check(rc, what):
    if rc != 0:
        L = lib()
        raise LicomError(f"{what}: {L.licom_strerror(rc).decode()} ({rc}): "
                         f"{L.licom_last_error().decode()}")


def _d(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_DP)


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


# -- state I/O (host side) ------------------------------------------------
def _byte_order(byte_order):
    if byte_order not in ("big", "little"):
        raise LicomError(f"byte_order must be 'big' or 'little', not {byte_order!r}")
    return int(byte_order == "little")


def read_restart(path, nx, ny, km, byte_order="big"):
    """The restart stream of inirun.F90:598-700 as numpy arrays on the global
    grid (h0 and the surface fields [ny][nx], u, v, ws [km][ny][nx], at
    [2][km][ny][nx]) plus number_month, number_day.  Records are big-endian
    by default, as the reference's -convert big_endian build writes them."""
    out = {k: np.zeros((ny, nx)) for k in RESTART_2D}
    out.update(u=np.zeros((km, ny, nx)), v=np.zeros((km, ny, nx)), ws=np.zeros((km, ny, nx)),
               at=np.zeros((2, km, ny, nx)))
    R = Restart(**{k: _d(v) for k, v in out.items()}, little_endian=_byte_order(byte_order))
    check(lib().licom_restart_read(os.fsencode(path), nx, ny, km, ctypes.byref(R)),
          "licom_restart_read")
    out["number_month"], out["number_day"] = int(R.number_month), int(R.number_day)
    return out


def write_restart(path, fields, kmt, kmu, number_month=0, number_day=0, byte_order="big"):
    """Writes `fields` (the keys read_restart returns) as a restart stream
    (big-endian by default); land (below the global kmt / kmu [ny][nx]) is
    written as 0."""
    at = np.ascontiguousarray(fields["at"], np.float64)
    _, km, ny, nx = at.shape
    arrs = {k: np.ascontiguousarray(fields[k], np.float64) for k in RESTART_2D + ("u", "v", "ws")}
    arrs["at"] = at
    R = Restart(**{k: _d(v) for k, v in arrs.items()}, number_month=int(number_month),
                number_day=int(number_day), little_endian=_byte_order(byte_order))
    kt = np.ascontiguousarray(kmt, np.int32)
    ku = np.ascontiguousarray(kmu, np.int32)
    check(lib().licom_restart_write(os.fsencode(path), nx, ny, km, kt.ctypes.data_as(_IP),
                                    ku.ctypes.data_as(_IP), ctypes.byref(R)), "licom_restart_write")


def read_eos(path, km):
    """dncoef.h1 (constant_mod.F90:331-364): to, so, po [km] and the density
    coefficients [9][km] in kg/m^3 (C scaled by 1000), licom_grid.eosc."""
    to, so, po, c = np.zeros(km), np.zeros(km), np.zeros(km), np.zeros((9, km))
    check(lib().licom_eos_read(os.fsencode(path), km, _d(to), _d(so), _d(po), _d(c)), "licom_eos_read")
    return to, so, po, c


def global_to_halo(a):
    """A global-grid array [..., ny, nx] in the engine's layout [..., ny+4,
    nx+4]: two ghost cells, cyclic in i, 0 in the closed rows, as the
    reference's halo update after a restart read (inirun.F90:616-667)."""
    a = np.asarray(a, np.float64)
    out = np.zeros(a.shape[:-2] + (a.shape[-2] + 4, a.shape[-1] + 4))
    out[..., 2:-2, 2:-2] = a
    out[..., 2:-2, 0:2] = a[..., :, -2:]
    out[..., 2:-2, -2:] = a[..., :, 0:2]
    return out


def make_grid(case, keep):
    g = Grid()
    g.nx, g.ny, g.km, g.imt_global = case.nx, case.ny, case.km, case.imt_global
    cg = case.grid
    keep.append(cg)
    g.kmt = cg["kmt"].ctypes.data_as(_IP)
    for f in ("zkp", "pen", "htw", "hts", "hun", "hue", "tarea", "au0", "aus", "auw", "ausw",
              "to", "so", "eosc"):
        setattr(g, f, _d(cg[f]))
    for f in ("dxu", "dyu"):
        if f in cg:
            setattr(g, f, _d(cg[f]))
    return g


def make_params(p):
    P = Params()
    for f in ("dts", "onbc", "oncc", "dwndmix", "gamma", "od0", "od0cp", "ah"):
        setattr(P, f, float(p[f]))
    P.hmix = int(p["hmix"])
    P.adv = int(p.get("adv", ADV_TSPAS))
    P.boundary_restore = int(p.get("boundary_restore", 0))
    P.sssnorm = int(p.get("sssnorm", 1))
    P.nss = int(p.get("nss", 1))
    return P


STATE_KEYS = ("at", "atb", "akt", "utf", "vtf", "utl", "vtl", "ws", "h0f", "h0l")
FORCING_KEYS = ("swv", "nswv", "sst", "seaice", "sss", "fresh")
DIAG_KEYS = ("ax", "ay", "az", "dz", "tend", "dt_diff", "dt_conv", "penetrate", "net",
             "fw_norm2")


class TracerModel:
    """Device-resident tracer state plus the reference's entry points.

    ``state`` holds torch float64 tensors on the GPU with the reference
    shapes; ``diag`` holds the diagnostics the reference produces.
    """

    # arrays held in FP32 by the mixed-precision form
    MIXED_F32 = ("akt", "utf", "vtf", "utl", "vtl", "ws")

    def __init__(self, case, device="cuda:0", diagnostics=True, stream=None, precision="fp64"):
        import torch
        if precision not in ("fp64", "mixed"):
            raise LicomError(f"precision must be 'fp64' or 'mixed', not {precision!r}")
        self.precision = precision
        if not torch.cuda.is_available():
            raise LicomError("TracerModel needs a GPU (torch.cuda.is_available() is False)")
        self.torch = torch
        self.case = case
        self.device = torch.device(device)
        torch.cuda.set_device(self.device)
        L = lib()
        self._keep = []
        self._grid = make_grid(case, self._keep)
        self._params = make_params(case.params)
        h = _VP()
        check(L.licom_tracer_create(ctypes.byref(self._grid), ctypes.byref(self._params),
                                    ctypes.byref(h)), "licom_tracer_create")
        self.ctx = h
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.device)
        check(L.licom_tracer_set_stream(self.ctx, _VP(self.stream.cuda_stream)), "set_stream")
        mixed = precision == "mixed"
        self.state = {}
        for k, v in case.state.items():
            t = torch.from_numpy(np.ascontiguousarray(v)).to(self.device)
            self.state[k] = t.to(torch.float32) if (mixed and k in self.MIXED_F32) else t
        self.forcing = {k: (None if case.forcing.get(k) is None else
                            torch.from_numpy(np.ascontiguousarray(case.forcing[k])).to(self.device))
                        for k in FORCING_KEYS}
        r = case.forcing.get("restore")
        self.restore = None if r is None else torch.from_numpy(r).to(self.device)
        km, jmt, imt = case.km, case.jmt, case.imt
        self.diag = None
        if diagnostics:
            dt3 = torch.float32 if mixed else torch.float64
            z = lambda *s: torch.zeros(s, dtype=dt3, device=self.device)  # noqa: E731
            z64 = lambda *s: torch.zeros(s, dtype=torch.float64, device=self.device)  # noqa: E731
            self.diag = dict(ax=z(2, km, jmt, imt), ay=z(2, km, jmt, imt), az=z(2, km, jmt, imt),
                             dz=z(2, km, jmt, imt), tend=z(2, km, jmt, imt),
                             dt_diff=z(2, km, jmt, imt), dt_conv=z(2, km, jmt, imt),
                             penetrate=z(km, jmt, imt), net=z64(2, jmt, imt), fw_norm2=z64(1))
        # licomqice of icesnow_ (tracer_mod.F90), written by icesnow() and by
        # steps with LICOM_STEP_ICESNOW
        self.licomqice = torch.zeros((jmt, imt), dtype=torch.float64, device=self.device)
        check(L.licom_tracer_set_icesnow(self.ctx, _VP(self.licomqice.data_ptr())),
              "licom_tracer_set_icesnow")
        self.ist = 0

    # -- C structs -------------------------------------------------------
    def _fields(self):
        F = Fields()
        for k in STATE_KEYS:
            setattr(F, k, _ptr(self.state[k]))
        for k in FORCING_KEYS:
            setattr(F, k, _ptr(self.forcing[k]))
        F.restore = _ptr(self.restore)
        return F

    def _diag(self):
        D = Diag()
        if self.diag is not None:
            for k in DIAG_KEYS:
                setattr(D, k, _ptr(self.diag[k]))
        return D

    # -- reference entry points -----------------------------------------
    def tracer(self, ii=0, flags=LICOM_STEP_SIDE_EFFECTS):
        """tracer_(ii): one tracer step of T and S (tracer.cpp:1701-1903)."""
        F, D = self._fields(), self._diag()
        if self.precision == "mixed":
            check(lib().licom_tracer_step_mixed(self.ctx, ctypes.byref(F), ctypes.byref(D),
                                                int(flags)), "licom_tracer_step_mixed")
        else:
            check(lib().licom_tracer_step(self.ctx, ctypes.byref(F), ctypes.byref(D), int(flags)),
                  "licom_tracer_step")
        self.ist += 1

    def convadj(self):
        """convadj_() over every column (convadj.cpp:290-298)."""
        if self.precision == "mixed":
            raise LicomError("convadj() alone is FP64-only; use step() with LICOM_STEP_CONVADJ")
        F, D = self._fields(), self._diag()
        check(lib().licom_convadj(self.ctx, ctypes.byref(F), ctypes.byref(D)), "licom_convadj")

    def icesnow(self, reset=True):
        """icesnow_() over every column (icesnow.cpp:70-85): surface freezing
        adjustment and TBICE clamp of at; licomqice zeroed first when reset
        (the FRC_CORE build)."""
        if self.precision == "mixed":
            raise LicomError("icesnow() alone is FP64-only; use step() with LICOM_STEP_ICESNOW")
        check(lib().licom_icesnow(self.ctx, _VP(self.state["at"].data_ptr()),
                                  _VP(self.licomqice.data_ptr()), int(bool(reset))), "licom_icesnow")

    def step(self, flags=LICOM_STEP_SIDE_EFFECTS | LICOM_STEP_CONVADJ):
        """tracer_ followed by convadj_ in one launch sequence (the CPU
        driver's per-step order, licom_drv.F90:357-380)."""
        self.tracer(0, flags)

    def stepon(self, nss, convadj_every_step=True, icesnow=False):
        """Tracer part of stepon_ (stepon.cpp:57-77): convadj every step, or
        only at ii == 0 as the GPU driver does; with icesnow, icesnow_
        between the two (stepon.cpp:71), fused into the step."""
        ice = LICOM_STEP_ICESNOW if icesnow else 0
        for ii in range(nss):
            if convadj_every_step or (icesnow and ii == 0):
                self.tracer(ii, LICOM_STEP_SIDE_EFFECTS | LICOM_STEP_CONVADJ | ice)
            else:
                self.tracer(ii, LICOM_STEP_SIDE_EFFECTS | ice)
                if ii == 0:
                    self.convadj()

    def sync(self):
        check(lib().licom_tracer_sync(self.ctx), "licom_tracer_sync")

    def derived(self):
        c = self.case
        out = dict(kmt=np.zeros((c.jmt, c.imt), np.int32), kmu=np.zeros((c.jmt, c.imt), np.int32))
        for k in ("tarea_r", "ohbt", "ohbu", "dtn", "dts", "dte", "dtw", "ahf"):
            out[k] = np.zeros((c.jmt, c.imt))
        for k in ("dzp", "odzp", "odzt"):
            out[k] = np.zeros(c.km)
        G = GridDerived()
        G.kmt = out["kmt"].ctypes.data_as(_IP)
        G.kmu = out["kmu"].ctypes.data_as(_IP)
        for k in ("tarea_r", "ohbt", "ohbu", "dtn", "dts", "dte", "dtw", "ahf", "dzp", "odzp", "odzt"):
            setattr(G, k, _d(out[k]))
        check(lib().licom_grid_export(self.ctx, ctypes.byref(G)), "licom_grid_export")
        out["area_t"] = G.area_t
        return out

    def set_profiling(self, on=True):
        check(lib().licom_set_profiling(self.ctx, int(on)), "licom_set_profiling")

    def kernel_times(self):
        n = ctypes.c_int(0)
        names = (ctypes.c_char_p * 64)()
        ms = (ctypes.c_double * 64)()
        cnt = (ctypes.c_int32 * 64)()
        check(lib().licom_kernel_times(self.ctx, 64, names, ms, cnt, ctypes.byref(n)),
              "licom_kernel_times")
        return {names[k].decode(): (ms[k], cnt[k]) for k in range(n.value)}

    def mix_count(self):
        n = ctypes.c_int32(0)
        check(lib().licom_last_mix_count(self.ctx, ctypes.byref(n)), "licom_last_mix_count")
        return int(n.value)

    def limiter_flags(self, on=True):
        """From the next step on, every step writes its tspas limiter flags
        (uint8 [2][km][jmt][imt], device) into self.flags; on=False stops it."""
        if on:
            c = self.case
            self.flags = self.torch.zeros((2, c.km, c.jmt, c.imt), dtype=self.torch.uint8,
                                          device=self.device)
            check(lib().licom_tracer_set_flag_output(self.ctx, _VP(self.flags.data_ptr())),
                  "licom_tracer_set_flag_output")
        else:
            check(lib().licom_tracer_set_flag_output(self.ctx, None), "licom_tracer_set_flag_output")
            self.flags = None
        return self.flags

    def nonfinite(self):
        """Tracer columns whose new state held a NaN or an infinity in the last
        step (0 for a healthy model)."""
        n = ctypes.c_int32(0)
        check(lib().licom_last_nonfinite(self.ctx, ctypes.byref(n)), "licom_last_nonfinite")
        return int(n.value)

    def scratch_bytes(self):
        return int(lib().licom_scratch_bytes(self.ctx))

    def host_state(self):
        self.sync()
        return {k: v.double().cpu().numpy() for k, v in self.state.items()}

    def host_diag(self):
        self.sync()
        return {k: v.double().cpu().numpy() for k, v in self.diag.items()}

    def close(self):
        if getattr(self, "ctx", None):
            self.sync()
            lib().licom_tracer_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def stream_copy_bandwidth(nbytes=4 << 30, reps=5, device="cuda:0"):
    """Achieved GB/s (read + write) of a 16-byte vector copy of `nbytes`."""
    import torch
    src = torch.empty(nbytes // 8, dtype=torch.float64, device=device).fill_(1.0)
    dst = torch.empty_like(src)
    s = torch.cuda.current_stream()
    check(lib().licom_stream_copy(_ptr(dst), _ptr(src), nbytes, _VP(s.cuda_stream)), "copy")
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        check(lib().licom_stream_copy(_ptr(dst), _ptr(src), nbytes, _VP(s.cuda_stream)), "copy")
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    return 2.0 * nbytes / (ms * 1e-3) / 1e9
