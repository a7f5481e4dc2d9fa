"""Host-side mirror of the reference's Canuto mixing producer over the C ABI
(include/licom_turb.h, in ``liblicom_tracer.so``).

:class:`TurbProducer` mirrors the reference's calls:

* constructor  -> ``turbinit()`` (turb_init.cpp:37-88): uploads the canuto_mod
  lookup tables, turb_param and fcort
* ``run(f)``   -> the launch of ``turb_cu`` in ``readyc`` (readyc.cpp:1666-1678)
  with the clears before it: AKMT, AKT (both tracers), the tidal term, the
  mixed-layer depth and the tidal diagnostics

Field dictionaries hold torch float64 tensors on the GPU in LICOM's block
layout ([k][j][i], i fastest).  There is no CPU path: without the library
or a GPU every entry point raises.
"""
from __future__ import annotations

import ctypes

import numpy as np

import licom_tracer as lt

_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int32)
_VP = ctypes.c_void_p

MT, MT0, NTH = 1462, 400, 75      # canuto_mod.h: mt, mt0, n_theta_r_oct
W = 2 * MT + 1
NB = 4 * NTH + 1

TABLE_ARRAYS = ("rib", "irimax", "slq2b", "smb", "shb", "ssb", "and2on2a1", "amtaun2a1", "sma1",
                "sha1", "ssa1", "back_ra_r", "sm_r1", "sh_r1", "ss_r1", "slq2_r1")
TABLE_SCALARS = ("rri", "rnd2on2", "dri", "deltheta_r", "b1", "theta_rcrp", "theta_rcrn",
                 "dand2on2")
INPUTS = ("at", "pdensity", "rit", "ricdttms", "s2t", "rict", "fz_tide", "vit",
          "buoytur", "buoysol", "wave_dis")
OUTPUTS = ("akmt", "akt", "ak_tide", "amld", "richardson", "fztidal", "wp3_tidal")

EXPORTED = ["licom_turb_create", "licom_turb_set_stream", "licom_turb_run", "licom_turb_sync",
            "licom_turb_last_ms", "licom_turb_destroy"]


class Tables(ctypes.Structure):
    _fields_ = ([("rib", _DP), ("irimax", _IP)] +
                [(k, _DP) for k in TABLE_ARRAYS[2:]] +
                [(k, ctypes.c_double) for k in TABLE_SCALARS])


class TurbFields(ctypes.Structure):
    _fields_ = [(k, _VP) for k in INPUTS + OUTPUTS]


_BOUND = False


def lib():
    global _BOUND
    L = lt.lib()
    if not _BOUND:
        L.licom_turb_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _IP, _DP, _DP,
                                        _DP, ctypes.POINTER(Tables), ctypes.c_int, ctypes.c_double,
                                        ctypes.c_double, ctypes.POINTER(_VP)]
        L.licom_turb_set_stream.argtypes = [_VP, _VP]
        L.licom_turb_run.argtypes = [_VP, ctypes.POINTER(TurbFields)]
        L.licom_turb_sync.argtypes = [_VP]
        L.licom_turb_last_ms.argtypes = [_VP, _DP]
        L.licom_turb_destroy.argtypes = [_VP]
        L.licom_turb_destroy.restype = None
        _BOUND = True
    return L


def make_tables(tab: dict, keep: list) -> Tables:
    """ctypes view of a table dictionary (numpy arrays + scalars)."""
    T = Tables()
    for k in TABLE_ARRAYS:
        a = np.ascontiguousarray(tab[k], np.int32 if k == "irimax" else np.float64)
        keep.append(a)
        setattr(T, k, a.ctypes.data_as(_IP if k == "irimax" else _DP))
    for k in TABLE_SCALARS:
        setattr(T, k, float(tab[k]))
    return T


class TurbProducer:
    """Device-resident Canuto producer for one block (turbinit + turb_cu)."""

    def __init__(self, tc, device="cuda:0", stream=None):
        import torch
        if not torch.cuda.is_available():
            raise lt.LicomError("TurbProducer needs a GPU (torch.cuda.is_available() is False)")
        self.torch = torch
        self.device = torch.device(device)
        torch.cuda.set_device(self.device)
        L = lib()
        self.imt, self.jmt, self.km = tc.imt, tc.jmt, tc.km
        keep = []
        T = make_tables(tc.tables, keep)
        g = {k: np.ascontiguousarray(tc.grid[k], np.float64) for k in ("zkp", "dzp", "fcort")}
        g["kmt"] = np.ascontiguousarray(tc.grid["kmt"], np.int32)
        h = _VP()
        lt.check(L.licom_turb_create(self.imt, self.jmt, self.km, g["kmt"].ctypes.data_as(_IP),
                                     g["zkp"].ctypes.data_as(_DP), g["dzp"].ctypes.data_as(_DP),
                                     g["fcort"].ctypes.data_as(_DP), ctypes.byref(T),
                                     int(tc.params["ncc"]), float(tc.params["dfricmx"]),
                                     float(tc.params["dwndmix"]), ctypes.byref(h)),
                 "licom_turb_create")
        self.handle = h
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.device)
        lt.check(L.licom_turb_set_stream(self.handle, _VP(self.stream.cuda_stream)),
                 "licom_turb_set_stream")
        self.inputs = {}
        for k in INPUTS:
            v = tc.fields[k]
            if isinstance(v, torch.Tensor):
                assert v.dtype == torch.float64 and v.is_contiguous() and v.device == self.device
                self.inputs[k] = v
            else:
                self.inputs[k] = torch.from_numpy(np.ascontiguousarray(v)).to(self.device)
        km, jmt, imt = self.km, self.jmt, self.imt
        z = lambda *s: torch.zeros(s, dtype=torch.float64, device=self.device)  # noqa: E731
        self.outputs = dict(akmt=z(km, jmt, imt), akt=z(2, km, jmt, imt), ak_tide=z(km, jmt, imt),
                            amld=z(jmt, imt), richardson=z(km, jmt, imt), fztidal=z(km, jmt, imt),
                            wp3_tidal=z(km, jmt, imt))

    def _fields(self, akt=None):
        F = TurbFields()
        for k in INPUTS:
            setattr(F, k, _VP(self.inputs[k].data_ptr()))
        for k in OUTPUTS:
            t = akt if (k == "akt" and akt is not None) else self.outputs[k]
            setattr(F, k, _VP(t.data_ptr()))
        return F

    def run(self, akt=None):
        """turb_cu; AKT goes to ``akt`` (a [2][km][jmt][imt] float64 tensor,
        e.g. a TracerModel's state["akt"]) when given."""
        if akt is not None:
            assert akt.dtype == self.torch.float64 and akt.is_contiguous()
            assert tuple(akt.shape) == (2, self.km, self.jmt, self.imt)
        lt.check(lib().licom_turb_run(self.handle, ctypes.byref(self._fields(akt))),
                 "licom_turb_run")

    def sync(self):
        lt.check(lib().licom_turb_sync(self.handle), "licom_turb_sync")

    def last_ms(self):
        v = ctypes.c_double()
        lt.check(lib().licom_turb_last_ms(self.handle, ctypes.byref(v)), "licom_turb_last_ms")
        return v.value

    def host_outputs(self):
        self.sync()
        return {k: v.cpu().numpy() for k, v in self.outputs.items()}

    def close(self):
        if self.handle:
            lib().licom_turb_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
