"""Host-side mirror of the baroclinic momentum update bclinc_ over the C ABI
(include/licom_bclinc.h, in ``liblicom_tracer.so``).

:class:`Bclinc` mirrors the reference's calls:

* constructor -> the uploads of the grid arrays bclinc_ reads (cuda_init.cpp)
* ``stage(isc, s)`` -> the part of bclinc_ between two halo updates
  (bclinc.cpp:728-827): stage 0 = bclinc12, bclinc11, then bclinc2 +
  bclinc3 (isc >= 1) or bclinc6 + bclinc71 + bclinc72 (isc < 1); stage 1 =
  bclinc4 or bclinc81 + bclinc82; stage 2 = bclinc5 (isc >= 1).  The halo
  exchanges between stages (of wka, or of u and v on the first step) are the
  caller's, as LICOM's MPI halo library is.
* ``step(isc, halo)`` -> the whole of bclinc_, calling ``halo(names)`` with
  the names of the fields to exchange between stages

Fields are torch float64 tensors on the GPU in LICOM's block layout
([k][j][i], i fastest).  There is no CPU path.
"""
from __future__ import annotations

import ctypes

import numpy as np

import licom_tracer as lt

_DP = ctypes.POINTER(ctypes.c_double)
_VP = ctypes.c_void_p

GRID_SCALARS = ("od0", "c0f", "cag", "sag", "onbb", "dtc", "dtc2", "afc1", "afc2")
GRID_2D = ("snlat", "ohbt", "ohbu", "fcor", "dxur", "dyur", "epea", "epeb", "epla", "eplb")
GRID_1D = ("dzp", "odzt", "odzp", "zkt")
FIELDS = ("vit", "viv", "gg", "akmu", "su", "sv", "psa", "h0bl", "ub", "vb", "h0bf", "u", "v", "up",
          "vp", "dlu", "dlv", "wka", "utf", "vtf", "sbcx", "sbcy", "bbcx", "bbcy")
FIELDS_2D = ("su", "sv", "psa", "h0bl", "ub", "vb", "h0bf", "sbcx", "sbcy", "bbcx", "bbcy")
OUTPUTS = ("h0bf", "u", "v", "up", "vp", "dlu", "dlv", "wka", "utf", "vtf", "sbcx", "sbcy", "bbcx", "bbcy")

EXPORTED = ["licom_bclinc_create", "licom_bclinc_set_stream", "licom_bclinc_stage", "licom_bclinc_sync",
            "licom_bclinc_last_ms", "licom_bclinc_destroy"]


class BclincGrid(ctypes.Structure):
    _fields_ = ([(k, ctypes.c_int) for k in ("imt", "jmt", "km")] +
                [(k, ctypes.c_double) for k in GRID_SCALARS] +
                [("kmu", ctypes.POINTER(ctypes.c_int))] + [(k, _DP) for k in GRID_2D + GRID_1D])


class BclincFields(ctypes.Structure):
    _fields_ = [(k, _VP) for k in FIELDS]


_BOUND = False


def lib():
    global _BOUND
    L = lt.lib()
    if not _BOUND:
        L.licom_bclinc_create.argtypes = [ctypes.POINTER(BclincGrid), ctypes.POINTER(_VP)]
        L.licom_bclinc_set_stream.argtypes = [_VP, _VP]
        L.licom_bclinc_stage.argtypes = [_VP, ctypes.POINTER(BclincFields), ctypes.c_int, ctypes.c_int]
        L.licom_bclinc_sync.argtypes = [_VP]
        L.licom_bclinc_last_ms.argtypes = [_VP, _DP]
        L.licom_bclinc_destroy.argtypes = [_VP]
        L.licom_bclinc_destroy.restype = None
        _BOUND = True
    return L


def make_grid(imt, jmt, km, grid, keep):
    """licom_bclinc_grid from a dict of host arrays; ``keep`` holds the
    contiguous copies alive for the call."""
    G = BclincGrid()
    G.imt, G.jmt, G.km = int(imt), int(jmt), int(km)
    for k in GRID_SCALARS:
        setattr(G, k, float(grid[k]))
    kmu = np.ascontiguousarray(grid["kmu"], np.int32)
    keep.append(kmu)
    G.kmu = kmu.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    for k in GRID_2D + GRID_1D:
        a = np.ascontiguousarray(grid[k], np.float64)
        keep.append(a)
        setattr(G, k, a.ctypes.data_as(_DP))
    return G


class Bclinc:
    """bclinc_ on one block: ``grid`` holds the host arrays of
    licom_bclinc_grid; ``fields`` maps every licom_bclinc_fields name to a
    float64 CUDA tensor (used in place) or a numpy array (uploaded)."""

    def __init__(self, imt, jmt, km, grid, fields=None, device="cuda:0", stream=None):
        import torch
        if not torch.cuda.is_available():
            raise lt.LicomError("Bclinc needs a GPU (torch.cuda.is_available() is False)")
        self.torch = torch
        self.imt, self.jmt, self.km = int(imt), int(jmt), int(km)
        self.device = torch.device(device)
        torch.cuda.set_device(self.device)
        keep = []
        G = make_grid(imt, jmt, km, grid, keep)
        h = _VP()
        lt.check(lib().licom_bclinc_create(ctypes.byref(G), ctypes.byref(h)), "licom_bclinc_create")
        self.handle = h
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.device)
        lt.check(lib().licom_bclinc_set_stream(h, _VP(self.stream.cuda_stream)), "licom_bclinc_set_stream")
        s2, s3 = (self.jmt, self.imt), (self.km, self.jmt, self.imt)
        self.fields = {}
        for k in FIELDS:
            sh = s2 if k in FIELDS_2D else s3
            v = None if fields is None else fields.get(k)
            if v is None:
                v = torch.zeros(sh, dtype=torch.float64, device=self.device)
            elif not isinstance(v, torch.Tensor):
                v = torch.from_numpy(np.ascontiguousarray(v, np.float64)).to(self.device)
            if tuple(v.shape) != sh or v.dtype != torch.float64 or not v.is_contiguous():
                raise lt.LicomError(f"bclinc field {k}: need a contiguous float64 {sh} array")
            self.fields[k] = v

    def stage(self, isc, s):
        F = BclincFields(*[self.fields[k].data_ptr() for k in FIELDS])
        lt.check(lib().licom_bclinc_stage(self.handle, ctypes.byref(F), int(isc), int(s)), "licom_bclinc_stage")

    def step(self, isc, halo=None):
        """The whole of bclinc_; ``halo(names)`` exchanges the named fields
        between stages (a single block with no neighbours passes None)."""
        self.stage(isc, 0)
        if halo is not None:
            halo(("u", "v") if isc < 1 else ("wka",))
        self.stage(isc, 1)
        if isc >= 1:
            if halo is not None:
                halo(("wka",))
            self.stage(isc, 2)

    def sync(self):
        lt.check(lib().licom_bclinc_sync(self.handle), "licom_bclinc_sync")

    def last_ms(self):
        v = ctypes.c_double()
        lt.check(lib().licom_bclinc_last_ms(self.handle, ctypes.byref(v)), "licom_bclinc_last_ms")
        return v.value

    def host(self, names=None):
        self.sync()
        return {k: self.fields[k].cpu().numpy() for k in (names or FIELDS)}

    def close(self):
        if getattr(self, "handle", None):
            lib().licom_bclinc_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
