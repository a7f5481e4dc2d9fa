"""Host-side mirror of the reference's ``readyt_`` over the C ABI
(include/licom_readyt.h, in ``liblicom_tracer.so``).

:class:`Readyt` mirrors the reference's calls:

* constructor     -> the uploads of the grid and pconst arrays readyt_'s
  kernels read (cuda_init.cpp:708-864)
* ``readyt(ist)`` -> ``readyt_()`` (readyt.cpp:511-598): time-level shift of
  h0, utf, vtf; potential density, hydrostatic pressure integrals, thermal
  expansion and saline contraction; the Richardson-number terms the Canuto
  producer reads (rict, ricdttms); the baroclinic and barotropic pressure
  gradients; the clears of akt, rit (and ax, ay, az at ist == 0)

Fields are torch float64 tensors on the GPU in LICOM's block layout ([k][j][i],
i fastest).  There is no CPU path: without the library or a GPU every entry
point raises.
"""
from __future__ import annotations

import ctypes

import numpy as np

import licom_tracer as lt

_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int32)
_VP = ctypes.c_void_p

GRID_2D = ("au0", "aus", "auw", "ausw", "dxur", "dyur", "ohbu", "ohbt", "hbx", "hby")
GRID_1D = ("dzp", "odzt", "zkt", "to", "so", "po")
# licom_readyt_fields, in declaration order
INPUTS = ("vit", "viv", "at", "atb", "h0", "u", "v", "psa", "swv", "nswv")
SHIFTED = ("h0l", "h0f", "utl", "vtl", "utf", "vtf")
OUTPUTS = ("dlu", "dlv", "pdensity", "gg", "pp", "ppa", "ppb", "ppc", "alpha", "beta",
           "ric", "rict", "ricdttms", "ricdt", "rit", "akt", "buoytur", "buoysol", "pxb", "pyb",
           "wgp", "work", "whx", "why", "work1", "work2", "pax", "pay")
OPTIONAL = ("ax", "ay", "az")
FIELDS = INPUTS + SHIFTED + OUTPUTS + OPTIONAL

EXPORTED = ["licom_readyt_create", "licom_readyt_set_stream", "licom_readyt_run",
            "licom_readyt_sync", "licom_readyt_last_ms", "licom_readyt_destroy"]


class ReadytGrid(ctypes.Structure):
    _fields_ = ([("imt", ctypes.c_int), ("jmt", ctypes.c_int), ("km", ctypes.c_int), ("kmu", _IP)] +
                [(k, _DP) for k in GRID_2D + GRID_1D] +
                [("eosc", _DP), ("od0", ctypes.c_double), ("od0cp", ctypes.c_double)])


class ReadytFields(ctypes.Structure):
    _fields_ = [(k, _VP) for k in FIELDS]


_BOUND = False


def lib():
    global _BOUND
    L = lt.lib()
    if not _BOUND:
        L.licom_readyt_create.argtypes = [ctypes.POINTER(ReadytGrid), ctypes.POINTER(_VP)]
        L.licom_readyt_set_stream.argtypes = [_VP, _VP]
        L.licom_readyt_run.argtypes = [_VP, ctypes.POINTER(ReadytFields), ctypes.c_int]
        L.licom_readyt_sync.argtypes = [_VP]
        L.licom_readyt_last_ms.argtypes = [_VP, _DP]
        L.licom_readyt_destroy.argtypes = [_VP]
        L.licom_readyt_destroy.restype = None
        _BOUND = True
    return L


def field_shapes(imt, jmt, km):
    """Shape of every licom_readyt_fields array."""
    s2, s3, s4 = (jmt, imt), (km, jmt, imt), (2, km, jmt, imt)
    sh = {k: s2 for k in ("h0", "psa", "swv", "nswv", "h0l", "h0f", "buoytur", "buoysol", "pxb",
                          "pyb", "wgp", "work", "whx", "why", "work1", "work2", "pax", "pay")}
    sh.update({k: s3 for k in ("vit", "viv", "u", "v", "utl", "vtl", "utf", "vtf", "dlu", "dlv",
                               "pdensity", "gg", "pp", "ppa", "ppb", "ppc", "alpha", "beta")})
    sh.update({k: (km - 1, jmt, imt) for k in ("ric", "rict", "ricdttms", "ricdt", "rit")})
    sh.update({k: s4 for k in ("at", "akt", "ax", "ay", "az")})
    sh["atb"] = (2, km + 1, jmt, imt)
    return sh


class Readyt:
    """readyt_ on one block.  ``grid`` holds the host arrays of
    licom_readyt_grid (kmu, the 2-D and 1-D fields, eosc [9][km], od0,
    od0cp); ``fields`` maps every licom_readyt_fields name (ax, ay, az
    optional) to a float64 CUDA tensor or a numpy array, which is uploaded."""

    def __init__(self, imt, jmt, km, grid, fields=None, device="cuda:0", stream=None):
        import torch
        if not torch.cuda.is_available():
            raise lt.LicomError("Readyt needs a GPU (torch.cuda.is_available() is False)")
        self.torch = torch
        self.imt, self.jmt, self.km = int(imt), int(jmt), int(km)
        self.device = torch.device(device)
        torch.cuda.set_device(self.device)
        L = lib()
        self._keep = []

        def host(a, dt=np.float64):
            a = np.ascontiguousarray(a, dt)
            self._keep.append(a)
            return a

        G = ReadytGrid()
        G.imt, G.jmt, G.km = self.imt, self.jmt, self.km
        G.kmu = host(grid["kmu"], np.int32).ctypes.data_as(_IP)
        for k in GRID_2D + GRID_1D + ("eosc",):
            setattr(G, k, host(grid[k]).ctypes.data_as(_DP))
        G.od0, G.od0cp = float(grid["od0"]), float(grid["od0cp"])
        h = _VP()
        lt.check(L.licom_readyt_create(ctypes.byref(G), ctypes.byref(h)), "licom_readyt_create")
        self.handle = h
        self._keep = []
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.device)
        lt.check(L.licom_readyt_set_stream(h, _VP(self.stream.cuda_stream)), "licom_readyt_set_stream")
        self.fields = {}
        shapes = field_shapes(self.imt, self.jmt, self.km)
        for k in FIELDS:
            v = None if fields is None else fields.get(k)
            if v is None:
                if k in OPTIONAL:
                    self.fields[k] = None
                    continue
                v = torch.zeros(shapes[k], dtype=torch.float64, device=self.device)
            elif not isinstance(v, torch.Tensor):
                v = torch.from_numpy(np.ascontiguousarray(v, np.float64)).to(self.device)
            if tuple(v.shape) != shapes[k] or v.dtype != torch.float64 or not v.is_contiguous():
                raise lt.LicomError(f"readyt field {k}: need a contiguous float64 {shapes[k]} array")
            self.fields[k] = v

    def _struct(self):
        F = ReadytFields()
        for k in FIELDS:
            v = self.fields[k]
            setattr(F, k, None if v is None else v.data_ptr())
        return F

    def readyt(self, ist=1):
        """readyt_() (readyt.cpp:511-598); ist == 0 also clears ax, ay, az."""
        lt.check(lib().licom_readyt_run(self.handle, ctypes.byref(self._struct()), int(ist)),
                 "licom_readyt_run")

    def sync(self):
        lt.check(lib().licom_readyt_sync(self.handle), "licom_readyt_sync")

    def last_ms(self):
        """(whole run, column pass, pressure-gradient pass, surface pass) in ms."""
        v = (ctypes.c_double * 4)()
        lt.check(lib().licom_readyt_last_ms(self.handle, v), "licom_readyt_last_ms")
        return tuple(v)

    def host(self, names=None):
        self.sync()
        return {k: self.fields[k].cpu().numpy() for k in (names or FIELDS) if self.fields[k] is not None}

    def close(self):
        if getattr(self, "handle", None):
            lib().licom_readyt_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
