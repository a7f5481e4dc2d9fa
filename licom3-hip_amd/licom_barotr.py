"""Host-side mirror of the barotropic subcycle barotr_ over the C ABI
(include/licom_barotr.h, in ``liblicom_tracer.so``).

:class:`Barotr` mirrors the reference's calls:

* constructor -> the uploads of the grid arrays barotr_ reads (cuda_init.cpp)
* ``stage(nc, s)`` -> the launches of subcycle nc between two halo updates
  (barotr.cpp:901-1024): stage 0 = barotr1_cu + barotr1_cu1, stage 1 =
  barotr2_cu1 + the viscosity (del2 or del4), stage 2 = barotr3_cu (+
  barotr3_tdel4_1) + barotr3_cu1, stage 3 = barotr4_cu.  The halo exchanges
  (of work, of wka planes 2 and 3, of work) are the caller's, as LICOM's MPI
  halo library is.
* ``run(nbb, halo)`` -> the whole of barotr_, calling ``halo(names)`` with
  the names of the fields to exchange between stages

Fields are torch float64 tensors on the GPU, 2-D [j][i]; wka is [6][j][i].
There is no CPU path.
"""
from __future__ import annotations

import ctypes

import numpy as np

import licom_tracer as lt

_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)
_VP = ctypes.c_void_p

GRID_INTS = ("imt", "jmt", "ib", "ie", "jb", "je", "del4")
GRID_SCALARS = ("dtb", "am", "ah")
GRID_I2D = ("kmt", "kmu", "kmtn", "kmts", "kmte", "kmtw")
GRID_D2D = ("au0", "aus", "auw", "ausw", "htw", "hts", "tarea_r", "uarea_r", "dxu", "dyu", "dxur", "dyur",
            "fcor", "ebea", "ebeb", "amf", "ahf", "dtn", "dts", "dte", "dtw", "duc", "dum", "dun", "dus",
            "due", "duw", "dmc", "dmn", "dms", "dme", "dmw")
FIELDS = ("vit", "viv", "dzph", "dlub", "dlvb", "wgp", "pax", "pay", "pxb", "pyb", "whx", "why", "h0", "h0p",
          "ub", "vb", "ubp", "vbp", "h0f", "h0bf", "work", "wka")
OUTPUTS = ("h0", "h0p", "ub", "vb", "ubp", "vbp", "h0f", "h0bf", "work")
HALO = {0: ("work",), 1: ("wka",), 2: ("work",)}   # exchanged after stage s (wka: planes 2, 3)

EXPORTED = ["licom_barotr_create", "licom_barotr_set_stream", "licom_barotr_stage", "licom_barotr_sync",
            "licom_barotr_last_ms", "licom_barotr_destroy"]


class BarotrGrid(ctypes.Structure):
    _fields_ = ([(k, ctypes.c_int) for k in GRID_INTS] + [(k, ctypes.c_double) for k in GRID_SCALARS] +
                [(k, _IP) for k in GRID_I2D] + [(k, _DP) for k in GRID_D2D])


class BarotrFields(ctypes.Structure):
    _fields_ = [(k, _VP) for k in FIELDS]


_BOUND = False


def lib():
    global _BOUND
    L = lt.lib()
    if not _BOUND:
        L.licom_barotr_create.argtypes = [ctypes.POINTER(BarotrGrid), ctypes.POINTER(_VP)]
        L.licom_barotr_set_stream.argtypes = [_VP, _VP]
        L.licom_barotr_stage.argtypes = [_VP, ctypes.POINTER(BarotrFields), ctypes.c_int, ctypes.c_int]
        L.licom_barotr_sync.argtypes = [_VP]
        L.licom_barotr_last_ms.argtypes = [_VP, _DP]
        L.licom_barotr_destroy.argtypes = [_VP]
        L.licom_barotr_destroy.restype = None
        _BOUND = True
    return L


def make_grid(imt, jmt, grid, keep):
    """licom_barotr_grid from a dict of host arrays; ``keep`` holds the
    contiguous copies alive for the call."""
    G = BarotrGrid()
    G.imt, G.jmt = int(imt), int(jmt)
    for k in GRID_INTS[2:]:
        setattr(G, k, int(grid[k]))
    for k in GRID_SCALARS:
        setattr(G, k, float(grid[k]))
    for k in GRID_I2D:
        a = np.ascontiguousarray(grid[k], np.int32)
        keep.append(a)
        setattr(G, k, a.ctypes.data_as(_IP))
    for k in GRID_D2D:
        a = np.ascontiguousarray(grid[k], np.float64)
        keep.append(a)
        setattr(G, k, a.ctypes.data_as(_DP))
    return G


class Barotr:
    """barotr_ on one block: ``grid`` holds the host arrays of
    licom_barotr_grid; ``fields`` maps every licom_barotr_fields name to a
    float64 CUDA tensor (used in place) or a numpy array (uploaded)."""

    def __init__(self, imt, jmt, grid, fields=None, device="cuda:0", stream=None):
        import torch
        if not torch.cuda.is_available():
            raise lt.LicomError("Barotr needs a GPU (torch.cuda.is_available() is False)")
        self.torch = torch
        self.imt, self.jmt = int(imt), int(jmt)
        self.device = torch.device(device)
        torch.cuda.set_device(self.device)
        keep = []
        G = make_grid(imt, jmt, grid, keep)
        h = _VP()
        lt.check(lib().licom_barotr_create(ctypes.byref(G), ctypes.byref(h)), "licom_barotr_create")
        self.handle = h
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.device)
        lt.check(lib().licom_barotr_set_stream(h, _VP(self.stream.cuda_stream)), "licom_barotr_set_stream")
        s2 = (self.jmt, self.imt)
        self.fields = {}
        for k in FIELDS:
            sh = (6,) + s2 if k == "wka" else s2
            v = None if fields is None else fields.get(k)
            if v is None:
                v = torch.zeros(sh, dtype=torch.float64, device=self.device)
            elif not isinstance(v, torch.Tensor):
                v = torch.from_numpy(np.ascontiguousarray(v, np.float64)).to(self.device)
            if tuple(v.shape) != sh or v.dtype != torch.float64 or not v.is_contiguous():
                raise lt.LicomError(f"barotr field {k}: need a contiguous float64 {sh} array")
            self.fields[k] = v

    def stage(self, nc, s):
        F = BarotrFields(*[self.fields[k].data_ptr() for k in FIELDS])
        lt.check(lib().licom_barotr_stage(self.handle, ctypes.byref(F), int(nc), int(s)), "licom_barotr_stage")

    def run(self, nbb, halo=None):
        """nbb subcycles; ``halo(names)`` exchanges the named fields between
        stages (a single block with no neighbours passes None)."""
        for nc in range(nbb):
            for s in range(4):
                self.stage(nc, s)
                if halo is not None and s in HALO:
                    halo(HALO[s])

    def sync(self):
        lt.check(lib().licom_barotr_sync(self.handle), "licom_barotr_sync")

    def last_ms(self):
        v = ctypes.c_double()
        lt.check(lib().licom_barotr_last_ms(self.handle, ctypes.byref(v)), "licom_barotr_last_ms")
        return v.value

    def host(self, names=None):
        self.sync()
        return {k: self.fields[k].cpu().numpy() for k in (names or FIELDS)}

    def close(self):
        if getattr(self, "handle", None):
            lib().licom_barotr_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
