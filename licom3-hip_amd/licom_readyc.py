"""Host-side mirror of readyc_'s vertical-mixing preparation over the C ABI
(include/licom_readyc.h, in ``liblicom_tracer.so``).

:class:`Readyc` mirrors the reference's calls:

* constructor  -> the uploads of the U<->T weights and odzt (cuda_init.cpp)
* ``shear()``  -> the clears and the kernels readyc1 and readyc1_2 of
  ``readyc_`` (readyc.cpp:44-170, 1636-1643): h0 time levels, velocities
  averaged to T points, S^2 and the Richardson numbers rit (T grid) and riu
  (U grid) that the Canuto producer reads
* ``akmu(akmt)`` -> readyc2 after turb_cu (readyc.cpp:223-246, 1679-1685)
* ``momentum_grid(mgrid)`` -> the uploads of the momentum metrics
* ``momentum(mfields)`` -> the rest of readyc_ (readyc.cpp:1700-1793):
  upwell_cu_0 / upwell_cu, readyc3, advection_momentum_cu, the del2 or del4
  viscosity, readyc4, readyc4_2

Fields are torch float64 tensors on the GPU in LICOM's block layout
([k][j][i], i fastest).  There is no CPU path.
"""
from __future__ import annotations

import ctypes

import numpy as np

import licom_tracer as lt

_DP = ctypes.POINTER(ctypes.c_double)
_VP = ctypes.c_void_p

GRID_2D = ("at0", "atn", "ate", "atne", "au0", "aus", "auw", "ausw")
FIELDS = ("vit", "viv", "up", "vp", "h0", "rict", "ric", "h0bl", "h0bf",
          "wp12", "wp13", "riv1", "riv2", "s2t", "rit", "s2u", "riu")
OUTPUTS = ("h0bl", "h0bf", "wp12", "wp13", "riv1", "riv2", "s2t", "rit", "s2u", "riu")

MGRID_INTS = ("ib", "ie", "jb", "je", "del4")
MGRID_SCALARS = ("am", "c0f", "od0", "cag", "sag")
MGRID_2D = ("ohbu", "ohbt", "hts", "htw", "hue", "hun", "tarea_r", "uarea_r", "dxu", "dyu", "dxur",
            "dyur", "snlat", "amf", "duc", "dum", "dun", "dus", "due", "duw", "dmc", "dmn", "dms",
            "dme", "dmw")
MFIELDS = ("u", "v", "up", "vp", "h0", "su", "sv", "vit", "viv", "akmu", "ws", "dlu", "dlv", "wka",
           "dlub", "dlvb", "sbcx", "sbcy", "bbcx", "bbcy")
MOUTPUTS = ("ws", "dlu", "dlv", "wka", "dlub", "dlvb", "sbcx", "sbcy", "bbcx", "bbcy")

EXPORTED = ["licom_readyc_create", "licom_readyc_set_stream", "licom_readyc_shear",
            "licom_readyc_akmu", "licom_readyc_momentum_grid", "licom_readyc_momentum",
            "licom_readyc_sync", "licom_readyc_last_ms", "licom_readyc_destroy"]


class ReadycGrid(ctypes.Structure):
    _fields_ = ([("imt", ctypes.c_int), ("jmt", ctypes.c_int), ("km", ctypes.c_int)] +
                [(k, _DP) for k in GRID_2D] + [("odzt", _DP)])


class ReadycFields(ctypes.Structure):
    _fields_ = [(k, _VP) for k in FIELDS]


class ReadycMgrid(ctypes.Structure):
    _fields_ = ([(k, ctypes.c_int) for k in MGRID_INTS] + [(k, ctypes.c_double) for k in MGRID_SCALARS] +
                [("kmt", ctypes.POINTER(ctypes.c_int)), ("kmu", ctypes.POINTER(ctypes.c_int))] +
                [(k, _DP) for k in MGRID_2D] + [("dzp", _DP), ("odzp", _DP)])


class ReadycMfields(ctypes.Structure):
    _fields_ = [(k, _VP) for k in MFIELDS]


_BOUND = False


def lib():
    global _BOUND
    L = lt.lib()
    if not _BOUND:
        L.licom_readyc_create.argtypes = [ctypes.POINTER(ReadycGrid), ctypes.POINTER(_VP)]
        L.licom_readyc_set_stream.argtypes = [_VP, _VP]
        L.licom_readyc_shear.argtypes = [_VP, ctypes.POINTER(ReadycFields)]
        L.licom_readyc_akmu.argtypes = [_VP, _VP, _VP, _VP]
        L.licom_readyc_momentum_grid.argtypes = [_VP, ctypes.POINTER(ReadycMgrid)]
        L.licom_readyc_momentum.argtypes = [_VP, ctypes.POINTER(ReadycMfields)]
        L.licom_readyc_sync.argtypes = [_VP]
        L.licom_readyc_last_ms.argtypes = [_VP, _DP]
        L.licom_readyc_destroy.argtypes = [_VP]
        L.licom_readyc_destroy.restype = None
        _BOUND = True
    return L


def field_shapes(imt, jmt, km):
    s2, s3, sm = (jmt, imt), (km, jmt, imt), (km - 1, jmt, imt)
    sh = {k: s3 for k in ("vit", "viv", "up", "vp", "wp12", "wp13", "riv1", "riv2", "s2t", "s2u",
                          "riu")}
    sh.update({k: s2 for k in ("h0", "h0bl", "h0bf")})
    sh.update({k: sm for k in ("rict", "ric", "rit")})
    return sh


class Readyc:
    """readyc_'s mixing preparation on one block: ``grid`` holds the host
    arrays of licom_readyc_grid; ``fields`` maps every licom_readyc_fields
    name to a float64 CUDA tensor or a numpy array (uploaded)."""

    def __init__(self, imt, jmt, km, grid, fields=None, device="cuda:0", stream=None):
        import torch
        if not torch.cuda.is_available():
            raise lt.LicomError("Readyc needs a GPU (torch.cuda.is_available() is False)")
        self.torch = torch
        self.imt, self.jmt, self.km = int(imt), int(jmt), int(km)
        self.device = torch.device(device)
        torch.cuda.set_device(self.device)
        keep = []
        G = ReadycGrid()
        G.imt, G.jmt, G.km = self.imt, self.jmt, self.km
        for k in GRID_2D + ("odzt",):
            a = np.ascontiguousarray(grid[k], np.float64)
            keep.append(a)
            setattr(G, k, a.ctypes.data_as(_DP))
        h = _VP()
        lt.check(lib().licom_readyc_create(ctypes.byref(G), ctypes.byref(h)), "licom_readyc_create")
        self.handle = h
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.device)
        lt.check(lib().licom_readyc_set_stream(h, _VP(self.stream.cuda_stream)), "licom_readyc_set_stream")
        shapes = field_shapes(self.imt, self.jmt, self.km)
        self.fields = {}
        for k in FIELDS:
            v = None if fields is None else fields.get(k)
            if v is None:
                v = torch.zeros(shapes[k], dtype=torch.float64, device=self.device)
            elif not isinstance(v, torch.Tensor):
                v = torch.from_numpy(np.ascontiguousarray(v, np.float64)).to(self.device)
            if tuple(v.shape) != shapes[k] or v.dtype != torch.float64 or not v.is_contiguous():
                raise lt.LicomError(f"readyc field {k}: need a contiguous float64 {shapes[k]} array")
            self.fields[k] = v

    def shear(self):
        F = ReadycFields(*[self.fields[k].data_ptr() for k in FIELDS])
        lt.check(lib().licom_readyc_shear(self.handle, ctypes.byref(F)), "licom_readyc_shear")

    def akmu(self, akmt, out=None):
        """readyc2: akmt [km][j][i] (a CUDA tensor) averaged to U points."""
        if out is None:
            out = self.torch.empty_like(akmt)
        lt.check(lib().licom_readyc_akmu(self.handle, _VP(akmt.data_ptr()),
                                         _VP(self.fields["viv"].data_ptr()), _VP(out.data_ptr())),
                 "licom_readyc_akmu")
        return out

    def momentum_grid(self, mgrid):
        """Uploads licom_readyc_mgrid: ``mgrid`` holds ib, ie, jb, je (1-based
        block bounds), del4, the scalars, kmt / kmu and the 2-D metrics
        (host arrays), dzp and odzp."""
        keep = []
        M = ReadycMgrid()
        for k in MGRID_INTS:
            setattr(M, k, int(mgrid[k]))
        for k in MGRID_SCALARS:
            setattr(M, k, float(mgrid[k]))
        for k in ("kmt", "kmu"):
            a = np.ascontiguousarray(mgrid[k], np.int32)
            if a.shape != (self.jmt, self.imt):
                raise lt.LicomError(f"readyc momentum grid {k}: need a ({self.jmt}, {self.imt}) array")
            keep.append(a)
            setattr(M, k, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        for k in MGRID_2D + ("dzp", "odzp"):
            a = np.ascontiguousarray(mgrid[k], np.float64)
            want = (self.km,) if k in ("dzp", "odzp") else (self.jmt, self.imt)
            if a.shape != want:
                raise lt.LicomError(f"readyc momentum grid {k}: need a {want} array")
            keep.append(a)
            setattr(M, k, a.ctypes.data_as(_DP))
        lt.check(lib().licom_readyc_momentum_grid(self.handle, ctypes.byref(M)),
                 "licom_readyc_momentum_grid")

    def momentum_fields(self, fields=None):
        """licom_readyc_mfields as float64 CUDA tensors: the given ones (CUDA
        tensors used in place, numpy arrays uploaded), zeros for the rest."""
        torch = self.torch
        s2, s3 = (self.jmt, self.imt), (self.km, self.jmt, self.imt)
        out = {}
        for k in MFIELDS:
            sh = s2 if k in ("h0", "su", "sv", "dlub", "dlvb", "sbcx", "sbcy", "bbcx", "bbcy") else s3
            v = None if fields is None else fields.get(k)
            if v is None:
                v = torch.zeros(sh, dtype=torch.float64, device=self.device)
            elif not isinstance(v, torch.Tensor):
                v = torch.from_numpy(np.ascontiguousarray(v, np.float64)).to(self.device)
            if tuple(v.shape) != sh or v.dtype != torch.float64 or not v.is_contiguous():
                raise lt.LicomError(f"readyc momentum field {k}: need a contiguous float64 {sh} array")
            out[k] = v
        return out

    def momentum(self, mfields):
        """licom_readyc_momentum on a dict from momentum_fields()."""
        F = ReadycMfields(*[mfields[k].data_ptr() for k in MFIELDS])
        lt.check(lib().licom_readyc_momentum(self.handle, ctypes.byref(F)), "licom_readyc_momentum")

    def sync(self):
        lt.check(lib().licom_readyc_sync(self.handle), "licom_readyc_sync")

    def last_ms(self):
        v = ctypes.c_double()
        lt.check(lib().licom_readyc_last_ms(self.handle, ctypes.byref(v)), "licom_readyc_last_ms")
        return v.value

    def host(self, names=None):
        self.sync()
        return {k: self.fields[k].cpu().numpy() for k in (names or FIELDS)}

    def close(self):
        if getattr(self, "handle", None):
            lib().licom_readyc_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
