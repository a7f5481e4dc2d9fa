"""Synthetic workloads for the tracer step (host side, numpy + ctypes).

Binds ``liblicom_synth.so`` (include/licom_synth.h) and packages one workload
as a :class:`Case`: grid metrics, static inputs, the initial state and the
scalar parameters, all as numpy arrays in the reference layout
(i fastest, ``imt = nx+4``, ``jmt = ny+4``).  The same Case feeds the HIP
library and the CPU checker of the test suite.

Workload sizes, seeds and physics follow the reference's tracer path
configuration: tspas advection (``tracer.cpp:1712-1714``), del2 with
``AH_HOR`` or del4 with ``AH_BIHAR`` (``hmix_del2.F90`` / ``hmix_del4.F90``),
``dwndmix = 10e-4`` (``constant_mod.F90:279``), ``OD0``/``OD0CP`` from
``D0 = 1026``, ``CP = 3996`` (``constant_mod.F90:164-173``).
"""
from __future__ import annotations

import ctypes
import dataclasses
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

D0 = 1026.0
CP = 3996.0
# advection schemes by the reference's Str code (tracer.cpp:1712-1714)
ADV_CENTERED, ADV_FLUX, ADV_TSPAS = 0, 1, 2


class SynthCfg(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("km", ctypes.c_int),
        ("lat_mode", ctypes.c_int),
        ("dlon", ctypes.c_double), ("lat_north", ctypes.c_double),
        ("dlat", ctypes.c_double), ("lat_c", ctypes.c_double),
        ("seed", ctypes.c_uint64),
        ("depth_min", ctypes.c_double), ("depth_max", ctypes.c_double),
        ("sigma", ctypes.c_double), ("land_depth", ctypes.c_double),
        ("dz_top", ctypes.c_double), ("dz_bot", ctypes.c_double),
        ("depth_total", ctypes.c_double),
        ("umax", ctypes.c_double), ("div_frac", ctypes.c_double),
        ("t_noise", ctypes.c_double), ("s_noise", ctypes.c_double),
        ("t_pert", ctypes.c_double),
        ("akt_lo", ctypes.c_double), ("akt_hi", ctypes.c_double),
    ]


_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int32)


class SynthGrid(ctypes.Structure):
    _fields_ = [("kmt", _IP), ("zkp", _DP), ("pen", _DP),
                ("htw", _DP), ("hts", _DP), ("hun", _DP), ("hue", _DP),
                ("tarea", _DP), ("dxu", _DP), ("dyu", _DP),
                ("au0", _DP), ("aus", _DP), ("auw", _DP), ("ausw", _DP),
                ("to", _DP), ("so", _DP), ("eosc", _DP), ("lat_t", _DP),
                ("imt_global", ctypes.c_int)]


class SynthState(ctypes.Structure):
    _fields_ = [("at", _DP), ("atb", _DP), ("akt", _DP),
                ("utf", _DP), ("vtf", _DP), ("utl", _DP), ("vtl", _DP),
                ("ws", _DP), ("h0f", _DP), ("h0l", _DP)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liblicom_synth.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run __graft_entry__.build() first")
        _LIB = ctypes.CDLL(path)
        _LIB.licom_synth_defaults.argtypes = [ctypes.POINTER(SynthCfg)]
        _LIB.licom_synth_make_grid.argtypes = [ctypes.POINTER(SynthCfg), ctypes.POINTER(SynthGrid)]
        _LIB.licom_synth_make_state.argtypes = [ctypes.POINTER(SynthCfg), _IP, _DP,
                                                ctypes.POINTER(SynthState)]
    return _LIB


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_DP)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_IP)


# Workload presets.  Grid rows run north to south as in LICOM.
#  C1 regional 0.5 deg centred on 30N, del2 AH=1e3, dt=3600 (10 steps)
#  C2 1 deg global-like, rows uniform over 87N..78S, del2, dt=3600
#  C3 0.1 deg, Mercator rows (square cells, +-74.7 deg), del4 AH=-1e10, dt=600
WORKLOADS = {
    "C1": dict(nx=128, ny=96, km=30, lat_mode=0, dlon=0.5, dlat=0.5, lat_north=54.0,
               seed=101, hmix=2, ah=1.0e3, dts=3600.0),
    "C2": dict(nx=360, ny=218, km=30, lat_mode=0, dlon=1.0, dlat=165.0 / 218.0,
               lat_north=87.0, seed=202, hmix=2, ah=1.0e3, dts=3600.0),
    "C3": dict(nx=3600, ny=2302, km=55, lat_mode=1, dlon=0.1, lat_c=0.0,
               seed=303, hmix=4, ah=-1.0e10, dts=600.0),
}

_CFG_KEYS = {f for f, _ in SynthCfg._fields_}


@dataclasses.dataclass
class Case:
    name: str
    nx: int
    ny: int
    km: int
    imt: int
    jmt: int
    imt_global: int
    grid: dict        # primary grid arrays (numpy)
    state: dict       # at, atb, akt, utf, vtf, utl, vtl, ws, h0f, h0l
    forcing: dict     # swv, nswv, sst, seaice, sss, fresh (2-D), restore (or None)
    params: dict      # scalars

    @property
    def n2(self):
        return self.imt * self.jmt

    @property
    def n3(self):
        return self.imt * self.jmt * self.km

    def copy_state(self):
        return {k: v.copy() for k, v in self.state.items()}


def default_params(hmix=2, ah=1.0e3, dts=3600.0):
    return dict(dts=float(dts), onbc=0.25, oncc=0.5, dwndmix=10.0e-4, gamma=0.0,
                od0=1.0 / D0, od0cp=1.0 / (D0 * CP), ah=float(ah), hmix=int(hmix),
                boundary_restore=0, sssnorm=1, convadj=1, adv=ADV_TSPAS, nss=1)


def make_cfg(**kw) -> SynthCfg:
    cfg = SynthCfg()
    lib().licom_synth_defaults(ctypes.byref(cfg))
    for k, v in kw.items():
        if k in _CFG_KEYS:
            setattr(cfg, k, v)
    return cfg


def make_case(name: str = "C1", **overrides) -> Case:
    spec = dict(WORKLOADS.get(name, {}))
    spec.update(overrides)
    cfg = make_cfg(**spec)
    nx, ny, km = cfg.nx, cfg.ny, cfg.km
    imt, jmt = nx + 4, ny + 4
    n2, n3 = imt * jmt, imt * jmt * km
    z2 = lambda: np.zeros((jmt, imt), np.float64)  # noqa: E731
    grid = dict(kmt=np.zeros((jmt, imt), np.int32), zkp=np.zeros(km + 1), pen=np.zeros(km - 1),
                htw=z2(), hts=z2(), hun=z2(), hue=z2(), tarea=z2(), dxu=z2(), dyu=z2(),
                au0=z2(), aus=z2(), auw=z2(), ausw=z2(), to=np.zeros(km), so=np.zeros(km),
                eosc=np.zeros((9, km)), lat_t=np.zeros(jmt))
    g = SynthGrid()
    for f, _ in SynthGrid._fields_:
        if f == "imt_global":
            continue
        a = grid[f]
        setattr(g, f, iptr(a) if a.dtype == np.int32 else dptr(a))
    rc = lib().licom_synth_make_grid(ctypes.byref(cfg), ctypes.byref(g))
    if rc != 0:
        raise ValueError(f"licom_synth_make_grid failed ({rc})")
    state = dict(at=np.zeros((2, km, jmt, imt)), atb=np.zeros((2, km + 1, jmt, imt)),
                 akt=np.zeros((2, km, jmt, imt)), utf=np.zeros((km, jmt, imt)),
                 vtf=np.zeros((km, jmt, imt)), utl=np.zeros((km, jmt, imt)),
                 vtl=np.zeros((km, jmt, imt)), ws=np.zeros((km + 1, jmt, imt)),
                 h0f=z2(), h0l=z2())
    s = SynthState(**{k: dptr(v) for k, v in state.items()})
    rc = lib().licom_synth_make_state(ctypes.byref(cfg), iptr(grid["kmt"]), dptr(grid["zkp"]),
                                      ctypes.byref(s))
    if rc != 0:
        raise ValueError(f"licom_synth_make_state failed ({rc})")
    forcing = dict(swv=z2(), nswv=z2(), sst=z2(), seaice=z2(), sss=z2(), fresh=z2(), restore=None)
    params = default_params(spec.get("hmix", 2), spec.get("ah", 1.0e3), spec.get("dts", 3600.0))
    for k in list(params):
        if k in overrides:
            params[k] = overrides[k]
    del n2, n3
    return Case(name=name, nx=nx, ny=ny, km=km, imt=imt, jmt=jmt, imt_global=g.imt_global,
                grid=grid, state=state, forcing=forcing, params=params)
