"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  See oracle.h for the reference
statements it restates.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int32)


class OrcGrid(ctypes.Structure):
    _fields_ = [("imt", ctypes.c_int), ("jmt", ctypes.c_int), ("km", ctypes.c_int),
                ("imt_global", ctypes.c_int), ("hmix", ctypes.c_int),
                ("kmt", _IP), ("zkp", _DP), ("pen", _DP),
                ("htw", _DP), ("hts", _DP), ("hun", _DP), ("hue", _DP), ("tarea", _DP),
                ("au0", _DP), ("aus", _DP), ("auw", _DP), ("ausw", _DP),
                ("to", _DP), ("so", _DP), ("eosc", _DP),
                ("kmu", _IP), ("vit", _DP), ("viv", _DP),
                ("tarea_r", _DP), ("ohbt", _DP), ("ohbu", _DP),
                ("dtn", _DP), ("dts", _DP), ("dte", _DP), ("dtw", _DP), ("ahf", _DP),
                ("dzp", _DP), ("odzp", _DP), ("odzt", _DP),
                ("area_t", ctypes.c_double), ("dxu", _DP), ("dyu", _DP)]


class OrcParams(ctypes.Structure):
    _fields_ = [("dts", ctypes.c_double), ("onbc", ctypes.c_double), ("oncc", ctypes.c_double),
                ("dwndmix", ctypes.c_double), ("gamma", ctypes.c_double),
                ("od0", ctypes.c_double), ("od0cp", ctypes.c_double), ("ah", ctypes.c_double),
                ("boundary_restore", ctypes.c_int), ("sssnorm", ctypes.c_int),
                ("convadj", ctypes.c_int), ("adv", ctypes.c_int), ("nss", ctypes.c_int)]


class OrcFields(ctypes.Structure):
    _fields_ = [("at", _DP), ("atb", _DP), ("akt", _DP),
                ("utf", _DP), ("vtf", _DP), ("utl", _DP), ("vtl", _DP), ("ws", _DP),
                ("h0f", _DP), ("h0l", _DP),
                ("swv", _DP), ("nswv", _DP), ("sst", _DP), ("seaice", _DP), ("sss", _DP),
                ("fresh", _DP), ("restore", _DP)]


class OrcDiag(ctypes.Structure):
    _fields_ = [("ax", _DP), ("ay", _DP), ("az", _DP), ("dz", _DP),
                ("tend", _DP), ("dt_diff", _DP), ("dt_conv", _DP),
                ("penetrate", _DP), ("net", _DP), ("fw_norm2", ctypes.c_double),
                ("tf", _DP)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run `make -C oracle`")
        _LIB = ctypes.CDLL(path)
        _LIB.orc_derive.argtypes = [ctypes.POINTER(OrcGrid)]
        _LIB.orc_tracer_step.argtypes = [ctypes.POINTER(OrcGrid), ctypes.POINTER(OrcParams),
                                         ctypes.POINTER(OrcFields), ctypes.POINTER(OrcDiag)]
        _LIB.orc_convadj.argtypes = _LIB.orc_tracer_step.argtypes
        _LIB.orc_invtri_column.argtypes = [ctypes.c_int, _DP, _DP, _DP, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_double, _DP]
        _LIB.orc_dens.argtypes = [_DP, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        _LIB.orc_dens.restype = ctypes.c_double
        _LIB.orc_halo_fill_f64.argtypes = [_DP, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        _LIB.orc_icesnow.argtypes = [ctypes.POINTER(OrcGrid), _DP, _DP, ctypes.c_int]
        _LIB.orc_nthreads.restype = ctypes.c_int
        _LIB.orc_set_nthreads.argtypes = [ctypes.c_int]
        _LIB.orc_advect.argtypes = [ctypes.POINTER(OrcGrid), ctypes.c_double, _DP, _DP, _DP, _DP, _DP]
        _LIB.orc_advect_cf.argtypes = [ctypes.POINTER(OrcGrid), ctypes.c_int, ctypes.c_int,
                                       _DP, _DP, _DP, _DP, _DP, _DP, _DP, _DP]
        _LIB.orc_faces.argtypes = [ctypes.POINTER(OrcGrid), ctypes.c_int, _DP, _DP, _DP, _DP]
        _LIB.orc_del2_level.argtypes = [ctypes.POINTER(OrcGrid), _DP, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_double, _DP]
    return _LIB


def _d(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous, a.dtype
    return a.ctypes.data_as(_DP)


def _i(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_IP)


class Oracle:
    """Derived grid + step driver over numpy arrays of one Case."""

    def __init__(self, case):
        self.case = case
        imt, jmt, km = case.imt, case.jmt, case.km
        g = case.grid
        self.kmt = g["kmt"].copy()
        z2 = lambda: np.zeros((jmt, imt))  # noqa: E731
        self.d = dict(kmu=np.zeros((jmt, imt), np.int32), vit=np.zeros((km, jmt, imt)),
                      viv=np.zeros((km, jmt, imt)), tarea_r=z2(), ohbt=z2(), ohbu=z2(),
                      dtn=z2(), dts=z2(), dte=z2(), dtw=z2(), ahf=z2(),
                      dzp=np.zeros(km), odzp=np.zeros(km), odzt=np.zeros(km))
        gr = OrcGrid()
        gr.imt, gr.jmt, gr.km = imt, jmt, km
        gr.imt_global = case.imt_global
        gr.hmix = int(case.params["hmix"])
        gr.kmt = _i(self.kmt)
        for f in ("zkp", "pen", "htw", "hts", "hun", "hue", "tarea", "au0", "aus", "auw", "ausw",
                  "to", "so", "eosc", "dxu", "dyu"):
            setattr(gr, f, _d(g[f]))
        gr.kmu = _i(self.d["kmu"])
        for f in ("vit", "viv", "tarea_r", "ohbt", "ohbu", "dtn", "dts", "dte", "dtw", "ahf",
                  "dzp", "odzp", "odzt"):
            setattr(gr, f, _d(self.d[f]))
        self.g = gr
        lib().orc_derive(ctypes.byref(gr))
        self.area_t = gr.area_t

    def advect(self, T, uw, vs, ws, dts):
        adv = np.zeros_like(T)
        args = [np.ascontiguousarray(a, np.float64) for a in (T, uw, vs, ws)]
        lib().orc_advect(ctypes.byref(self.g), float(dts), *[_d(a) for a in args], _d(adv))
        return adv

    def faces(self, adv, wkd, wkb):
        """advection_tracer_1_0 face transports (uw, vs) of scheme adv."""
        uw, vs = np.zeros_like(wkd), np.zeros_like(wkb)
        a, b = (np.ascontiguousarray(x, np.float64) for x in (wkd, wkb))
        lib().orc_faces(ctypes.byref(self.g), int(adv), _d(a), _d(b), _d(uw), _d(vs))
        return uw, vs

    def advect_cf(self, adv, T, uw, vs, ws, nss=1, ax=None, ay=None, az=None):
        """Centered (adv=0) or flux (adv=1) advection of one tracer; ax, ay,
        az (or None) accumulate the centered increments / nss in place."""
        out = np.zeros_like(T)
        args = [np.ascontiguousarray(a, np.float64) for a in (T, uw, vs, ws)]
        lib().orc_advect_cf(ctypes.byref(self.g), int(adv), int(nss), *[_d(a) for a in args], _d(out),
                            _d(ax), _d(ay), _d(az))
        return out

    def del2_level(self, X, k, use_ahf=False, ah=1.0):
        out = np.zeros_like(X)
        Xc = np.ascontiguousarray(X, np.float64)
        lib().orc_del2_level(ctypes.byref(self.g), _d(Xc), int(k), int(use_ahf), float(ah), _d(out))
        return out

    def params(self, **over):
        p = dict(self.case.params)
        p.update(over)
        op = OrcParams()
        for f, _ in OrcParams._fields_:
            setattr(op, f, p[f])
        return op

    @staticmethod
    def new_diag(case):
        km, jmt, imt = case.km, case.jmt, case.imt
        z = lambda *s: np.zeros(s)  # noqa: E731
        return dict(ax=z(2, km, jmt, imt), ay=z(2, km, jmt, imt), az=z(2, km, jmt, imt),
                    dz=z(2, km, jmt, imt), tend=z(2, km, jmt, imt), dt_diff=z(2, km, jmt, imt),
                    dt_conv=z(2, km, jmt, imt), penetrate=z(km, jmt, imt), net=z(2, jmt, imt))

    def _fields(self, st, forcing):
        f = OrcFields()
        for k in ("at", "atb", "akt", "utf", "vtf", "utl", "vtl", "ws", "h0f", "h0l"):
            setattr(f, k, _d(st[k]))
        for k in ("swv", "nswv", "sst", "seaice", "sss", "fresh"):
            setattr(f, k, _d(forcing[k]))
        f.restore = _d(forcing.get("restore"))
        return f

    def step(self, st, diag=None, forcing=None, **over):
        """Advance `st` (dict of numpy arrays, modified in place) by one step."""
        forcing = forcing or self.case.forcing
        diag = diag if diag is not None else self.new_diag(self.case)
        od = OrcDiag(**{k: _d(v) for k, v in diag.items() if k != "fw_norm2"})
        lib().orc_tracer_step(ctypes.byref(self.g), ctypes.byref(self.params(**over)),
                              ctypes.byref(self._fields(st, forcing)), ctypes.byref(od))
        diag["fw_norm2"] = od.fw_norm2
        return diag

    def icesnow(self, st, licomqice, reset=True):
        """icesnow_ (icesnow.cpp:15-85) on st["at"] in place; licomqice
        [jmt][imt] is zeroed first when reset (the FRC_CORE build)."""
        assert st["at"].flags.c_contiguous and licomqice.flags.c_contiguous
        lib().orc_icesnow(ctypes.byref(self.g), _d(st["at"]), _d(licomqice), int(bool(reset)))
        return licomqice

    def convadj(self, st, diag=None, **over):
        diag = diag if diag is not None else self.new_diag(self.case)
        od = OrcDiag(**{k: _d(v) for k, v in diag.items() if k != "fw_norm2"})
        lib().orc_convadj(ctypes.byref(self.g), ctypes.byref(self.params(**over)),
                          ctypes.byref(self._fields(st, self.case.forcing)), ctypes.byref(od))
        return diag


def invtri_column(kz, wkc, odzt, odzp, dts, stf, vit0, d):
    x = np.ascontiguousarray(d, dtype=np.float64).copy()
    lib().orc_invtri_column(int(kz), _d(np.ascontiguousarray(wkc, np.float64)),
                            _d(np.ascontiguousarray(odzt, np.float64)),
                            _d(np.ascontiguousarray(odzp, np.float64)), float(dts), float(stf),
                            float(vit0), _d(x))
    return x


def dens(eosc, k, t, s):
    e = np.ascontiguousarray(eosc, np.float64)
    return lib().orc_dens(_d(e), e.shape[1], int(k), float(t), float(s))


def halo_fill(a):
    a = np.ascontiguousarray(a)
    jmt, imt = a.shape[-2], a.shape[-1]
    nz = int(np.prod(a.shape[:-2])) if a.ndim > 2 else 1
    lib().orc_halo_fill_f64(_d(a), imt, jmt, nz)
    return a


def content(case, oracle, st):
    """Compensated T and S content sums sum(dzp*tarea*vit*at) over the
    physical cells (the quantity energy_c.c:81-82 prints), via math.fsum."""
    import math
    j0, j1, i0, i1 = 2, case.jmt - 2, 2, case.imt - 2
    w = (oracle.d["dzp"][:, None, None] * case.grid["tarea"][None] * oracle.d["vit"])[:, j0:j1, i0:i1]
    out = []
    for n in range(2):
        out.append(math.fsum((w * st["at"][n][:, j0:j1, i0:i1]).ravel().tolist()))
    return out


# ---------------------------------------------------------------- Canuto mixing (turb.h)
TURB_MT, TURB_MT0, TURB_NTH = 1462, 400, 75
TURB_COUNTERS = ("pure", "table2d", "below", "deep", "epsy", "deardorff", "rafglt", "theta_p",
                 "theta_n", "exit")
_TURB_ARRAYS = ("rib", "irimax", "slq2b", "smb", "shb", "ssb", "and2on2a1", "amtaun2a1", "sma1",
                "sha1", "ssa1", "back_ra_r", "sm_r1", "sh_r1", "ss_r1", "slq2_r1")
_TURB_SCALARS = ("rri", "rnd2on2", "dri", "deltheta_r", "b1", "theta_rcrp", "theta_rcrn",
                 "dand2on2")
_TURB_IN = ("at", "pdensity", "rit", "ricdttms", "s2t", "rict", "fz_tide", "vit",
            "buoytur", "buoysol", "fcort", "wave_dis")
_TURB_OUT = ("akmt", "akt", "ak_tide", "amld", "richardson", "fztidal", "wp3_tidal")


class OrcTurbTables(ctypes.Structure):
    _fields_ = ([("rib", _DP), ("irimax", _IP)] + [(k, _DP) for k in _TURB_ARRAYS[2:]] +
                [(k, ctypes.c_double) for k in _TURB_SCALARS])


class OrcTurbFields(ctypes.Structure):
    _fields_ = [(k, _DP) for k in _TURB_IN + _TURB_OUT]


def _turb_lib():
    L = lib()
    if not getattr(L, "_turb_bound", False):
        T = ctypes.POINTER(OrcTurbTables)
        L.orc_turb_formld.argtypes = [ctypes.c_int, _DP, _DP]
        L.orc_turb_formld.restype = ctypes.c_double
        L.orc_turb_eplatidepend.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_turb_eplatidepend.restype = ctypes.c_double
        L.orc_turb_interp1d.argtypes = [T, ctypes.c_double, _DP]
        L.orc_turb_interp2d.argtypes = [T, _DP, _DP, _DP]
        L.orc_turb_column.argtypes = [T, ctypes.c_int, ctypes.c_int, _DP, _DP, _DP, _DP, _DP, _DP,
                                      _DP, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double, _DP, _DP, _DP, _DP,
                                      ctypes.POINTER(ctypes.c_long)]
        L.orc_turb_run.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _IP, _DP, _DP, T,
                                   ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                   ctypes.POINTER(OrcTurbFields), ctypes.POINTER(ctypes.c_long)]
        L._turb_bound = True
    return L


class TurbOracle:
    """The oracle's view of one set of Canuto tables."""

    def __init__(self, tables):
        self._keep = []
        T = OrcTurbTables()
        for k in _TURB_ARRAYS:
            a = np.ascontiguousarray(tables[k], np.int32 if k == "irimax" else np.float64)
            self._keep.append(a)
            setattr(T, k, _i(a) if k == "irimax" else _d(a))
        for k in _TURB_SCALARS:
            setattr(T, k, float(tables[k]))
        self.T = T
        self.L = _turb_lib()

    def interp1d(self, x):
        o = np.zeros(4)
        st = self.L.orc_turb_interp1d(ctypes.byref(self.T), float(x), _d(o))
        return st, o

    def interp2d(self, ri, rid):
        a, b, o = ctypes.c_double(ri), ctypes.c_double(rid), np.zeros(4)
        st = self.L.orc_turb_interp2d(ctypes.byref(self.T), ctypes.byref(a), ctypes.byref(b), _d(o))
        return st, o, a.value, b.value

    def column(self, n, nmax, z, t, rh, ri, rid, s2, an2, fricmx, wndmix, buoytur, buoysol, coriol):
        arr = [np.ascontiguousarray(v, np.float64).copy() for v in (z, t, rh, ri, rid, s2, an2)]
        akm, akh, aks = np.zeros(nmax), np.zeros(nmax), np.zeros(nmax)
        amld = ctypes.c_double()
        cnt = (ctypes.c_long * len(TURB_COUNTERS))()
        ex = self.L.orc_turb_column(ctypes.byref(self.T), n, nmax, *[_d(a) for a in arr], fricmx,
                                    wndmix, buoytur, buoysol, coriol, ctypes.byref(amld), _d(akm),
                                    _d(akh), _d(aks), cnt)
        return dict(exit=ex, amld=amld.value, akm=akm, akh=akh, aks=aks, s2=arr[5],
                    counts=dict(zip(TURB_COUNTERS, list(cnt))))

    def run(self, tc, outputs=None):
        """turb_cu over a TurbCase; returns (outputs, branch counts).  Pass
        ``outputs`` to start from given output arrays (akt, richardson,
        fztidal and wp3_tidal keep what the call does not write)."""
        km, jmt, imt = tc.km, tc.jmt, tc.imt
        if outputs is None:
            outputs = dict(akmt=np.zeros((km, jmt, imt)), akt=np.zeros((2, km, jmt, imt)),
                           ak_tide=np.zeros((km, jmt, imt)), amld=np.zeros((jmt, imt)),
                           richardson=np.zeros((km, jmt, imt)), fztidal=np.zeros((km, jmt, imt)),
                           wp3_tidal=np.zeros((km, jmt, imt)))
        src = dict(tc.fields)
        src["fcort"] = tc.grid["fcort"]
        keep = {k: np.ascontiguousarray(src[k], np.float64) for k in _TURB_IN}
        F = OrcTurbFields(*([_d(keep[k]) for k in _TURB_IN] + [_d(outputs[k]) for k in _TURB_OUT]))
        kmt = np.ascontiguousarray(tc.grid["kmt"], np.int32)
        zkp = np.ascontiguousarray(tc.grid["zkp"], np.float64)
        dzp = np.ascontiguousarray(tc.grid["dzp"], np.float64)
        cnt = (ctypes.c_long * len(TURB_COUNTERS))()
        self.L.orc_turb_run(imt, jmt, km, _i(kmt), _d(zkp), _d(dzp), ctypes.byref(self.T),
                            int(tc.params["ncc"]), float(tc.params["dfricmx"]),
                            float(tc.params["dwndmix"]), ctypes.byref(F), cnt)
        return outputs, dict(zip(TURB_COUNTERS, list(cnt)))


def turb_formld(z, t):
    z = np.ascontiguousarray(z, np.float64)
    t = np.ascontiguousarray(t, np.float64)
    return _turb_lib().orc_turb_formld(len(z), _d(z), _d(t))


def turb_eplatidepend(f, an):
    return _turb_lib().orc_turb_eplatidepend(float(f), float(an))


# ------------------------------------------------------------------ readyt_
READYT_GRID_2D = ("au0", "aus", "auw", "ausw", "dxur", "dyur", "ohbu", "ohbt", "hbx", "hby")
READYT_GRID_1D = ("dzp", "odzt", "zkt", "to", "so", "po")
READYT_FIELDS = ("vit", "viv", "at", "atb", "h0", "u", "v", "psa", "swv", "nswv",
                 "h0l", "h0f", "utl", "vtl", "utf", "vtf",
                 "dlu", "dlv", "pdensity", "gg", "pp", "ppa", "ppb", "ppc", "alpha", "beta",
                 "ric", "rict", "ricdttms", "ricdt", "rit", "akt", "buoytur", "buoysol", "pxb", "pyb",
                 "wgp", "work", "whx", "why", "work1", "work2", "pax", "pay", "ax", "ay", "az")


class OrcReadytGrid(ctypes.Structure):
    _fields_ = ([("imt", ctypes.c_int), ("jmt", ctypes.c_int), ("km", ctypes.c_int), ("kmu", _IP)] +
                [(k, _DP) for k in READYT_GRID_2D + READYT_GRID_1D] +
                [("eosc", _DP), ("od0", ctypes.c_double), ("od0cp", ctypes.c_double)])


class OrcReadytFields(ctypes.Structure):
    _fields_ = [(k, _DP) for k in READYT_FIELDS]


def _readyt_lib():
    L = lib()
    if not getattr(L, "_readyt_bound", False):
        L.orc_readyt.argtypes = [ctypes.POINTER(OrcReadytGrid), ctypes.POINTER(OrcReadytFields),
                                 ctypes.c_int]
        L.orc_thermal.argtypes = [ctypes.c_double] * 5 + [_DP, _DP]
        L._readyt_bound = True
    return L


def readyt(imt, jmt, km, grid, fields, ist=1, threads=None):
    """readyt_ (readyt.cpp:511-598) on numpy arrays: returns a dict with a
    copy of every field after the call (inputs unchanged, in/out and output
    arrays as the reference leaves them; ax/ay/az only if given)."""
    L = _readyt_lib()
    need = [k for k in READYT_FIELDS[:-3] if fields.get(k) is None]
    if need:
        raise ValueError(f"readyt: missing fields {need}")
    keep = []
    G = OrcReadytGrid()
    G.imt, G.jmt, G.km = int(imt), int(jmt), int(km)
    kmu = np.ascontiguousarray(grid["kmu"], np.int32)
    keep.append(kmu)
    G.kmu = _i(kmu)
    for k in READYT_GRID_2D + READYT_GRID_1D + ("eosc",):
        a = np.ascontiguousarray(grid[k], np.float64)
        keep.append(a)
        setattr(G, k, _d(a))
    G.od0, G.od0cp = float(grid["od0"]), float(grid["od0cp"])
    out = {k: (None if fields.get(k) is None else np.array(fields[k], np.float64, order="C"))
           for k in READYT_FIELDS}
    F = OrcReadytFields(*[_d(out[k]) for k in READYT_FIELDS])
    old = None
    if threads is not None:
        old = L.orc_nthreads()
        L.orc_set_nthreads(int(threads))
    try:
        L.orc_readyt(ctypes.byref(G), ctypes.byref(F), int(ist))
    finally:
        if old is not None:
            L.orc_set_nthreads(old)
    return {k: v for k, v in out.items() if v is not None}


def thermal(tt, ss, ppa, vit, od0):
    """thermal_cu's alpha, beta at one cell (readyt.cpp:272-294)."""
    readyt_lib = _readyt_lib()
    a, b = ctypes.c_double(), ctypes.c_double()
    readyt_lib.orc_thermal(float(tt), float(ss), float(ppa), float(vit), float(od0),
                           ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value



# ------------------------------------------------------- readyc_ (mixing part)
READYC_GRID = ("at0", "atn", "ate", "atne", "au0", "aus", "auw", "ausw")
READYC_FIELDS = ("vit", "viv", "up", "vp", "h0", "rict", "ric", "h0bl", "h0bf",
                 "wp12", "wp13", "riv1", "riv2", "s2t", "rit", "s2u", "riu")


class OrcReadycGrid(ctypes.Structure):
    _fields_ = ([("imt", ctypes.c_int), ("jmt", ctypes.c_int), ("km", ctypes.c_int)] +
                [(k, _DP) for k in READYC_GRID] + [("odzt", _DP)])


class OrcReadycFields(ctypes.Structure):
    _fields_ = [(k, _DP) for k in READYC_FIELDS]


def _readyc_lib():
    L = lib()
    if not getattr(L, "_readyc_bound", False):
        L.orc_readyc_shear.argtypes = [ctypes.POINTER(OrcReadycGrid), ctypes.POINTER(OrcReadycFields)]
        L.orc_readyc_akmu.argtypes = [ctypes.POINTER(OrcReadycGrid), _DP, _DP, _DP]
        L._readyc_bound = True
    return L


def _readyc_grid(imt, jmt, km, grid, keep):
    G = OrcReadycGrid()
    G.imt, G.jmt, G.km = int(imt), int(jmt), int(km)
    for k in READYC_GRID + ("odzt",):
        a = np.ascontiguousarray(grid[k], np.float64)
        keep.append(a)
        setattr(G, k, _d(a))
    return G


def readyc_shear(imt, jmt, km, grid, fields):
    """readyc1 + readyc1_2 with the clears before them (readyc.cpp:44-170,
    1636-1643) on numpy arrays: a dict with a copy of every field after the
    call."""
    L = _readyc_lib()
    keep = []
    G = _readyc_grid(imt, jmt, km, grid, keep)
    missing = [k for k in READYC_FIELDS if fields.get(k) is None]
    if missing:
        raise ValueError(f"readyc_shear: missing fields {missing}")
    out = {k: np.array(fields[k], np.float64, order="C") for k in READYC_FIELDS}
    L.orc_readyc_shear(ctypes.byref(G), ctypes.byref(OrcReadycFields(*[_d(out[k]) for k in READYC_FIELDS])))
    return out


def readyc_akmu(imt, jmt, km, grid, akmt, viv):
    """readyc2 with the clear of akmu (readyc.cpp:223-246, 1679)."""
    L = _readyc_lib()
    keep = []
    G = _readyc_grid(imt, jmt, km, grid, keep)
    a = np.ascontiguousarray(akmt, np.float64)
    v = np.ascontiguousarray(viv, np.float64)
    out = np.full(a.shape, np.nan)
    L.orc_readyc_akmu(ctypes.byref(G), _d(a), _d(v), _d(out))
    return out


# ------------------------------------------------- readyc_ (momentum part)
READYC_MGRID_SCALARS = ("am", "c0f", "od0", "cag", "sag")
READYC_MGRID_2D = ("ohbu", "ohbt", "hts", "htw", "hue", "hun", "tarea_r", "uarea_r", "dxu", "dyu",
                   "dxur", "dyur", "snlat", "amf", "duc", "dum", "dun", "dus", "due", "duw", "dmc",
                   "dmn", "dms", "dme", "dmw")
READYC_MFIELDS_IN = ("u", "v", "up", "vp", "h0", "su", "sv", "vit", "viv", "akmu")
READYC_MFIELDS_OUT = ("ws", "dlu", "dlv", "wka", "dlub", "dlvb", "sbcx", "sbcy", "bbcx", "bbcy")


class OrcReadycMgrid(ctypes.Structure):
    _fields_ = ([(k, ctypes.c_int) for k in ("ib", "ie", "jb", "je", "del4")] +
                [(k, ctypes.c_double) for k in READYC_MGRID_SCALARS] +
                [("kmt", _IP), ("kmu", _IP)] + [(k, _DP) for k in READYC_MGRID_2D] +
                [("dzp", _DP), ("odzp", _DP)])


class OrcReadycMfields(ctypes.Structure):
    _fields_ = [(k, _DP) for k in READYC_MFIELDS_IN + READYC_MFIELDS_OUT]


READYC_MOMENTUM_PIECES = ("upwell", "advect", "hdiff", "stress", "vdiff")


def _readyc_mlib():
    L = _readyc_lib()
    if not getattr(L, "_readyc_mbound", False):
        for name in ["orc_readyc_momentum"] + ["orc_rm_" + p for p in READYC_MOMENTUM_PIECES]:
            getattr(L, name).argtypes = [ctypes.POINTER(OrcReadycGrid), ctypes.POINTER(OrcReadycMgrid),
                                         ctypes.POINTER(OrcReadycMfields)]
        L._readyc_mbound = True
    return L


def readyc_momentum(imt, jmt, km, grid, mgrid, fields, pieces=None):
    """The momentum tendencies of readyc_ after readyc2 (readyc.cpp:1700-1793)
    on numpy arrays: ``grid`` as for readyc_shear, ``mgrid`` the momentum
    grid (block bounds ib, ie, jb, je 1-based, del4, the scalars, kmt, kmu,
    the 2-D metrics, dzp, odzp), ``fields`` the inputs and the initial
    outputs (ws is read at level 0).  ``pieces`` runs only the named
    reference kernels in launch order (upwell, advect, hdiff, stress, vdiff)
    on the fields given.  Returns a dict with a copy of every output."""
    L = _readyc_mlib()
    keep = []
    G = _readyc_grid(imt, jmt, km, grid, keep)
    M = OrcReadycMgrid()
    for k in ("ib", "ie", "jb", "je", "del4"):
        setattr(M, k, int(mgrid[k]))
    for k in READYC_MGRID_SCALARS:
        setattr(M, k, float(mgrid[k]))
    for k in ("kmt", "kmu"):
        a = np.ascontiguousarray(mgrid[k], np.int32)
        keep.append(a)
        setattr(M, k, _i(a))
    for k in READYC_MGRID_2D + ("dzp", "odzp"):
        a = np.ascontiguousarray(mgrid[k], np.float64)
        keep.append(a)
        setattr(M, k, _d(a))
    n2, n3 = (jmt, imt), (km, jmt, imt)
    shapes = {k: n2 for k in ("h0", "su", "sv", "dlub", "dlvb", "sbcx", "sbcy", "bbcx", "bbcy")}
    ins = {}
    for k in READYC_MFIELDS_IN:
        a = np.ascontiguousarray(fields[k], np.float64)
        if a.shape != shapes.get(k, n3):
            raise ValueError(f"readyc_momentum: {k} has shape {a.shape}")
        keep.append(a)
        ins[k] = a
    out = {}
    for k in READYC_MFIELDS_OUT:
        v = fields.get(k)
        out[k] = (np.zeros(shapes.get(k, n3)) if v is None else np.array(v, np.float64, order="C"))
    F = OrcReadycMfields(*([_d(ins[k]) for k in READYC_MFIELDS_IN] + [_d(out[k]) for k in READYC_MFIELDS_OUT]))
    if pieces is None:
        L.orc_readyc_momentum(ctypes.byref(G), ctypes.byref(M), ctypes.byref(F))
    else:
        for p in pieces:
            getattr(L, "orc_rm_" + p)(ctypes.byref(G), ctypes.byref(M), ctypes.byref(F))
    return out


# ------------------------------------------------------------------ bclinc_
BCLINC_SCALARS = ("od0", "c0f", "cag", "sag", "onbb", "dtc", "dtc2", "afc1", "afc2")
BCLINC_GRID_2D = ("snlat", "ohbt", "ohbu", "fcor", "dxur", "dyur", "epea", "epeb", "epla", "eplb")
BCLINC_GRID_1D = ("dzp", "odzt", "odzp", "zkt")
BCLINC_IN = ("vit", "viv", "gg", "akmu", "su", "sv", "psa", "h0bl", "ub", "vb")
BCLINC_INOUT = ("h0bf", "u", "v", "up", "vp", "dlu", "dlv", "wka", "utf", "vtf", "sbcx", "sbcy", "bbcx", "bbcy")
BCLINC_2D = ("su", "sv", "psa", "h0bl", "ub", "vb", "h0bf", "sbcx", "sbcy", "bbcx", "bbcy")


class OrcBclincGrid(ctypes.Structure):
    _fields_ = ([(k, ctypes.c_int) for k in ("imt", "jmt", "km")] +
                [(k, ctypes.c_double) for k in BCLINC_SCALARS] + [("kmu", _IP)] +
                [(k, _DP) for k in BCLINC_GRID_2D + BCLINC_GRID_1D])


class OrcBclincFields(ctypes.Structure):
    _fields_ = [(k, _DP) for k in BCLINC_IN + BCLINC_INOUT]


def bclinc_stage(imt, jmt, km, grid, fields, isc, stages=(0, 1, 2)):
    """bclinc_ (bclinc.cpp:728-827) on numpy arrays, the given stages in
    order with no halo exchange between them (stage 0: bclinc12, bclinc11,
    then bclinc2 + bclinc3 or, at isc < 1, bclinc6 + bclinc71 + bclinc72;
    stage 1: bclinc4 or bclinc81 + bclinc82; stage 2: bclinc5).  Returns a
    dict with a copy of every in/out field after the last stage."""
    L = lib()
    if not getattr(L, "_bclinc_bound", False):
        L.orc_bclinc_stage.argtypes = [ctypes.POINTER(OrcBclincGrid), ctypes.POINTER(OrcBclincFields),
                                       ctypes.c_int, ctypes.c_int]
        L._bclinc_bound = True
    keep = []
    G = OrcBclincGrid()
    G.imt, G.jmt, G.km = int(imt), int(jmt), int(km)
    for k in BCLINC_SCALARS:
        setattr(G, k, float(grid[k]))
    kmu = np.ascontiguousarray(grid["kmu"], np.int32)
    keep.append(kmu)
    G.kmu = _i(kmu)
    for k in BCLINC_GRID_2D + BCLINC_GRID_1D:
        a = np.ascontiguousarray(grid[k], np.float64)
        keep.append(a)
        setattr(G, k, _d(a))
    s2, s3 = (jmt, imt), (km, jmt, imt)
    ins = {}
    for k in BCLINC_IN:
        a = np.ascontiguousarray(fields[k], np.float64)
        if a.shape != (s2 if k in BCLINC_2D else s3):
            raise ValueError(f"bclinc_stage: {k} has shape {a.shape}")
        keep.append(a)
        ins[k] = a
    out = {k: np.array(fields[k], np.float64, order="C") for k in BCLINC_INOUT}
    F = OrcBclincFields(*([_d(ins[k]) for k in BCLINC_IN] + [_d(out[k]) for k in BCLINC_INOUT]))
    for s in stages:
        L.orc_bclinc_stage(ctypes.byref(G), ctypes.byref(F), int(isc), int(s))
    return out


# ------------------------------------------------------------------ barotr_
BAROTR_INTS = ("ib", "ie", "jb", "je", "del4")
BAROTR_SCALARS = ("dtb", "am", "ah")
BAROTR_IGRID = ("kmt", "kmu", "kmtn", "kmts", "kmte", "kmtw")
BAROTR_DGRID = ("au0", "aus", "auw", "ausw", "htw", "hts", "tarea_r", "uarea_r", "dxu", "dyu", "dxur",
                "dyur", "fcor", "ebea", "ebeb", "amf", "ahf", "dtn", "dts", "dte", "dtw", "duc", "dum",
                "dun", "dus", "due", "duw", "dmc", "dmn", "dms", "dme", "dmw")
BAROTR_IN = ("vit", "viv", "dzph", "dlub", "dlvb", "wgp", "pax", "pay", "pxb", "pyb", "whx", "why")
BAROTR_INOUT = ("h0", "h0p", "ub", "vb", "ubp", "vbp", "h0f", "h0bf", "work", "wka")


class OrcBarotrGrid(ctypes.Structure):
    _fields_ = ([(k, ctypes.c_int) for k in ("imt", "jmt") + BAROTR_INTS] +
                [(k, ctypes.c_double) for k in BAROTR_SCALARS] + [(k, _IP) for k in BAROTR_IGRID] +
                [(k, _DP) for k in BAROTR_DGRID])


class OrcBarotrFields(ctypes.Structure):
    _fields_ = [(k, _DP) for k in BAROTR_IN + BAROTR_INOUT]


def barotr_stage(imt, jmt, grid, fields, nc, stages=(0, 1, 2, 3)):
    """One subcycle nc of barotr_ (barotr.cpp:901-1024) on numpy arrays, the
    given stages in order with no halo exchange between them.  Returns a dict
    with a copy of every in/out field (wka: 6 planes)."""
    L = lib()
    if not getattr(L, "_barotr_bound", False):
        L.orc_barotr_stage.argtypes = [ctypes.POINTER(OrcBarotrGrid), ctypes.POINTER(OrcBarotrFields),
                                       ctypes.c_int, ctypes.c_int]
        L._barotr_bound = True
    keep = []
    G = OrcBarotrGrid()
    G.imt, G.jmt = int(imt), int(jmt)
    for k in BAROTR_INTS:
        setattr(G, k, int(grid[k]))
    for k in BAROTR_SCALARS:
        setattr(G, k, float(grid[k]))
    for k in BAROTR_IGRID:
        a = np.ascontiguousarray(grid[k], np.int32)
        keep.append(a)
        setattr(G, k, _i(a))
    for k in BAROTR_DGRID:
        a = np.ascontiguousarray(grid[k], np.float64)
        keep.append(a)
        setattr(G, k, _d(a))
    ins = [np.ascontiguousarray(fields[k], np.float64) for k in BAROTR_IN]
    keep += ins
    out = {k: np.array(fields[k], np.float64, order="C") for k in BAROTR_INOUT}
    if out["wka"].shape != (6, jmt, imt):
        raise ValueError(f"barotr_stage: wka has shape {out['wka'].shape}")
    F = OrcBarotrFields(*([_d(a) for a in ins] + [_d(out[k]) for k in BAROTR_INOUT]))
    for s in stages:
        L.orc_barotr_stage(ctypes.byref(G), ctypes.byref(F), int(nc), int(s))
    return out
