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


# ---------------------------------------------------------------- Canuto mixing
TURB_MT, TURB_MT0, TURB_NTH = 1462, 400, 75   # canuto_mod.h: mt, mt0, n_theta_r_oct


@dataclasses.dataclass
class TurbCase:
    """Inputs of the Canuto producer (turbinit tables + one turb_cu call)."""
    imt: int
    jmt: int
    km: int
    grid: dict      # kmt [jmt][imt], zkp km+1, dzp km, fcort [jmt][imt]
    tables: dict    # canuto_mod arrays and turb_param scalars (licom_turb.TABLE_*)
    fields: dict    # turb_cu inputs ([k][j][i] / [j][i])
    params: dict    # ncc, dfricmx, dwndmix


def turb_tables(b1=16.6, theta_rcrp=1.1, theta_rcrn=1.7):
    """Synthetic Canuto tables on the reference's grids.

    The independent variables follow turb_ini: the Ri (= Ri_d) and
    N_d^2/N^2 tables are linear with step dri = -ri0/mt0 = 0.01 up to index
    mt0 and grow geometrically by rri = rib(mt0)/rib(mt0-1) beyond it
    (canuto_ini_2002.F90:415-471, 509-533); the background tables sit on
    theta = itheta * (pi/4)/n_theta_r_oct, itheta = -n..3n (:700).  The
    table values are smooth positive functions, not the Canuto 2002
    closure: the realizable region ends at a critical Ri_c(Ri_d) = 0.6 +
    0.35 tanh(Ri_d) (irimax), beyond which slq2b is negative, as turb_ini
    leaves it."""
    mt, mt0, nth = TURB_MT, TURB_MT0, TURB_NTH
    dri = 4.0 / mt0
    idx = np.arange(-mt, mt + 1)
    rib = idx * dri
    rri = rib[mt + mt0] / rib[mt + mt0 - 1]
    far = np.abs(idx) > mt0
    rib[far] = np.sign(idx[far]) * rib[mt + mt0] * rri ** (np.abs(idx[far]) - mt0).astype(np.float64)
    ri = rib[None, :]
    rid = rib[:, None]
    ri_c = 0.6 + 0.35 * np.tanh(rib)
    irimax = (np.searchsorted(rib, ri_c, side="right") - 1 - mt).astype(np.int32)
    lim = rib[mt + irimax][:, None]
    slq2b = np.where(ri <= lim, 3.0 * (ri_c[:, None] - ri + 0.02) / (1.0 + 0.1 * np.abs(ri)), -1.0)
    smb = 0.08 + 0.3 / (1.0 + np.exp(np.minimum(2.0 * ri, 700.0))) + 0.0 * rid
    shb = smb * (1.3 + 0.2 * np.tanh(rid))
    ssb = shb * (0.9 + 0.1 * np.tanh(0.5 * rid))
    x = rib.copy()
    th = np.arange(-nth, 3 * nth + 1) * ((np.pi / 4.0) / nth)
    return dict(
        rib=rib, irimax=irimax,
        slq2b=np.ascontiguousarray(slq2b), smb=np.ascontiguousarray(smb),
        shb=np.ascontiguousarray(shb), ssb=np.ascontiguousarray(ssb),
        and2on2a1=x, amtaun2a1=1.5 + 0.5 * np.tanh(x), sma1=0.1 + 0.05 * np.tanh(x),
        sha1=0.15 + 0.05 * np.tanh(0.5 * x), ssa1=0.12 + 0.04 * np.tanh(x),
        back_ra_r=0.12 + 0.05 * np.cos(2.0 * th), sm_r1=0.05 + 0.02 * np.sin(th),
        sh_r1=0.07 + 0.03 * np.cos(th), ss_r1=0.06 + 0.02 * np.cos(2.0 * th),
        slq2_r1=2.0 + 0.5 * np.sin(2.0 * th),
        rri=rri, rnd2on2=rri, dri=dri, deltheta_r=(np.pi / 4.0) / nth, b1=b1,
        theta_rcrp=theta_rcrp, theta_rcrn=theta_rcrn, dand2on2=dri)


def make_turb_case(case: Case, seed: int = 7, tables: dict | None = None) -> TurbCase:
    """turb_cu inputs over a tracer case's grid (kmt, zkp, latitudes).

    Profiles: a mixed layer of 10-120 m over a thermocline, density rising
    with depth, N^2 mostly stable with unstable near-surface levels in a
    third of the columns, S^2 log-uniform, and a few columns of strong
    convection with almost no shear (the zero-shear table), an unstable
    level below stable ones in a tenth of the columns; the equator row has
    f = 0 exactly."""
    rng = np.random.default_rng(seed)
    km, jmt, imt = case.km, case.jmt, case.imt
    kmt = case.grid["kmt"].astype(np.int32)
    zkp = case.grid["zkp"].astype(np.float64)
    dzp = zkp[:-1] - zkp[1:]
    depth = -0.5 * (zkp[:-1] + zkp[1:])                       # T-level depths (m)
    lev = np.arange(km)[:, None, None]
    vit = (lev < kmt[None]).astype(np.float64)
    lat = np.deg2rad(case.grid["lat_t"])
    fcort = np.repeat((2.0 * 7.292e-5 * np.sin(lat))[:, None], imt, axis=1)
    fcort[jmt // 2, :] = 0.0
    s3 = (km, jmt, imt)
    d3 = depth[:, None, None]
    mld = rng.uniform(10.0, 120.0, (jmt, imt))
    ts = 28.0 - 20.0 * np.abs(np.sin(lat))[:, None] + rng.normal(0.0, 0.5, (jmt, imt))
    at = ts[None] - 8.0 * (1.0 - np.exp(-np.maximum(d3 - mld[None], 0.0) / 400.0))
    at = at + rng.normal(0.0, 0.01, s3)
    pden = 1022.0 + 5.0 * (1.0 - np.exp(-d3 / 800.0)) + rng.normal(0.0, 0.002, s3)
    s2t = 10.0 ** rng.uniform(-8.0, -4.0, s3)
    rict = 10.0 ** rng.uniform(-7.0, -4.0, s3)
    conv = (rng.random((jmt, imt)) < 0.33)[None] & (lev < 3)
    rict = np.where(conv, -(10.0 ** rng.uniform(-7.0, -5.0, s3)), rict)
    # an unstable level below stable ones in a tenth of the columns
    inv = (rng.random((jmt, imt)) < 0.1)[None] & (lev == rng.integers(3, max(km - 1, 4), (jmt, imt))[None])
    rict = np.where(inv, -(10.0 ** rng.uniform(-7.0, -5.0, s3)), rict)
    calm = (rng.random((jmt, imt)) < 0.03)[None] & (lev < 4)
    s2t = np.where(calm, 1.0e-10, s2t)
    rict = np.where(calm, -1.0e-6, rict)
    rit = rict / (s2t + 1.0e-25)
    ricdttms = rit * rng.uniform(-1.5, 1.5, s3)
    bottom = np.take(np.concatenate([[0.0], -zkp[1:]]), kmt)  # column depth (m)
    fz = np.exp(-np.maximum(bottom[None] - d3, 0.0) / 500.0) * vit
    fields = dict(at=at, pdensity=pden, rit=rit, ricdttms=ricdttms, s2t=s2t, rict=rict,
                  fz_tide=fz, vit=vit,
                  buoytur=rng.normal(0.0, 1.0e-8, (jmt, imt)),
                  buoysol=-np.abs(rng.normal(0.0, 5.0e-9, (jmt, imt))),
                  wave_dis=10.0 ** rng.uniform(-4.0, -2.0, (jmt, imt)))
    fields = {k: np.ascontiguousarray(v, np.float64) for k, v in fields.items()}
    params = dict(ncc=int(round(1.0 / case.params["oncc"])), dfricmx=100.0e-4, dwndmix=10.0e-4)
    return TurbCase(imt=imt, jmt=jmt, km=km,
                    grid=dict(kmt=kmt, zkp=zkp, dzp=dzp, fcort=fcort),
                    tables=tables if tables is not None else turb_tables(), fields=fields,
                    params=params)


def make_turb_case_device(case: Case, torch, device, seed: int = 7,
                          tables: dict | None = None) -> TurbCase:
    """make_turb_case's profiles drawn on the device with torch's generator
    (for grids too large to draw on the host); the fields are float64
    device tensors, the grid and tables stay numpy."""
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    km, jmt, imt = case.km, case.jmt, case.imt
    f64 = dict(dtype=torch.float64, device=device)
    kmt_np = case.grid["kmt"].astype(np.int32)
    zkp = case.grid["zkp"].astype(np.float64)
    dzp = zkp[:-1] - zkp[1:]
    kmt = torch.from_numpy(kmt_np).to(device)
    depth = torch.from_numpy(-0.5 * (zkp[:-1] + zkp[1:])).to(device)[:, None, None]
    lev = torch.arange(km, device=device)[:, None, None]
    lat = np.deg2rad(case.grid["lat_t"])
    fcort = np.repeat((2.0 * 7.292e-5 * np.sin(lat))[:, None], imt, axis=1)
    fcort[jmt // 2, :] = 0.0
    sinlat = torch.from_numpy(np.abs(np.sin(lat))).to(device)[:, None]

    def u(lo, hi, shape):
        return torch.rand(shape, generator=g, **f64) * (hi - lo) + lo

    def nrm(sd, shape):
        return torch.randn(shape, generator=g, **f64) * sd

    s3, s2 = (km, jmt, imt), (jmt, imt)
    fields = {}
    fields["vit"] = (lev < kmt[None]).to(torch.float64)
    mld = u(10.0, 120.0, s2)
    ts = 28.0 - 20.0 * sinlat + nrm(0.5, s2)
    fields["at"] = (ts[None] - 8.0 * (1.0 - torch.exp(-torch.clamp(depth - mld[None], min=0.0) / 400.0))
                    + nrm(0.01, s3))
    del mld, ts
    fields["pdensity"] = 1022.0 + 5.0 * (1.0 - torch.exp(-depth / 800.0)) + nrm(0.002, s3)
    s2t = 10.0 ** u(-8.0, -4.0, s3)
    rict = 10.0 ** u(-7.0, -4.0, s3)
    conv = (torch.rand(s2, generator=g, device=device) < 0.33)[None] & (lev < 3)
    rict = torch.where(conv, -(10.0 ** u(-7.0, -5.0, s3)), rict)
    del conv
    lk = torch.randint(3, max(km - 1, 4), s2, generator=g, device=device)
    inv = (torch.rand(s2, generator=g, device=device) < 0.1)[None] & (lev == lk[None])
    rict = torch.where(inv, -(10.0 ** u(-7.0, -5.0, s3)), rict)
    del inv, lk
    calm = (torch.rand(s2, generator=g, device=device) < 0.03)[None] & (lev < 4)
    s2t = torch.where(calm, torch.full_like(s2t, 1.0e-10), s2t)
    rict = torch.where(calm, torch.full_like(rict, -1.0e-6), rict)
    del calm
    fields["rit"] = rict / (s2t + 1.0e-25)
    fields["ricdttms"] = fields["rit"] * u(-1.5, 1.5, s3)
    fields["s2t"], fields["rict"] = s2t, rict
    bottom = torch.from_numpy(np.take(np.concatenate([[0.0], -zkp[1:]]), kmt_np)).to(device)
    fields["fz_tide"] = torch.exp(-torch.clamp(bottom[None] - depth, min=0.0) / 500.0) * fields["vit"]
    fields["buoytur"] = nrm(1.0e-8, s2)
    fields["buoysol"] = -nrm(5.0e-9, s2).abs()
    fields["wave_dis"] = 10.0 ** u(-4.0, -2.0, s2)
    fields = {k: v.contiguous() for k, v in fields.items()}
    params = dict(ncc=int(round(1.0 / case.params["oncc"])), dfricmx=100.0e-4, dwndmix=10.0e-4)
    return TurbCase(imt=imt, jmt=jmt, km=km, grid=dict(kmt=kmt_np, zkp=zkp, dzp=dzp, fcort=fcort),
                    tables=tables if tables is not None else turb_tables(), fields=fields,
                    params=params)


# ---------------------------------------------------------------- readyt_
def readyt_grid(case: Case, seed: int = 11) -> dict:
    """licom_readyt_grid of a tracer case's block: kmu (the minimum kmt of
    the four T cells around a U point), T->U weights, 1/dxu, 1/dyu, the
    inverse column depths at U and T points, a smooth synthetic hbx/hby, dzp,
    odzt and the T-level depths from zkp, the EOS table (to, so, C) and a
    synthetic density offset po rising 4.5 kg/m^3 per 1000 m."""
    rng = np.random.default_rng(seed)
    km, jmt, imt = case.km, case.jmt, case.imt
    g = case.grid
    kmt = g["kmt"].astype(np.int32)
    kmu = np.zeros_like(kmt)
    kmu[:-1, 1:] = np.minimum(np.minimum(kmt[:-1, :-1], kmt[:-1, 1:]),
                              np.minimum(kmt[1:, :-1], kmt[1:, 1:]))
    zkp = g["zkp"].astype(np.float64)
    dzp = zkp[:-1] - zkp[1:]
    zkt = 0.5 * (zkp[:-1] + zkp[1:])
    odzt = np.empty(km)
    odzt[0] = 2.0 / dzp[0]
    odzt[1:] = 1.0 / (zkt[:-1] - zkt[1:])
    cum = np.concatenate([[0.0], np.cumsum(dzp)])
    hu, ht = cum[kmu], cum[kmt]
    inv = lambda h: np.where(h > 0.0, 1.0 / np.where(h > 0.0, h, 1.0), 0.0)  # noqa: E731
    recip = lambda a: np.where(a > 0.0, 1.0 / np.where(a > 0.0, a, 1.0), 0.0)  # noqa: E731

    def smooth(amp):
        a = rng.normal(0.0, 1.0, (jmt, imt))
        for _ in range(3):
            a = 0.25 * (np.roll(a, 1, 1) + np.roll(a, -1, 1) + np.roll(a, 1, 0) + np.roll(a, -1, 0))
        return amp * a

    return dict(kmu=kmu, au0=g["au0"].copy(), aus=g["aus"].copy(), auw=g["auw"].copy(),
                ausw=g["ausw"].copy(), dxur=recip(g["dxu"]), dyur=recip(g["dyu"]),
                ohbu=inv(hu), ohbt=inv(ht), hbx=smooth(1.0e-3), hby=smooth(1.0e-3),
                dzp=dzp, odzt=odzt, zkt=zkt, to=g["to"].copy(), so=g["so"].copy(),
                po=4.5e-3 * (-zkt), eosc=g["eosc"].copy(),
                od0=float(case.params["od0"]), od0cp=float(case.params["od0cp"]))


def readyt_fields(case: Case, grid: dict, seed: int = 12, sentinel: float = 7.0) -> dict:
    """licom_readyt_fields over a tracer case: the wet masks from kmt / kmu,
    the case's at / atb and time levels, new velocities and surface height
    (the current ones perturbed by 1%), a smooth surface pressure, surface
    heat fluxes; outputs start at `sentinel` so that a test sees what a call
    leaves alone."""
    rng = np.random.default_rng(seed)
    km, jmt, imt = case.km, case.jmt, case.imt
    lev = np.arange(km)[:, None, None]
    vit = (lev < case.grid["kmt"][None]).astype(np.float64)
    viv = (lev < grid["kmu"][None]).astype(np.float64)
    st = case.state
    s2 = (jmt, imt)
    f = dict(vit=vit, viv=viv, at=st["at"].copy(), atb=st["atb"].copy(),
             h0=rng.normal(0.0, 0.3, s2),
             u=st["utf"] * (1.0 + rng.normal(0.0, 0.01, st["utf"].shape)),
             v=st["vtf"] * (1.0 + rng.normal(0.0, 0.01, st["vtf"].shape)),
             psa=1.0e3 * np.cos(np.linspace(0.0, 6.0, imt))[None, :] * np.sin(np.linspace(0.0, 3.0, jmt))[:, None]
             + rng.normal(0.0, 5.0, s2),
             swv=rng.uniform(0.0, 3.0e-5, s2), nswv=rng.normal(-2.0e-5, 1.0e-5, s2),
             h0l=st["h0l"].copy(), h0f=st["h0f"].copy() + rng.normal(0.0, 0.1, s2),
             utl=st["utl"].copy(), vtl=st["vtl"].copy(), utf=st["utf"].copy(), vtf=st["vtf"].copy())
    shapes = dict(dlu=(km, jmt, imt), dlv=(km, jmt, imt), pdensity=(km, jmt, imt), gg=(km, jmt, imt),
                  pp=(km, jmt, imt), ppa=(km, jmt, imt), ppb=(km, jmt, imt), ppc=(km, jmt, imt),
                  alpha=(km, jmt, imt), beta=(km, jmt, imt), ric=(km - 1, jmt, imt),
                  rict=(km - 1, jmt, imt), ricdttms=(km - 1, jmt, imt), ricdt=(km - 1, jmt, imt),
                  rit=(km - 1, jmt, imt), akt=(2, km, jmt, imt))
    for k in ("buoytur", "buoysol", "pxb", "pyb", "wgp", "work", "whx", "why", "work1", "work2",
              "pax", "pay"):
        shapes[k] = s2
    for k, sh in shapes.items():
        f[k] = np.full(sh, sentinel)
    return {k: np.ascontiguousarray(v, np.float64) for k, v in f.items()}


# ------------------------------------------------- readyc_ (mixing part)
def readyc_grid(case: Case, rgrid: dict, seed: int = 13) -> dict:
    """licom_readyc_grid of a tracer case's block: U->T weights near 1/4
    (perturbed, so that the weighting is exercised), the T->U weights and
    odzt of its readyt_ grid."""
    rng = np.random.default_rng(seed)
    s2 = (case.jmt, case.imt)
    w = lambda: 0.25 * (1.0 + rng.uniform(-0.2, 0.2, s2))  # noqa: E731
    return dict(at0=w(), atn=w(), ate=w(), atne=w(), au0=rgrid["au0"], aus=rgrid["aus"],
                auw=rgrid["auw"], ausw=rgrid["ausw"], odzt=rgrid["odzt"])


def readyc_inputs(case: Case, rgrid: dict, rfields: dict, seed: int = 13,
                  sentinel: float = 7.0) -> tuple:
    """licom_readyc_grid (readyc_grid) and licom_readyc_fields over a tracer
    case and its readyt_ block: velocities with vertical shear, rict / ric
    and the masks of the readyt_ fields, outputs starting at `sentinel`."""
    grid = readyc_grid(case, rgrid, seed)
    rng = np.random.default_rng(seed + 1)
    km, jmt, imt = case.km, case.jmt, case.imt
    s2 = (jmt, imt)
    st = case.state
    shear = 1.0 + 0.3 * rng.normal(0.0, 1.0, (km, 1, 1))
    up = st["utf"] * shear + rng.normal(0.0, 0.01, st["utf"].shape) * rfields["viv"]
    vp = st["vtf"] * shear[::-1] + rng.normal(0.0, 0.01, st["vtf"].shape) * rfields["viv"]
    f = dict(vit=rfields["vit"], viv=rfields["viv"], up=up * rfields["viv"], vp=vp * rfields["viv"],
             h0=rng.normal(0.0, 0.3, s2), rict=rfields["rict"], ric=rfields["ric"],
             h0bl=np.full(s2, sentinel), h0bf=rng.normal(0.0, 0.3, s2))
    for k in ("wp12", "wp13", "riv1", "riv2", "s2t", "s2u", "riu"):
        f[k] = np.full((km, jmt, imt), sentinel)
    f["rit"] = np.full((km - 1, jmt, imt), sentinel)
    return grid, {k: np.ascontiguousarray(v, np.float64) for k, v in f.items()}


# ---------------------------------------------- readyc_ (momentum part)
def readyc_mgrid(case: Case, rgrid: dict, del4: bool = False, seed: int = 15) -> dict:
    """licom_readyc_mgrid of a tracer case's block: POP block bounds of a
    two-cell halo (ib = jb = 3, ie = imt-2, je = jmt-2, 1-based), the case's
    kmt and metrics with the readyt_ grid's kmu, ohbu, ohbt, dzp, a
    Cartesian-like momentum Laplacian (du* about 1/dx^2 around a negative
    centre) with small metric and rotation terms (dum, dm*), a del4 factor
    amf spread over three decades so that the am_f limits 1 and 40 and the
    range between are all met, hemisphere signs of both kinds, and a
    turning angle of 10 degrees."""
    rng = np.random.default_rng(seed)
    km, jmt, imt = case.km, case.jmt, case.imt
    g = case.grid
    s2 = (jmt, imt)
    dxu, dyu = g["dxu"], g["dyu"]
    uarea = dxu * dyu
    pert = lambda: 1.0 + rng.uniform(-0.1, 0.1, s2)  # noqa: E731
    due, duw = pert() / (dxu * dxu), pert() / (dxu * dxu)
    dun, dus = pert() / (dyu * dyu), pert() / (dyu * dyu)
    duc = -(due + duw + dun + dus)
    small = lambda s: s * rng.normal(0.0, 1.0, s2) / (dxu * dyu)  # noqa: E731
    dum = small(1.0e-3)
    dmc, dmn, dms, dme, dmw = (small(1.0e-2) for _ in range(5))
    dzp = rgrid["dzp"]
    theta = np.deg2rad(10.0)
    jj = np.arange(jmt)[:, None] * np.ones((1, imt))
    return dict(ib=3, ie=imt - 2, jb=3, je=jmt - 2, del4=int(bool(del4)),
                am=-1.0e14 if del4 else 2.0e4, c0f=2.6e-3, od0=float(case.params["od0"]),
                cag=float(np.cos(theta)), sag=float(np.sin(theta)),
                kmt=g["kmt"].astype(np.int32), kmu=rgrid["kmu"].astype(np.int32),
                ohbu=rgrid["ohbu"], ohbt=rgrid["ohbt"], hts=g["hts"], htw=g["htw"], hue=g["hue"],
                hun=g["hun"], tarea_r=1.0 / g["tarea"], uarea_r=1.0 / uarea, dxu=dxu, dyu=dyu,
                dxur=rgrid["dxur"], dyur=rgrid["dyur"],
                snlat=np.where(jj < jmt // 3, -1.0, 1.0), amf=10.0 ** rng.uniform(-2.0, 1.0, s2),
                duc=duc, dum=dum, dun=dun, dus=dus, due=due, duw=duw,
                dmc=dmc, dmn=dmn, dms=dms, dme=dme, dmw=dmw, dzp=dzp, odzp=1.0 / dzp)


def readyc_minputs(case: Case, rfields: dict, cfields: dict, akmu: np.ndarray, seed: int = 16,
                   sentinel: float = 7.0) -> dict:
    """licom_readyc_mfields over a tracer case, its readyt_ fields and its
    readyc_ mixing inputs: current velocities (the previous ones of the
    mixing pass perturbed by 5%), wind stress, AKMU, ws with a random
    surface level, outputs starting at `sentinel`."""
    rng = np.random.default_rng(seed)
    km, jmt, imt = case.km, case.jmt, case.imt
    s2, s3 = (jmt, imt), (km, jmt, imt)
    viv = rfields["viv"]
    up, vp = cfields["up"], cfields["vp"]
    f = dict(u=up * (1.0 + rng.normal(0.0, 0.05, s3)) * viv,
             v=vp * (1.0 + rng.normal(0.0, 0.05, s3)) * viv,
             up=up, vp=vp, h0=cfields["h0"], su=rng.normal(0.0, 0.1, s2), sv=rng.normal(0.0, 0.1, s2),
             vit=rfields["vit"], viv=viv, akmu=akmu, ws=np.full(s3, sentinel))
    f["ws"][0] = rng.normal(0.0, 1.0e-6, s2)
    for k in ("dlu", "dlv", "wka"):
        f[k] = np.full(s3, sentinel)
    for k in ("dlub", "dlvb", "sbcx", "sbcy", "bbcx", "bbcy"):
        f[k] = np.full(s2, sentinel)
    return {k: np.ascontiguousarray(v, np.float64) for k, v in f.items()}


# ------------------------------------------------------------------ bclinc_
def bclinc_grid(case: Case, rgrid: dict, dtc: float = 1200.0, seed: int = 17) -> dict:
    """licom_bclinc_grid of a tracer case's block: the readyt_ grid's kmu,
    ohbu, ohbt, dzp, odzt and 1/dxu, 1/dyu, T-level depths zkt (negative
    down), a Coriolis parameter from the case's latitudes with both signs,
    the Coriolis factors of the model's formula (grid.F90's EPEA..EPLB with
    eps = f dtc / 2 and f dtc), LICOM's filter afc1 = 0.43, nbb = 3."""
    rng = np.random.default_rng(seed)
    jmt, imt = case.jmt, case.imt
    lat = np.deg2rad(case.grid["lat_t"])[:, None] * np.ones((1, imt))
    lat = np.where(np.arange(jmt)[:, None] < jmt // 3, -lat, lat)
    fcor = 2.0 * 7.292e-5 * np.sin(lat) * (1.0 + rng.uniform(-0.01, 0.01, (jmt, imt)))
    eps_e, eps_l = 0.5 * fcor * dtc, fcor * dtc
    dzp = rgrid["dzp"]
    afc1 = 0.43
    theta = np.deg2rad(10.0)
    return dict(od0=float(case.params["od0"]), c0f=2.6e-3, cag=float(np.cos(theta)),
                sag=float(np.sin(theta)), onbb=1.0 / 4.0, dtc=dtc, dtc2=2.0 * dtc, afc1=afc1,
                afc2=1.0 - 2.0 * afc1, kmu=rgrid["kmu"].astype(np.int32), snlat=np.sign(lat) + (lat == 0),
                ohbt=rgrid["ohbt"], ohbu=rgrid["ohbu"], fcor=fcor, dxur=rgrid["dxur"], dyur=rgrid["dyur"],
                epea=1.0 / (1.0 + eps_e * eps_e), epeb=eps_e / (1.0 + eps_e * eps_e),
                epla=1.0 / (1.0 + eps_l * eps_l), eplb=eps_l / (1.0 + eps_l * eps_l),
                dzp=dzp, odzt=rgrid["odzt"], odzp=1.0 / dzp, zkt=rgrid["zkt"])


def bclinc_fields(case: Case, rfields: dict, seed: int = 18) -> dict:
    """licom_bclinc_fields over a tracer case and its readyt_ fields: two
    time levels of velocities from the case's (perturbed), buoyancy gg and
    surface pressure of a smooth synthetic state, AKMU, wind stress,
    barotropic velocities, readyc_-like tendencies, time sums."""
    rng = np.random.default_rng(seed)
    km, jmt, imt = case.km, case.jmt, case.imt
    s2, s3 = (jmt, imt), (km, jmt, imt)
    viv, vit = rfields["viv"], rfields["vit"]
    st = case.state
    u = st["utf"] * viv
    v = st["vtf"] * viv
    f = dict(vit=vit, viv=viv, gg=rng.normal(-1.0e-2, 2.0e-3, s3) * vit,
             akmu=rng.uniform(1.0e-4, 5.0e-2, s3) * viv,
             su=rng.normal(0.0, 0.1, s2), sv=rng.normal(0.0, 0.1, s2),
             psa=1.0e3 + rng.normal(0.0, 5.0, s2), h0bl=rng.normal(0.0, 0.3, s2),
             ub=rng.normal(0.0, 0.05, s2), vb=rng.normal(0.0, 0.05, s2), h0bf=rng.normal(0.0, 1.2, s2),
             u=u, v=v, up=u * (1.0 + rng.normal(0.0, 0.02, s3)), vp=v * (1.0 + rng.normal(0.0, 0.02, s3)),
             dlu=rng.normal(0.0, 1.0e-6, s3) * viv, dlv=rng.normal(0.0, 1.0e-6, s3) * viv,
             wka=rng.normal(0.0, 1.0, s3), utf=rng.normal(0.0, 0.1, s3), vtf=rng.normal(0.0, 0.1, s3),
             sbcx=np.full(s2, 7.0), sbcy=np.full(s2, 7.0), bbcx=np.full(s2, 7.0), bbcy=np.full(s2, 7.0))
    return {k: np.ascontiguousarray(val, np.float64) for k, val in f.items()}


# ------------------------------------------------------------------ barotr_
def barotr_grid(case: Case, rgrid: dict, del4: bool = True, dtb: float = 20.0, seed: int = 19) -> dict:
    """barotr_'s grid over a tracer case's block: the readyc_ momentum
    grid's metrics and block bounds, kmt's neighbour levels kmtn..kmtw, the
    tracer del2 coefficients dtn..dtw (the case's own, Cartesian-like when
    it has none), the barotropic Coriolis factors ebea/ebeb (grid.F90's
    formula with eps = f dtb / 2), am and ah for the scheme."""
    rng = np.random.default_rng(seed)
    jmt, imt = case.jmt, case.imt
    mg = readyc_mgrid(case, rgrid, del4=del4, seed=seed)
    kmt = case.grid["kmt"].astype(np.int32)
    pad = np.zeros_like(kmt)
    kmtn, kmts, kmte, kmtw = pad.copy(), pad.copy(), pad.copy(), pad.copy()
    kmtn[1:] = kmt[:-1]
    kmts[:-1] = kmt[1:]
    kmte[:, :-1] = kmt[:, 1:]
    kmtw[:, 1:] = kmt[:, :-1]
    dx = case.grid["dxu"]
    tl = lambda: (1.0 + rng.uniform(-0.1, 0.1, (jmt, imt))) / (dx * dx)  # noqa: E731
    lat = np.deg2rad(case.grid["lat_t"])[:, None] * np.ones((1, imt))
    fcor = 2.0 * 7.292e-5 * np.sin(np.where(np.arange(jmt)[:, None] < jmt // 3, -lat, lat))
    eps = 0.5 * fcor * dtb
    g = {k: mg[k] for k in ("ib", "ie", "jb", "je", "del4", "kmu", "htw", "hts", "tarea_r", "uarea_r", "dxu",
                            "dyu", "dxur", "dyur", "amf", "duc", "dum", "dun", "dus", "due", "duw", "dmc",
                            "dmn", "dms", "dme", "dmw")}
    g.update(dtb=dtb, am=mg["am"], ah=-2.0e13 if del4 else 1.0e3, kmt=kmt, kmtn=kmtn, kmts=kmts, kmte=kmte,
             kmtw=kmtw, au0=rgrid["au0"], aus=rgrid["aus"], auw=rgrid["auw"], ausw=rgrid["ausw"], fcor=fcor,
             ebea=1.0 / (1.0 + eps * eps), ebeb=eps / (1.0 + eps * eps),
             ahf=10.0 ** rng.uniform(-1.0, 1.0, (jmt, imt)), dtn=tl(), dts=tl(), dte=tl(), dtw=tl())
    return g


def barotr_fields(case: Case, rfields: dict, seed: int = 20) -> dict:
    """barotr_'s fields over a tracer case: level 0 of the wet masks, layer
    depth dzph with the surface height, barotropic velocities at two levels,
    the forcing terms bclinc_/readyt_ hand over (dlub, pax, pxb, whx, wgp
    ...), running sums h0f/h0bf; work and wka start at a sentinel."""
    rng = np.random.default_rng(seed)
    jmt, imt = case.jmt, case.imt
    s2 = (jmt, imt)
    viv, vit = rfields["viv"][0], rfields["vit"][0]
    h0p = rng.normal(0.0, 0.5, s2) * vit
    ubp, vbp = rng.normal(0.0, 0.1, s2) * viv, rng.normal(0.0, 0.1, s2) * viv
    f = dict(vit=vit, viv=viv, dzph=(3000.0 + rng.uniform(-500.0, 500.0, s2)) * viv,
             dlub=rng.normal(0.0, 1.0e-6, s2), dlvb=rng.normal(0.0, 1.0e-6, s2), wgp=rng.uniform(0.9, 1.0, s2),
             pax=rng.normal(0.0, 1.0e-7, s2), pay=rng.normal(0.0, 1.0e-7, s2), pxb=rng.normal(0.0, 1.0e-6, s2),
             pyb=rng.normal(0.0, 1.0e-6, s2), whx=rng.normal(0.0, 1.0e-5, s2), why=rng.normal(0.0, 1.0e-5, s2),
             h0=h0p + rng.normal(0.0, 0.01, s2), h0p=h0p, ub=ubp + rng.normal(0.0, 0.01, s2), vb=vbp.copy(),
             ubp=ubp, vbp=vbp, h0f=rng.normal(0.0, 1.0, s2), h0bf=rng.normal(0.0, 1.0, s2),
             work=np.full(s2, 7.0), wka=np.full((6, jmt, imt), 7.0))
    return {k: np.ascontiguousarray(v, np.float64) for k, v in f.items()}
