"""Plays LICOM's side of the link-compatible shim (licom3-hip_amd/shim):
fills the module symbols and cuda_data.h device pointers that
libshim_check.so defines from a synthetic case, then calls tracer_, convadj_
and icesnow_ the way stepon_ does (stepon.cpp:69-77).

Run as a script (a child process: the shim exits the process on an error,
as the reference does):  python shim_host.py <mode>
  mode "nogpu": tracer_(0) on a machine without a GPU; the shim must print
                the library's error and exit(1)
  mode "gpu":   three stepon-style steps through the shim and through the C
                ABI directly (licom_tracer.TracerModel) on the same inputs;
                prints one JSON line with the mismatch counts
"""
import ctypes
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("licom3-hip_amd", "oracle"):
    sys.path.insert(0, os.path.join(ROOT, sub))

import licom_synth as ls  # noqa: E402

NX, NY, KM = 128, 96, 30        # the block libshim_check.so is built for


def load():
    return ctypes.CDLL(os.path.join(ROOT, "licom3-hip_amd", "libshim_check.so"))


def set_module_symbols(L, case, keep):
    p, g = case.params, case.grid
    for name in ("dts", "onbc", "oncc", "dwndmix", "gamma", "od0", "od0cp"):
        ctypes.c_double.in_dll(L, f"pconst_mod_mp_{name}_").value = float(p[name])
    ctypes.c_int.in_dll(L, "pconst_mod_mp_nss_").value = int(p.get("nss", 1))
    ctypes.c_int.in_dll(L, "pconst_mod_mp_boundary_restore_").value = int(p["boundary_restore"])
    scheme = {ls.ADV_CENTERED: b"centered", ls.ADV_FLUX: b"flux", ls.ADV_TSPAS: b"tspas"}[p["adv"]]
    buf = (ctypes.c_char * 80).in_dll(L, "pconst_mod_mp_adv_tracer_")
    buf.value = scheme
    ctypes.c_double.in_dll(L, "hmix_del2_mp_ah_").value = float(p["ah"])
    for sym, arr in (("pconst_mod_mp_zkp_", g["zkp"]), ("pconst_mod_mp_to_", g["to"]),
                     ("pconst_mod_mp_so_", g["so"]), ("pmix_mod_mp_pen_", g["pen"]),
                     ("pconst_mod_mp_c_", g["eosc"])):
        a = np.ascontiguousarray(arr, np.float64).ravel()
        (ctypes.c_double * a.size).in_dll(L, sym)[:] = a.tolist()
    for f in ("kmt", "htw", "hts", "hun", "hue", "tarea", "dxu", "dyu"):
        a = np.ascontiguousarray(g[f])
        keep.append(a)
        ctypes.c_void_p.in_dll(L, f"grid_mp_{f}_").value = a.ctypes.data


def set_device_pointers(L, model):
    ptr = lambda t: ctypes.c_void_p(t.data_ptr() if t is not None else 0)  # noqa: E731
    for k, t in model.state.items():
        ctypes.c_void_p.in_dll(L, f"d_{k}").value = ptr(t).value
    for k, t in model.forcing.items():
        ctypes.c_void_p.in_dll(L, f"d_{k}").value = ptr(t).value
    for k in ("ax", "ay", "az", "dz", "tend", "dt_diff", "dt_conv", "penetrate", "net"):
        ctypes.c_void_p.in_dll(L, f"d_{k}").value = ptr(model.diag[k]).value
    ctypes.c_void_p.in_dll(L, "d_licomqice").value = ptr(model.licomqice).value


def case_for_shim():
    c = ls.make_case("C1", nx=NX, ny=NY, km=KM, seed=31, depth_min=-3000.0, sigma=1.5)
    c.params["sssnorm"] = 1
    return c


def main(mode):
    if mode == "gpu":
        # the HIP runtime torch loads must be the process's first, as in every
        # other caller here (torch initialises the device before the library
        # is loaded); loading the shim first would bring in a second runtime
        import torch
        torch.cuda.init()
    L = load()
    keep = []
    case = case_for_shim()
    set_module_symbols(L, case, keep)
    if mode == "nogpu":
        L.tracer_(ctypes.c_int(0))
        print("tracer_ returned without a GPU", flush=True)
        return 0
    import licom_tracer as lt
    a = lt.TracerModel(case)                 # driven through the shim's globals
    b = lt.TracerModel(case)                 # driven through the C ABI directly
    set_device_pointers(L, a)
    for s in range(3):
        L.tracer_(ctypes.c_int(s))           # stepon.cpp:69
        L.icesnow_()                          # stepon.cpp:71
        L.convadj_()                          # stepon.cpp:77
        b.tracer(s, lt.LICOM_STEP_SIDE_EFFECTS | lt.LICOM_STEP_VTL_SCRATCH)
        b.icesnow(reset=True)
        b.convadj()
    b.sync()
    sa, sb = a.host_state(), b.host_state()
    da, db = a.host_diag(), b.host_diag()
    out = {k: int(np.count_nonzero(sa[k] != sb[k])) for k in sa}
    out.update({f"diag_{k}": int(np.count_nonzero(da[k] != db[k])) for k in da})
    out["licomqice"] = int(np.count_nonzero(a.licomqice.cpu().numpy() != b.licomqice.cpu().numpy()))
    out["ist"] = ctypes.c_int.in_dll(L, "pconst_mod_mp_ist_").value
    out["mixed_columns"] = b.mix_count()
    print(json.dumps(out), flush=True)
    a.close()
    b.close()
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1]))
