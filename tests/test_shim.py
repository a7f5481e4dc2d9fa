"""The link-compatible shim (licom3-hip_amd/shim/licom_shim.cpp): it builds
for the biharmonic and the Laplacian build, defines tracer_, convadj_ and
icesnow_ with the reference's C linkage, needs exactly LICOM's symbols, and
fails as the reference does (message, exit status 1) -- here, where no GPU
is visible."""
import os
import subprocess
import sys

import pytest

from conftest import ROOT, gpu_available

PKG = os.path.join(ROOT, "licom3-hip_amd")


def _symbols(path, dynamic=False):
    out = subprocess.run(["nm"] + (["-D"] if dynamic else []) + [path], check=True,
                         capture_output=True, text=True).stdout
    defined, undefined = set(), set()
    for line in out.splitlines():
        parts = line.split()
        if len(parts) == 2 and parts[0] == "U":
            undefined.add(parts[1])
        elif len(parts) == 3:
            defined.add(parts[2])
    return defined, undefined


def test_shim_object_links_to_licom_and_the_abi():
    obj = os.path.join(PKG, "_build", "licom_shim_bihar.o")
    assert os.path.exists(obj), "build() compiles the shim"
    defined, undefined = _symbols(obj)
    assert {"tracer_", "convadj_", "icesnow_"} <= defined
    licom = {s for s in undefined if s.startswith(("pconst_mod_mp_", "grid_mp_", "hmix_", "pmix_", "d_"))}
    assert "hmix_del4_mp_ah_" in licom and "hmix_del2_mp_ah_" not in licom
    for s in ("pconst_mod_mp_dts_", "pconst_mod_mp_adv_tracer_", "pconst_mod_mp_nss_",
              "grid_mp_kmt_", "grid_mp_dxu_", "d_at", "d_atb", "d_vtl", "d_licomqice"):
        assert s in licom, s
    abi = {s for s in undefined if s.startswith("licom_")}
    assert abi == {"licom_tracer_create", "licom_tracer_step", "licom_convadj", "licom_icesnow",
                   "licom_tracer_sync", "licom_strerror", "licom_last_error"}


@pytest.mark.skipif(gpu_available(), reason="checks the failure path where no GPU is visible")
def test_shim_fails_like_the_reference_without_a_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "shim_host.py"), "nogpu"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 1, (r.returncode, r.stdout, r.stderr)
    assert "licom tracer_ setup" in r.stderr and "no HIP device" in r.stderr, r.stderr
    assert "returned" not in r.stdout
