"""GPU: three stepon-style steps (tracer_, icesnow_, convadj_,
stepon.cpp:69-77) through the link-compatible shim, which reads LICOM's
module symbols and device pointers, equal bit for bit the same steps through
the C ABI.  Runs in a child process: the shim exits the process on an
error, as the reference does."""
import json
import os
import subprocess
import sys

import pytest

from conftest import ROOT
from test_gpu_parity import announce

pytestmark = pytest.mark.gpu


def test_shim_steps_equal_the_abi_steps(need_gpu):
    announce("shim: 3 x (tracer_, icesnow_, convadj_) on 128x96x30 vs the C ABI")
    r = subprocess.run([sys.executable, "-u", os.path.join(ROOT, "tests", "shim_host.py"), "gpu"],
                       capture_output=True, text=True, timeout=240)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stderr[-2000:]
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out.pop("ist") == 3
    assert out.pop("mixed_columns") >= 0
    assert all(v == 0 for v in out.values()), out
