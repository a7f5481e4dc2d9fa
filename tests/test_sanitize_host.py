"""Host AddressSanitizer / UndefinedBehaviorSanitizer run of the oracle and
the input generator (oracle/sanitize_driver.cpp, built by `make -C oracle
sanitize` with -fsanitize=address,undefined): every oracle entry point over
small seeded cases with land, del2/del4, forcing, restoring, icesnow and
convadj.  CPU only."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None,
                    reason="needs the host compilers")
def test_oracle_and_generator_are_clean_under_asan_and_ubsan():
    b = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "sanitize"],
                       capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stdout[-2000:] + b.stderr[-2000:]
    env = dict(os.environ,
               ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([os.path.join(ROOT, "oracle", "_san", "sanitize_driver")], capture_output=True,
                       text=True, timeout=300, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "runtime error" not in out and "AddressSanitizer" not in out, out[-4000:]
    assert out.count("ok ") == 3
