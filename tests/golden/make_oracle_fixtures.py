"""Writes tests/golden/oracle_regression.json: SHA-256 digests (and a few
sample values) of the oracle's outputs on small seeded cases, so that any
later change to the oracle that alters a single bit is caught by
tests/test_oracle_regression.py.  The reference ships no golden vectors for
this path; these fixtures freeze the oracle as pinned by test_oracle_pins.py.

Run from the repository root:  python tests/golden/make_oracle_fixtures.py
"""
import hashlib
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "licom3-hip_amd"), os.path.join(ROOT, "oracle")]

import numpy as np  # noqa: E402

import licom_synth as ls  # noqa: E402
import orc  # noqa: E402

CASES = {
    "c1_land_del2": dict(name="C1", nx=40, ny=24, km=12, seed=31, depth_min=-2000.0, sigma=1.5),
    "c1_del4": dict(name="C1", nx=32, ny=20, km=10, seed=32, hmix=4, ah=-1e10, dts=600.0),
}
STEPS = 3


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def run(spec):
    spec = dict(spec)
    c = ls.make_case(spec.pop("name"), **spec)
    o = orc.Oracle(c)
    st = c.copy_state()
    dg = orc.Oracle.new_diag(c)
    for _ in range(STEPS):
        o.step(st, dg, convadj=1)
    out = {k: st[k] for k in ("at", "atb", "ws", "utf", "akt")}
    out.update({k: dg[k] for k in ("ax", "az", "tend", "dt_diff", "dt_conv")})
    return out


def main():
    fx = {"steps": STEPS, "cases": {}}
    for name, spec in CASES.items():
        out = run(spec)
        fx["cases"][name] = {"spec": spec,
                             "sha256": {k: digest(v) for k, v in out.items()},
                             "at_sample": [float(x) for x in out["at"].ravel()[::997][:16]]}
    with open(os.path.join(os.path.dirname(__file__), "oracle_regression.json"), "w") as f:
        json.dump(fx, f, indent=1)


if __name__ == "__main__":
    main()
