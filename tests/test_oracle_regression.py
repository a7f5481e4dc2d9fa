"""The oracle reproduces its frozen outputs bit for bit (fixtures written by
tests/golden/make_oracle_fixtures.py)."""
import json
import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_oracle_fixtures as mk  # noqa: E402

with open(os.path.join(HERE, "golden", "oracle_regression.json")) as f:
    FIX = json.load(f)


@pytest.mark.parametrize("name", sorted(FIX["cases"]))
def test_oracle_matches_frozen_digests(name):
    case = FIX["cases"][name]
    out = mk.run(case["spec"])
    for k, h in case["sha256"].items():
        assert mk.digest(out[k]) == h, f"{name}: {k} changed"
