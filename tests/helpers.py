"""Comparison helpers for the parity tests."""
import numpy as np

INNER = (slice(2, -2), slice(2, -2))


def inner(a):
    """Physical cells [2,jmt-2) x [2,imt-2) of a [..., jmt, imt] array."""
    return a[(Ellipsis,) + INNER]


def rel_err(a, b):
    """max |a-b| / max |b| (0 when both vanish)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.max(np.abs(a - b)) if a.size else 0.0
    s = np.max(np.abs(b)) if b.size else 0.0
    return 0.0 if d == 0.0 else (d / s if s > 0 else np.inf)


def mismatch(a, b):
    """Number of entries that differ in value (+0 and -0 are equal)."""
    return int(np.count_nonzero(np.asarray(a) != np.asarray(b)))


def report(name, a, b):
    n = mismatch(a, b)
    return f"{name}: {n} mismatches, rel err {rel_err(a, b):.3e}"
