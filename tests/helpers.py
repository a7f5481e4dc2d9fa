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


TBICE = -1.8   # freezing point of sea water (constant_mod.F90:168)


def plant_cold(case, st, seed, frac=0.05):
    """Cools a random `frac` of the cells of T (every column, ghosts and land
    included) to between 1.2 and 0.01 K below TBICE, and a larger share of
    the surface, so that each branch of icesnow has cases.  In place."""
    rng = np.random.default_rng(seed)
    T = st["at"][0]
    cold = rng.random(T.shape) < frac
    cold[0] |= rng.random(T.shape[1:]) < 4 * frac
    T[cold] = TBICE - rng.uniform(0.01, 1.2, int(cold.sum()))
    return st
