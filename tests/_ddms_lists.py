"""Adversarial distance lists for the DDMS term (128-byte records: the count,
then the values, then 255 padding; value 255 = missing, values >= 20 share
the pair-table row of 20).

Kinds, each aimed at one part of the per-cell update:
  const      one value repeated: long segments, equal pair scores, so equal
             sums at neighbouring cells (first-maximum selection);
  twenty     values >= 20 with a few small ones: pair scores near zero, sums
             that stay flat over many cells;
  alternate  two values whose pair scores have opposite signs: frequent
             restarts of the running sum;
  runs       blocks of equal values of random lengths;
  missing    many 255 entries (pair score -1);
  sorted     increasing values, as real distance lists are;
  random     uniform values with a few missing.
"""
import numpy as np

KINDS = ("const", "twenty", "alternate", "runs", "missing", "sorted", "random")


def values(rng, kind, n):
    if kind == "const":
        v = np.full(n, rng.integers(0, 31))
    elif kind == "twenty":
        v = rng.integers(20, 40, n)
        v[rng.random(n) < 0.1] = rng.integers(0, 8)
    elif kind == "alternate":
        a, b = int(rng.integers(0, 6)), int(rng.integers(14, 31))
        v = np.where(np.arange(n) % 2 == 0, a, b)
    elif kind == "runs":
        v = np.empty(n, np.int64)
        k = 0
        while k < n:
            r = int(rng.integers(1, 12))
            v[k:k + r] = rng.integers(0, 31)
            k += r
    elif kind == "missing":
        v = rng.integers(0, 31, n)
        v[rng.random(n) < 0.4] = 255
    elif kind == "sorted":
        v = np.sort(rng.integers(2, 40, n))
    else:
        v = rng.integers(0, 31, n)
        v[rng.random(n) < 0.05] = 255
    return np.asarray(v, np.uint8)


def record(rng, kind, n):
    rec = np.full(128, 255, np.uint8)
    rec[0] = n
    rec[1:n + 1] = values(rng, kind, n)
    return rec


def fill(pset, rng, lengths=None):
    """Overwrite every position's distance list of pset with adversarial
    records (kind and length drawn per position; lengths from the given
    choices, default 0..99 with the DSTSEGM boundary 5..9 over-represented)."""
    dd = pset.field("dd")
    for p in range(dd.shape[0]):
        kind = KINDS[int(rng.integers(0, len(KINDS)))]
        if lengths is not None:
            n = int(rng.choice(lengths))
        elif rng.random() < 0.3:
            n = int(rng.integers(5, 10))
        else:
            n = int(rng.integers(0, 100))
        dd[p] = record(rng, kind, n)
