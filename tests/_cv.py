"""Informative context vectors for test profile sets.

The synthetic generator gives every position the reference's uninformative
context vector (all 19 entries 0.17302947, prior and squared norm 0:
IOProfileModelBin.cu:277,434), so the CVS term is constant there.  This helper
writes random vectors, log priors and squared norms instead, as the text
profile format carries them ("CV: prior norm2 19 values", scaled integers),
quantised to 1/scale so that they survive a text round trip exactly.
"""
import numpy as np


def _q(x, scale):
    return (np.rint(np.asarray(x, np.float64) * scale).astype(np.int64).astype(np.float32) /
            np.float32(scale)).astype(np.float32)


def set_informative_cv(ps, seed, scale=10000, frac=1.0):
    """Give a fraction `frac` of the positions of ps random context vectors:
    entries N(0, s^2) with a per-position spread s in [0.02, 0.5], the squared
    norm of the quantised vector, and a log prior in [-40, 4] (which puts the
    CVS log-odds below, inside and above the range of the cvs2s map)."""
    rng = np.random.default_rng(seed)
    n = ps.npos
    cv, pr, n2 = ps.field("cv"), ps.field("cvprior"), ps.field("cvnorm2")
    pick = rng.random(n) < frac
    s = rng.uniform(0.02, 0.5, n)
    v = _q(rng.normal(0.0, 1.0, (19, n)) * s, scale)
    cv[:, pick] = v[:, pick]
    n2[pick] = _q((v.astype(np.float64) ** 2).sum(axis=0), scale)[pick]
    pr[pick] = _q(rng.uniform(-40.0, 4.0, n), scale)[pick]
    return ps
