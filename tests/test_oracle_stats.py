"""Pin the oracle's statistics stage (CPU only).

- special functions against scipy.special (betainc, gammaincc) in float64;
- lambda, E, H of the score distribution against their defining equations
  (sum_k p_k e^{lambda s_k} = 1, E = sum p s, H = lambda sum p s e^{lambda s});
- K against a textbook Karlin-Altschul computation (sigma series over the
  exact convolution P_j = P_{j-1} * P_1, scores limited to [-32, 32]) on
  score ranges where the reference's 21x32 grid layout reads every cell at
  its own score (j*max < 16 until the series stops), and the K special cases;
- the search-space length adjustment against its fixed-point equation.
"""
import math
import os

import numpy as np
import pytest
import scipy.special as sp


def test_special_functions_vs_scipy(oracle):
    for a in [1.0, 1.5, 3.0, 11.0, 57.0, 400.0]:
        for b in [0.1, 0.5, 2.0]:
            for x in [0.05, 0.3, 0.757213, 0.95]:
                v = sp.betainc(a, b, x)
                if v < 1e-290:
                    continue
                want = math.log(v)
                got = oracle.L.orc_lbetainc(a, b, x)
                assert abs(got - want) < 1e-9 * max(1, abs(want)), (a, b, x, got, want)
    for a in [0.5, 1.346206, 4.0]:
        for z in [1e-3, 0.3, 1.0, 2.5, 7.0, 30.0, 200.0]:
            q = sp.gammaincc(a, z)
            want = math.log(q) if q > 0 else None
            got = oracle.L.orc_lgammaincq(a, z)
            if want is not None:
                assert abs(got - want) < 1e-9 * max(1, abs(want)), (a, z, got, want)


def hist_from_unscaled(p_unscaled, rng, n=200000):
    """counts over the 641 scaled bins from a distribution on unscaled scores
    (each unscaled score u lands in its central bin u*32)."""
    counts = np.zeros(641)
    for u, pu in p_unscaled.items():
        counts[u * 32 + 320] += round(pu * n)
    return counts


def textbook(counts):
    p = counts / counts.sum()
    s = np.arange(641) - 320.0
    E = float((p * s).sum())
    nz = p > 0
    p, s = p[nz], s[nz]
    from scipy.optimize import brentq
    f = lambda lam: float((p * np.exp(lam * s)).sum() - 1.0)
    lam_s = brentq(f, 1e-6, 2.0, xtol=1e-14)
    H = float((p * lam_s * s * np.exp(lam_s * s)).sum())
    return E / 32, lam_s * 32, H


def textbook_K(pu, lam, H):
    """sigma series over exact convolutions on unscaled integer scores."""
    minu, maxu = min(pu), max(pu)
    P1 = {u: v for u, v in pu.items()}
    Pj = {0: 1.0}
    sigma = 0.0
    for j in range(1, 33):
        nxt = {}
        for s0, v0 in Pj.items():
            for u, v in P1.items():
                s = s0 + u
                if -32 <= s <= 32:
                    nxt[s] = nxt.get(s, 0.0) + v0 * v
        Pj = nxt
        tot = sum(v * (math.exp(lam * s) if s < 0 else 1.0) for s, v in Pj.items())
        sigma += tot / j
        if tot <= 1e-4:
            break
    return lam * math.exp(-2 * sigma) / (H * (1 - math.exp(-lam)))


@pytest.mark.parametrize("dist", [
    {-4: 0.3, -3: 0.25, -2: 0.2, -1: 0.15, 0: 0.05, 1: 0.04, 2: 0.01},
    {-5: 0.4, -2: 0.3, -1: 0.2, 1: 0.08, 2: 0.02},
    {-3: 0.5, -2: 0.3, 0: 0.15, 1: 0.05},
])
def test_lambda_E_H_K_textbook(oracle, dist):
    counts = hist_from_unscaled(dist, None)
    out = oracle.stat_params(counts)
    umin, umax, E, H, lam, K = out
    assert umin == min(dist) and umax == max(dist)
    tE, tlam, tH = textbook(counts)
    assert abs(E - tE) < 1e-12
    assert abs(lam - tlam) < 1e-6 * tlam   # Newton stops at |dx| < 1e-4 (scaled); the root is reached closely
    assert abs(H - tH) < 1e-5 * tH
    pu = {u: counts[u * 32 + 320] / counts.sum() for u in dist}
    if max(dist) == 1 and min(dist) != -1:
        tK = H / lam * (1 - math.exp(-lam))  # Karlin-Altschul closed form when the high score is the gcd
    else:
        tK = textbook_K(pu, lam, H)
    assert abs(K - tK) < 1e-9 * tK, (K, tK)


def test_K_special_cases(oracle):
    counts = hist_from_unscaled({-1: 0.7, 0: 0.1, 1: 0.2}, None)
    umin, umax, E, H, lam, K = oracle.stat_params(counts)
    p = counts / counts.sum()
    pm, pp = p[-32 + 320], p[32 + 320]
    assert abs(K - (pm - pp) ** 2 / pm) < 1e-12
    counts = hist_from_unscaled({-3: 0.6, 0: 0.2, 1: 0.2}, None)
    umin, umax, E, H, lam, K = oracle.stat_params(counts)
    assert abs(K - H / lam * (1 - math.exp(-lam))) < 1e-12
    # a non-negative expected score fails lambda and K
    counts = hist_from_unscaled({-1: 0.2, 1: 0.8}, None)
    out = oracle.stat_params(counts)
    assert out[4] == -1.0 * 32 and out[5] == -1.0


def test_significance_monotone_and_finite(oracle):
    prev = None
    for score in [5.0, 10.0, 20.0, 40.0, 80.0]:
        lmbdest, kest, logE, logEpA = oracle.significance(2, 250, 220, 7.0, 5.0, 3.0e9, score, score / 3, 0.3, 0.05)
        assert math.isfinite(logE) and math.isfinite(logEpA)
        assert abs(logEpA - (math.log(0.05 * 250 * 220) - 0.3 * score)) < 1e-9
        assert abs(lmbdest - 0.3 * 0.561132 / 1.116276) < 1e-6 and abs(kest - 0.05 * 0.001488 / 0.115812) < 1e-7
        if prev is not None:
            assert logE < prev
        prev = logE


def test_length_adjustment_fixed_point(oracle):
    for m, n, N in [(150, 447922, 2000), (250, 14_000_000, 60000), (40, 10000, 50), (1000, 14_000_000, 60000)]:
        ss, dl = oracle.length_adjustment(m, n, N)
        lam, K, alpha, beta = 0.561132, 0.001488, 0.5, -1.0
        if m * n - max(m, n) / K < 0:  # no length satisfies K(m-l)(n-Nl) >= max(m,n)
            assert dl == 0 and ss == m * n
            continue
        rhs = alpha / lam * math.log(K * (m - dl) * (n - N * dl)) + beta
        assert dl <= rhs + 1e-3 and rhs - dl < 1.0 + 1e-3
        # the search space is that of the last length iterate, between dl and dl+1
        assert (m - dl - 1) * (n - N * (dl + 1)) * (1 - 1e-6) <= ss <= (m - dl) * (n - N * dl) * (1 + 1e-6)


LENADJ = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "lenadj_ref")


@pytest.mark.skipif(not os.path.exists(LENADJ), reason="reference length adjustment not built (oracle/Makefile.ref)")
def test_length_adjustment_equals_compiled_reference(oracle):
    """The search space and length adjustment of the oracle and of the
    product's host code (cth_search_space) equal those of the reference's own
    AbstractScoreMatrix::ComputeLengthAdjustment (AbstractScoreMatrix.cpp:
    303-395, built from source into oracle/_ref/lenadj_ref), on query lengths,
    database sizes and counts across the configurations and beyond, including
    the cases with no admissible length (c < 0) and those that need the
    bisection."""
    import subprocess
    import cother_amd
    rng = np.random.default_rng(77)
    cases = [(150, 447922, 2000), (250, 14074887, 60000), (1000, 14074887, 60000), (2000, 9950000, 5000),
             (40, 10000, 50), (1, 1, 1), (5000, 10 ** 9, 10 ** 6)]
    for _ in range(3000):
        m = int(rng.integers(1, 5001))
        n = int(10 ** rng.uniform(2, 9.5))
        N = int(max(1, min(n, 10 ** rng.uniform(0, 6.5))))
        cases.append((m, n, N))
    inp = "".join(f"{m} {n} {N}\n" for m, n, N in cases)
    r = subprocess.run([LENADJ, "0.561132", "0.001488", "0.5", "-1"], input=inp, capture_output=True, text=True,
                       check=True)
    ref = [tuple(int(x) for x in l.split()) for l in r.stdout.splitlines()]
    assert len(ref) == len(cases)
    adjusted = 0
    for (m, n, N), (rss, rdl) in zip(cases, ref):
        assert cother_amd.search_space(m, n, N) == (rss, rdl), (m, n, N)
        assert oracle.length_adjustment(m, n, N) == (rss, rdl), (m, n, N)
        adjusted += rdl > 0
    assert adjusted > 500  # most cases take the iteration, not the early return
