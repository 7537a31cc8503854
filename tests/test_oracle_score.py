"""Pin the oracle's score-matrix stage (CPU only).

Checks, each chosen so a plausible restatement mistake fails one of them:
  - orc_logf against the correctly rounded logarithm (<= 1 ulp);
  - the table readers against entries of the reference-held files and the
    5-point running-mean definition of the CVS map smoothing;
  - the DDMS dynamic programme against the reference's own CPU DDMS
    (cother-scores-optimizer/ddms.cxx, built from source into oracle/_ref);
  - the DDMS pair score against its closed form;
  - the Init, SS and CVS terms against float64 restatements of their
    formulas (cusco/CuBatchScoreMatrix_Init.cu:129-185,
    cusco/CuBatchScoreMatrix_ssss.cuh:64-116, cusco/CuBatchScoreMatrix_cvs2s.cuh:80-113).
"""
import ctypes
import math
import os

import numpy as np
import pytest

import cother_amd
from tests import _oracle

ENV = cother_amd.ENVDIR


def ulp_diff(a, b):
    ia = np.array(a, np.float32).view(np.int32).astype(np.int64)
    ib = np.array(b, np.float32).view(np.int32).astype(np.int64)
    return np.abs(ia - ib)


def test_logf_within_one_ulp(oracle):
    rng = np.random.default_rng(1)
    xs = np.concatenate([
        rng.uniform(1e-30, 1e30, 2000).astype(np.float32),
        np.exp(rng.uniform(-80, 80, 20000)).astype(np.float32),
        rng.uniform(0.5, 2.0, 20000).astype(np.float32),
        np.array([1.0, 2.0, 0.5, 1e-40, 3.4e38, 1.1754944e-38], np.float32),
    ])
    got = np.array([oracle.logf(float(x)) for x in xs], np.float32)
    ref = np.log(xs.astype(np.float64)).astype(np.float32)
    assert ulp_diff(got, ref).max() <= 1
    assert oracle.logf(1.0) == 0.0
    assert math.isinf(oracle.logf(0.0)) and oracle.logf(0.0) < 0


def _file_rows(name):
    rows = []
    for ln in open(os.path.join(ENV, name)):
        if ln.strip() and not ln.lstrip().startswith("#"):
            rows.append(ln.split())
    return rows


def test_ss_table_matches_file(oracle):
    rows = _file_rows("ssss")
    tri = [r for r in rows if r and all(c.replace(".", "").replace("-", "").isdigit() for c in r)][1:]
    assert len(tri) == 30
    t = np.ctypeslib.as_array(oracle.tables.ss)
    k = 2
    for r in range(30):
        assert len(tri[r]) == r + 1
        for c in range(r + 1):
            assert t[k] == np.float32(float(tri[r][c]))
            k += 1


def test_cvs_map_is_running_mean(oracle):
    t = oracle.tables
    rows = _file_rows("cvs2s")
    keys = [float(x) for x in rows[5]]
    raw = [float(x) for x in rows[6] if x != "+"]
    card = t.cvcard
    assert card == len(keys) == 93 and t.cvshift == 5
    vals = np.ctypeslib.as_array(t.cvvals)[:card]
    # interior entries are the mean of the 5 raw scores centred on them
    for i in range(2, len(raw) - 2):
        assert abs(vals[i] - np.mean(raw[i - 2:i + 3])) < 2e-6
    # the ends take the first / last window mean
    assert abs(vals[0] - np.mean(raw[:5])) < 2e-6
    assert abs(vals[card - 1] - np.mean(raw[-5:])) < 2e-6
    assert np.ctypeslib.as_array(t.cvkeys)[0] == -2.5


def test_cvs_constants(oracle):
    t = oracle.tables
    dim, k0, nu0 = 19, 0.1, 19.0
    cterm = (0.5 * dim - 0.5 * (nu0 + 2)) * (math.log(k0) - math.log(k0 + 2))
    cterm += sum(math.log(nu0 + 1 - i) - math.log(2) for i in range(1, dim + 1))
    assert abs(t.cvcterm - cterm) < 1e-4
    assert t.cvkappa0inv == 10.0 and t.cvpowernu0 == -10.5


def test_ddms_pair_score_closed_form(oracle):
    tab = np.ctypeslib.as_array(oracle.tables.dms).reshape(22, 22)
    for a in range(21):
        for b in range(21):
            r, c = max(a, b), min(a, b)
            assert abs(tab[a, b] - (0.35 + 0.02 * c - 0.0375 * r)) < 1e-6
        assert tab[a, 21] == -1.0 and tab[21, a] == -1.0
    assert abs(tab[0, 0] - 0.35) < 1e-7 and abs(tab[20, 0] + 0.4) < 1e-7 and abs(tab[20, 20]) < 1e-7


def _ref_style_table():
    """Pair scores as the reference CPU DDMS builds them (separate mul/add)."""
    f32 = np.float32
    slope = (f32(-0.4) - f32(0.35)) * f32(0.05)
    intpt = [f32(0.35) - (f32(d) * f32(-0.4)) * f32(0.05) for d in range(21)]
    tab = np.full((22, 22), -1.0, np.float32)
    for a in range(21):
        for b in range(21):
            r, c = max(a, b), min(a, b)
            tab[a, b] = f32(r) * slope + intpt[c]
    return tab


def _random_list(rng, n, pmiss=0.03):
    v = rng.integers(0, 31, n)
    v[rng.random(n) < pmiss] = 255
    rec = np.full(128, 255, np.uint8)
    rec[0] = n
    rec[1:n + 1] = v
    return rec


@pytest.mark.skipif(_oracle.load_ref_ddms() is None, reason="reference CPU DDMS not built (oracle/_ref)")
def test_ddms_matches_reference_cpu_ddms(oracle):
    ref = _oracle.load_ref_ddms()
    tab = _ref_style_table()
    rng = np.random.default_rng(7)
    nchk = 0
    for it in range(3000):
        n1, n2 = int(rng.integers(24, 100)), int(rng.integers(24, 100))
        q, t = _random_list(rng, n1), _random_list(rng, n2)
        if it % 3 == 0:  # correlated lists give long high-scoring segments
            t[1:min(n1, n2) + 1] = q[1:min(n1, n2) + 1]
            t[1 + rng.integers(0, min(n1, n2), 5)] = rng.integers(0, 31, 5)
        qi = np.zeros(128 + 8, np.int32)
        ti = np.zeros(128 + 8, np.int32)
        qi[:128], ti[:128] = q, t
        qi[128:], ti[128:] = 255, 255
        want = ref.ref_ddms_calculate(qi.ctypes.data, ti.ctypes.data)
        _, m, _ = oracle.ddms_tab(tab, q, t)
        s = np.float32(0)
        for k in range(3):
            s = np.float32(s + m[k])
        got = np.float32(np.rint(np.float32(s * np.float32(16.0))))
        assert got == want, (it, n1, n2, got, want)
        nchk += 1
    assert nchk == 3000


@pytest.mark.skipif(_oracle.load_ref_ddms() is None, reason="reference CPU DDMS not built (oracle/_ref)")
def test_ddms_matches_reference_on_adversarial_lists(oracle):
    """The same comparison on hand-shaped lists (tests/_ddms_lists.py):
    constant, near-zero, alternating, run-structured, missing-heavy and
    sorted lists, including pairs of identical lists."""
    from tests import _ddms_lists as dl
    ref = _oracle.load_ref_ddms()
    tab = _ref_style_table()
    rng = np.random.default_rng(17)
    for it in range(2000):
        n1, n2 = int(rng.integers(24, 100)), int(rng.integers(24, 100))
        k1, k2 = dl.KINDS[it % len(dl.KINDS)], dl.KINDS[(it // len(dl.KINDS)) % len(dl.KINDS)]
        q, t = dl.record(rng, k1, n1), dl.record(rng, k2, n2)
        if it % 5 == 0:
            t = q.copy()
        qi = np.zeros(128 + 8, np.int32)
        ti = np.zeros(128 + 8, np.int32)
        qi[:128], ti[:128] = q, t
        qi[128:], ti[128:] = 255, 255
        want = ref.ref_ddms_calculate(qi.ctypes.data, ti.ctypes.data)
        _, m, _ = oracle.ddms_tab(tab, q, t)
        s = np.float32(0)
        for k in range(3):
            s = np.float32(s + m[k])
        got = np.float32(np.rint(np.float32(s * np.float32(16.0))))
        assert got == want, (it, k1, k2, n1, n2, got, want)


def test_ddms_short_lists_score_zero(oracle):
    rng = np.random.default_rng(3)
    q = _random_list(rng, 5)
    t = _random_list(rng, 80)
    assert oracle.L.orc_cell_ddms(ctypes.byref(oracle.tables), q.ctypes.data, t.ctypes.data) == 0.0


def test_ddms_no_segment_gives_intercept(oracle):
    # all-missing lists: every pair scores -1, no segment -> (0.848*0 - 1.235) * 0.2
    q = np.full(128, 255, np.uint8); q[0] = 40
    t = np.full(128, 255, np.uint8); t[0] = 40
    v = oracle.L.orc_cell_ddms(ctypes.byref(oracle.tables), q.ctypes.data, t.ctypes.data)
    assert abs(v - (-1.235 * 0.2)) < 1e-6


@pytest.fixture(scope="module")
def small_sets():
    q = cother_amd.synth(2, 0, lmin=30, lmax=60, dist_mode=0, sort_desc=0, seed=11)
    db = cother_amd.synth(6, 0, lmin=20, lmax=70, dist_mode=1, sort_desc=1, seed=12)
    return q, db


def test_init_ss_cvs_terms_float64(oracle, small_sets):
    q, db = small_sets
    T = oracle.tables
    qt, dt = q.field("trg"), db.field("trg")
    qb, dbk = q.field("bkg"), db.field("bkg")
    qs, ds = q.field("ss"), db.field("ss")
    qcv, dcv = q.field("cv"), db.field("cv")
    sst = np.ctypeslib.as_array(T.ss)
    dd, dl = db.dists(), db.lens()
    rng = np.random.default_rng(5)
    vp = ctypes.c_void_p
    for _ in range(300):
        qi = int(rng.integers(0, q.npos)); qp = int(np.searchsorted(q.dists(), qi, side="right") - 1)
        tj = int(rng.integers(0, db.npos)); tp = int(np.searchsorted(dd, tj, side="right") - 1)
        got = oracle.L.orc_cell_init(ctypes.cast(q.ptr, vp), qi, qp, ctypes.cast(db.ptr, vp), tj, tp)
        w = 0.5 * (qb[qp].astype(np.float64) + dbk[tp])
        want = math.log(float(np.sum(qt[:, qi].astype(np.float64) * dt[:, tj] / w))) - 0.002
        assert abs(got - want) < 2e-5 * max(1.0, abs(want))
        # SS
        s = 0.0
        for i in range(3):
            r = i * 10 + int(np.float32(qs[i, qi]) * np.float32(9.999))
            for j in range(3):
                c = j * 10 + int(np.float32(ds[j, tj]) * np.float32(9.999))
                a, b = max(r, c), min(r, c)
                s += float(qs[i, qi]) * float(ds[j, tj]) * float(sst[2 + a * (a + 1) // 2 + b])
        s = (s + (qs[2, qi] + 0.1) * (ds[2, tj] + 0.1) / 9.0 - 0.05) * 0.12
        got = oracle.L.orc_cell_ss(ctypes.byref(T), ctypes.cast(q.ptr, vp), qi, ctypes.cast(db.ptr, vp), tj)
        assert abs(got - s) < 1e-5
        # CVS (uninformative vectors: constant)
        dot = float(np.dot(qcv[:, qi].astype(np.float64), dcv[:, tj])) + 10.0
        det = 11.0 * 11.0 - dot * dot
        x = T.cvcterm + T.cvpowernu0 * math.log(det)
        want = np.ctypeslib.as_array(T.cvvals)[T.cvshift + int(np.rint(2 * x))] * 0.15
        got = oracle.L.orc_cell_cvs(ctypes.byref(T), ctypes.cast(q.ptr, vp), qi, ctypes.cast(db.ptr, vp), tj)
        assert abs(got - want) < 1e-6


def test_score_matrix_composition(oracle, small_sets):
    """S = Init + (SS+CVS) + DDMS; mod = (SS+CVS-0.075) + (DDMS-0.04) when DDMS != 0."""
    q, db = small_sets
    sc, md = oracle.score_matrix(q, 0, db, 0, db.nprof)
    T = oracle.tables
    vp = ctypes.c_void_p
    dd, qdd = db.field("dd"), q.field("dd")
    rng = np.random.default_rng(9)
    for _ in range(200):
        r = int(rng.integers(0, sc.shape[0])); c = int(rng.integers(0, sc.shape[1]))
        tp = int(np.searchsorted(db.dists(), c, side="right") - 1)
        i0 = oracle.L.orc_cell_init(ctypes.cast(q.ptr, vp), r, 0, ctypes.cast(db.ptr, vp), c, tp)
        s2 = np.float32(oracle.L.orc_cell_ss(ctypes.byref(T), ctypes.cast(q.ptr, vp), r, ctypes.cast(db.ptr, vp), c)) + \
            np.float32(oracle.L.orc_cell_cvs(ctypes.byref(T), ctypes.cast(q.ptr, vp), r, ctypes.cast(db.ptr, vp), c))
        d = np.float32(oracle.L.orc_cell_ddms(ctypes.byref(T), qdd[r].ctypes.data, dd[c].ctypes.data))
        want = np.float32(np.float32(i0) + s2)
        if d != 0:
            want = np.float32(want + d)
        assert sc[r, c] == want
        assert abs(md[r, c] - ((s2 - 0.075) + ((d - 0.04) if d != 0 else 0))) < 1e-6


def _cvs_float64(T, qcv, qn2, qpr, dcv, dn2, dpr):
    """The CVS term in float64 from the formula of CalcCVS2ScoreSMEMUnroll2x
    (cusco/CuBatchScoreMatrix_cvs2s.cuh:80-113; GetScoreE1S1Step2Boundary,
    cupro/SerializedCVS2Scores.cuh:377-420): the log-odds of the two vectors
    under the normal-Wishart model, det = (|q|^2 + 1 + 1/k0)(|t|^2 + 1 + 1/k0)
    - (q.t + 1/k0)^2, mapped through the smoothed cvs2s table; returns the
    mapped score x 0.15 and the log-odds before the map."""
    k = float(T.cvkappa0inv)
    dot = float(np.dot(qcv.astype(np.float64), dcv.astype(np.float64))) + k
    det = (float(qn2) + 1.0 + k) * (float(dn2) + 1.0 + k) - dot * dot
    if det <= 0.0:  # only for inconsistent inputs (a vector that does not match its norm)
        return math.nan, math.nan
    x = float(T.cvcterm) + float(T.cvpowernu0) * math.log(det) - float(qpr) - float(dpr)
    keys, vals, card = np.ctypeslib.as_array(T.cvkeys), np.ctypeslib.as_array(T.cvvals), T.cvcard
    if x <= keys[0]:
        v = vals[0]
    elif keys[card - 1] <= x:
        v = vals[card - 1]
    else:
        v = vals[T.cvshift + int(np.rint(2.0 * x))]
    return float(v) * 0.15, x


def test_cvs_term_informative_vectors(oracle):
    """orc_cell_cvs on random context vectors (entries, squared norms and log
    priors as text profiles carry them) against the float64 formula, over all
    three branches of the map; a transposed vector stride or a dropped prior
    gives other values."""
    from tests._cv import set_informative_cv
    q = set_informative_cv(cother_amd.synth(3, 0, lmin=30, lmax=60, dist_mode=2, sort_desc=0, seed=71), 72)
    db = set_informative_cv(cother_amd.synth(8, 0, lmin=30, lmax=90, dist_mode=2, sort_desc=0, seed=73), 74)
    T = oracle.tables
    vp = ctypes.c_void_p
    qcv, qn2, qpr = q.field("cv"), q.field("cvnorm2"), q.field("cvprior")
    dcv, dn2, dpr = db.field("cv"), db.field("cvnorm2"), db.field("cvprior")
    keys = np.ctypeslib.as_array(T.cvkeys)
    rng = np.random.default_rng(75)
    branch = [0, 0, 0]
    checked = stride_bad = prior_bad = 0
    for _ in range(3000):
        qi, tj = int(rng.integers(0, q.npos)), int(rng.integers(0, db.npos))
        want, x = _cvs_float64(T, qcv[:, qi], qn2[qi], qpr[qi], dcv[:, tj], dn2[tj], dpr[tj])
        if abs(2.0 * x - np.floor(2.0 * x) - 0.5) < 1e-3:
            continue  # single precision may round to the neighbouring entry
        got = oracle.L.orc_cell_cvs(ctypes.byref(T), ctypes.cast(q.ptr, vp), qi, ctypes.cast(db.ptr, vp), tj)
        assert abs(got - want) < 1e-6, (qi, tj, x)
        checked += 1
        branch[0 if x <= keys[0] else (2 if x >= keys[T.cvcard - 1] else 1)] += 1
        # planted mistakes, in the float64 formula
        tv = qcv.reshape(-1)[qi * 19:(qi + 1) * 19] if (qi + 1) * 19 <= qcv.size else qcv[:, qi]
        stride_bad += not abs(_cvs_float64(T, tv, qn2[qi], qpr[qi], dcv[:, tj], dn2[tj], dpr[tj])[0] - want) <= 1e-4
        prior_bad += not abs(_cvs_float64(T, qcv[:, qi], qn2[qi], 0.0, dcv[:, tj], dn2[tj], dpr[tj])[0] - want) <= 1e-4
    assert checked > 2500
    assert min(branch) > 50, branch
    assert stride_bad > checked // 2 and prior_bad > checked // 2
