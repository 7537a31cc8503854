"""The SS18 E-value of the oracle (oracle/orc_stats.c), pinned step by step
(CalcSignificanceSS18DynPlmProfile, cuss/CuBatchSS_dynplm.cu:447-580;
PredictEVDParametersSS18 and CalculateEvalueSS18, cuss/CuBatchSS_SS18.cuh:
119-351):

- the network weights and the adjustment parameters equal the reference's
  table (cupro/CuDeviceMemory.cu:43-77), read from its text at test time, and
  the GPU kernel's copy of it (cother_amd/csrc/stats.hip);
- the three networks evaluate as the table's comment documents them ("rows --
  units, columns -- inputs (0th -- biases)", one output row of bias and unit
  weights, activation 2/(1+e^-x) - 1 = tanh(x/2)), in float64 with numpy;
- mu and scale combine the networks with the adjustment parameters taken by
  their names in the reference's enum (TSS18hypers,
  cupro/CuBatchProcessing_com.cuh:61-72), for both variants;
- the score's p-value is the Gumbel tail (scipy.stats.gumbel_r);
- the positive-count p-value is the regularized incomplete beta
  (scipy.special.betainc; I_x(1, b) = 1 - (1 - x)^b in closed form);
- Brown's combination is the regularized upper gamma (scipy.special.gammaincc);
- the p-value to E-value step is E = -log(1 - p), scaled to the search space.

Constants are read from the reference's #defines (cuss/CuBatchSS_SS18.cuh:
17-41).  A transposed weight matrix, swapped adjustment indices, a dropped
K term or the wrong incomplete-beta argument order fail these checks (each
is planted below in the test's own evaluation and shown to differ).
"""
import ctypes
import math
import os
import re

import numpy as np
import pytest
from scipy import special, stats

REF = "/root/reference/src/libmycu"
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
needs_ref = pytest.mark.skipif(not os.path.isdir(REF), reason="reference sources absent")


def _ref_weights():
    src = open(os.path.join(REF, "cupro", "CuDeviceMemory.cu")).read()
    body = src[src.index("SSE_NN_WEIGHTS_[]"):]
    body = body[body.index("{") + 1:body.index("};")]
    body = "\n".join(l.split("//")[0] for l in body.splitlines())
    w = np.array([float(x) for x in re.findall(r"[-+]?\d+(?:\.\d*)?(?:[eE][-+]?\d+)?", body)], np.float32)
    assert len(w) == 93
    return w


def _ref_defines():
    src = open(os.path.join(REF, "cuss", "CuBatchSS_SS18.cuh")).read()
    d = {m.group(1): float(np.float32(float(m.group(2))))
         for m in re.finditer(r"#define\s+(CUSS_SS18_\w+)\s+([-+]?\d+(?:\.\d*)?(?:[eE][-+]?\d+)?)f?", src)}
    com = open(os.path.join(REF, "cuss", "CuBatchSS_com.h")).read()
    m = re.search(r"#define\s+CUSS_PSTS_MIN_LENPRODUCT\s+([\d.]+)", com)
    d["CUSS_PSTS_MIN_LENPRODUCT"] = float(m.group(1))
    return d


def _ref_hyper_names():
    src = open(os.path.join(REF, "cupro", "CuBatchProcessing_com.cuh")).read()
    body = src[src.index("enum TSS18hypers"):]
    body = body[body.index("{") + 1:body.index("}")]
    names = [l.split("//")[0].strip().rstrip(",") for l in body.splitlines()]
    names = [n for n in names if n]
    assert names[-1] == "nTSS18hypers"
    return {n: k for k, n in enumerate(names[:-1])}


class _L:
    def __init__(self, oracle):
        L = self.L = oracle.L
        d = ctypes.c_double
        vp = ctypes.c_void_p
        L.orc_ss18_weights.argtypes = [vp]
        L.orc_ss18_nn.restype = d
        L.orc_ss18_nn.argtypes = [ctypes.c_int, vp]
        L.orc_ss18_input_e.argtypes = [ctypes.c_int, ctypes.c_int, d, d, vp]
        L.orc_ss18_input_l.argtypes = [ctypes.c_int, ctypes.c_int, d, vp]
        L.orc_ss18_predict.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, d, d, d, ctypes.POINTER(d),
                                       ctypes.POINTER(d)]
        L.orc_ss18_score_logp.restype = d
        L.orc_ss18_score_logp.argtypes = [ctypes.c_int, d, d, d, d]
        L.orc_ss18_psts_logp.restype = d
        L.orc_ss18_psts_logp.argtypes = [ctypes.c_int, ctypes.c_int, d]
        L.orc_ss18_combine.restype = d
        L.orc_ss18_combine.argtypes = [d, d]
        L.orc_ss18_logevalue.restype = d
        L.orc_ss18_logevalue.argtypes = [d, d, ctypes.c_int, ctypes.c_int]

    def weights(self):
        w = np.zeros(93, np.float32)
        self.L.orc_ss18_weights(w.ctypes.data)
        return w

    def nn(self, which, x):
        x = np.ascontiguousarray(x, np.float64)
        return self.L.orc_ss18_nn(which, x.ctypes.data)

    def input_e(self, lq, lt, qe, te):
        x = np.zeros(4)
        self.L.orc_ss18_input_e(lq, lt, qe, te, x.ctypes.data)
        return x

    def input_l(self, lq, lt, lam):
        x = np.zeros(4)
        self.L.orc_ss18_input_l(lq, lt, lam, x.ctypes.data)
        return x

    def predict(self, model, lq, lt, qe, te, lam):
        mu, sc = ctypes.c_double(), ctypes.c_double()
        self.L.orc_ss18_predict(model, lq, lt, qe, te, lam, ctypes.byref(mu), ctypes.byref(sc))
        return mu.value, sc.value


@pytest.fixture(scope="module")
def ss(oracle):
    return _L(oracle)


def _net(w, x, factor, transposed=False):
    """The documented layout in float64: W1 [4 units][bias + 4 inputs], then
    [bias + 4 unit weights]; activation tanh(x/2)."""
    w = w.astype(np.float64)
    W1 = w[:20].reshape(4, 5)
    if transposed:  # a plausible mistake: units and inputs swapped
        W1 = np.concatenate([W1[:, :1], W1[:, 1:].T], axis=1)
    h = np.tanh(0.5 * (W1[:, 0] + W1[:, 1:] @ x))
    return np.tanh(0.5 * (w[20] + w[21:25] @ h)) * factor


@needs_ref
def test_weight_tables_equal_reference_text(ss):
    ref = _ref_weights()
    assert (ss.weights().view(np.int32) == ref.view(np.int32)).all()
    # the device table of the statistics kernel
    src = open(os.path.join(ROOT, "cother_amd", "csrc", "stats.hip")).read()
    body = src[src.index("c_nnw[93]"):]
    body = body[body.index("{") + 1:body.index("};")]
    dev = np.array([float(x) for x in re.findall(r"[-+]?\d+(?:\.\d*)?(?:[eE][-+]?\d+)?", body)], np.float32)
    assert (dev.view(np.int32) == ref.view(np.int32)).all()


@needs_ref
def test_networks_follow_documented_layout(ss):
    w = _ref_weights()
    rng = np.random.default_rng(5)
    for which, factor in ((0, 30.0), (1, 30.0), (2, 5.0)):
        ww = w[25 * which:25 * which + 25]
        worst = 0.0
        for _ in range(200):
            x = rng.uniform(-1.0, 1.5, 4)
            if which:
                x[3] = 0.0  # configuration L has three inputs and a zero-padded fourth (CuDeviceMemory.cu:57,67)
            got, want = ss.nn(which, x), _net(ww, x, factor)
            assert abs(got - want) <= 1e-12 * max(1.0, abs(want)), (which, x)
            worst = max(worst, abs(_net(ww, x, factor, transposed=True) - want))
        assert worst > 1e-2  # the transposed layout is told apart


@needs_ref
def test_network_inputs(ss):
    """FormatInputConfigurationE/L (cuss/CuBatchSS_NN.cuh:150-202): the pair
    ordered by ENO (then length), lengths as log(n)/log(1000); lambda rounded
    to one decimal and clamped to [0.1, 10], lengths ordered."""
    ln1k = math.log(1000.0)
    x = ss.input_e(100, 300, 4.0, 9.0)  # the target has the larger ENO: swapped
    np.testing.assert_allclose(x, [0.05 * 9.0, math.log(300) / ln1k, 0.05 * 4.0, math.log(100) / ln1k], rtol=1e-7)
    x = ss.input_e(100, 300, 6.0, 6.0)  # equal ENOs: the longer first
    np.testing.assert_allclose(x, [0.3, math.log(300) / ln1k, 0.3, math.log(100) / ln1k], rtol=1e-7)
    x = ss.input_l(80, 40, 0.26)
    np.testing.assert_allclose(x, [0.2 * 0.3, math.log(80) / ln1k, math.log(40) / ln1k, 0.0], rtol=1e-7)
    assert ss.input_l(80, 40, -1.0)[0] == pytest.approx(0.2 * 0.1)
    assert ss.input_l(80, 40, 37.0)[0] == pytest.approx(0.2 * 10.0)


def _predict_by_name(ss, w, names, d, model, lq, lt, qe, te, lam, swap=None):
    """mu and scale from the networks and the named adjustment parameters
    (PredictEVDParametersSS18 steps)."""
    emu = ss.nn(0, ss.input_e(lq, lt, qe, te))
    xl = ss.input_l(lq, lt, lam)
    lmu, lsc = ss.nn(1, xl), ss.nn(2, xl)
    adj = w[75:84] if model == 1 else w[84:93]
    a = {n: float(adj[k]) for n, k in names.items()}
    if swap:
        a[swap[0]], a[swap[1]] = a[swap[1]], a[swap[0]]
    mn, smin, smax = d["CUSS_SS18_MINMU"], d["CUSS_SS18_MINSCALE"], d["CUSS_SS18_MAXSCALE"]
    emu, lmu = max(emu, mn), max(lmu, mn)
    lsc = min(max(a["ss18scaleR_F"] * (math.exp(max(lsc, smin)) - 1.0), smin), smax)
    scale_e = min(max(a["ss18scaleE_S"] * emu + a["ss18scaleE_I"], smin), smax)
    emu = a["ss18muE_S"] * emu + a["ss18muE_I"]
    lmu = a["ss18muR_S"] * lmu + a["ss18muR_I"]
    alpha, ibeta = a["ss18alpha"], a["ss18ibeta"]
    if lam > 0:
        r = min(10.0, 1.0 / lam)
        alpha -= alpha * float(np.float32(0.1)) * r
        ibeta += (1.0 - ibeta) * float(np.float32(0.1)) * r
    return alpha * emu + (1 - alpha) * lmu, (1 - ibeta) * scale_e + ibeta * lsc


@needs_ref
@pytest.mark.parametrize("model", [1, 2])
def test_mu_scale_from_named_parameters(ss, model):
    w, names, d = _ref_weights(), _ref_hyper_names(), _ref_defines()
    assert len(names) == 9
    rng = np.random.default_rng(11 + model)
    moved = 0.0
    for _ in range(100):
        lq, lt = int(rng.integers(30, 1500)), int(rng.integers(30, 1500))
        qe, te, lam = rng.uniform(1, 14), rng.uniform(1, 14), rng.uniform(-0.1, 2.0)
        got = ss.predict(model, lq, lt, qe, te, lam)
        want = _predict_by_name(ss, w, names, d, model, lq, lt, qe, te, lam)
        np.testing.assert_allclose(got, want, rtol=1e-12)
        bad = _predict_by_name(ss, w, names, d, model, lq, lt, qe, te, lam, swap=("ss18muE_S", "ss18muE_I"))
        moved = max(moved, abs(bad[0] - want[0]))
    assert moved > 1.0


@needs_ref
@pytest.mark.parametrize("model", [1, 2])
def test_score_pvalue_is_gumbel_tail(ss, model):
    d = _ref_defines()
    kref = float(np.float32(0.115812))
    rng = np.random.default_rng(21 + model)
    n_exact = n_approx = 0
    for _ in range(400):
        mu, scale = rng.uniform(3, 40), rng.uniform(0.1, 5)
        K = rng.choice([0.05, 0.3, 0.9])
        loc = mu + (scale * math.log(K / kref) if K > kref else 0.0)
        if model == 2:
            loc += scale * d["CUSS_SS18_LOGINVNPAIRS"]
        score = loc + scale * rng.uniform(-3, 20)
        got = ss.L.orc_ss18_score_logp(model, score, K, mu, scale)
        logexp = (loc - score) / scale  # log of the expected count exp(-(s - loc)/scale)
        if logexp >= d["CUSS_SS18_LOG001"]:
            want = stats.gumbel_r.logsf(score, loc=loc, scale=scale)
            assert abs(got - want) <= 1e-9 * max(1.0, abs(want)), (mu, scale, score, K)
            n_exact += 1
        else:
            # below an expected count of 0.01 the count stands for the p-value
            assert abs(got - logexp) <= 1e-9 * abs(logexp)
            assert abs(got - stats.gumbel_r.logsf(score, loc=loc, scale=scale)) < 0.006
            n_approx += 1
        if K > kref:  # a dropped K term moves the result
            assert abs(ss.L.orc_ss18_score_logp(model, score, K, mu, scale) -
                       ss.L.orc_ss18_score_logp(model, score, kref * 0.5, mu, scale)) > 1e-6 or logexp > 2
    assert n_exact > 50 and n_approx > 50


@needs_ref
def test_positive_count_pvalue(ss):
    d = _ref_defines()
    a_shape, x = d["CUSS_SS18_PSTS_NBD_SHAPE_EXP"], d["CUSS_SS18_PSTS_NBD_PROB_EXP"]
    # normalised count 0: I_x(1, b) = 1 - (1 - x)^b
    got = ss.L.orc_ss18_psts_logp(1000, 1000, 1.0)  # 1e5 / (1e6)^1.5 -> 0
    assert abs(got - math.log(1.0 - (1.0 - x) ** a_shape)) < 1e-12
    # n = rint(1e5 psts / max(Lq Lt, 625)^1.5), rounding half to even
    for lq, lt, psts, n in [(100, 100, 25.0, 2), (100, 100, 35.0, 4), (100, 100, 7.0, 1), (10, 10, 3.0, 19),
                            (20, 25, 60.0, 384), (300, 200, 80.0, 1), (300, 200, 20.0, 0)]:
        got = ss.L.orc_ss18_psts_logp(lq, lt, psts)
        want = math.log(special.betainc(n + 1.0, a_shape, x))
        assert abs(got - want) <= 1e-10 * max(1.0, abs(want)), (lq, lt, psts, n)
        # the argument order matters: I_x(b, n+1) is another value
        assert abs(got - math.log(special.betainc(a_shape, n + 1.0, x))) > 1e-3


@needs_ref
def test_brown_combination_is_upper_gamma(ss):
    d = _ref_defines()
    k, c = d["CUSS_SS18_EBM_DFo2"], d["CUSS_SS18_EBM_C"]
    rng = np.random.default_rng(31)
    for _ in range(300):
        lp1, lp2 = -rng.exponential(8.0), -rng.exponential(3.0)
        got = ss.L.orc_ss18_combine(lp1, lp2)
        want = math.log(special.gammaincc(k, (-lp1 - lp2) / c))
        assert abs(got - want) <= 1e-10 * max(1.0, abs(want)), (lp1, lp2)
    # independent p-values of 1 combine to 1
    assert ss.L.orc_ss18_combine(0.0, 0.0) == 0.0


@needs_ref
def test_pvalue_to_evalue(ss):
    d = _ref_defines()
    for logp in [-30.0, -8.0, -4.7, -4.6, -2.0, -0.5, -1e-3]:
        for lq, lt, sspace in [(250, 220, 3.1e9), (40, 1500, 5.5e8)]:
            l1l2 = float(np.float32(lq * lt))
            got = ss.L.orc_ss18_logevalue(logp, sspace, lq, lt)
            if logp >= d["CUSS_SS18_LOG001"]:
                want = math.log(-math.log1p(-math.exp(logp))) + math.log(sspace / l1l2)
            else:
                want = logp + math.log(sspace / l1l2)
            assert abs(got - want) <= 1e-10 * max(1.0, abs(want)), (logp, lq, lt)
    assert ss.L.orc_ss18_logevalue(0.0, 1e6, 10, 10) == pytest.approx(77.0 + math.log(1e6 / 100))


@needs_ref
@pytest.mark.parametrize("model", [1, 2])
def test_composition_with_library_routines(oracle, ss, model):
    """The whole E-value of orc_significance equals the composition of the
    steps above with scipy's routines in place of the oracle's."""
    d = _ref_defines()
    w, names = _ref_weights(), _ref_hyper_names()
    rng = np.random.default_rng(41 + model)
    kref = float(np.float32(0.115812))
    for _ in range(200):
        lq, lt = int(rng.integers(30, 1200)), int(rng.integers(30, 1200))
        qe, te = rng.uniform(1, 14), rng.uniform(1, 14)
        lam, K = rng.uniform(0.05, 1.5), rng.uniform(0.01, 0.5)
        score, psts = rng.uniform(5, 200), float(rng.integers(0, 120))
        sspace = rng.uniform(1e6, 1e10)
        out = oracle.significance(model, lq, lt, qe, te, sspace, score, psts, lam, K)
        mu, scale = _predict_by_name(ss, w, names, d, model, lq, lt, qe, te, lam)
        loc = mu + (scale * math.log(K / kref) if K > kref else 0.0)
        if model == 2:
            loc += scale * d["CUSS_SS18_LOGINVNPAIRS"]
        logexp = (loc - score) / scale
        lps = stats.gumbel_r.logsf(score, loc=loc, scale=scale) if logexp >= d["CUSS_SS18_LOG001"] else logexp
        l1l2 = float(np.float32(lq * lt))
        n = np.rint(d["CUSS_SS18_PSTS_FACTOR"] * psts / max(l1l2, 625.0) ** d["CUSS_SS18_PSTS_EXP"])
        lpp = math.log(special.betainc(n + 1.0, d["CUSS_SS18_PSTS_NBD_SHAPE_EXP"], d["CUSS_SS18_PSTS_NBD_PROB_EXP"]))
        q = special.gammaincc(d["CUSS_SS18_EBM_DFo2"], (-lps - lpp) / d["CUSS_SS18_EBM_C"])
        lq_ = math.log(q) if q > 0 else -1e300
        if lq_ >= d["CUSS_SS18_LOG001"]:
            le = 77.0 if q >= 1.0 else math.log(-math.log1p(-q))
        else:
            le = lq_
        want = le + math.log(sspace / l1l2)
        if q > 1e-300:
            assert abs(out[2] - want) <= 1e-8 * max(1.0, abs(want)), (lq, lt, score, psts)
