"""The benchmark configurations' generators on the GPU, at sizes the CPU
oracle finishes in seconds:

- batching workload (queries lognormal median 250 clipped 40-1000 with
  predicted distance maps; templates lognormal median 220 clipped 30-1500
  with C-alpha chain distances): two of its queries against 600 of its
  templates, hit lists equal to the oracle's;
- the reduced-precision contraction (BF16 and FP16 operands on the matrix
  cores, FP32 accumulation) on the same generators, at a permissive E-value
  that passes over 100 hits per query: per-cell error of the log-odds term
  against FP64, at least 99% of each top-100 hit list shared with the FP32
  path, scores of shared hits within a few unit roundoffs of the operand
  format;
- the long-profile workload's BF16 variant: a 2000-position query with a
  full predicted distance map against 1000-3000-position templates and
  itself: per-cell error and the same hits as FP32.
The FP32 path is the parity path; the contraction has no reference
counterpart and is held to the tolerances written here."""
import numpy as np
import pytest

import cother_amd
from tests.test_gpu_contraction import TOL, _init64

pytestmark = pytest.mark.gpu

FIELDS = ("profile", "score", "psts", "beg", "end", "mapbeg", "mapend", "idnts", "ngaps", "alnlen", "pstvs_map",
          "mapscore")
STATS = ("logE", "logEpA", "lambda_est", "K_est", "minsc", "maxsc", "expected", "H", "lambda_", "K")


def batch_queries(n, seed):
    return cother_amd.synth(n, 1, 250.0, 0.35, 40, 1000, dist_mode=0, sort_desc=0, seed=seed)


def batch_templates(n, seed):
    return cother_amd.synth(n, 1, 220.0, 0.35, 30, 1500, dist_mode=1, sort_desc=1, seed=seed)


@pytest.mark.timeout(600)
def test_batching_generators_match_oracle(ctx, oracle):
    q = batch_queries(2, 20260)
    db = cother_amd.concat([cother_amd.subset(q, 0, 1), batch_templates(600, 20261)])
    ctx.upload_queries(q)
    ctx.upload_db(db)
    print(f"queries {list(q.lens())} vs {db.nprof} templates ({db.npos} positions)", flush=True)
    hits, nh, st = ctx.search(eval=10.0)
    off = 0
    for qi in range(q.nprof):
        got = hits[off:off + nh[qi]]
        off += nh[qi]
        want = oracle.search_query(q, qi, db, eval_=10.0)
        print(f"query {qi}: {len(got)} hits (oracle {len(want)})", flush=True)
        assert len(got) == len(want)
        for f in FIELDS:
            assert (got[f] == want[f]).all(), f
        for f in STATS:
            np.testing.assert_allclose(got[f], want[f], rtol=2e-6, atol=1e-30, err_msg=f)
    assert nh[0] >= 1 and hits[0]["profile"] == 0


@pytest.fixture(scope="module")
def big_sets():
    q = batch_queries(4, 20262)
    db = cother_amd.concat([cother_amd.subset_list(q, [0, 1]), batch_templates(3000, 20263)])
    return q, db


@pytest.mark.parametrize("prec", [1, 2])
def test_contraction_error_on_batching_generators(ctx, big_sets, prec):
    q, db = big_sets
    ctx.upload_queries(q)
    ctx.upload_db(db)
    worst = 0.0
    for qi in range(q.nprof):
        s32, _ = ctx.score_block(qi, 0, 300)
        sv, _ = ctx.score_block(qi, 0, 300, contraction=prec)
        init64 = _init64(q, qi, db, 0, 300)
        other = s32.astype(np.float64) - init64
        err = np.abs(sv.astype(np.float64) - (init64 + other))
        print(f"prec {prec} query {qi} (Lq {q.lens()[qi]}): per-cell error max {err.max():.2e} mean {err.mean():.2e}",
              flush=True)
        worst = max(worst, err.max())
    assert worst < TOL[prec]


@pytest.mark.parametrize("prec", [1, 2])
def test_contraction_top100_on_batching_generators(ctx, big_sets, prec):
    q, db = big_sets
    ctx.upload_queries(q)
    ctx.upload_db(db)
    eval_ = 1e6
    h32, n32, _ = ctx.search(eval=eval_)
    hv, nv, _ = ctx.search(eval=eval_, contraction=prec)
    o32 = np.concatenate([[0], np.cumsum(n32)])
    ov = np.concatenate([[0], np.cumsum(nv)])
    for qi in range(q.nprof):
        a = h32[o32[qi]:o32[qi + 1]]
        b = hv[ov[qi]:ov[qi + 1]]
        assert len(a) >= 100 and len(b) >= 100, (qi, len(a), len(b))
        top_a, top_b = set(a["profile"][:100]), set(b["profile"][:100])
        shared = len(top_a & top_b)
        def logE_diff(profiles):
            common = [p for p in profiles if p in set(b["profile"])]
            ea = np.array([a["logE"][a["profile"] == p][0] for p in common], np.float64)
            eb = np.array([b["logE"][b["profile"] == p][0] for p in common], np.float64)
            return np.abs(ea - eb) / np.maximum(1.0, np.abs(ea))
        top = logE_diff(a["profile"][:100])
        every = logE_diff(a["profile"])
        print(f"prec {prec} query {qi}: {len(a)} hits, top-100 shared {shared}, log E rel. difference "
              f"top-100 max {top.max():.2e} mean {top.mean():.2e} p95 {np.percentile(top, 95):.2e}; "
              f"all hits max {every.max():.2e} mean {every.mean():.2e}", flush=True)
        # the top-100 hit that moved most, in both forms
        common = [p for p in a["profile"][:100] if p in set(b["profile"])]
        pw = common[int(np.argmax(top))]
        for tag, h in (("fp32", a), ("variant", b)):
            r = h[h["profile"] == pw][0]
            print(f"  worst profile {pw} {tag}: score {r['score']:.4f} beg {r['beg']} end {r['end']} "
                  f"logE {r['logE']:.4f} lambda_est {r['lambda_est']:.5f} K_est {r['K_est']:.5f} "
                  f"expected {r['expected']:.5f} minsc {r['minsc']:.3f} maxsc {r['maxsc']:.3f}", flush=True)
        assert shared >= 99, (qi, sorted(top_a ^ top_b))
        # The contraction perturbs the alignment score; the E-value follows
        # from it (E = K m n exp(-lambda S)), so a top-100 hit near E = 1
        # shows a large relative log E change for a small score change.  The
        # shared top-100 hits' scores: within 8 unit roundoffs of the operand
        # format relative (BF16 2^-8, FP16 2^-11; the per-cell error is about
        # 2 of them), on average within one.
        u = 2.0 ** -8 if prec == 1 else 2.0 ** -11
        common = [p for p in a["profile"][:100] if p in set(b["profile"])]
        sa = np.array([a["score"][a["profile"] == p][0] for p in common], np.float64)
        sb = np.array([b["score"][b["profile"] == p][0] for p in common], np.float64)
        rs = np.abs(sb - sa) / np.abs(sa)
        print(f"  score rel. difference top-100 max {rs.max():.2e} mean {rs.mean():.2e} ({rs.max() / u:.2f} u, "
              f"{rs.mean() / u:.2f} u); variant higher in {int((sb > sa).sum())} of {len(sa)}", flush=True)
        assert rs.max() <= 8 * u and rs.mean() <= u
    # the queries' copies are their first hits in both forms
    assert h32[0]["profile"] == 0 and hv[0]["profile"] == 0


@pytest.mark.timeout(600)
def test_long_profile_bf16_variant(ctx):
    q = cother_amd.synth(1, 2, lmin=2000, lmax=2000, dist_mode=0, sort_desc=0, seed=20264)
    db = cother_amd.concat([cother_amd.synth(6, 0, lmin=1000, lmax=3000, dist_mode=1, sort_desc=1, seed=20265),
                            cother_amd.subset(q, 0, 1)])
    ctx.upload_queries(q)
    ctx.upload_db(db)
    s32, _ = ctx.score_block(0, 0, 2)
    sv, _ = ctx.score_block(0, 0, 2, contraction=1)
    init64 = _init64(q, 0, db, 0, 2)
    err = np.abs(sv.astype(np.float64) - (init64 + (s32.astype(np.float64) - init64)))
    print(f"long profile bf16: per-cell error max {err.max():.2e} mean {err.mean():.2e}", flush=True)
    assert err.max() < TOL[1]
    h32, _, _ = ctx.search(eval=10.0)
    hv, _, _ = ctx.search(eval=10.0, contraction=1)
    print(f"fp32 hits {list(h32['profile'])}, bf16 hits {list(hv['profile'])}", flush=True)
    assert h32[0]["profile"] == hv[0]["profile"] == db.nprof - 1
    assert abs(float(h32[0]["logE"]) - float(hv[0]["logE"])) <= 1e-3 * abs(float(h32[0]["logE"]))
