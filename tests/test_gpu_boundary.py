"""The search entry's inputs that come from the reference's batch interface:
the significance model (SSEMODEL 0, 1, 2) and the whole database's size and
count when the resident set is one part of it (CuBatchProcessing::
SetDbDetails, cupro/CuBatchProcessing.cuh:73)."""
import numpy as np
import pytest

import cother_amd

pytestmark = pytest.mark.gpu

FIELDS = ("profile", "score", "psts", "beg", "end", "mapbeg", "mapend", "idnts", "ngaps", "alnlen", "pstvs_map",
          "mapscore")
STATS = ("logE", "logEpA", "lambda_est", "K_est", "minsc", "maxsc", "expected", "H", "lambda_", "K")


def _same(got, want):
    assert len(got) == len(want)
    for f in FIELDS:
        assert (got[f] == want[f]).all(), f
    for f in STATS:
        np.testing.assert_allclose(got[f], want[f], rtol=2e-6, atol=1e-30, err_msg=f)


@pytest.mark.parametrize("model", [0, 1, 2])
def test_significance_models_match_oracle(ctx, oracle, model):
    q = cother_amd.synth(2, 0, lmin=60, lmax=200, dist_mode=0, sort_desc=0, seed=301)
    db = cother_amd.concat([cother_amd.subset(q, 0, 2),
                            cother_amd.synth(150, 0, lmin=20, lmax=260, dist_mode=1, sort_desc=1, seed=302)])
    ctx.upload_queries(q)
    ctx.upload_db(db)
    hits, nh, _ = ctx.search(eval=1e3, ssemodel=model)
    off = 0
    for qi in range(q.nprof):
        got = hits[off:off + nh[qi]]
        off += nh[qi]
        want = oracle.search_query(q, qi, db, eval_=1e3, ssemodel=model)
        print(f"model {model} query {qi}: {len(got)} hits (oracle {len(want)})", flush=True)
        assert len(want) > 0
        _same(got, want)
        if model == 0:
            # observed statistics: log E = log(K_g * searchspace) - lambda_g * score (CalculateEvalueObs)
            ss, _ = cother_amd.search_space(int(q.lens()[qi]), db.npos, db.nprof)
            lg, kg = np.float64(np.float32(0.561132)), np.float64(np.float32(0.001488))
            np.testing.assert_allclose(got["logE"], np.log(kg * ss) - lg * got["score"].astype(np.float64),
                                       rtol=1e-6, atol=1e-6)


def test_models_differ(ctx):
    """The three models give different E-values for the same hits (a model
    silently replaced by another would not)."""
    q = cother_amd.synth(1, 0, lmin=120, lmax=120, dist_mode=0, sort_desc=0, seed=303)
    db = cother_amd.concat([q, cother_amd.synth(60, 0, lmin=40, lmax=200, dist_mode=1, sort_desc=1, seed=304)])
    ctx.upload_queries(q)
    ctx.upload_db(db)
    logE = {}
    for m in (0, 1, 2):
        h, _, _ = ctx.search(eval=1e6, ssemodel=m)
        h = h[np.argsort(h["profile"], kind="stable")]
        logE[m] = h["logE"]
    assert len(logE[0]) == len(logE[1]) == len(logE[2]) > 0
    assert not np.allclose(logE[0], logE[1]) and not np.allclose(logE[1], logE[2])


def test_database_in_two_parts_with_global_size(oracle):
    """The database split into two resident parts, each searched with the whole
    database's size and count, gives the whole database's hits: merged by
    (log E, global index) the lists are identical, field by field."""
    q = cother_amd.synth(2, 0, lmin=50, lmax=180, dist_mode=0, sort_desc=0, seed=311)
    db = cother_amd.concat([cother_amd.subset(q, 0, 2),
                            cother_amd.synth(300, 0, lmin=20, lmax=300, dist_mode=1, sort_desc=1, seed=312)])
    half = db.nprof // 2
    parts = [cother_amd.subset(db, 0, half), cother_amd.subset(db, half, db.nprof)]
    whole = cother_amd.Context(0)
    whole.upload_queries(q)
    whole.upload_db(db)
    want_h, want_n, _ = whole.search(eval=1e3)
    whole.close()
    got = [[] for _ in range(q.nprof)]
    for k, part in enumerate(parts):
        c = cother_amd.Context(0)
        c.upload_queries(q)
        c.upload_db(part)
        h, n, _ = c.search(eval=1e3, dbsize=db.npos, ndb=db.nprof)
        c.close()
        h = h.copy()
        h["profile"] += half * k
        off = 0
        for qi in range(q.nprof):
            got[qi].append(h[off:off + n[qi]])
            off += n[qi]
    off = 0
    for qi in range(q.nprof):
        m = np.concatenate(got[qi])
        m = m[np.lexsort((m["profile"], m["logE"]))]
        w = want_h[off:off + want_n[qi]]
        off += want_n[qi]
        print(f"query {qi}: {len(m)} hits from the parts, {len(w)} from the whole", flush=True)
        assert len(w) > 0
        _same(m, w)
        # and the whole equals the oracle's
        _same(w, oracle.search_query(q, qi, db, eval_=1e3))
    # without the global size the parts' E-values are those of smaller databases
    c = cother_amd.Context(0)
    c.upload_queries(q)
    c.upload_db(parts[0])
    h, _, _ = c.search(eval=1e3)
    c.close()
    hg = [x for x in got[0][0] if x["profile"] == 0]
    hl = [x for x in h if x["query"] == 0 and x["profile"] == 0]
    assert hg and hl and hl[0]["logE"] < hg[0]["logE"]
