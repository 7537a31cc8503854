"""The oracle's whole-query composition (orc_search_query) against its stages.

The stages are pinned by their own tests; this checks the glue: the score
threshold, the search space from the length adjustment, the E-value filter and
the hit ordering, on a case small enough for the stage-by-stage path.
"""
import numpy as np

import cother_amd


def test_search_composition_matches_stages(oracle):
    q = cother_amd.synth(1, 0, lmin=50, lmax=90, dist_mode=0, sort_desc=0, seed=7)
    db = cother_amd.synth(30, 0, lmin=20, lmax=120, dist_mode=1, sort_desc=1, seed=8)
    eval_ = 1e4  # permissive: keeps many hits so the ordering is exercised
    hits = oracle.search_query(q, 0, db, eval_=eval_, do_map=False)
    lq = int(q.lens()[0])
    thld, logethld = oracle.score_threshold(eval_, lq, db.npos)
    sspace, _ = oracle.length_adjustment(lq, db.npos, db.nprof)
    sc, _ = oracle.score_matrix(q, 0, db, 0, db.nprof)
    res, _ = oracle.phase1(q, 0, db, 0, db.nprof, sc, thld)
    d0, ln = db.dists(), db.lens()
    qaa = q.field("aa")[q.dists()[0]:]
    want = []
    for k in np.flatnonzero(res["passed"]):
        lt = int(ln[k])
        sub = np.ascontiguousarray(sc[:, d0[k]:d0[k] + lt])
        counts = np.zeros(641)
        oracle.L.orc_histogram(sub.ctypes.data, lt, 0, lq, lt, qaa.ctypes.data,
                               db.field("aa")[d0[k]:d0[k] + lt].ctypes.data, counts.ctypes.data)
        par = oracle.stat_params(counts)
        sig = oracle.significance(2, lq, lt, float(q.field("eno")[0]), float(db.field("eno")[k]), float(sspace),
                                  float(res["score"][k]), float(res["psts"][k]), par[4], par[5])
        loge = np.float32(sig[2])
        if not logethld < loge:
            want.append((loge, int(k), res["score"][k]))
    want.sort(key=lambda w: (w[0], w[1]))
    assert len(hits) == len(want) and len(want) > 3
    assert [int(h) for h in hits["profile"]] == [w[1] for w in want]
    assert (hits["logE"] == np.array([w[0] for w in want], np.float32)).all()
    assert (hits["score"] == np.array([w[2] for w in want], np.float32)).all()
    assert (np.diff(hits["logE"]) >= 0).all()


def test_search_strict_threshold_filters(oracle):
    q = cother_amd.synth(1, 0, lmin=50, lmax=90, dist_mode=0, sort_desc=0, seed=7)
    db = cother_amd.synth(30, 0, lmin=20, lmax=120, dist_mode=1, sort_desc=1, seed=8)
    loose = oracle.search_query(q, 0, db, eval_=1e4, do_map=False)
    strict = oracle.search_query(q, 0, db, eval_=1e-3, do_map=False)
    assert len(strict) <= len(loose)
    assert set(strict["profile"]) <= set(loose["profile"])
    assert (strict["logE"] <= np.float32(np.log(np.float32(1e-3)))).all()


def test_search_with_map_keeps_aligned_hits(oracle):
    q = cother_amd.synth(1, 0, lmin=50, lmax=90, dist_mode=0, sort_desc=0, seed=7)
    db = cother_amd.synth(30, 0, lmin=20, lmax=120, dist_mode=1, sort_desc=1, seed=8)
    plain = oracle.search_query(q, 0, db, eval_=1e4, do_map=False)
    mapped = oracle.search_query(q, 0, db, eval_=1e4, do_map=True)
    # MAP drops hits without an alignment and changes nothing else
    assert set(mapped["profile"]) <= set(plain["profile"])
    assert (mapped["alnlen"] >= 1).all()
    byp = {int(h["profile"]): h for h in plain}
    for h in mapped:
        w = byp[int(h["profile"])]
        assert h["logE"] == w["logE"] and h["score"] == w["score"]
        assert h["idnts"] + h["pstvs_map"] + h["ngaps"] <= h["alnlen"]
