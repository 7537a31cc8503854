"""cth_search (the whole per-query pipeline over the resident database) against
the oracle's composition of the same stages."""
import numpy as np
import pytest

import cother_amd

pytestmark = pytest.mark.gpu

FIELDS = ("profile", "score", "psts", "beg", "end", "mapbeg", "mapend", "idnts", "ngaps", "alnlen", "pstvs_map",
          "mapscore")
STATS = ("logE", "logEpA", "lambda_est", "K_est", "minsc", "maxsc", "expected", "H", "lambda_", "K")


@pytest.mark.parametrize("budget", [0, 1])  # 1 byte: the minimum chunk, several chunks per query
def test_search_matches_oracle(ctx, oracle, budget):
    q = cother_amd.synth(3, 0, lmin=40, lmax=220, dist_mode=0, sort_desc=0, seed=51)
    db = cother_amd.synth(120, 0, lmin=15, lmax=300, dist_mode=1, sort_desc=1, seed=52)
    ctx.upload_queries(q)
    ctx.upload_db(db)
    eval_ = 1e3
    hits, nh, st = ctx.search(eval=eval_, mem_budget=budget, timing=True)
    assert st.cells == float(q.lens().sum()) * db.npos
    if budget:
        assert st.nchunks > q.nprof
    off = 0
    for qi in range(q.nprof):
        got = hits[off:off + nh[qi]]
        off += nh[qi]
        want = oracle.search_query(q, qi, db, eval_=eval_)
        print(f"query {qi}: {len(got)} hits (oracle {len(want)})", flush=True)
        assert len(got) == len(want) and len(want) > 0
        assert (got["query"] == qi).all()
        for f in FIELDS:
            assert (got[f] == want[f]).all(), f
        for f in STATS:
            np.testing.assert_allclose(got[f], want[f], rtol=2e-6, atol=1e-30, err_msg=f)


@pytest.mark.timeout(400)
def test_search_at_the_cpu_runnable_size(ctx, oracle):
    """One query of 150 positions (predicted distance map) against 2000
    templates of 50-400 positions (C-alpha chains), the workload the CPU path
    can run whole (~7e7 cells), plus the query itself as template 0 so a long
    alignment is among the hits: the hit list equals the oracle's."""
    q = cother_amd.synth(1, 2, lmin=150, lmax=150, dist_mode=0, sort_desc=0, seed=57)
    rest = cother_amd.synth(2000, 0, lmin=50, lmax=400, dist_mode=1, sort_desc=1, seed=58)
    db = cother_amd.concat([cother_amd.subset(q, 0, 1), rest])
    ctx.upload_queries(q)
    ctx.upload_db(db)
    print(f"search: Lq=150 vs {db.nprof} templates ({db.npos} positions)", flush=True)
    hits, nh, st = ctx.search(eval=10.0)
    want = oracle.search_query(q, 0, db, eval_=10.0)
    print(f"{len(hits)} hits (oracle {len(want)})", flush=True)
    assert len(hits) == len(want) >= 1
    assert hits[0]["profile"] == 0 and hits[0]["alnlen"] >= 140
    for f in FIELDS:
        assert (hits[f] == want[f]).all(), f
    for f in STATS:
        np.testing.assert_allclose(hits[f], want[f], rtol=2e-6, atol=1e-30, err_msg=f)
    # the score matrix over the first 60 templates, bit for bit
    sc, md = ctx.score_block(0, 0, 60)
    osc, omd = oracle.score_matrix(q, 0, db, 0, 60)
    assert (sc.view(np.int32) == osc.view(np.int32)).all()
    assert (md.view(np.int32) == omd.view(np.int32)).all()


def test_search_over_several_dp_groups(ctx, oracle):
    """More than 2048 templates in one chunk: the first phase-1 pass runs in
    groups of 2048 profiles (each launch sized for its own longest profile,
    records written at the group's offset); 4500 short templates, sorted by
    length as databases are, give three groups.  Hits (many, at a permissive
    E-value) equal the oracle's."""
    q = cother_amd.synth(1, 0, lmin=40, lmax=40, dist_mode=0, sort_desc=0, seed=61)
    db = cother_amd.synth(4500, 0, lmin=8, lmax=70, dist_mode=1, sort_desc=1, seed=62)
    ctx.upload_queries(q)
    ctx.upload_db(db)
    print(f"search: Lq=40 vs {db.nprof} templates", flush=True)
    hits, nh, st = ctx.search(eval=1e5)
    want = oracle.search_query(q, 0, db, eval_=1e5)
    print(f"{len(hits)} hits (oracle {len(want)})", flush=True)
    assert len(hits) == len(want) >= 50
    assert hits["profile"].max() >= 4096  # hits from the third group
    for f in FIELDS:
        assert (hits[f] == want[f]).all(), f
    for f in STATS:
        np.testing.assert_allclose(hits[f], want[f], rtol=2e-6, atol=1e-30, err_msg=f)
