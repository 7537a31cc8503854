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
