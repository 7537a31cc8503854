"""Edge cases of the whole search on the GPU against the oracle: tiny profiles
(shorter than the DDMS segment, single positions, empty distance lists),
query lengths around the 32-row block and 64-row stripe sizes, and long
profiles (the long-profile workload's shape: Lq = 2000 against templates of
1000-3000 positions, with a self-hit whose MAP alignment spans 32 stripes)."""
import numpy as np
import pytest

import cother_amd

pytestmark = pytest.mark.gpu

FIELDS = ("profile", "score", "psts", "beg", "end", "mapbeg", "mapend", "idnts", "ngaps", "alnlen", "pstvs_map",
          "mapscore")
STATS = ("logE", "logEpA", "lambda_est", "K_est", "minsc", "maxsc", "expected", "H", "lambda_", "K")


def _compare(ctx, oracle, q, db, eval_, **kw):
    ctx.upload_queries(q)
    ctx.upload_db(db)
    hits, nh, _ = ctx.search(eval=eval_, **kw)
    off = 0
    total = 0
    for qi in range(q.nprof):
        got = hits[off:off + nh[qi]]
        off += nh[qi]
        want = oracle.search_query(q, qi, db, eval_=eval_)
        print(f"query {qi} (Lq={q.lens()[qi]}): {len(got)} hits, oracle {len(want)}", flush=True)
        assert len(got) == len(want)
        for f in FIELDS:
            assert (got[f] == want[f]).all(), (qi, f)
        for f in STATS:
            np.testing.assert_allclose(got[f], want[f], rtol=2e-6, atol=1e-30, err_msg=f)
        total += len(got)
    return total


def test_tiny_profiles(ctx, oracle):
    q = cother_amd.synth(3, 0, lmin=1, lmax=9, dist_mode=0, sort_desc=0, seed=81)
    db = cother_amd.synth(40, 0, lmin=1, lmax=8, dist_mode=1, sort_desc=0, seed=82)
    _compare(ctx, oracle, q, db, 1e6)


@pytest.mark.parametrize("lq", [7, 31, 32, 33, 63, 64, 65, 97])
def test_query_lengths_around_block_sizes(ctx, oracle, lq):
    q = cother_amd.synth(1, 2, lmin=lq, lmax=lq, dist_mode=0, sort_desc=0, seed=90 + lq)
    db = cother_amd.synth(30, 0, lmin=5, lmax=140, dist_mode=1, sort_desc=1, seed=91 + lq)
    _compare(ctx, oracle, q, db, 1e4)


def test_long_profiles_with_self_hit(ctx, oracle):
    q = cother_amd.synth(1, 2, lmin=2000, lmax=2000, dist_mode=1, sort_desc=0, seed=95)
    # template 0 is the query itself; two more long random templates
    same = cother_amd.synth(1, 2, lmin=2000, lmax=2000, dist_mode=1, sort_desc=0, seed=95)
    other = cother_amd.synth(2, 0, lmin=1000, lmax=3000, dist_mode=1, sort_desc=0, seed=96)
    db = cother_amd.concat([same, other])
    n = _compare(ctx, oracle, q, db, 10.0)
    assert n >= 1
    hits, _, _ = ctx.search(eval=10.0)
    assert hits[0]["profile"] == 0 and hits[0]["alnlen"] > 1900


def test_empty_database_and_profiles_without_distances(ctx, oracle):
    q = cother_amd.synth(2, 0, lmin=30, lmax=80, dist_mode=1, sort_desc=0, seed=97)
    empty = cother_amd.synth(0, 0, lmin=30, lmax=80, dist_mode=1, sort_desc=0, seed=98)
    ctx.upload_queries(q)
    ctx.upload_db(empty)
    print("edges: empty database", flush=True)
    hits, nh, st = ctx.search(eval=10.0)
    assert len(hits) == 0 and list(nh) == [0, 0] and st.cells == 0
    # templates without distance lists: the DDMS term is skipped (lists shorter than DSTSEGM)
    nodist = cother_amd.synth(5, 0, lmin=30, lmax=200, dist_mode=2, sort_desc=0, seed=99)
    print("edges: templates without distance lists", flush=True)
    assert _compare(ctx, oracle, q, nodist, 100.0) >= 0


def test_distance_lists_over_99_values_refused(ctx):
    """The reference's lists hold at most 99 values (MAX_N_IR_DISTANCE_VALUES,
    libpro/srcpro/MOptions.h:15), and the score kernel's scan relies on it:
    a longer list is refused on upload, for queries and for the database."""
    ps = cother_amd.synth(3, 0, lmin=20, lmax=30, dist_mode=1, sort_desc=0, seed=5)
    dd = ps.field("dd")
    assert dd.shape[1] == 128
    dd[7, 0] = 99
    ctx.upload_db(ps)
    ctx.upload_queries(ps)
    dd[7, 0] = 100
    for up in (ctx.upload_db, ctx.upload_queries):
        with pytest.raises(RuntimeError, match="position 7: distance list of 100 values"):
            up(ps)
