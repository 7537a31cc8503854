"""cth_search_batch (many queries per launch, device-side selection of each
stage's pairs, graph capture and replay, autotuned persistent grids)
against cth_search: every hit field, the hit order and the alignment rows
identical; also against the oracle directly, across several batches of one
call, without MAP, for each significance model, for templates without
distance lists, over graph replays with other searches in between, and for
databases the batched form hands to cth_search (a profile over the LDS row
limit)."""
import numpy as np
import pytest

import cother_amd

pytestmark = pytest.mark.gpu

FIELDS = ("query", "profile", "score", "psts", "beg", "end", "mapbeg", "mapend", "idnts", "ngaps", "alnlen",
          "pstvs_map", "mapscore", "logE", "logEpA", "lambda_est", "K_est", "minsc", "maxsc", "expected", "H",
          "lambda_", "K")


def same(a, b):
    assert len(a) == len(b)
    for f in FIELDS:
        assert (a[f].view(np.uint32) == b[f].view(np.uint32)).all(), f


@pytest.fixture(scope="module")
def sets():
    q = cother_amd.synth(6, 0, lmin=40, lmax=150, dist_mode=0, sort_desc=0, seed=401)
    db = cother_amd.concat([cother_amd.subset_list(q, [0, 3]),
                            cother_amd.synth(500, 0, lmin=30, lmax=400, dist_mode=1, sort_desc=1, seed=402)])
    return q, db


def test_batch_equals_per_query_search(ctx, oracle, sets):
    q, db = sets
    ctx.upload_queries(q)
    ctx.upload_db(db)
    a, na, sa, ra = ctx.search(eval=1e3, alignments=True)
    b, nb, sb, rb = ctx.search(eval=1e3, alignments=True, batch=True)
    print(f"{len(a)} hits per-query, {len(b)} batched; batches {sb.nchunks:.0f}", flush=True)
    assert len(a) > 50 and (na == nb).all()
    same(a, b)
    assert ra == rb
    # and against the oracle for two of the queries
    off = np.concatenate([[0], np.cumsum(nb)])
    for qi in (0, 5):
        want = oracle.search_query(q, qi, db, eval_=1e3)
        got = b[off[qi]:off[qi + 1]]
        assert len(got) == len(want)
        for f in ("profile", "score", "beg", "end", "mapbeg", "mapend", "alnlen", "idnts"):
            assert (got[f] == want[f]).all(), f


def test_graph_replays_with_other_work_between(ctx, sets):
    q, db = sets
    ctx.upload_queries(q)
    ctx.upload_db(db)
    c0, r0 = ctx.graph_counts()
    first, n1, _ = ctx.search(eval=1e3, batch=True)
    for k in range(3):
        # eager searches on the same stream between the replays
        ctx.search(1, 3, eval=10.0)
        again, n2, _ = ctx.search(eval=1e3, batch=True)
        same(first, again)
    c1, r1 = ctx.graph_counts()
    print(f"captures {c1 - c0}, replays {r1 - r0}", flush=True)
    assert c1 - c0 <= 1 and r1 - r0 >= 3


def test_several_batches_and_a_query_over_the_budget(ctx, sets):
    q, db = sets
    ctx.upload_queries(q)
    ctx.upload_db(db)
    a, na, _ = ctx.search(eval=1e3)
    # a budget that holds two queries' buffers: three or more batches
    ld = (db.npos + 63) // 64 * 64
    budget = int(2 * 150 * ld * 13)
    b, nb, sb = ctx.search(eval=1e3, batch=True, mem_budget=budget)
    assert sb.nchunks >= 3
    same(a, b)
    # a budget below one query's buffers: each query runs in chunks through cth_search
    c, nc, sc = ctx.search(eval=1e3, batch=True, mem_budget=4096)
    same(a, c)


@pytest.mark.parametrize("ssemodel,do_map", [(0, True), (1, True), (2, False)])
def test_batch_models_and_without_map(ctx, sets, ssemodel, do_map):
    q, db = sets
    ctx.upload_queries(q)
    ctx.upload_db(db)
    a, na, _ = ctx.search(eval=1e3, ssemodel=ssemodel, do_map=do_map)
    b, nb, _ = ctx.search(eval=1e3, ssemodel=ssemodel, do_map=do_map, batch=True)
    same(a, b)


def test_templates_without_distances(ctx):
    q = cother_amd.synth(8, 0, lmin=40, lmax=150, dist_mode=0, sort_desc=0, seed=403)
    db = cother_amd.concat([cother_amd.subset(q, 0, 1),
                            cother_amd.synth(800, 0, lmin=30, lmax=400, dist_mode=2, sort_desc=1, seed=404)])
    ctx.upload_queries(q)
    ctx.upload_db(db)
    a, na, _ = ctx.search(eval=1e3)
    b, nb, _ = ctx.search(eval=1e3, batch=True)
    assert len(a) > 0
    same(a, b)


def test_long_profile_database_runs_per_query(ctx):
    q = cother_amd.synth(2, 0, lmin=60, lmax=120, dist_mode=0, sort_desc=0, seed=405)
    db = cother_amd.concat([cother_amd.synth(1, 2, lmin=8500, lmax=8500, dist_mode=2, sort_desc=0, seed=406),
                            cother_amd.synth(50, 0, lmin=30, lmax=300, dist_mode=1, sort_desc=1, seed=407)])
    ctx.upload_queries(q)
    ctx.upload_db(db)
    a, na, _ = ctx.search(eval=1e3)
    b, nb, _ = ctx.search(eval=1e3, batch=True)
    same(a, b)
