"""GPU phase 1 where the reference's buffer of maxima decides the reported
maximum and its cell: exactly equal maxima and templates under 32 positions.

The reference keeps each 32x32 block thread's running maximum in a
per-column buffer and reduces the buffer's last 64 columns
(cudp/CuBatchDP_init_btck.cu:246-263,436-464; cudp/CuBatchDP_final.cu:
100-142); the oracle replays that flow and is checked against the replay of
the block schedule (tests/test_oracle_ties.py).  Here profiles whose
positions are all copies of one query position (and whose transitions are
equal) give a constant positive score matrix, so many cells carry exactly
equal DP values (along the last row of a template longer than the query,
for one); short templates drop rows as the reference does.  The GPU must report the oracle's maximum,
cell, pass decision, score and begin coordinates, and the same traceback."""
import numpy as np
import pytest

import cother_amd

pytestmark = pytest.mark.gpu


def _uniform(ps, src):
    """every position of every profile of ps a copy of position 0 of src
    (another set or ps itself), and every transition vector the same"""
    for name in ("trg", "ss", "cv"):
        ps.field(name)[:] = src.field(name)[:, :1]
    for name in ("cvprior", "cvnorm2", "aa", "sss"):
        ps.field(name)[:] = src.field(name)[0]
    ps.field("dd")[:] = src.field("dd")[:1]
    ps.field("bkg")[:] = src.field("bkg")[:1]
    t = ps.field("trn")
    t[:] = t[:, :1]


@pytest.mark.parametrize("lq,lts", [(20, (7, 31, 45, 64)), (90, (5, 12, 30, 33, 100)), (150, (16, 31, 70, 140))])
def test_equal_maxima_and_short_templates(ctx, oracle, lq, lts):
    q = cother_amd.synth(1, 2, lmin=lq, lmax=lq, dist_mode=0, sort_desc=0, seed=700 + lq)
    db = cother_amd.concat([cother_amd.synth(1, 2, lmin=lt, lmax=lt, dist_mode=1, sort_desc=0, seed=800 + lt)
                            for lt in lts])
    _uniform(q, q)
    _uniform(db, q)
    ctx.upload_queries(q)
    ctx.upload_db(db)
    res, bt = ctx.phase1_block(0, 0, db.nprof, 0.0, want_btck=True)
    sc, _ = oracle.score_matrix(q, 0, db, 0, db.nprof)
    ores, obt = oracle.phase1(q, 0, db, 0, db.nprof, sc, 0.0)
    assert (bt == obt).all()
    for f in ("maxsc", "maxxy", "passed", "psts", "beg"):
        assert (res[f] == ores[f]).all(), (f, res[f], ores[f])
    assert (res["score"].view(np.int32) == ores["score"].view(np.int32)).all()
    cells = [(int(v) >> 16, int(v) & 0xFFFF) for v in res["maxxy"]]
    print(f"lq {lq}: templates {list(lts)}, maxima {res['maxsc'].round(3).tolist()} at (x, y) {cells}", flush=True)


@pytest.mark.parametrize("seed", [41, 42])
def test_random_short_templates(ctx, oracle, seed):
    """random profiles: a query of 40-150 positions against templates of
    6-40 positions (rows dropped by the buffer whenever the template is
    shorter than 32)"""
    q = cother_amd.synth(2, 0, lmin=40, lmax=150, dist_mode=0, sort_desc=0, seed=seed)
    db = cother_amd.synth(60, 0, lmin=6, lmax=40, dist_mode=1, sort_desc=1, seed=seed + 50)
    ctx.upload_queries(q)
    ctx.upload_db(db)
    for qi in range(q.nprof):
        res, bt = ctx.phase1_block(qi, 0, db.nprof, 0.0, want_btck=True)
        sc, _ = oracle.score_matrix(q, qi, db, 0, db.nprof)
        ores, obt = oracle.phase1(q, qi, db, 0, db.nprof, sc, 0.0)
        assert (bt == obt).all()
        for f in ("maxsc", "maxxy", "passed", "psts", "beg"):
            assert (res[f] == ores[f]).all(), (qi, f)


def test_search_realigns_short_templates(ctx, oracle):
    """whole searches at a permissive E-value, so that MAP realigns many hits
    against templates of 6-40 positions (the MAP alignment DP's buffer drops
    rows there too): hit lists and MAP coordinates as the oracle's"""
    q = cother_amd.synth(2, 0, lmin=40, lmax=120, dist_mode=0, sort_desc=0, seed=61)
    db = cother_amd.concat([cother_amd.subset(q, 0, 1),
                            cother_amd.synth(200, 0, lmin=6, lmax=40, dist_mode=1, sort_desc=1, seed=62)])
    ctx.upload_queries(q)
    ctx.upload_db(db)
    hits, nh, _ = ctx.search(eval=1e4)
    off = np.concatenate([[0], np.cumsum(nh)])
    realigned = 0
    for qi in range(q.nprof):
        want = oracle.search_query(q, qi, db, eval_=1e4)
        got = hits[off[qi]:off[qi + 1]]
        assert len(got) == len(want)
        for f in ("profile", "score", "beg", "end", "mapbeg", "mapend", "alnlen", "idnts", "ngaps"):
            assert (got[f] == want[f]).all(), (qi, f)
        realigned += int((got["alnlen"] > 0).sum())
    print(f"{len(hits)} hits, {realigned} realigned", flush=True)
    assert realigned > 20
