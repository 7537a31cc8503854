"""GPU score-matrix stage against the oracle: bit-exact in FP32."""
import numpy as np
import pytest

import cother_amd

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sets():
    q = cother_amd.synth(3, 0, lmin=60, lmax=150, dist_mode=0, sort_desc=0, seed=101)
    db = cother_amd.synth(40, 0, lmin=20, lmax=400, dist_mode=1, sort_desc=1, seed=102)
    return q, db


def test_score_matrix_bit_exact(ctx, oracle, sets):
    q, db = sets
    ctx.upload_queries(q)
    ctx.upload_db(db)
    for qi in range(q.nprof):
        sc, md = ctx.score_block(qi, 0, db.nprof)
        osc, omd = oracle.score_matrix(q, qi, db, 0, db.nprof)
        assert sc.shape == osc.shape
        bad = np.argwhere(sc.view(np.int32) != osc.view(np.int32))
        assert bad.size == 0, f"query {qi}: {len(bad)} cells differ, first {bad[:5].tolist()}"
        assert (md.view(np.int32) == omd.view(np.int32)).all()


def test_score_matrix_subrange(ctx, oracle, sets):
    q, db = sets
    ctx.upload_queries(q)
    ctx.upload_db(db)
    sc, md = ctx.score_block(1, 7, 19)
    osc, omd = oracle.score_matrix(q, 1, db, 7, 19)
    assert (sc.view(np.int32) == osc.view(np.int32)).all()
    assert (md.view(np.int32) == omd.view(np.int32)).all()
