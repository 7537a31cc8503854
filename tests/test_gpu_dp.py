"""GPU phase 1 (DP with traceback + finalisation) against the oracle."""
import numpy as np
import pytest

import cother_amd

pytestmark = pytest.mark.gpu


def test_dpp_lane_layout(ctx):
    out = np.zeros(64, np.int32)
    assert cother_amd.lib().cth_selftest_dpp(out.ctypes.data) == 0
    assert out[0] == -1 and (out[1:] == np.arange(63)).all()


@pytest.mark.parametrize("seed,nq,lmin,lmax,thld", [(31, 2, 40, 150, 0.0), (32, 1, 130, 300, 3.0),
                                                     (33, 2, 20, 70, 1.0)])
def test_phase1_bit_exact(ctx, oracle, seed, nq, lmin, lmax, thld):
    q = cother_amd.synth(nq, 0, lmin=lmin, lmax=lmax, dist_mode=0, sort_desc=0, seed=seed)
    db = cother_amd.synth(24, 0, lmin=10, lmax=260, dist_mode=1, sort_desc=1, seed=seed + 100)
    ctx.upload_queries(q)
    ctx.upload_db(db)
    for qi in range(nq):
        res, bt = ctx.phase1_block(qi, 0, db.nprof, thld, want_btck=True)
        sc, _ = oracle.score_matrix(q, qi, db, 0, db.nprof)
        ores, obt = oracle.phase1(q, qi, db, 0, db.nprof, sc, thld)
        bad = np.argwhere(bt != obt)
        assert bad.size == 0, f"q{qi}: {len(bad)} traceback bytes differ, first {bad[:5].tolist()}"
        for f in ("maxsc", "maxxy", "passed", "psts", "beg"):
            assert (res[f] == ores[f]).all(), (f, res[f], ores[f])
        assert (res["score"].view(np.int32) == ores["score"].view(np.int32)).all()
