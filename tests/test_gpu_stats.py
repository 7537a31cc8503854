"""GPU statistics of passed profiles against the oracle (double precision)."""
import numpy as np
import pytest

import cother_amd

pytestmark = pytest.mark.gpu


def test_stats_match_oracle(ctx, oracle):
    q = cother_amd.synth(2, 0, lmin=60, lmax=200, dist_mode=0, sort_desc=0, seed=41)
    db = cother_amd.synth(40, 0, lmin=20, lmax=300, dist_mode=1, sort_desc=1, seed=42)
    ctx.upload_queries(q)
    ctx.upload_db(db)
    sspace = 3.0e9
    for qi in range(q.nprof):
        res, plist, st = ctx.phase2_block(qi, 0, db.nprof, 0.0, sspace, 2)
        assert len(plist) > 3
        sc, _ = oracle.score_matrix(q, qi, db, 0, db.nprof)
        d0 = db.dists()
        qaa = q.field("aa")[q.dists()[qi]:]
        for j, k in enumerate(plist):
            lt = int(db.lens()[k])
            counts = np.zeros(641)
            taa = db.field("aa")[d0[k]:d0[k] + lt]
            sub = np.ascontiguousarray(sc[:, d0[k]:d0[k] + lt])
            oracle.L.orc_histogram(sub.ctypes.data, lt, 0, sub.shape[0], lt, qaa.ctypes.data, taa.ctypes.data,
                                   counts.ctypes.data)
            par = oracle.stat_params(counts)  # min max E H lambda K
            sig = oracle.significance(2, int(q.lens()[qi]), lt, float(q.field("eno")[qi]), float(db.field("eno")[k]),
                                      sspace, float(res["score"][k]), float(res["psts"][k]), par[4], par[5])
            want = np.concatenate([sig, par])
            np.testing.assert_allclose(st[j], want, rtol=1e-9, atol=1e-12)
