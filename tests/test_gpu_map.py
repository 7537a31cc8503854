"""GPU MAP stage (forward, backward, partition function, MAP DP, traceback)
against the oracle, bit for bit on every output field."""
import numpy as np
import pytest

import cother_amd

pytestmark = pytest.mark.gpu

FIELDS = ("maxsc", "end", "beg", "idnts", "psts", "ngaps", "alnlen", "logz", "maxlog")


def test_dpp_shl_lane_layout(ctx):
    out = np.zeros(64, np.int32)
    assert cother_amd.lib().cth_selftest_dpp_shl(out.ctypes.data) == 0
    assert out[63] == -1 and (out[:63] == np.arange(1, 64)).all()


def _check(ctx, oracle, q, db, qi, p0, p1, profiles, logepa):
    res = ctx.map_block(qi, p0, p1, profiles, logepa)
    sc, md = oracle.score_matrix(q, qi, db, p0, p1)
    d = db.dists()
    for j, p in enumerate(profiles):
        want = oracle.map_profile(q, qi, db, int(p), sc, md, int(d[p] - d[p0]), float(logepa[j]))
        for f in FIELDS:
            g, w = res[j][f], want[f]
            assert np.asarray(g).tobytes() == np.asarray(w).tobytes(), (p, f, g, w)
        # the reference's MAP threads also compute cells outside the matrix;
        # each adds exp(-log Z) - minpp to its predecessor, so none of them
        # can hold a maximum (neither model has them) while log Z > log(1/minpp)
        assert res[j]["logz"] > np.log(1.0 / cother_amd.MINPP), (p, res[j]["logz"])
    return res


@pytest.mark.parametrize("seed,lmin,lmax", [(61, 20, 140), (62, 100, 330)])
def test_map_random_profiles(ctx, oracle, seed, lmin, lmax):
    q = cother_amd.synth(2, 0, lmin=lmin, lmax=lmax, dist_mode=0, sort_desc=0, seed=seed)
    db = cother_amd.synth(16, 0, lmin=10, lmax=300, dist_mode=1, sort_desc=1, seed=seed + 50)
    ctx.upload_queries(q)
    ctx.upload_db(db)
    rng = np.random.default_rng(seed)
    for qi in range(q.nprof):
        prof = np.arange(db.nprof, dtype=np.int32)
        # log pair E-values both above and below -22 exercise both score adjustments
        epa = rng.choice([-5.0, -21.0, -25.0, -40.0, -120.0], size=db.nprof).astype(np.float32)
        print(f"seed {seed} query {qi}", flush=True)
        _check(ctx, oracle, q, db, qi, 0, db.nprof, prof, epa)


def test_map_self_alignments(ctx, oracle):
    # the same profiles as queries and templates: long MAP alignments with a
    # strong diagonal, traced back over many cells
    q = cother_amd.synth(3, 0, lmin=60, lmax=260, dist_mode=1, sort_desc=0, seed=71)
    db = cother_amd.synth(3, 0, lmin=60, lmax=260, dist_mode=1, sort_desc=0, seed=71)
    ctx.upload_queries(q)
    ctx.upload_db(db)
    for qi in range(3):
        res = _check(ctx, oracle, q, db, qi, 0, 3, np.arange(3, dtype=np.int32), np.full(3, -30.0, np.float32))
        assert res[qi]["alnlen"] > 0.8 * q.lens()[qi], res[qi]


def test_alignment_rows_match_oracle(ctx, oracle):
    # self hits (long alignments) plus chance hits among random profiles
    q = cother_amd.synth(2, 0, lmin=60, lmax=200, dist_mode=1, sort_desc=0, seed=72)
    same = cother_amd.synth(2, 0, lmin=60, lmax=200, dist_mode=1, sort_desc=0, seed=72)
    other = cother_amd.synth(12, 0, lmin=30, lmax=300, dist_mode=1, sort_desc=0, seed=73)
    db = cother_amd.concat([same, other])
    ctx.upload_queries(q)
    ctx.upload_db(db)
    hits, nh, _, alns = ctx.search(eval=100.0, alignments=True)
    assert len(hits) >= 2
    d = db.dists()
    for h, rows in zip(hits, alns):
        qi, p = int(h["query"]), int(h["profile"])
        sc, md = oracle.score_matrix(q, qi, db, 0, db.nprof)
        res, want = oracle.map_profile(q, qi, db, p, sc, md, int(d[p]), float(h["logEpA"]), want_aln=True)
        assert int(res["alnlen"]) == int(h["alnlen"]) and rows == want, (qi, p)
    assert any(int(h["alnlen"]) > 50 for h in hits)
