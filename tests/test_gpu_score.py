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


def test_score_matrix_informative_context_vectors(ctx, oracle):
    """Random context vectors, norms and priors (the CVS term over all three
    branches of its map): bit-exact."""
    from tests._cv import set_informative_cv
    q = set_informative_cv(cother_amd.synth(2, 0, lmin=50, lmax=120, dist_mode=0, sort_desc=0, seed=111), 112)
    db = set_informative_cv(cother_amd.synth(30, 0, lmin=20, lmax=300, dist_mode=1, sort_desc=1, seed=113), 114)
    ctx.upload_queries(q)
    ctx.upload_db(db)
    for qi in range(q.nprof):
        sc, md = ctx.score_block(qi, 0, db.nprof)
        osc, omd = oracle.score_matrix(q, qi, db, 0, db.nprof)
        bad = np.argwhere(sc.view(np.int32) != osc.view(np.int32))
        assert bad.size == 0, f"query {qi}: {len(bad)} cells differ, first {bad[:5].tolist()}"
        assert (md.view(np.int32) == omd.view(np.int32)).all()
    # the modular scores (SS + CVS terms) are not constant across cells
    assert np.unique(md).size > 1000


def test_score_matrix_subrange(ctx, oracle, sets):
    q, db = sets
    ctx.upload_queries(q)
    ctx.upload_db(db)
    sc, md = ctx.score_block(1, 7, 19)
    osc, omd = oracle.score_matrix(q, 1, db, 7, 19)
    assert (sc.view(np.int32) == osc.view(np.int32)).all()
    assert (md.view(np.int32) == omd.view(np.int32)).all()


def test_score_matrix_adversarial_distance_lists(ctx, oracle):
    """Every position's distance list replaced by a hand-shaped one
    (tests/_ddms_lists.py: constant, near-zero, alternating, run-structured,
    missing-heavy, sorted; lengths 0..99 with 5..9 over-represented): the
    segment restarts, length test, first-maximum selection and top-3
    register of the DDMS term, bit for bit against the oracle."""
    from tests import _ddms_lists as dl
    q = cother_amd.synth(2, 0, lmin=60, lmax=130, dist_mode=0, sort_desc=0, seed=141)
    db = cother_amd.synth(30, 0, lmin=8, lmax=160, dist_mode=1, sort_desc=1, seed=142)
    rng = np.random.default_rng(143)
    dl.fill(q, rng)
    dl.fill(db, rng)
    # a few templates carry the query's own lists: long top segments
    qdd, ddd = q.field("dd"), db.field("dd")
    n = min(int(q.lens()[0]), int(db.lens()[1]))
    ddd[int(db.dists()[1]):int(db.dists()[1]) + n] = qdd[:n]
    ctx.upload_queries(q)
    ctx.upload_db(db)
    for qi in range(q.nprof):
        print(f"adversarial lists: query {qi}", flush=True)
        sc, md = ctx.score_block(qi, 0, db.nprof)
        osc, omd = oracle.score_matrix(q, qi, db, 0, db.nprof)
        bad = np.argwhere(sc.view(np.int32) != osc.view(np.int32))
        assert bad.size == 0, f"query {qi}: {len(bad)} cells differ, first {bad[:5].tolist()}"
        assert (md.view(np.int32) == omd.view(np.int32)).all()


@pytest.mark.parametrize("case", ["templates without lists", "query without lists", "lists shorter than DSTSEGM",
                                  "operands outside the quotient range"])
def test_score_matrix_without_ddms_bit_exact(ctx, oracle, case):
    """Where no cell can have a DDMS term (every template's or the query's
    distance lists shorter than DSTSEGM) the plain score kernel runs: its
    tall tiles (64 query rows; queries of 1-100 positions cover partial and
    several tiles) must give the oracle's matrices bit for bit, with
    informative context vectors.  The kernel's log-odds quotients take a
    division-free form for frequencies of 0 or in [2^-32, 2] and denominators
    in [2^-30, 2^30]; the last case plants zeros (admitted), frequencies of
    1e-12 and 3e-11 in some query rows and template columns and a background
    of 1e-12 in one query and one template profile (not admitted: those rows
    and wavefronts divide), so both forms run beside each other."""
    from tests._cv import set_informative_cv
    qmode, tmode = (2, 1) if case == "query without lists" else (0, 1) if case == "lists shorter than DSTSEGM" else (0, 2)
    q = set_informative_cv(cother_amd.synth(4, 0, lmin=1, lmax=100, dist_mode=qmode, sort_desc=0, seed=121), 122)
    db = set_informative_cv(cother_amd.synth(70, 0, lmin=5, lmax=300, dist_mode=tmode, sort_desc=1, seed=123), 124)
    if case == "lists shorter than DSTSEGM":
        dd = db.field("dd")
        dd[:, 0] = np.minimum(dd[:, 0], 5)
    if case == "operands outside the quotient range":
        qt, tt = q.field("trg"), db.field("trg")
        qt[3, ::7] = 0.0
        qt[5, ::11] = 1e-12
        tt[2, ::5] = 0.0
        tt[4, 3::97] = 3e-11
        q.field("bkg")[1, 6] = 1e-12
        db.field("bkg")[3, 6] = 1e-12
    ctx.upload_queries(q)
    ctx.upload_db(db)
    for qi in range(q.nprof):
        sc, md = ctx.score_block(qi, 0, db.nprof)
        osc, omd = oracle.score_matrix(q, qi, db, 0, db.nprof)
        assert sc.shape == osc.shape
        bad = np.argwhere(sc.view(np.int32) != osc.view(np.int32))
        assert bad.size == 0, f"{case}, query {qi}: {len(bad)} cells differ, first {bad[:5].tolist()}"
        assert (md.view(np.int32) == omd.view(np.int32)).all()
