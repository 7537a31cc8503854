"""Pin the oracle's phase-1 DP and finalisation (CPU only).

- Cell values and traceback bytes against a step-by-step replay of the
  reference's 32x32 oblique-block schedule (tests/_blocksim.py), which also
  checks the oracle's derived boundary rules (row starts at -1-(q mod 32),
  zero boundaries, q mod 32 == 0 rows seeing zeros above negative columns).
- The maximum and its coordinates against the block schedule's own
  max-buffer flow and FinalizeDPPro reduction.
- The correlation bonus against its closed form sum_k s_k (s_{k-1}+..+s_{k-4})
  over the DIAG scores of the traced path (independent of the 32-score
  chunking), and the begin coordinates against a direct traceback.
- The score threshold against the float64 closed form.
"""
import math

import numpy as np
import pytest

import cother_amd
from tests import _blocksim

f32 = np.float32


@pytest.fixture(scope="module")
def sets():
    q = cother_amd.synth(3, 0, lmin=35, lmax=110, dist_mode=0, sort_desc=0, seed=21)
    db = cother_amd.synth(8, 0, lmin=12, lmax=95, dist_mode=1, sort_desc=1, seed=22)
    return q, db


def planted(lq, lt, rng):
    S = rng.normal(-0.6, 0.8, (lq, lt)).astype(np.float32)
    k = min(lq, lt)
    off = int(rng.integers(0, max(1, lt - k // 2)))
    for i in range(min(lq, lt - off)):
        if rng.random() < 0.9:
            S[i, off + i] = f32(rng.uniform(0.5, 2.5))
    return S


def run_both(oracle, q, qp, db, p, S, thld=0.0):
    res, bt = oracle.phase1(q, qp, db, p, p + 1, S, thld)
    qt = q.field("trn").reshape(-1)
    dt = db.field("trn").reshape(-1)
    lq, lt = int(q.lens()[qp]), int(db.lens()[p])
    sbt, smax, sxy = _blocksim.simulate(S, qt, int(q.dists()[qp] + qp), q.npos + q.nprof, lq,
                                        dt, int(db.dists()[p] + p), db.npos + db.nprof, lt)
    return res[0], bt, sbt, smax, sxy


def test_dp_matches_block_schedule_real_scores(oracle, sets):
    q, db = sets
    for qp in range(q.nprof):
        for p in range(0, db.nprof, 2):
            S, _ = oracle.score_matrix(q, qp, db, p, p + 1)
            r, bt, sbt, smax, sxy = run_both(oracle, q, qp, db, p, S)
            assert (bt == sbt).all(), (qp, p, np.argwhere(bt != sbt)[:5])
            assert r["maxsc"] == smax
            if smax > 0:
                assert r["maxxy"] == sxy


def test_dp_matches_block_schedule_planted(oracle, sets):
    q, db = sets
    rng = np.random.default_rng(4)
    for qp in range(q.nprof):
        for p in range(db.nprof):
            lq, lt = int(q.lens()[qp]), int(db.lens()[p])
            S = planted(lq, lt, rng)
            r, bt, sbt, smax, sxy = run_both(oracle, q, qp, db, p, S)
            assert (bt == sbt).all(), (qp, p, np.argwhere(bt != sbt)[:5])
            assert r["maxsc"] == smax and smax > 0
            assert r["maxxy"] == sxy


def test_finalize_closed_form(oracle, sets):
    q, db = sets
    rng = np.random.default_rng(8)
    n = 0
    for qp in range(q.nprof):
        for p in range(db.nprof):
            lq, lt = int(q.lens()[qp]), int(db.lens()[p])
            S = planted(lq, lt, rng)
            res, bt = oracle.phase1(q, qp, db, p, p + 1, S, 0.5)
            r = res[0]
            if not r["passed"] and r["maxsc"] <= 0.5:
                continue
            x, y = int(r["maxxy"] >> 16), int(r["maxxy"] & 0xffff)
            s = []
            while x >= 0 and y >= 0:
                b = bt[y, x]
                if b == 0:
                    break
                if b == 2:
                    y -= 1
                elif b == 1:
                    x -= 1
                else:
                    s.append(float(S[y, x])); x -= 1; y -= 1
            corr = sum(s[k] * sum(s[max(0, k - 4):k]) for k in range(len(s)))
            bonus = math.copysign(math.sqrt(abs(0.5 * corr)), corr)
            assert abs(float(r["score"]) - (float(r["maxsc"]) + bonus)) < 1e-4 * max(1, abs(bonus))
            assert r["psts"] == sum(1 for v in s if v > 0)
            assert r["beg"] == (((x + 1) << 16) | (y + 1))
            assert r["passed"] == (r["score"] > 0.5)
            n += 1
    assert n > 5


def test_score_threshold_closed_form(oracle):
    for ev, lq, dbs in [(10.0, 150, 447922), (0.5, 250, 14_000_000), (1e-3, 40, 1000)]:
        s, le = oracle.score_threshold(ev, lq, dbs)
        e2 = ev * (10 if ev < 1 else 3)
        want = max(0.0, -math.log(e2 / (0.001488 * lq * dbs)) / 0.561132)
        assert abs(s - want) < 1e-4 * max(1, want)
        assert abs(le - math.log(ev)) < 1e-6
