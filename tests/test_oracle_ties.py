"""Exactly equal alignment maxima at different cells, and the maxima the
reference's buffer drops (CPU only).

The reference keeps the phase-1 maximum in a per-column buffer that each
32x32 oblique block reads at its start and writes at its end, and reduces
the buffer's last 64 columns (ExecDP_Btck_Unroll32x,
cudp/CuBatchDP_init_btck.cu:246-263,436-464; FinalizeDPPro,
cudp/CuBatchDP_final.cu:100-142), so among cells with exactly the same
maximum its winner follows that schedule, and for templates under 32
positions the rows whose block-end column falls before the matrix do not
reach the buffer at all.  The oracle (and the GPU path) replay that flow;
these tests check it against the replay of the block schedule
(tests/_blocksim.py) where the flow decides: planted ties, integer scores,
short templates."""
import numpy as np
import pytest

import cother_amd
from tests import _blocksim

f32 = np.float32


def _flat_transitions(ps):
    """every transition vector of the set the same (so equal scores give equal DP values)"""
    t = ps.field("trn")
    for k in range(t.shape[0]):
        t[k, :] = t[k, 0]


def test_planted_ties(oracle):
    q = cother_amd.synth(1, 2, lmin=90, lmax=90, dist_mode=2, sort_desc=0, seed=501)
    db = cother_amd.synth(1, 2, lmin=120, lmax=120, dist_mode=2, sort_desc=0, seed=502)
    _flat_transitions(q)
    _flat_transitions(db)
    lq, lt = 90, 120
    qt = q.field("trn").reshape(-1)
    dt = db.field("trn").reshape(-1)
    rng = np.random.default_rng(503)
    edge = _blocksim._corner(lq, lt)
    trials, differ = 0, 0
    for t in range(40):
        S = np.full((lq, lt), f32(-3.0))
        k = int(rng.integers(2, 5))
        cells = set()
        while len(cells) < k:
            r, c = int(rng.integers(2, lq - 2)), int(rng.integers(2, lt - 2))
            # inside the valid area (CellYXinValidArea, cudp/CuBatchDP_init_btck.cuh:118-170)
            if _blocksim._valid(lq, lt, edge, r, c) and all(abs(r - a) > 3 or abs(c - b) > 3 for a, b in cells):
                cells.add((r, c))
        for r, c in cells:
            S[r, c] = f32(4.0)
        res, _ = oracle.phase1(q, 0, db, 0, 1, S, 0.0)
        sbt, smax, sxy = _blocksim.simulate(S, qt, 0, q.npos + 1, lq, dt, 0, db.npos + 1, lt)
        r = res[0]
        assert r["maxsc"] == smax > 0
        mine = (int(r["maxxy"]) & 0xFFFF, int(r["maxxy"]) >> 16)
        theirs = (sxy & 0xFFFF, sxy >> 16)
        # the cells in the valid area attaining the maximum
        tied = sorted(cells)
        assert mine in cells and theirs in cells, (mine, theirs, tied)
        assert mine == theirs, (t, mine, theirs, tied)
        trials += 1
        differ += mine != min(cells)
    print(f"planted ties: {trials} matrices with 2-4 tied cells; the reference's pick is not the first in "
          f"row-major order in {differ}", flush=True)
    assert trials == 40 and differ > 0


def _run(oracle, q, db, S):
    lq, lt = S.shape
    res, bt = oracle.phase1(q, 0, db, 0, 1, S, 0.0)
    sbt, smax, sxy = _blocksim.simulate(S, q.field("trn").reshape(-1), 0, q.npos + 1, lq,
                                        db.field("trn").reshape(-1), 0, db.npos + 1, lt)
    return res[0], bt, sbt, smax, sxy


@pytest.mark.parametrize("lq,lt", [(3, 2), (12, 5), (40, 7), (70, 20), (33, 31), (65, 30), (20, 33), (100, 64)])
def test_buffer_flow_against_block_schedule(oracle, lq, lt):
    """integer-valued scores (many equal cell values) and templates under 32
    positions: maximum, its cell and the traceback bytes as the schedule
    replay gives them"""
    q = cother_amd.synth(1, 2, lmin=lq, lmax=lq, dist_mode=2, sort_desc=0, seed=lq * 131 + lt)
    db = cother_amd.synth(1, 2, lmin=lt, lmax=lt, dist_mode=2, sort_desc=0, seed=lq * 137 + lt)
    _flat_transitions(q)
    _flat_transitions(db)
    rng = np.random.default_rng(lq * 1009 + lt)
    dropped = 0
    for t in range(6):
        S = rng.integers(-3, 4, (lq, lt)).astype(f32)
        r, bt, sbt, smax, sxy = _run(oracle, q, db, S)
        assert (bt == sbt).all()
        assert r["maxsc"] == smax, (t, r["maxsc"], smax)
        assert int(r["maxxy"]) == sxy, (t, int(r["maxxy"]) & 0xFFFF, int(r["maxxy"]) >> 16, sxy & 0xFFFF, sxy >> 16)
        full = np.max(np.where(S > 0, S, 0))  # a positive cell's own MM is at least its score
        dropped += int(smax < full)
    print(f"lq {lq} lt {lt}: reported maximum below a single positive cell's score in {dropped} of 6", flush=True)
