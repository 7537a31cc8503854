"""Exactly equal alignment maxima at different cells (CPU only).

The reference keeps the phase-1 maximum in per-column buffers updated by
each 32x32 oblique block and reduces them at the end (FinalizeDPPro,
cudp/CuBatchDP_final.cu:100-142), so among cells with exactly the same
maximum its winner follows that schedule; this build (oracle and GPU) takes
the first maximum in row-major order.  Here the tie is planted: constant
transitions and isolated equal scores make several cells' MM values exactly
equal.  Both rules must find the same maximum and pick one of the tied cells;
how often their picks differ is printed (it is a documented deviation, not a
parity failure: untied maxima agree, tests/test_oracle_dp.py)."""
import numpy as np

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
        assert mine == min(cells)  # first in row-major order
        trials += 1
        differ += mine != theirs
    print(f"planted ties: {trials} matrices with 2-4 tied cells, the reference schedule picks another tied "
          f"cell in {differ}", flush=True)
    assert trials == 40
