"""The phase-1 DP of the oracle (orc_dp.c, ExecDP_Btck_Unroll32x,
cudp/CuBatchDP_init_btck.cu:65-494) pinned by explicit path enumeration on
tiny profiles (Lq, Lt <= 6).

The DP is a max-plus recursion over five pair states (MM, MI, IM, DN, ND):
its value at a cell and state is the best weight of any path of state moves
ending there, a path starting either at a cell the DP does not compute
(value 0: row -1, the columns left of a row's first computed column) or,
for MM, afresh at the cell itself (the floor at 0).  Here every path is
enumerated explicitly and its weight summed in float64:

- the maximum over all paths ending in MM at a cell of the matrix equals the
  DP's maximum, and the DP's argmax cell attains it;
- the traceback byte of every cell equals the reference's rule applied to the
  enumerated per-state maxima (STOP or DIAG for MM, then UP for MI, LEFT for
  IM, UP for DN, LEFT for ND when that state's value is strictly larger,
  CuBatchDP_init_btck.cu:304-360), wherever the decision is not a near-tie.

The move weights are the transition sums of the recursion (the query's and
the template's log transitions into the cell's positions, with MI/II read at
position+1, DPLocCacheTrnProbs, cudp/CuBatchDP_init.cuh:136-155) and the cell
score for MM (-9 outside the matrix and in the two ignored corner triangles,
CellYXinValidArea, cudp/CuBatchDP_init_btck.cuh:118-170).  What this pins
beyond the oracle's own loop: the predecessor cell of each state, the floor,
the order of the traceback decisions and the argmax; a swapped U/D neighbour
or a floor applied before the score fails it (planted below).
"""
import numpy as np
import pytest

import cother_amd
from tests import _blocksim

LOG0 = -32768.0
MM, MI, IM, DN, ND = range(5)
tMM, tMI, tMD, tIM, tII, tDM, tDD = range(7)
STOP, LEFT, UP, DIAG = 0, 1, 2, 3


def _trn(ps, prof, pos, length):
    if pos < 0 or pos >= length:
        return [LOG0] * 7
    t = ps.field("trn").astype(np.float64)
    b = int(ps.dists()[prof]) + prof + pos
    return [t[tMM, b], t[tMI, b + 1], t[tMD, b], t[tIM, b], t[tII, b + 1], t[tDM, b], t[tDD, b]]


def _valid(lq, lt, y, x):
    edge = min(60, int(np.rint(np.float32(0.34) * np.float32(max(lq, lt)))))
    corners = (lq <= y + edge and x <= edge + y - lq) or (lt - edge + y <= x and y < edge)
    return (not corners) and 0 <= y < lq and 0 <= x < lt


class Paths:
    """Explicit enumeration of the DP's paths (weights in float64)."""

    def __init__(self, q, db, tp, sc, planted=None):
        self.lq, self.lt = int(q.lens()[0]), int(db.lens()[tp])
        self.qt = {y: _trn(q, 0, y, self.lq) for y in range(-1, self.lq)}
        self.dt = {c: _trn(db, tp, c, self.lt) for c in range(-40, self.lt)}
        self.sc = sc
        self.planted = planted

    def computed(self, y, c):
        return 0 <= y < self.lq and -1 - (y % 32) <= c < self.lt

    def S(self, y, c):
        return float(self.sc[y, c]) if _valid(self.lq, self.lt, y, c) else -9.0

    def weights(self, y, c, s):
        """All path weights ending at (y, c, s)."""
        if not self.computed(y, c):
            return [0.0]
        qt, dt = self.qt[y], self.dt[c]
        out = []
        if s == MM:
            out.append(0.0)  # a new alignment starts here (the floor)
            dy, dc = (y - 1, c - 1) if self.planted != "updiag" else (y - 1, c)
            for p, w in ((MI, qt[tMM] + dt[tIM]), (IM, qt[tIM] + dt[tMM]), (DN, qt[tDM] + dt[tMM]),
                         (ND, qt[tMM] + dt[tDM]), (MM, qt[tMM] + dt[tMM])):
                out += [v + w + self.S(y, c) for v in self.weights(dy, dc, p)]
            if self.planted == "floor_first":
                out = [max(v, self.S(y, c)) for v in out]
        elif s == MI:
            out += [v + qt[tMM] + dt[tMI] for v in self.weights(y - 1, c, MM)]
            out += [v + qt[tMM] + dt[tII] for v in self.weights(y - 1, c, MI)]
        elif s == IM:
            out += [v + qt[tMI] + dt[tMM] for v in self.weights(y, c - 1, MM)]
            out += [v + qt[tII] + dt[tMM] for v in self.weights(y, c - 1, IM)]
        elif s == DN:
            out += [v + qt[tMD] for v in self.weights(y - 1, c, MM)]
            out += [v + qt[tDD] for v in self.weights(y - 1, c, DN)]
        else:
            out += [v + dt[tMD] for v in self.weights(y, c - 1, MM)]
            out += [v + dt[tDD] for v in self.weights(y, c - 1, ND)]
        return out

    def best(self, y, c, s):
        return max(self.weights(y, c, s))


def _case(seed, lq, lt):
    rng = np.random.default_rng(seed)
    q = cother_amd.synth(1, 2, lmin=lq, lmax=lq, dist_mode=2, sort_desc=0, seed=seed)
    db = cother_amd.synth(1, 2, lmin=lt, lmax=lt, dist_mode=2, sort_desc=0, seed=seed + 1000)
    # scores in [-2, 4] so that alignments of several cells and gaps compete
    sc = rng.uniform(-2.0, 4.0, (lq, lt)).astype(np.float32)
    return q, db, sc


def _expected_byte(v, tol):
    """The traceback decision on the per-state maxima v[5]; None for a near-tie."""
    m, bk = v[MM], (DIAG if v[MM] > tol else STOP)
    if 0.0 < v[MM] <= tol:
        return None
    for s, d in ((MI, UP), (IM, LEFT), (DN, UP), (ND, LEFT)):
        if abs(v[s] - m) <= tol:
            return None
        if m < v[s]:
            m, bk = v[s], d
    return bk


@pytest.mark.parametrize("lq,lt", [(1, 1), (2, 3), (3, 2), (3, 5), (4, 4), (5, 3), (5, 6), (6, 6)])
def test_dp_equals_path_enumeration(oracle, lq, lt):
    tol = 2e-3
    for seed in range(6):
        q, db, sc = _case(1000 * lq + 10 * lt + seed, lq, lt)
        res, bt = oracle.phase1(q, 0, db, 0, 1, sc, 1e9)
        P = Paths(q, db, 0, sc)
        best = {(y, c): [P.best(y, c, s) for s in range(5)] for y in range(lq) for c in range(lt)}
        # the reference reports the maximum that its per-column buffer keeps
        # (for templates under 32 positions, rows whose block-end column lies
        # before the matrix drop out), applied here to the enumerated values
        mm = np.array([[max(0.0, best[(y, c)][MM]) for c in range(lt)] for y in range(lq)], np.float32)
        bmax, _ = _blocksim.flow_max(mm, lq, lt)
        assert abs(float(res["maxsc"][0]) - float(bmax)) <= tol, (seed, res["maxsc"][0], bmax)
        xy = int(res["maxxy"][0])
        x, y = xy >> 16, xy & 0xFFFF
        if bmax > 0:
            assert abs(best[(y, x)][MM] - bmax) <= tol
        checked = 0
        for (y, c), v in best.items():
            e = _expected_byte(v, tol)
            if e is None:
                continue
            assert bt[y, c] == e, (seed, y, c, v, bt[y, c])
            checked += 1
        assert checked >= lq * lt // 2


@pytest.mark.parametrize("planted", ["updiag", "floor_first"])
def test_planted_recursion_mistakes_are_detected(oracle, planted):
    """A diagonal move read from the cell above, or the floor taken before the
    score is added, gives path maxima the oracle does not have."""
    differs = 0
    for seed in range(8):
        q, db, sc = _case(77 + seed, 4, 5)
        res, _ = oracle.phase1(q, 0, db, 0, 1, sc, 1e9)
        P = Paths(q, db, 0, sc, planted=planted)
        bmax = max(max(0.0, P.best(y, c, MM)) for y in range(4) for c in range(5))
        differs += abs(float(res["maxsc"][0]) - bmax) > 1e-2
    assert differs >= 4
