"""The MAP forward/backward of the oracle (orc_map.c; ExecMAPDP_Fwd_Unroll32x
and ExecMAPDP_Bwd_Unroll32x, cumapdp/CuBatchMAPDP_fwd.cu:82-478,
CuBatchMAPDP_bwd.cu:82-545) against an exact twin in float64 on tiny
profiles.

The twin restates the two recursions with exact exp/log (numpy logaddexp)
over the same moves and transition sums: forward, a five-state log-sum-exp
over paths ending at each cell (a path may start afresh at any MM cell with
log-weight 0: the "log 1" term); backward, over paths starting at each cell
(the end term log 1).  Then:

- forward total = backward total: sum over cells of exp(forward MM) equals
  the sum of exp(backward MM), and both equal the brute-force sum over every
  explicit path (enumerated) -- which fixes that the forward and backward
  recursions use the same move weights (a transition read at another
  position in one of them breaks it, planted below);
- the oracle's forward matrix, log partition function and posteriors
  exp(fwd + bwd - log Z) agree with the twin run on the oracle's own
  boundary (the cells left of column 0 of the 32x32 block schedule, log 0 =
  -99) to within the error of Schraudolph's exp/log approximations, which
  the oracle (like the reference, CUMAPDP_USEFASTAPPROXMATH) uses.
"""
import itertools
import math

import numpy as np
import pytest

import cother_amd

LP0 = -99.0
MM, MI, IM, DN, ND = range(5)
tMM, tMI, tMD, tIM, tII, tDM, tDD = range(7)


def _trn_fwd(t, base, pos, n, lo):
    if pos < 0 or pos >= n:
        return [lo] * 7
    b = base + pos
    return [t[tMM, b], t[tMI, b + 1], t[tMD, b], t[tIM, b], t[tII, b + 1], t[tDM, b], t[tDD, b]]


def _trn_bwd(t, base, pos, n, lo):
    if pos < 0 or pos >= n:
        return [lo] * 7
    return [t[k, base + pos + 1] for k in range(7)]


def _lse(vals):
    m = max(vals)
    if m == -math.inf:
        return -math.inf
    return m + math.log(sum(math.exp(v - m) for v in vals))


class Twin:
    """Exact forward/backward.  boundary=True reproduces the oracle's extended
    grid (rows computed from column -1 - (y mod 32), log 0 = -99); False is
    the plain recursion over the matrix with log 0 = -inf."""

    def __init__(self, q, db, S, boundary, shift_bwd_mi=False):
        self.lq, self.lt = int(q.lens()[0]), int(db.lens()[0])
        self.S = S.astype(np.float64)
        self.boundary = boundary
        self.lo = LP0 if boundary else -math.inf
        tq, tt = q.field("trn").astype(np.float64), db.field("trn").astype(np.float64)
        self.qf = {y: _trn_fwd(tq, 0, y, self.lq, self.lo) for y in range(-1, self.lq + 1)}
        self.df = {c: _trn_fwd(tt, 0, c, self.lt, self.lo) for c in range(-40, self.lt + 1)}
        self.qb = {y: _trn_bwd(tq, 0, y, self.lq, self.lo) for y in range(-1, self.lq + 1)}
        self.db_ = {c: _trn_bwd(tt, 0, c, self.lt, self.lo) for c in range(-1, self.lt + 1)}
        if shift_bwd_mi:  # a planted mistake: the backward reads MI one position later
            for c in range(-1, self.lt):
                self.db_[c] = list(self.db_[c])
                self.db_[c][tMI] = self.db_[c + 1][tMI] if c + 1 < self.lt else self.lo

    def s(self, y, c):
        return self.S[y, c] if 0 <= y < self.lq and 0 <= c < self.lt else self.lo

    def forward(self):
        lq, lt, lo = self.lq, self.lt, self.lo
        F = {}

        def get(y, c):
            return F.get((y, c), [lo] * 5)

        for y in range(lq):
            c0 = -1 - (y % 32) if self.boundary else 0
            for c in range(c0, lt):
                qt, dt, s = self.qf[y], self.df[c], self.s(y, c)
                G, U, L = get(y - 1, c - 1), get(y - 1, c), get(y, c - 1)
                o = [0.0] * 5
                o[MM] = _lse([G[MI] + qt[tMM] + dt[tIM] + s, G[IM] + qt[tIM] + dt[tMM] + s,
                              G[DN] + qt[tDM] + dt[tMM] + s, G[ND] + qt[tMM] + dt[tDM] + s,
                              G[MM] + qt[tMM] + dt[tMM] + s, 0.0])
                o[MI] = _lse([U[MM] + qt[tMM] + dt[tMI], U[MI] + qt[tMM] + dt[tII]])
                o[IM] = _lse([L[MM] + qt[tMI] + dt[tMM], L[IM] + qt[tII] + dt[tMM]])
                o[DN] = _lse([U[MM] + qt[tMD], U[DN] + qt[tDD]])
                o[ND] = _lse([L[MM] + dt[tMD], L[ND] + dt[tDD]])
                F[(y, c)] = o
        return np.array([[F[(y, c)][MM] for c in range(lt)] for y in range(lq)])

    def backward(self):
        lq, lt, lo = self.lq, self.lt, self.lo
        B = {}

        def get(y, c):
            return B.get((y, c), [lo] * 5)

        for y in range(lq - 1, -1, -1):
            for c in range(lt - 1, -1, -1):
                qt, dt = self.qb[y], self.db_[c]
                s = self.s(y + 1, c + 1)
                Dn, Rt, Dg = get(y + 1, c), get(y, c + 1), get(y + 1, c + 1)
                o = [0.0] * 5
                o[MM] = _lse([Dn[MI] + qt[tMM] + dt[tMI], Rt[IM] + qt[tMI] + dt[tMM], Dn[DN] + qt[tMD],
                              Rt[ND] + dt[tMD], Dg[MM] + qt[tMM] + dt[tMM] + s, 0.0])
                o[MI] = _lse([Dg[MM] + qt[tMM] + dt[tIM] + s, Dn[MI] + qt[tMM] + dt[tII]])
                o[IM] = _lse([Dg[MM] + qt[tIM] + dt[tMM] + s, Rt[IM] + qt[tII] + dt[tMM]])
                o[DN] = _lse([Dg[MM] + qt[tDM] + dt[tMM] + s, Dn[DN] + qt[tDD]])
                o[ND] = _lse([Dg[MM] + qt[tMM] + dt[tDM] + s, Rt[ND] + dt[tDD]])
                B[(y, c)] = o
        return np.array([[B[(y, c)][MM] for c in range(lt)] for y in range(lq)])

    def brute_total(self):
        """log of the sum over every explicit path (start at an MM cell with
        log-weight 0, state moves as the forward recursion, end at an MM cell)."""
        lq, lt = self.lq, self.lt
        out = []

        def ext(y, c, st, w):
            # the path is at (y, c) in state st with weight w
            if st == MM:
                out.append(w)
            nxt = []
            if st == MM:
                nxt = [(y + 1, c + 1, MM, lambda qt, dt: qt[tMM] + dt[tMM]), (y + 1, c, MI, lambda qt, dt: qt[tMM] + dt[tMI]),
                       (y, c + 1, IM, lambda qt, dt: qt[tMI] + dt[tMM]), (y + 1, c, DN, lambda qt, dt: qt[tMD]),
                       (y, c + 1, ND, lambda qt, dt: dt[tMD])]
            elif st == MI:
                nxt = [(y + 1, c + 1, MM, lambda qt, dt: qt[tMM] + dt[tIM]), (y + 1, c, MI, lambda qt, dt: qt[tMM] + dt[tII])]
            elif st == IM:
                nxt = [(y + 1, c + 1, MM, lambda qt, dt: qt[tIM] + dt[tMM]), (y, c + 1, IM, lambda qt, dt: qt[tII] + dt[tMM])]
            elif st == DN:
                nxt = [(y + 1, c + 1, MM, lambda qt, dt: qt[tDM] + dt[tMM]), (y + 1, c, DN, lambda qt, dt: qt[tDD])]
            else:
                nxt = [(y + 1, c + 1, MM, lambda qt, dt: qt[tMM] + dt[tDM]), (y, c + 1, ND, lambda qt, dt: dt[tDD])]
            for ny, nc, ns, f in nxt:
                if ny >= lq or nc >= lt:
                    continue
                # forward transition sums at the destination cell's positions
                w2 = w + f(self.qf[ny], self.df[nc]) + (self.S[ny, nc] if ns == MM else 0.0)
                ext(ny, nc, ns, w2)

        for y, c in itertools.product(range(lq), range(lt)):
            ext(y, c, MM, 0.0)
        return _lse(out), len(out)


def _case(seed, lq, lt):
    rng = np.random.default_rng(seed)
    q = cother_amd.synth(1, 2, lmin=lq, lmax=lq, dist_mode=2, sort_desc=0, seed=seed)
    db = cother_amd.synth(1, 2, lmin=lt, lmax=lt, dist_mode=2, sort_desc=0, seed=seed + 500)
    S = rng.uniform(-2.0, 2.5, (lq, lt)).astype(np.float32)
    return q, db, S


@pytest.mark.parametrize("lq,lt", [(1, 1), (2, 3), (3, 3), (3, 4), (4, 3), (4, 4)])
def test_forward_total_equals_backward_total_and_path_sum(lq, lt):
    for seed in range(4):
        q, db, S = _case(3000 + 100 * lq + 10 * lt + seed, lq, lt)
        T = Twin(q, db, S, boundary=False)
        F, B = T.forward(), T.backward()
        zf = _lse(list(F.ravel()))
        zb = _lse(list(B.ravel()))
        zp, npaths = T.brute_total()
        assert npaths >= lq * lt
        assert abs(zf - zb) <= 1e-12 * max(1.0, abs(zf)), (zf, zb)
        assert abs(zf - zp) <= 1e-12 * max(1.0, abs(zf)), (zf, zp)
        # per-cell posteriors are probabilities of passing through the cell
        post = np.exp(F + B - zf)
        assert (post > 0).all() and (post <= 1 + 1e-12).all()


def test_planted_backward_transition_shift_breaks_the_identity():
    bad = 0
    for seed in range(4):
        q, db, S = _case(3900 + seed, 4, 4)
        T = Twin(q, db, S, boundary=False, shift_bwd_mi=True)
        zf, zb = _lse(list(T.forward().ravel())), _lse(list(T.backward().ravel()))
        bad += abs(zf - zb) > 1e-6
    assert bad >= 3


@pytest.mark.parametrize("lq,lt", [(3, 4), (5, 5), (6, 4), (8, 9)])
def test_oracle_forward_backward_within_approximation_error(oracle, lq, lt):
    """The oracle (Schraudolph exp/log in float32) against the exact twin on
    the oracle's boundary: per-cell forward values within 0.05 + 0.02 per
    cell of path length (the approximation's error in log space, accumulated
    over the recursion), log Z within 0.15, posteriors within 0.05."""
    import ctypes
    L = oracle.L
    vp = ctypes.c_void_p
    L.orc_map_fwd.restype = ctypes.c_float
    L.orc_map_fwd.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int, vp, vp]
    L.orc_map_bwd.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int, vp, vp, ctypes.c_float, vp, vp,
                              ctypes.POINTER(ctypes.c_double)]
    worst_f = worst_p = 0.0
    for seed in range(4):
        q, db, S = _case(4000 + 100 * lq + 10 * lt + seed, lq, lt)
        S = np.ascontiguousarray(S)
        fwd = np.zeros((lq, lt), np.float32)
        maxlog = L.orc_map_fwd(ctypes.cast(q.ptr, vp), 0, ctypes.cast(db.ptr, vp), 0, S.ctypes.data, fwd.ctypes.data)
        post = np.zeros((lq, lt), np.float32)
        bwd = np.zeros((lq, lt), np.float32)
        z = ctypes.c_double()
        L.orc_map_bwd(ctypes.cast(q.ptr, vp), 0, ctypes.cast(db.ptr, vp), 0, S.ctypes.data, fwd.ctypes.data,
                      maxlog, post.ctypes.data, bwd.ctypes.data, ctypes.byref(z))
        logz = math.log(z.value) + maxlog
        T = Twin(q, db, S, boundary=True)
        F, B = T.forward(), T.backward()
        zt = _lse(list(F.ravel()))
        steps = np.add.outer(np.arange(lq), np.arange(lt)) + 1
        err_f = np.abs(fwd - F)
        assert (err_f <= 0.05 + 0.02 * steps).all(), err_f.max()
        assert abs(logz - zt) <= 0.15, (logz, zt)
        p_or = np.exp(post.astype(np.float64) - logz)
        p_tw = np.exp(F + B - zt)
        assert np.abs(p_or - p_tw).max() <= 0.05
        worst_f, worst_p = max(worst_f, err_f.max()), max(worst_p, np.abs(p_or - p_tw).max())
    print(f"Lq={lq} Lt={lt}: max |fwd - exact| {worst_f:.4f}, max |posterior - exact| {worst_p:.4f}")
