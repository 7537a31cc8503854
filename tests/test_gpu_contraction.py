"""The reduced-precision log-odds contraction (BF16 / FP16 operands on the
matrix cores, FP32 accumulation): lane layout of the MFMA instruction,
per-cell score error against an FP64 evaluation of the term, and agreement
of the hit lists with the FP32 path.  The FP32 path itself is the parity
path (bit-exact against the oracle, test_gpu_score.py); this variant has no
reference counterpart and is held to the stated tolerances."""
import numpy as np
import pytest

import cother_amd

pytestmark = pytest.mark.gpu

# BF16 keeps 8 significant bits: each operand is within 2^-9 relative, so a
# sum of positive products is within ~2^-8 and its logarithm within ~0.004;
# FP16 (11 bits) within ~5e-4
TOL = {1: 8e-3, 2: 1e-3}


def _init64(q, qi, db, p0, p1):
    """The log-odds term in FP64 (CuBatchScoreMatrix_Init.cu:129-144): [Lq][ncols]."""
    qd, ql = q.dists(), q.lens()
    dd, dl = db.dists(), db.lens()
    Q = q.field("trg")[:, qd[qi]:qd[qi] + ql[qi]].astype(np.float64)  # [20][Lq]
    bq = q.field("bkg")[qi].astype(np.float64)
    cols = []
    for p in range(p0, p1):
        T = db.field("trg")[:, dd[p]:dd[p] + dl[p]].astype(np.float64)
        bt = db.field("bkg")[p].astype(np.float64)
        cols.append(T / (0.5 * (bq + bt))[:, None])
    Tw = np.concatenate(cols, axis=1)  # [20][ncols]
    return np.log(Q.T @ Tw) - 0.002


def test_mfma_lane_layout():
    assert cother_amd.selftest_mfma() == 0


@pytest.mark.parametrize("prec", [1, 2])
def test_score_error_against_fp64(ctx, prec):
    q = cother_amd.synth(3, 0, lmin=40, lmax=200, dist_mode=1, sort_desc=0, seed=151)
    db = cother_amd.synth(40, 0, lmin=30, lmax=300, dist_mode=1, sort_desc=0, seed=152)
    ctx.upload_queries(q)
    ctx.upload_db(db)
    for qi in range(q.nprof):
        s32, m32 = ctx.score_block(qi, 0, db.nprof)
        sv, mv = ctx.score_block(qi, 0, db.nprof, contraction=prec)
        init64 = _init64(q, qi, db, 0, db.nprof)
        # the other terms are the same FP32 values in both paths, so the
        # variant's error against FP64 is its log-odds term's error
        other = s32.astype(np.float64) - init64
        err = np.abs(sv.astype(np.float64) - (init64 + other))
        assert err.max() < TOL[prec], (qi, err.max())
        # the FP32 path's own log-odds error is at the FP32 rounding level
        assert np.abs(s32 - (init64 + other)).max() < 1e-5
        # the modular scores do not contain the log-odds term
        assert (mv.view(np.int32) == m32.view(np.int32)).all()
        print(f"prec {prec} query {qi}: max {err.max():.2e} mean {err.mean():.2e}", flush=True)


@pytest.mark.parametrize("prec", [1, 2])
def test_hit_lists_agree_with_fp32(ctx, prec):
    q = cother_amd.synth(4, 0, lmin=50, lmax=200, dist_mode=1, sort_desc=0, seed=161)
    same = cother_amd.subset_list(q, [0, 2])
    other = cother_amd.synth(200, 1, 220.0, 0.35, 30, 800, dist_mode=1, sort_desc=1, seed=162)
    db = cother_amd.concat([same, other])
    ctx.upload_queries(q)
    ctx.upload_db(db)
    h32, n32, _ = ctx.search(eval=100.0)
    hv, nv, _ = ctx.search(eval=100.0, contraction=prec)
    o32 = np.concatenate([[0], np.cumsum(n32)])
    ov = np.concatenate([[0], np.cumsum(nv)])
    for qi in range(q.nprof):
        a = h32[o32[qi]:o32[qi + 1]]
        b = hv[ov[qi]:ov[qi + 1]]
        top_a, top_b = set(a["profile"][:100]), set(b["profile"][:100])
        if top_a:
            assert len(top_a & top_b) >= 0.99 * len(top_a) - 1, (qi, sorted(top_a ^ top_b))
        common = np.intersect1d(a["profile"], b["profile"])
        for p in common:
            ea = float(a["logE"][a["profile"] == p][0])
            eb = float(b["logE"][b["profile"] == p][0])
            assert abs(ea - eb) <= 0.05 * max(1.0, abs(ea)), (qi, p, ea, eb)
