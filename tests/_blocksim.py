"""Step-by-step simulation of the reference's phase-1 DP block schedule.

A test helper (CPU, small matrices only) that replays, in numpy float32,
the data movement of ExecDP_Btck_Unroll32x (cudp/CuBatchDP_init_btck.cu:
65-494) launched over block diagonals as CuBatchDP::PerformCompleteDynProg
Device does (cudp/CuBatchDP.cu:191-312), followed by FinalizeDPPro's max
reduction (cudp/CuBatchDP_final.cu:100-142): oblique 32x32 blocks, their
diagonal / bottom / max buffers and the per-thread max tracking.  It exists
to check the oracle's boundary rules (which it derives from this schedule)
and its max/argmax against what the block schedule itself produces.
"""
import numpy as np

f32 = np.float32
D = 32
LOG0 = f32(-32768.0)
MM, MI, IM, DN, ND = range(5)
tMM, tMI, tMD, tIM, tII, tDM, tDD = range(7)


def _trn(trn, base, pos, length, ld):
    if pos < 0 or pos >= length:
        return [LOG0] * 7
    b = base + pos
    return [trn[tMM * ld + b], trn[tMI * ld + b + 1], trn[tMD * ld + b], trn[tIM * ld + b],
            trn[tII * ld + b + 1], trn[tDM * ld + b], trn[tDD * ld + b]]


def _corner(lq, lt):
    m = int(np.rint(f32(0.34) * f32(max(lq, lt))))
    return min(m, 60)


def _valid(lq, lt, e, y, x):
    c = (lq <= y + e and x <= e + y - lq) or (lt - e + y <= x and y < e)
    return (not c) and 0 <= y < lq and 0 <= x < lt


def _maxiters(x, y, lq, lt):
    x = lt - x
    if x <= 0:
        return 0
    y -= lq - 1
    if 0 < y:
        x -= y
    if x <= 0:
        return 0
    return min(32, x)


def simulate(S, qtrn, qbase, qld, lq, dtrn, dbase, dld, lt):
    """S: [lq][lt] float32 scores.  Returns (btck [lq][lt], maxsc, maxxy)."""
    qt = [[f32(v) for v in _trn(qtrn, qbase, r, lq, qld)] for r in range(-64, lq + 64)]
    dtc = {}

    def dt(c):
        if c not in dtc:
            dtc[c] = [f32(v) for v in _trn(dtrn, dbase, c, lt, dld)]
        return dtc[c]

    diag1 = np.zeros((5, lt), f32)
    diag2 = np.zeros((5, lt), f32)
    diagm = np.zeros(lt, f32)
    bot = np.zeros((5, lt), f32)
    mco = np.zeros(lt, np.uint32)
    btck = np.zeros((lq, lt), np.uint8)
    e = _corner(lq, lt)
    nblk = ((lt + lq) + 31) // 32 + (lq - 1) // 32
    last = 2 * ((lq - 1) // 32)
    for d in range(nblk):
        for bx in range(nblk + 2):
            if d <= last:
                x0 = (2 * bx - (0 if d & 1 else 1)) * D
                y = ((d >> 1) + 1 - bx) * D - 1
            else:
                x0 = (2 * bx + (d - last - 1)) * D
                y = ((last >> 1) + 1 - bx) * D - 1
            ilim = _maxiters(x0, y, lq, lt)
            if y < 0 or lq <= y + 1 - D or lt <= x0 or ilim < 1:
                continue
            qtr = [qt[y - tx + 64] for tx in range(D)]
            sc = [[f32(-9.0)] * D for _ in range(D)]  # sc[tx][i]
            for tx in range(D):
                for i in range(D):
                    r, cc = y - tx, x0 + tx + i
                    if _valid(lq, lt, e, r, cc):
                        sc[tx][i] = f32(S[r][cc])
            p1 = np.zeros((5, D + 1), f32)
            p2 = np.zeros((5, D + 1), f32)
            mxs = [f32(0)] * D
            mxc = [0] * D
            bottm = np.zeros((5, D), f32)
            for tx in range(D):
                x = x0 + tx
                if 0 <= x - 1 < lt:
                    p1[:, tx] = diag1[:, x - 1]
                    p2[:, tx + 1] = diag2[:, x - 1]
                    mxs[tx] = diagm[x - 1]
                    mxc[tx] = int(mco[x - 1])
                if D <= y and 0 <= x + D - 1 < lt:
                    bottm[:, tx] = bot[:, x + D - 1]
            bt = np.zeros((D, D), np.uint8)
            for i in range(ilim):
                p1[:, D] = bottm[:, i]
                new = np.zeros((5, D), f32)
                for tx in range(D):
                    q = qtr[tx]
                    dd = dt(x0 + tx + i)
                    g = p2[:, tx + 1]   # diagonal
                    u = p1[:, tx + 1]   # up
                    l = p1[:, tx]       # left
                    v1 = f32(f32(g[MI] + q[tMM]) + dd[tIM])
                    v2 = f32(f32(g[IM] + q[tIM]) + dd[tMM]); v1 = v2 if v1 < v2 else v1
                    v2 = f32(f32(g[DN] + q[tDM]) + dd[tMM]); v1 = v2 if v1 < v2 else v1
                    v2 = f32(f32(g[ND] + q[tMM]) + dd[tDM]); v1 = v2 if v1 < v2 else v1
                    v2 = f32(f32(g[MM] + q[tMM]) + dd[tMM]); v1 = v2 if v1 < v2 else v1
                    v1 = f32(v1 + sc[tx][i]); mm = f32(0) if v1 < 0 else v1
                    m = mm; b = 3 if mm else 0
                    if mxs[tx] < mm:
                        mxs[tx] = mm; mxc[tx] = ((x0 + tx + i) << 16) | (y - tx)
                    a1 = f32(f32(u[MM] + q[tMM]) + dd[tMI]); a2 = f32(f32(u[MI] + q[tMM]) + dd[tII])
                    mi = a2 if a1 < a2 else a1
                    if m < mi: m, b = mi, 2
                    a1 = f32(f32(l[MM] + q[tMI]) + dd[tMM]); a2 = f32(f32(l[IM] + q[tII]) + dd[tMM])
                    im = a2 if a1 < a2 else a1
                    if m < im: m, b = im, 1
                    a1 = f32(u[MM] + q[tMD]); a2 = f32(u[DN] + q[tDD])
                    dn = a2 if a1 < a2 else a1
                    if m < dn: m, b = dn, 2
                    a1 = f32(l[MM] + dd[tMD]); a2 = f32(l[ND] + dd[tDD])
                    nd = a2 if a1 < a2 else a1
                    if m < nd: m, b = nd, 1
                    new[:, tx] = (mm, mi, im, dn, nd)
                    bt[tx][i] = b
                p2[:, :D] = new
                bottm[:, i] = new[:, 0]
                p1, p2 = p2, p1
            for tx in range(D):
                x = x0 + tx
                if 0 <= x + D - 1 < lt:
                    diag1[:, x + D - 1] = p1[:, tx]
                    diag2[:, x + D - 1] = p2[:, tx + 1]
                ndx = D - 1
                xtl = x0 + D - 1
                if xtl >= lt:
                    ndx = 0
                elif xtl + ndx >= lt:
                    ndx = lt - xtl - 1
                if 0 <= x + ndx < lt:
                    if D <= y:
                        if diagm[x + ndx] < mxs[tx]:
                            diagm[x + ndx] = mxs[tx]; mco[x + ndx] = mxc[tx]
                    else:
                        diagm[x + ndx] = mxs[tx]; mco[x + ndx] = mxc[tx]
                if 0 <= x < lt:
                    bot[:, x] = bottm[:, tx]
                for i in range(D):
                    r, cc = y - i, x + i
                    if 0 <= r < lq and 0 <= cc < lt:
                        btck[r][cc] = bt[i][tx]
    # FinalizeDPPro max reduction over the last 64 columns
    ms = [f32(0)] * 64
    for tx in range(32):
        if tx < lt:
            ms[tx] = diagm[lt - 1 - tx]
        if tx + 32 < lt:
            ms[tx + 32] = diagm[lt - 1 - tx - 32]
    nd_ = [0] * 64
    for tx in range(32):
        if ms[tx] < ms[tx + 32]:
            ms[tx] = ms[tx + 32]; nd_[tx] = tx + 32
        else:
            nd_[tx] = tx
    i = 16
    while i >= 1:
        old_ms, old_nd = list(ms), list(nd_)
        for tx in range(32):
            if old_ms[tx] < old_ms[tx + i]:
                ms[tx] = old_ms[tx + i]; nd_[tx] = old_nd[tx + i]
        i >>= 1
    return btck, ms[0], int(mco[lt - 1 - nd_[0]])


def flow_max(mm, lq, lt):
    """The reference's maximum and its cell (x << 16 | y) for given MM values
    mm [lq][lt]: the per-thread running maximum and the per-column buffer of
    ExecDP_Btck_Unroll32x (cudp/CuBatchDP_init_btck.cu:246-263, 328, 436-464)
    in the launch order of the block diagonals (as simulate() replays them),
    then FinalizeDPPro's reduction of the last 64 columns
    (cudp/CuBatchDP_final.cu:100-142)."""
    diagm = np.zeros(lt, f32)
    mco = np.zeros(lt, np.uint32)
    nblk = ((lt + lq) + 31) // 32 + (lq - 1) // 32
    last = 2 * ((lq - 1) // 32)
    for d in range(nblk):
        reads, writes = [], []
        for bx in range(nblk + 2):
            if d <= last:
                x0 = (2 * bx - (0 if d & 1 else 1)) * D
                y = ((d >> 1) + 1 - bx) * D - 1
            else:
                x0 = (2 * bx + (d - last - 1)) * D
                y = ((last >> 1) + 1 - bx) * D - 1
            ilim = _maxiters(x0, y, lq, lt)
            if y < 0 or lq <= y + 1 - D or lt <= x0 or ilim < 1:
                continue
            mxs = [f32(0)] * D
            mxc = [0] * D
            for tx in range(D):
                x = x0 + tx
                if 0 <= x - 1 < lt:
                    mxs[tx] = diagm[x - 1]
                    mxc[tx] = int(mco[x - 1])
            for tx in range(D):
                r = y - tx
                for i in range(ilim):
                    c = x0 + tx + i
                    if 0 <= r < lq and 0 <= c < lt and mxs[tx] < mm[r][c]:
                        mxs[tx] = f32(mm[r][c])
                        mxc[tx] = (c << 16) | r
            ndx = D - 1
            xtl = x0 + D - 1
            if xtl >= lt:
                ndx = 0
            elif xtl + ndx >= lt:
                ndx = lt - xtl - 1
            for tx in range(D):
                w = x0 + tx + ndx
                if 0 <= w < lt:
                    writes.append((w, D <= y, mxs[tx], mxc[tx]))
        # the blocks of one launch touch disjoint buffer columns: all reads
        # (above) precede all writes within a block, and blocks do not
        # share columns
        for w, merge, v, cxy in writes:
            if not merge or diagm[w] < v:
                diagm[w] = v
                mco[w] = cxy
    ms = [f32(0)] * 64
    for tx in range(32):
        if tx < lt:
            ms[tx] = diagm[lt - 1 - tx]
        if tx + 32 < lt:
            ms[tx + 32] = diagm[lt - 1 - tx - 32]
    nd_ = [0] * 64
    for tx in range(32):
        if ms[tx] < ms[tx + 32]:
            ms[tx] = ms[tx + 32]
            nd_[tx] = tx + 32
        else:
            nd_[tx] = tx
    i = 16
    while i >= 1:
        old_ms, old_nd = list(ms), list(nd_)
        for tx in range(32):
            if old_ms[tx] < old_ms[tx + i]:
                ms[tx] = old_ms[tx + i]
                nd_[tx] = old_nd[tx + i]
        i >>= 1
    return ms[0], int(mco[lt - 1 - nd_[0]])
