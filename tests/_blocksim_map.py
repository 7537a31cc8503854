"""Step-by-step replay of the reference's MAP-alignment block schedule.

A test helper (CPU, small matrices only) that replays in float32 the data
movement of the three MAP kernels as CuBatchMAPDP::CalculateFwdBwdProbs
MAPDynProgDevice launches them over block diagonals
(cumapdp/CuBatchMAPDP.cu:182-392):
  - ExecMAPDP_Fwd_Unroll32x (cumapdp/CuBatchMAPDP_fwd.cu:82-478): forward
    log probabilities, the running maximum kept in the per-column max buffer;
  - ExecMAPDP_Bwd_Unroll32x (cumapdp/CuBatchMAPDP_bwd.cu:82-545): backward
    log probabilities added to the forward ones, with the sum of
    exp(forward - max) accumulated for the partition function
    (MAPDP_UpdateForTotalProbability, CuBatchMAPDP_bwd.cuh:225-262);
  - ExecMAPDP_MAP_Btck_Unroll32x (cumapdp/CuBatchMAPDP_map_btck.cu:75-386):
    the MAP alignment DP with traceback bytes and the max buffer,
followed by the max reduction of FinalizeMAPDP_MAP
(cumapdp/CuBatchMAPDP_map_final.cu:167-210): oblique 32x32 blocks, their
diagonal / bottom / max buffers and per-thread state.  It exists to check the
oracle's boundary rules, which are derived from this schedule.

Schraudolph's exponential and its inverse (Neural Computation 11, 1999),
as the reference parametrises them (cucom/myfastapproxmath.cuh:12-41), are
evaluated here with exact rational arithmetic for the fused multiply-add.
"""
from fractions import Fraction

import numpy as np

f32 = np.float32
D = 32
LP0 = f32(-99.0)          # log of probability 0 (MAPDPLocGetLogProb0)
MM, MI, IM, DN, ND = range(5)
tMM, tMI, tMD, tIM, tII, tDM, tDD = range(7)
_C1 = 12102203            # float32(12102203.16)
_C0 = 1064631168          # float32(1064631197.0)
_LF = f32(8.2629582959339e-08)


def _round_f32(s):
    """Round a positive Fraction to the nearest float32 (ties to even)."""
    e = s.numerator.bit_length() - s.denominator.bit_length()
    if Fraction(2) ** e > s:
        e -= 1
    ulp = Fraction(2) ** (e - 23)
    q = s / ulp
    n = q.numerator // q.denominator
    rem = q - n
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and n % 2 == 1):
        n += 1
    return n * ulp


def fastexp(a):
    """MAPDPexpf: 0 unless -16 < a, else the bits int(fma(12102203.16, a, 1064631197))."""
    a = f32(a)
    if not (f32(-16.0) < a):
        return f32(0.0)
    s = _round_f32(Fraction(float(a)) * _C1 + _C0)
    i = int(s)  # truncation toward zero (s > 0)
    return np.array([i], np.int32).view(np.float32)[0]


def fastlog(a):
    """MAPDPlogf: -99 unless 3.8e-11 < a, else (bits(a) - 1064631197) * 8.2629582959339e-08."""
    a = f32(a)
    if not (f32(3.8e-11) < a):
        return f32(-99.0)
    bits = int(np.array([a], np.float32).view(np.int32)[0])
    return f32(f32(bits - 1064631197) * _LF)


def _mx(a, b):
    return b if a < b else a


def _trn_fwd(trn, base, pos, length, ld):
    """MAPDPLocCacheTrnProbs (CuBatchMAPDP_fwd.cuh:97-137): p = vector pos, c = pos+1."""
    if pos < 0 or pos >= length:
        return [LP0] * 7
    b = base + pos
    return [f32(trn[tMM * ld + b]), f32(trn[tMI * ld + b + 1]), f32(trn[tMD * ld + b]), f32(trn[tIM * ld + b]),
            f32(trn[tII * ld + b + 1]), f32(trn[tDM * ld + b]), f32(trn[tDD * ld + b])]


def _trn_bwd(trn, base, pos, length, ld):
    """MAPDP_Bwd_LocCacheTrnProbs (CuBatchMAPDP_bwd.cuh:93-133): all at vector pos+1."""
    if pos < 0 or pos >= length:
        return [LP0] * 7
    b = base + pos + 1
    return [f32(trn[k * ld + b]) for k in range(7)]


def _corner(lq, lt):
    return min(int(np.rint(f32(0.34) * f32(max(lq, lt)))), 60)


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


def _blocks(lq, lt, reverse=False):
    """Participating blocks (d, x0, y, ilim) in launch order."""
    nblk = ((lt + lq) + 31) // 32 + (lq - 1) // 32
    last = 2 * ((lq - 1) // 32)
    ds = range(nblk - 1, -1, -1) if reverse else range(nblk)
    for d in ds:
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
            yield d, nblk, x0, y, ilim


def _lse6(vals):
    m = f32(0.0)
    for v in vals:
        m = _mx(m, v)
    s = f32(0.0)
    for v in vals:
        s = f32(s + fastexp(f32(v - m)))
    s = f32(s + fastexp(f32(-m)))
    return f32(fastlog(s) + m)


def _lse2(v1, v2):
    m = _mx(v1, v2)
    s = f32(fastexp(f32(v1 - m)) + fastexp(f32(v2 - m)))
    return f32(fastlog(s) + m)


def simulate_fwd(S, qtrn, qbase, qld, lq, dtrn, dbase, dld, lt):
    """S: [lq][lt] float32 MAP scores.  Returns (fwd MM [lq][lt], diagM buffer)."""
    qt = {r: _trn_fwd(qtrn, qbase, r, lq, qld) for r in range(-64, lq + 64)}
    dtc = {}

    def dt(c):
        if c not in dtc:
            dtc[c] = _trn_fwd(dtrn, dbase, c, lt, dld)
        return dtc[c]

    diag1 = np.zeros((5, lt), f32)
    diag2 = np.zeros((5, lt), f32)
    diagm = np.full(lt, f32(12345.0), f32)  # leftover content of the buffer: must be overwritten
    bot = np.zeros((5, lt), f32)
    fwd = np.zeros((lq, lt), f32)
    for d, nblk, x0, y, ilim in _blocks(lq, lt):
        sc = [[LP0] * D for _ in range(D)]  # sc[tx][i] for cell (y-tx, x0+tx+i)
        for tx in range(D):
            for i in range(D):
                r, cc = y - tx, x0 + tx + i
                if 0 <= r < lq and 0 <= cc < lt:
                    sc[tx][i] = f32(S[r][cc])
        p1 = np.full((5, D + 1), LP0, f32)
        p2 = np.full((5, D + 1), LP0, f32)
        mxl = [f32(0.0)] * D
        bottm = np.full((5, D), LP0, f32)
        for tx in range(D):
            x = x0 + tx
            if 0 <= x - 1 and x < lt:
                p1[:, tx] = diag1[:, x - 1]
                p2[:, tx + 1] = diag2[:, x - 1]
                mxl[tx] = diagm[x - 1]
            if D <= y and 0 <= x + D - 1 < lt:
                bottm[:, tx] = bot[:, x + D - 1]
        res = {}
        for i in range(ilim):
            p1[:, D] = bottm[:, i]
            new = np.zeros((5, D), f32)
            for tx in range(D):
                q = qt[y - tx]
                dd = dt(x0 + tx + i)
                g = p2[:, tx + 1]
                u = p1[:, tx + 1]
                l_ = p1[:, tx]
                s = sc[tx][i]
                v1 = f32(f32(f32(g[MI] + q[tMM]) + dd[tIM]) + s)
                v2 = f32(f32(f32(g[IM] + q[tIM]) + dd[tMM]) + s)
                v3 = f32(f32(f32(g[DN] + q[tDM]) + dd[tMM]) + s)
                v4 = f32(f32(f32(g[ND] + q[tMM]) + dd[tDM]) + s)
                v5 = f32(f32(f32(g[MM] + q[tMM]) + dd[tMM]) + s)
                mm = _lse6([v1, v2, v3, v4, v5])
                mxl[tx] = _mx(mxl[tx], mm)
                mi = _lse2(f32(f32(u[MM] + q[tMM]) + dd[tMI]), f32(f32(u[MI] + q[tMM]) + dd[tII]))
                im = _lse2(f32(f32(l_[MM] + q[tMI]) + dd[tMM]), f32(f32(l_[IM] + q[tII]) + dd[tMM]))
                dn = _lse2(f32(u[MM] + q[tMD]), f32(u[DN] + q[tDD]))
                nd = _lse2(f32(l_[MM] + dd[tMD]), f32(l_[ND] + dd[tDD]))
                new[:, tx] = (mm, mi, im, dn, nd)
                res[(tx, i)] = mm
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
                    if diagm[x + ndx] < mxl[tx]:
                        diagm[x + ndx] = mxl[tx]
                else:
                    diagm[x + ndx] = mxl[tx]
            if 0 <= x < lt:
                bot[:, x] = bottm[:, tx]
            for i in range(D):
                r, cc = y - i, x + i
                if 0 <= r < lq and 0 <= cc < lt:
                    fwd[r][cc] = res[(i, tx)]
    return fwd, diagm


def simulate_bwd(S, qtrn, qbase, qld, lq, dtrn, dbase, dld, lt, fwd, diagm):
    """Backward pass over the forward matrix.  Returns (bwd MM [lq][lt],
    post = fwd + bwd [lq][lt], maxlogprob, zsum in the launch order)."""
    qt = {r: _trn_bwd(qtrn, qbase, r, lq, qld) for r in range(-64, lq + 64)}
    dtc = {}

    def dt(c):
        if c not in dtc:
            dtc[c] = _trn_bwd(dtrn, dbase, c, lt, dld)
        return dtc[c]

    diag1 = np.full((5, lt), f32(777.0), f32)
    diag2 = np.full((5, lt), f32(777.0), f32)
    bot = np.full((5, lt), f32(777.0), f32)
    bwd = np.zeros((lq, lt), f32)
    post = fwd.copy()
    maxlog = None
    zsum = f32(0.0)
    nprof_diags = ((lt + lq) + 31) // 32 + (lq - 1) // 32
    for d, nblk, x0, y, ilim in _blocks(lq, lt, reverse=True):
        first = d + 1 == nprof_diags
        if first:
            lanes = [f32(0.0)] * D
            for tx in range(D):
                if tx < lt:
                    lanes[tx] = diagm[lt - 1 - tx]
                if tx + D < lt:
                    lanes[tx] = _mx(lanes[tx], diagm[lt - 1 - tx - D])
            m = lanes[0]
            for v in lanes[1:]:
                m = _mx(m, v)
            maxlog = m
        # MAPDP_UpdateForTotalProbability: per line a shfl_down tree, then sequential over lines
        bsum = f32(0.0)
        for i in range(D):
            lv = [f32(0.0)] * D
            for tx in range(D):
                r, cc = y - i, x0 + tx + i
                if 0 <= r < lq and 0 <= cc < lt:
                    lv[tx] = fastexp(f32(fwd[r][cc] - maxlog))
            off = 16
            while off >= 1:
                lv = [f32(lv[t] + lv[t + off]) if t + off < D else lv[t] for t in range(D)]
                off >>= 1
            bsum = f32(bsum + lv[0])
        zsum = f32(zsum + bsum)
        sc = [[LP0] * D for _ in range(D)]  # sc[tx][i] = score of (y+1-tx, x0+tx+1+i)
        for tx in range(D):
            for i in range(D):
                r, cc = y + 1 - tx, x0 + tx + 1 + i
                if 0 <= r < lq and 0 <= cc < lt:
                    sc[tx][i] = f32(S[r][cc])
        p1 = np.full((5, D + 1), LP0, f32)
        p2 = np.full((5, D + 1), LP0, f32)
        bottm = np.full((5, D), LP0, f32)
        for tx in range(D):
            x = x0 + tx
            if not first and 0 <= x + D < lt:
                p1[:, tx + 1] = diag1[:, x + D]
                p2[:, tx + 1] = diag2[:, x + D]
            if y + 1 < lq and 0 <= x < lt:
                bottm[:, tx] = bot[:, x]
        if not first and y + 1 < lq and 0 <= x0 + D < lt:
            p2[:, 0] = bot[:, x0 + D]
        res = {}
        for i in range(ilim - 1, -1, -1):
            p1[:, 0] = bottm[:, i]
            new = np.zeros((5, D), f32)
            for tx in range(D):
                q = qt[y - tx]
                dd = dt(x0 + tx + i)
                dn_ = p1[:, tx]       # (r+1, c)
                rt = p1[:, tx + 1]    # (r, c+1)
                dg = p2[:, tx]        # (r+1, c+1)
                s = sc[tx][i]
                v1 = f32(f32(dn_[MI] + q[tMM]) + dd[tMI])
                v2 = f32(f32(rt[IM] + q[tMI]) + dd[tMM])
                v3 = f32(dn_[DN] + q[tMD])
                v4 = f32(rt[ND] + dd[tMD])
                v5 = f32(f32(f32(dg[MM] + q[tMM]) + dd[tMM]) + s)
                mm = _lse6([v1, v2, v3, v4, v5])
                mi = _lse2(f32(f32(f32(dg[MM] + q[tMM]) + dd[tIM]) + s), f32(f32(dn_[MI] + q[tMM]) + dd[tII]))
                im = _lse2(f32(f32(f32(dg[MM] + q[tIM]) + dd[tMM]) + s), f32(f32(rt[IM] + q[tII]) + dd[tMM]))
                dn = _lse2(f32(f32(f32(dg[MM] + q[tDM]) + dd[tMM]) + s), f32(dn_[DN] + q[tDD]))
                nd = _lse2(f32(f32(f32(dg[MM] + q[tMM]) + dd[tDM]) + s), f32(rt[ND] + dd[tDD]))
                new[:, tx] = (mm, mi, im, dn, nd)
                res[(tx, i)] = mm
            p2[:, 1:] = new
            bottm[:, i] = new[:, D - 1]
            p1, p2 = p2, p1
        for tx in range(D):
            x = x0 + tx
            if 0 <= x < lt:
                diag1[:, x] = p1[:, tx + 1]
                diag2[:, x] = p2[:, tx + 1]
            if 0 <= x + D - 1 < lt:
                bot[:, x + D - 1] = bottm[:, tx]
            for i in range(D):
                r, cc = y - i, x + i
                if 0 <= r < lq and 0 <= cc < lt and (i, tx) in res:
                    bwd[r][cc] = res[(i, tx)]
                    post[r][cc] = f32(post[r][cc] + res[(i, tx)])
    return bwd, post, maxlog, zsum


def simulate_map(post, lq, lt, logz, minpp=f32(0.28)):
    """MAP alignment DP over the posterior matrix.  Returns (btck [lq][lt],
    maxsc, maxcoords) after FinalizeMAPDP_MAP's reduction."""
    gap = f32(minpp * f32(0.5))
    e = _corner(lq, lt)
    diag1 = np.zeros(lt, f32)
    diag2 = np.zeros(lt, f32)
    diagm = np.full(lt, f32(-5.0), f32)
    mco = np.zeros(lt, np.uint32)
    bot = np.zeros(lt, f32)
    btck = np.zeros((lq, lt), np.uint8)
    for d, nblk, x0, y, ilim in _blocks(lq, lt):
        sc = [[f32(0.0)] * D for _ in range(D)]
        for tx in range(D):
            for i in range(D):
                r, cc = y - tx, x0 + tx + i
                if _valid(lq, lt, e, r, cc):
                    sc[tx][i] = f32(post[r][cc])
        p1 = np.zeros(D + 1, f32)
        p2 = np.zeros(D + 1, f32)
        mxs = [f32(0.0)] * D
        mxc = [0] * D
        bottm = np.zeros(D, f32)
        for tx in range(D):
            x = x0 + tx
            if 0 <= x - 1 and x < lt:
                p1[tx] = diag1[x - 1]
                p2[tx + 1] = diag2[x - 1]
                mxs[tx] = diagm[x - 1]
                mxc[tx] = int(mco[x - 1])
            if D <= y and 0 <= x + D - 1 < lt:
                bottm[tx] = bot[x + D - 1]
        bt = np.zeros((D, D), np.uint8)
        for i in range(ilim):
            p1[D] = bottm[i]
            new = np.zeros(D, f32)
            for tx in range(D):
                v = f32(f32(p2[tx + 1] + fastexp(f32(sc[tx][i] - logz))) - minpp)
                b = 3
                if not (v > f32(0.0)):
                    v, b = f32(0.0), 0
                if mxs[tx] < v:
                    mxs[tx] = v
                    mxc[tx] = ((x0 + tx + i) << 16) | ((y - tx) & 0xffff)
                w = f32(p1[tx + 1] - gap)
                if v < w:
                    v, b = w, 2
                w = f32(p1[tx] - gap)
                if v < w:
                    v, b = w, 1
                new[tx] = v
                bt[tx][i] = b
            p2[:D] = new
            bottm[i] = new[0]
            p1, p2 = p2, p1
        for tx in range(D):
            x = x0 + tx
            if 0 <= x + D - 1 < lt:
                diag1[x + D - 1] = p1[tx]
                diag2[x + D - 1] = p2[tx + 1]
            ndx = D - 1
            xtl = x0 + D - 1
            if xtl >= lt:
                ndx = 0
            elif xtl + ndx >= lt:
                ndx = lt - xtl - 1
            if 0 <= x + ndx < lt:
                cy, cx = mxc[tx] & 0xffff, mxc[tx] >> 16
                if not _valid(lq, lt, e, cy, cx):
                    mxs[tx], mxc[tx] = f32(0.0), 0
                if D <= y:
                    if diagm[x + ndx] < mxs[tx]:
                        diagm[x + ndx] = mxs[tx]; mco[x + ndx] = mxc[tx]
                else:
                    diagm[x + ndx] = mxs[tx]; mco[x + ndx] = mxc[tx]
            if 0 <= x < lt:
                bot[x] = bottm[tx]
            for i in range(D):
                r, cc = y - i, x + i
                if 0 <= r < lq and 0 <= cc < lt:
                    btck[r][cc] = bt[i][tx]
    # FinalizeMAPDP_MAP: max with index propagation over the last 64 columns
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
