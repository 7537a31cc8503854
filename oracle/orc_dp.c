/* orc_dp.c: oracle restatement of phase-1 DP and its finalisation (TEST
 * INFRASTRUCTURE ONLY; see oracle.h).
 *
 * ExecDP_Btck_Unroll32x (cudp/CuBatchDP_init_btck.cu:65-494), launched over
 * block diagonals by CuBatchDP::PerformCompleteDynProgDevice
 * (cudp/CuBatchDP.cu:149-320), computes a five-state local DP (MM, MI, IM,
 * DN, ND) cell by cell.  Its cell values do not depend on the launch order,
 * but they do depend on what the 32x32 oblique blocks feed in at the
 * boundaries, which this restatement reproduces:
 *   - row -1 is all zeros (bottom buffer initialised for the first block row,
 *     :273-287);
 *   - row q is computed from column c0(q) = -1 - (q mod 32) onwards (the
 *     first block of each block row starts at x = -32, :113-121); its left
 *     neighbour at c0(q)-1 and its up neighbour at c0(q) are zero
 *     (DPLocInitCache, :262-266);
 *   - a row q with q mod 32 == 0 sees zeros from row q-1 at columns < 0
 *     (its bottom-buffer entries for negative columns are initialised, :276-286);
 *   - a row q with q mod 32 == 31 sees zeros to its left and upper left at
 *     column 0 (it is thread 0 of the block starting at x = 0, whose
 *     diagonal-buffer entries at x-1 = -1 are initialised, :262-266);
 *   - transitions of a column outside [0, Lt) and of a row outside [0, Lq)
 *     are log(0) = -32768 (DPLocInitTrnProbs, :191-212), and the score of a
 *     cell outside the valid area (outside the matrix or in the two ignored
 *     corner triangles, CellYXinValidArea, cudp/CuBatchDP_init_btck.cuh:118-170)
 *     is -9 (:241-251).
 * The maximum MM value and its cell follow the reference's flow through its
 * per-column buffer, so that equal maxima resolve as there (orc_flow_max):
 * each thread of a 32x32 block keeps the running maximum of the cells it
 * computes, seeded from the buffer at the column before its first cell and
 * written back at the block's end (:452-480), and FinalizeDPPro
 * (cudp/CuBatchDP_final.cu:100-142) reduces the buffer's last 64 columns.
 *
 * FinalizeDPPro (cudp/CuBatchDP_final.cu:61-292) is restated for the
 * pass decision, the traceback, the correlation bonus (32-score chunks
 * reduced by a 5-level xor butterfly, :194-232) and the output record.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define LOG0 (-32768.0f)
enum { MM = 0, MI, IM, DN, ND, NST };
enum { tMM = 0, tMI, tMD, tIM, tII, tDM, tDD };
enum { BT_STOP = 0, BT_LEFT = 1, BT_UP = 2, BT_DIAG = 3 };

static int corner_maxlen(int lq, int lt) {
  int m = (int)rintf(0.34f * (float)(lq > lt ? lq : lt));
  return m > 60 ? 60 : m;
}

static int in_valid_area(int lq, int lt, int edge, int y, int x) {
  int corners = (lq <= y + edge && x <= edge + y - lq) || (lt - edge + y <= x && y < edge);
  return !corners && 0 <= y && y < lq && 0 <= x && x < lt;
}

/* transitions used at one DP position: MMp, MIc(pos+1), MDp, IMp, IIc(pos+1), DMp, DDp
 * (DPLocCacheTrnProbs, cudp/CuBatchDP_init.cuh:136-155) */
static void trn_at(const orc_pset* s, int prof, int pos, int len, float out[7]) {
  if (pos < 0 || pos >= len) {
    for (int t = 0; t < 7; t++) out[t] = LOG0;
    return;
  }
  const int64_t ld = s->npos + s->nprof;
  const int64_t b = s->dist[prof] + prof + pos;
  out[tMM] = s->trn[tMM * ld + b];
  out[tMI] = s->trn[tMI * ld + b + 1];
  out[tMD] = s->trn[tMD * ld + b];
  out[tIM] = s->trn[tIM * ld + b];
  out[tII] = s->trn[tII * ld + b + 1];
  out[tDM] = s->trn[tDM * ld + b];
  out[tDD] = s->trn[tDD * ld + b];
}

static inline float mx(float a, float b) { return a < b ? b : a; }

/* MaxIterationsForBlock (cudp/CuBatchDP_init_btck.cuh): the columns a block
 * starting at x0 in the block row ending at query row y iterates over */
static int block_iters(int x0, int y, int lq, int lt) {
  int x = lt - x0;
  if (x <= 0) return 0;
  const int yy = y - (lq - 1);
  if (0 < yy) x -= yy;
  if (x <= 0) return 0;
  return x < 32 ? x : 32;
}

/* The reference's maximum of the values mm [lq][lt] and its cell
 * (x << 16 | y).  A block row (query rows 32b .. 32b+31) runs its blocks
 * x0 = -32, 0, 32, ... left to right; thread tx of a block takes query row
 * 32b+31-tx, columns x0+tx .. x0+tx+iters-1.  At the block's start each
 * thread reads the buffer at column x0+tx-1 (zero when outside [0, lt)), it
 * keeps the first cell of a larger value (strict <), and at the block's end
 * it writes the buffer at column x0+tx+ndx, ndx = 31 but 0 for a block
 * whose last column x0+31 is past the end and lt-x0-32 for the block that
 * covers the end (:452-480); the first block row overwrites, later ones keep
 * the buffer's value unless it is smaller.  For every buffer column the
 * block rows' writes and reads alternate top to bottom, so block rows are
 * replayed one after another.  FinalizeDPPro then reduces the last 64
 * columns pairwise (column lt-1-t against lt-33-t), then over halves of
 * 32 .. 1, a tie keeping the lower index, i.e. the later column.
 * The MAP alignment DP (cumapdp/CuBatchMAPDP_map_btck.cu:232-244,341-370)
 * differs in two rules, taken when edge >= 0: a thread whose first column
 * is at or past the end starts from zero, and before the write a maximum
 * whose cell lies outside the valid area (the ignored corners, edge = their
 * side) is reset to zero at cell 0.  (Its threads also carry the values of
 * cells outside the matrix, which are not represented here.) */
void orc_flow_max(const float* mm, int lq, int lt, int edge, float* maxsc, uint32_t* maxxy) {
  float* buf = (float*)calloc((size_t)lt, sizeof(float));
  uint32_t* bco = (uint32_t*)calloc((size_t)lt, sizeof(uint32_t));
  for (int b = 0; 32 * b < lq; b++) {
    const int y = 32 * b + 31;
    for (int x0 = -32; x0 < lt; x0 += 32) {
      const int iters = block_iters(x0, y, lq, lt);
      if (iters < 1) continue;
      float ms[32];
      uint32_t mc[32];
      for (int tx = 0; tx < 32; tx++) {
        const int xp = x0 + tx - 1;
        const int in = 0 <= xp && (edge < 0 ? xp < lt : xp + 1 < lt);
        ms[tx] = in ? buf[xp] : 0.0f;
        mc[tx] = in ? bco[xp] : 0u;
      }
      for (int tx = 0; tx < 32; tx++) {
        const int r = y - tx;
        if (r < 0 || lq <= r) continue;
        for (int i = 0; i < iters; i++) {
          const int c = x0 + tx + i;
          if (c < 0 || lt <= c) continue;
          const float v = mm[(size_t)r * lt + c];
          if (ms[tx] < v) { ms[tx] = v; mc[tx] = ((uint32_t)c << 16) | (uint32_t)r; }
        }
      }
      int ndx = 31;
      const int xtl = x0 + 31;
      if (lt <= xtl) ndx = 0;
      else if (lt <= xtl + ndx) ndx = lt - xtl - 1;
      for (int tx = 0; tx < 32; tx++) {
        const int w = x0 + tx + ndx;
        if (w < 0 || lt <= w) continue;
        if (0 <= edge && !in_valid_area(lq, lt, edge, (int)(mc[tx] & 0xffff), (int)(mc[tx] >> 16))) {
          ms[tx] = 0.0f;
          mc[tx] = 0u;
        }
        if (b == 0 || buf[w] < ms[tx]) { buf[w] = ms[tx]; bco[w] = mc[tx]; }
      }
    }
  }
  float v[64];
  int ix[64];
  for (int t = 0; t < 32; t++) {
    v[t] = t < lt ? buf[lt - 1 - t] : 0.0f;
    v[t + 32] = t + 32 < lt ? buf[lt - 33 - t] : 0.0f;
  }
  for (int t = 0; t < 32; t++) {
    ix[t] = t;
    if (v[t] < v[t + 32]) { v[t] = v[t + 32]; ix[t] = t + 32; }
  }
  for (int h = 16; h >= 1; h >>= 1) {
    float nv[32];
    int ni[32];
    for (int t = 0; t < 32; t++) {
      nv[t] = v[t];
      ni[t] = ix[t];
      if (t + h < 32 && v[t] < v[t + h]) { nv[t] = v[t + h]; ni[t] = ix[t + h]; }
    }
    memcpy(v, nv, sizeof nv);
    memcpy(ix, ni, sizeof ni);
  }
  *maxsc = v[0];
  *maxxy = bco[lt - 1 - ix[0]];
  free(buf);
  free(bco);
}

/* one template profile: DP with traceback bytes and the max */
static void dp_profile(const orc_pset* q, int qp, const orc_pset* db, int tp, const float* sc, int64_t ldsc,
                       int64_t col, uint8_t* bt, float* maxsc, uint32_t* maxxy) {
  const int lq = q->len[qp], lt = db->len[tp];
  const int edge = corner_maxlen(lq, lt);
  const int W = lt + 33; /* columns -33 .. lt-1 */
  float* prev = (float*)calloc((size_t)W * NST, sizeof(float));
  float* cur = (float*)calloc((size_t)W * NST, sizeof(float));
  float* dtr = (float*)malloc(sizeof(float) * 7 * (size_t)W);
  for (int c = -33; c < lt; c++) trn_at(db, tp, c, lt, dtr + 7 * (c + 33));
  float* mmv = (float*)calloc((size_t)lq * (size_t)lt, sizeof(float));
  for (int y = 0; y < lq; y++) {
    float qt[7];
    trn_at(q, qp, y, lq, qt);
    const int c0 = -1 - (y % 32);
    if (y % 32 == 0)
      for (int c = -33; c < 0; c++) memset(prev + NST * (c + 33), 0, sizeof(float) * NST);
    memset(cur, 0, sizeof(float) * NST * (size_t)W);
    for (int c = c0; c < lt; c++) {
      static const float Z[NST] = {0, 0, 0, 0, 0};
      const int zl = (c == 0 && y % 32 == 31);
      const float* U = prev + NST * (c + 33);
      const float* D = zl ? Z : prev + NST * (c + 32);
      const float* L = zl ? Z : cur + NST * (c + 32);
      float* O = cur + NST * (c + 33);
      const float* dt = dtr + 7 * (c + 33);
      const float S = in_valid_area(lq, lt, edge, y, c) ? sc[(int64_t)y * ldsc + col + c] : -9.0f;
      float v1, v2, m;
      int bk;
      v1 = D[MI] + qt[tMM] + dt[tIM];
      v2 = D[IM] + qt[tIM] + dt[tMM];
      v1 = mx(v1, v2);
      v2 = D[DN] + qt[tDM] + dt[tMM];
      v1 = mx(v1, v2);
      v2 = D[ND] + qt[tMM] + dt[tDM];
      v1 = mx(v1, v2);
      v2 = D[MM] + qt[tMM] + dt[tMM];
      v1 = mx(v1, v2);
      v1 = mx(v1 + S, 0.0f);
      O[MM] = v1;
      m = v1;
      bk = v1 ? BT_DIAG : BT_STOP;
      if (0 <= c) mmv[(size_t)y * lt + c] = v1;
      v1 = U[MM] + qt[tMM] + dt[tMI];
      v2 = U[MI] + qt[tMM] + dt[tII];
      O[MI] = mx(v1, v2);
      if (m < O[MI]) { m = O[MI]; bk = BT_UP; }
      v1 = L[MM] + qt[tMI] + dt[tMM];
      v2 = L[IM] + qt[tII] + dt[tMM];
      O[IM] = mx(v1, v2);
      if (m < O[IM]) { m = O[IM]; bk = BT_LEFT; }
      v1 = U[MM] + qt[tMD];
      v2 = U[DN] + qt[tDD];
      O[DN] = mx(v1, v2);
      if (m < O[DN]) { m = O[DN]; bk = BT_UP; }
      v1 = L[MM] + dt[tMD];
      v2 = L[ND] + dt[tDD];
      O[ND] = mx(v1, v2);
      if (m < O[ND]) { m = O[ND]; bk = BT_LEFT; }
      if (0 <= c && bt) bt[(int64_t)y * ldsc + col + c] = (uint8_t)bk;
    }
    float* t = prev; prev = cur; cur = t;
  }
  orc_flow_max(mmv, lq, lt, -1, maxsc, maxxy);
  free(mmv); free(prev); free(cur); free(dtr);
}

/* butterfly sum of 32 lane values with xor offsets 16,8,4,2,1; lane 0's result */
static float xor_tree32(const float v[32]) {
  float a[32];
  memcpy(a, v, sizeof a);
  for (int off = 16; off >= 1; off >>= 1) {
    float b[32];
    for (int l = 0; l < 32; l++) b[l] = a[l] + a[l ^ off];
    memcpy(a, b, sizeof a);
  }
  return a[0];
}

/* FinalizeDPPro for one profile; returns 1 if passed; rec gets
 * [score, orgpro, newdst, psts, beg, end] (floats / uint bit patterns) */
static int finalize_profile(const float* sc, const uint8_t* bt, int64_t ldsc, int64_t col, float maxsc,
                            uint32_t xx, float scorethld, float* score, int* psts_out, uint32_t* beg) {
  if (maxsc <= scorethld) return 0;
  int x = (int)(xx >> 16), y = (int)(xx & 0xffff);
  int psts = 0;
  float cache[36];
  memset(cache, 0, sizeof cache);
  float sum = 0.0f;
  int btck = BT_DIAG;
  while (btck != BT_STOP) {
    for (int k = 0; k < 4; k++) cache[k] = cache[32 + k];
    int ndx;
    for (ndx = 4; ndx < 36;) {
      if (x < 0 || y < 0) { btck = BT_STOP; break; }
      const int64_t qpos = (int64_t)y * ldsc + col + x;
      btck = bt[qpos];
      if (btck == BT_STOP) break;
      if (btck == BT_UP) { y--; continue; }
      if (btck == BT_LEFT) { x--; continue; }
      cache[ndx] = sc[qpos];
      if (cache[ndx] > 0.0f) psts++;
      x--; y--;
      ndx++;
    }
    float v[32];
    for (int l = 0; l < 32; l++) {
      if (ndx <= l + 4) v[l] = 0.0f;
      else v[l] = cache[l + 4] * (((cache[l] + cache[l + 1]) + cache[l + 2]) + cache[l + 3]);
    }
    sum += xor_tree32(v);
  }
  sum *= 0.5f;
  sum = sum < 0 ? -sqrtf(-sum) : sqrtf(sum);
  maxsc += sum;
  *psts_out = psts;
  *beg = ((uint32_t)(x + 1) << 16) | (uint32_t)(y + 1);
  *score = maxsc;
  return scorethld < maxsc;
}

/* Phase 1 for query qp against template profiles [p0,p1).
 * sc: [Lq][ncols] score matrix (column 0 = first position of p0);
 * bt (optional, may be NULL -> internal): [Lq][ncols] traceback bytes;
 * per profile k = p - p0: maxsc[k], maxxy[k] (x<<16|y), passed[k],
 * score[k] (with the correlation bonus), psts[k], beg[k]. */
void orc_dp_phase1(const orc_pset* q, int qp, const orc_pset* db, int p0, int p1, const float* sc, uint8_t* bt,
                   float scorethld, float* maxsc, uint32_t* maxxy, int* passed, float* score, int* psts,
                   uint32_t* beg, int nthreads) {
  const int lq = q->len[qp];
  const int64_t c0 = db->dist[p0];
  const int64_t ncols = db->dist[p1 - 1] + db->len[p1 - 1] - c0;
  uint8_t* btl = bt;
  if (!btl) btl = (uint8_t*)calloc((size_t)lq * ncols + 1, 1);
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (int p = p0; p < p1; p++) {
    const int k = p - p0;
    const int64_t col = db->dist[p] - c0;
    dp_profile(q, qp, db, p, sc, ncols, col, btl, &maxsc[k], &maxxy[k]);
    score[k] = 0.0f; psts[k] = 0; beg[k] = 0;
    passed[k] = finalize_profile(sc, btl, ncols, col, maxsc[k], maxxy[k], scorethld, &score[k], &psts[k], &beg[k]);
  }
  if (!bt) free(btl);
}

/* JobDispatcher::CalculateScoreThreshold (cualn/JobDispatcher.cpp:806-860) with
 * EVAL, gapped lambda and K */
void orc_score_threshold(float eval, float lambda, float K, uint64_t nqyposs, uint64_t dbsize, float* scorethld,
                         float* logevthld) {
  uint64_t sspace = nqyposs * dbsize;
  float ev = eval;
  if (ev <= 0.0f) { *scorethld = 999.0f; *logevthld = -99.0f; return; }
  *logevthld = logf(ev);
  if (ev < 1.0f) ev *= 10.0f; else ev *= 3.0f;
  ev /= K * (float)sspace;
  if (ev <= 0.0f) { *scorethld = 999.0f; return; }
  ev = logf(ev);
  float s = -ev / lambda;
  if (s < 0.0f) s = 0.0f;
  *scorethld = s;
}

/* AbstractScoreMatrix::ComputeLengthAdjustment (libpro/srcsco/AbstractScoreMatrix.cpp:303-395):
 * Gertz/Altschul length correction; returns the search space, delta length in *dl */
uint64_t orc_length_adjustment(float lambda, float K, float alpha, float beta, uint64_t m, uint64_t n, uint64_t N,
                               uint64_t* dl) {
  *dl = 0;
  if (lambda <= 0.0f || K <= 0.0f || N < 1) return m * n;
  const int maxit = 20;
  const float logK = logf(K);
  const float alratio = alpha / lambda;
  const float max_mn = (float)((m > n) ? m : n);
  float length = 0.0f, min_length = 0.0f, max_length = 0.0f, nxt_length = 0.0f;
  float sspace = 0.0f, space2 = 0.0f;
  int converged = 0;
  const float a = (float)N;
  const float b = (float)(n + m * N);
  const float c = (float)(m * n) - max_mn / K;
  if (c < 0.0f) return m * n;
  max_length = 2.0f * c / (b + sqrtf(b * b - 4.0f * a * c));
  for (int j = 0; j < maxit; j++) {
    length = nxt_length;
    sspace = (m - length) * (n - N * length);
    nxt_length = alratio * (logK + logf(sspace)) + beta;
    if (length <= nxt_length) {
      min_length = length;
      if (nxt_length - length <= 1.0f) { converged = 1; break; }
      if (min_length == max_length) break;
    } else
      max_length = length;
    if (nxt_length < min_length || max_length < nxt_length)
      nxt_length = (!j) ? max_length : (min_length + max_length) * 0.5f;
  }
  if (converged) {
    length = ceilf(min_length);
    if (length <= max_length) {
      space2 = (m - length) * (n - N * length);
      nxt_length = alratio * (logK + logf(space2)) + beta;
      if (length <= nxt_length) { min_length = length; sspace = space2; }
    }
  }
  *dl = (uint64_t)min_length;
  return (uint64_t)sspace;
}
