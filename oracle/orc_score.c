/* orc_score.c: oracle restatement of the score-matrix stage (TEST
 * INFRASTRUCTURE ONLY; see oracle.h).
 *
 * Reference kernels restated here:
 *   CalcSM_Init_SMEMUnroll2x             cusco/CuBatchScoreMatrix_Init.cu:45-186
 *   CalcSSSScoreSMEMUnroll2x             cusco/CuBatchScoreMatrix_ssss.cuh:49-131
 *   CalcCVS2ScoreSMEMUnroll2x            cusco/CuBatchScoreMatrix_cvs2s.cuh:49-136
 *   CalcSM_SSSS_CVS2S_SMEMUnroll2x_2     cusco/CuBatchScoreMatrix_ssss_cvs2s.cu:59-270
 *   CalcSM_DDMS2S_SMEMUnroll2x_h2<2,1,2> cusco/CuBatchScoreMatrix_ddms2s_h2.cu:61-243,
 *                                        cusco/CuBatchScoreMatrix_ddms2s_h2.cuh:55-500,
 *                                        cusco/CuBatchScoreMatrix_ddms2s.cuh:113-261
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <stdlib.h>

#include "oracle.h"

#define NAA 20
#define NCV 19
#define NSS 3
#define DDSTRIDE 128

/* ---- logarithm ------------------------------------------------------------
 * Project-written single-precision natural logarithm used in place of CUDA's
 * __logf.  Method: x = 2^e m with m in [sqrt(1/2), sqrt(2)), f = m-1 (exact),
 * s = f/(2+f), log(1+f) = 2 atanh(s) = f - s (f - R), R the truncated series
 * sum_{k>=1} 2 s^{2k}/(2k+1) up to s^10 (truncation < 1e-10 relative for
 * |s| <= 0.1716); then e ln2 added in two parts (ln2 = HI + LO, HI with 16
 * significant bits so e*HI is exact). */
float orc_logf(float x) {
  union { float f; uint32_t u; } v;
  v.f = x;
  if (!(x > 0.0f)) return x == 0.0f ? -INFINITY : NAN;
  if (x == INFINITY) return x;
  int e = 0;
  if (v.u < 0x00800000u) { v.f = x * 8388608.0f; e = -23; }
  e += (int)(v.u >> 23) - 127;
  v.u = (v.u & 0x007fffffu) | 0x3f800000u;
  if (v.u > 0x3fb504f3u) { v.f = v.f * 0.5f; e += 1; }
  const float f = v.f - 1.0f;
  const float s = f / (2.0f + f);
  const float z = s * s;
  float r = fmaf(z, 0.18181818f, 0.22222222f);
  r = fmaf(z, r, 0.28571429f);
  r = fmaf(z, r, 0.4f);
  r = fmaf(z, r, 0.66666669f);
  r = z * r;
  const float l1p = fmaf(-s, f - r, f);
  const float fe = (float)e;
  float res = fmaf(fe, 1.42860682e-06f, l1p);
  return fmaf(fe, 0.693145751953125f, res);
}

/* ---- Init: log-odds term (CuBatchScoreMatrix_Init.cu:129-144,173-185) ---- */
float orc_cell_init(const orc_pset* q, int64_t qi, int qp, const orc_pset* db, int64_t tj, int tp) {
  float s = 0.0f;
  for (int a = 0; a < NAA; a++) {
    const float f1 = q->trg[(int64_t)a * q->npos + qi];
    const float f2 = q->bkg[(int64_t)qp * NAA + a];
    const float num = f1 * db->trg[(int64_t)a * db->npos + tj];
    const float den = 0.5f * (f2 + db->bkg[(int64_t)tp * NAA + a]);
    s = s + num / den;
  }
  return orc_logf(s) - 0.002f; /* CONSTINITSHIFT_MAPALN1, Init.cuh:15 */
}

/* ---- SS term (CuBatchScoreMatrix_ssss.cuh:25-31,64-116;
 *      SerializedScoresSM.cuh:103-120) ----------------------------------- */
static float ss_lookup(const orc_tables* t, int row, int col) {
  if (row < col) { int x = row; row = col; col = x; }
  if (row) col += row + ((row * (row - 1)) >> 1);
  return t->ss[col + 2];
}

float orc_cell_ss(const orc_tables* t, const orc_pset* q, int64_t qi, const orc_pset* db, int64_t tj) {
  float sc = 0.0f;
  for (int i = 0; i < NSS; i++) {
    const float f1 = q->ss[(int64_t)i * q->npos + qi];
    const int row = i * 10 + (int)(f1 * 9.999f);
    for (int j = 0; j < NSS; j++) {
      const float f2 = db->ss[(int64_t)j * db->npos + tj];
      const int col = j * 10 + (int)(f2 * 9.999f);
      sc = fmaf(f1 * f2, ss_lookup(t, row, col), sc);
    }
  }
  const float f1 = q->ss[(int64_t)(NSS - 1) * q->npos + qi];
  if (-99.f < sc) {
    const float f2 = db->ss[(int64_t)(NSS - 1) * db->npos + tj];
    sc = fmaf((f1 + 0.1f) * (f2 + 0.1f), 0.1111111f, sc);
    sc -= 0.05f;   /* CONSTSSSSSHIFT */
    sc *= 0.12f;   /* SSSWGT */
  }
  return sc;
}

/* ---- CVS term (CuBatchScoreMatrix_cvs2s.cuh:80-113;
 *      SerializedCVS2Scores.cuh:377-409) ------------------------------------ */
float orc_cell_cvs(const orc_tables* t, const orc_pset* q, int64_t qi, const orc_pset* db, int64_t tj) {
  float sc = 0.0f;
  for (int i = 0; i < NCV; i++)
    sc = fmaf(q->cv[(int64_t)i * q->npos + qi], db->cv[(int64_t)i * db->npos + tj], sc);
  sc += t->cvkappa0inv;
  const float f1 = q->cvnorm2[qi];
  const float a = (f1 + 1.0f + t->cvkappa0inv) * (db->cvnorm2[tj] + 1.0f + t->cvkappa0inv);
  sc = fmaf(-sc, sc, a);
  sc = fmaf(t->cvpowernu0, orc_logf(sc), t->cvcterm) - q->cvprior[qi] - db->cvprior[tj];
  /* GetScoreE1S1Step2Boundary */
  float v;
  if (sc <= t->cvkeys[0])
    v = t->cvvals[0];
  else if (t->cvkeys[t->cvcard - 1] <= sc)
    v = t->cvvals[t->cvcard - 1];
  else
    v = t->cvvals[t->cvshift + (int)rintf(2.0f * sc)];
  return v * 0.15f; /* CVSWGT */
}

/* ---- DDMS term --------------------------------------------------------------
 * pair score s(d1,d2) of DMS_AS_2 (SerializedDstMatchScores.cuh:350-378):
 * the larger distance r, the smaller c (both saturated at 20), 255 -> -1 */
static inline float dms(const float* tab, int d1, int d2) {
  return tab[d1 * 22 + d2];
}
static inline int sat(int d) { return d == 255 ? 21 : (d < 20 ? d : 20); }

/* byte-wise max/min of packed [QB QE PB PE] coordinates (CoordsIntersect,
 * CuBatchScoreMatrix_ddms2s.cuh:141-153) */
static int coords_intersect(uint32_t c1, uint32_t c2) {
  uint32_t mx = 0, mn = 0;
  for (int b = 0; b < 32; b += 8) {
    uint32_t x = (c1 >> b) & 0xff, y = (c2 >> b) & 0xff;
    mx |= (x > y ? x : y) << b;
    mn |= (x < y ? x : y) << b;
  }
  return ((mn >> 16) & 0xff) >= ((mx >> 24) & 0xff) || (mn & 0xff) >= ((mx >> 8) & 0xff);
}

/* RegisterMaxDDMScore (CuBatchScoreMatrix_ddms2s.cuh:168-261) */
static void register_max(float m[3], uint32_t c[3], float s, uint32_t sc) {
  if (!s) return;
  if (m[0] <= s) {
    m[2] = m[1]; c[2] = c[1];
    m[1] = m[0]; c[1] = c[0];
    m[0] = s; c[0] = sc;
    if (m[1] && coords_intersect(c[0], c[1])) m[1] = 0.0f;
    if (m[2] && coords_intersect(c[0], c[2])) m[2] = 0.0f;
    if (m[1] == 0.0f) { m[1] = m[2]; c[1] = c[2]; m[2] = 0.0f; }
  } else if (m[1] <= s) {
    if (coords_intersect(sc, c[0])) return;
    m[2] = m[1]; c[2] = c[1];
    m[1] = s; c[1] = sc;
    if (coords_intersect(c[1], c[2])) m[2] = 0.0f;
  } else if (m[2] <= s) {
    if (coords_intersect(sc, c[0]) || coords_intersect(sc, c[1])) return;
    m[2] = s; c[2] = sc;
  }
}

#define MKC(qb, qe, pb, pe) \
  ((((uint32_t)(qb) & 0xff) << 24) | (((uint32_t)(qe) & 0xff) << 16) | (((uint32_t)(pb) & 0xff) << 8) | ((uint32_t)(pe) & 0xff))

/* distance value k (0-based) of a list record; the list is padded with 255 */
static inline int dv(const uint8_t* rec, int k) { return k < 127 ? sat(rec[1 + k]) : 21; }

/* CalcDDMS2ScoreSMEMh2<2,true,2> (CuBatchScoreMatrix_ddms2s_h2.cuh:435-500) with
 * ProcessDiagonalUpperh2/Lowerh2 (:282-406) and UpdateSegScore{Upper,Lower}h2 (:133-181) */
float orc_ddms_tab(const float* t, const uint8_t* qrec, const uint8_t* trec, float m[3], uint32_t c[3]) {
  const int len1 = qrec[0], len2 = trec[0];
  const int dstsegm = 6;
  m[0] = m[1] = m[2] = 0.0f;
  c[0] = c[1] = c[2] = 0xffffffffu;
  if (len1 < dstsegm || len2 < dstsegm) return 0.0f;
  float sc[5];
  /* upper part: diagonals j = len2>>2 ... 0, step 2 */
  for (int j = len2 >> 2; 0 <= j; j -= 2) {
    int segqb = 0, segpb = j;
    sc[0] = dms(t, dv(qrec, 0), dv(trec, j));
    float segsco = sc[0] < 0.0f ? 0.0f : sc[0];
    for (int i = 0; i < len1 && j + i < len2; i += 4) {
      float scom = 0.0f;
      int im = 0, jim = 0, ib = 0, jib = 0;
      sc[2] = dms(t, dv(qrec, i + 2), dv(trec, j + i + 2));
      sc[4] = dms(t, dv(qrec, i + 4), dv(trec, j + i + 4));
      sc[3] = 0.5f * (sc[2] + sc[4]);
      sc[1] = 0.5f * (sc[0] + sc[2]);
      for (int k = 1; k <= 4; k++) {
        if (!segsco) { segqb = i + k; segpb = j + i + k; }
        const float x = segsco + sc[k];
        segsco = x < 0.0f ? 0.0f : x;
        if (segsco && scom < segsco && dstsegm <= i + k - segqb + 1) {
          scom = segsco; im = i + k; jim = j + i + k; ib = segqb; jib = segpb;
        }
      }
      register_max(m, c, scom, MKC(ib, im, jib, jim));
      sc[0] = sc[4];
    }
  }
  /* lower part: diagonals i = len1>>2 ... 1, step 2 */
  for (int i = len1 >> 2; 0 < i; i -= 2) {
    int segqb = i, segpb = 0;
    sc[0] = dms(t, dv(qrec, i), dv(trec, 0));
    float segsco = sc[0] < 0.0f ? 0.0f : sc[0];
    for (int j = 0; j < len2 && i + j < len1; j += 4) {
      float scom = 0.0f;
      int ijm = 0, jm = 0, ijb = 0, jb = 0;
      sc[2] = dms(t, dv(qrec, i + j + 2), dv(trec, j + 2));
      sc[4] = dms(t, dv(qrec, i + j + 4), dv(trec, j + 4));
      sc[3] = 0.5f * (sc[2] + sc[4]);
      sc[1] = 0.5f * (sc[0] + sc[2]);
      for (int k = 1; k <= 4; k++) {
        if (!segsco) { segqb = i + j + k; segpb = j + k; }
        const float x = segsco + sc[k];
        segsco = x < 0.0f ? 0.0f : x;
        if (segsco && scom < segsco && dstsegm <= j + k - segpb + 1) {
          scom = segsco; ijm = i + j + k; jm = j + k; ijb = segqb; jb = segpb;
        }
      }
      register_max(m, c, scom, MKC(ijb, ijm, jb, jm));
      sc[0] = sc[4];
    }
  }
  float s = 0.0f;
  s += m[0];
  s += m[1];
  s += m[2];
  s = fmaf(0.053f * 16.0f, s, -1.235f);
  return s * 0.2f; /* DDMSWGT */
}

float orc_cell_ddms_ex(const orc_tables* t, const uint8_t* qrec, const uint8_t* trec, float m[3], uint32_t c[3]) {
  return orc_ddms_tab(t->dms, qrec, trec, m, c);
}

float orc_cell_ddms(const orc_tables* t, const uint8_t* qrec, const uint8_t* trec) {
  float m[3];
  uint32_t c[3];
  return orc_cell_ddms_ex(t, qrec, trec, m, c);
}

/* ---- full matrix (kernel sequence of CuBatchScoreMatrix.cu:236-396) ------ */
void orc_score_matrix(const orc_tables* t, const orc_pset* q, int qp, const orc_pset* db, int p0, int p1,
                      float* scores, float* modscores, int nthreads) {
  const int lq = q->len[qp];
  const int64_t q0 = q->dist[qp];
  const int64_t c0 = db->dist[p0];
  const int64_t ncols = (p1 > p0) ? db->dist[p1 - 1] + db->len[p1 - 1] - c0 : 0;
  if (nthreads < 1) nthreads = 1;
  /* profile index of every column */
  int* prof = (int*)malloc(sizeof(int) * (size_t)(ncols > 0 ? ncols : 1));
  for (int p = p0; p < p1; p++)
    for (int m = 0; m < db->len[p]; m++) prof[db->dist[p] + m - c0] = p;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
  for (int64_t col = 0; col < ncols; col++) {
    const int p = prof[col];
    const int64_t tj = c0 + col;
    for (int r = 0; r < lq; r++) {
      const int64_t qi = q0 + r;
      float s = orc_cell_init(q, qi, qp, db, tj, p);
      const float ssc = orc_cell_ss(t, q, qi, db, tj) + orc_cell_cvs(t, q, qi, db, tj);
      s = s + ssc;
      float ms = fmaf(-0.5f, 0.15f, ssc); /* score - CONSTCVSSHIFT*cvswgt */
      const float d = orc_cell_ddms(t, q->dd + (size_t)qi * DDSTRIDE, db->dd + (size_t)tj * DDSTRIDE);
      if (d) {
        s = s + d;
        ms = ms + fmaf(-0.2f, 0.2f, d); /* score - CONSTDDMSSHIFT*ddmswgt */
      }
      scores[(int64_t)r * ncols + col] = s;
      modscores[(int64_t)r * ncols + col] = ms;
    }
  }
  free(prof);
}
