/* orc_tables.c: oracle readers for the reference's score tables (TEST
 * INFRASTRUCTURE ONLY; see oracle.h).
 *
 *   ssss  : SSSScores::ReadLevels/ReadScoreTables (libpro/srcpro/SSSScores.cpp:66-222),
 *           VirtScores::ReadScoresHelper (libpro/srcpro/VirtScores.cpp:44-140),
 *           serialized as SerializedScoresCtor does (2 header cells, then the
 *           lower triangle row by row).
 *   cvs2s : CVS2Scores::ReadLevels/ReadScoreTables/ReadScoresHelper/SmoothScores
 *           (libpro/srcpro/CVS2Scores.cpp:95-466).
 *   CVS constants: _TCTXVECT (libpro/srcpro/TCTXVECT.cpp:26-60).
 *   DDMS pair scores: SerializedDstMatchScoresSM<float,DMS_AS_2>::GetScore
 *           (cupro/SerializedDstMatchScores.cuh:350-378).
 */
#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* next non-comment, non-empty line (liblib/myfiler.cpp skip_comments, '#') */
static int next_line(FILE* fp, char* buf, int n) {
  while (fgets(buf, n, fp)) {
    char* p = buf;
    while (*p == ' ' || *p == '\t') p++;
    if (*p == '#' || *p == '\n' || *p == '\r' || *p == 0) continue;
    return 1;
  }
  return 0;
}

static int load_ssss(const char* path, orc_tables* t) {
  FILE* fp = fopen(path, "r");
  char buf[8192];
  if (!fp) return -1;
  /* "ssss= 0 +1": one probability level, one ENO level -> one table "0000 11" */
  if (!next_line(fp, buf, sizeof buf) || !strstr(buf, "ssss=")) { fclose(fp); return -2; }
  if (!next_line(fp, buf, sizeof buf) || strncmp(buf, "0000 11", 7)) { fclose(fp); return -3; }
  if (!next_line(fp, buf, sizeof buf) || !strstr(buf, "Cardinality =")) { fclose(fp); return -4; }
  int card = atoi(strstr(buf, "=") + 1);
  if (card != 30) { fclose(fp); return -5; }
  t->ss[0] = 0.0f;  /* probability level value */
  t->ss[1] = 1.0f;  /* ENO level value */
  int k = 2;
  for (int r = 0; r < card; r++) {
    if (!next_line(fp, buf, sizeof buf)) { fclose(fp); return -6; }
    char* p = buf;
    for (int c = 0; c <= r; c++) {
      char* e;
      float v = strtof(p, &e);
      if (e == p) { fclose(fp); return -7; }
      t->ss[k++] = v;
      p = e;
    }
  }
  fclose(fp);
  return 0;
}

/* SmoothScores (CVS2Scores.cpp:366-466) for step 2 */
static void smooth_cvs(float* keys, float* scos, int* n, int* card, int* shift) {
  const float plusval = 9999.0f;
  const float incr = 1.0f / 2.0f;
  const int nSWIN = 5, nSWINh = nSWIN >> 1, nSLAST = 9999;
  float smth[256], rwin[5] = {0, 0, 0, 0, 0};
  float rsum = 0.0f, avg = 0.0f, sm = 0.0f, pkk = 0.0f;
  int sz = *n;
  int nv = 0, rn = 0, nn = 0;
  for (int i = 0; i < sz; i++) smth[i] = 0.0f;
  for (int i = 0; i < sz; i++) {
    float sc = scos[i];
    if (nv < nSWINh) nn++;
    if (sc != plusval) {
      if (nSWIN <= rn) rn = 0;
      rsum -= rwin[rn];
      rwin[rn++] = sc;
      rsum += sc;
      avg = rsum / (float)nSWIN;
      nv++;
    }
    if (nSWIN <= nv) {
      if (sc != plusval) {
        if (nSWIN == nv)
          for (int p = 0; p < nn; p++) smth[p] = avg;
        smth[nn] = avg;
        for (++nn; scos[nn] == plusval && nn < i; nn++) smth[nn] = avg;
      }
      if (sz == i + 1)
        for (int p = nn; p < sz; p++) smth[p] = avg;
    }
  }
  for (int i = sz - 1; 0 <= i && sz - nSLAST <= i; i--) {
    sm = smth[i];
    scos[i] = sm;
  }
  /* insert missing keys (none for a gapless key line) and replace '+' */
  for (int i = 0; i < sz; i++) {
    float kk = keys[i], sc = scos[i];
    if (i && incr < kk - pkk) {
      pkk += incr;
      memmove(keys + i + 1, keys + i, sizeof(float) * (sz - i));
      memmove(scos + i + 1, scos + i, sizeof(float) * (sz - i));
      memmove(smth + i + 1, smth + i, sizeof(float) * (sz - i));
      keys[i] = pkk; scos[i] = sm; smth[i] = sm;
      sz++;
      continue;
    }
    pkk = kk;
    sm = smth[i];
    if (sc == plusval) scos[i] = sm;
  }
  *n = sz;
  *card = sz;
  *shift = 0;
  for (int i = 0; i < sz; i++) {
    if (0.0f <= keys[i]) break;
    (*shift)++;
  }
}

static int load_cvs2s(const char* path, orc_tables* t) {
  FILE* fp = fopen(path, "r");
  static char buf[65536];
  if (!fp) return -1;
  if (!next_line(fp, buf, sizeof buf) || !strstr(buf, "cvs2s=")) { fclose(fp); return -2; }
  if (!next_line(fp, buf, sizeof buf) || strncmp(buf, "11", 2)) { fclose(fp); return -3; }
  if (!next_line(fp, buf, sizeof buf)) { fclose(fp); return -4; }
  if (strstr(buf, "Scale =")) { if (!next_line(fp, buf, sizeof buf)) { fclose(fp); return -4; } }
  if (!strstr(buf, "Dimensions =")) { fclose(fp); return -5; }
  if (!next_line(fp, buf, sizeof buf) || !strstr(buf, "Cardinality =")) { fclose(fp); return -6; }
  int rcard = atoi(strstr(buf, "=") + 1);
  if (rcard <= 0 || rcard > 100) { fclose(fp); return -7; }
  float keys[256], scos[256];
  if (!next_line(fp, buf, sizeof buf)) { fclose(fp); return -8; }
  char* p = buf;
  for (int c = 0; c < rcard; c++) {
    char* e;
    keys[c] = strtof(p, &e);
    if (e == p) { fclose(fp); return -9; }
    p = e;
  }
  if (!next_line(fp, buf, sizeof buf)) { fclose(fp); return -10; }
  p = buf;
  for (int c = 0; c < rcard; c++) {
    while (*p == ' ' || *p == '\t') p++;
    if (*p == '+') { scos[c] = 9999.0f; p++; continue; }
    char* e;
    scos[c] = strtof(p, &e);
    if (e == p) { fclose(fp); return -11; }
    p = e;
  }
  fclose(fp);
  int n = rcard, card = 0, shift = 0;
  smooth_cvs(keys, scos, &n, &card, &shift);
  if (card > 128) return -12;
  t->cvcard = card;
  t->cvshift = shift;
  for (int i = 0; i < card; i++) { t->cvkeys[i] = keys[i]; t->cvvals[i] = scos[i]; }
  return 0;
}

/* _TCTXVECT constants: DIM 19, KAPPA0 0.1, NU0 19 (TCTXVECT.cpp:36-54) */
static void cvs_consts(orc_tables* t) {
  const int DIM = 19;
  const float KAPPA0 = 0.1f, NU0 = (float)DIM;
  t->cvkappa0inv = 1.f / KAPPA0;
  t->cvpowernu0 = -0.5f * (NU0 + 2.f);
  float cterm = (0.5f * DIM + t->cvpowernu0) * (logf(KAPPA0) - logf(KAPPA0 + 2.f));
  float sum = 0.0f;
  for (int ii = 1; ii <= DIM; ii++) sum += logf(NU0 + 1.0f - ii) - 0.69314718055994530942f;
  t->cvcterm = cterm + sum;
}

/* pair scores, DMS_AS_2: r = larger, c = smaller distance (saturated at 20);
 * intercept = 0.35 - c*(-0.4)*0.05, slope = (-0.4-0.35)*0.05 (constant-folded
 * in single precision), score = r*slope + intercept; 255 in either -> -1 */
static void dms_table(orc_tables* t) {
  const float slope = (-0.4f - 0.35f) * 0.05f;
  for (int a = 0; a < 22; a++)
    for (int b = 0; b < 22; b++) {
      float v;
      if (a == 21 || b == 21)
        v = -1.0f;
      else {
        int r = a, c = b;
        if (r < c) { int x = r; r = c; c = x; }
        const float intpt = fmaf(-((float)c * -0.4f), 0.05f, 0.35f);
        v = fmaf((float)r, slope, intpt);
      }
      t->dms[a * 22 + b] = v;
    }
}

int orc_tables_load(const char* envdir, orc_tables* t) {
  char path[4096];
  memset(t, 0, sizeof *t);
  snprintf(path, sizeof path, "%s/ssss", envdir);
  int rc = load_ssss(path, t);
  if (rc) return rc;
  snprintf(path, sizeof path, "%s/cvs2s", envdir);
  rc = load_cvs2s(path, t);
  if (rc) return 100 + rc;
  cvs_consts(t);
  dms_table(t);
  return 0;
}
