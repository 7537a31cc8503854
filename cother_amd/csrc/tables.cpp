// tables.cpp: loading of the score tables the score-matrix stage consumes.
//
//  - ssss: the 30x30 secondary-structure score table (file format read by
//    SSSScores::ReadLevels/ReadScoreTables, libpro/srcpro/SSSScores.cpp:66-222,
//    and VirtScores::ReadScoresHelper, libpro/srcpro/VirtScores.cpp:44-140),
//    kept in the serialized order SerializedScoresCtor produces: two header
//    cells followed by the lower triangle row by row
//    (cupro/SerializedScoresSM.cuh:103-120 indexes it that way).
//  - cvs2s: the map from context-vector log-odds to scores, read and
//    smoothed as CVS2Scores::ReadScoresHelper/SmoothScores do
//    (libpro/srcpro/CVS2Scores.cpp:207-466): a 5-point running mean replaces
//    every score and "+" entries take the running mean.
//  - CVS constants of _TCTXVECT (libpro/srcpro/TCTXVECT.cpp:26-60).
//  - the 22x22 distance-match pair-score table of DMS_AS_2
//    (cupro/SerializedDstMatchScores.cuh:350-378); index 21 is the missing
//    distance 255, which scores -1.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "cth_tables.h"

namespace {

bool read_content_line(FILE* fp, std::string& out) {
  char buf[16384];
  while (fgets(buf, sizeof buf, fp)) {
    const char* p = buf;
    while (*p == ' ' || *p == '\t') p++;
    if (*p == '#' || *p == '\n' || *p == '\r' || *p == 0) continue;  // skip_comments
    out = buf;
    return true;
  }
  return false;
}

int parse_ssss(const std::string& path, cth_tables* t) {
  FILE* fp = fopen(path.c_str(), "r");
  if (!fp) return -1;
  std::string ln;
  int rc = 0;
  if (!read_content_line(fp, ln) || ln.find("ssss=") == std::string::npos) rc = -2;
  else if (!read_content_line(fp, ln) || ln.compare(0, 7, "0000 11")) rc = -3;
  else if (!read_content_line(fp, ln) || ln.find("Cardinality =") == std::string::npos) rc = -4;
  else if (atoi(ln.c_str() + ln.find('=') + 1) != 30) rc = -5;
  if (rc) { fclose(fp); return rc; }
  t->ss[0] = 0.0f;  // probability level (0/100)
  t->ss[1] = 1.0f;  // ENO level
  int k = 2;
  for (int r = 0; r < 30 && !rc; r++) {
    if (!read_content_line(fp, ln)) { rc = -6; break; }
    const char* p = ln.c_str();
    for (int c = 0; c <= r; c++) {
      char* e;
      float v = strtof(p, &e);
      if (e == p) { rc = -7; break; }
      t->ss[k++] = v;
      p = e;
    }
  }
  fclose(fp);
  return rc;
}

// 5-point running-mean smoothing of the map's scores, step 2
void smooth_map(std::vector<float>& keys, std::vector<float>& scos, int* shift) {
  const float kPlus = 9999.0f;
  const int win = 5, winh = win >> 1;
  const size_t sz = scos.size();
  std::vector<float> smth(sz + 1, 0.0f);
  float ring[win] = {0, 0, 0, 0, 0};
  float rsum = 0.0f, avg = 0.0f;
  int nv = 0, rn = 0;
  size_t nn = 0;
  scos.push_back(0.0f);  // sentinel read by the '+'-run scan below
  for (size_t n = 0; n < sz; n++) {
    const float sc = scos[n];
    if (nv < winh) nn++;
    if (sc != kPlus) {
      if (win <= rn) rn = 0;
      rsum -= ring[rn];
      ring[rn++] = sc;
      rsum += sc;
      avg = rsum / (float)win;
      nv++;
    }
    if (win <= nv) {
      if (sc != kPlus) {
        if (nv == win)
          for (size_t p = 0; p < nn; p++) smth[p] = avg;
        smth[nn] = avg;
        for (++nn; scos[nn] == kPlus && nn < n; nn++) smth[nn] = avg;
      }
      if (sz == n + 1)
        for (size_t p = nn; p < sz; p++) smth[p] = avg;
    }
  }
  scos.pop_back();
  for (size_t n = 0; n < sz; n++) scos[n] = smth[n];  // every entry is replaced by its smoothed value
  *shift = 0;
  for (size_t n = 0; n < keys.size(); n++) {
    if (0.0f <= keys[n]) break;
    (*shift)++;
  }
}

int parse_cvs2s(const std::string& path, cth_tables* t) {
  FILE* fp = fopen(path.c_str(), "r");
  if (!fp) return -1;
  std::string ln;
  int rc = 0, card = 0;
  if (!read_content_line(fp, ln) || ln.find("cvs2s=") == std::string::npos) rc = -2;
  else if (!read_content_line(fp, ln) || ln.compare(0, 2, "11")) rc = -3;
  else if (!read_content_line(fp, ln)) rc = -4;
  if (!rc && ln.find("Scale =") != std::string::npos && !read_content_line(fp, ln)) rc = -4;
  if (!rc && ln.find("Dimensions =") == std::string::npos) rc = -5;
  if (!rc && (!read_content_line(fp, ln) || ln.find("Cardinality =") == std::string::npos)) rc = -6;
  if (!rc) card = atoi(ln.c_str() + ln.find('=') + 1);
  if (!rc && (card <= 0 || card > CTH_CVMAXCARD)) rc = -7;
  std::vector<float> keys, scos;
  if (!rc && read_content_line(fp, ln)) {
    const char* p = ln.c_str();
    for (int c = 0; c < card; c++) {
      char* e;
      float v = strtof(p, &e);
      if (e == p) { rc = -9; break; }
      keys.push_back(v);
      p = e;
    }
  } else if (!rc) rc = -8;
  if (!rc && read_content_line(fp, ln)) {
    const char* p = ln.c_str();
    for (int c = 0; c < card; c++) {
      while (*p == ' ' || *p == '\t') p++;
      if (*p == '+') { scos.push_back(9999.0f); p++; continue; }
      char* e;
      float v = strtof(p, &e);
      if (e == p) { rc = -11; break; }
      scos.push_back(v);
      p = e;
    }
  } else if (!rc) rc = -10;
  fclose(fp);
  if (rc) return rc;
  // keys are on a gapless 0.5 grid in the shipped table; a gap would need the
  // key insertion of SmoothScores, which this loader rejects instead
  for (size_t i = 1; i < keys.size(); i++)
    if (keys[i] - keys[i - 1] > 0.5f) return -12;
  int shift = 0;
  smooth_map(keys, scos, &shift);
  t->cvcard = card;
  t->cvshift = shift;
  for (int i = 0; i < card; i++) { t->cvkeys[i] = keys[i]; t->cvvals[i] = scos[i]; }
  return 0;
}

}  // namespace

extern "C" int cth_tables_load(const char* envdir, cth_tables* t) {
  memset(t, 0, sizeof *t);
  std::string d(envdir);
  int rc = parse_ssss(d + "/ssss", t);
  if (rc) return rc;
  rc = parse_cvs2s(d + "/cvs2s", t);
  if (rc) return 100 + rc;
  // _TCTXVECT: DIM=19, KAPPA0=0.1, NU0=DIM
  const int dim = 19;
  const float kappa0 = 0.1f, nu0 = (float)dim;
  t->cvkappa0inv = 1.f / kappa0;
  t->cvpowernu0 = -0.5f * (nu0 + 2.f);
  float cterm = (0.5f * dim + t->cvpowernu0) * (logf(kappa0) - logf(kappa0 + 2.f));
  float sum = 0.0f;
  for (int ii = 1; ii <= dim; ii++) sum += logf(nu0 + 1.0f - ii) - 0.69314718055994530942f;
  t->cvcterm = cterm + sum;
  // pair scores: larger distance r, smaller c; slope constant-folded in float
  const float slope = (CTH_DMS_S200 - CTH_DMS_S00) * CTH_DMS_MAXDSTINV;
  for (int a = 0; a < 22; a++)
    for (int b = 0; b < 22; b++) {
      float v = -1.0f;
      if (a < 21 && b < 21) {
        const int r = a > b ? a : b, c = a > b ? b : a;
        const float intpt = fmaf(-((float)c * CTH_DMS_S200), CTH_DMS_MAXDSTINV, CTH_DMS_S00);
        v = fmaf((float)r, slope, intpt);
      }
      t->dms[a * 22 + b] = v;
    }
  // the score kernel's DDMS scan tells a positive running sum from zero, and
  // two sums apart, by scaling with 2^126 and 2^100 and clamping to [0, 1]:
  // exact when every nonzero pair score is at least 2^-75 in magnitude (all
  // sums of such scores and of their halves are then multiples of 2^-99);
  // these entries are multiples of 2^-48, the smallest nonzero near 2^-25
  for (float v : t->dms)
    if (v != 0.0f && fabsf(v) < 0x1p-75f) return 200;
  return 0;
}
