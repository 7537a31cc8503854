// dstio.cpp: distance files into per-position distance lists (adddist).
//
// adddist (src/adddist/adddist.cpp) reads a file of residue-pair distances
// and stores, for each position of a profile, the list the DDMS term reads.
// The file (ReadDistances, adddist.cpp:267-367; GetPositionPairHelper,
// :385-436; GetPairDistanceHelper, :438-545): one pair per line, "i j" (1-based,
// i < j, lines ordered by i) followed by columns; lines whose first
// non-blank character is '#' or that are empty are skipped.  The distance of
// a line is the mean of the selected columns (1-based numbers >= 3); a '-'
// in a selected column is a missing value; with a probability threshold
// each selected distance is followed by its probability, and a line counts
// only when every selected distance meets the threshold.  Distances must lie
// in [0, 64], probabilities in [0, 1].  The list of a position is compiled
// when the first line of the next position arrives (and after the last
// line), from its own lines and those of earlier positions naming it
// (cth_adddist.h); positions that never start a line keep an empty list, as
// in adddist.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <string>
#include <vector>

#include "cth_adddist.h"
#include "cth_pset.h"

void cth_set_error(const char* fmt, ...);

namespace {

// the next line that is not a comment or blank (skip_comments,
// liblib/myfiler.cpp:131-181); false at the end of the file
bool next_line(FILE* fp, std::string& line) {
  line.clear();
  char buf[1024];
  bool full = true;
  while (fgets(buf, sizeof buf, fp)) {
    const size_t len = strlen(buf);
    if (full) {
      const char* p = buf;
      while (*p == ' ' || *p == '\t') p++;
      if (*p != '#' && *p != '\n' && *p != '\r' && len) {
        line = buf;
        if (buf[len - 1] != '\n' && buf[len - 1] != '\r') {
          while (fgets(buf, sizeof buf, fp)) {
            const size_t l2 = strlen(buf);
            line += buf;
            if (l2 && (buf[l2 - 1] == '\n' || buf[l2 - 1] == '\r')) break;
          }
        }
        return true;
      }
    }
    // as in the reference, once a line longer than the buffer was skipped no
    // later line of this call is taken
    if (len && buf[len - 1] != '\n' && buf[len - 1] != '\r') full = false;
  }
  return false;
}

// one numeric token at p (leading blanks skipped; the token ends at a blank,
// ';', '(', ')', end of line and, for floats, ','); the reference's
// read_integer / read_float (liblib/myfiler.cpp:324-425)
bool token(const std::string& s, size_t& pos, std::string& tok, bool is_float) {
  size_t p = pos;
  while (p < s.size() && (s[p] == ' ' || s[p] == '\t')) p++;
  const size_t b = p;
  while (p < s.size() && s[p] && s[p] != ' ' && s[p] != '\t' && s[p] != ';' && s[p] != '(' && s[p] != ')' &&
         s[p] != '\n' && s[p] != '\r' && !(is_float && s[p] == ','))
    p++;
  if (p == b) return false;
  tok.assign(s, b, p - b);
  pos = p;
  return true;
}

bool read_int(const std::string& s, size_t& pos, int* v) {
  std::string t;
  if (!token(s, pos, t, false)) return false;
  char* e;
  errno = 0;
  const long x = strtol(t.c_str(), &e, 10);
  if (errno || *e) return false;
  *v = (int)x;
  return true;
}

bool read_flt(const std::string& s, size_t& pos, float* v) {
  std::string t;
  if (!token(s, pos, t, true)) return false;
  char* e;
  errno = 0;
  const float x = strtof(t.c_str(), &e);
  if (errno || *e) return false;
  *v = x;
  return true;
}

bool fail(const char* path, int line, const char* what) {
  cth_set_error("distance file %s, line %d (comments excluded): %s", path, line, what);
  return false;
}

bool read_distances(const char* path, int prolen, const std::vector<int>& cols, float prbthrhld,
                    std::vector<std::vector<float>>& dgram) {
  dgram.assign(prolen, std::vector<float>());
  if (cols.empty()) return true;
  FILE* fp = fopen(path, "r");
  if (!fp) {
    cth_set_error("cannot open distance file %s", path);
    return false;
  }
  std::vector<std::vector<std::pair<float, int>>> tri(prolen);
  std::vector<float> vdsts;
  std::vector<int> vposs;
  int prevpos = 1;
  std::string s;
  bool ok = true;
  const int coldstlast = cols.back(), colast = prbthrhld > 0.0f ? coldstlast + 1 : coldstlast;
  for (int ll = 0; ok && next_line(fp, s); ll++) {
    size_t pos = 0;
    int rp[2];
    for (int c = 0; c < 2 && ok; c++) {
      if (s.size() <= pos) ok = fail(path, ll, "missing a pair of residue indices");
      else if (!read_int(s, pos, &rp[c])) ok = fail(path, ll, "invalid residue index");
      else if (rp[c] < 1 || 10000 < rp[c]) ok = fail(path, ll, "residue index out of bounds");
    }
    if (!ok) break;
    // the line's distance: mean of the selected columns
    float dst = -1.f, dstsum = 0.f;
    int ndx = 0, ndsts = 0;
    for (int c = 2; c <= colast && ndx < (int)cols.size() && ok; c++) {
      if (s.size() <= pos) { ok = fail(path, ll, "missing a distance value"); break; }
      while (pos < s.size() && (s[pos] == ' ' || s[pos] == '\t')) pos++;
      size_t end = pos;
      while (end < s.size() && s[end] && s[end] != ' ' && s[end] != '\t' && s[end] != '\n' && s[end] != '\r') end++;
      if (c < cols[ndx]) { pos = end; continue; }
      ndx++;
      if (pos < s.size() && s[pos] == '-') { pos = end; continue; }
      float v;
      if (!read_flt(s, pos, &v)) { ok = fail(path, ll, "invalid distance value"); break; }
      if (v < 0.f || 64.f < v) { ok = fail(path, ll, "distance outside [0,64]"); break; }
      const float d = v;
      if (prbthrhld > 0.0f) {
        if (++c > colast) { ok = fail(path, ll, "missing a probability value"); break; }
        if (!read_flt(s, pos, &v)) { ok = fail(path, ll, "invalid probability value"); break; }
        if (v < 0.f || 1.f < v) { ok = fail(path, ll, "probability outside [0,1]"); break; }
        if (v < prbthrhld) continue;
      }
      dstsum += d;
      ndsts++;
    }
    if (!ok) break;
    if (ndx < (int)cols.size()) { ok = fail(path, ll, "missing distance values"); break; }
    if (ndsts && (!prbthrhld || ndsts == (int)cols.size())) dst = dstsum / ndsts;
    if (rp[0] < prevpos || rp[1] <= rp[0]) { ok = fail(path, ll, "positions not increasing"); break; }
    // adddist checks the profile length only when a new position begins; a
    // pair beyond the profile is an error here on every line
    if (prolen < rp[0] || prolen < rp[1]) { ok = fail(path, ll, "position beyond the profile length"); break; }
    if (prevpos < rp[0]) {
      cth_adddist_position(prevpos - 1, prolen, vdsts, vposs, tri, dgram);
      vdsts.clear();
      vposs.clear();
      prevpos = rp[0];
    }
    if (dst >= 0.f) {
      vdsts.push_back(dst);
      vposs.push_back(rp[1] - 1);
      tri[rp[1] - 1].push_back({dst, rp[0] - 1});
    }
  }
  fclose(fp);
  if (!ok) return false;
  if (prolen > 0) cth_adddist_position(prevpos - 1, prolen, vdsts, vposs, tri, dgram);
  return true;
}

bool columns(const int* cols, int ncols, std::vector<int>& out) {
  out.clear();
  if (!cols || ncols < 1) {
    out.push_back(2);  // the default column 3
    return true;
  }
  for (int k = 0; k < ncols; k++) {
    if (cols[k] < 3 || 100 < cols[k]) {
      cth_set_error("distance column %d outside [3,100]", cols[k]);
      return false;
    }
    out.push_back(cols[k] - 1);
  }
  std::sort(out.begin(), out.end());
  out.erase(std::unique(out.begin(), out.end()), out.end());
  return true;
}

}  // namespace

// lists: [prolen][CTH_MAXDST + 1] floats, entry 0 the count, then the values
extern "C" int cth_read_distances(const char* path, int prolen, const int* cols, int ncols, float prbthrhld,
                                  float* lists) {
  std::vector<int> cv;
  if (prolen < 0 || !lists || !columns(cols, ncols, cv)) {
    if (prolen < 0 || !lists) cth_set_error("read_distances: invalid arguments");
    return -1;
  }
  if (prbthrhld < 0.0f || 1.0f <= prbthrhld) {
    cth_set_error("read_distances: probability threshold outside [0,1)");
    return -1;
  }
  std::vector<std::vector<float>> dg;
  if (!read_distances(path, prolen, cv, prbthrhld, dg)) return -1;
  for (int m = 0; m < prolen; m++) {
    float* o = lists + (size_t)m * (CTH_MAXDST + 1);
    o[0] = (float)dg[m].size();
    for (size_t k = 0; k < dg[m].size(); k++) o[1 + k] = dg[m][k];
  }
  return 0;
}

// Replace the distance lists of profile k of ps by those of a distance file,
// as adddist writes them to a profile: count, then each value rounded to an
// integer (TextWriteProfileCondensed, PMProfileModel.cpp:2063-2067).
extern "C" int cth_add_distances(cth_pset* ps, int k, const char* path, const int* cols, int ncols,
                                 float prbthrhld) {
  if (!ps || k < 0 || k >= ps->nprof) {
    cth_set_error("add_distances: profile %d out of range", k);
    return -1;
  }
  const int L = ps->len[k];
  std::vector<float> lists((size_t)L * (CTH_MAXDST + 1));
  if (cth_read_distances(path, L, cols, ncols, prbthrhld, lists.data())) return -1;
  for (int m = 0; m < L; m++) {
    const float* l = &lists[(size_t)m * (CTH_MAXDST + 1)];
    uint8_t* rec = ps->dd + (size_t)(ps->dist[k] + m) * CTH_DDSTRIDE;
    memset(rec, 0xff, CTH_DDSTRIDE);
    const int n = (int)l[0];
    rec[0] = (uint8_t)n;
    for (int t = 0; t < n; t++) {
      const int v = (int)rintf(l[1 + t]);
      rec[1 + t] = (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v);
    }
  }
  return 0;
}
