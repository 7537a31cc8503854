// proio.cpp: the reference's profile and database file formats.
//
//  - text profiles, "COTHER profile v2.5" (written by TextWriteProfileCondensed,
//    libpro/srcpro/PMProfileModel.cpp:1856-2080; read by TextReadProfile,
//    :3527-4620): one or more profiles per file, as single .pro files or the
//    makedb text database (.tprd, libpro/srcpro/Db.cpp:590-670);
//  - the binary profile database, "COMER bin profile v2.5" layout v2.2
//    (CalculateAndWriteAddressTable_v2_2 and BinaryWriteProfile_v2_2,
//    PMProfileModel.cpp:1235-1817), which the reference's search reads
//    section by section into its packed device fields
//    (libmycu/cupro/TdDataReader.cpp:385-800).
//
// Both readers produce a cth_pset with the values the reference's device
// code sees: transitions as logarithms (x0.4 for the states after MD, -32768
// for a zero probability), target/background/SS probabilities and context
// vectors as read, distances as the count followed by up to 99 values (255
// for missing), the uninformative context vector 0.17302947 where a profile
// has none.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <cmath>
#include <fstream>
#include <string>
#include <vector>

#include "cth_consts.h"
#include "cth_pset.h"

extern "C" cth_pset* cth_pset_alloc(int nprof, const int32_t* lens);
extern "C" void cth_pset_free(cth_pset* ps);
extern "C" int cth_pset_set_text(cth_pset* ps, int k, const char* name, const char* desc);
extern "C" const char* cth_pset_name(const cth_pset* ps, int k);
extern "C" const char* cth_pset_desc(const cth_pset* ps, int k);
void cth_set_error(const char* fmt, ...);

namespace {

const char kAlphabet[] = "ARNDCQEGHILKMFPSTWYVBZJX";  // liblib/alpha.cpp:11
const char kSSAlphabet[] = "CEH";                     // liblib/alpha.cpp:13 (text files)
const char kSSLower[] = " eh";                        // liblib/alpha.cpp:14 (the device's SS state letters)
constexpr int kNStates = 9;                           // P_MM..P_DD (libpro/srcpro/TRANSPROBS.h:14-19)
constexpr int kPID = 5, kPDI = 7, kPMD = 2;
constexpr float kLogFct = 0.4f;                       // MAIN_TRNPRB_LOGFCT (TRANSPROBS.h:10)
constexpr int kIntScale = 1000;                       // INTSCALE (PMProfileModel.cpp:1226)
constexpr int kMaxDst = 99;                           // MAX_N_IR_DISTANCE_VALUES (MOptions.h:15)
constexpr float kUninfCV = 0.17302947f;
constexpr int kV22Sections = 179;  // PMProfileModel::v2_2_NSECTIONS (libpro/srcpro/PMProfileModel.h:42-48)

// the stored (logarithmic) form of a transition probability
// (BinaryWriteProfile_v2_2, PMProfileModel.cpp:1450-1463)
inline float trn_log(float p, int t9) {
  if (!p) return -32768.0f;
  return kPMD < t9 ? logf(p) * kLogFct : logf(p);
}

// inverse for writing: the scaled integer whose quotient gives log value v
inline int trn_int(float v, int t9, int scale) {
  if (v <= -32768.0f) return 0;
  const double p = kPMD < t9 ? exp((double)v / kLogFct) : exp((double)v);
  return (int)lrint(scale * p);
}

int aa_code(char c) {
  const char* p = strchr(kAlphabet, c);
  return p && c ? (int)(p - kAlphabet) : -1;
}

struct Lines {
  std::vector<std::string> v;
  size_t i = 0;
  bool next(std::string& s) {
    while (i < v.size()) {
      s = v[i++];
      // skip comments and blank lines (skip_comments, liblib/myfiler.cpp)
      size_t k = s.find_first_not_of(" \t\r");
      if (k == std::string::npos || s[k] == '#') continue;
      return true;
    }
    return false;
  }
  bool peek(std::string& s) {
    size_t save = i;
    bool ok = next(s);
    i = save;
    return ok;
  }
};

// integers after `tag` (or from the line start when tag is NULL)
bool ints_after(const std::string& s, const char* tag, std::vector<long>& out, size_t want = 0) {
  out.clear();
  size_t pos = 0;
  if (tag) {
    pos = s.find(tag);
    if (pos == std::string::npos) return false;
    pos += strlen(tag);
  }
  const char* p = s.c_str() + pos;
  char* e;
  for (;;) {
    while (*p == ' ' || *p == '\t') p++;
    if (!*p || *p == '\r' || *p == '\n') break;
    long v = strtol(p, &e, 10);
    if (e == p) break;
    out.push_back(v);
    p = e;
  }
  return want == 0 || out.size() >= want;
}

struct TextProfile {
  std::string desc, file;
  int len = 0, nos = 1, scale = 0;
  float effnos = 0.f;
  float bkg[CTH_NAA];
  float trnbeg[kNStates];
  std::string aa, sss;
  std::vector<float> trg, trn, ss, cv, cvprior, cvnorm2;  // trg [len][20], trn [len][9], ss [len][3], cv [len][19]
  std::vector<std::vector<int>> dst;
  bool has_ss = false, has_cv = false;
};

bool fail(const char* what, int pos = -1) {
  if (pos >= 0) cth_set_error("text profile: %s at position %d", what, pos + 1);
  else cth_set_error("text profile: %s", what);
  return false;
}

// one profile (TextReadProfile, PMProfileModel.cpp:3527-4620); false at the end of input
bool read_one(Lines& L, TextProfile& P, bool* err) {
  *err = false;
  std::string s;
  std::vector<long> v;
  if (!L.next(s)) return false;
  *err = true;
  if (s.find("COTHER profile v") == std::string::npos) return fail("no version line");
  if (!L.next(s) || s.find("DESC:") == std::string::npos) return fail("no DESC");
  P.desc = s.substr(s.find("DESC:") + 5);
  if (!L.next(s) || s.find("FILE:") == std::string::npos) return fail("no FILE");
  P.file = s.substr(s.find("FILE:") + 5);
  for (std::string* t : {&P.desc, &P.file}) {
    size_t a = t->find_first_not_of(" \t");
    *t = a == std::string::npos ? "" : t->substr(a);
    while (!t->empty() && (t->back() == '\r' || t->back() == '\n')) t->pop_back();
  }
  if (!L.next(s) || s.find("CMD:") == std::string::npos) return fail("no CMD");
  if (!L.next(s) || !ints_after(s, "LEN:", v, 1)) return fail("no LEN");
  P.len = (int)v[0];
  if (P.len < 1) return fail("invalid LEN");
  if (!L.next(s) || !ints_after(s, "NOS:", v, 1)) return fail("no NOS");
  P.nos = (int)v[0];
  if (!L.next(s) || s.find("EFFNOS:") == std::string::npos) return fail("no EFFNOS");
  P.effnos = strtof(s.c_str() + s.find("EFFNOS:") + 7, nullptr);
  if (!L.next(s) || !ints_after(s, "SCALE:", v, 1)) return fail("no SCALE");
  P.scale = (int)v[0];
  if (P.scale < 1) return fail("invalid SCALE");
  const float fs = (float)P.scale;
  if (!L.next(s)) return fail("no alphabet line");  // residue letters
  if (!L.next(s) || !ints_after(s, "NULL", v, CTH_NAA)) return fail("no background probabilities");
  for (int r = 0; r < CTH_NAA; r++) P.bkg[r] = (float)v[r] / fs;
  if (!L.next(s) || !ints_after(s, "POST", v, CTH_NAA)) return fail("no posterior probabilities");
  if (!L.next(s) || !ints_after(s, nullptr, v, kNStates)) return fail("no beginning transitions");
  for (int t = 0; t < kNStates; t++) P.trnbeg[t] = (float)v[t] / fs;
  if (!L.next(s) || !ints_after(s, nullptr, v, 3)) return fail("no beginning MID observations");
  const int n = P.len;
  P.aa.assign(n, 'X');
  P.sss.assign(n, 0);
  P.trg.assign((size_t)n * CTH_NAA, 0.f);
  P.trn.assign((size_t)n * kNStates, 0.f);
  P.ss.assign((size_t)n * 3, 0.f);
  P.cv.assign((size_t)n * CTH_NCV, kUninfCV);
  P.cvprior.assign(n, 0.f);
  P.cvnorm2.assign(n, 0.f);
  P.dst.assign(n, {});
  P.has_ss = P.has_cv = false;
  for (int m = 0; m < n; m++) {
    // "<n> <residue> 20 target probabilities"
    if (!L.next(s)) return fail("missing position", m);
    {
      const char* p = s.c_str();
      char* e;
      long num = strtol(p, &e, 10);
      if (e == p || num != m + 1) return fail("wrong numbering", m);
      p = e;
      while (*p == ' ' || *p == '\t') p++;
      if (aa_code(*p) < 0) return fail("unknown residue", m);
      P.aa[m] = *p;
      if (!ints_after(std::string(p + 1), nullptr, v, CTH_NAA)) return fail("incomplete target probabilities", m);
      for (int r = 0; r < CTH_NAA; r++) P.trg[(size_t)m * CTH_NAA + r] = (float)v[r] / fs;
    }
    if (!L.next(s) || !ints_after(s, nullptr, v, kNStates)) return fail("incomplete transitions", m);
    for (int t = 0; t < kNStates; t++) P.trn[(size_t)m * kNStates + t] = (float)v[t] / fs;
    if (!L.next(s) || !ints_after(s, nullptr, v, 3)) return fail("incomplete MID observations", m);
    // HDP1: background posterior, count, then (count > 0) probabilities and indices
    if (!L.next(s) || !ints_after(s, nullptr, v, 2)) return fail("no HDP1 posteriors", m);
    if (v[1] > 0) {
      if (!L.next(s) || !L.next(s)) return fail("incomplete HDP1 posteriors", m);
    }
    // optional lines, in the writer's order
    std::string t;
    if (L.peek(t) && t.find("CT:") != std::string::npos) L.next(s);
    if (L.peek(t) && t.find("CV:") != std::string::npos) {
      L.next(s);
      if (!ints_after(s, "CV:", v, 3)) return fail("incomplete context vector", m);
      const int vs = (int)v[2];
      if (vs != CTH_NCV || (int)v.size() < 3 + vs) return fail("context vector size", m);
      P.has_cv = true;
      P.cvprior[m] = (float)v[0] / fs;
      P.cvnorm2[m] = (float)v[1] / fs;
      for (int k = 0; k < vs; k++) P.cv[(size_t)m * CTH_NCV + k] = (float)v[3 + k] / fs;
    }
    if (L.peek(t) && t.find("SS:") != std::string::npos) {
      L.next(s);
      const size_t k = s.find("SS:") + 3;
      const char* ss = strchr(kSSAlphabet, s[k]);
      if (!ss || !s[k]) return fail("unknown SS state", m);
      P.sss[m] = kSSLower[ss - kSSAlphabet];
      if (!ints_after(s.substr(k + 1), nullptr, v, 3)) return fail("incomplete SS probabilities", m);
      P.has_ss = true;
      for (int q = 0; q < 3; q++) P.ss[(size_t)m * 3 + q] = (float)v[q] / fs;
    }
    if (L.peek(t) && t.find("D:") != std::string::npos && t.find("SS:") == std::string::npos) {
      L.next(s);
      if (!ints_after(s, "D:", v, 1)) return fail("no distance count", m);
      const int nd = (int)v[0];
      if (nd < 0 || (int)v.size() < 1 + nd) return fail("incomplete distances", m);
      P.dst[m].assign(v.begin() + 1, v.begin() + 1 + nd);
    }
  }
  // statistical parameters and the end mark
  while (L.next(s)) {
    size_t k = s.find_first_not_of(" \t");
    if (k != std::string::npos && s[k] == '*') return true;
  }
  return fail("no end mark");
}

cth_pset* to_pset(const std::vector<TextProfile>& ps) {
  std::vector<int32_t> lens;
  for (const auto& p : ps) lens.push_back(p.len);
  cth_pset* o = cth_pset_alloc((int)ps.size(), lens.data());
  if (!o) return nullptr;
  const int64_t n = o->npos, ntr = o->npos + o->nprof;
  for (int p = 0; p < o->nprof; p++) {
    const TextProfile& P = ps[p];
    const int64_t d0 = o->dist[p];
    if (cth_pset_set_text(o, p, P.file.c_str(), P.desc.c_str())) { cth_pset_free(o); return nullptr; }
    o->eno[p] = P.effnos;
    for (int r = 0; r < CTH_NAA; r++) o->bkg[p * CTH_NAA + r] = P.bkg[r];
    for (int t9 = 0, t = 0; t9 < kNStates; t9++) {
      if (t9 == kPID || t9 == kPDI) continue;
      o->trn[t * ntr + d0 + p] = trn_log(P.trnbeg[t9], t9);
      for (int m = 0; m < P.len; m++) o->trn[t * ntr + d0 + p + m + 1] = trn_log(P.trn[(size_t)m * kNStates + t9], t9);
      t++;
    }
    for (int m = 0; m < P.len; m++) {
      const int64_t i = d0 + m;
      o->aa[i] = P.aa[m];
      o->sss[i] = P.has_ss ? P.sss[m] : 0;
      for (int r = 0; r < CTH_NAA; r++) o->trg[r * n + i] = P.trg[(size_t)m * CTH_NAA + r];
      for (int q = 0; q < 3; q++) o->ss[q * n + i] = P.has_ss ? P.ss[(size_t)m * 3 + q] : 0.f;
      for (int k = 0; k < CTH_NCV; k++) o->cv[k * n + i] = P.has_cv ? P.cv[(size_t)m * CTH_NCV + k] : kUninfCV;
      o->cvprior[i] = P.has_cv ? P.cvprior[m] : 0.f;
      o->cvnorm2[i] = P.has_cv ? P.cvnorm2[m] : 0.f;
      uint8_t* rec = o->dd + (size_t)i * CTH_DDSTRIDE;
      memset(rec, 255, CTH_DDSTRIDE);
      const int nd = (int)P.dst[m].size() < kMaxDst ? (int)P.dst[m].size() : kMaxDst;
      rec[0] = (uint8_t)nd;
      for (int k = 0; k < nd; k++) {
        const int d = P.dst[m][k];
        rec[1 + k] = (uint8_t)(d < 0 ? 0 : (d > 255 ? 255 : d));
      }
    }
  }
  return o;
}

}  // namespace

// Read every profile of a text profile file or text database (.pro, .tprd).
extern "C" cth_pset* cth_read_text_profiles(const char* path) {
  std::ifstream f(path);
  if (!f) { cth_set_error("cannot open %s", path); return nullptr; }
  Lines L;
  std::string s;
  while (std::getline(f, s)) L.v.push_back(s);
  // a text database starts with its own header (Db::WriteTextHeader,
  // libpro/srcpro/Db.cpp): skip to the first profile
  while (L.i < L.v.size() && L.v[L.i].find("COTHER profile v") == std::string::npos) L.i++;
  std::vector<TextProfile> ps;
  for (;;) {
    TextProfile P;
    bool err = false;
    if (!read_one(L, P, &err)) {
      if (err) return nullptr;
      break;
    }
    ps.push_back(std::move(P));
  }
  if (ps.empty()) { cth_set_error("%s: no profiles", path); return nullptr; }
  return to_pset(ps);
}

// Write profiles in the text format of TextWriteProfileCondensed
// (PMProfileModel.cpp:1856-2080); a NULL name or description is written empty.
extern "C" int cth_write_text_profiles(const cth_pset* ps, const char* path, int scale) {
  if (scale < 1) { cth_set_error("write_text: invalid scale"); return -1; }
  FILE* fp = fopen(path, "w");
  if (!fp) { cth_set_error("cannot write %s", path); return -1; }
  const int64_t n = ps->npos, ntr = ps->npos + ps->nprof;
  for (int p = 0; p < ps->nprof; p++) {
    const int L = ps->len[p];
    const int64_t d0 = ps->dist[p];
    fprintf(fp, "COTHER profile v2.5\n");
    const char* desc = cth_pset_desc(ps, p);
    const char* name = cth_pset_name(ps, p);
    fprintf(fp, "DESC: %s\n", desc ? desc : "");
    fprintf(fp, "FILE: %s\n", name ? name : "");
    fprintf(fp, "CMD: \n");
    fprintf(fp, "LEN: %d\n", L);
    fprintf(fp, "NOS: %d\n", 1);
    fprintf(fp, "EFFNOS: %.1f\n", ps->eno[p]);
    fprintf(fp, "SCALE: %d\n", scale);
    for (int r = 0; r < CTH_NAA; r++) fprintf(fp, " %c", kAlphabet[r]);
    fprintf(fp, "\nNULL");
    for (int r = 0; r < CTH_NAA; r++) fprintf(fp, " %d", (int)rintf(scale * ps->bkg[p * CTH_NAA + r]));
    fprintf(fp, "\nPOST");
    for (int r = 0; r < CTH_NAA; r++) fprintf(fp, " %d", (int)rintf(scale * ps->bkg[p * CTH_NAA + r]));
    auto trn9 = [&](int64_t vec) {
      for (int t9 = 0, t = 0; t9 < kNStates; t9++) {
        if (t9 == kPID || t9 == kPDI) { fprintf(fp, " %d", 0); continue; }
        fprintf(fp, " %d", trn_int(ps->trn[t * ntr + vec], t9, scale));
        t++;
      }
    };
    fprintf(fp, "\n");
    trn9(d0 + p);
    fprintf(fp, "\n %d %d %d", kIntScale, 0, 0);
    const bool has_ss = L > 0 && ps->sss[d0] != 0;
    for (int m = 0; m < L; m++) {
      const int64_t i = d0 + m;
      fprintf(fp, "\n%d %c", m + 1, ps->aa[i]);
      for (int r = 0; r < CTH_NAA; r++) fprintf(fp, " %d", (int)rintf(scale * ps->trg[r * n + i]));
      fprintf(fp, "\n");
      trn9(d0 + p + m + 1);
      fprintf(fp, "\n %d %d %d", kIntScale, 0, 0);
      // HDP1: one uninformative cluster (the binary writer requires >= 1)
      fprintf(fp, "\n %d %d\n %d\n %d", 0, 1, 0, -1);
      bool uninf = ps->cvprior[i] == 0.f && ps->cvnorm2[i] == 0.f;
      for (int k = 0; k < CTH_NCV && uninf; k++) uninf = ps->cv[k * n + i] == kUninfCV;
      if (!uninf) {
        fprintf(fp, "\n CV: %d %d %d", (int)rintf(scale * ps->cvprior[i]), (int)rintf(scale * ps->cvnorm2[i]),
                CTH_NCV);
        for (int k = 0; k < CTH_NCV; k++) fprintf(fp, " %d", (int)rintf(scale * ps->cv[k * n + i]));
      }
      if (has_ss) {
        const char* lc = strchr(kSSLower, ps->sss[i]);
        fprintf(fp, "\n SS:%c", lc ? kSSAlphabet[lc - kSSLower] : 'C');
        // floor, as the reference writes SS probabilities
        for (int q = 0; q < 3; q++) fprintf(fp, " %d", (int)(scale * ps->ss[q * n + i]));
      }
      const uint8_t* rec = ps->dd + (size_t)i * CTH_DDSTRIDE;
      if (rec[0]) {
        fprintf(fp, "\n D: %d", rec[0]);
        for (int k = 0; k < rec[0]; k++) fprintf(fp, " %d", rec[1 + k]);
      }
    }
    fprintf(fp, "\n");
    fprintf(fp, "K, Lambda:\n");
    fprintf(fp, "Computed  ungapped, %6.4f %6.4f\n", 0.0, 0.0);
    fprintf(fp, "Reference ungapped, %6.4f %6.4f\n", 0.0, 0.0);
    fprintf(fp, "Entropy, %6.4f; Expected, %6.4f\n", 0.0, 0.0);
    fprintf(fp, "*\n");
  }
  if (fclose(fp)) { cth_set_error("write failed: %s", path); return -1; }
  return 0;
}

namespace {

// Field indices of the reference's packed profile model, TPM2DVectorFields
// (cupro/PM2DVectorFields.h:53-87): 181 fields, one array each, over
// profiles (background probabilities, ENO, length, offset), over the
// npos + nprof transition vectors, or over positions.
enum {
  fBkg = 0, fENO = fBkg + CTH_NAA, fLen, fDist, fTrn, fTrg = fTrn + CTH_NTRN, fCV = fTrg + CTH_NAA,
  fCVprior = fCV + CTH_NCV, fCVnorm2, fDD, fSS = fDD + kMaxDst + 1, fHDPprb = fSS + 3, fHDPind, fAddrPro, fAddrCV,
  fAddrSS, fAA, fSSstate, fTot
};
static_assert(fTot == 181, "pmv2DTotFlds");

// A profile set from per-field arrays in the packed layout (f[k]: field k's
// first element; lengths already checked to be >= 1 and to sum to npos).
cth_pset* decode_fields(const char* const* f, size_t nprof, size_t npos, const int32_t* lens) {
  cth_pset* o = cth_pset_alloc((int)nprof, lens);
  if (!o) return nullptr;
  auto f32 = [&](int k, size_t i) {
    float v;
    memcpy(&v, f[k] + 4 * i, 4);
    return v;
  };
  const size_t ntr = npos + nprof;
  for (size_t p = 0; p < nprof; p++) {
    o->eno[p] = f32(fENO, p);
    for (int r = 0; r < CTH_NAA; r++) o->bkg[p * CTH_NAA + r] = f32(fBkg + r, p);
  }
  for (int t = 0; t < CTH_NTRN; t++)
    for (size_t k = 0; k < ntr; k++) o->trn[t * ntr + k] = f32(fTrn + t, k);
  for (size_t i = 0; i < npos; i++) {
    for (int r = 0; r < CTH_NAA; r++) o->trg[r * npos + i] = f32(fTrg + r, i);
    for (int k = 0; k < CTH_NCV; k++) o->cv[k * npos + i] = f32(fCV + k, i);
    o->cvprior[i] = f32(fCVprior, i);
    o->cvnorm2[i] = f32(fCVnorm2, i);
    for (int q = 0; q < 3; q++) o->ss[q * npos + i] = f32(fSS + q, i);
    o->aa[i] = f[fAA][i];
    o->sss[i] = f[fSSstate][i];
    uint8_t* rec = o->dd + i * CTH_DDSTRIDE;
    memset(rec, 255, CTH_DDSTRIDE);
    const int nd = (int)f32(fDD, i);
    rec[0] = (uint8_t)(nd < 0 ? 0 : (nd > kMaxDst ? kMaxDst : nd));
    for (int k = 0; k < rec[0]; k++) {
      const float d = f32(fDD + 1 + k, i);
      rec[1 + k] = (uint8_t)(d < 0.f ? 0 : (d > 255.f ? 255 : (int)d));
    }
  }
  // a profile without SS predictions has blank SS states and zero probabilities
  for (size_t p = 0; p < nprof; p++) {
    const int64_t d0 = o->dist[p];
    bool any = false;
    for (int m = 0; m < o->len[p] && !any; m++)
      any = o->ss[0 * npos + d0 + m] != 0.f || o->ss[1 * npos + d0 + m] != 0.f || o->ss[2 * npos + d0 + m] != 0.f;
    if (!any)
      for (int m = 0; m < o->len[p]; m++) o->sss[d0 + m] = 0;
  }
  return o;
}

}  // namespace

// Read a binary profile database ("COTHER profile bin database v2.5", layout
// v2.2: Db::WriteBinaryHeader, libpro/srcpro/Db.cpp:708-717, then the
// address table of CalculateAndWriteAddressTable_v2_2 and the sections).
// Its sections hold the packed fields' arrays, which the reference's reader
// copies into the chunk buffer as they are (TdDataReader::ReadDataChunkHelper,
// libmycu/cupro/TdDataReader.cpp:703-925).
extern "C" cth_pset* cth_read_binary_db(const char* path) {
  FILE* fp = fopen(path, "rb");
  if (!fp) { cth_set_error("cannot open %s", path); return nullptr; }
  std::vector<char> buf;
  fseek(fp, 0, SEEK_END);
  const long sz = ftell(fp);
  fseek(fp, 0, SEEK_SET);
  buf.resize(sz > 0 ? (size_t)sz : 0);
  const bool ok = sz > 0 && fread(buf.data(), 1, buf.size(), fp) == buf.size();
  fclose(fp);
  if (!ok) { cth_set_error("cannot read %s", path); return nullptr; }
  const char* sig = "COTHER profile bin database v2.5";
  const size_t ls = strlen(sig);
  if (buf.size() < ls + 24 || memcmp(buf.data(), sig, ls) != 0) {
    cth_set_error("%s: not a COTHER binary database v2.5", path);
    return nullptr;
  }
  size_t nprof;
  uint64_t dbsize, ndst;
  memcpy(&nprof, buf.data() + ls, 8);
  memcpy(&dbsize, buf.data() + ls + 8, 8);
  memcpy(&ndst, buf.data() + ls + 16, 8);
  const size_t tab = ls + 24;
  if (buf.size() < tab + 8 * (size_t)kV22Sections) { cth_set_error("%s: truncated address table", path); return nullptr; }
  std::vector<uint64_t> addr(kV22Sections);
  memcpy(addr.data(), buf.data() + tab, 8 * (size_t)kV22Sections);
  // section indices in the order of the address table
  enum { sLen = 0, sENO = 1, sBkg = 2, sTrn = sBkg + 20, sTrg = sTrn + 7, sCV = sTrg + 20, sCVprior = sCV + 19,
         sCVnorm2, sDst, sSS = sDst + 100, sHDPprb = sSS + 3, sHDPind, sAA, sSSstate, sDescEnd, sDesc };
  static_assert(sDesc + 1 == kV22Sections, "address table size");
  const size_t npos = (size_t)dbsize;
  // every section must lie inside the file (sizes checked before they are
  // multiplied, offsets before they are added)
  if (nprof > buf.size() / 4 || npos > buf.size() / 4 || nprof > (size_t)INT32_MAX) {
    cth_set_error("%s: header counts %zu profiles, %zu positions", path, nprof, npos);
    return nullptr;
  }
  auto need = [&](int s, size_t bytes) { return addr[s] <= buf.size() && bytes <= buf.size() - addr[s]; };
  const size_t ntr = npos + nprof;
  bool inside = need(sLen, 4 * nprof) && need(sENO, 4 * nprof) && need(sAA, npos) && need(sSSstate, npos);
  for (int r = 0; r < CTH_NAA; r++) inside = inside && need(sBkg + r, 4 * nprof) && need(sTrg + r, 4 * npos);
  for (int t = 0; t < CTH_NTRN; t++) inside = inside && need(sTrn + t, 4 * ntr);
  for (int k = 0; k < CTH_NCV; k++) inside = inside && need(sCV + k, 4 * npos);
  inside = inside && need(sCVprior, 4 * npos) && need(sCVnorm2, 4 * npos);
  for (int k = 0; k <= kMaxDst; k++) inside = inside && need(sDst + k, 4 * npos);
  for (int q = 0; q < 3; q++) inside = inside && need(sSS + q, 4 * npos);
  if (!inside) { cth_set_error("%s: truncated sections", path); return nullptr; }
  std::vector<int32_t> lens(nprof);
  memcpy(lens.data(), buf.data() + addr[sLen], 4 * nprof);
  size_t tot = 0;
  for (int32_t l : lens) {
    if (l < 1) { cth_set_error("%s: profile length %d", path, (int)l); return nullptr; }
    tot += (size_t)l;
  }
  if (tot != npos) { cth_set_error("%s: lengths sum to %zu, header says %zu", path, tot, npos); return nullptr; }
  // the packed fields' arrays in the file (the position-address fields are
  // not stored: the reader fills them, TdDataReader.cpp:953-978)
  const char* f[fTot] = {};
  const char* b = buf.data();
  for (int r = 0; r < CTH_NAA; r++) { f[fBkg + r] = b + addr[sBkg + r]; f[fTrg + r] = b + addr[sTrg + r]; }
  f[fENO] = b + addr[sENO];
  f[fLen] = b + addr[sLen];
  for (int t = 0; t < CTH_NTRN; t++) f[fTrn + t] = b + addr[sTrn + t];
  for (int k = 0; k < CTH_NCV; k++) f[fCV + k] = b + addr[sCV + k];
  f[fCVprior] = b + addr[sCVprior];
  f[fCVnorm2] = b + addr[sCVnorm2];
  for (int k = 0; k <= kMaxDst; k++) f[fDD + k] = b + addr[sDst + k];
  for (int q = 0; q < 3; q++) f[fSS + q] = b + addr[sSS + q];
  f[fAA] = b + addr[sAA];
  f[fSSstate] = b + addr[sSSstate];
  cth_pset* o = decode_fields(f, nprof, npos, lens.data());
  if (!o) return nullptr;
  // descriptions: NUL-terminated strings from addr[sDesc] on, profile k ending
  // (after its NUL) at the file address held in entry k of the end-address section
  if (need(sDescEnd, 8 * nprof)) {
    uint64_t beg = addr[sDesc];
    for (size_t p = 0; p < nprof; p++) {
      uint64_t end;
      memcpy(&end, buf.data() + addr[sDescEnd] + 8 * p, 8);
      if (end < beg || end > buf.size()) { cth_pset_free(o); cth_set_error("%s: bad description address", path); return nullptr; }
      const std::string d(buf.data() + beg, strnlen(buf.data() + beg, end - beg));
      if (cth_pset_set_text(o, (int)p, nullptr, d.c_str())) { cth_pset_free(o); return nullptr; }
      beg = end;
    }
  }
  return o;
}

// A profile set from the reference's packed profile-model buffers: pmbeg[k]
// points at field k's array (181 fields, cupro/PM2DVectorFields.h:53-87, as
// TdDataReader::ReadDataChunkHelper fills them for a database chunk,
// libmycu/cupro/TdDataReader.cpp:703-980, and PMBatchProData for the query
// batch).  nprof < 0: every profile, counted from pmend (the end of each
// field's array); nprof >= 0: the first nprof profiles from pmbeg, as a
// worker reads one query after PMBatchProData::PMDataSkipNPros has advanced
// the pointers (cualn/DevCommThread.cpp:314-320; pmend may then be NULL, and
// when given must bound at least those profiles).  szfields (optional)
// holds each field's byte offset from pmbeg[0] (szCpm2dvfields), the offsets
// at which CuDeviceMemory::CopyCPMDataToDevice addresses the fields of the
// one contiguous buffer (cupro/CuDeviceMemory.cu:1095-1141); they are
// checked against pmbeg.  descs (optional) holds one description per profile.
extern "C" cth_pset* cth_pset_from_packed(char* const* pmbeg, char* const* pmend, const size_t* szfields, int nprof,
                                          const char* const* descs) {
  if (!pmbeg || (nprof < 0 && !pmend)) { cth_set_error("packed: null field pointers"); return nullptr; }
  for (int k = 0; k < fTot; k++)
    if (!pmbeg[k] || (pmend && (!pmend[k] || pmend[k] < pmbeg[k]))) {
      cth_set_error("packed: field %d unset", k);
      return nullptr;
    }
  if (szfields)
    for (int k = 0; k < fTot; k++)
      if (pmbeg[k] != pmbeg[0] + szfields[k]) {
        cth_set_error("packed: field %d at offset %lld, szfields says %zu", k, (long long)(pmbeg[k] - pmbeg[0]),
                      szfields[k]);
        return nullptr;
      }
  auto bytes = [&](int k) { return (size_t)(pmend[k] - pmbeg[k]); };
  size_t np = 0;
  if (nprof < 0) {
    if (bytes(fLen) % 4) { cth_set_error("packed: length field of %zu bytes", bytes(fLen)); return nullptr; }
    np = bytes(fLen) / 4;
  } else {
    np = (size_t)nprof;
  }
  if (np > (size_t)INT32_MAX) { cth_set_error("packed: %zu profiles", np); return nullptr; }
  if (pmend && bytes(fLen) < 4 * np) { cth_set_error("packed: %zu profiles beyond the length field", np); return nullptr; }
  std::vector<int32_t> lens(np);
  if (np) memcpy(lens.data(), pmbeg[fLen], 4 * np);
  size_t npos = 0;
  uint32_t d0 = 0;
  for (size_t p = 0; p < np; p++) {
    if (lens[p] < 1 || lens[p] > CTH_MAXCOLUMNS) {
      cth_set_error("packed: profile %zu has length %d", p, (int)lens[p]);
      return nullptr;
    }
    // the offsets count positions from the first profile of the buffer
    uint32_t d;
    memcpy(&d, pmbeg[fDist] + 4 * p, 4);
    if (p == 0) d0 = d;
    if (d - d0 != npos) {
      cth_set_error("packed: profile %zu at position %u, the lengths give %zu", p, d - d0, npos);
      return nullptr;
    }
    npos += (size_t)lens[p];
  }
  // each field's size (FPTYPE, INTYPE and LNTYPE are 4 bytes, CHTYPE 1)
  if (pmend)
    for (int k = 0; k < fTot; k++) {
      size_t want = 4 * npos;
      if (k < fTrn) want = 4 * np;
      else if (k < fTrg) want = 4 * (npos + np);
      else if (k == fAA || k == fSSstate) want = npos;
      if (nprof < 0 ? bytes(k) != want : bytes(k) < want) {
        cth_set_error("packed: field %d holds %zu bytes, %zu profiles of %zu positions need %zu", k, bytes(k), np,
                      npos, want);
        return nullptr;
      }
    }
  const char* f[fTot];
  for (int k = 0; k < fTot; k++) f[k] = pmbeg[k];
  cth_pset* o = decode_fields(f, np, npos, lens.data());
  if (!o) return nullptr;
  if (descs)
    for (size_t p = 0; p < np; p++)
      if (cth_pset_set_text(o, (int)p, nullptr, descs[p])) { cth_pset_free(o); return nullptr; }
  return o;
}

extern "C" int cth_pset_nprof(const cth_pset* ps) { return ps ? ps->nprof : 0; }
extern "C" int64_t cth_pset_npos(const cth_pset* ps) { return ps ? ps->npos : 0; }
