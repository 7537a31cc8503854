// resout.cpp: the reference's plain-text results file for one query.
//
// The reference formats every reported profile into two records in its
// finalizer thread (CuBatchProcessingFinalizer::CompressResultsPlain,
// libmycu/cupro/CuBatchProcessingFinalizerPlain.cpp:36-403): a one-line
// annotation for the hit list and an alignment record (description, scores,
// the alignment in blocks of ALNWIDTH columns, statistical parameters), and
// its writer thread assembles the file from them (AlnWriter::WriteResultsPlain
// and helpers, libmycu/cualn/AlnWriterPlain.cpp:19-398; merge and file name,
// libmycu/cualn/AlnWriter.cpp:250-376): program name, query description,
// database summary, the first NOHITS annotations, the first NOALNS alignment
// records, reference statistics.  Descriptions are wrapped by the reference's
// word-wrapping rule (liblib/fmtdescription.h:35-83), restated here as
// wrap_text() because its line breaks and truncation are part of the format.
//
// The JSON format follows the same two stages (CompressResultsJSON,
// cupro/CuBatchProcessingFinalizerJSON.cpp:35-383; AlnWriter::
// WriteResultsJSON, cualn/AlnWriterJSON.cpp:19-429), with the reference's
// escaping rule for strings (FormatDescriptionJSON, liblib/fmtdescription.h:
// 94-128) restated as json_text().  Kept as the reference writes it: the
// separator after the last alignment record is chosen by NOHITS, not NOALNS
// (AlnWriterJSON.cpp:112-118), so NOALNS < NOHITS leaves a trailing comma.
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "cth_consts.h"
#include "cth_pset.h"
#include "cother_hip.h"

void cth_set_error(const char* fmt, ...);

namespace {

constexpr int kIndent = 13;          // OUTPUTINDENT (liblib/mylimits.h:56)
constexpr int kAnnotLen = 75;        // ANNOTATIONLEN (liblib/mylimits.h:59)
constexpr int kWriterBuffer = 4096;  // AlnWriter::szWriterBuffer (cualn/AlnWriter.h:38-41)
constexpr int kTmpBuffer = 1024;     // AlnWriter::szTmpBuffer
constexpr float kLn2 = 0.69314718055994530941723212146f;  // SLC_LN2 (extsp/psl.h:29)
constexpr float kLogDpMax = 7.0978271289338397e+02f;      // SLC_LOG_DP_MAX (extsp/psl.h:49)

void appendf(std::string& s, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void appendf(std::string& s, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  const int n = vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (n < (int)sizeof buf) {
    s.append(buf, n);
    return;
  }
  std::vector<char> big(n + 1);
  va_start(ap, fmt);
  vsnprintf(big.data(), big.size(), fmt, ap);
  va_end(ap);
  s.append(big.data(), n);
}

inline bool blank(char c) { return c == ' ' || c == '\t'; }

// The description layout rule: leading white space dropped, every run of
// blanks becomes one separator which is a space or, when the next word would
// reach `width` on the current line (and width < maxlen), a line break
// followed by `indent` spaces; output stops once `outpos` reaches `maxlen`,
// and then its last three characters become "...".  `outpos` and `linepos`
// carry over between calls.
void wrap_text(std::string& out, const char* text, int maxlen, int indent, int width, int& outpos,
               int& linepos, char lead) {
  while (*text == ' ' || *text == '\t' || *text == '\n' || *text == '\r') text++;
  if (lead) {
    out += lead;
    outpos++;
    linepos++;
  }
  while (*text && outpos < maxlen) {
    if (blank(*text)) {
      while (blank(*text)) text++;
      bool brk = false;
      if (width < maxlen) {
        int col = linepos + 1;
        for (const char* w = text; *w && !blank(*w) && col < width; w++) col++;
        brk = width <= col;
      }
      if (brk) {
        out += '\n';
        out.append(indent, ' ');
        outpos += 1 + indent;
        linepos = indent;
      } else {
        out += ' ';
        outpos++;
        linepos++;
      }
      continue;
    }
    out += *text++;
    outpos++;
    linepos++;
  }
  if (maxlen <= outpos && 3 < maxlen)
    for (size_t k = 1; k <= 3 && k <= out.size(); k++) out[out.size() - k] = '.';
}

// The JSON string rule: leading white space dropped, each run of blanks one
// space, '"', '\\' and '/' escaped with a backslash; output stops once `outpos`
// reaches `maxlen`, and then the last three characters, and any backslashes
// before them, become '.'.
void json_text(std::string& out, const char* text, int maxlen, int& outpos) {
  while (*text == ' ' || *text == '\t' || *text == '\n' || *text == '\r') text++;
  while (*text && outpos < maxlen) {
    if (blank(*text)) {
      while (blank(*text)) text++;
      out += ' ';
      outpos++;
      continue;
    } else if (*text == '"' || *text == '\\' || *text == '/') {
      out += '\\';
      outpos++;
      if (maxlen <= outpos) break;
    }
    out += *text++;
    outpos++;
  }
  if (maxlen <= outpos && 3 < maxlen) {
    int i = 1;
    for (; i <= 3 && i <= (int)out.size(); i++) out[out.size() - i] = '.';
    for (; i <= outpos && i <= (int)out.size() && out[out.size() - i] == '\\'; i++) out[out.size() - i] = '.';
  }
}

inline unsigned coord_x(uint32_t xy) { return (xy >> 16) & 0xffff; }
inline unsigned coord_y(uint32_t xy) { return xy & 0xffff; }

bool has_ss(const cth_pset* ps, int k) { return ps->sss && ps->len[k] > 0 && ps->sss[ps->dist[k]] != 0; }

const char* text_or(const char* s, const char* dflt) { return s ? s : dflt; }

struct Meta {
  int qlen;
  float qeno;
  bool qss;
  const char* qname;
  const char* qdesc;
  float sspace;      // effective search space as the reference passes it (float)
  uint64_t deltalen; // length adjustment
};

// CuBatchProcessingFinalizer::GetEvalue/GetPvalue/GetBitScore
// (cupro/CuBatchProcessingFinalizer.h:179-188)
double evalue_of(float loge) { return kLogDpMax < loge ? 1.0e+308 : exp((double)loge); }
double pvalue_of(double e) { return e < 0.01 ? e : 1.0 - exp(-e); }
float bits_of(float sspace, float loge) { return sspace <= 0.0f ? 0.0f : (logf(sspace) - loge) / kLn2; }

// MakeAnnotationPlain (CuBatchProcessingFinalizerPlain.cpp:161-186)
void annotation(std::string& out, const char* desc, float score, double evalue) {
  const size_t start = out.size();
  int outpos = 0, linepos = 0;
  if (desc) wrap_text(out, desc, kAnnotLen, 0, kAnnotLen, outpos, linepos, 0);
  const int written = (int)(out.size() - start);
  if (written < kAnnotLen) out.append(kAnnotLen - written, ' ');
  appendf(out, " %7.0f %7.1g\n", score, evalue);
}

// The alignment record of one hit (CompressResultsPlain with FormatScoresPlain,
// FormatAlignmentPlain and FormatFooterPlain, CuBatchProcessingFinalizerPlain.cpp:
// 110-137, 188-403)
void record(std::string& out, const Meta& m, const cth_pset* db, const cth_hit& h, const char* rows,
            const cth_outinfo& o) {
  const int p = h.profile;
  const int tlen = db->len[p];
  const char* desc = cth_pset_desc(db, p);
  const int width = kIndent < o.dscwidth ? o.dscwidth : 1;
  int outpos = 0, linepos = 0;
  if (desc) wrap_text(out, desc, o.dsclen, kIndent, width, outpos, linepos, '>');
  out += '\n';
  // scores
  const unsigned alnlen = (unsigned)h.alnlen;
  const double evalue = evalue_of(h.logE);
  appendf(out, "  Length/ENO: Query = %d/%.1f, Sbjct = %d/%.1f\n\n", m.qlen, m.qeno, tlen, db->eno[p]);
  appendf(out, " Score = %.2f (%.1f bits),  Expect = %.2g, P-value = %.2g\n", h.score,
          bits_of(m.sspace, h.logE), evalue, pvalue_of(evalue));
  const int idts = h.idnts, psts = h.pstvs_map, gaps = h.ngaps;
  if (idts > 0) appendf(out, " Identities = %d/%d (%d%%)", idts, alnlen, (int)(idts * 100u / alnlen));
  if (psts > 0) {
    if (idts) out += ',';
    appendf(out, " Positives = %d/%d (%d%%)", psts, alnlen, (int)(psts * 100u / alnlen));
  }
  if (gaps > 0) {
    if (idts || psts) out += ',';
    appendf(out, " Gaps = %d/%d (%d%%)", gaps, alnlen, (int)(gaps * 100u / alnlen));
  }
  out += "\n\n";
  // alignment blocks
  const bool ss = o.ssswgt > 0.0f && m.qss && has_ss(db, p);
  unsigned qpos = coord_y(h.mapbeg) + 1, tpos = coord_x(h.mapbeg) + 1;
  const char* rq = rows;
  const char* rm = rows + alnlen;
  const char* rt = rows + 2 * (size_t)alnlen;
  const char* rqs = rows + 3 * (size_t)alnlen;
  const char* rts = rows + 4 * (size_t)alnlen;
  for (int f = 0; f < (int)alnlen; f += o.alnwidth) {
    const int nb = std::min<int>(o.alnwidth, (int)alnlen - f);
    auto seq_line = [&](const char* label, const char* r, unsigned& pos) {
      appendf(out, "%s %5u ", label, pos);
      out.append(r + f, nb);
      const int g = (int)std::count(r + f, r + f + nb, '-');
      pos += nb - g;
      appendf(out, " %-5d\n", nb <= g ? (int)pos : (int)pos - 1);
    };
    if (ss) {
      appendf(out, "%-13s", "struct");
      out.append(rqs + f, nb);
      out += '\n';
    }
    seq_line("Query:", rq, qpos);
    out.append(13, ' ');
    out.append(rm + f, nb);
    out += '\n';
    seq_line("Sbjct:", rt, tpos);
    if (ss) {
      appendf(out, "%-13s", "struct");
      out.append(rts + f, nb);
      out += '\n';
    }
    out += '\n';
  }
  // statistical parameters
  if (o.show) {
    auto val = [](float v, bool ok) {
      char b[32];
      if (ok && v > 0.0f) snprintf(b, sizeof b, "%6.4f", v);
      else snprintf(b, sizeof b, "n/a");
      return std::string(b);
    };
    const bool ok = h.expected < 0.0f;
    appendf(out, "%25c  %-6s   %-6s\n", ' ', "K", "Lambda");
    appendf(out, "%-25s  %6s   %6s\n", "Computed  ungapped,", val(h.K, ok).c_str(), val(h.lambda, ok).c_str());
    appendf(out, "%-25s  %6s   %6s\n", "Estimated gapped,", val(h.K_est, ok).c_str(), val(h.lambda_est, ok).c_str());
    appendf(out, "Entropy, %6.4f; Expected, %6.4f; Min/Max, %.0f/%-.0f\n\n", h.H, h.expected, h.minsc, h.maxsc);
  }
  out += "\n\n";
}

// WritePrognamePlain (AlnWriterPlain.cpp:131-173); no reference list is printed
void progname(std::string& out, const cth_outinfo& o) {
  const size_t start = out.size();
  if (o.progname) out += o.progname;
  if (o.progversion) appendf(out, " %s", o.progversion);
  if (o.progname || o.progversion) out += "\n\n";
  if (kWriterBuffer - (int)(out.size() - start) - 4 < 1) return;
  out += '\n';
}

// WriteQueryDescriptionPlain (AlnWriterPlain.cpp:187-228)
void query_description(std::string& out, int maxsize, const Meta& m, const cth_outinfo& o) {
  if (maxsize < 40) return;
  const size_t start = out.size();
  appendf(out, "\n Query (%d positions):\n", m.qlen);
  maxsize -= (int)(out.size() - start);
  if (maxsize < 1) return;
  int outpos = 0, linepos = 0;
  if (m.qdesc) wrap_text(out, m.qdesc, maxsize, 0, o.dscwidth, outpos, linepos, 0);
}

std::string base_name(const char* path) {
  const char* s = strrchr(path, '/');
  return s ? s + 1 : path;
}

// WriteSearchInformationPlain (AlnWriterPlain.cpp:244-325), its text bounded
// by a 1 KiB buffer
void search_information(std::string& out, const cth_outinfo& o, bool found) {
  const int maxsize = kTmpBuffer;
  const size_t start = out.size();
  auto size = [&] { return (int)(out.size() - start); };
  auto full = [&](int delta) {
    if (maxsize < size() + delta) {
      if (size() + 5 < maxsize) out += "...\n\n";
      return true;
    }
    return false;
  };
  if (full(30)) return;
  out += "\n\n Database(s):\n";
  for (int k = 0; k < o.ndbnames; k++) {
    const std::string name = base_name(o.dbnames[k]);
    size_t n = std::min<size_t>(name.size(), 100);
    if (full((int)n + 2)) return;
    out.append(name, 0, n);
    if (n < name.size() && 3 < n)
      for (int i = 1; i <= 3; i++) out[out.size() - i] = '.';
    out += '\n';
  }
  if (full(kIndent + 40)) return;
  out.append(kIndent, ' ');
  appendf(out, "%llu profiles\n", (unsigned long long)o.ndbentries);
  if (full(kIndent + 50)) return;
  out.append(kIndent, ' ');
  appendf(out, "%llu total positions\n\n\n", (unsigned long long)o.dbsize);
  if (full(kAnnotLen * 2)) return;
  if (found) {
    const char* head = " Profiles found below the e-value threshold:";
    out += head;
    for (int w = (int)strlen(head); w < kAnnotLen; w++) out += ' ';
    appendf(out, " %7s %7s\n\n", "Score", "E-value");
  } else {
    appendf(out, " No profiles found below an e-value threshold of %g.\n\n\n", exp(o.logevthld));
  }
}

// WriteSummaryPlain (AlnWriterPlain.cpp:340-397)
void summary(std::string& out, const Meta& m, const cth_outinfo& o) {
  int effq = m.qlen;
  uint64_t effdb = o.dbsize;
  const int dl = (int)m.deltalen;
  if (dl < m.qlen) effq -= dl;
  if (o.ndbentries * (uint64_t)dl < o.dbsize) effdb -= o.ndbentries * (uint64_t)dl;
  appendf(out, "%-20s  %-6s   %-6s\n", "Reference values of", "K", "Lambda");
  appendf(out, "%-20s  %6.4f   %6.4f\n", "Ungapped", o.ref_k, o.ref_lambda);
  appendf(out, "%-20s  %6.4f   %6.4f\n\n", "Gapped", o.gap_k, o.gap_lambda);
  appendf(out, "Length of query, %d\n", m.qlen);
  appendf(out, "Length of database, %llu\n", (unsigned long long)o.dbsize);
  appendf(out, "Effective length of query, %d\n", effq);
  appendf(out, "Effective length of database, %llu\n", (unsigned long long)effdb);
  appendf(out, "Effective search space, %llu\n\n", (unsigned long long)(size_t)m.sspace);
}

bool make_meta(const cth_pset* qs, int qidx, const cth_outinfo* o, Meta& m) {
  if (!qs || qidx < 0 || qidx >= qs->nprof) {
    cth_set_error("results: query index %d out of range", qidx);
    return false;
  }
  m.qlen = qs->len[qidx];
  m.qeno = qs->eno[qidx];
  m.qss = has_ss(qs, qidx);
  m.qname = o->qname ? o->qname : text_or(cth_pset_name(qs, qidx), "query");
  m.qdesc = o->qdesc ? o->qdesc : cth_pset_desc(qs, qidx);
  uint64_t ss = 0, dl = 0;
  cth_search_space((uint64_t)m.qlen, o->dbsize, o->ndbentries, &ss, &dl);
  m.sspace = (float)ss;
  m.deltalen = dl;
  return true;
}

// hits of the query below the threshold with an alignment, by ascending log E-value
// (the finalizer's filter and sort, CuBatchProcessingFinalizerPlain.cpp:87-101 and
// CuBatchProcessingFinalizer.cpp:276-292; equal values keep their input order)
std::vector<int> reported(const cth_hit* hits, int nhits, int qidx, const cth_outinfo& o) {
  std::vector<int> idx;
  for (int k = 0; k < nhits; k++)
    if (hits[k].query == qidx && !(o.logevthld < hits[k].logE) && hits[k].alnlen >= 1) idx.push_back(k);
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return hits[a].logE < hits[b].logE; });
  return idx;
}

bool check_hits(const cth_pset* db, const cth_hit* hits, const std::vector<int>& idx, const char* alnbuf,
                uint64_t alncap) {
  for (int k : idx) {
    const cth_hit& h = hits[k];
    if (h.profile < 0 || h.profile >= db->nprof) {
      cth_set_error("results: hit %d names profile %d outside the database", k, h.profile);
      return false;
    }
    if (!alnbuf || h.alnoff + 5 * (uint64_t)h.alnlen > alncap) {
      cth_set_error("results: hit %d: alignment rows outside the buffer", k);
      return false;
    }
  }
  return true;
}

char* to_c(const std::string& s, uint64_t* len) {
  char* p = (char*)malloc(s.size() + 1);
  if (!p) {
    cth_set_error("results: out of memory");
    return nullptr;
  }
  memcpy(p, s.data(), s.size());
  p[s.size()] = 0;
  if (len) *len = s.size();
  return p;
}

// MakeAnnotationJSON (CuBatchProcessingFinalizerJSON.cpp:161-187)
void annotation_json(std::string& out, const char* desc, float score, double evalue) {
  out += "    {\"summary_entry\": {\n      \"description\": \"";
  int outpos = 0;
  if (desc) json_text(out, desc, kAnnotLen, outpos);
  out += "\",\n";
  appendf(out, "      \"score\": %.0f,\n      \"evalue\": %.2g\n    }}", score, evalue);
}

// The JSON alignment record of one hit (CompressResultsJSON with
// FormatScoresJSON, FormatAlignmentJSON and FormatFooterJSON,
// CuBatchProcessingFinalizerJSON.cpp:116-145, 197-383)
void record_json(std::string& out, const Meta& m, const cth_pset* db, const cth_hit& h, const char* rows,
                 const cth_outinfo& o) {
  const int p = h.profile;
  const char* desc = cth_pset_desc(db, p);
  out += "    {\"hit_record\": {\n      \"target_description\": \"";
  int outpos = 0;
  if (desc) json_text(out, desc, o.dsclen, outpos);
  out += "\",\n";
  appendf(out,
          "      \"query_length\": %d,\n      \"query_eno\": %.1f,\n      \"target_length\": %d,\n"
          "      \"target_eno\": %.1f,\n      \"alignment\": {\n",
          m.qlen, m.qeno, db->len[p], db->eno[p]);
  const double evalue = evalue_of(h.logE);
  appendf(out, "        \"score\": %.2f,\n        \"bit_score\": %.1f,\n        \"evalue\": %.2g,\n"
               "        \"pvalue\": %.2g,\n",
          h.score, bits_of(m.sspace, h.logE), evalue, pvalue_of(evalue));
  const unsigned alnlen = (unsigned)h.alnlen;
  appendf(out, "        \"n_identities\": %d,\n        \"n_positives\": %d,\n        \"n_gaps\": %d,\n"
               "        \"aln_length\": %d,\n",
          h.idnts, h.pstvs_map, h.ngaps, (int)alnlen);
  const char* rq = rows;
  const char* rm = rows + alnlen;
  const char* rt = rows + 2 * (size_t)alnlen;
  const unsigned qbeg = coord_y(h.mapbeg) + 1, tbeg = coord_x(h.mapbeg) + 1;
  const int n = (int)alnlen;
  const int gq = (int)std::count(rq, rq + n, '-'), gt = (int)std::count(rt, rt + n, '-');
  const unsigned qend = qbeg + n - gq, tend = tbeg + n - gt;
  // SS strings whenever both profiles have SS predictions (the JSON writer does not consult SSSWGT)
  const int ss = m.qss && has_ss(db, p);
  appendf(out, "        \"query_from\": %u,\n        \"query_to\": %u,\n        \"target_from\": %u,\n"
               "        \"target_to\": %u,\n",
          qbeg, n <= gq ? qend : qend - 1, tbeg, n <= gt ? tend : tend - 1);
  appendf(out, "        \"query_secstr_available\": %d,\n        \"target_secstr_available\": %d,\n"
               "        \"query_secstr\": \"",
          ss, ss);
  if (ss) out.append(rows + 3 * (size_t)alnlen, n);
  out += "\",\n        \"target_secstr\": \"";
  if (ss) out.append(rows + 4 * (size_t)alnlen, n);
  out += "\",\n        \"query_aln\": \"";
  out.append(rq, n);
  out += "\",\n        \"target_aln\": \"";
  out.append(rt, n);
  out += "\",\n        \"middle\": \"";
  out.append(rm, n);
  out += "\",\n";
  auto val = [](float v, bool ok) {
    char b[32];
    if (ok && v > 0.0f) snprintf(b, sizeof b, "%6.4f", v);
    else snprintf(b, sizeof b, "NA");
    return std::string(b);
  };
  const bool ok = h.expected < 0.0f;
  appendf(out, "        \"computed_K_ungapped\": \"%s\",\n        \"computed_lambda_ungapped\": \"%s\",\n",
          val(h.K, ok).c_str(), val(h.lambda, ok).c_str());
  appendf(out,
          "        \"estimated_K_gapped\": \"%s\",\n        \"estimated_lambda_gapped\": \"%s\",\n"
          "        \"entropy\": %.4f,\n        \"expected_score\": %.4f,\n        \"min_score\": %.0f,\n"
          "        \"max_score\": %.0f\n",
          val(h.K_est, ok).c_str(), val(h.lambda_est, ok).c_str(), h.H, h.expected, h.minsc, h.maxsc);
  out += "      }\n    }}";
}

// WritePrognameJSON (AlnWriterJSON.cpp:154-201); an empty reference list
// gives one empty string, as the reference's does
void progname_json(std::string& out, const cth_outinfo& o) {
  appendf(out, "{\"cother_search\": {\n  \"program\": \"%s\",\n  \"version\": \"%s\",\n",
          o.progname ? o.progname : "", o.progversion ? o.progversion : "");
  if (kWriterBuffer - (int)out.size() - 1024 < 1) return;
  out += "  \"references\": [\n    \"\"\n  ],\n";
}

// WriteQueryDescriptionJSON (AlnWriterJSON.cpp:215-266)
void query_description_json(std::string& out, int maxsize, const Meta& m) {
  if (maxsize < 60) return;
  const size_t start = out.size();
  appendf(out, "  \"query\": {\n    \"length\": %d,\n", m.qlen);
  maxsize -= (int)(out.size() - start) + 256;  // BUF_MAX (liblib/mylimits.h:41)
  if (maxsize < 1) return;
  int outpos = 0;
  const char* name_open = "    \"name\": \"";
  out += name_open;
  outpos += (int)strlen(name_open);
  const char* close = "\",\n";
  out += close;
  outpos += (int)strlen(close);
  const char* desc_open = "    \"description\": \"";
  out += desc_open;
  outpos += (int)strlen(desc_open);
  if (m.qdesc) json_text(out, m.qdesc, maxsize, outpos);
  out += "\"\n  },\n";
}

// WriteSearchInformationJSON (AlnWriterJSON.cpp:281-360) in a 1 KiB budget
void search_information_json(std::string& out, const cth_outinfo& o, bool found) {
  int maxsize = kTmpBuffer;
  auto full = [&](int delta) {
    maxsize -= delta;
    return maxsize < 1;
  };
  if (full(40 * 7)) return;
  out += "  \"database\": {\n    \"names\": [\n";
  for (int k = 0; k < o.ndbnames; k++) {
    const std::string name = base_name(o.dbnames[k]);
    const size_t n = std::min<size_t>(name.size(), 99);
    if (full((int)n + 15)) {
      out += "      \"...\"\n";
      break;
    }
    out += "      \"";
    int outpos = 0;
    json_text(out, name.c_str(), (int)n + 1, outpos);
    out += (k + 1 < o.ndbnames) ? "\",\n" : "\"\n";
  }
  appendf(out, "    ],\n    \"number_of_profiles\": %llu,\n    \"number_of_positions\": %llu\n  },\n",
          (unsigned long long)o.ndbentries, (unsigned long long)o.dbsize);
  if (full(90 * 4)) return;
  if (found) out += "  \"message\": \"Profiles found below the e-value threshold:\",\n";
  else
    appendf(out, "  \"message\": \"No profiles found below an e-value threshold of %g\",\n", exp(o.logevthld));
}

// WriteSummaryJSON (AlnWriterJSON.cpp:375-429)
void summary_json(std::string& out, const Meta& m, const cth_outinfo& o) {
  int effq = m.qlen;
  uint64_t effdb = o.dbsize;
  const int dl = (int)m.deltalen;
  if (dl < m.qlen) effq -= dl;
  if (o.ndbentries * (uint64_t)dl < o.dbsize) effdb -= o.ndbentries * (uint64_t)dl;
  appendf(out,
          "  \"search_statistics\": {\n    \"reference_K_ungapped\": %.4f,\n"
          "    \"reference_lambda_ungapped\": %.4f,\n    \"reference_K_gapped\": %.4f,\n"
          "    \"reference_lambda_gapped\": %.4f,\n    \"query_length\": %d,\n    \"database_size\": %llu,\n"
          "    \"effective_query_length\": %d,\n    \"effective_database_size\": %llu,\n"
          "    \"effective_search_space\": %llu\n  }\n}}\n",
          o.ref_k, o.ref_lambda, o.gap_k, o.gap_lambda, m.qlen, (unsigned long long)o.dbsize, effq,
          (unsigned long long)effdb, (unsigned long long)(size_t)m.sspace);
}

bool format_file(const cth_pset* qs, int qidx, const cth_pset* db, const cth_hit* hits, int nhits,
                 const char* alnbuf, uint64_t alncap, const cth_outinfo* o, std::string& out, Meta& m,
                 bool json = false) {
  if (!db || !o) {
    cth_set_error("results: null database or options");
    return false;
  }
  if (o->alnwidth < 1) {
    cth_set_error("results: ALNWIDTH must be positive");
    return false;
  }
  if (!make_meta(qs, qidx, o, m)) return false;
  const std::vector<int> idx = reported(hits, nhits, qidx, *o);
  if (!check_hits(db, hits, idx, alnbuf, alncap)) return false;
  if (json) {
    // AlnWriter::WriteResultsJSON (AlnWriterJSON.cpp:19-146)
    progname_json(out, *o);
    const int left = std::min<int>(kWriterBuffer - (int)out.size(), o->dsclen);
    query_description_json(out, left, m);
    const bool found = !idx.empty();
    search_information_json(out, *o, found);
    out += "  \"search_summary\": [\n";
    const size_t n = idx.size(), nh = (size_t)std::max(0, o->nohits), na = (size_t)std::max(0, o->noalns);
    for (size_t i = 0; i < n && i < nh; i++) {
      const cth_hit& h = hits[idx[i]];
      annotation_json(out, cth_pset_desc(db, h.profile), h.score, evalue_of(h.logE));
      out += (i + 1 < n && i + 1 < nh) ? ",\n" : "\n";
    }
    out += "  ],\n  \"search_hits\": [\n";
    for (size_t i = 0; i < n && i < na; i++) {
      const cth_hit& h = hits[idx[i]];
      record_json(out, m, db, h, alnbuf + h.alnoff, *o);
      out += (i + 1 < n && i + 1 < nh) ? ",\n" : "\n";
    }
    out += "  ],\n";
    summary_json(out, m, *o);
    return true;
  }
  // AlnWriter::WriteResultsPlain (AlnWriterPlain.cpp:19-121)
  progname(out, *o);
  const int left = std::min<int>(kWriterBuffer - (int)out.size(), o->dsclen);
  query_description(out, left, m, *o);
  const bool found = !idx.empty();
  search_information(out, *o, found);
  if (found) {
    for (size_t i = 0; i < idx.size() && (int64_t)i < o->nohits; i++) {
      const cth_hit& h = hits[idx[i]];
      annotation(out, cth_pset_desc(db, h.profile), h.score, evalue_of(h.logE));
    }
    out += "\n\n\n";
    for (size_t i = 0; i < idx.size() && (int64_t)i < o->noalns; i++) {
      const cth_hit& h = hits[idx[i]];
      record(out, m, db, h, alnbuf + h.alnoff, *o);
    }
  }
  summary(out, m, *o);
  return true;
}

}  // namespace

extern "C" void cth_outinfo_init(cth_outinfo* o) {
  memset(o, 0, sizeof *o);
  o->logevthld = logf(10.0f);  // EVAL = 10
  o->nohits = 700;             // options.txt defaults (libpro/srcpro/MOptions.cpp)
  o->noalns = 700;
  o->dsclen = 1024;
  o->dscwidth = 80;
  o->alnwidth = 60;
  o->show = 1;
  o->ssswgt = 0.12f;
  o->ref_lambda = CTH_UNGAP_LAMBDA;
  o->ref_k = CTH_UNGAP_K;
  o->gap_lambda = CTH_GAPPED_LAMBDA;
  o->gap_k = CTH_GAPPED_K;
  o->progname = "cother";
  o->progversion = "2.03.03";  // cother/cother.h:23
}

namespace {

int format_hit(const cth_pset* qs, int qidx, const cth_pset* db, const cth_hit* hit, const char* rows,
               const cth_outinfo* o, char** annot, char** rec, bool json) {
  Meta m;
  if (!db || !o || !hit || !rows) {
    cth_set_error("results: null argument");
    return -1;
  }
  if (!make_meta(qs, qidx, o, m)) return -1;
  if (hit->profile < 0 || hit->profile >= db->nprof || hit->alnlen < 1 || o->alnwidth < 1) {
    cth_set_error("results: hit without an alignment or outside the database");
    return -1;
  }
  std::string a, r;
  if (json) {
    annotation_json(a, cth_pset_desc(db, hit->profile), hit->score, evalue_of(hit->logE));
    record_json(r, m, db, *hit, rows, *o);
  } else {
    annotation(a, cth_pset_desc(db, hit->profile), hit->score, evalue_of(hit->logE));
    record(r, m, db, *hit, rows, *o);
  }
  *annot = to_c(a, nullptr);
  *rec = to_c(r, nullptr);
  if (!*annot || !*rec) {
    free(*annot);
    free(*rec);
    return -1;
  }
  return 0;
}

int format_results(const cth_pset* qs, int qidx, const cth_pset* db, const cth_hit* hits, int nhits,
                   const char* alnbuf, uint64_t alncap, const cth_outinfo* o, char** text, uint64_t* len, bool json) {
  std::string out;
  Meta m;
  if (!format_file(qs, qidx, db, hits, nhits, alnbuf, alncap, o, out, m, json)) return -1;
  *text = to_c(out, len);
  return *text ? 0 : -1;
}

// AlnWriter::GetOutputFilename (cualn/AlnWriter.cpp:352-376):
// <outdir>/<query name without a short extension>__<serial>.out (.json)
int write_results(const cth_pset* qs, int qidx, const cth_pset* db, const cth_hit* hits, int nhits,
                  const char* alnbuf, uint64_t alncap, const cth_outinfo* o, const char* outdir, char* path,
                  int pathcap, bool json) {
  std::string out;
  Meta m;
  if (!format_file(qs, qidx, db, hits, nhits, alnbuf, alncap, o, out, m, json)) return -1;
  std::string fn = outdir ? outdir : "";
  if (!fn.empty() && fn.back() != '/') fn += '/';
  std::string qn = base_name(m.qname);
  const size_t dot = qn.rfind('.');
  if (dot != std::string::npos && qn.size() - dot <= 4) qn = qn.substr(0, dot);
  char tail[32];
  snprintf(tail, sizeof tail, "__%d.%s", o->qrysernr, json ? "json" : "out");
  fn += qn + tail;
  FILE* fp = fopen(fn.c_str(), "wb");
  if (!fp) {
    cth_set_error("results: cannot open %s for writing", fn.c_str());
    return -1;
  }
  const bool ok = fwrite(out.data(), 1, out.size(), fp) == out.size();
  if (fclose(fp) != 0 || !ok) {
    cth_set_error("results: write to %s failed", fn.c_str());
    return -1;
  }
  if (path && pathcap > 0) snprintf(path, pathcap, "%s", fn.c_str());
  return 0;
}

}  // namespace

extern "C" int cth_format_hit_plain(const cth_pset* qs, int qidx, const cth_pset* db, const cth_hit* hit,
                                    const char* rows, const cth_outinfo* o, char** annot, char** rec) {
  return format_hit(qs, qidx, db, hit, rows, o, annot, rec, false);
}
extern "C" int cth_format_hit_json(const cth_pset* qs, int qidx, const cth_pset* db, const cth_hit* hit,
                                   const char* rows, const cth_outinfo* o, char** annot, char** rec) {
  return format_hit(qs, qidx, db, hit, rows, o, annot, rec, true);
}

extern "C" int cth_format_results_plain(const cth_pset* qs, int qidx, const cth_pset* db, const cth_hit* hits,
                                        int nhits, const char* alnbuf, uint64_t alncap, const cth_outinfo* o,
                                        char** text, uint64_t* len) {
  return format_results(qs, qidx, db, hits, nhits, alnbuf, alncap, o, text, len, false);
}
extern "C" int cth_format_results_json(const cth_pset* qs, int qidx, const cth_pset* db, const cth_hit* hits,
                                       int nhits, const char* alnbuf, uint64_t alncap, const cth_outinfo* o,
                                       char** text, uint64_t* len) {
  return format_results(qs, qidx, db, hits, nhits, alnbuf, alncap, o, text, len, true);
}

extern "C" int cth_write_results_plain(const cth_pset* qs, int qidx, const cth_pset* db, const cth_hit* hits,
                                       int nhits, const char* alnbuf, uint64_t alncap, const cth_outinfo* o,
                                       const char* outdir, char* path, int pathcap) {
  return write_results(qs, qidx, db, hits, nhits, alnbuf, alncap, o, outdir, path, pathcap, false);
}
extern "C" int cth_write_results_json(const cth_pset* qs, int qidx, const cth_pset* db, const cth_hit* hits,
                                      int nhits, const char* alnbuf, uint64_t alncap, const cth_outinfo* o,
                                      const char* outdir, char* path, int pathcap) {
  return write_results(qs, qidx, db, hits, nhits, alnbuf, alncap, o, outdir, path, pathcap, true);
}

extern "C" void cth_free_text(char* p) { free(p); }
