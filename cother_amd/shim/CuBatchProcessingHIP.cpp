// CuBatchProcessingHIP.cpp: the reference-side binding of libcother_hip
// (see CuBatchProcessingHIP.h).
#include "CuBatchProcessingHIP.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <memory>
#include <string>
#include <utility>

#include "liblib/myexception.h"
#include "liblib/mystring.h"
#include "libpro/srcpro/CLOptions.h"
#include "libpro/srcpro/MOptions.h"

namespace {

// the alignment rows of one search: grown when a search reports them full
constexpr size_t kInitialAlnCap = (size_t)64 << 20;

}  // namespace

void CuBatchProcessingHIP::Check(int rc, const char* what) const {
  if (rc) throw MYRUNTIME_ERROR(mystring("CuBatchProcessingHIP::") + what + ": " + cth_last_error());
}

CuBatchProcessingHIP::CuBatchProcessingHIP(int device, const char* envdir, int /*dareano*/,
                                           Configuration /*config*/[NoCTypes], AlnWriter* writer,
                                           const mystring* queryfnames, const mystring* querydescs,
                                           const char** bdb1descs)
    : writer_(writer), queryfnames_(queryfnames), querydescs_(querydescs), bdb1descs_(bdb1descs) {
  Check(cth_ctx_create(device, envdir, &ctx_), "constructor");
  alnbuf_.resize(kInitialAlnCap);
}

CuBatchProcessingHIP::~CuBatchProcessingHIP() {
  cth_ctx_destroy(ctx_);
  cth_pset_free(db_);
  cth_pset_free(query_);
}

void CuBatchProcessingHIP::TransferCPMData(char** bdbCpmbeg, char** bdbCpmend, size_t* szCpm2dvfields) {
  cth_pset* db = cth_pset_from_packed(bdbCpmbeg, bdbCpmend, szCpm2dvfields, -1, nullptr);
  if (!db) Check(-1, "TransferCPMData");
  cth_pset_free(db_);
  db_ = db;
  db_is_bdb1_ = false;
  db_descs_set_ = false;
  Check(cth_upload_db(ctx_, db_), "TransferCPMData");
}

void CuBatchProcessingHIP::ProcessScoreMatrix(int qrysernr, size_t nqyposs, char** querypmbeg, char** bdb1pmbeg,
                                              char** bdb1pmend, const char** bdbCdesc, char** bdbCpmbeg,
                                              char** /*bdbCpmend*/, size_t* /*szCpm2dvfields*/, TSCounterVar* cnt) {
  if (bdbCpmbeg && bdb1pmbeg)
    throw MYRUNTIME_ERROR("CuBatchProcessingHIP::ProcessScoreMatrix: a cached database and a chunk at once");
  // the database: the chunk TransferCPMData made resident, or the cached one
  if (!bdbCpmbeg) {
    if (!bdb1pmbeg || !bdb1pmend) throw MYRUNTIME_ERROR("CuBatchProcessingHIP::ProcessScoreMatrix: no database");
    if (!db_is_bdb1_) {
      cth_pset* db = cth_pset_from_packed(bdb1pmbeg, bdb1pmend, nullptr, -1, bdb1descs_);
      if (!db) Check(-1, "ProcessScoreMatrix");
      cth_pset_free(db_);
      db_ = db;
      Check(cth_upload_db(ctx_, db_), "ProcessScoreMatrix");
      db_is_bdb1_ = true;
      db_descs_set_ = true;
    }
  } else if (!db_) {
    throw MYRUNTIME_ERROR("CuBatchProcessingHIP::ProcessScoreMatrix: no chunk transferred");
  } else if (!db_descs_set_ && bdbCdesc) {
    const int n = cth_pset_nprof(db_);
    for (int k = 0; k < n; k++) Check(cth_pset_set_text(db_, k, nullptr, bdbCdesc[k]), "ProcessScoreMatrix");
    db_descs_set_ = true;
  }
  // the query: its packed fields begin at querypmbeg (PMDataSkipNPros has
  // advanced them to query qrysernr, DevCommThread.cpp:314-320)
  cth_pset* q = cth_pset_from_packed(querypmbeg, nullptr, nullptr, 1, nullptr);
  if (!q) Check(-1, "ProcessScoreMatrix");
  cth_pset_free(query_);
  query_ = q;
  if ((size_t)cth_pset_npos(query_) != nqyposs)
    throw MYRUNTIME_ERROR("CuBatchProcessingHIP::ProcessScoreMatrix: query length differs from nqyposs");
  Check(cth_upload_queries(ctx_, query_), "ProcessScoreMatrix");
  // the search with the whole database's size and count (SetDbDetails)
  cth_search_opts o;
  memset(&o, 0, sizeof o);
  o.eval = MOptions::GetEVAL();
  o.ssemodel = MOptions::GetSSEMODEL();
  o.do_map = MOptions::GetMAPALN();
  o.minpp = MOptions::GetMINPP();
  o.dbsize = dblength_;
  o.ndb = ndbsequences_;
  const int ndb = cth_pset_nprof(db_);
  hits_.resize(ndb > 0 ? (size_t)ndb : 1);
  int nh = 0;
  for (;;) {
    o.alnbuf = alnbuf_.data();
    o.alncap = alnbuf_.size();
    if (cth_search(ctx_, 0, 1, &o, hits_.data(), (int)hits_.size(), &nh, nullptr) == 0) break;
    if (!strstr(cth_last_error(), "alignment buffer full")) Check(-1, "ProcessScoreMatrix");
    alnbuf_.resize(alnbuf_.size() * 2);
  }
  // the finalizer's per-hit records (CompressResultsPlain / CompressResultsJSON)
  cth_outinfo oi;
  cth_outinfo_init(&oi);
  const std::string qname = queryfnames_ ? queryfnames_[qrysernr].c_str() : "";
  const std::string qdesc = querydescs_ ? querydescs_[qrysernr].c_str() : "";
  oi.qname = qname.c_str();
  oi.qdesc = qdesc.c_str();
  oi.qrysernr = qrysernr;
  oi.logevthld = logevthld_;
  oi.dbsize = dblength_;
  oi.ndbentries = ndbsequences_;
  oi.nohits = MOptions::GetNOHITS();
  oi.noalns = MOptions::GetNOALNS();
  oi.dsclen = MOptions::GetDSCLEN();
  oi.dscwidth = MOptions::GetDSCWIDTH();
  oi.alnwidth = MOptions::GetALNWIDTH();
  oi.show = MOptions::GetSHOW();
  oi.ssswgt = MOptions::GetSSSWGT();
  const bool json = CLOptions::GetB_FMT() == CLOptions::ofJSON;
  std::vector<std::pair<char*, char*>> recs;
  size_t szannot = 0, szalns = 0;
  for (int k = 0; k < nh; k++) {
    const cth_hit& h = hits_[k];
    if (logevthld_ < h.logE || h.alnlen < 1) continue;
    char *annot = nullptr, *record = nullptr;
    const char* rows = alnbuf_.data() + h.alnoff;
    const int rc = json ? cth_format_hit_json(query_, 0, db_, &h, rows, &oi, &annot, &record)
                        : cth_format_hit_plain(query_, 0, db_, &h, rows, &oi, &annot, &record);
    Check(rc, "ProcessScoreMatrix");
    szannot += strlen(annot) + 1;
    szalns += strlen(record) + 1;
    recs.emplace_back(annot, record);
  }
  std::unique_ptr<char, MyDataDestroyer> annotations, alignments;
  std::unique_ptr<std::vector<int>> srtindxs(new std::vector<int>);
  std::unique_ptr<std::vector<float>> logevalues(new std::vector<float>);
  std::unique_ptr<std::vector<char*>> alnptrs(new std::vector<char*>);
  std::unique_ptr<std::vector<char*>> annotptrs(new std::vector<char*>);
  if (!recs.empty()) {
    annotations.reset((char*)malloc(szannot));
    alignments.reset((char*)malloc(szalns));
    if (!annotations || !alignments) throw MYRUNTIME_ERROR("CuBatchProcessingHIP::ProcessScoreMatrix: no memory");
    char* ap = annotations.get();
    char* op = alignments.get();
    int sernr = 0;
    for (int k = 0; k < nh; k++) {
      const cth_hit& h = hits_[k];
      if (logevthld_ < h.logE || h.alnlen < 1) continue;
      auto& r = recs[sernr];
      // hits come sorted by log E-value (ties by database index)
      srtindxs->push_back(sernr++);
      logevalues->push_back(h.logE);
      annotptrs->push_back(ap);
      alnptrs->push_back(op);
      const size_t la = strlen(r.first) + 1, lr = strlen(r.second) + 1;
      memcpy(ap, r.first, la);
      memcpy(op, r.second, lr);
      ap += la;
      op += lr;
      cth_free_text(r.first);
      cth_free_text(r.second);
    }
  }
  uint64_t sspace = 0, deltalen = 0;
  cth_search_space(nqyposs, dblength_, ndbsequences_, &sspace, &deltalen);
  if (writer_)
    writer_->PushPartOfResults(qrysernr, (int)nqyposs, queryfnames_ ? queryfnames_[qrysernr].c_str() : nullptr,
                               querydescs_ ? querydescs_[qrysernr].c_str() : nullptr, (unsigned int)deltalen,
                               (float)sspace, logevthld_, std::move(annotations), std::move(alignments),
                               std::move(srtindxs), std::move(logevalues), std::move(alnptrs), std::move(annotptrs));
  if (cnt) cnt->dec();
}
