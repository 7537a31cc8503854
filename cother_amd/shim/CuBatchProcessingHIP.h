// CuBatchProcessingHIP.h: the reference-side binding of libcother_hip.
//
// A C++ class with the method set through which the reference's worker
// threads drive a search (class CuBatchProcessing, src/libmycu/cupro/
// CuBatchProcessing.cuh:47-112, called by DevCommThread::Execute and
// DevCommThread::ProcessScoreMatrix, src/libmycu/cualn/DevCommThread.cpp:
// 155-165,296-356), forwarding to the C-ABI of include/cother_hip.h.  It
// compiles against the reference's own headers (no CUDA): a maintainer
// builds it into the reference in place of src/libmycu/cupro/
// CuBatchProcessing.cu and links libcother_hip.so; the worker's one
// construction line passes the device id (dmem_->GetDeviceProp().devid_)
// and the directory of the score tables instead of the CuDeviceMemory
// object.  The per-hit records go to the reference's AlnWriter as its
// finalizer thread hands them over (CuBatchProcessingFinalizer::
// PassResultsToWriter, src/libmycu/cupro/CuBatchProcessingFinalizer.cpp:
// 248-270), so the writer, the job dispatcher and the command line stay as
// they are.
#ifndef COTHER_AMD_CUBATCHPROCESSINGHIP_H
#define COTHER_AMD_CUBATCHPROCESSINGHIP_H

#include <stddef.h>

#include <vector>

#include "liblib/mybase.h"
#include "libmycu/cualn/AlnWriter.h"
#include "libpro/srcpro/Configuration.h"
#include "tsafety/TSCounterVar.h"

#include "cother_hip.h"

class CuBatchProcessingHIP {
public:
  // device: the HIP device of this worker; envdir: the directory holding the
  // score tables (ssss, cvs2s); the other arguments are those of
  // CuBatchProcessing's constructor (CuBatchProcessing.cuh:50-58).  The
  // statistical configuration is the library's (the reference's cother.par
  // values), so config is not consulted.
  CuBatchProcessingHIP(int device, const char* envdir, int dareano, Configuration config[NoCTypes],
                       AlnWriter* writer, const mystring* queryfnames, const mystring* querydescs,
                       const char** bdb1descs);
  ~CuBatchProcessingHIP();

  // CuDeviceMemory::TransferCPMData (CuBatchProcessing.cuh:62-70): the packed
  // chunk becomes the resident database of this worker's context
  void TransferCPMData(char** bdbCpmbeg, char** bdbCpmend, size_t* szCpm2dvfields);

  void SetDbDetails(uint64_mt db_len, size_t no_sequences) { dblength_ = db_len; ndbsequences_ = no_sequences; }
  size_t GetQueryLen() const { return querylength_; }
  void SetQueryLen(size_t value) { querylength_ = value; }
  float GetScoreThreshold() const { return scorethld_; }
  void SetScoreThreshold(float value) { scorethld_ = value; }
  float GetLogEThreshold() const { return logevthld_; }
  void SetLogEThreshold(float value) { logevthld_ = value; }

  // The search of query qrysernr (its packed fields at querypmbeg) against
  // the database chunk (CuBatchProcessing::ProcessScoreMatrix,
  // CuBatchProcessing.cu:189-226): one cth_search, then the per-hit records
  // in the output format of the command line (plain or JSON) handed to the
  // writer.  bdb1pm* is the database cached on the device when no chunk is
  // given (bdbCpm* NULL).
  void ProcessScoreMatrix(int qrysernr, size_t nqyposs, char** querypmbeg, char** bdb1pmbeg, char** bdb1pmend,
                          const char** bdbCdesc, char** bdbCpmbeg, char** bdbCpmend, size_t* szCpm2dvfields,
                          TSCounterVar* cnt);

  // Results are handed to the writer before ProcessScoreMatrix returns.
  void WaitForIdleChilds() {}

private:
  void Check(int rc, const char* what) const;

  cth_ctx* ctx_ = nullptr;
  AlnWriter* writer_;
  const mystring* queryfnames_;
  const mystring* querydescs_;
  const char** bdb1descs_;
  cth_pset* db_ = nullptr;     // host copy of the resident database (descriptions for the records)
  cth_pset* query_ = nullptr;  // the query being searched
  bool db_is_bdb1_ = false;
  bool db_descs_set_ = false;
  std::vector<cth_hit> hits_;
  std::vector<char> alnbuf_;
  uint64_mt dblength_ = 0;
  size_t ndbsequences_ = 0;
  size_t querylength_ = 0;
  float scorethld_ = 0.0f;
  float logevthld_ = 0.0f;
};

#endif
