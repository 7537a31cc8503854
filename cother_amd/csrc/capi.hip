// capi.hip: the C-ABI of libcother_hip (declared in include/cother_hip.h).
//
// A context owns one device, one HIP stream, the score tables, the uploaded
// query set and the resident template database, plus scratch buffers that
// grow on demand.  Errors are reported as negative return codes with a
// message available from cth_last_error(); nothing throws across the ABI.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "cother_hip.h"
#include "cth_ctx.h"
#include "cth_dp.h"
#include "cth_map.h"
#include "cth_stats.h"
#include <vector>

namespace {
thread_local char g_err[1024] = "";
}

void cth_set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

extern "C" const char* cth_last_error(void) { return g_err; }

extern "C" int cth_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int cth_ctx_reserve(cth_ctx* c, void** p, size_t* cap, size_t need) {
  if (need <= *cap) return 0;
  if (*p) CTH_HIP_CHECK(hipFree(*p));
  *p = nullptr;
  *cap = 0;
  size_t n = need + need / 8 + 4096;
  CTH_HIP_CHECK(hipMalloc(p, n));
  *cap = n;
  (void)c;
  return 0;
}

extern "C" int cth_ctx_create(int device, const char* envdir, cth_ctx** out) {
  *out = nullptr;
  int ndev = 0;
  CTH_HIP_CHECK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) { cth_set_error("device %d not present (%d devices)", device, ndev); return -1; }
  CTH_HIP_CHECK(hipSetDevice(device));
  cth_ctx* c = (cth_ctx*)calloc(1, sizeof(cth_ctx));
  if (!c) { cth_set_error("out of host memory"); return -1; }
  new (&c->nshared) std::atomic<int>(0);
  new (&c->tune_mu) std::mutex();
  c->device = device;
  int rc = cth_tables_load(envdir, &c->htab);
  if (rc) { cth_set_error("cannot load score tables from %s (code %d)", envdir, rc); free(c); return -1; }
  if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
    cth_set_error("hipStreamCreate failed"); free(c); return -1;
  }
  if (hipMalloc((void**)&c->dtab, sizeof(cth_tables)) != hipSuccess ||
      hipMemcpy(c->dtab, &c->htab, sizeof(cth_tables), hipMemcpyHostToDevice) != hipSuccess) {
    cth_set_error("cannot upload score tables"); cth_ctx_destroy(c); return -1;
  }
  // the device address of cth_search_noalloc's scratch array (once per device)
  if (cth_scratch_init(device)) { cth_ctx_destroy(c); return -1; }
  *out = c;
  return 0;
}

// A context on the parent's device with its own stream and scratch that
// reads the parent's tables, query set and resident database (as the
// reference's device-memory areas of one CuDeviceMemory share its cached
// data, CuDeviceMemory.cuh:148-154).  The parent must outlive it, and
// uploads to the parent after this call are not seen by it.
extern "C" int cth_ctx_create_shared(cth_ctx* parent, cth_ctx** out) {
  *out = nullptr;
  if (!parent || parent->borrowed) { cth_set_error("shared context: invalid parent"); return -1; }
  CTH_HIP_CHECK(hipSetDevice(parent->device));
  cth_ctx* c = (cth_ctx*)calloc(1, sizeof(cth_ctx));
  if (!c) { cth_set_error("out of host memory"); return -1; }
  new (&c->nshared) std::atomic<int>(0);
  new (&c->tune_mu) std::mutex();
  c->device = parent->device;
  c->borrowed = 1;
  c->parent = parent;
  c->htab = parent->htab;
  c->dtab = parent->dtab;
  c->q = parent->q;
  c->db = parent->db;
  c->has_q = parent->has_q;
  c->has_db = parent->has_db;
  if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
    cth_set_error("hipStreamCreate failed"); free(c); return -1;
  }
  parent->nshared++;
  *out = c;
  return 0;
}

extern "C" void cth_ctx_trim(cth_ctx* c, size_t keep) {
  struct { void** p; size_t* cap; } bufs[] = {
      {(void**)&c->scores, &c->cap_scores}, {(void**)&c->mods, &c->cap_mods}, {(void**)&c->btck, &c->cap_btck},
      {(void**)&c->mapbuf, &c->cap_mapbuf}, {&c->work, &c->cap_work}, {(void**)&c->grow, &c->cap_grow}};
  for (auto& b : bufs)
    if (*b.p && *b.cap > keep) {
      (void)hipFree(*b.p);
      *b.p = nullptr;
      *b.cap = 0;
    }
}

extern "C" void cth_ctx_destroy(cth_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->st) (void)hipStreamSynchronize(c->st);
  if (c->borrowed && c->parent) c->parent->nshared--;
  if (!c->borrowed) {
    cth_devset_free(&c->q);
    cth_devset_free(&c->db);
    if (c->dtab) (void)hipFree(c->dtab);
  }
  void* bufs[] = {c->scores, c->mods, c->btck, c->dpout, c->work, c->mapbuf, c->grow};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  if (c->h_pin) (void)hipHostFree(c->h_pin);
  if (c->has_ev)
    for (int k = 0; k < 4; k++) (void)hipEventDestroy(c->ev[k]);
  if (c->nbev == 5) (void)hipEventDestroy(c->ev[4]);
  if (c->has_bgraph) (void)hipGraphExecDestroy(c->bgraph);
  if (c->bcells) (void)hipFree(c->bcells);
  if (c->bsmall) (void)hipFree(c->bsmall);
  if (c->st) (void)hipStreamDestroy(c->st);
  c->tune_mu.~mutex();
  free(c);
}

// the uploaded sets' sizes: number of queries and their lengths (host copy
// owned by the context), database profiles and positions
extern "C" int cth_ctx_shape(const cth_ctx* c, int* nqueries, const int32_t** qlens, int* ndb, int64_t* dbpos) {
  if (!c->has_q || !c->has_db) { cth_set_error("queries or database not uploaded"); return -1; }
  if (nqueries) *nqueries = c->q.d.nprof;
  if (qlens) *qlens = c->q.h_len;
  if (ndb) *ndb = c->db.d.nprof;
  if (dbpos) *dbpos = c->db.d.npos;
  return 0;
}

extern "C" int cth_upload_queries(cth_ctx* c, const cth_pset* qs) {
  if (c->borrowed) { cth_set_error("upload on a shared context"); return -1; }
  if (c->nshared > 0) { cth_set_error("upload while %d shared contexts read this one", c->nshared.load()); return -1; }
  CTH_HIP_CHECK(hipSetDevice(c->device));
  cth_devset_free(&c->q);
  c->has_q = 0;
  if (cth_devset_upload(qs, &c->q)) return -1;
  if (cth_devset_make_ddoff(qs, &c->q)) { cth_devset_free(&c->q); return -1; }
  c->has_q = 1;
  return 0;
}

extern "C" int cth_upload_db(cth_ctx* c, const cth_pset* db) {
  if (c->borrowed) { cth_set_error("upload on a shared context"); return -1; }
  if (c->nshared > 0) { cth_set_error("upload while %d shared contexts read this one", c->nshared.load()); return -1; }
  CTH_HIP_CHECK(hipSetDevice(c->device));
  cth_devset_free(&c->db);
  c->has_db = 0;
  if (cth_devset_upload(db, &c->db)) return -1;
  c->has_db = 1;
  return 0;
}

extern "C" cth_pset* cth_pset_from_packed(char* const* pmbeg, char* const* pmend, const size_t* szfields, int nprof,
                                          const char* const* descs);
extern "C" void cth_pset_free(cth_pset* ps);

// The reference's packed chunk buffers become the resident database
// (CuDeviceMemory::TransferCPMData, cupro/CuDeviceMemory.cu:267-274,
// 1095-1157, called through CuBatchProcessing::TransferCPMData,
// cupro/CuBatchProcessing.cuh:62-70).
extern "C" int cth_upload_db_packed(cth_ctx* c, char* const* pmbeg, char* const* pmend, const size_t* szfields) {
  cth_pset* db = cth_pset_from_packed(pmbeg, pmend, szfields, -1, nullptr);
  if (!db) return -1;
  const int rc = cth_upload_db(c, db);
  cth_pset_free(db);
  return rc;
}

// The query batch from the reference's packed buffers (the query part of
// CuDeviceMemory::CacheCompleteData, cualn/JobDispatcher.cpp:325-355); nprof
// as in cth_pset_from_packed.
extern "C" int cth_upload_queries_packed(cth_ctx* c, char* const* pmbeg, char* const* pmend, const size_t* szfields,
                                         int nprof) {
  cth_pset* q = cth_pset_from_packed(pmbeg, pmend, szfields, nprof, nullptr);
  if (!q) return -1;
  const int rc = cth_upload_queries(c, q);
  cth_pset_free(q);
  return rc;
}

extern "C" int cth_launch_score(const cth_dpset* q, const int32_t* qoff, int qp, const cth_dpset* db, int64_t col0, int ncols,
                                int64_t ldsc, const cth_tables* dtabs, int lq, float* scores, float* mods, int contraction,
                                int ddms, hipStream_t st);

// whether a score stage of query qidx against profiles [p0, p1) can have a
// nonzero DDMS term
static int ddms_of(const cth_ctx* c, int qidx, int p0, int p1) {
  return cth_devset_has_ddms(&c->q, qidx, qidx + 1) && cth_devset_has_ddms(&c->db, p0, p1);
}

// Score matrix of query qidx against template profiles [p0,p1), copied to
// host arrays [Lq][ncols] (ncols = positions of those profiles).
extern "C" int cth_score_block_ex(cth_ctx* c, int qidx, int p0, int p1, int contraction, float* h_scores,
                                  float* h_mods) {
  if (contraction < 0 || contraction > 2) { cth_set_error("score_block: contraction %d", contraction); return -1; }
  if (!c->has_q || !c->has_db) { cth_set_error("score_block: queries or database not uploaded"); return -1; }
  if (qidx < 0 || qidx >= c->q.d.nprof) { cth_set_error("score_block: query %d out of range", qidx); return -1; }
  if (p0 < 0 || p1 > c->db.d.nprof || p0 >= p1) { cth_set_error("score_block: bad profile range [%d,%d)", p0, p1); return -1; }
  CTH_HIP_CHECK(hipSetDevice(c->device));
  const int lq = c->q.h_len[qidx];
  const int64_t col0 = c->db.h_dist[p0];
  const int64_t ncols = c->db.h_dist[p1 - 1] + c->db.h_len[p1 - 1] - col0;
  if (ncols > (1ll << 30)) { cth_set_error("score_block: too many columns"); return -1; }
  const int64_t ldsc = (ncols + 63) / 64 * 64;
  const size_t need = sizeof(float) * (size_t)lq * ldsc;
  if (cth_ctx_reserve(c, (void**)&c->scores, &c->cap_scores, need)) return -1;
  if (cth_ctx_reserve(c, (void**)&c->mods, &c->cap_mods, need)) return -1;
  if (cth_launch_score(&c->q.d, c->q.ddoff, qidx, &c->db.d, col0, (int)ncols, ldsc, c->dtab, lq, c->scores, c->mods,
                       contraction, ddms_of(c, qidx, p0, p1), c->st))
    return -1;
  CTH_HIP_CHECK(hipStreamSynchronize(c->st));
  CTH_HIP_CHECK(hipMemcpy2D(h_scores, sizeof(float) * ncols, c->scores, sizeof(float) * ldsc, sizeof(float) * ncols,
                            lq, hipMemcpyDeviceToHost));
  CTH_HIP_CHECK(hipMemcpy2D(h_mods, sizeof(float) * ncols, c->mods, sizeof(float) * ldsc, sizeof(float) * ncols,
                            lq, hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int cth_score_block(cth_ctx* c, int qidx, int p0, int p1, float* h_scores, float* h_mods) {
  return cth_score_block_ex(c, qidx, p0, p1, 0, h_scores, h_mods);
}

extern "C" int cth_launch_dp(const cth_dpset* q, int qp, const cth_dpset* db, int p0, int np, int ltmax,
                             int64_t col0, int64_t ldsc, const float* scores, uint8_t* btck, float scorethld,
                             cth_dpres* res, float* grow, hipStream_t st);
extern "C" size_t cth_dp_row_bytes(int np, int ltmax);
extern "C" size_t cth_map_row_bytes(int njobs, int ltmax);

// Phase 1 (score matrix, DP, finalisation) of query qidx against template
// profiles [p0,p1); per-profile results to res (p1-p0 records of 6 words:
// maxsc, maxxy, passed, score, psts, beg) and, when btck is not NULL, the
// traceback bytes [Lq][ncols] to host.
extern "C" int cth_phase1_block(cth_ctx* c, int qidx, int p0, int p1, float scorethld, void* res, uint8_t* h_btck) {
  if (!c->has_q || !c->has_db) { cth_set_error("phase1_block: queries or database not uploaded"); return -1; }
  if (qidx < 0 || qidx >= c->q.d.nprof) { cth_set_error("phase1_block: query %d out of range", qidx); return -1; }
  if (p0 < 0 || p1 > c->db.d.nprof || p0 >= p1) { cth_set_error("phase1_block: bad profile range [%d,%d)", p0, p1); return -1; }
  CTH_HIP_CHECK(hipSetDevice(c->device));
  const int lq = c->q.h_len[qidx];
  const int64_t col0 = c->db.h_dist[p0];
  const int64_t ncols = c->db.h_dist[p1 - 1] + c->db.h_len[p1 - 1] - col0;
  if (ncols > (1ll << 30)) { cth_set_error("phase1_block: too many columns"); return -1; }
  const int64_t ldsc = (ncols + 63) / 64 * 64;
  const size_t need = sizeof(float) * (size_t)lq * ldsc;
  int ltmax = 0;
  for (int p = p0; p < p1; p++) ltmax = c->db.h_len[p] > ltmax ? c->db.h_len[p] : ltmax;
  if (cth_ctx_reserve(c, (void**)&c->scores, &c->cap_scores, need)) return -1;
  if (cth_ctx_reserve(c, (void**)&c->mods, &c->cap_mods, need)) return -1;
  if (cth_ctx_reserve(c, (void**)&c->btck, &c->cap_btck, (size_t)lq * ldsc)) return -1;
  if (cth_ctx_reserve(c, (void**)&c->dpout, &c->cap_dpout, sizeof(cth_dpres) * (size_t)(p1 - p0))) return -1;
  if (cth_launch_score(&c->q.d, c->q.ddoff, qidx, &c->db.d, col0, (int)ncols, ldsc, c->dtab, lq, c->scores, c->mods, 0,
                       ddms_of(c, qidx, p0, p1), c->st))
    return -1;
  // boundary rows in global memory for profiles over the LDS row limit
  const size_t rows = cth_dp_row_bytes(p1 - p0, ltmax);
  if (rows && cth_ctx_reserve(c, (void**)&c->grow, &c->cap_grow, rows)) return -1;
  if (cth_launch_dp(&c->q.d, qidx, &c->db.d, p0, p1 - p0, ltmax, col0, ldsc, c->scores, c->btck, scorethld,
                    (cth_dpres*)c->dpout, rows ? c->grow : nullptr, c->st))
    return -1;
  CTH_HIP_CHECK(hipStreamSynchronize(c->st));
  CTH_HIP_CHECK(hipMemcpy(res, c->dpout, sizeof(cth_dpres) * (size_t)(p1 - p0), hipMemcpyDeviceToHost));
  if (h_btck)
    CTH_HIP_CHECK(hipMemcpy2D(h_btck, ncols, c->btck, ldsc, ncols, lq, hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int cth_launch_stats(const cth_dpset* q, int qp, const cth_dpset* db, int p0, int64_t col0, int64_t ldsc,
                                const float* scores, const int* plist_abs, const int* plist_rel,
                                const cth_dpres* dpres, int np2, double searchspace, int ssemodel, int* hist,
                                double* out, hipStream_t st);

// Phase 1 and the statistics of the passed profiles for query qidx against
// template profiles [p0,p1).  res: p1-p0 phase-1 records (as cth_phase1_block);
// plist: indices (relative to p0) of the passed profiles, *npass their number;
// stats: npass x 10 doubles {lambda_est, K_est, logE, logEpA, min, max, E, H,
// lambda, K}.  plist and stats must hold p1-p0 entries.
extern "C" int cth_phase2_block(cth_ctx* c, int qidx, int p0, int p1, float scorethld, double searchspace,
                                int ssemodel, void* res, int* plist, int* npass, double* stats) {
  // SSEMODEL takes 0..2 (libpro/srcpro/MOptions.cpp:123)
  if (ssemodel < 0 || ssemodel > 2) { cth_set_error("phase2_block: ssemodel %d (0, 1 or 2)", ssemodel); return -1; }
  if (cth_phase1_block(c, qidx, p0, p1, scorethld, res, nullptr)) return -1;
  const int n = p1 - p0;
  const cth_dpres* r = (const cth_dpres*)res;
  std::vector<int> rel, abs_;
  for (int k = 0; k < n; k++)
    if (r[k].passed) { rel.push_back(k); abs_.push_back(p0 + k); }
  *npass = (int)rel.size();
  for (size_t i = 0; i < rel.size(); i++) plist[i] = rel[i];
  if (rel.empty()) return 0;
  const int np2 = (int)rel.size();
  const int lq = c->q.h_len[qidx];
  const int64_t col0 = c->db.h_dist[p0];
  const int64_t ncols = c->db.h_dist[p1 - 1] + c->db.h_len[p1 - 1] - col0;
  const int64_t ldsc = (ncols + 63) / 64 * 64;
  (void)lq;
  const size_t need = sizeof(int) * 2 * np2 + sizeof(int) * 641 * (size_t)np2 + sizeof(double) * 10 * (size_t)np2 + 256;
  if (cth_ctx_reserve(c, &c->work, &c->cap_work, need)) return -1;
  int* d_abs = (int*)c->work;
  int* d_rel = d_abs + np2;
  int* d_hist = d_rel + np2;
  double* d_out = (double*)(((uintptr_t)(d_hist + 641 * (size_t)np2) + 15) & ~(uintptr_t)15);
  CTH_HIP_CHECK(hipMemcpyAsync(d_abs, abs_.data(), sizeof(int) * np2, hipMemcpyHostToDevice, c->st));
  CTH_HIP_CHECK(hipMemcpyAsync(d_rel, rel.data(), sizeof(int) * np2, hipMemcpyHostToDevice, c->st));
  if (cth_launch_stats(&c->q.d, qidx, &c->db.d, p0, col0, ldsc, c->scores, d_abs, d_rel, (const cth_dpres*)c->dpout,
                       np2, searchspace, ssemodel, d_hist, d_out, c->st))
    return -1;
  CTH_HIP_CHECK(hipStreamSynchronize(c->st));
  CTH_HIP_CHECK(hipMemcpy(stats, d_out, sizeof(double) * 10 * np2, hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int cth_launch_map(const cth_dpset* q, int qp, int lq, const cth_dpset* db, const cth_mapjob* jobs,
                              int njobs, int ltmax, int64_t col0, int64_t ldsc, const float* scores,
                              const float* mods, float* post, uint8_t* btck, double* zrow, float minpp,
                              char* alnout, int rowcap, cth_mapres* res, float* grow, hipStream_t st);

// Score matrix of query qidx against template profiles [p0,p1), then the MAP
// stage for njobs of them (absolute indices profiles[], log pair E-values
// logepa[]); res receives njobs records of cth_mapres.
extern "C" int cth_map_block(cth_ctx* c, int qidx, int p0, int p1, const int32_t* profiles, const float* logepa,
                             int njobs, float minpp, void* res) {
  if (!c->has_q || !c->has_db) { cth_set_error("map_block: queries or database not uploaded"); return -1; }
  if (qidx < 0 || qidx >= c->q.d.nprof) { cth_set_error("map_block: query %d out of range", qidx); return -1; }
  if (p0 < 0 || p1 > c->db.d.nprof || p0 >= p1) { cth_set_error("map_block: bad profile range [%d,%d)", p0, p1); return -1; }
  if (njobs <= 0) return 0;
  std::vector<cth_mapjob> jobs(njobs);
  std::vector<char> seen(p1 - p0, 0);
  int ltmax = 0;
  for (int j = 0; j < njobs; j++) {
    if (profiles[j] < p0 || profiles[j] >= p1) { cth_set_error("map_block: profile %d outside the block", profiles[j]); return -1; }
    // each job writes its profile's columns of the posterior buffer: no duplicates
    if (seen[profiles[j] - p0]++) { cth_set_error("map_block: profile %d listed twice", profiles[j]); return -1; }
    jobs[j].profile = profiles[j];
    jobs[j].logepa = logepa[j];
    ltmax = c->db.h_len[profiles[j]] > ltmax ? c->db.h_len[profiles[j]] : ltmax;
  }
  CTH_HIP_CHECK(hipSetDevice(c->device));
  const int lq = c->q.h_len[qidx];
  const int64_t col0 = c->db.h_dist[p0];
  const int64_t ncols = c->db.h_dist[p1 - 1] + c->db.h_len[p1 - 1] - col0;
  if (ncols > (1ll << 30)) { cth_set_error("map_block: too many columns"); return -1; }
  const int64_t ldsc = (ncols + 63) / 64 * 64;
  const size_t need = sizeof(float) * (size_t)lq * ldsc;
  if (cth_ctx_reserve(c, (void**)&c->scores, &c->cap_scores, need)) return -1;
  if (cth_ctx_reserve(c, (void**)&c->mods, &c->cap_mods, need)) return -1;
  if (cth_ctx_reserve(c, (void**)&c->mapbuf, &c->cap_mapbuf, need)) return -1;
  if (cth_ctx_reserve(c, (void**)&c->btck, &c->cap_btck, (size_t)lq * ldsc)) return -1;
  const size_t wneed = sizeof(double) * (size_t)njobs * lq + sizeof(cth_mapjob) * njobs +
                       sizeof(cth_mapres) * njobs + 512;
  if (cth_ctx_reserve(c, &c->work, &c->cap_work, wneed)) return -1;
  double* d_zrow = (double*)c->work;
  cth_mapjob* d_jobs = (cth_mapjob*)(d_zrow + (size_t)njobs * lq);
  cth_mapres* d_res = (cth_mapres*)(((uintptr_t)(d_jobs + njobs) + 15) & ~(uintptr_t)15);
  CTH_HIP_CHECK(hipMemcpyAsync(d_jobs, jobs.data(), sizeof(cth_mapjob) * njobs, hipMemcpyHostToDevice, c->st));
  if (cth_launch_score(&c->q.d, c->q.ddoff, qidx, &c->db.d, col0, (int)ncols, ldsc, c->dtab, lq, c->scores, c->mods, 0,
                       ddms_of(c, qidx, p0, p1), c->st))
    return -1;
  const size_t rows = cth_map_row_bytes(njobs, ltmax);
  if (rows && cth_ctx_reserve(c, (void**)&c->grow, &c->cap_grow, rows)) return -1;
  if (cth_launch_map(&c->q.d, qidx, lq, &c->db.d, d_jobs, njobs, ltmax, col0, ldsc, c->scores, c->mods,
                     (float*)c->mapbuf, c->btck, d_zrow, minpp, nullptr, lq + ltmax, d_res, rows ? c->grow : nullptr,
                     c->st))
    return -1;
  CTH_HIP_CHECK(hipMemcpyAsync(res, d_res, sizeof(cth_mapres) * njobs, hipMemcpyDeviceToHost, c->st));
  CTH_HIP_CHECK(hipStreamSynchronize(c->st));
  return 0;
}

namespace {
__global__ void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
}  // namespace

// Measurement aid (no reference counterpart): copies `bytes` (a multiple of
// 16) between two device buffers `reps` times with a plain streaming kernel,
// so that HBM traffic counters can be scaled against a kernel of known
// traffic (bytes read and written per launch); *ms receives the mean time.
extern "C" int cth_bench_copy(int device, size_t bytes, int reps, float* ms) {
  if (bytes == 0 || bytes % 16 || reps < 1) { cth_set_error("bench_copy: bad size"); return -1; }
  CTH_HIP_CHECK(hipSetDevice(device));
  void *a = nullptr, *b = nullptr;
  CTH_HIP_CHECK(hipMalloc(&a, bytes));
  if (hipMalloc(&b, bytes) != hipSuccess) { (void)hipFree(a); cth_set_error("bench_copy: out of memory"); return -1; }
  hipEvent_t e0, e1;
  CTH_HIP_CHECK(hipEventCreate(&e0));
  CTH_HIP_CHECK(hipEventCreate(&e1));
  const size_t n = bytes / 16;
  hipLaunchKernelGGL(k_copy16, dim3(4096), dim3(256), 0, 0, (const uint4*)a, (uint4*)b, n);
  CTH_HIP_CHECK(hipEventRecord(e0, 0));
  for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy16, dim3(4096), dim3(256), 0, 0, (const uint4*)a, (uint4*)b, n);
  CTH_HIP_CHECK(hipEventRecord(e1, 0));
  CTH_HIP_CHECK(hipEventSynchronize(e1));
  float t = 0.0f;
  CTH_HIP_CHECK(hipEventElapsedTime(&t, e0, e1));
  if (ms) *ms = t / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  CTH_HIP_CHECK(hipFree(a));
  CTH_HIP_CHECK(hipFree(b));
  return 0;
}
