// search.hip: the per-query search pipeline over the HBM-resident database.
//
// This is the device half of CuBatchProcessing::BatchProcessScoreMatrixDevice
// (cupro/CuBatchProcessing.cu:230-575) for one query: for each chunk of the
// resident database whose score buffers fit the memory budget, the score
// matrix, phase-1 DP and its finalisation run on the context's stream; the
// passed profiles of the chunk then get their statistics (and, with MAP
// enabled, the MAP alignment stage).  Host work is limited to the score
// threshold (JobDispatcher::CalculateScoreThreshold,
// cualn/JobDispatcher.cpp:806-860), the search-space length adjustment
// (AbstractScoreMatrix::ComputeLengthAdjustment,
// libpro/srcsco/AbstractScoreMatrix.cpp:303-395), the passed-profile lists and
// the final hit ordering.
//
// Deterministic by construction: passed profiles are taken in database order
// (the reference appends them in atomic order, cudp/CuBatchDP_final.cu:260-264)
// and hits are sorted by (log E-value, database index) (the reference's sort is
// unstable, cupro/CuBatchProcessingFinalizer.cpp:289-292).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "cth_ctx.h"
#include "cth_dp.h"
#include "cth_map.h"
#include "cth_search.h"
#include "cth_stats.h"

extern "C" int cth_launch_score(const cth_dpset* q, const int32_t* qoff, int qp, const cth_dpset* db, int64_t col0, int ncols,
                                int64_t ldsc, const cth_tables* dtabs, int lq, float* scores, float* mods, int contraction,
                                int ddms, hipStream_t st);
extern "C" int cth_launch_dp_list(const cth_dpset* q, int qp, const cth_dpset* db, int p0, const int* plist, int np,
                                  int ltmax, int64_t col0, int64_t ldsc, const float* scores, uint8_t* btck,
                                  float scorethld, int finalize, cth_dpres* res, float* grow, hipStream_t st);
extern "C" size_t cth_dp_row_bytes(int np, int ltmax);
extern "C" size_t cth_map_row_bytes(int njobs, int ltmax);
extern "C" int cth_launch_stats(const cth_dpset* q, int qp, const cth_dpset* db, int p0, int64_t col0, int64_t ldsc,
                                const float* scores, const int* plist_abs, const int* plist_rel,
                                const cth_dpres* dpres, int np2, double searchspace, int ssemodel, int* hist,
                                double* out, hipStream_t st);
extern "C" int cth_launch_map(const cth_dpset* q, int qp, int lq, const cth_dpset* db, const cth_mapjob* jobs,
                              int njobs, int ltmax, int64_t col0, int64_t ldsc, const float* scores,
                              const float* mods, float* post, uint8_t* btck, double* zrow, float minpp,
                              char* alnout, int rowcap, cth_mapres* res, float* grow, hipStream_t st);

namespace {

// JobDispatcher::CalculateScoreThreshold (cualn/JobDispatcher.cpp:806-860)
void score_threshold(float eval, uint64_t nqyposs, uint64_t dbsize, float* scorethld, float* logevthld) {
  const uint64_t sspace = nqyposs * dbsize;
  float ev = eval;
  if (ev <= 0.0f) { *scorethld = 999.0f; *logevthld = -99.0f; return; }
  *logevthld = logf(ev);
  ev *= ev < 1.0f ? 10.0f : 3.0f;
  ev /= CTH_GAPPED_K * (float)sspace;
  if (ev <= 0.0f) { *scorethld = 999.0f; return; }
  const float s = -logf(ev) / CTH_GAPPED_LAMBDA;
  *scorethld = s < 0.0f ? 0.0f : s;
}

// AbstractScoreMatrix::ComputeLengthAdjustment (libpro/srcsco/AbstractScoreMatrix.cpp:303-395),
// Gertz's iteration for the edge-effect length l with l = a/lambda ln(K(m-l)(n-Nl)) + b
uint64_t length_adjustment(uint64_t m, uint64_t n, uint64_t N, uint64_t* deltalen = nullptr) {
  const float lambda = CTH_GAPPED_LAMBDA, K = CTH_GAPPED_K, alpha = CTH_GAPPED_ALPHA, beta = CTH_GAPPED_BETA;
  if (deltalen) *deltalen = 0;
  if (N < 1) return m * n;
  const float logK = logf(K), alratio = alpha / lambda, max_mn = (float)(m > n ? m : n);
  float length = 0.0f, min_length = 0.0f, max_length = 0.0f, nxt = 0.0f, sspace = 0.0f;
  bool converged = false;
  const float a = (float)N, b = (float)(n + m * N), c = (float)(m * n) - max_mn / K;
  if (c < 0.0f) return m * n;
  max_length = 2.0f * c / (b + sqrtf(b * b - 4.0f * a * c));
  for (int j = 0; j < 20; j++) {
    length = nxt;
    sspace = (m - length) * (n - N * length);
    nxt = alratio * (logK + logf(sspace)) + beta;
    if (length <= nxt) {
      min_length = length;
      if (nxt - length <= 1.0f) { converged = true; break; }
      if (min_length == max_length) break;
    } else
      max_length = length;
    if (nxt < min_length || max_length < nxt) nxt = (!j) ? max_length : (min_length + max_length) * 0.5f;
  }
  if (converged) {
    length = ceilf(min_length);
    if (length <= max_length) {
      const float space2 = (m - length) * (n - N * length);
      nxt = alratio * (logK + logf(space2)) + beta;
      if (length <= nxt) {
        min_length = length;
        sspace = space2;
      }
    }
  }
  if (deltalen) *deltalen = (uint64_t)min_length;
  return (uint64_t)sspace;
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Working memory of cth_search_noalloc: one fixed array that every device
// running the library holds, at its own address, until the process ends.
// In host code `scratch` names a host-side stand-in, not device memory: the
// device address comes from hipGetSymbolAddress (scratch_init) and only that
// address is passed to kernels.  No buffer carved from it needs to start at
// zero: every kernel writes the cells, records and lists it reads later in
// the same call.
static __device__ __align__(256) uint8_t scratch[CTH_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch_addr[kMaxDevices];  // device address of scratch per device (null until set up)

uint8_t* scratch_of(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return (device >= 0 && device < kMaxDevices) ? g_scratch_addr[device] : nullptr;
}

// Device buffers of one database chunk of the search.  cth_search grows the
// context's own buffers (allocating when a chunk needs more than they hold);
// cth_search_noalloc lays them out in the fixed scratch region: scores,
// modular scores, traceback bytes, phase-1 records and MAP posteriors at the
// front, then the work area that each later stage reuses for its lists and
// records (a stage's work contents are dead once the next stage takes it).
struct chunk_bufs {
  cth_ctx* c;
  uint8_t* fixed;       // the fixed region, or null to grow the context's buffers
  size_t fixed_cap;
  const char* who;      // entry-point name for error messages
  size_t work_off = 0;  // fixed mode: start of the work area
  float* scores = nullptr;
  float* mods = nullptr;
  uint8_t* btck = nullptr;
  cth_dpres* dpout = nullptr;
  float* mapbuf = nullptr;

  chunk_bufs(cth_ctx* ctx, uint8_t* base, size_t cap, const char* name)
      : c(ctx), fixed(base), fixed_cap(cap), who(name) {}

  static size_t up(size_t n) { return (n + 255) & ~(size_t)255; }
  // fixed mode: the share of the region the chunk buffers may take (the rest is the work area)
  size_t chunk_share() const { return fixed_cap / 4 * 3; }
  // fixed mode: the most columns a chunk may span so that its layout fits the
  // share: with ldsc <= cols + 63 and at most one profile per column, the
  // layout needs at most lq*(cols+63)*bpc + 24*cols + 5*256 bytes
  int64_t max_chunk_cols(int lq, size_t bpc) const {
    const int64_t fixed_part = 5 * 256 + 63 * (int64_t)lq * (int64_t)bpc;
    const int64_t room = (int64_t)chunk_share() - fixed_part;
    return room > 0 ? room / ((int64_t)lq * (int64_t)bpc + (int64_t)sizeof(cth_dpres)) : 0;
  }

  int chunk(int qi, int lq, int ltchunk, int64_t ldsc, int np, bool map) {
    const size_t cells = (size_t)lq * (size_t)ldsc;
    if (!fixed) {
      if (cth_ctx_reserve(c, (void**)&c->scores, &c->cap_scores, 4 * cells)) return -1;
      if (cth_ctx_reserve(c, (void**)&c->mods, &c->cap_mods, 4 * cells)) return -1;
      if (cth_ctx_reserve(c, (void**)&c->btck, &c->cap_btck, cells)) return -1;
      if (cth_ctx_reserve(c, (void**)&c->dpout, &c->cap_dpout, sizeof(cth_dpres) * (size_t)np)) return -1;
      scores = c->scores; mods = c->mods; btck = c->btck; dpout = (cth_dpres*)c->dpout;
      mapbuf = nullptr;
      return 0;
    }
    const size_t need = up(4 * cells) * (map ? 3 : 2) + up(cells) + up(sizeof(cth_dpres) * (size_t)np);
    if (need > chunk_share()) {
      cth_set_error("%s: query %d (%d positions) against a database chunk of %lld columns (longest profile %d) "
                    "needs %zu bytes of buffers; the scratch allows %zu",
                    who, qi, lq, (long long)ldsc, ltchunk, need, chunk_share());
      return -1;
    }
    size_t o = 0;
    auto take = [&](size_t n) { uint8_t* p = fixed + o; o += up(n); return p; };
    scores = (float*)take(4 * cells);
    mods = (float*)take(4 * cells);
    btck = take(cells);
    dpout = (cth_dpres*)take(sizeof(cth_dpres) * (size_t)np);
    mapbuf = map ? (float*)take(4 * cells) : nullptr;
    work_off = o;
    return 0;
  }

  // the MAP posterior buffer of the current chunk
  float* map_buffer(int lq, int64_t ldsc) {
    if (fixed) {
      if (!mapbuf) cth_set_error("%s: no MAP buffer in this chunk's layout", who);
      return mapbuf;
    }
    if (cth_ctx_reserve(c, (void**)&c->mapbuf, &c->cap_mapbuf, sizeof(float) * (size_t)lq * (size_t)ldsc))
      return nullptr;
    return mapbuf = c->mapbuf;
  }

  // How many profiles of one launch get global boundary rows of per_item
  // bytes each when the stage's lists take `used` bytes of the work area (at
  // most `want`; 0 with the error set when not even one fits).  Growing
  // buffers hold up to kRowBudget of rows per launch.
  int row_batch(size_t per_item, size_t used, int want) {
    constexpr size_t kRowBudget = (size_t)1 << 30;
    size_t room = kRowBudget;
    if (fixed) room = fixed_cap - work_off > used + 256 ? fixed_cap - work_off - used - 256 : 0;
    const size_t n = per_item ? room / per_item : (size_t)want;
    if (n < 1) {
      cth_set_error("%s: boundary rows of %zu bytes per profile do not fit the scratch", who, per_item);
      return 0;
    }
    return (int)std::min<size_t>(n, (size_t)want);
  }

  // the work area of a stage (null with the error set when it does not fit)
  void* work(size_t need) {
    if (!fixed) {
      if (cth_ctx_reserve(c, &c->work, &c->cap_work, need)) return nullptr;
      return c->work;
    }
    if (need > fixed_cap - work_off) {
      cth_set_error("%s: a stage needs %zu bytes of lists and records; %zu are left after the chunk buffers", who,
                    need, fixed_cap - work_off);
      return nullptr;
    }
    return fixed + work_off;
  }
};

}  // namespace

// Set up this device's scratch address (once per device, with the device
// current); contexts call it at creation.
extern "C" int cth_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) { cth_set_error("scratch: device %d out of range", device); return -1; }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_addr[device]) return 0;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    cth_set_error("scratch: hipGetSymbolAddress on device %d: %s", device, hipGetErrorString(e));
    return -1;
  }
  g_scratch_addr[device] = (uint8_t*)p;
  return 0;
}

extern "C" void cth_search_space(uint64_t qlen, uint64_t dbsize, uint64_t ndb, uint64_t* sspace, uint64_t* deltalen) {
  uint64_t dl = 0;
  const uint64_t ss = length_adjustment(qlen, dbsize, ndb, &dl);
  if (sspace) *sspace = ss;
  if (deltalen) *deltalen = dl;
}

// A query's score threshold, log E-value threshold and length-adjusted
// search space against dbsize positions in ndb profiles.
extern "C" void cth_query_thresholds(float eval, uint64_t lq, uint64_t dbsize, uint64_t ndb, float* scorethld,
                                     float* logethld, double* sspace) {
  score_threshold(eval, lq, dbsize, scorethld, logethld);
  *sspace = (double)length_adjustment(lq, dbsize, ndb);
}

// Option checks shared by every entry point that takes cth_search_opts.
extern "C" int cth_search_opts_check(const cth_search_opts* opt, const char* who) {
  if (!opt) { cth_set_error("%s: no options", who); return -1; }
  if (opt->contraction < 0 || opt->contraction > 2) { cth_set_error("%s: contraction %d", who, opt->contraction); return -1; }
  // SSEMODEL takes 0..2 (libpro/srcpro/MOptions.cpp:123)
  if (opt->ssemodel < 0 || opt->ssemodel > 2) {
    cth_set_error("%s: ssemodel %d (0, 1 or 2)", who, opt->ssemodel);
    return -1;
  }
  if ((opt->dbsize == 0) != (opt->ndb == 0)) {
    cth_set_error("%s: dbsize and ndb are given together (both 0: the resident set's)", who);
    return -1;
  }
  if (opt->ndb > opt->dbsize) { cth_set_error("%s: ndb %llu > dbsize %llu", who, (unsigned long long)opt->ndb,
                                              (unsigned long long)opt->dbsize); return -1; }
  return 0;
}

namespace {

// The whole per-query pipeline; its per-chunk device buffers come from bufs.
int search_run(cth_ctx* c, int q0, int q1, const cth_search_opts* opt, cth_hit* hits, int maxhits, int* nhits,
               cth_search_stats* stats, chunk_bufs& bufs) {
  const char* who = bufs.who;
  if (!c->has_q || !c->has_db) { cth_set_error("%s: queries or database not uploaded", who); return -1; }
  if (q0 < 0 || q1 > c->q.d.nprof || q0 >= q1) { cth_set_error("%s: bad query range [%d,%d)", who, q0, q1); return -1; }
  if (cth_search_opts_check(opt, who)) return -1;
  CTH_HIP_CHECK(hipSetDevice(c->device));
  const cth_devset& db = c->db;
  const int nprof = db.d.nprof;
  // the database size and count the statistics use: the whole database's
  // when the resident set is one part of it (CuBatchProcessing::SetDbDetails,
  // cupro/CuBatchProcessing.cuh:73, set per worker at cualn/DevCommThread.cpp:165)
  const uint64_t dbsize = opt->dbsize ? opt->dbsize : (uint64_t)db.d.npos;
  const uint64_t ndbent = opt->ndb ? opt->ndb : (uint64_t)nprof;
  // 64 GiB by default: a quarter of the device's HBM, so that long queries
  // run in few, wide chunks (more template profiles per DP launch)
  const size_t budget = opt->mem_budget ? opt->mem_budget : ((size_t)64 << 30);
  const bool timing = opt->timing != 0;
  if (timing && !c->has_ev) {
    for (int k = 0; k < 4; k++) CTH_HIP_CHECK(hipEventCreate(&c->ev[k]));
    c->has_ev = 1;
  }
  int nh = 0;
  uint64_t alnused = 0;
  cth_search_stats st{};
  for (int qi = q0; qi < q1; qi++) {
    const int lq = c->q.h_len[qi];
    float scorethld, logethld;
    score_threshold(opt->eval, (uint64_t)lq, dbsize, &scorethld, &logethld);
    const double sspace = (double)length_adjustment((uint64_t)lq, dbsize, ndbent);
    // chunk columns so that scores + modscores + traceback (+ MAP posteriors) fit the budget
    // (and at most 6M columns: the phase-1 DP reads 64 score rows per step,
    // which is slower when the rows are much further apart)
    const size_t bytes_per_cell = opt->do_map ? 13 : 9;
    constexpr int64_t kMaxChunkCols = 6 << 20;
    const int64_t chunkcols =
        bufs.fixed ? std::min<int64_t>(kMaxChunkCols, bufs.max_chunk_cols(lq, bytes_per_cell))
                   : std::max<int64_t>(4096, std::min<int64_t>(kMaxChunkCols,
                                                               (int64_t)(budget / ((size_t)lq * bytes_per_cell))));
    const int qh0 = nh;
    for (int p0 = 0; p0 < nprof;) {
      int p1 = p0;
      const int64_t col0 = db.h_dist[p0];
      while (p1 < nprof && (db.h_dist[p1] + db.h_len[p1] - col0 <= chunkcols || p1 == p0)) p1++;
      const int64_t ncols = db.h_dist[p1 - 1] + db.h_len[p1 - 1] - col0;
      const int64_t ldsc = (ncols + 63) / 64 * 64;
      int ltchunk = 0;
      for (int p = p0; p < p1; p++) ltchunk = std::max(ltchunk, db.h_len[p]);
      if (bufs.chunk(qi, lq, ltchunk, ldsc, p1 - p0, opt->do_map != 0)) return -1;
      float* const scores = bufs.scores;
      float* const mods = bufs.mods;
      uint8_t* const btck = bufs.btck;
      cth_dpres* const dpout = bufs.dpout;
      const double t0 = now_ms();
      if (timing) CTH_HIP_CHECK(hipEventRecord(c->ev[0], c->st));
      // no DDMS term anywhere in the chunk when the query's or every template's
      // distance lists are shorter than DSTSEGM
      const int ddms = cth_devset_has_ddms(&c->q, qi, qi + 1) && cth_devset_has_ddms(&db, p0, p1);
      if (cth_launch_score(&c->q.d, c->q.ddoff, qi, &db.d, col0, (int)ncols, ldsc, c->dtab, lq, scores, mods,
                           opt->contraction, ddms, c->st))
        return -1;
      if (timing) CTH_HIP_CHECK(hipEventRecord(c->ev[1], c->st));
      // phase 1 in two passes: the DP of every profile without traceback,
      // then, for the profiles whose maximum exceeds the score threshold (the
      // only ones FinalizeDPPro traces back, cudp/CuBatchDP_final.cu:144),
      // the same DP with traceback and the finalisation.  The first pass runs
      // in groups of consecutive profiles, each launch reserving LDS for its
      // own longest profile, so that groups of shorter profiles (databases
      // are sorted by length) run more wavefronts per CU.
      // Profiles longer than CTH_DP_GLOBAL_FROM keep their stripe boundary rows
      // in the work area, in launches of as many profiles as their rows fit.
      constexpr int kDpGroup = 2048;
      for (int a = p0; a < p1; a += kDpGroup) {
        const int b = std::min(p1, a + kDpGroup);
        int lg = 0;
        for (int p = a; p < b; p++) lg = std::max(lg, db.h_len[p]);
        int nb = b - a;
        float* grow = nullptr;
        if (lg > CTH_DP_GLOBAL_FROM) {
          nb = bufs.row_batch(cth_dp_row_bytes(1, lg), 0, b - a);
          if (nb < 1) return -1;
          grow = (float*)bufs.work(cth_dp_row_bytes(nb, lg));
          if (!grow) return -1;
        }
        for (int a2 = a; a2 < b; a2 += nb) {
          const int b2 = std::min(b, a2 + nb);
          if (cth_launch_dp_list(&c->q.d, qi, &db.d, a2, nullptr, b2 - a2, lg, col0, ldsc, scores, nullptr, scorethld,
                                 0, dpout + (a2 - p0), grow, c->st))
            return -1;
        }
      }
      if (timing) CTH_HIP_CHECK(hipEventRecord(c->ev[2], c->st));
      std::vector<cth_dpres> res(p1 - p0);
      CTH_HIP_CHECK(hipMemcpyAsync(res.data(), dpout, sizeof(cth_dpres) * res.size(), hipMemcpyDeviceToHost, c->st));
      CTH_HIP_CHECK(hipStreamSynchronize(c->st));
      if (timing) {
        float a = 0.0f, b = 0.0f;
        CTH_HIP_CHECK(hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
        CTH_HIP_CHECK(hipEventElapsedTime(&b, c->ev[1], c->ev[2]));
        st.ms_k_score += a;
        st.ms_k_dp += b;
      }
      // candidates for the traceback pass: profiles within the LDS row limit
      // first, then the longer ones (global boundary rows, in batches)
      std::vector<int> cand, cand_abs, cand_long;
      int ltc = 0, ltl = 0;
      for (int k = 0; k < p1 - p0; k++) {
        res[k].passed = 0; res[k].score = 0.0f; res[k].psts = 0; res[k].beg = 0;
        if (scorethld < res[k].maxsc) {
          const int lt = db.h_len[p0 + k];
          if (lt > CTH_DP_GLOBAL_FROM) {
            cand_long.push_back(k);
            ltl = std::max(ltl, lt);
          } else {
            cand.push_back(k);
            cand_abs.push_back(p0 + k);
            ltc = std::max(ltc, lt);
          }
        }
      }
      const int ns = (int)cand.size();
      for (int k : cand_long) {
        cand.push_back(k);
        cand_abs.push_back(p0 + k);
      }
      if (!cand.empty()) {
        const int nc = (int)cand.size(), nl = nc - ns;
        const size_t lists = ((sizeof(int) * nc + 15) & ~(size_t)15) + sizeof(cth_dpres) * nc + 256;
        int nbl = nl;
        if (nl) {
          nbl = bufs.row_batch(cth_dp_row_bytes(1, ltl), lists, nl);
          if (nbl < 1) return -1;
        }
        int* d_cand = (int*)bufs.work(lists + (nl ? cth_dp_row_bytes(nbl, ltl) : 0));
        if (!d_cand) return -1;
        cth_dpres* d_res2 = (cth_dpres*)(((uintptr_t)(d_cand + nc) + 15) & ~(uintptr_t)15);
        float* grow = nl ? (float*)(((uintptr_t)(d_res2 + nc) + 255) & ~(uintptr_t)255) : nullptr;
        CTH_HIP_CHECK(hipMemcpyAsync(d_cand, cand_abs.data(), sizeof(int) * nc, hipMemcpyHostToDevice, c->st));
        if (timing) CTH_HIP_CHECK(hipEventRecord(c->ev[1], c->st));
        if (cth_launch_dp_list(&c->q.d, qi, &db.d, p0, d_cand, ns, ltc, col0, ldsc, scores, btck, scorethld, 1,
                               d_res2, nullptr, c->st))
          return -1;
        for (int j = 0; j < nl; j += nbl)
          if (cth_launch_dp_list(&c->q.d, qi, &db.d, p0, d_cand + ns + j, std::min(nbl, nl - j), ltl, col0, ldsc,
                                 scores, btck, scorethld, 1, d_res2 + ns + j, grow, c->st))
            return -1;
        if (timing) CTH_HIP_CHECK(hipEventRecord(c->ev[2], c->st));
        std::vector<cth_dpres> res2(nc);
        CTH_HIP_CHECK(hipMemcpyAsync(res2.data(), d_res2, sizeof(cth_dpres) * nc, hipMemcpyDeviceToHost, c->st));
        CTH_HIP_CHECK(hipStreamSynchronize(c->st));
        for (int k = 0; k < nc; k++) res[cand[k]] = res2[k];
        // the statistics kernel reads the phase-1 records of the chunk from dpout
        CTH_HIP_CHECK(hipMemcpyAsync(dpout, res.data(), sizeof(cth_dpres) * res.size(), hipMemcpyHostToDevice,
                                     c->st));
        if (timing) {
          float b = 0.0f;
          CTH_HIP_CHECK(hipEventElapsedTime(&b, c->ev[1], c->ev[2]));
          st.ms_k_dp += b;
        }
      }
      const double t1 = now_ms();
      std::vector<int> rel, abs_;
      for (int k = 0; k < p1 - p0; k++)
        if (res[k].passed) { rel.push_back(k); abs_.push_back(p0 + k); }
      const int np2 = (int)rel.size();
      std::vector<double> sv((size_t)np2 * CTH_STATS_NOUT);
      if (np2) {
        const size_t wneed = sizeof(int) * 2 * np2 + sizeof(int) * 641 * (size_t)np2 +
                             sizeof(double) * CTH_STATS_NOUT * (size_t)np2 + 256;
        int* d_abs = (int*)bufs.work(wneed);
        if (!d_abs) return -1;
        int* d_rel = d_abs + np2;
        int* d_hist = d_rel + np2;
        double* d_out = (double*)(((uintptr_t)(d_hist + 641 * (size_t)np2) + 15) & ~(uintptr_t)15);
        CTH_HIP_CHECK(hipMemcpyAsync(d_abs, abs_.data(), sizeof(int) * np2, hipMemcpyHostToDevice, c->st));
        CTH_HIP_CHECK(hipMemcpyAsync(d_rel, rel.data(), sizeof(int) * np2, hipMemcpyHostToDevice, c->st));
        if (timing) CTH_HIP_CHECK(hipEventRecord(c->ev[3], c->st));
        if (cth_launch_stats(&c->q.d, qi, &db.d, p0, col0, ldsc, scores, d_abs, d_rel, dpout, np2, sspace,
                             opt->ssemodel, d_hist, d_out, c->st))
          return -1;
        if (timing) CTH_HIP_CHECK(hipEventRecord(c->ev[0], c->st));
        CTH_HIP_CHECK(hipMemcpyAsync(sv.data(), d_out, sizeof(double) * sv.size(), hipMemcpyDeviceToHost, c->st));
        CTH_HIP_CHECK(hipStreamSynchronize(c->st));
        if (timing) {
          float a = 0.0f;
          CTH_HIP_CHECK(hipEventElapsedTime(&a, c->ev[3], c->ev[0]));
          st.ms_k_stats += a;
        }
      }
      // MAP stage for the passed profiles within the E-value threshold (the
      // reference's MAP kernels return early for the others,
      // cumapdp/CuBatchMAPDP_fwd.cu:127-130)
      // (jobs within the LDS row limit first, then the longer ones)
      std::vector<int> sel, sel_long;
      for (int j = 0; j < np2; j++)
        if (!(logethld < (float)sv[(size_t)j * CTH_STATS_NOUT + CTH_ST_LOGE]))
          (db.h_len[abs_[j]] > CTH_LDS_ROW_MAX ? sel_long : sel).push_back(j);
      const int nms = (int)sel.size();
      sel.insert(sel.end(), sel_long.begin(), sel_long.end());
      std::vector<cth_mapres> mres(sel.size());
      std::vector<char> hrows;  // alignment rows of the jobs, in traceback order
      int rowcap = 0;
      if (opt->do_map && !sel.empty()) {
        const int nj = (int)sel.size(), nml = nj - nms;
        std::vector<cth_mapjob> jobs(nj);
        int ltm = 0, ltms = 0;
        for (int k = 0; k < nj; k++) {
          jobs[k].profile = abs_[sel[k]];
          jobs[k].logepa = (float)sv[(size_t)sel[k] * CTH_STATS_NOUT + CTH_ST_LOGEPA];
          ltm = std::max(ltm, db.h_len[jobs[k].profile]);
          if (k < nms) ltms = ltm;
        }
        float* const post = bufs.map_buffer(lq, ldsc);
        if (!post) return -1;
        rowcap = lq + ltm;
        const size_t alnbytes = opt->alnbuf ? (size_t)nj * 5 * rowcap : 0;
        const size_t wneed = sizeof(double) * (size_t)nj * lq + sizeof(cth_mapjob) * nj + sizeof(cth_mapres) * nj +
                             alnbytes + 512;
        int nbm = nml;
        if (nml) {
          nbm = bufs.row_batch(cth_map_row_bytes(1, ltm), wneed, nml);
          if (nbm < 1) return -1;
        }
        double* d_zrow = (double*)bufs.work(wneed + (nml ? cth_map_row_bytes(nbm, ltm) + 256 : 0));
        if (!d_zrow) return -1;
        cth_mapjob* d_jobs = (cth_mapjob*)(d_zrow + (size_t)nj * lq);
        cth_mapres* d_res = (cth_mapres*)(((uintptr_t)(d_jobs + nj) + 15) & ~(uintptr_t)15);
        char* d_aln = alnbytes ? (char*)(d_res + nj) : nullptr;
        float* grow = nml ? (float*)(((uintptr_t)((char*)(d_res + nj) + alnbytes) + 255) & ~(uintptr_t)255) : nullptr;
        CTH_HIP_CHECK(hipMemcpyAsync(d_jobs, jobs.data(), sizeof(cth_mapjob) * nj, hipMemcpyHostToDevice, c->st));
        if (timing) CTH_HIP_CHECK(hipEventRecord(c->ev[3], c->st));
        const float minpp = opt->minpp > 0.0f ? opt->minpp : CTH_MINPP;
        if (cth_launch_map(&c->q.d, qi, lq, &db.d, d_jobs, nms, ltms, col0, ldsc, scores, mods, post, btck, d_zrow,
                           minpp, d_aln, rowcap, d_res, nullptr, c->st))
          return -1;
        for (int j = nms; j < nj; j += nbm) {
          const int n = std::min(nbm, nj - j);
          if (cth_launch_map(&c->q.d, qi, lq, &db.d, d_jobs + j, n, ltm, col0, ldsc, scores, mods, post, btck,
                             d_zrow + (size_t)j * lq, minpp, d_aln ? d_aln + (size_t)j * 5 * rowcap : nullptr, rowcap,
                             d_res + j, grow, c->st))
            return -1;
        }
        if (timing) CTH_HIP_CHECK(hipEventRecord(c->ev[0], c->st));
        CTH_HIP_CHECK(hipMemcpyAsync(mres.data(), d_res, sizeof(cth_mapres) * nj, hipMemcpyDeviceToHost, c->st));
        if (alnbytes) {
          hrows.resize(alnbytes);
          CTH_HIP_CHECK(hipMemcpyAsync(hrows.data(), d_aln, alnbytes, hipMemcpyDeviceToHost, c->st));
        }
        CTH_HIP_CHECK(hipStreamSynchronize(c->st));
        if (timing) {
          float a = 0.0f;
          CTH_HIP_CHECK(hipEventElapsedTime(&a, c->ev[3], c->ev[0]));
          st.ms_k_map += a;
        }
        for (int k = 0; k < nj; k++) st.map_cells += (double)lq * (double)db.h_len[jobs[k].profile];
      }
      const double t2 = now_ms();
      st.ms_phase1 += t1 - t0;
      st.ms_phase2 += t2 - t1;
      st.cells += (double)lq * (double)ncols;
      st.npassed += np2;
      st.nchunks += 1;
      for (size_t k = 0; k < sel.size(); k++) {
        const int j = sel[k];
        const double* s = &sv[(size_t)j * CTH_STATS_NOUT];
        // the host keeps hits with an alignment (CompressResultsPlain,
        // cupro/CuBatchProcessingFinalizerPlain.cpp:87-142)
        if (opt->do_map && mres[k].alnlen < 1) continue;
        if (nh >= maxhits) { cth_set_error("%s: hit buffer full (%d)", who, maxhits); return -1; }
        cth_hit& h = hits[nh++];
        memset(&h, 0, sizeof h);
        const cth_dpres& r = res[rel[j]];
        h.query = qi;
        h.profile = abs_[j];
        h.score = r.score;
        h.psts = r.psts;
        h.beg = r.beg;
        h.end = r.maxxy;
        h.logE = (float)s[CTH_ST_LOGE];
        h.logEpA = (float)s[CTH_ST_LOGEPA];
        h.lambda_est = (float)s[CTH_ST_LMBDEST];
        h.K_est = (float)s[CTH_ST_KEST];
        h.minsc = (float)s[CTH_ST_MIN];
        h.maxsc = (float)s[CTH_ST_MAX];
        h.expected = (float)s[CTH_ST_E];
        h.H = (float)s[CTH_ST_H];
        h.lambda = (float)s[CTH_ST_LMBD];
        h.K = (float)s[CTH_ST_K];
        if (opt->do_map) {
          h.mapbeg = mres[k].beg;
          h.mapend = mres[k].end;
          h.idnts = mres[k].idnts;
          h.ngaps = mres[k].ngaps;
          h.alnlen = mres[k].alnlen;
          h.pstvs_map = mres[k].psts;
          h.mapscore = mres[k].maxsc;
          if (opt->alnbuf) {
            // five rows of alnlen letters each, in reading order
            const size_t n = (size_t)mres[k].alnlen;
            if (alnused + 5 * n > opt->alncap) {
              cth_set_error("%s: alignment buffer full (%llu bytes)", who, (unsigned long long)opt->alncap);
              return -1;
            }
            h.alnoff = alnused;
            const char* src = hrows.data() + k * 5 * (size_t)rowcap;
            for (int r = 0; r < 5; r++)
              for (size_t m = 0; m < n; m++) opt->alnbuf[alnused + r * n + m] = src[r * (size_t)rowcap + n - 1 - m];
            alnused += 5 * n;
          }
        }
      }
      p0 = p1;
    }
    std::sort(hits + qh0, hits + nh, [](const cth_hit& a, const cth_hit& b) {
      return a.logE < b.logE || (a.logE == b.logE && a.profile < b.profile);
    });
    if (nhits) nhits[qi - q0] = nh - qh0;
  }
  if (stats) *stats = st;
  return 0;
}

}  // namespace

extern "C" int cth_search(cth_ctx* c, int q0, int q1, const cth_search_opts* opt, cth_hit* hits, int maxhits,
                          int* nhits, cth_search_stats* stats) {
  chunk_bufs bufs(c, nullptr, 0, "search");
  return search_run(c, q0, q1, opt, hits, maxhits, nhits, stats, bufs);
}

// The no-allocation variant: the same pipeline with every per-chunk device
// buffer carved out of this device's fixed scratch array.
extern "C" int cth_search_noalloc(cth_ctx* c, int q0, int q1, const cth_search_opts* opt, cth_hit* hits, int maxhits,
                                  int* nhits, cth_search_stats* stats) {
  if (!c) { cth_set_error("search_noalloc: no context"); return -1; }
  uint8_t* base = scratch_of(c->device);
  if (!base) {
    cth_set_error("search_noalloc: device %d has no scratch address (contexts set it up at creation)", c->device);
    return -1;
  }
  chunk_bufs bufs(c, base, CTH_NOALLOC_SCRATCH_BYTES, "search_noalloc");
  return search_run(c, q0, q1, opt, hits, maxhits, nhits, stats, bufs);
}
