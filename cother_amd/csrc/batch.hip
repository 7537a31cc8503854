// batch.hip: the batched search -- many queries against the resident
// database in one pass of fused, persistent launches.
//
// cth_search runs the pipeline query by query and chunk by chunk, with the
// host choosing the next stage's profiles after each stage (four
// synchronisations per chunk).  For a batch of queries whose buffers fit the
// memory budget together, cth_search_batch instead runs every stage once
// over all (query, profile) pairs of the batch, and decides on the device
// which pairs go on to the next stage:
//   1. scores of every query of the batch (one launch, row tiles of all
//      queries);
//   2. phase-1 DP of every pair without traceback (persistent wavefronts
//      taking pairs, longest first, from a device counter);
//   3. pairs over their query's score threshold -> list (device compaction);
//   4. DP with traceback and finalisation of the listed pairs (persistent);
//   5. passed pairs -> list; their statistics (histogram in LDS, then the
//      parameters and E-value; persistent);
//   6. pairs within their query's E-value threshold -> list; MAP forward,
//      backward, MAP DP and traceback of each in one persistent kernel;
//   7. the hit records of the listed pairs gathered into one array.
// The host uploads the batch's tables, launches the stages -- captured once
// as a HIP graph and replayed while the batch shape (sizes, buffers,
// options) stays the same; with stage timing requested they run eagerly
// between events -- and copies back the counts and the hit records: one
// synchronisation per batch.  Each pair's arithmetic is the per-query
// kernels' (the same device functions), so the hits equal cth_search's; the
// lists' order varies, the hit order is restored by the (log E-value,
// profile) sort.
//
// The persistent grids' sizes (wavefronts per compute unit of the phase-1
// DP and of the later stages) are chosen by a small autotuner: the first
// batch of a shape runs once per candidate with events around it, and the
// fastest candidate is kept for that shape (database size and pair count).
//
// Covered: the FP32 contraction and databases whose profiles fit the LDS
// boundary rows (CTH_DP_LDS_ROW_MAX positions); other searches, and queries
// whose buffers alone exceed the budget, run through cth_search.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>
#include <vector>

#include "cth_batch.h"
#include "cth_ctx.h"
#include "cth_dp.h"
#include "cth_map.h"
#include "cth_search.h"
#include "cth_stats.h"

extern "C" int cth_launch_score_batch(const cth_dpset* q, const int32_t* qoff, const cth_dpset* db, int ncols,
                                      int64_t ldsc, const cth_tables* dtabs, const cth_bytile* ytab, int nytiles,
                                      const cth_bquery* bq, float* scores, float* mods, int ddms, hipStream_t st);
extern "C" int cth_score_rows_per_tile(int ddms);
extern "C" int cth_launch_dp_batch(const cth_dpset* q, const cth_dpset* db, const cth_bquery* bq,
                                   const cth_bpair* pairs, const int* list, const int* count, int n, int* next,
                                   int ltmax, int64_t ldsc, const float* scores, uint8_t* btck, cth_dpres* res,
                                   int grid, hipStream_t st);
extern "C" int cth_launch_stats_batch(const cth_dpset* q, const cth_dpset* db, const cth_bquery* bq,
                                      const cth_bpair* pairs, const int* list, const int* count, int64_t ldsc,
                                      const float* scores, const cth_dpres* dpres, int ssemodel, double* out,
                                      int grid, hipStream_t st);
extern "C" int cth_launch_map_batch(const cth_dpset* q, const cth_dpset* db, const cth_bquery* bq,
                                    const cth_bpair* pairs, const int* list, const int* count, int ltmax, int lqmax,
                                    int64_t ldsc, const float* scores, const float* mods, float* post, uint8_t* btck,
                                    const double* stats, double* zrow, float minpp, char* alnout, int rowcap,
                                    cth_mapres* res, int grid, hipStream_t st);

namespace {

enum { kCtrNext1 = 0, kCtrNext2, kCtrCand, kCtrPassed, kCtrHits, kNCtr = 8 };

__global__ void k_batch_reset(int* ctr) {
  if (threadIdx.x < kNCtr) ctr[threadIdx.x] = 0;
}

// Pairs whose phase-1 maximum exceeds their query's score threshold go to
// the traceback pass (the only ones FinalizeDPPro traces back,
// cudp/CuBatchDP_final.cu:144); the others get an empty record.
__global__ void k_select_cand(const cth_bquery* __restrict__ bq, const cth_bpair* __restrict__ pairs, int npairs,
                              cth_dpres* __restrict__ res, int* __restrict__ list, int* __restrict__ count) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < npairs; i += gridDim.x * blockDim.x) {
    cth_dpres r = res[i];
    if (bq[pairs[i].b].scorethld < r.maxsc) {
      list[atomicAdd(count, 1)] = i;
    } else {
      r.passed = 0; r.score = 0.0f; r.psts = 0; r.beg = 0;
      res[i] = r;
    }
  }
}

// passed pairs of the traceback pass -> list
__global__ void k_select_passed(const cth_dpres* __restrict__ res, const int* __restrict__ cand,
                                const int* __restrict__ ncand, int* __restrict__ list, int* __restrict__ count) {
  const int n = *ncand;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const int i = cand[k];
    if (res[i].passed) list[atomicAdd(count, 1)] = i;
  }
}

// passed pairs within their query's E-value threshold -> list (the profiles
// the MAP stage realigns, cumapdp/CuBatchMAPDP_fwd.cu:127-130)
__global__ void k_select_hits(const cth_bquery* __restrict__ bq, const cth_bpair* __restrict__ pairs,
                              const double* __restrict__ stats, const int* __restrict__ passed,
                              const int* __restrict__ npassed, int* __restrict__ list, int* __restrict__ count) {
  const int n = *npassed;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const int i = passed[k];
    const float loge = (float)stats[(int64_t)i * CTH_STATS_NOUT + CTH_ST_LOGE];
    if (!(bq[pairs[i].b].logethld < loge)) list[atomicAdd(count, 1)] = i;
  }
}

// the hit records of the listed pairs, with the fields cth_search fills;
// alnoff holds the pair's slot in the list (its rows' place in alnout)
__global__ void k_gather_hits(const cth_bquery* __restrict__ bq, const cth_bpair* __restrict__ pairs,
                              const cth_dpres* __restrict__ res, const double* __restrict__ stats,
                              const cth_mapres* __restrict__ mres, int do_map, const int* __restrict__ list,
                              const int* __restrict__ count, cth_hit* __restrict__ hits) {
  const int n = *count;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const int i = list[k];
    const cth_bpair pr = pairs[i];
    const cth_dpres r = res[i];
    const double* s = stats + (int64_t)i * CTH_STATS_NOUT;
    cth_hit h;
    memset(&h, 0, sizeof h);
    h.query = bq[pr.b].qp;
    h.profile = pr.p;
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
    if (do_map) {
      const cth_mapres m = mres[i];
      h.mapbeg = m.beg;
      h.mapend = m.end;
      h.idnts = m.idnts;
      h.ngaps = m.ngaps;
      h.alnlen = m.alnlen;
      h.pstvs_map = m.psts;
      h.mapscore = m.maxsc;
    }
    h.alnoff = (uint64_t)k;
    hits[k] = h;
  }
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

size_t up256(size_t n) { return (n + 255) & ~(size_t)255; }

// FNV-1a over the bytes of a plain struct
uint64_t fnv(const void* p, size_t n) {
  uint64_t h = 1469598103934665603ull;
  for (size_t k = 0; k < n; k++) h = (h ^ ((const uint8_t*)p)[k]) * 1099511628211ull;
  return h;
}

// The device layout of one batch: cell buffers (per query [Lq][ldsc], at the
// queries' offsets) and the small tables, records, lists and counters.
struct batch_layout {
  // cells
  float* scores; float* mods; float* post; uint8_t* btck;
  // small
  cth_bquery* bq; cth_bytile* ytab; cth_bpair* pairs; cth_dpres* res; double* stats; cth_mapres* mres;
  int* cand; int* passed; int* hitl; int* ctr; double* zrow; cth_hit* hits; char* alnout;
};

// One batch's description on the host.
struct batch_plan {
  int q0, q1;                       // entries [q0,q1) of the query list
  std::vector<cth_bquery> bq;
  std::vector<cth_bytile> ytab;
  std::vector<cth_bpair> pairs;
  int64_t cells = 0;
  int lqmax = 0;
  int ddms = 1;                     // 0: no cell of the batch has a DDMS term
};

struct launch_shape {
  int dp_grid, small_grid;
};

int capture_or_launch(cth_ctx* c, const batch_plan& P, const batch_layout& L, const cth_search_opts* opt,
                      int64_t ldsc, int ltmax, int rowcap, const launch_shape& S, bool capture, bool timing) {
  const cth_devset& db = c->db;
  const int npairs = (int)P.pairs.size();
  const int ntiles = (int)P.ytab.size();
  hipStream_t st = c->st;
  hipGraph_t graph = nullptr;
  if (capture) CTH_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  auto launches = [&]() -> int {
    if (timing) CTH_HIP_CHECK(hipEventRecord(c->ev[0], st));
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, st, L.ctr);
    CTH_HIP_CHECK(hipGetLastError());
    if (cth_launch_score_batch(&c->q.d, c->q.ddoff, &db.d, (int)db.d.npos, ldsc, c->dtab, L.ytab, ntiles, L.bq,
                               L.scores, L.mods, P.ddms, st))
      return -1;
    if (timing) CTH_HIP_CHECK(hipEventRecord(c->ev[1], st));
    if (cth_launch_dp_batch(&c->q.d, &db.d, L.bq, L.pairs, nullptr, nullptr, npairs, L.ctr + kCtrNext1, ltmax, ldsc,
                            L.scores, nullptr, L.res, std::min(S.dp_grid, npairs), st))
      return -1;
    const int sel_grid = std::max(1, std::min(1024, (npairs + 255) / 256));
    hipLaunchKernelGGL(k_select_cand, dim3(sel_grid), dim3(256), 0, st, L.bq, L.pairs, npairs, L.res, L.cand,
                       L.ctr + kCtrCand);
    CTH_HIP_CHECK(hipGetLastError());
    if (cth_launch_dp_batch(&c->q.d, &db.d, L.bq, L.pairs, L.cand, L.ctr + kCtrCand, 0, L.ctr + kCtrNext2, ltmax,
                            ldsc, L.scores, L.btck, L.res, std::min(S.small_grid, npairs), st))
      return -1;
    hipLaunchKernelGGL(k_select_passed, dim3(sel_grid), dim3(256), 0, st, L.res, L.cand, L.ctr + kCtrCand, L.passed,
                       L.ctr + kCtrPassed);
    CTH_HIP_CHECK(hipGetLastError());
    if (timing) CTH_HIP_CHECK(hipEventRecord(c->ev[2], st));
    if (cth_launch_stats_batch(&c->q.d, &db.d, L.bq, L.pairs, L.passed, L.ctr + kCtrPassed, ldsc, L.scores, L.res,
                               opt->ssemodel, L.stats, std::min(S.small_grid, npairs), st))
      return -1;
    hipLaunchKernelGGL(k_select_hits, dim3(sel_grid), dim3(256), 0, st, L.bq, L.pairs, L.stats, L.passed,
                       L.ctr + kCtrPassed, L.hitl, L.ctr + kCtrHits);
    CTH_HIP_CHECK(hipGetLastError());
    if (timing) CTH_HIP_CHECK(hipEventRecord(c->ev[3], st));
    if (opt->do_map &&
        cth_launch_map_batch(&c->q.d, &db.d, L.bq, L.pairs, L.hitl, L.ctr + kCtrHits, ltmax, P.lqmax, ldsc, L.scores,
                             L.mods, L.post, L.btck, L.stats, L.zrow, opt->minpp > 0.0f ? opt->minpp : CTH_MINPP,
                             L.alnout, rowcap, L.mres, std::min(S.small_grid, npairs), st))
      return -1;
    hipLaunchKernelGGL(k_gather_hits, dim3(sel_grid), dim3(256), 0, st, L.bq, L.pairs, L.res, L.stats, L.mres,
                       opt->do_map, L.hitl, L.ctr + kCtrHits, L.hits);
    CTH_HIP_CHECK(hipGetLastError());
    if (timing) CTH_HIP_CHECK(hipEventRecord(c->ev[4], st));
    return 0;
  };
  const int rc = launches();
  if (capture) {
    const hipError_t e = hipStreamEndCapture(st, &graph);
    if (rc) { if (graph) (void)hipGraphDestroy(graph); return -1; }
    if (e != hipSuccess) { cth_set_error("batch: graph capture: %s", hipGetErrorString(e)); return -1; }
    if (c->has_bgraph) { (void)hipGraphExecDestroy(c->bgraph); c->has_bgraph = 0; }
    const hipError_t ei = hipGraphInstantiate(&c->bgraph, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) { cth_set_error("batch: graph instantiation: %s", hipGetErrorString(ei)); return -1; }
    c->has_bgraph = 1;
    CTH_HIP_CHECK(hipGraphLaunch(c->bgraph, st));
  }
  return rc;
}

}  // namespace

namespace {

// the batched search over the queries ql[0..nq) (distinct indices), hits per
// query in list order; nhits[k] for query ql[k]
int search_batch_list(cth_ctx* c, const int* ql, int nq, const cth_search_opts* opt, cth_hit* hits, int maxhits,
                      int* nhits, cth_search_stats* stats) {
  CTH_HIP_CHECK(hipSetDevice(c->device));
  const cth_devset& db = c->db;
  const int nprof = db.d.nprof;
  const int64_t npos = db.d.npos;
  const int ltmax = nprof ? *std::max_element(db.h_len, db.h_len + nprof) : 0;
  cth_search_stats st{};
  // the batched form covers the FP32 contraction and profiles within the LDS
  // row limit; otherwise the queries run through cth_search
  if (opt->contraction != 0 || ltmax > CTH_DP_LDS_ROW_MAX || nprof == 0) {
    int nh = 0;
    uint64_t used = 0;
    for (int k = 0; k < nq; k++) {
      cth_search_opts o1 = *opt;
      if (opt->alnbuf) {
        o1.alnbuf = opt->alnbuf + used;
        o1.alncap = opt->alncap - used;
      }
      int n1 = 0;
      cth_search_stats s1{};
      if (cth_search(c, ql[k], ql[k] + 1, &o1, hits + nh, maxhits - nh, &n1, &s1)) return -1;
      for (int j = 0; j < n1; j++) {
        if (opt->alnbuf) hits[nh + j].alnoff += used;
        if (opt->alnbuf && opt->do_map) used += 5 * (uint64_t)hits[nh + j].alnlen;
      }
      if (nhits) nhits[k] = n1;
      nh += n1;
      st.cells += s1.cells; st.npassed += s1.npassed; st.nchunks += s1.nchunks; st.map_cells += s1.map_cells;
      st.ms_phase1 += s1.ms_phase1; st.ms_phase2 += s1.ms_phase2; st.ms_k_score += s1.ms_k_score;
      st.ms_k_dp += s1.ms_k_dp; st.ms_k_stats += s1.ms_k_stats; st.ms_k_map += s1.ms_k_map;
    }
    if (stats) *stats = st;
    return 0;
  }
  const uint64_t dbsize = opt->dbsize ? opt->dbsize : (uint64_t)npos;
  const uint64_t ndbent = opt->ndb ? opt->ndb : (uint64_t)nprof;
  const size_t budget = opt->mem_budget ? opt->mem_budget : ((size_t)64 << 30);
  const int64_t ldsc = (npos + 63) / 64 * 64;
  const size_t bytes_per_cell = opt->do_map ? 13 : 9;
  const bool timing = opt->timing != 0;
  if (timing && !c->has_ev) {
    for (int k = 0; k < 4; k++) CTH_HIP_CHECK(hipEventCreate(&c->ev[k]));
    c->has_ev = 1;
  }
  if (timing && c->nbev < 5) {
    CTH_HIP_CHECK(hipEventCreate(&c->ev[4]));
    c->nbev = 5;
  }
  int ncu = 0;
  CTH_HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
  // profiles in decreasing length: the longest DPs start first
  std::vector<int> porder(nprof);
  std::iota(porder.begin(), porder.end(), 0);
  std::stable_sort(porder.begin(), porder.end(), [&](int a, int b) { return db.h_len[a] > db.h_len[b]; });

  int nh = 0;
  uint64_t alnused = 0;
  // batch-local index of each query of the current batch (for its hits)
  std::vector<int> local(c->q.d.nprof, -1);
  const int rows4 = cth_score_rows_per_tile(1);
  for (int qa = 0; qa < nq;) {
    // the batch: consecutive entries of the list whose buffers fit the
    // budget together
    batch_plan P;
    P.q0 = qa;
    size_t bytes = 0;
    int qb = qa;
    int tiles = 0;
    while (qb < nq) {
      const int lq = c->q.h_len[ql[qb]];
      const size_t qbytes = (size_t)lq * (size_t)ldsc * bytes_per_cell;
      // row tiles of the 4-row score kernel (the plain kernel's taller tiles are fewer)
      const int qt = (lq + rows4 - 1) / rows4;
      if (qb > qa && (bytes + qbytes > budget || tiles + qt > 65535)) break;
      bytes += qbytes;
      tiles += qt;
      qb++;
    }
    P.q1 = qb;
    // a query whose buffers alone exceed the budget runs in chunks through
    // cth_search (its alignment rows after those already written)
    if (qb == qa + 1 && bytes > budget) {
      cth_search_opts o1 = *opt;
      if (opt->alnbuf) {
        o1.alnbuf = opt->alnbuf + alnused;
        o1.alncap = opt->alncap - alnused;
      }
      int n1 = 0;
      cth_search_stats s1{};
      if (cth_search(c, ql[qa], ql[qa] + 1, &o1, hits + nh, maxhits - nh, &n1, &s1)) return -1;
      for (int k = 0; k < n1; k++) {
        if (opt->alnbuf) hits[nh + k].alnoff += alnused;
        if (opt->alnbuf && opt->do_map) alnused += 5 * (uint64_t)hits[nh + k].alnlen;
      }
      if (nhits) nhits[qa] = n1;
      nh += n1;
      st.cells += s1.cells; st.npassed += s1.npassed; st.nchunks += s1.nchunks; st.map_cells += s1.map_cells;
      st.ms_phase1 += s1.ms_phase1; st.ms_phase2 += s1.ms_phase2; st.ms_k_score += s1.ms_k_score;
      st.ms_k_dp += s1.ms_k_dp; st.ms_k_stats += s1.ms_k_stats; st.ms_k_map += s1.ms_k_map;
      qa++;
      continue;
    }
    // no DDMS term in the batch when the database's or every query's
    // distance lists are shorter than DSTSEGM: the plain score kernel and
    // its taller row tiles
    P.ddms = 0;
    if (cth_devset_has_ddms(&db, 0, nprof))
      for (int k = P.q0; k < P.q1 && !P.ddms; k++) P.ddms = cth_devset_has_ddms(&c->q, ql[k], ql[k] + 1);
    const int rows = cth_score_rows_per_tile(P.ddms);
    for (int k = P.q0; k < P.q1; k++) {
      const int qi = ql[k];
      local[qi] = k - P.q0;
      cth_bquery Q{};
      Q.qp = qi;
      Q.lq = c->q.h_len[qi];
      Q.off = P.cells;
      cth_query_thresholds(opt->eval, (uint64_t)Q.lq, dbsize, ndbent, &Q.scorethld, &Q.logethld, &Q.sspace);
      const int b = (int)P.bq.size();
      P.bq.push_back(Q);
      for (int r0 = 0; r0 < Q.lq; r0 += rows) P.ytab.push_back(cth_bytile{b, r0});
      P.cells += (int64_t)Q.lq * ldsc;
      P.lqmax = std::max(P.lqmax, Q.lq);
    }
    // pairs: longest DPs (query length x profile length) first, equal
    // sizes in query-then-profile order (a stable sort; counting sort on the
    // size where the sizes are small enough for a table, ~4x faster on the
    // host per batch of 128,000 pairs)
    {
      std::vector<cth_bpair> src;
      src.reserve((size_t)P.bq.size() * porder.size());
      int64_t maxsz = 0;
      for (int b = 0; b < (int)P.bq.size(); b++) {
        for (int p : porder) src.push_back(cth_bpair{b, p});
        if (!porder.empty()) maxsz = std::max(maxsz, (int64_t)P.bq[b].lq * db.h_len[porder[0]]);
      }
      auto size_of = [&](const cth_bpair& x) { return (int64_t)P.bq[x.b].lq * db.h_len[x.p]; };
      if (maxsz < (int64_t)1 << 22) {
        std::vector<int> at((size_t)maxsz + 2, 0);
        for (const cth_bpair& x : src) at[(size_t)(maxsz - size_of(x)) + 1]++;
        for (size_t k = 1; k < at.size(); k++) at[k] += at[k - 1];
        P.pairs.resize(src.size());
        for (const cth_bpair& x : src) P.pairs[at[(size_t)(maxsz - size_of(x))]++] = x;
      } else {
        std::stable_sort(src.begin(), src.end(),
                         [&](const cth_bpair& x, const cth_bpair& y) { return size_of(x) > size_of(y); });
        P.pairs.swap(src);
      }
    }
    const int npairs = (int)P.pairs.size();
    const int rowcap = P.lqmax + ltmax;
    const size_t alnbytes = opt->alnbuf && opt->do_map ? (size_t)npairs * 5 * (size_t)rowcap : 0;
    // device buffers
    const size_t cells = (size_t)P.cells;
    const size_t cbytes = up256(4 * cells) * (opt->do_map ? 3 : 2) + up256(cells);
    const int small_max = std::max(1, std::min(npairs, ncu * 32));
    const size_t sbytes = up256(sizeof(cth_bquery) * P.bq.size()) + up256(sizeof(cth_bytile) * P.ytab.size()) +
                          up256(sizeof(cth_bpair) * npairs) + up256(sizeof(cth_dpres) * npairs) +
                          up256(sizeof(double) * CTH_STATS_NOUT * npairs) + up256(sizeof(cth_mapres) * npairs) +
                          3 * up256(sizeof(int) * npairs) + up256(sizeof(int) * kNCtr) +
                          up256(sizeof(double) * (size_t)small_max * P.lqmax) + up256(sizeof(cth_hit) * npairs) +
                          up256(alnbytes);
    if (cth_ctx_reserve(c, &c->bcells, &c->cap_bcells, cbytes)) return -1;
    if (cth_ctx_reserve(c, &c->bsmall, &c->cap_bsmall, sbytes)) return -1;
    batch_layout L{};
    {
      uint8_t* o = (uint8_t*)c->bcells;
      auto take = [&](size_t n) { uint8_t* p = o; o += up256(n); return p; };
      L.scores = (float*)take(4 * cells);
      L.mods = (float*)take(4 * cells);
      L.post = opt->do_map ? (float*)take(4 * cells) : nullptr;
      L.btck = take(cells);
      uint8_t* s = (uint8_t*)c->bsmall;
      auto takes = [&](size_t n) { uint8_t* p = s; s += up256(n); return p; };
      L.bq = (cth_bquery*)takes(sizeof(cth_bquery) * P.bq.size());
      L.ytab = (cth_bytile*)takes(sizeof(cth_bytile) * P.ytab.size());
      L.pairs = (cth_bpair*)takes(sizeof(cth_bpair) * npairs);
      L.res = (cth_dpres*)takes(sizeof(cth_dpres) * npairs);
      L.stats = (double*)takes(sizeof(double) * CTH_STATS_NOUT * npairs);
      L.mres = (cth_mapres*)takes(sizeof(cth_mapres) * npairs);
      L.cand = (int*)takes(sizeof(int) * npairs);
      L.passed = (int*)takes(sizeof(int) * npairs);
      L.hitl = (int*)takes(sizeof(int) * npairs);
      L.ctr = (int*)takes(sizeof(int) * kNCtr);
      L.zrow = (double*)takes(sizeof(double) * (size_t)small_max * P.lqmax);
      L.hits = (cth_hit*)takes(sizeof(cth_hit) * npairs);
      L.alnout = alnbytes ? (char*)takes(alnbytes) : nullptr;
    }
    // the batch's tables (outside the graph: a replayed graph reads the current ones)
    CTH_HIP_CHECK(hipMemcpyAsync(L.bq, P.bq.data(), sizeof(cth_bquery) * P.bq.size(), hipMemcpyHostToDevice, c->st));
    CTH_HIP_CHECK(hipMemcpyAsync(L.ytab, P.ytab.data(), sizeof(cth_bytile) * P.ytab.size(), hipMemcpyHostToDevice,
                                 c->st));
    CTH_HIP_CHECK(hipMemcpyAsync(L.pairs, P.pairs.data(), sizeof(cth_bpair) * npairs, hipMemcpyHostToDevice, c->st));
    const double t0 = now_ms();
    // persistent grid sizes: tuned once per class of shapes (database size
    // and pair count rounded up to a power of two), so that batches of
    // similar size (a serving loop's ticks) share the tuning instead of each
    // paying for its candidates; shared contexts use their parent's table
    uint64_t pbucket = 1;
    while (pbucket < (uint64_t)npairs) pbucket <<= 1;
    const uint64_t tkey = ((uint64_t)npos << 24) ^ pbucket;
    cth_ctx* tc = c->borrowed ? c->parent : c;
    launch_shape S{ncu * 20, ncu * 8};
    bool tuned = false;
    {
      std::lock_guard<std::mutex> g(tc->tune_mu);
      for (int k = 0; k < tc->nbtune && k < 32; k++)
        if (tc->btune[k].key == tkey) {
          S = launch_shape{tc->btune[k].dp_grid, tc->btune[k].small_grid};
          tuned = true;
        }
    }
    if (!tuned) {
      // wavefronts per compute unit for the phase-1 DP: measured on 128,000
      // pairs of <= 150 x <= 400 positions (LDS for 14 per unit), 4: 47.5 ms,
      // 6: 32.5, 8: 25.9, 10: 23.8, 12: 28.2, 14: 29.4.  The batch runs once
      // with the first candidate (its results stand); then the DP's first
      // pass alone runs again per candidate on the scores in place (it
      // rewrites the same maxima), so that tuning a class costs one batch and
      // a few DP passes, not a batch per candidate (with distance lists the
      // score stage is most of a batch)
      const int dp_cand[] = {ncu * 8, ncu * 10, ncu * 20};
      hipEvent_t e0, e1;
      CTH_HIP_CHECK(hipEventCreate(&e0));
      CTH_HIP_CHECK(hipEventCreate(&e1));
      S = launch_shape{dp_cand[0], ncu * 4};
      if (capture_or_launch(c, P, L, opt, ldsc, ltmax, rowcap, S, false, false)) return -1;
      // (each candidate twice, the faster run counting: single runs of
      // candidates a few percent apart are not told apart reliably)
      float best = 1e30f;
      for (int rep = 0; rep < 2; rep++)
        for (const int g : dp_cand) {
          CTH_HIP_CHECK(hipMemsetAsync(L.ctr + kCtrNext1, 0, sizeof(int), c->st));
          CTH_HIP_CHECK(hipEventRecord(e0, c->st));
          if (cth_launch_dp_batch(&c->q.d, &db.d, L.bq, L.pairs, nullptr, nullptr, npairs, L.ctr + kCtrNext1,
                                  ltmax, ldsc, L.scores, nullptr, L.res, std::min(g, npairs), c->st))
            return -1;
          CTH_HIP_CHECK(hipEventRecord(e1, c->st));
          CTH_HIP_CHECK(hipEventSynchronize(e1));
          float ms = 0.0f;
          CTH_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
          if (ms < best) { best = ms; S.dp_grid = g; }
        }
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      // the table keeps the last 32 classes (another slot may have added
      // this one meanwhile: then its entry stays)
      std::lock_guard<std::mutex> g(tc->tune_mu);
      bool have = false;
      for (int k = 0; k < tc->nbtune && k < 32; k++) have |= tc->btune[k].key == tkey;
      if (!have) {
        tc->btune[tc->nbtune % 32] = cth_btune{tkey, S.dp_grid, S.small_grid, best};
        tc->nbtune++;
      }
    }
    // the graph: replayed when everything its launches captured is the same
    // (buffers and their layout, sizes, options, the uploaded sets)
    uint32_t mp;
    memcpy(&mp, &opt->minpp, sizeof mp);
    uint64_t key[12] = {(uint64_t)(uintptr_t)c->bcells, (uint64_t)(uintptr_t)c->bsmall, (uint64_t)npairs,
                        (uint64_t)P.ytab.size() | ((uint64_t)P.bq.size() << 32),
                        (uint64_t)P.cells, (uint64_t)P.lqmax | ((uint64_t)rowcap << 32),
                        (uint64_t)S.dp_grid | ((uint64_t)S.small_grid << 32),
                        (uint64_t)opt->do_map | ((uint64_t)opt->ssemodel << 8) | ((uint64_t)(alnbytes != 0) << 17) |
                            ((uint64_t)P.ddms << 18) | ((uint64_t)mp << 32),
                        fnv(&db.d, sizeof db.d), fnv(&c->q.d, sizeof c->q.d),
                        (uint64_t)(uintptr_t)c->q.ddoff, (uint64_t)(uintptr_t)c->dtab};
    if (timing) {
      // stage events: the launches run eagerly
      if (capture_or_launch(c, P, L, opt, ldsc, ltmax, rowcap, S, false, true)) return -1;
    } else if (c->has_bgraph && memcmp(key, c->bkey, sizeof key) == 0) {
      CTH_HIP_CHECK(hipGraphLaunch(c->bgraph, c->st));
      c->nbgraph_replays++;
    } else {
      if (capture_or_launch(c, P, L, opt, ldsc, ltmax, rowcap, S, true, false)) return -1;
      memcpy(c->bkey, key, sizeof key);
      c->nbgraph_captures++;
    }
    // results: counts, then the hit records (and their alignment rows)
    int ctr[kNCtr];
    CTH_HIP_CHECK(hipMemcpyAsync(ctr, L.ctr, sizeof ctr, hipMemcpyDeviceToHost, c->st));
    CTH_HIP_CHECK(hipStreamSynchronize(c->st));
    const int nb = ctr[kCtrHits];
    std::vector<cth_hit> bh(nb);
    std::vector<char> rows_h;
    if (nb) CTH_HIP_CHECK(hipMemcpyAsync(bh.data(), L.hits, sizeof(cth_hit) * nb, hipMemcpyDeviceToHost, c->st));
    if (nb && alnbytes) {
      rows_h.resize((size_t)nb * 5 * rowcap);
      CTH_HIP_CHECK(hipMemcpyAsync(rows_h.data(), L.alnout, rows_h.size(), hipMemcpyDeviceToHost, c->st));
    }
    CTH_HIP_CHECK(hipStreamSynchronize(c->st));
    const double t1 = now_ms();
    if (timing) {
      float a = 0, b = 0, s2 = 0, m = 0;
      CTH_HIP_CHECK(hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
      CTH_HIP_CHECK(hipEventElapsedTime(&b, c->ev[1], c->ev[2]));
      CTH_HIP_CHECK(hipEventElapsedTime(&s2, c->ev[2], c->ev[3]));
      CTH_HIP_CHECK(hipEventElapsedTime(&m, c->ev[3], c->ev[4]));
      st.ms_k_score += a;
      st.ms_k_dp += b;
      st.ms_k_stats += s2;
      st.ms_k_map += m;
    }
    // hits per query: those with an alignment (with MAP), sorted by (log E, profile)
    std::vector<std::vector<cth_hit>> per(P.q1 - P.q0);
    for (const cth_hit& h : bh) {
      if (opt->do_map) st.map_cells += (double)c->q.h_len[h.query] * (double)db.h_len[h.profile];
      if (opt->do_map && h.alnlen < 1) continue;
      per[local[h.query]].push_back(h);
    }
    for (int b = 0; b < P.q1 - P.q0; b++) {
      auto& v = per[b];
      std::sort(v.begin(), v.end(), [](const cth_hit& x, const cth_hit& y) {
        return x.logE < y.logE || (x.logE == y.logE && x.profile < y.profile);
      });
      if (nh + (int)v.size() > maxhits) { cth_set_error("search_batch: hit buffer full (%d)", maxhits); return -1; }
      for (cth_hit h : v) {
        if (alnbytes) {
          const size_t n = (size_t)h.alnlen;
          if (alnused + 5 * n > opt->alncap) {
            cth_set_error("search_batch: alignment buffer full (%llu bytes)", (unsigned long long)opt->alncap);
            return -1;
          }
          const char* src = rows_h.data() + h.alnoff * 5 * (size_t)rowcap;
          for (int r = 0; r < 5; r++)
            for (size_t m = 0; m < n; m++) opt->alnbuf[alnused + r * n + m] = src[r * (size_t)rowcap + n - 1 - m];
          h.alnoff = alnused;
          alnused += 5 * n;
        } else {
          h.alnoff = 0;
        }
        hits[nh++] = h;
      }
      if (nhits) nhits[P.q0 + b] = (int)v.size();
    }
    st.cells += (double)P.cells;
    st.npassed += ctr[kCtrPassed];
    st.nchunks += 1;
    st.ms_phase1 += t1 - t0;
    qa = P.q1;
  }
  if (stats) *stats = st;
  return 0;
}

}  // namespace

extern "C" int cth_search_batch(cth_ctx* c, int q0, int q1, const cth_search_opts* opt, cth_hit* hits, int maxhits,
                                int* nhits, cth_search_stats* stats) {
  if (!c->has_q || !c->has_db) { cth_set_error("search_batch: queries or database not uploaded"); return -1; }
  if (q0 < 0 || q1 > c->q.d.nprof || q0 >= q1) { cth_set_error("search_batch: bad query range [%d,%d)", q0, q1); return -1; }
  if (cth_search_opts_check(opt, "search_batch")) return -1;
  std::vector<int> ql(q1 - q0);
  std::iota(ql.begin(), ql.end(), q0);
  return search_batch_list(c, ql.data(), q1 - q0, opt, hits, maxhits, nhits, stats);
}

extern "C" int cth_search_batch_list(cth_ctx* c, const int32_t* qidx, int nq, const cth_search_opts* opt,
                                     cth_hit* hits, int maxhits, int* nhits, cth_search_stats* stats) {
  if (!c->has_q || !c->has_db) { cth_set_error("search_batch_list: queries or database not uploaded"); return -1; }
  if (!qidx || nq < 1) { cth_set_error("search_batch_list: empty query list"); return -1; }
  if (cth_search_opts_check(opt, "search_batch_list")) return -1;
  std::vector<char> seen(c->q.d.nprof, 0);
  std::vector<int> ql(nq);
  for (int k = 0; k < nq; k++) {
    const int q = qidx[k];
    if (q < 0 || q >= c->q.d.nprof) { cth_set_error("search_batch_list: query %d out of range", q); return -1; }
    if (seen[q]) { cth_set_error("search_batch_list: query %d listed twice", q); return -1; }
    seen[q] = 1;
    ql[k] = q;
  }
  return search_batch_list(c, ql.data(), nq, opt, hits, maxhits, nhits, stats);
}

// Counts of graph captures and replays of the batched search on a context
// (measurement aid).
extern "C" int cth_batch_graph_counts(const cth_ctx* c, int64_t* captures, int64_t* replays) {
  if (captures) *captures = c->nbgraph_captures;
  if (replays) *replays = c->nbgraph_replays;
  return 0;
}
