// serve.cpp: a serving loop over the resident database.
//
// The reference runs queries in batches: JobDispatcher reads the query set,
// splits device memory into areas and hands each (query batch, database
// chunk) to device threads that call CuBatchProcessing serially
// (libmycu/cualn/JobDispatcher.cpp:967-1157, DevCommThread.cpp:149-356).
// Here the database stays resident and requests arrive one query at a time:
//
//  - slots: each is a worker thread with its own context (own HIP stream and
//    scratch buffers) that reads the parent context's tables, queries and
//    database, so requests on different slots overlap on the device;
//  - admission: a request's footprint is its score/DP/MAP buffer bytes,
//    Lq x chunk columns x bytes per cell (13 with MAP, 9 without, as the
//    search chunks them); a waiting request is admitted when a slot is free
//    and the footprints in flight plus its own fit the memory budget.  The
//    queue is scanned in arrival order and the first request that fits is
//    admitted, so a long query that does not fit yet does not hold back
//    shorter ones behind it;
//  - ticks: with batch > 1 a free slot takes, besides the first request
//    that fits, further waiting requests of distinct queries that fit next
//    to it (up to batch of them), and runs them as one batched search
//    (cth_search_batch_list: each stage once over all their pairs), so
//    queries of mixed lengths that arrive together share launches; with a
//    batching window (batch_wait_us) the slot waits that long from its first
//    request for more, while no other slot starts a tick, so that a burst
//    becomes one tick rather than one per idle slot;
//  - a slot keeps its scratch between requests (freeing device memory
//    synchronizes the device, which would stall the other slots), except
//    that scratch beyond both its share of the budget (mem_budget / nslots)
//    and its next request's footprint is returned first, so the slots hold
//    at most about max(mem_budget, one footprint per slot).
//
// Results are those of cth_search for the same query: requests do not
// share buffers, and every kernel's output depends only on its own inputs.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "cother_hip.h"

void cth_set_error(const char* fmt, ...);
extern "C" int cth_ctx_create_shared(cth_ctx* parent, cth_ctx** out);
extern "C" int cth_search_opts_check(const cth_search_opts* opt, const char* who);
extern "C" void cth_ctx_trim(cth_ctx* c, size_t keep);
extern "C" int cth_ctx_shape(const cth_ctx* c, int* nqueries, const int32_t** qlens, int* ndb, int64_t* dbpos);

namespace {

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Request {
  uint64_t ticket = 0;
  int qidx = 0;
  uint64_t footprint = 0;
  double t_submit = 0, t_start = 0, t_end = 0;
  int slot = -1;
  int group = 1;
  bool done = false;
  int rc = 0;
  std::string err;
  std::vector<cth_hit> hits;
  std::vector<char> aln;
  cth_search_stats stats{};
};

}  // namespace

struct cth_server {
  cth_ctx* parent = nullptr;
  cth_server_opts opt{};
  int nq = 0, ndb = 0;
  int64_t dbpos = 0;
  const int32_t* qlens = nullptr;
  std::vector<cth_ctx*> slots;
  std::vector<std::thread> workers;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> pending;
  std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
  uint64_t next_ticket = 1;
  uint64_t inflight = 0;
  int busy = 0;
  bool stop = false;
  bool gathering = false;  // a slot is in its batching window: the others do not start a tick

  uint64_t bytes_per_cell() const { return opt.search.do_map ? 13 : 9; }
  uint64_t chunk_cols() const {
    // the search pads a chunk to 64 columns per profile boundary at most
    const uint64_t whole = (uint64_t)dbpos + 64;
    return opt.chunk_cols ? std::min<uint64_t>(opt.chunk_cols, whole) : whole;
  }
  uint64_t footprint(int qidx) const { return (uint64_t)qlens[qidx] * chunk_cols() * bytes_per_cell(); }

  // first waiting request that fits, and whose query is not in the tick's
  // group yet (caller holds mu)
  Request* admissible(const std::vector<Request*>& group) {
    for (auto it = pending.begin(); it != pending.end(); ++it) {
      if (inflight + (*it)->footprint > opt.mem_budget) continue;
      bool dup = false;
      for (const Request* g : group) dup |= g->qidx == (*it)->qidx;
      if (dup) continue;
      Request* r = *it;
      pending.erase(it);
      return r;
    }
    return nullptr;
  }

  // one request: cth_search of its query
  void run_one(cth_ctx* c, Request* r) {
    cth_search_opts so = opt.search;
    so.mem_budget = r->footprint;
    so.alnbuf = nullptr;
    so.alncap = 0;
    if (opt.alncap) {
      r->aln.resize(opt.alncap);
      so.alnbuf = r->aln.data();
      so.alncap = opt.alncap;
    }
    r->hits.resize(std::max(1, ndb));
    int nh = 0;
    r->rc = cth_search(c, r->qidx, r->qidx + 1, &so, r->hits.data(), (int)r->hits.size(), &nh, &r->stats);
    if (r->rc) r->err = cth_last_error();
    r->hits.resize(r->rc ? 0 : nh);
  }

  // a tick's group of requests (distinct queries) in one batched search,
  // within the sum of their footprints; each request gets its query's hits
  // and alignment rows, and the tick's statistics
  void run_group(cth_ctx* c, const std::vector<Request*>& g, uint64_t budget) {
    const int n = (int)g.size();
    std::vector<int32_t> ql(n);
    for (int k = 0; k < n; k++) ql[k] = g[k]->qidx;
    cth_search_opts so = opt.search;
    so.mem_budget = budget;
    std::vector<char> aln(opt.alncap ? (size_t)opt.alncap * n : 0);
    so.alnbuf = opt.alncap ? aln.data() : nullptr;
    so.alncap = opt.alncap ? aln.size() : 0;
    std::vector<cth_hit> hits((size_t)std::max(1, ndb) * n);
    std::vector<int> nh(n, 0);
    cth_search_stats st{};
    const int rc = cth_search_batch_list(c, ql.data(), n, &so, hits.data(), (int)hits.size(), nh.data(), &st);
    if (rc) {
      // a failure of the batch (one request's rows not fitting, say) is
      // reported per request: each runs alone, as it would without batching
      for (Request* r : g) run_one(c, r);
      return;
    }
    size_t at = 0;
    for (int k = 0; k < n; k++) {
      Request* r = g[k];
      r->rc = 0;
      r->err.clear();
      r->stats = st;
      r->hits.assign(hits.begin() + at, hits.begin() + at + nh[k]);
      at += nh[k];
      if (opt.alncap) {
        // this request's rows, with offsets into its own buffer; rows over
        // the per-request capacity fail the request as cth_search fails it
        uint64_t used = 0;
        for (cth_hit& h : r->hits) used += 5 * (uint64_t)h.alnlen;
        if (used > opt.alncap) {
          char msg[96];
          snprintf(msg, sizeof msg, "search: alignment buffer full (%llu bytes)", (unsigned long long)opt.alncap);
          r->rc = -1;
          r->err = msg;
          r->hits.clear();
          r->aln.clear();
          continue;
        }
        r->aln.resize(std::max<uint64_t>(used, 1));
        used = 0;
        for (cth_hit& h : r->hits) {
          const uint64_t sz = 5 * (uint64_t)h.alnlen;
          memcpy(r->aln.data() + used, aln.data() + h.alnoff, sz);
          h.alnoff = used;
          used += sz;
        }
      }
    }
  }

  void worker(int slot) {
    cth_ctx* c = slots[slot];
    const size_t maxgroup = opt.batch > 1 ? (size_t)opt.batch : 1;
    for (;;) {
      std::vector<Request*> group;
      uint64_t fp = 0;
      {
        std::unique_lock<std::mutex> lk(mu);
        Request* r = nullptr;
        cv_work.wait(lk, [&] { return stop || (!gathering && (r = admissible(group)) != nullptr); });
        if (!r) return;  // stopping
        // the tick: this request and, up to the batch size, the next waiting
        // ones that fit next to it; with a batching window, the slot waits up
        // to batch_wait_us from its first request for more to arrive, the
        // only slot gathering meanwhile (so that a burst is not split)
        const bool window = maxgroup > 1 && opt.batch_wait_us > 0;
        const auto until = std::chrono::steady_clock::now() + std::chrono::microseconds(opt.batch_wait_us);
        if (window) gathering = true;
        for (;;) {
          inflight += r->footprint;
          fp += r->footprint;
          r->slot = slot;
          r->t_start = now_ms();
          group.push_back(r);
          if (group.size() >= maxgroup) break;
          r = admissible(group);
          if (!r && window) cv_work.wait_until(lk, until, [&] { return stop || (r = admissible(group)) != nullptr; });
          if (!r) break;
        }
        if (window) {
          gathering = false;
          cv_work.notify_all();
        }
        busy++;
      }
      cth_ctx_trim(c, std::max<uint64_t>(fp, opt.mem_budget / slots.size()));
      if (group.size() == 1) run_one(c, group[0]);
      else run_group(c, group, fp);
      {
        std::lock_guard<std::mutex> lk(mu);
        const double t = now_ms();
        for (Request* r : group) {
          r->t_end = t;
          r->group = (int)group.size();
          r->done = true;
        }
        inflight -= fp;
        busy--;
      }
      cv_done.notify_all();
      cv_work.notify_all();
    }
  }
};

extern "C" void cth_server_opts_init(cth_server_opts* o) {
  memset(o, 0, sizeof *o);
  o->nslots = 2;
  o->mem_budget = (uint64_t)64 << 30;
  o->chunk_cols = (uint64_t)4 << 20;
  o->search.eval = 10.0f;
  o->search.ssemodel = 2;
  o->search.do_map = 1;
  o->search.minpp = 0.28f;
}

extern "C" int cth_server_create(cth_ctx* parent, const cth_server_opts* o, cth_server** out) {
  *out = nullptr;
  if (!parent || !o || o->nslots < 1 || o->nslots > 64) {
    cth_set_error("server: invalid arguments");
    return -1;
  }
  if (o->batch_wait_us > 1000000u) {
    cth_set_error("server: batching window of %u us (at most 1 s)", o->batch_wait_us);
    return -1;
  }
  if (cth_search_opts_check(&o->search, "server")) return -1;
  std::unique_ptr<cth_server> s(new cth_server);
  s->parent = parent;
  s->opt = *o;
  if (cth_ctx_shape(parent, &s->nq, &s->qlens, &s->ndb, &s->dbpos)) return -1;
  for (int k = 0; k < o->nslots; k++) {
    cth_ctx* c = nullptr;
    if (cth_ctx_create_shared(parent, &c)) {
      for (cth_ctx* x : s->slots) cth_ctx_destroy(x);
      return -1;
    }
    s->slots.push_back(c);
  }
  for (int k = 0; k < o->nslots; k++) s->workers.emplace_back(&cth_server::worker, s.get(), k);
  *out = s.release();
  return 0;
}

extern "C" int cth_server_footprint(const cth_server* s, int qidx, uint64_t* bytes) {
  if (qidx < 0 || qidx >= s->nq) {
    cth_set_error("server: query %d out of range", qidx);
    return -1;
  }
  *bytes = s->footprint(qidx);
  return 0;
}

extern "C" int cth_server_submit(cth_server* s, int qidx, uint64_t* ticket) {
  if (qidx < 0 || qidx >= s->nq) {
    cth_set_error("server: query %d out of range", qidx);
    return -1;
  }
  const uint64_t fp = s->footprint(qidx);
  if (fp > s->opt.mem_budget) {
    cth_set_error("server: query %d needs %llu bytes, budget %llu", qidx, (unsigned long long)fp,
                  (unsigned long long)s->opt.mem_budget);
    return -1;
  }
  {
    std::lock_guard<std::mutex> lk(s->mu);
    std::unique_ptr<Request> r(new Request);
    r->ticket = s->next_ticket++;
    r->qidx = qidx;
    r->footprint = fp;
    r->t_submit = now_ms();
    *ticket = r->ticket;
    s->pending.push_back(r.get());
    s->reqs[r->ticket] = std::move(r);
  }
  s->cv_work.notify_all();
  return 0;
}

extern "C" int cth_server_wait(cth_server* s, uint64_t ticket, cth_hit* hits, int maxhits, int* nhits, char* alnbuf,
                               uint64_t alncap, cth_request_info* info) {
  std::unique_ptr<Request> r;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) {
      cth_set_error("server: unknown ticket %llu", (unsigned long long)ticket);
      return -1;
    }
    Request* p = it->second.get();
    s->cv_done.wait(lk, [&] { return p->done; });
    r = std::move(it->second);
    s->reqs.erase(it);
  }
  if (info) {
    info->qidx = r->qidx;
    info->slot = r->slot;
    info->footprint = r->footprint;
    info->ms_queued = r->t_start - r->t_submit;
    info->ms_latency = r->t_end - r->t_submit;
    info->ms_service = r->t_end - r->t_start;
    info->stats = r->stats;
    info->group = r->group;
  }
  if (r->rc) {
    cth_set_error("%s", r->err.c_str());
    return r->rc;
  }
  const int n = (int)r->hits.size();
  if (n > maxhits) {
    cth_set_error("server: %d hits, room for %d", n, maxhits);
    return -1;
  }
  uint64_t used = 0;
  for (int k = 0; k < n; k++) {
    cth_hit h = r->hits[k];
    if (alnbuf && !r->aln.empty()) {
      const uint64_t sz = 5 * (uint64_t)h.alnlen;
      if (used + sz > alncap) {
        cth_set_error("server: alignment buffer full");
        return -1;
      }
      memcpy(alnbuf + used, r->aln.data() + h.alnoff, sz);
      h.alnoff = used;
      used += sz;
    }
    hits[k] = h;
  }
  *nhits = n;
  return 0;
}

extern "C" void cth_server_destroy(cth_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto& t : s->workers) t.join();
  for (cth_ctx* c : s->slots) cth_ctx_destroy(c);
  delete s;
}
