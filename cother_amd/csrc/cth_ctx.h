// cth_ctx.h: the device context behind the C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>

#include "cth_dev.h"
#include "cth_devdb.h"
#include "cth_tables.h"

// an autotuned launch shape of the batched search, per batch shape
struct cth_btune {
  uint64_t key;
  int dp_grid, small_grid;
  float ms;  // the chosen phase-1 DP pass's time when tuned
};

struct cth_ctx {
  int device;
  hipStream_t st;
  cth_tables htab;
  cth_tables* dtab;
  cth_devset q, db;
  int has_q, has_db;
  // scratch, grown on demand
  float* scores; size_t cap_scores;
  float* mods; size_t cap_mods;
  uint8_t* btck; size_t cap_btck;
  float* dpout; size_t cap_dpout;
  void* work; size_t cap_work;
  float* mapbuf; size_t cap_mapbuf;  // MAP posteriors [Lq][ldsc]
  float* grow; size_t cap_grow;      // DP/MAP stripe boundary rows of profiles over the LDS row limit
  void* h_pin; size_t cap_pin;
  // stage events of cth_search's timing mode (created on first use; the
  // batched search uses a fifth)
  hipEvent_t ev[5]; int has_ev; int nbev;
  // batched search (batch.hip): cell buffers and tables of a batch, the
  // captured graph of the last batch shape and the tuned grid sizes
  void* bcells; size_t cap_bcells;
  void* bsmall; size_t cap_bsmall;
  hipGraphExec_t bgraph; int has_bgraph; uint64_t bkey[12];
  int64_t nbgraph_captures, nbgraph_replays;
  // tuned grid sizes; a shared context reads and adds to its parent's
  // table (under the parent's tune_mu), so the slots of a serving loop tune
  // each class of shapes once between them
  cth_btune btune[32]; int nbtune;
  std::mutex tune_mu;
  // a context made by cth_ctx_create_shared reads its parent's tables,
  // queries and database and frees only its own stream and scratch
  int borrowed;
  cth_ctx* parent;             // the context a shared one reads from
  std::atomic<int> nshared;    // live shared contexts reading this one's data
};

// release scratch buffers larger than keep bytes (the next allocation sizes them anew)
extern "C" void cth_ctx_trim(cth_ctx* c, size_t keep);

int cth_ctx_reserve(cth_ctx* c, void** p, size_t* cap, size_t need);
// set up the device address of cth_search_noalloc's scratch array (once per device, device current)
extern "C" int cth_scratch_init(int device);
extern "C" void cth_ctx_destroy(cth_ctx* c);
