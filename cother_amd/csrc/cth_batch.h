// cth_batch.h: work descriptors of the batched search (many queries per
// launch, see batch.hip).
#pragma once

#include <stdint.h>

// one query of a batch: its index in the uploaded query set, its length, the
// element offset of its [Lq][ldsc] score / modular-score / traceback /
// posterior blocks in the batch buffers, and its thresholds and search space
struct cth_bquery {
  int32_t qp;
  int32_t lq;
  int64_t off;
  float scorethld;   // phase-1 score threshold
  float logethld;    // log E-value threshold
  double sspace;     // length-adjusted search space
};

// one (query, template profile) pair of a batch: batch-local query, database profile
struct cth_bpair {
  int32_t b;
  int32_t p;
};

// one row tile of the batched score kernel: batch-local query and its first row
struct cth_bytile {
  int32_t b;
  int32_t row0;
};
