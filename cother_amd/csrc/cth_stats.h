// cth_stats.h: layout of the per-profile statistics record (double precision).
//
// Fields follow TDP2OutputAlnData's phase-2 part (cuss/CuBatchSS_com.h:54-66):
// estimated gapped lambda and K, log E-value, log pair E-value, min/max
// unscaled score, expected score, relative entropy, lambda, K.
#pragma once

enum {
  CTH_ST_LMBDEST = 0, CTH_ST_KEST, CTH_ST_LOGE, CTH_ST_LOGEPA,
  CTH_ST_MIN, CTH_ST_MAX, CTH_ST_E, CTH_ST_H, CTH_ST_LMBD, CTH_ST_K,
  CTH_STATS_NOUT
};
