// cth_dp.h: phase-1 DP results per template profile.
#pragma once

#include <stdint.h>

// per-profile phase-1 result; the record fields follow TDP2OutputAlnData
// (cuss/CuBatchSS_com.h:45-52): score (with the correlation bonus), number of
// positive substitution scores, begin coordinates (x+1)<<16|(y+1), end
// coordinates x<<16|y of the maximum
struct cth_dpres {
  float maxsc;      // maximum MM value of the DP
  uint32_t maxxy;   // its coordinates, column<<16 | row
  int32_t passed;   // 1 if the profile passes to phase 2
  float score;      // maxsc + signed sqrt of half the correlation sum
  int32_t psts;     // positive DIAG scores along the traceback
  uint32_t beg;     // begin coordinates
};
