// cth_devdb.h: a device-resident profile set and its host-side index.
#pragma once

#include "cth_dev.h"
#include "cth_pset.h"

struct cth_devset {
  cth_dpset d;        // device pointers (kernel argument)
  int32_t* h_len;     // host copy of profile lengths
  int64_t* h_dist;    // host copy of profile offsets
  int32_t* ddoff;     // query sets: [npos][128] distance lists as score-table byte offsets
                      // (entry 0 = count, entry k = 22 * 4 * value k), staged by the score kernel
  int32_t* h_maxdd;   // host: per profile, its longest distance list (values)
};

// whether any cell of profiles [p0, p1) of s can have a nonzero DDMS term:
// some profile has a list of at least DSTSEGM values (shorter lists score
// zero, CuBatchScoreMatrix_ddms2s_h2.cuh:450)
int cth_devset_has_ddms(const cth_devset* s, int p0, int p1);

extern "C" int cth_devset_make_ddoff(const cth_pset* h, cth_devset* s);

extern "C" int cth_devset_upload(const cth_pset* h, cth_devset* s);
extern "C" void cth_devset_free(cth_devset* s);
