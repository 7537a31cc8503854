// cth_map.h: MAP-stage job and result records, and its constants.
#pragma once

#include <stdint.h>

// CuBatchMAPDP_com.h:21-26; SLC_LOG_SP_MIN extsp/psl.h:43; MINPP options.txt:165
#define CTH_MAP_EPA_LOGFACTOR 22.0f
#define CTH_MAP_DYNS_OFFSET   0.05f
#define CTH_MAP_DYNS_MULTP    0.9f
#define CTH_LOG_SP_MIN        (-8.7336545e+01f)
#define CTH_MINPP             0.28f

// one profile to realign: database profile index and its log pair E-value
struct cth_mapjob {
  int32_t profile;
  float logepa;
};

// per-job result (the MAP fields of TDP2OutputAlnData, cuss/CuBatchSS_com.h:45-72)
struct cth_mapres {
  float maxsc;        // MAP alignment score at the end cell
  uint32_t end, beg;  // x<<16|y of the end cell and of the first aligned cell
  int32_t idnts, psts, ngaps, alnlen;
  float logz, maxlog; // log partition function, max log forward probability
};
