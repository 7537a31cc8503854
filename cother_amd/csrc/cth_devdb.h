// cth_devdb.h: a device-resident profile set and its host-side index.
#pragma once

#include "cth_dev.h"
#include "cth_pset.h"

struct cth_devset {
  cth_dpset d;        // device pointers (kernel argument)
  int32_t* h_len;     // host copy of profile lengths
  int64_t* h_dist;    // host copy of profile offsets
};

extern "C" int cth_devset_upload(const cth_pset* h, cth_devset* s);
extern "C" void cth_devset_free(cth_devset* s);
