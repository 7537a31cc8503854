// cth_search.h: the public search types (defined in include/cother_hip.h).
#pragma once

#include "cother_hip.h"

// checks of the option fields every search entry point shares (0 or -1 with the error set)
extern "C" int cth_search_opts_check(const cth_search_opts* opt, const char* who);
// a query's score threshold, log E-value threshold and length-adjusted search space
extern "C" void cth_query_thresholds(float eval, uint64_t lq, uint64_t dbsize, uint64_t ndb, float* scorethld,
                                     float* logethld, double* sspace);
