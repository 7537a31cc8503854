// cth_search.h: the public search types (defined in include/cother_hip.h).
#pragma once

#include "cother_hip.h"

// checks of the option fields every search entry point shares (0 or -1 with the error set)
extern "C" int cth_search_opts_check(const cth_search_opts* opt, const char* who);
