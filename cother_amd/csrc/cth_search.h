// cth_search.h: the public search types (defined in include/cother_hip.h).
#pragma once

#include "cother_hip.h"
