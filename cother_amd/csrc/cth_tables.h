// cth_tables.h: score tables consumed by the score-matrix kernels.
#pragma once

#include "cth_consts.h"

#define CTH_CVMAXCARD 128

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cth_tables {
  float ss[467];               // 2 header cells + 30x30 lower triangle
  int   cvcard, cvshift;       // CVS map cardinality and index of key 0
  float cvkeys[CTH_CVMAXCARD];
  float cvvals[CTH_CVMAXCARD];
  float cvkappa0inv, cvpowernu0, cvcterm;
  float dms[22 * 22];          // distance-match pair scores, index 21 = missing
} cth_tables;

int cth_tables_load(const char* envdir, cth_tables* t);

#ifdef __cplusplus
}
#endif
