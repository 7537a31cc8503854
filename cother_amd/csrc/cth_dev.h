// cth_dev.h: device-resident profile sets and common device helpers.
//
// A profile set (the query set or the template database) lives in HBM as
// structure-of-arrays: one array per field, each padded to a multiple of 64
// positions (ld) so that field f of position i is at base[f*ld + i] and a
// wavefront of 64 consecutive positions reads one aligned 256-byte line per
// float field.  Distance lists are 128-byte records per position (count in
// byte 0, then the distances saturated at 20, with 21 standing for the
// missing value 255, padded with 21).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cth_consts.h"
#include "cth_tables.h"

struct cth_dpset {
  int32_t nprof;
  int64_t npos;
  int64_t ld;              // padded leading dimension of position fields
  int64_t ldtr;            // padded leading dimension of transition fields (npos + nprof)
  const int32_t* len;      // [nprof]
  const int64_t* dist;     // [nprof]
  const float*   eno;      // [nprof]
  const float*   bkg;      // [nprof*20]
  const char*    aa;       // [ld]
  const char*    sss;      // [ld] SS state letters (0: the profile has no SS prediction)
  const int32_t* pro;      // [ld] profile index of each position
  const float*   trg;      // [20*ld]
  const float*   trn;      // [7*ldtr]
  const float*   ss;       // [3*ld]
  const float*   cv;       // [19*ld]
  const float*   cvprior;  // [ld]
  const float*   cvnorm2;  // [ld]
  const uint8_t* dd;       // [ld*128]
};

#define CTH_DD_MISSING 21

#define CTH_HIP_CHECK(expr)                                                        \
  do {                                                                             \
    hipError_t _e = (expr);                                                        \
    if (_e != hipSuccess) {                                                        \
      cth_set_error("%s:%d: %s: %s", __FILE__, __LINE__, #expr, hipGetErrorString(_e)); \
      return -1;                                                                   \
    }                                                                              \
  } while (0)

void cth_set_error(const char* fmt, ...);
