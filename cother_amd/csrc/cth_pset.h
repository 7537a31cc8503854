// cth_pset.h: host-side columnar profile set (queries or a template database).
//
// The field set and value conventions follow the reference's packed profile
// model (cupro/PM2DVectorFields.h:53-87) as filled by its text/binary readers
// (cupro/IOProfileModelTxt.cu:296-1030; the binary database sections of
// libpro/srcpro/PMProfileModel.cpp:1235-1817): target frequencies and background
// probabilities are k/10000, transitions are stored as log values (x0.4 for
// IM, II, DM, DD; -32768 for a zero probability), one extra "beginning"
// transition vector precedes the positions of every profile, context vectors
// of profiles without CV data hold the uninformative value 0.17302947, and
// distance lists hold the count n followed by n integer distances, 255 marking
// a missing value and padding the rest of the list.
//
// Storage is structure-of-arrays so that a lane per template position reads
// each field coalesced; distance lists are stored as 128-byte records so a
// lane loads its whole list with eight 16-byte loads.
#pragma once

#include <stdint.h>
#include <stddef.h>
#include "cth_consts.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cth_pset {
  int32_t  nprof;     // number of profiles
  int64_t  npos;      // total number of positions
  int32_t* len;       // [nprof] profile lengths
  int64_t* dist;      // [nprof] index of the first position of each profile
  float*   eno;       // [nprof] effective number of observations
  float*   bkg;       // [nprof*20] background probabilities (profile-major)
  char*    aa;        // [npos] residues
  char*    sss;       // [npos] SS state letter as the device sees it (' ', 'e', 'h' for C, E, H;
                      // liblib/alpha.cpp:14); 0 throughout a profile without SS predictions
  float*   trg;       // [20*npos] target frequencies, field-major
  float*   trn;       // [7*(npos+nprof)] log transitions, field-major; profile p vector k at dist[p]+p+k
  float*   ss;        // [3*npos] SS state probabilities, field-major
  float*   cv;        // [19*npos] context vectors, field-major
  float*   cvprior;   // [npos] log prior of the context vector
  float*   cvnorm2;   // [npos] squared norm of the context vector
  uint8_t* dd;        // [npos*128] distance lists: n, d1..dn, 255 padding
  char**   name;      // [nprof] profile names (the FILE field of text profiles), NULL entries allowed
  char**   desc;      // [nprof] profile descriptions (DESC field, database description section)
} cth_pset;

#ifdef __cplusplus
}
#endif
