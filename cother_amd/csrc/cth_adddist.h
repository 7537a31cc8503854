// cth_adddist.h: adddist's per-position list transform.
//
// adddist (src/adddist/adddist.cpp) turns pairwise distances into the
// per-position distance lists the DDMS term reads.  For one position, with
// the distances to later positions read from its own lines and those to
// earlier positions collected from their lines (ProcessDistancesForPosition,
// adddist.cpp:560-651): keep the MAX_N_IR_DISTANCE_VALUES partners nearest
// in sequence, in sequence order; drop partners within DSTFREE of the
// position (the main-diagonal band); where partners are not adjacent,
// insert a penalty distance (DSTGAPL when the gap or the skipped band
// reaches DSTNPOS positions from the position, else DSTSEP6), repeated for
// gaps of up to DSTSEGM positions; finally cut the list to the maximum
// length, from the front near the end of the profile, else from the back.
// Option values: env/options.txt DSTSEGM, DSTNPOS, DSTSEP6, DSTGAPL, DSTFREE.
#pragma once

#include <stdlib.h>

#include <algorithm>
#include <numeric>
#include <utility>
#include <vector>

#include "cth_consts.h"

namespace cth_adddist {
constexpr int kDstSegm = 6, kDstNpos = 12, kDstFree = 3;
constexpr float kDstSep6 = 12.f, kDstGapl = 24.f;
}  // namespace cth_adddist

// position: 0-based; vdsts/vposs: the distances (and 0-based partner
// positions) of this position's own lines, in file order; tri[position]:
// (distance, earlier partner) pairs collected from earlier lines
inline void cth_adddist_position(int position, int L, std::vector<float>& vdsts, std::vector<int>& vposs,
                                 std::vector<std::vector<std::pair<float, int>>>& tri,
                                 std::vector<std::vector<float>>& dgram) {
  using namespace cth_adddist;
  for (const auto& dn : tri[position]) { vdsts.push_back(dn.first); vposs.push_back(dn.second); }
  std::vector<int> ndx(vdsts.size());
  std::iota(ndx.begin(), ndx.end(), 0);
  std::sort(ndx.begin(), ndx.end(), [position, &vposs](size_t n1, size_t n2) {
    return abs(position - vposs[n1]) < abs(position - vposs[n2]);
  });
  if ((int)ndx.size() > CTH_MAXDST) ndx.resize(CTH_MAXDST);
  std::sort(ndx.begin(), ndx.end(), [&vposs](size_t n1, size_t n2) { return vposs[n1] < vposs[n2]; });
  std::vector<float> out;
  out.reserve(2 * CTH_MAXDST);
  if (ndx.size() && kDstFree < abs(position - vposs[ndx[0]])) out.push_back(vdsts[ndx[0]]);
  int maxmainsep = 1;
  float dstpen = 0.f;
  for (size_t i = 1; i < ndx.size(); i++) {
    int posep = vposs[ndx[i]] - vposs[ndx[i - 1]];
    int mainsepp = abs(position - vposs[ndx[i - 1]]);
    int mainsep = abs(position - vposs[ndx[i]]);
    int mainsp2 = std::max(mainsepp, mainsep);
    if (mainsep <= kDstFree) {
      if (posep > 1 && maxmainsep < mainsp2) maxmainsep = mainsp2;
      continue;
    }
    if ((posep > 1 && mainsp2 >= kDstNpos) || maxmainsep >= kDstNpos)
      out.push_back(dstpen = kDstGapl);
    else if (posep > 1 || maxmainsep > 1)
      out.push_back(dstpen = kDstSep6);
    for (int k = 2; k < posep && k < kDstSegm; k++) out.push_back(dstpen);
    out.push_back(vdsts[ndx[i]]);
    maxmainsep = 1;
  }
  if ((int)out.size() > CTH_MAXDST) {
    if (position + CTH_MAXDST > L)
      out.erase(out.begin(), out.begin() + (out.size() - CTH_MAXDST));
    else
      out.resize(CTH_MAXDST);
  }
  dgram[position] = out;
}
