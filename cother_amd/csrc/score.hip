// score.hip: the score-matrix stage on gfx950.
//
// One kernel computes, for a query profile against a contiguous range of
// template positions, every cell of
//   S[i,j]   = Init + (SS + CVS) + DDMS
//   mod[i,j] = (SS + CVS - 0.5*CVSWGT) + (DDMS - 0.2*DDMSWGT)
// in the order the reference's three kernels accumulate them
// (cusco/CuBatchScoreMatrix.cu:236-396: CalcSM_Init_SMEMUnroll2x writes S,
// CalcSM_SSSS_CVS2S_SMEMUnroll2x_2 adds SS+CVS and writes mod,
// CalcSM_DDMS2S_SMEMUnroll2x_h2 adds DDMS to both when it is non-zero).
//
// Tiling: a workgroup of 256 lanes covers 4 query rows x 64 template
// columns; wave w owns query row w and lane l column l, so the query's
// distance list is uniform across a wavefront and each lane walks its own
// template list.  The 64 template lists (132-byte padded rows, conflict-free
// byte reads) and the 22x22 pair-score table with its halved copy are staged
// in LDS; a wavefront's query list is read by scalar loads, as pair-table
// byte offsets.  The per-cell DDMS dynamic programme (the dominant cost,
// ~2200 cell updates per score cell for full lists) runs from LDS, scalar
// registers and vector registers.
#include <float.h>

#include <hip/hip_runtime.h>

#include "cth_batch.h"
#include "cth_dev.h"
#include "cth_math.h"

namespace {

constexpr int kRows = 4;        // query rows per workgroup (one per wavefront)
constexpr int kCols = 64;       // template columns per workgroup (one per lane)
constexpr int kListPitch = 132; // LDS bytes per template list

constexpr int kQlEntries = 112;  // query record entries the DDMS scan reads (0..111)

struct ScoreLds {
  uint8_t tl[kCols * kListPitch];  // template lists, values scaled by 4 (table byte columns)
  int32_t ql[kRows][kQlEntries];   // each wavefront's query list as table byte offsets
  float dms[22 * 22];
  float dmsh[22 * 22];  // the pair scores halved (exact), right after dms
  float ss[467];
  float cvvals[CTH_CVMAXCARD];
};

__device__ __forceinline__ float ss_lookup(const float* ss, int row, int col) {
  if (row < col) { int x = row; row = col; col = x; }
  if (row) col += row + ((row * (row - 1)) >> 1);
  return ss[col + 2];
}

// A segment's coordinates as packed words: begin B = QB | PB << 11 and end
// E = QE | PE << 11.  Every index is below 256, so with kEndBias = 0x200 in
// each 11-bit half, (E + kEndBias) - B = 0x200 + (end - begin) per half with
// no borrow between the halves, and bit 9 of the half is set exactly when
// end >= begin.  Packed words stay below 2^22, so a float 2^23 + W holds W
// exactly and its bits are kMagic + W: the kernel computes a candidate's
// begin as one fma on floats (an instruction that dual-issues) and takes its
// bits, and every word it compares carries the same kMagic, which cancels in
// each difference of a test.  The register keeps each entry as
// nb = kEndBias - B and ce = E + kEndBias (with kMagic), so that both
// differences of a test are one add or subtract from a segment's B and E.
constexpr uint32_t kPack = 0x801u;        // one index step on both axes
constexpr float kPackF = 2049.0f;
constexpr uint32_t kEndBias = 0x00100200u;
constexpr uint32_t kMagic = 0x4b000000u;  // the bits of 2^23
// entries that intersect nothing: begins 255, ends 0
constexpr uint32_t kNoBeginN = kEndBias - (kMagic + 0x0007f8ffu), kNoEnd = kMagic + kEndBias;

// whether segment s (begin sb, end se) intersects register entry c (nb, ce):
// on the query or the template axis the byte-wise max of the begins is <=
// the min of the ends (CuBatchScoreMatrix_ddms2s.cuh:141-153), i.e. (begin
// <= end within each segment) ce >= sb and se >= cb
__device__ __forceinline__ bool meets(uint32_t sb, uint32_t se, uint32_t nb, uint32_t ce) {
  // the three-input and as one instruction (left to itself the compiler
  // shares (ce - sb) & kEndBias between tests at the cost of an extra and)
  return __builtin_amdgcn_bitop3_b32(ce - sb, se + nb, kEndBias, 0x80) != 0u;
}

// RegisterMaxDDMScore (CuBatchScoreMatrix_ddms2s.cuh:168-261): keep the
// three best segment scores m0 >= m1 >= m2 whose coordinates do not
// intersect.  Called only for s >= max(m2, FLT_MIN).  The reference's three
// cases, as selects on the two intersection tests with the current m0 and m1:
//  s >= m0: s becomes m0; the old m0 stays (as m1) unless it intersects s,
//           the old m1 stays (next) unless it intersects s, the old m2 drops;
//  m1 <= s < m0, s not intersecting m0: s becomes m1, the old m1 becomes m2
//           unless it intersects s;
//  m2 <= s < m1, s intersecting neither m0 nor m1: s becomes m2.
// Only m0's and m1's coordinates are ever tested, so m2's are not kept.  An
// entry whose value is 0 may keep stale coordinates: every intersection test
// with it only decides between two zeros, and no segment intersects the
// initial entries (kNoBeginN, kNoEnd).  lt0: s < m0; i0, i1: whether s
// intersects the current m0, m1.  The register changes only if
// register_changes(); register_max is called only then.
__device__ __forceinline__ bool register_changes(bool lt0, float m1, float s, bool i0, bool i1) {
  return !lt0 | (!i0 & ((m1 <= s) | !i1));
}
__device__ __forceinline__ void register_max(float& m0, float& m1, float& m2, uint32_t& nb0, uint32_t& e0,
                                             uint32_t& nb1, uint32_t& e1, float s, bool lt0, uint32_t sb,
                                             uint32_t se, bool i0, bool i1) {
  // one block per case, so that a wavefront runs only the cases its lanes
  // have (most often only the first: the best segment growing)
  if (!lt0) {
    // s becomes m0
    const float ym1 = i1 ? 0.0f : m1;
    m2 = i0 ? 0.0f : ym1;
    m1 = i0 ? ym1 : m0;
    nb1 = i0 ? nb1 : nb0;
    e1 = i0 ? e1 : e0;
    m0 = s;
    nb0 = kEndBias - sb;
    e0 = se + kEndBias;
    // new values rather than the candidate's registers, so that the join
    // keeps m0 and its coordinates where they live without copies
    asm("" : "+v"(m0), "+v"(nb0), "+v"(e0));
  } else if (m1 <= s) {
    // s becomes m1 (s does not intersect m0 when the register changes)
    m2 = i1 ? 0.0f : m1;
    m1 = s;
    nb1 = kEndBias - sb;
    e1 = se + kEndBias;
  } else {
    // s becomes m2 (s intersects neither m0 nor m1 when the register changes)
    m2 = s;
  }
}

// CalcDDMS2ScoreSMEMh2<2,true,2> (cusco/CuBatchScoreMatrix_ddms2s_h2.cuh:435-500):
// upper diagonals j = len2>>2 .. 0 step 2, lower diagonals i = len1>>2 .. 1
// step 2; per diagonal a Kadane scan in steps of 4 cells where cells 2 and 4
// are exact pair scores and cells 1 and 3 their means; per 4-cell step the
// best qualifying segment (length >= DSTSEGM) goes to the top-3 register.
//
// The per-cell update in arithmetic that issues on the VALU's second
// (add/mul) slot: compares, selects and min/max issue one per cycle, plain
// FP32 adds, multiplies and fmas (source modifiers and clamp included) two.
// The running sum is kept doubled, u = 2 segsco (exact), so that the next
// sum is one fma.  Per cell, with x the cell's pair score (or, for the
// interpolated cells, 0.5 (a + b): the reference's 0.5f * (a + b), exact),
// dd = the segment's begin minus the cell index (an integer value, <= 0) and
// big = 2^126:
//   nz   = clamp(u * big)          1 when the previous sum is > 0, else 0
//                                  (positive sums of pair scores are far
//                                  above 2^-126)
//   dd   = fma(nz, dd, -nz)        dd - 1, or 0 where the segment restarts
//   t    = fma(0.5, u, x)          segsco + x with the reference's one
//                                  rounding (0.5 u is exact)
//   u    = t + |t|                 2 max(t, 0), exact
//   cand = u * clamp(-dd - 4)      u if the segment has at least DSTSEGM = 6
//                                  cells (dd <= -5), else 0
// and the step's best is the max of its four candidates: the reference takes
// a cell when its sum exceeds the best so far, i.e. the first cell reaching
// the max.  The clamp at 0 differs from the reference's select only in the
// sign of a zero sum, and a zero is never registered.  The top-3 register
// holds doubled sums too (doubling commutes with every comparison and with
// the rounding of the final sum, which is halved exactly).  It is entered
// only when the step's best can displace its smallest entry:
// scom >= max(m2, FLT_MIN) (positive sums are far above FLT_MIN, so this is
// scom > 0 && scom >= m2; m2 is finite, so the max is the median of m2,
// FLT_MIN and FLT_MAX, one instruction).
#define CTH_DDMS_CELL(X, COUT)                                                \
  {                                                                           \
    const float nz = __builtin_amdgcn_fmed3f(u * big, 0.0f, 1.0f);            \
    dd = fmaf(nz, dd, -nz);                                                   \
    const float t = fmaf(0.5f, u, X);                                         \
    u = t + fabsf(t);                                                         \
    COUT = u * __builtin_amdgcn_fmed3f(-dd - 4.0f, 0.0f, 1.0f);               \
  }
// one 4-cell step, cells K0+1..K0+4 relative to the base ib, with pair
// scores A2, A4 (cells K0+2 and K0+4) and the halves H0, H2, H4 of those of
// cells K0 (the previous step's last cell), K0+2 and K0+4, and FB the float
// 2^23 + the packed coordinates of cell K0+1 on both axes: the reference's
// interpolated sum 0.5f * (a + b) is h(a) + h(b), halving being exact and
// commuting with the rounding of the sum.  The begin and end of the step's
// best are recovered only when it can be registered.  Three levels: the
// step's best is a candidate when it reaches max(m2, FLT_MIN) (m2 is zero
// most of the time: a growing best segment drops it); a candidate whose
// segment already meets m0 with its end at the step's first cell and that
// scores below m0 is dropped after one intersection test, before its end cell
// is selected (most candidates are of this kind: segments of neighbouring
// diagonals through the same region); the rest take the full test, and the
// update runs for a wavefront only when one of its lanes really changes its
// register.  The empty asm statements keep the candidate's packed words in
// registers: without them the compiler folds their sums into each test as
// scalar-operand adds, which issue one per cycle.
#define CTH_DDMS_4CELL(H0, A2, H2, A4, H4, FB)                                \
  {                                                                           \
    float cd1, cd2, cd3, cd4;                                                 \
    CTH_DDMS_CELL(H0 + H2, cd1)                                               \
    const float dd1 = dd;                                                     \
    CTH_DDMS_CELL(A2, cd2)                                                    \
    CTH_DDMS_CELL(H2 + H4, cd3)                                               \
    CTH_DDMS_CELL(A4, cd4)                                                    \
    const float scom = fmaxf(fmaxf(cd1, cd2), fmaxf(cd3, cd4));               \
    if (scom >= m2f) {                                                        \
      /* the step's first cell on both axes, packed (with kMagic) */          \
      const uint32_t base = __float_as_uint(FB);                              \
      /* begin: the candidate's segment has DSTSEGM or more cells at a cell   \
         of this step, so it runs at the step's first cell, whose dd gives    \
         its begin: 2^23 + first cell + dd1 (1 + 2^11), exact in float */     \
      uint32_t sb = __float_as_uint(fmaf(dd1, packf, FB));                    \
      asm("" : "+v"(sb));                                                     \
      /* the end lies in cells 1..4 and an intersection, once there, stays as \
         the end grows: a segment that meets m0 ending at cell 1 meets it at  \
         its end, and below m0 it cannot change the register */               \
      const bool lt0 = scom < m0;                                             \
      if (!(meets(sb, base, nb0, ce0) && lt0)) {                              \
        /* end: the first of cells 1..4 reaching scom, the cell k + 1 after   \
           the first, as 2^23 + first cell + k (1 + 2^11); n_k = 1 where cell \
           k is below scom and 0 at scom (scom >= cd_k, and sums of the pair  \
           scores are multiples of 2^-99, so a difference is 0 or scaled by   \
           2^100 at least 2), so k = n1 (1 + n2 (1 + n3)):                    \
           fmas that dual-issue, where compares and selects do not */         \
        const float sbig = scom * big2;                                       \
        const float n1 = __builtin_amdgcn_fmed3f(fmaf(cd1, -big2, sbig), 0.0f, 1.0f); \
        const float n2 = __builtin_amdgcn_fmed3f(fmaf(cd2, -big2, sbig), 0.0f, 1.0f); \
        const float n3 = __builtin_amdgcn_fmed3f(fmaf(cd3, -big2, sbig), 0.0f, 1.0f); \
        uint32_t se = __float_as_uint(fmaf(n1, fmaf(n2, fmaf(n3, packf, packf), packf), FB)); \
        asm("" : "+v"(se));                                                   \
        const bool i0 = meets(sb, se, nb0, ce0);                              \
        const bool i1 = meets(sb, se, nb1, ce1);                              \
        if (register_changes(lt0, m1, scom, i0, i1)) {                        \
          register_max(m0, m1, m2, nb0, ce0, nb1, ce1, scom, lt0, sb, se, i0, \
                       i1);                                                   \
          m2f = __builtin_amdgcn_fmed3f(m2, FLT_MIN, FLT_MAX);                \
        }                                                                     \
      }                                                                       \
    }                                                                         \
  }

// The scan reads everything from the workgroup's LDS block by byte offsets
// held in vector registers, so that each table address is one vector add
// (which dual-issues) and the rest is the instruction's immediate offset.
// kDms, kDmsh: the pair-score table and its halves in the block
constexpr int kDms = (int)offsetof(ScoreLds, dms), kDmsh = (int)offsetof(ScoreLds, dmsh);
__device__ __forceinline__ int lds_i32(const char* lds, int off) { return *reinterpret_cast<const int32_t*>(lds + off); }
__device__ __forceinline__ int lds_u8(const char* lds, int off) { return *reinterpret_cast<const uint8_t*>(lds + off); }
// the pair score of list values whose table byte offsets are qb (query,
// d1*88) and tv (template, d2*4): dmsT[d1][d2], and its half
__device__ __forceinline__ float pair(const char* lds, int qb, int tv) {
  return *reinterpret_cast<const float*>(lds + kDms + qb + tv);
}
__device__ __forceinline__ float pairh(const char* lds, int qb, int tv) {
  return *reinterpret_cast<const float*>(lds + kDmsh + qb + tv);
}

// lds: the workgroup's block; qr: byte offset of the wavefront's query list
// (int32 table byte offsets, entry 0 = the list length); tr: byte offset of
// the lane's template list (byte 0 = length, then values times 4)
__device__ __forceinline__ float ddms_cell(const char* lds, int qr, int tr) {
  const int len1 = lds_i32(lds, qr), len2 = lds_u8(lds, tr);
  if (len1 < CTH_DSTSEGM || len2 < CTH_DSTSEGM) return 0.0f;
  float m0 = 0.0f, m1 = 0.0f, m2 = 0.0f, m2f = FLT_MIN;
  uint32_t nb0 = kNoBeginN, ce0 = kNoEnd, nb1 = kNoBeginN, ce1 = kNoEnd;
  // 2^126, 1 + 2^11 and 2^100 in vector registers: as scalar operands they
  // would keep the multiplies and fmas off the second issue slot
  float big, packf, big2;
  asm("v_mov_b32 %0, 0x7e800000" : "=v"(big));
  asm("v_mov_b32 %0, 0x71800000" : "=v"(big2));  // 2^100
  asm("v_mov_b32 %0, %1" : "=v"(packf) : "i"(__builtin_bit_cast(uint32_t, kPackF)));
  // upper part: diagonal j pairs query index i with template index j + i
  for (int j = len2 >> 2; 0 <= j; j -= 2) {
    const int iend = min(len1, len2 - j);
    // byte offsets of query entry 1 and template entry j + 1, advanced by
    // one iteration (8 entries) at a time in vector registers
    int qo = qr + 4, to = tr + 1 + j;
    const float s0 = pair(lds, lds_i32(lds, qo), lds_u8(lds, to));
    float u = s0 + fabsf(s0);  // 2 max(s0, 0)
    float dd = 0.0f;
    // 2^23 + the packed coordinates of cell 1 (query 1, template j + 1)
    float fb = __uint_as_float(kMagic + kPack + ((uint32_t)j << 11));
    float h0 = pairh(lds, lds_i32(lds, qo), lds_u8(lds, to));
    float s2 = pair(lds, lds_i32(lds, qo + 8), lds_u8(lds, to + 2)),
          h2 = pairh(lds, lds_i32(lds, qo + 8), lds_u8(lds, to + 2));
    float s4 = pair(lds, lds_i32(lds, qo + 16), lds_u8(lds, to + 4)),
          h4 = pairh(lds, lds_i32(lds, qo + 16), lds_u8(lds, to + 4));
    for (int ib = 0; ib < iend; ib += 8) {
      // two steps per iteration; the pair scores of the step after are loaded
      // one step ahead (lists are padded)
      const int q6 = lds_i32(lds, qo + 24), q8 = lds_i32(lds, qo + 32);
      const int t6 = lds_u8(lds, to + 6), t8 = lds_u8(lds, to + 8);
      const float n2 = pair(lds, q6, t6), nh2 = pairh(lds, q6, t6);
      const float n4 = pair(lds, q8, t8), nh4 = pairh(lds, q8, t8);
      CTH_DDMS_4CELL(h0, s2, h2, s4, h4, fb)
      const int q10 = lds_i32(lds, qo + 40), q12 = lds_i32(lds, qo + 48);
      const int t10 = lds_u8(lds, to + 10), t12 = lds_u8(lds, to + 12);
      const float p2 = pair(lds, q10, t10), ph2 = pairh(lds, q10, t10);
      const float p4 = pair(lds, q12, t12), ph4 = pairh(lds, q12, t12);
      if (ib + 4 < iend) CTH_DDMS_4CELL(h4, n2, nh2, n4, nh4, fb + 4.0f * kPackF)
      fb += 8.0f * kPackF;
      qo += 32;
      to += 8;
      asm("" : "+v"(qo), "+v"(to));
      h0 = nh4;
      s2 = p2;
      h2 = ph2;
      s4 = p4;
      h4 = ph4;
    }
  }
  // lower part: diagonal i pairs query index i + j with template index j
  for (int i = len1 >> 2; 0 < i; i -= 2) {
    const int jend = min(len2, len1 - i);
    int qo = qr + 4 * (1 + i), to = tr + 1;
    const float s0 = pair(lds, lds_i32(lds, qo), lds_u8(lds, to));
    float u = s0 + fabsf(s0);  // 2 max(s0, 0)
    float dd = 0.0f;
    // 2^23 + the packed coordinates of cell 1 (query i + 1, template 1)
    float fb = __uint_as_float(kMagic + kPack + (uint32_t)i);
    float h0 = pairh(lds, lds_i32(lds, qo), lds_u8(lds, to));
    float s2 = pair(lds, lds_i32(lds, qo + 8), lds_u8(lds, to + 2)),
          h2 = pairh(lds, lds_i32(lds, qo + 8), lds_u8(lds, to + 2));
    float s4 = pair(lds, lds_i32(lds, qo + 16), lds_u8(lds, to + 4)),
          h4 = pairh(lds, lds_i32(lds, qo + 16), lds_u8(lds, to + 4));
    for (int ib = 0; ib < jend; ib += 8) {
      const int q6 = lds_i32(lds, qo + 24), q8 = lds_i32(lds, qo + 32);
      const int t6 = lds_u8(lds, to + 6), t8 = lds_u8(lds, to + 8);
      const float n2 = pair(lds, q6, t6), nh2 = pairh(lds, q6, t6);
      const float n4 = pair(lds, q8, t8), nh4 = pairh(lds, q8, t8);
      CTH_DDMS_4CELL(h0, s2, h2, s4, h4, fb)
      const int q10 = lds_i32(lds, qo + 40), q12 = lds_i32(lds, qo + 48);
      const int t10 = lds_u8(lds, to + 10), t12 = lds_u8(lds, to + 12);
      const float p2 = pair(lds, q10, t10), ph2 = pairh(lds, q10, t10);
      const float p4 = pair(lds, q12, t12), ph4 = pairh(lds, q12, t12);
      if (ib + 4 < jend) CTH_DDMS_4CELL(h4, n2, nh2, n4, nh4, fb + 4.0f * kPackF)
      fb += 8.0f * kPackF;
      qo += 32;
      to += 8;
      asm("" : "+v"(qo), "+v"(to));
      h0 = nh4;
      s2 = p2;
      h2 = ph2;
      s4 = p4;
      h4 = ph4;
    }
  }
  float s = 0.0f;
  s += m0;
  s += m1;
  s += m2;
  s *= 0.5f;  // the register holds doubled sums
  s = fmaf(CTH_DDMS_SLOPE * CTH_DDMS_SCALE, s, CTH_DDMS_ICEPT);
  return s * CTH_DDMSWGT;
}
#undef CTH_DDMS_CELL
#undef CTH_DDMS_4CELL

// kPreInit: the log-odds term is already in scores (the reduced-precision
// contraction of init_mfma.hip); otherwise it is computed here in FP32.
// kBatch: the row tiles of several queries in one launch, blockIdx.y indexing
// ytab (query of the batch and first row), each query's blocks at its offset.
template <bool kPreInit, bool kBatch = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_score(cth_dpset q, int qp, cth_dpset db, int64_t col0, int ncols,
                                               int64_t ldsc, const cth_tables* __restrict__ tabs,
                                               float* __restrict__ scores, float* __restrict__ mods,
                                               const int32_t* __restrict__ qoff,
                                               const cth_bytile* __restrict__ ytab = nullptr,
                                               const cth_bquery* __restrict__ bq = nullptr) {
  __shared__ ScoreLds L;
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  int row = blockIdx.y * kRows + w;
  if constexpr (kBatch) {
    const cth_bytile t = ytab[blockIdx.y];
    const cth_bquery Q = bq[t.b];
    qp = Q.qp;
    row = t.row0 + w;
    scores += Q.off;
    mods += Q.off;
  }
  const int lq = q.len[qp];
  const int64_t q0 = q.dist[qp];
  const int col = blockIdx.x * kCols + lane;

  // stage tables
  for (int k = tid; k < 22 * 22; k += 256) {
    L.dms[k] = tabs->dms[k];
    L.dmsh[k] = 0.5f * tabs->dms[k];
  }
  for (int k = tid; k < 467; k += 256) L.ss[k] = tabs->ss[k];
  for (int k = tid; k < tabs->cvcard; k += 256) L.cvvals[k] = tabs->cvvals[k];
  // stage 64 template lists (32 dwords each), values (<= 21) times 4: the
  // pair table's byte column; the length byte stays
  for (int k = tid; k < kCols * 32; k += 256) {
    const int c = k >> 5, d = k & 31;
    const int cc = blockIdx.x * kCols + c;
    uint32_t v = 0x15151515u;  // padding value 21
    if (cc < ncols) v = reinterpret_cast<const uint32_t*>(db.dd + (size_t)(col0 + cc) * CTH_DDSTRIDE)[d];
    if (cc >= ncols && d == 0) v = 0x15151500u;  // empty list
    v = d == 0 ? (v & 0xffu) | ((v & 0xffffff00u) << 2) : v << 2;
    *reinterpret_cast<uint32_t*>(&L.tl[c * kListPitch + d * 4]) = v;
  }
  // each wavefront's query list (table byte offsets; rows past the query
  // stage its last row, whose results are not stored)
  {
    const int rr = min(row, lq - 1);
    const int32_t* qsrc = qoff + (size_t)(q0 + rr) * CTH_DDSTRIDE;
    for (int k = lane; k < kQlEntries; k += 64) L.ql[w][k] = qsrc[k];
  }
  __syncthreads();
  if (row >= lq || col >= ncols) return;

  // DDMS first, while nothing else is live
  const float d = ddms_cell(reinterpret_cast<const char*>(&L), (int)offsetof(ScoreLds, ql) + w * kQlEntries * 4,
                            lane * kListPitch);

  const int64_t qi = q0 + row;
  const int64_t tj = col0 + col;
  const int tp = db.pro[tj];

  // Init: log-odds (CuBatchScoreMatrix_Init.cu:129-144,173-185)
  float s = 0.0f;
  if constexpr (kPreInit) {
    s = scores[row * ldsc + col];
  } else {
  #pragma unroll
    for (int a = 0; a < CTH_NAA; a++) {
      const float f1 = q.trg[a * q.ld + qi];
      const float f2 = q.bkg[qp * CTH_NAA + a];
      const float num = f1 * db.trg[a * db.ld + tj];
      const float den = 0.5f * (f2 + db.bkg[tp * CTH_NAA + a]);
      s = s + num / den;
    }
    s = cth_logf(s) - CTH_INITSHIFT;
  }

  // SS (CuBatchScoreMatrix_ssss.cuh:64-116)
  float sss = 0.0f;
#pragma unroll
  for (int i = 0; i < CTH_NSS; i++) {
    const float f1 = q.ss[i * q.ld + qi];
    const int r = i * 10 + (int)(f1 * 9.999f);
#pragma unroll
    for (int j = 0; j < CTH_NSS; j++) {
      const float f2 = db.ss[j * db.ld + tj];
      const int c = j * 10 + (int)(f2 * 9.999f);
      sss = fmaf(f1 * f2, ss_lookup(L.ss, r, c), sss);
    }
  }
  {
    const float f1 = q.ss[(CTH_NSS - 1) * q.ld + qi];
    if (-99.f < sss) {
      const float f2 = db.ss[(CTH_NSS - 1) * db.ld + tj];
      sss = fmaf((f1 + 0.1f) * (f2 + 0.1f), 0.1111111f, sss);
      sss -= CTH_SSSHIFT;
      sss *= CTH_SSSWGT;
    }
  }
  // CVS (CuBatchScoreMatrix_cvs2s.cuh:80-113)
  float cvs = 0.0f;
#pragma unroll
  for (int i = 0; i < CTH_NCV; i++) cvs = fmaf(q.cv[i * q.ld + qi], db.cv[i * db.ld + tj], cvs);
  const float kap = tabs->cvkappa0inv;
  cvs += kap;
  {
    const float a = (q.cvnorm2[qi] + 1.0f + kap) * (db.cvnorm2[tj] + 1.0f + kap);
    cvs = fmaf(-cvs, cvs, a);
    cvs = fmaf(tabs->cvpowernu0, cth_logf(cvs), tabs->cvcterm) - q.cvprior[qi] - db.cvprior[tj];
    const int card = tabs->cvcard;
    float v;
    if (cvs <= tabs->cvkeys[0]) v = L.cvvals[0];
    else if (tabs->cvkeys[card - 1] <= cvs) v = L.cvvals[card - 1];
    else v = L.cvvals[tabs->cvshift + (int)rintf(2.0f * cvs)];
    cvs = v * CTH_CVSWGT;
  }
  const float sc = sss + cvs;
  s = s + sc;
  float ms = fmaf(-CTH_CVSSHIFT, CTH_CVSWGT, sc);

  if (d != 0.0f) {
    s = s + d;
    ms = ms + fmaf(-CTH_DDMSSHIFT, CTH_DDMSWGT, d);
  }
  scores[row * ldsc + col] = s;
  mods[row * ldsc + col] = ms;
}

// The score stage where no cell can have a DDMS term (every query or every
// template list shorter than DSTSEGM: ddms_cell returns 0 and k_score adds
// nothing), with the same per-cell arithmetic as k_score's log-odds, SS and
// CVS terms.  Without the DDMS scan the cell is a few hundred instructions
// and k_score's 4-row tiles would re-read each template column's fields per
// 4 rows; here a workgroup covers kPlainRows query rows (every fourth row
// per wavefront) of 64 template columns, each lane holding its column's fields
// in registers.  The workgroup's query rows are staged in LDS, one 48-float
// record per row, and a wavefront reads its row by broadcast: its values
// are vector operands (scalar-register operands would keep the FP32
// arithmetic from dual-issuing) and a row costs no scalar loads or address
// arithmetic.
constexpr int kPlainRows = 64;
constexpr int kQrec = 48;  // floats per staged query row
constexpr int kQvals = 44; // 20 frequencies, 3 SS, 19 context-vector entries, norm^2, prior

// a factor of a log-odds numerator the quotient correction below admits:
// +0 or [2^-32, 2], compared on the bit pattern
__device__ __forceinline__ bool in_quot_range(float v) {
  const uint32_t b = __builtin_bit_cast(uint32_t, v);
  return b == 0u || b - 0x2F800000u <= 0x10800000u;
}

// value v (0..kQvals-1) of the staged record of query position qi
__device__ __forceinline__ float query_value(const cth_dpset& q, int v, int64_t qi) {
  if (v < CTH_NAA) return q.trg[v * q.ld + qi];
  if (v < CTH_NAA + CTH_NSS) return q.ss[(v - CTH_NAA) * q.ld + qi];
  if (v < CTH_NAA + CTH_NSS + CTH_NCV) return q.cv[(v - CTH_NAA - CTH_NSS) * q.ld + qi];
  return v == kQvals - 2 ? q.cvnorm2[qi] : q.cvprior[qi];
}

template <bool kBatch = false>
// Five wavefronts per SIMD (at most 96 VGPRs): the template frequencies are
// read from LDS per row, so that the column's denominators and their
// reciprocals fit in registers beside the row's values.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_score_plain(
    cth_dpset q, int qp, cth_dpset db, int64_t col0, int ncols, int64_t ldsc, const cth_tables* __restrict__ tabs,
    float* __restrict__ scores, float* __restrict__ mods, const cth_bytile* __restrict__ ytab = nullptr,
    const cth_bquery* __restrict__ bq = nullptr) {
  // the SS table square (30 x 30, symmetric: entry r*30+c is ss_lookup(r, c)),
  // so that a lookup is one index add
  __shared__ float Lsq[30 * 30];
  __shared__ float Lcv[CTH_CVMAXCARD];
  __shared__ float4 Ltcv[5][kCols];
  __shared__ float4 Ltf[5][kCols];
  __shared__ __align__(16) float Lq[kPlainRows][kQrec];
  __shared__ int Lbad[kPlainRows];  // a row has a frequency outside in_quot_range
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  int row0 = blockIdx.y * kPlainRows;
  if constexpr (kBatch) {
    const cth_bytile t = ytab[blockIdx.y];
    const cth_bquery Q = bq[t.b];
    qp = Q.qp;
    row0 = t.row0;
    scores += Q.off;
    mods += Q.off;
  }
  const int lq = q.len[qp];
  const int64_t q0 = q.dist[qp];
  const int col = blockIdx.x * kCols + lane;
  const bool colin = col < ncols;
  for (int k = tid; k < 30 * 30; k += 256) Lsq[k] = ss_lookup(tabs->ss, k / 30, k % 30);
  for (int k = tid; k < tabs->cvcard; k += 256) Lcv[k] = tabs->cvvals[k];
  if (tid < kPlainRows) Lbad[tid] = 0;
  __syncthreads();

  // the query rows, read along the rows (rows past the query's end as zeros)
  for (int k = tid; k < kPlainRows * kQvals; k += 256) {
    const int r = k % kPlainRows, v = k / kPlainRows;
    const float x = row0 + r < lq ? query_value(q, v, q0 + row0 + r) : 0.0f;
    Lq[r][v] = x;
    if (v < CTH_NAA && !in_quot_range(x)) Lbad[r] = 1;
  }

  // this lane's template column (lanes past the last column read column 0
  // and leave after the staging): frequencies and the log-odds denominators
  // (CuBatchScoreMatrix_Init.cu:129-144,173-185: 0.5 (bq(a) + bt(a)), per
  // query and template profile), SS probabilities and their table columns,
  // context vector, its prior and squared norm.  The frequencies and the
  // context vectors go to LDS, staged by wavefront 0 for the workgroup's 64
  // columns, four values a lane per 16-byte slot.
  const int64_t tj = col0 + (colin ? col : 0);
  const int tp = db.pro[tj];
  float tf[CTH_NAA], den[CTH_NAA];
#pragma unroll
  for (int a = 0; a < CTH_NAA; a++) {
    tf[a] = db.trg[a * db.ld + tj];
    den[a] = 0.5f * (q.bkg[qp * CTH_NAA + a] + db.bkg[tp * CTH_NAA + a]);
  }
  if (w == 0) {
#pragma unroll
    for (int i = 0; i < 5; i++) {
      float4 v;
      v.x = db.cv[(4 * i) * db.ld + tj];
      v.y = db.cv[(4 * i + 1) * db.ld + tj];
      v.z = db.cv[(4 * i + 2) * db.ld + tj];
      v.w = 4 * i + 3 < CTH_NCV ? db.cv[(4 * i + 3) * db.ld + tj] : 0.0f;
      Ltcv[i][lane] = v;
      Ltf[i][lane] = make_float4(tf[4 * i], tf[4 * i + 1], tf[4 * i + 2], tf[4 * i + 3]);
    }
  }
  float ts[CTH_NSS];
  int tc[CTH_NSS];
#pragma unroll
  for (int j = 0; j < CTH_NSS; j++) {
    ts[j] = db.ss[j * db.ld + tj];
    tc[j] = j * 10 + (int)(ts[j] * 9.999f);
  }
  const float kap = tabs->cvkappa0inv;
  const float tn = db.cvnorm2[tj] + 1.0f + kap;
  const float tprior = db.cvprior[tj];
  // The log-odds quotients num / den[a] by Markstein's correction: with
  // y = RN(1 / den) (an IEEE division, once per lane), q0 = RN(num y),
  // r = fma(-den, q0, num) (exact) and q = RN(q0 + r y) equal RN(num / den)
  // when no step underflows or overflows (P. Markstein, IBM J. Res. Dev.
  // 34:111-119, 1990): here when den lies in [2^-30, 2^30] and both factors
  // of num are 0 or in [2^-32, 2].  A wavefront whose columns or query row
  // have an operand outside those ranges divides.
  float y[CTH_NAA];
  bool lane_ok = true;
#pragma unroll
  for (int a = 0; a < CTH_NAA; a++) {
    y[a] = 1.0f / den[a];
    lane_ok &= in_quot_range(tf[a]) && __builtin_bit_cast(uint32_t, den[a]) - 0x30800000u <= 0x1E000000u;
  }
  __syncthreads();
  if (!colin) return;
  const bool cols_ok = __builtin_amdgcn_ballot_w64(!lane_ok) == 0;

  for (int k = 0; k < kPlainRows / 4; k++) {
    const int rl = k * 4 + w;  // rows interleaved over the wavefronts: partial tiles stay balanced
    const int row = row0 + rl;
    if (row >= lq) break;
    const float* const R = Lq[rl];
    // log-odds
    float tfr[CTH_NAA];
#pragma unroll
    for (int i4 = 0; i4 < 5; i4++) {
      const float4 v = Ltf[i4][lane];
      tfr[4 * i4] = v.x;
      tfr[4 * i4 + 1] = v.y;
      tfr[4 * i4 + 2] = v.z;
      tfr[4 * i4 + 3] = v.w;
    }
    float s = 0.0f;
    if (cols_ok && __builtin_amdgcn_readfirstlane(Lbad[rl]) == 0) {
#pragma unroll
      for (int a = 0; a < CTH_NAA; a++) {
        const float num = R[a] * tfr[a];
        const float q0 = num * y[a];
        s = s + fmaf(fmaf(-den[a], q0, num), y[a], q0);
      }
    } else {
#pragma unroll
      for (int a = 0; a < CTH_NAA; a++) s = s + R[a] * tfr[a] / den[a];
    }
    s = cth_logf(s) - CTH_INITSHIFT;
    // SS (CuBatchScoreMatrix_ssss.cuh:64-116)
    float sss = 0.0f;
#pragma unroll
    for (int i = 0; i < CTH_NSS; i++) {
      const float f1 = R[CTH_NAA + i];
      const int r30 = (i * 10 + (int)(f1 * 9.999f)) * 30;
#pragma unroll
      for (int j = 0; j < CTH_NSS; j++) sss = fmaf(f1 * ts[j], Lsq[r30 + tc[j]], sss);
    }
    {
      const float f1 = R[CTH_NAA + CTH_NSS - 1];
      if (-99.f < sss) {
        sss = fmaf((f1 + 0.1f) * (ts[CTH_NSS - 1] + 0.1f), 0.1111111f, sss);
        sss -= CTH_SSSHIFT;
        sss *= CTH_SSSWGT;
      }
    }
    // CVS (CuBatchScoreMatrix_cvs2s.cuh:80-113)
    float cvs = 0.0f;
#pragma unroll
    for (int i4 = 0; i4 < 5; i4++) {
      const float4 v = Ltcv[i4][lane];
      const float* const c = R + CTH_NAA + CTH_NSS + 4 * i4;
      cvs = fmaf(c[0], v.x, cvs);
      cvs = fmaf(c[1], v.y, cvs);
      cvs = fmaf(c[2], v.z, cvs);
      if (4 * i4 + 3 < CTH_NCV) cvs = fmaf(c[3], v.w, cvs);
    }
    cvs += kap;
    {
      const float a = (R[kQvals - 2] + 1.0f + kap) * tn;
      cvs = fmaf(-cvs, cvs, a);
      cvs = fmaf(tabs->cvpowernu0, cth_logf(cvs), tabs->cvcterm) - R[kQvals - 1] - tprior;
      const int card = tabs->cvcard;
      float v;
      if (cvs <= tabs->cvkeys[0]) v = Lcv[0];
      else if (tabs->cvkeys[card - 1] <= cvs) v = Lcv[card - 1];
      else v = Lcv[tabs->cvshift + (int)rintf(2.0f * cvs)];
      cvs = v * CTH_CVSWGT;
    }
    const float sc = sss + cvs;
    s = s + sc;
    scores[row * ldsc + col] = s;
    mods[row * ldsc + col] = fmaf(-CTH_CVSSHIFT, CTH_CVSWGT, sc);
  }
}

}  // namespace

// Launch the score stage for query profile qp against template positions
// [col0, col0+ncols) of db; outputs are [Lq][ldsc].
// contraction 0: the FP32 log-odds term of the reference; 1 / 2: the term
// from the BF16 / FP16 matrix-core product (init_mfma.hip) first.  ddms 0:
// no cell can have a DDMS term (the query's or every template's lists are
// shorter than DSTSEGM), so the FP32 stage runs k_score_plain.
extern "C" int cth_launch_init_mfma(const cth_dpset* q, int qp, const cth_dpset* db, int64_t col0, int ncols,
                                    int64_t ldsc, int lq, int prec, float* out, hipStream_t st);
extern "C" int cth_launch_score(const cth_dpset* q, const int32_t* qoff, int qp, const cth_dpset* db, int64_t col0,
                                int ncols, int64_t ldsc, const cth_tables* dtabs, int lq, float* scores, float* mods,
                                int contraction, int ddms, hipStream_t st) {
  if (ncols <= 0 || lq <= 0) return 0;
  if (!qoff) { cth_set_error("score: query distance offsets not built"); return -1; }
  if (ldsc < ncols) { cth_set_error("score: ldsc %lld < ncols %d", (long long)ldsc, ncols); return -1; }
  if (!ddms && !contraction) {
    dim3 pgrid((ncols + kCols - 1) / kCols, (lq + kPlainRows - 1) / kPlainRows);
    hipLaunchKernelGGL(k_score_plain<false>, pgrid, dim3(256), 0, st, *q, qp, *db, col0, ncols, ldsc, dtabs, scores,
                       mods);
    CTH_HIP_CHECK(hipGetLastError());
    return 0;
  }
  dim3 grid((ncols + kCols - 1) / kCols, (lq + kRows - 1) / kRows);
  if (contraction) {
    if (cth_launch_init_mfma(q, qp, db, col0, ncols, ldsc, lq, contraction, scores, st)) return -1;
    hipLaunchKernelGGL(k_score<true>, grid, dim3(256), 0, st, *q, qp, *db, col0, ncols, ldsc, dtabs, scores, mods, qoff);
  } else {
    hipLaunchKernelGGL(k_score<false>, grid, dim3(256), 0, st, *q, qp, *db, col0, ncols, ldsc, dtabs, scores, mods,
                       qoff);
  }
  CTH_HIP_CHECK(hipGetLastError());
  return 0;
}

// The score stage of a batch of queries in one launch: nytiles row tiles
// (ytab: query of the batch and first row, cth_score_rows_per_tile(ddms)
// rows each) against the whole resident database (ncols positions from
// column 0); each query's scores and modular scores at its offset
// (bq[].off, row stride ldsc); ddms 0 as in cth_launch_score.
extern "C" int cth_launch_score_batch(const cth_dpset* q, const int32_t* qoff, const cth_dpset* db, int ncols,
                                      int64_t ldsc, const cth_tables* dtabs, const cth_bytile* ytab, int nytiles,
                                      const cth_bquery* bq, float* scores, float* mods, int ddms, hipStream_t st) {
  if (ncols <= 0 || nytiles <= 0) return 0;
  if (!qoff) { cth_set_error("score batch: query distance offsets not built"); return -1; }
  if (ldsc < ncols) { cth_set_error("score batch: ldsc %lld < ncols %d", (long long)ldsc, ncols); return -1; }
  if (nytiles > 65535) { cth_set_error("score batch: %d row tiles (at most 65535)", nytiles); return -1; }
  dim3 grid((ncols + kCols - 1) / kCols, nytiles);
  if (!ddms)
    hipLaunchKernelGGL((k_score_plain<true>), grid, dim3(256), 0, st, *q, 0, *db, (int64_t)0, ncols, ldsc, dtabs,
                       scores, mods, ytab, bq);
  else
    hipLaunchKernelGGL((k_score<false, true>), grid, dim3(256), 0, st, *q, 0, *db, (int64_t)0, ncols, ldsc, dtabs,
                       scores, mods, qoff, ytab, bq);
  CTH_HIP_CHECK(hipGetLastError());
  return 0;
}

// rows of one tile of the score kernel (ddms 0: of k_score_plain)
extern "C" int cth_score_rows_per_tile(int ddms) { return ddms ? kRows : kPlainRows; }
