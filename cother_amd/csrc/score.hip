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
// byte reads), the 4 query lists and the 22x22 pair-score table are staged in
// LDS.  The per-cell DDMS dynamic programme (the dominant cost, ~2200 cell
// updates per score cell for full lists) runs entirely from LDS and
// registers.
#include <hip/hip_runtime.h>

#include "cth_dev.h"
#include "cth_math.h"

namespace {

constexpr int kRows = 4;        // query rows per workgroup (one per wavefront)
constexpr int kCols = 64;       // template columns per workgroup (one per lane)
constexpr int kListPitch = 132; // LDS bytes per template list

struct ScoreLds {
  uint8_t tl[kCols * kListPitch];
  int32_t qo[kRows][128];  // query lists: entry 0 = length, then 88 * value
  float dms[22 * 22];
  float ss[467];
  float cvvals[CTH_CVMAXCARD];
};

__device__ __forceinline__ float ss_lookup(const float* ss, int row, int col) {
  if (row < col) { int x = row; row = col; col = x; }
  if (row) col += row + ((row * (row - 1)) >> 1);
  return ss[col + 2];
}

__device__ __forceinline__ bool coords_intersect(uint32_t c1, uint32_t c2) {
  // byte-wise max/min of [QB QE PB PE] (CuBatchScoreMatrix_ddms2s.cuh:141-153)
  const uint32_t qb = max(c1 >> 24, c2 >> 24);
  const uint32_t qe = min((c1 >> 16) & 0xff, (c2 >> 16) & 0xff);
  const uint32_t pb = max((c1 >> 8) & 0xff, (c2 >> 8) & 0xff);
  const uint32_t pe = min(c1 & 0xff, c2 & 0xff);
  return qe >= qb || pe >= pb;
}

// RegisterMaxDDMScore (CuBatchScoreMatrix_ddms2s.cuh:168-261)
__device__ __forceinline__ void register_max(float& m0, float& m1, float& m2, uint32_t& c0, uint32_t& c1,
                                             uint32_t& c2, float s, uint32_t sc) {
  if (!s) return;
  if (m0 <= s) {
    m2 = m1; c2 = c1;
    m1 = m0; c1 = c0;
    m0 = s; c0 = sc;
    if (m1 != 0.0f && coords_intersect(c0, c1)) m1 = 0.0f;
    if (m2 != 0.0f && coords_intersect(c0, c2)) m2 = 0.0f;
    if (m1 == 0.0f) { m1 = m2; c1 = c2; m2 = 0.0f; }
  } else if (m1 <= s) {
    if (coords_intersect(sc, c0)) return;
    m2 = m1; c2 = c1;
    m1 = s; c1 = sc;
    if (coords_intersect(c1, c2)) m2 = 0.0f;
  } else if (m2 <= s) {
    if (coords_intersect(sc, c0) || coords_intersect(sc, c1)) return;
    m2 = s; c2 = sc;
  }
}

__device__ __forceinline__ uint32_t mkc(int qb, int qe, int pb, int pe) {
  return ((uint32_t)(qb & 0xff) << 24) | ((uint32_t)(qe & 0xff) << 16) | ((uint32_t)(pb & 0xff) << 8) |
         (uint32_t)(pe & 0xff);
}

// CalcDDMS2ScoreSMEMh2<2,true,2> (cusco/CuBatchScoreMatrix_ddms2s_h2.cuh:435-500):
// upper diagonals j = len2>>2 .. 0 step 2, lower diagonals i = len1>>2 .. 1
// step 2; per diagonal a Kadane scan in steps of 4 cells where cells 2 and 4
// are exact pair scores and cells 1 and 3 their means; per 4-cell step the
// best qualifying segment (length >= DSTSEGM) goes to the top-3 register.
//
// Branch-free form of the reference's per-cell update: a segment's begin
// index is kept on one axis only (the other is that index plus the diagonal
// offset) and relative to the current 4-cell step, so that the per-cell
// indices are the literals 1..4; "score > 0 and better than the step's best" reduces to
// "better than the step's best" because the best starts at 0; and the top-3
// register is entered only when the step's best can displace its smallest
// entry (the register stays ordered m0 >= m1 >= m2 >= 0).  The clamp at 0
// is a max: it differs from the reference's select only in the sign of a zero
// sum, and a zero is never registered.  The query list is
// the same for the whole wavefront, so its values are read into scalar
// registers.
#define CTH_DDMS_UPD(K, SK, IDX)                                               \
  {                                                                           \
    segr = (segsco != 0.0f) ? segr : (K);                                     \
    segsco = fmaxf(segsco + (SK), 0.0f);                                      \
    const bool take = (scom < segsco) & (segr <= (K) + 1 - CTH_DSTSEGM);      \
    scom = take ? segsco : scom;                                              \
    bb = take ? segr : bb;                                                    \
    be = take ? (K) : be;                                                     \
  }

// the pair score of list values whose table byte offsets are qb (query,
// d1*88) and tv (template value d2): dmsT[d1][d2]
__device__ __forceinline__ float pair(const float* __restrict__ dmsT, int qb, int tv) {
  return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(dmsT) + qb + (tv << 2));
}

// q: the query list as table byte offsets (entry k = 88 * value k, entry 0 =
// the list length); t: the template list (byte 0 = length, then values)
__device__ float ddms_cell(const float* __restrict__ dmsT, const int32_t* __restrict__ q,
                           const uint8_t* __restrict__ t) {
  const int len1 = q[0], len2 = t[0];
  if (len1 < CTH_DSTSEGM || len2 < CTH_DSTSEGM) return 0.0f;
  float m0 = 0.0f, m1 = 0.0f, m2 = 0.0f;
  uint32_t c0 = 0xffffffffu, c1 = 0xffffffffu, c2 = 0xffffffffu;
  // upper part: diagonal j pairs query index i with template index j + i;
  // i + segr is the query index where the running segment begins
  for (int j = len2 >> 2; 0 <= j; j -= 2) {
    const uint8_t* tj = t + 1 + j;
    int segr = 4;  // relative to the first step's index 0 after the step's "segr -= 4"
    float s0 = pair(dmsT, q[1], tj[0]);
    float segsco = s0 < 0.0f ? 0.0f : s0;
    const int iend = min(len1, len2 - j);
    float s2 = pair(dmsT, q[1 + 2], tj[2]);
    float s4 = pair(dmsT, q[1 + 4], tj[4]);
    for (int i = 0; i < iend; i += 4) {
      // the next step's pair scores, loaded one step ahead (lists are padded)
      const float n2 = pair(dmsT, q[1 + i + 6], tj[i + 6]);
      const float n4 = pair(dmsT, q[1 + i + 8], tj[i + 8]);
      const float s3 = 0.5f * (s2 + s4);
      const float s1 = 0.5f * (s0 + s2);
      float scom = 0.0f;
      int bb = 0, be = 0;
      segr -= 4;
      CTH_DDMS_UPD(1, s1, i)
      CTH_DDMS_UPD(2, s2, i)
      CTH_DDMS_UPD(3, s3, i)
      CTH_DDMS_UPD(4, s4, i)
      if (scom != 0.0f && m2 <= scom)
        register_max(m0, m1, m2, c0, c1, c2, scom, mkc(bb + i, be + i, bb + i + j, be + i + j));
      s0 = s4;
      s2 = n2;
      s4 = n4;
    }
  }
  // lower part: diagonal i pairs query index i + j with template index j;
  // j + segr is the template index where the running segment begins
  for (int i = len1 >> 2; 0 < i; i -= 2) {
    const int32_t* qi = q + 1 + i;
    int segr = 4;  // relative to the first step's index 0 after the step's "segr -= 4"
    float s0 = pair(dmsT, qi[0], t[1]);
    float segsco = s0 < 0.0f ? 0.0f : s0;
    const int jend = min(len2, len1 - i);
    float s2 = pair(dmsT, qi[2], t[1 + 2]);
    float s4 = pair(dmsT, qi[4], t[1 + 4]);
    for (int j = 0; j < jend; j += 4) {
      const float n2 = pair(dmsT, qi[j + 6], t[1 + j + 6]);
      const float n4 = pair(dmsT, qi[j + 8], t[1 + j + 8]);
      const float s3 = 0.5f * (s2 + s4);
      const float s1 = 0.5f * (s0 + s2);
      float scom = 0.0f;
      int bb = 0, be = 0;
      segr -= 4;
      CTH_DDMS_UPD(1, s1, j)
      CTH_DDMS_UPD(2, s2, j)
      CTH_DDMS_UPD(3, s3, j)
      CTH_DDMS_UPD(4, s4, j)
      if (scom != 0.0f && m2 <= scom)
        register_max(m0, m1, m2, c0, c1, c2, scom, mkc(bb + j + i, be + j + i, bb + j, be + j));
      s0 = s4;
      s2 = n2;
      s4 = n4;
    }
  }
  float s = 0.0f;
  s += m0;
  s += m1;
  s += m2;
  s = fmaf(CTH_DDMS_SLOPE * CTH_DDMS_SCALE, s, CTH_DDMS_ICEPT);
  return s * CTH_DDMSWGT;
}
#undef CTH_DDMS_UPD

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_score(cth_dpset q, int qp, cth_dpset db, int64_t col0, int ncols,
                                               int64_t ldsc, const cth_tables* __restrict__ tabs,
                                               float* __restrict__ scores, float* __restrict__ mods) {
  __shared__ ScoreLds L;
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int lq = q.len[qp];
  const int64_t q0 = q.dist[qp];
  const int row = blockIdx.y * kRows + w;
  const int col = blockIdx.x * kCols + lane;

  // stage tables
  for (int k = tid; k < 22 * 22; k += 256) L.dms[k] = tabs->dms[k];
  for (int k = tid; k < 467; k += 256) L.ss[k] = tabs->ss[k];
  for (int k = tid; k < tabs->cvcard; k += 256) L.cvvals[k] = tabs->cvvals[k];
  // stage 64 template lists (32 dwords each) and 4 query lists
  for (int k = tid; k < kCols * 32; k += 256) {
    const int c = k >> 5, d = k & 31;
    const int cc = blockIdx.x * kCols + c;
    uint32_t v = 0x15151515u;  // padding value 21
    if (cc < ncols) v = reinterpret_cast<const uint32_t*>(db.dd + (size_t)(col0 + cc) * CTH_DDSTRIDE)[d];
    if (cc >= ncols && d == 0) v = 0x15151500u;  // empty list
    *reinterpret_cast<uint32_t*>(&L.tl[c * kListPitch + d * 4]) = v;
  }
  for (int k = tid; k < kRows * 128; k += 256) {
    const int r = k >> 7, d = k & 127;
    const int rr = blockIdx.y * kRows + r;
    int v = (d == 0) ? 0 : CTH_DD_MISSING;
    if (rr < lq) v = q.dd[(size_t)(q0 + rr) * CTH_DDSTRIDE + d];
    L.qo[r][d] = (d == 0) ? v : v * 22 * (int)sizeof(float);
  }
  __syncthreads();
  if (row >= lq || col >= ncols) return;

  const int64_t qi = q0 + row;
  const int64_t tj = col0 + col;
  const int tp = db.pro[tj];

  // Init: log-odds (CuBatchScoreMatrix_Init.cu:129-144,173-185)
  float s = 0.0f;
#pragma unroll
  for (int a = 0; a < CTH_NAA; a++) {
    const float f1 = q.trg[a * q.ld + qi];
    const float f2 = q.bkg[qp * CTH_NAA + a];
    const float num = f1 * db.trg[a * db.ld + tj];
    const float den = 0.5f * (f2 + db.bkg[tp * CTH_NAA + a]);
    s = s + num / den;
  }
  s = cth_logf(s) - CTH_INITSHIFT;

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

  const float d = ddms_cell(L.dms, L.qo[w], &L.tl[lane * kListPitch]);
  if (d != 0.0f) {
    s = s + d;
    ms = ms + fmaf(-CTH_DDMSSHIFT, CTH_DDMSWGT, d);
  }
  scores[row * ldsc + col] = s;
  mods[row * ldsc + col] = ms;
}

}  // namespace

// Launch the score stage for query profile qp against template positions
// [col0, col0+ncols) of db; outputs are [Lq][ldsc].
extern "C" int cth_launch_score(const cth_dpset* q, int qp, const cth_dpset* db, int64_t col0, int ncols,
                                int64_t ldsc, const cth_tables* dtabs, int lq, float* scores, float* mods,
                                hipStream_t st) {
  if (ncols <= 0 || lq <= 0) return 0;
  if (ldsc < ncols) { cth_set_error("score: ldsc %lld < ncols %d", (long long)ldsc, ncols); return -1; }
  dim3 grid((ncols + kCols - 1) / kCols, (lq + kRows - 1) / kRows);
  hipLaunchKernelGGL(k_score, grid, dim3(256), 0, st, *q, qp, *db, col0, ncols, ldsc, dtabs, scores, mods);
  CTH_HIP_CHECK(hipGetLastError());
  return 0;
}
