// dp.hip: phase-1 profile-profile DP with traceback, and its finalisation.
//
// k_dp computes, for one template profile per wavefront, the five-state
// local DP of ExecDP_Btck_Unroll32x (cudp/CuBatchDP_init_btck.cu:65-494)
// over the whole Lq x Lt matrix.  MI355X layout: the 64 lanes of a
// wavefront own 64 consecutive query rows (a stripe) and sweep the
// anti-diagonals of the stripe; each step a lane takes its upper
// neighbour's state (and the template column's transitions) from lane-1 by a
// DPP wave_shr:1 move, so the wavefront advances one anti-diagonal per step
// with no LDS round trip.  The bottom row of a stripe is kept in LDS for the
// next stripe (in a global-memory row of the wavefront's own for profiles
// longer than CTH_DP_GLOBAL_FROM positions in the per-query search, whose
// five rows and buffer of maxima would leave a CU one or two single-wavefront
// workgroups in LDS, and over CTH_DP_LDS_ROW_MAX, the LDS capacity, in the
// batched search; the row of column x is read at step x + 31 and
// overwritten at step x + 95 of the same wavefront, so one row serves both
// stripes as in LDS).  Cell values
// equal the reference's: rows start at column -1-(q mod 32) with zero
// boundaries, rows q = 0 mod 32 see zeros above at
// negative columns, rows q = 31 mod 32 see zeros left and upper-left at
// column 0, out-of-range transitions are log(0) = -32768 and cells
// outside the valid area score -9 -- the boundaries the reference's 32x32
// oblique blocks produce.  The maximum and its cell are the ones the
// reference's per-column buffer of maxima yields (dp_pair below), so equal
// maxima resolve as there, and for templates under 32 positions the rows
// whose block-end column lies before the matrix drop out as there.
//
// k_dpfin restates FinalizeDPPro (cudp/CuBatchDP_final.cu:61-292): pass
// test against the score threshold, traceback from the maximum, correlation
// bonus over DIAG scores in 32-score chunks (each chunk reduced in the
// reference's xor-butterfly order), signed square root, positive-score count
// and begin coordinates.
#include <hip/hip_runtime.h>

#include "cth_batch.h"
#include "cth_flow.h"
#include "cth_dev.h"
#include "cth_dp.h"

namespace {

constexpr float kLog0 = CTH_LOG0;

__device__ __forceinline__ float dpp_shr1(float old_lane0, float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old_lane0),
                                                               __builtin_bit_cast(int, v), 0x138, 0xf, 0xf, false));
}

__device__ __forceinline__ float mx(float a, float b) { return a < b ? b : a; }

__device__ __forceinline__ int corner_maxlen(int lq, int lt) {
  int m = (int)rintf(CTH_CORNERPRC * (float)(lq > lt ? lq : lt));
  return m > CTH_MAXCORNER ? CTH_MAXCORNER : m;
}

__device__ __forceinline__ bool in_valid_area(int lq, int lt, int edge, int y, int x) {
  const bool corners = (lq <= y + edge && x <= edge + y - lq) || (lt - edge + y <= x && y < edge);
  return !corners && 0 <= y && y < lq && 0 <= x && x < lt;
}

// The DP of query qp against template profile p, whose scores begin at
// scores[colp] (row stride ldsc); traceback bytes to btck at the same offsets
// (NULL: none); bot: [7][lt] (LDS or global): the stripe boundary rows [5]
// and the reference's buffer of maxima, values and cells.  The maximum and
// its cell follow that buffer (see below) and reach every lane.
//
// The reference (ExecDP_Btck_Unroll32x, cudp/CuBatchDP_init_btck.cu:246-263,
// 328, 436-464) computes 32x32 oblique blocks; thread tx of a block starting
// at column x0 in the block row of query rows 32b..32b+31 takes row
// 32b+31-tx, columns x0+tx .. x0+tx+31, keeps the first cell of a larger MM
// value (strict <), seeded from the buffer at column x0+tx-1, and at the
// block's end writes the buffer at column x0+tx+ndx (ndx = 31, but 0 for a
// block whose column x0+31 is past the end and lt-x0-32 for the block that
// covers the end; block row 0 overwrites, later ones keep the buffer's
// value unless it is smaller).  FinalizeDPPro (cudp/CuBatchDP_final.cu:
// 100-142) reduces the last 64 columns.  Here lane l of stripe s is row
// 64s+l of block row 2s + l/32 with tx = 31 - l%32, and it reaches column c at
// step c + l + 32: every lane's blocks start at steps = 31 and end at steps
// = 30 (mod 32).  At a block end the lower half-wavefront writes before the
// upper one (its block row's blocks run two launches ahead, so where the
// right-edge writes of both meet a column, the lower one's comes first);
// every buffer column then sees its block rows' writes and reads in the
// reference's order.  Lanes past the query's end take part as the
// reference's threads of those rows do.
__device__ __forceinline__ void dp_pair(const cth_dpset& q, int qp, const cth_dpset& db, int p, int64_t colp,
                                        int64_t ldsc, const float* __restrict__ scores, uint8_t* __restrict__ btck,
                                        float* bot, float& best_out, uint32_t& bestxy_out) {
  const int lane = threadIdx.x;
  const int lq = q.len[qp], lt = db.len[p];
  const int64_t qb = q.dist[qp] + qp;      // query transition vector 0
  const int64_t tb = db.dist[p] + p;       // template transition vector 0
  const int edge = corner_maxlen(lq, lt);
  const int ltmax = lt;  // row stride of bot (the launch reserves 7 x its longest profile per wavefront)
  float* const bmv = bot + 5 * ltmax;
  uint32_t* const bmc = reinterpret_cast<uint32_t*>(bot + 6 * ltmax);
  const int nstripes = (lq + 63) >> 6;
  const int c0 = -1 - (lane & 31);
  const int hi = lane & 32;           // 32 for the upper half-wavefront
  const int tx = 31 - (lane & 31);    // the reference's thread in its block
  for (int s = 0; s < nstripes; s++) {
    const int y = (s << 6) + lane;
    const bool rowok = y < lq;
    const int band = 2 * s + (hi >> 5);     // block row of this lane
    const int yb = 32 * band + 31;          // its last query row
    const bool bandok = 32 * band < lq;     // the block row takes part
    float mv = 0.0f;                        // the lane's running maximum
    uint32_t mc = 0;                        // and its cell
    float qt[CTH_NTRN];
    if (rowok) {
      qt[CTH_MM] = q.trn[CTH_MM * q.ldtr + qb + y];
      qt[CTH_MI] = q.trn[CTH_MI * q.ldtr + qb + y + 1];
      qt[CTH_MD] = q.trn[CTH_MD * q.ldtr + qb + y];
      qt[CTH_IM] = q.trn[CTH_IM * q.ldtr + qb + y];
      qt[CTH_II] = q.trn[CTH_II * q.ldtr + qb + y + 1];
      qt[CTH_DM] = q.trn[CTH_DM * q.ldtr + qb + y];
      qt[CTH_DD] = q.trn[CTH_DD * q.ldtr + qb + y];
    } else {
#pragma unroll
      for (int i = 0; i < CTH_NTRN; i++) qt[i] = kLog0;
    }
    // wait for the query transitions here, once per stripe: a wait inside
    // the step loop would also wait for the prefetches issued in it
#pragma unroll
    for (int i = 0; i < CTH_NTRN; i++) asm volatile("" ::"v"(qt[i]));
    float cMM = 0, cMI = 0, cIM = 0, cDN = 0, cND = 0;     // state of this lane's last column
    float pMM = 0, pMI = 0, pIM = 0, pDN = 0, pND = 0;     // up-state received last step (= diagonal now)
    float dt[CTH_NTRN];
#pragma unroll
    for (int i = 0; i < CTH_NTRN; i++) dt[i] = kLog0;
    // Loads run ahead of their use (the step itself is a few hundred cycles
    // of dependent arithmetic): the template column's transitions and the
    // previous stripe's bottom row one step ahead, this lane's score two
    // steps ahead.  Step t uses template column t - 32 and lane column
    // t - lane - 32.
    float ntn[CTH_NTRN], bn[5];
    auto load_col = [&](int cu, float* nt, float* b) {
      if (0 <= cu && cu < lt) {
        nt[CTH_MM] = db.trn[CTH_MM * db.ldtr + tb + cu];
        nt[CTH_MI] = db.trn[CTH_MI * db.ldtr + tb + cu + 1];
        nt[CTH_MD] = db.trn[CTH_MD * db.ldtr + tb + cu];
        nt[CTH_IM] = db.trn[CTH_IM * db.ldtr + tb + cu];
        nt[CTH_II] = db.trn[CTH_II * db.ldtr + tb + cu + 1];
        nt[CTH_DM] = db.trn[CTH_DM * db.ldtr + tb + cu];
        nt[CTH_DD] = db.trn[CTH_DD * db.ldtr + tb + cu];
      } else {
#pragma unroll
        for (int i = 0; i < CTH_NTRN; i++) nt[i] = kLog0;
      }
      if (s > 0 && 0 <= cu && cu < lt) {
        b[0] = bot[0 * ltmax + cu]; b[1] = bot[1 * ltmax + cu]; b[2] = bot[2 * ltmax + cu];
        b[3] = bot[3 * ltmax + cu]; b[4] = bot[4 * ltmax + cu];
      } else {
        b[0] = b[1] = b[2] = b[3] = b[4] = 0.0f;
      }
    };
    // the raw score at lane column c (0 outside the profile; the valid-area
    // test is applied where it is used)
    const float* __restrict__ srow = scores + (int64_t)y * ldsc + colp;
    auto load_sc = [&](int c) { return (rowok && 0 <= c && c < lt) ? srow[c] : 0.0f; };
    // The scores come in groups of four columns per lane, group G of lane l
    // being columns 4G - l .. 4G - l + 3: at step t every lane is at element
    // (t - 32) mod 4 of group (t - 32) / 4, so the whole wavefront refills at
    // once, one 16-byte load (dword-aligned) per lane every four steps instead
    // of four 4-byte loads; each lane reads its own row, so a load touches a
    // cache line per lane, and fewer, wider loads take the texture path and
    // L1 far fewer requests per cell.  Groups reaching outside the profile
    // load their columns one by one (0 outside).
    typedef float float4u __attribute__((ext_vector_type(4), aligned(4)));
    auto load_grp = [&](int G, float* v) {
      const int cs = 4 * G - lane;
      if (rowok && 0 <= cs && cs + 3 < lt) {
        const float4u x = *reinterpret_cast<const float4u*>(srow + cs);
        v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
      } else {
        v[0] = load_sc(cs); v[1] = load_sc(cs + 1); v[2] = load_sc(cs + 2); v[3] = load_sc(cs + 3);
      }
    };
    // Steps 0..30 compute no cell (lane l's first cell, at column
    // -1-(l mod 32), comes at step 31 or later), write nothing and leave the
    // state as it starts (dt log 0, zero states; no block starts or ends
    // there), so the sweep starts at step 31 with its loads issued for it.
    // After lane 63's last column (step lt + 94) only the final block end
    // (step 32 ceil(lt/32) + 94, the next step = 30 mod 32, with no block
    // start before it) remains, so the sweep jumps there.
    constexpr int t0 = 31;
    load_col(t0 - 32, ntn, bn);
    float grp[4], grpn[4];  // this step's group and the next
    load_grp((t0 - 32) >> 2, grp);
    load_grp(((t0 - 32) >> 2) + 1, grpn);
    // the last block of the upper half ends at step 32 ceil(lt/32) + 94
    const int nsteps = 32 * ((lt + 31) >> 5) + 95;
    const int tlast = lt + 94;
    for (int t = t0; t < nsteps; t = t < tlast ? t + 1 : (t < nsteps - 1 ? nsteps - 1 : nsteps)) {
      const int c = t - lane - 32;
      const int cu = t - 32;
      if ((t & 31) == 31) {
        // a block starts: seed from the buffer at the column before
        const int x0 = t - 63 - hi;
        if (bandok && -32 <= x0 && x0 < lt) {
          const int xs = x0 + tx - 1;
          if (0 <= xs && xs < lt) {
            mv = bmv[xs];
            mc = bmc[xs];
          } else {
            mv = 0.0f;
            mc = 0u;
          }
        }
      }
      // this step's column data, then the loads for the steps after
      float nt[CTH_NTRN], bcu[5];
#pragma unroll
      for (int i = 0; i < CTH_NTRN; i++) nt[i] = ntn[i];
#pragma unroll
      for (int i = 0; i < 5; i++) bcu[i] = bn[i];
      const int e = (t - 32) & 3;
      const float scv = e == 0 ? grp[0] : e == 1 ? grp[1] : e == 2 ? grp[2] : grp[3];
      if (e == 3) {
#pragma unroll
        for (int i = 0; i < 4; i++) grp[i] = grpn[i];
        load_grp(((t - 32) >> 2) + 2, grpn);
      }
      load_col(cu + 1, ntn, bn);
      // template transitions: lane 0 takes column t-32, the others lane-1's
#pragma unroll
      for (int i = 0; i < CTH_NTRN; i++) dt[i] = dpp_shr1(nt[i], dt[i]);
      // up state: lane 0 reads the previous stripe's bottom row (zero for the
      // first stripe and for negative columns), the others take lane-1's
      float uMM = dpp_shr1(bcu[0], cMM), uMI = dpp_shr1(bcu[1], cMI), uIM = dpp_shr1(bcu[2], cIM);
      float uDN = dpp_shr1(bcu[3], cDN), uND = dpp_shr1(bcu[4], cND);
      if ((lane & 31) == 0 && c < 0) { uMM = uMI = uIM = uDN = uND = 0.0f; }
      if ((lane & 31) == 31 && c == 0) {  // thread 0 of the block at x = 0: zero left/diagonal
        cMM = cMI = cIM = cDN = cND = 0.0f;
        pMM = pMI = pIM = pDN = pND = 0.0f;
      }
      if (rowok && c0 <= c && c < lt) {
        const float S = in_valid_area(lq, lt, edge, y, c) ? scv : CTH_OUTSCORE;
        float v1, v2, m;
        int bk;
        v1 = pMI + qt[CTH_MM] + dt[CTH_IM];
        v2 = pIM + qt[CTH_IM] + dt[CTH_MM];
        v1 = mx(v1, v2);
        v2 = pDN + qt[CTH_DM] + dt[CTH_MM];
        v1 = mx(v1, v2);
        v2 = pND + qt[CTH_MM] + dt[CTH_DM];
        v1 = mx(v1, v2);
        v2 = pMM + qt[CTH_MM] + dt[CTH_MM];
        v1 = mx(v1, v2);
        const float nMM = mx(v1 + S, 0.0f);
        m = nMM;
        bk = nMM != 0.0f ? CTH_BT_DIAG : CTH_BT_STOP;
        v1 = uMM + qt[CTH_MM] + dt[CTH_MI];
        v2 = uMI + qt[CTH_MM] + dt[CTH_II];
        const float nMI = mx(v1, v2);
        if (m < nMI) { m = nMI; bk = CTH_BT_UP; }
        v1 = cMM + qt[CTH_MI] + dt[CTH_MM];
        v2 = cIM + qt[CTH_II] + dt[CTH_MM];
        const float nIM = mx(v1, v2);
        if (m < nIM) { m = nIM; bk = CTH_BT_LEFT; }
        v1 = uMM + qt[CTH_MD];
        v2 = uDN + qt[CTH_DD];
        const float nDN = mx(v1, v2);
        if (m < nDN) { m = nDN; bk = CTH_BT_UP; }
        v1 = cMM + dt[CTH_MD];
        v2 = cND + dt[CTH_DD];
        const float nND = mx(v1, v2);
        if (m < nND) { m = nND; bk = CTH_BT_LEFT; }
        cMM = nMM; cMI = nMI; cIM = nIM; cDN = nDN; cND = nND;
        if (0 <= c) {
          if (mv < nMM) { mv = nMM; mc = ((uint32_t)c << 16) | (uint32_t)y; }
          if (btck) btck[(int64_t)y * ldsc + colp + c] = (uint8_t)bk;
        }
      }
      pMM = uMM; pMI = uMI; pIM = uIM; pDN = uDN; pND = uND;
      if (lane == 63 && 0 <= c && c < lt) {
        bot[0 * ltmax + c] = cMM; bot[1 * ltmax + c] = cMI; bot[2 * ltmax + c] = cIM;
        bot[3 * ltmax + c] = cDN; bot[4 * ltmax + c] = cND;
      }
      if ((t & 31) == 30) {
        // a block ends: write the buffer, the lower half-wavefront first
        const int x0 = t - 94 - hi;
        const int w = flow_write_column(x0, tx, lt);
        const bool wr = bandok && -32 <= x0 && x0 < lt && flow_block_iters(x0, yb, lq, lt) >= 1 && 0 <= w && w < lt;
        if (wr && hi == 0 && (band == 0 || bmv[w] < mv)) {
          bmv[w] = mv;
          bmc[w] = mc;
        }
        flow_phase();
        if (wr && hi != 0 && bmv[w] < mv) {
          bmv[w] = mv;
          bmc[w] = mc;
        }
        flow_phase();
      }
    }
    __syncthreads();
  }
  // FinalizeDPPro's reduction of the buffer's last 64 columns
  flow_final(bmv, bmc, lt, best_out, bestxy_out);
}

template <bool kGlobalRow>
__global__ __launch_bounds__(64) void k_dp(cth_dpset q, int qp, cth_dpset db, int pbase,
                                           const int* __restrict__ plist, int64_t col0, int64_t ldsc,
                                           const float* __restrict__ scores, uint8_t* __restrict__ btck,
                                           cth_dpres* __restrict__ res, float* __restrict__ grow,
                                           int64_t growstride) {
  // [7][lt]: bottom row of the previous stripe and the buffer of maxima, in
  // LDS or in this wavefront's growstride floats of grow
  extern __shared__ float lds_row[];
  float* const bot = kGlobalRow ? grow + (int64_t)blockIdx.x * growstride : lds_row;
  const int p = plist ? plist[blockIdx.x] : pbase + blockIdx.x;
  float best;
  uint32_t bestxy;
  dp_pair(q, qp, db, p, db.dist[p] - col0, ldsc, scores, btck, bot, best, bestxy);
  if (threadIdx.x == 0) {
    res[blockIdx.x].maxsc = best;
    res[blockIdx.x].maxxy = bestxy;
  }
}

// Batched phase 1: a persistent grid works through n pairs (or the *count
// pairs listed in list), each wavefront taking the next pair from the device
// counter next; pair pairs[i] of query bq[pairs[i].b] reads that query's
// scores (and writes its traceback bytes) at the query's offset; res is per
// pair.  Pairs come in the order given (longest first), so the longest DPs
// start first.
//
// No branch in this loop depends on the lane: every lane takes part in the
// fetch (lane 0 adds 1, the others 0, so lane 0's old value is the pair) and
// every lane stores the wave-uniform maximum.  A fetch under
// `if (threadIdx.x == 0)` lets the compiler thread that loop-invariant branch
// across the back edge, which leaves lanes 1..63 in a second copy of the loop
// that never fetches again (they re-run the first pair forever).
template <bool kTrace>
__global__ __launch_bounds__(64) void k_dp_batch(cth_dpset q, cth_dpset db, const cth_bquery* __restrict__ bq,
                                                 const cth_bpair* __restrict__ pairs, const int* __restrict__ list,
                                                 const int* __restrict__ count, int n, int* __restrict__ next,
                                                 int64_t ldsc, const float* __restrict__ scores,
                                                 uint8_t* __restrict__ btck, cth_dpres* __restrict__ res) {
  extern __shared__ float lds_row[];
  const int nj = count ? *count : n;
  for (;;) {
    // uniform: scalar loads of the pair and query
    const int k = __builtin_amdgcn_readfirstlane(atomicAdd(next, threadIdx.x == 0 ? 1 : 0));
    if (k >= nj) break;
    const int i = list ? list[k] : k;
    const cth_bpair pr = pairs[i];
    const cth_bquery Q = bq[pr.b];
    float best;
    uint32_t bestxy;
    dp_pair(q, Q.qp, db, pr.p, db.dist[pr.p], ldsc, scores + Q.off, kTrace ? btck + Q.off : nullptr, lds_row, best,
            bestxy);
    // the butterfly in dp_pair leaves the same maximum in every lane
    res[i].maxsc = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, best)));
    res[i].maxxy = (uint32_t)__builtin_amdgcn_readfirstlane((int)bestxy);
  }
}

// FinalizeDPPro for one profile (lane 0): pass test against scorethld,
// traceback from the maximum, correlation bonus, signed square root,
// positive-score count and begin coordinates into r.
__device__ __forceinline__ void dpfin_pair(const cth_dpset& db, int p, int64_t colp, int64_t ldsc, const float* __restrict__ scores,
                           const uint8_t* __restrict__ btck, float scorethld, cth_dpres& r, float* cache,
                           float* tree) {
  r.passed = 0; r.score = 0.0f; r.psts = 0; r.beg = 0;
  if (r.maxsc <= scorethld) return;
  int x = (int)(r.maxxy >> 16), y = (int)(r.maxxy & 0xffff);
  int psts = 0;
  for (int k = 0; k < 36; k++) cache[k] = 0.0f;
  float sum = 0.0f;
  int bt = CTH_BT_DIAG;
  while (bt != CTH_BT_STOP) {
    for (int k = 0; k < 4; k++) cache[k] = cache[32 + k];
    int ndx;
    for (ndx = 4; ndx < 36;) {
      if (x < 0 || y < 0) { bt = CTH_BT_STOP; break; }
      const int64_t a = (int64_t)y * ldsc + colp + x;
      bt = btck[a];
      if (bt == CTH_BT_STOP) break;
      if (bt == CTH_BT_UP) { y--; continue; }
      if (bt == CTH_BT_LEFT) { x--; continue; }
      const float v = scores[a];
      cache[ndx] = v;
      if (v > 0.0f) psts++;
      x--; y--;
      ndx++;
    }
    for (int l = 0; l < 32; l++)
      tree[l] = (ndx <= l + 4) ? 0.0f
                               : cache[l + 4] * (((cache[l] + cache[l + 1]) + cache[l + 2]) + cache[l + 3]);
    // xor butterfly (offsets 16,8,4,2,1); lane 0's value
    for (int off = 16; off >= 1; off >>= 1) {
      float nt[32];
      for (int l = 0; l < 32; l++) nt[l] = tree[l] + tree[l ^ off];
      for (int l = 0; l < 32; l++) tree[l] = nt[l];
    }
    sum += tree[0];
  }
  sum *= 0.5f;
  sum = sum < 0.0f ? -sqrtf(-sum) : sqrtf(sum);
  const float score = r.maxsc + sum;
  r.psts = psts;
  r.beg = ((uint32_t)(x + 1) << 16) | (uint32_t)(y + 1);
  r.score = score;
  r.passed = scorethld < score ? 1 : 0;
}

// lane 0 of each wavefront finalises one profile
__global__ __launch_bounds__(64) void k_dpfin(cth_dpset db, int pbase, const int* __restrict__ plist, int64_t col0,
                                              int64_t ldsc, const float* __restrict__ scores,
                                              const uint8_t* __restrict__ btck, float scorethld,
                                              cth_dpres* __restrict__ res) {
  __shared__ float cache[36];
  __shared__ float tree[32];
  if (threadIdx.x != 0) return;
  const int p = plist ? plist[blockIdx.x] : pbase + blockIdx.x;
  cth_dpres r = res[blockIdx.x];
  dpfin_pair(db, p, db.dist[p] - col0, ldsc, scores, btck, scorethld, r, cache, tree);
  res[blockIdx.x] = r;
}

// Batched finalisation over the *count pairs listed in list (persistent,
// lane 0 of each wavefront), each against its query's threshold.
__global__ __launch_bounds__(64) void k_dpfin_batch(cth_dpset db, const cth_bquery* __restrict__ bq,
                                                    const cth_bpair* __restrict__ pairs, const int* __restrict__ list,
                                                    const int* __restrict__ count, int64_t ldsc,
                                                    const float* __restrict__ scores,
                                                    const uint8_t* __restrict__ btck, cth_dpres* __restrict__ res) {
  __shared__ float cache[36];
  __shared__ float tree[32];
  if (threadIdx.x != 0) return;
  const int nj = *count;
  for (int k = blockIdx.x; k < nj; k += gridDim.x) {
    const int i = list[k];
    const cth_bpair pr = pairs[i];
    const cth_bquery Q = bq[pr.b];
    cth_dpres r = res[i];
    dpfin_pair(db, pr.p, db.dist[pr.p], ldsc, scores + Q.off, btck + Q.off, Q.scorethld, r, cache, tree);
    res[i] = r;
  }
}

__global__ void k_selftest_dpp(int* out) {
  const float v = (float)threadIdx.x;
  out[threadIdx.x] = (int)dpp_shr1(-1.0f, v);
}

}  // namespace

// Lane-layout check of the DPP wave_shr:1 move the DP kernel relies on:
// out[l] = l-1 for l >= 1, out[0] = -1.
extern "C" int cth_selftest_dpp(int* host_out64) {
  int* d = nullptr;
  CTH_HIP_CHECK(hipMalloc(&d, 64 * sizeof(int)));
  hipLaunchKernelGGL(k_selftest_dpp, dim3(1), dim3(64), 0, 0, d);
  CTH_HIP_CHECK(hipGetLastError());
  CTH_HIP_CHECK(hipMemcpy(host_out64, d, 64 * sizeof(int), hipMemcpyDeviceToHost));
  CTH_HIP_CHECK(hipFree(d));
  return 0;
}

// Bytes of global boundary rows a launch over np profiles of at most ltmax
// positions needs (0 when the rows fit the LDS).
extern "C" size_t cth_dp_row_bytes(int np, int ltmax) {
  return ltmax > CTH_DP_GLOBAL_FROM ? sizeof(float) * 7 * (size_t)ltmax * (size_t)(np > 0 ? np : 0) : 0;
}

// Phase 1 for np profiles: p0 + k, or plist[k] (device array of absolute
// indices) when plist is not NULL; res[k] receives profile k's record.
// btck NULL: no traceback bytes (maximum and its coordinates only);
// finalize: run the finalisation (needs btck).  grow: cth_dp_row_bytes(np,
// ltmax) bytes of device memory when ltmax exceeds CTH_DP_GLOBAL_FROM (else
// unused).
extern "C" int cth_launch_dp_list(const cth_dpset* q, int qp, const cth_dpset* db, int p0, const int* plist, int np,
                                  int ltmax, int64_t col0, int64_t ldsc, const float* scores, uint8_t* btck,
                                  float scorethld, int finalize, cth_dpres* res, float* grow, hipStream_t st) {
  if (np <= 0) return 0;
  if (finalize && !btck) { cth_set_error("dp: finalisation needs the traceback"); return -1; }
  if (ltmax > CTH_MAXCOLUMNS) { cth_set_error("dp: profile length %d over %d", ltmax, CTH_MAXCOLUMNS); return -1; }
  if (ltmax > CTH_DP_GLOBAL_FROM) {
    if (!grow) { cth_set_error("dp: profile length %d needs global boundary rows", ltmax); return -1; }
    hipLaunchKernelGGL(k_dp<true>, dim3(np), dim3(64), 0, st, *q, qp, *db, p0, plist, col0, ldsc, scores, btck, res,
                       grow, (int64_t)7 * ltmax);
  } else {
    const size_t lds = sizeof(float) * 7 * (size_t)(ltmax > 0 ? ltmax : 1);
    hipLaunchKernelGGL(k_dp<false>, dim3(np), dim3(64), lds, st, *q, qp, *db, p0, plist, col0, ldsc, scores, btck,
                       res, nullptr, (int64_t)0);
  }
  CTH_HIP_CHECK(hipGetLastError());
  if (finalize) {
    hipLaunchKernelGGL(k_dpfin, dim3(np), dim3(64), 0, st, *db, p0, plist, col0, ldsc, scores, btck, scorethld, res);
    CTH_HIP_CHECK(hipGetLastError());
  }
  return 0;
}

extern "C" int cth_launch_dp(const cth_dpset* q, int qp, const cth_dpset* db, int p0, int np, int ltmax,
                             int64_t col0, int64_t ldsc, const float* scores, uint8_t* btck, float scorethld,
                             cth_dpres* res, float* grow, hipStream_t st) {
  return cth_launch_dp_list(q, qp, db, p0, nullptr, np, ltmax, col0, ldsc, scores, btck, scorethld, 1, res, grow, st);
}

// Batched phase 1 (k_dp_batch) on a persistent grid of `grid` wavefronts:
// the first pass over n pairs (list NULL, btck NULL), or the traceback pass
// over the *count pairs of list followed by their finalisation (btck not
// NULL).  next: a device counter that starts at zero.  Every profile of the
// database fits the LDS row limit (the batched path keeps boundary rows in
// LDS; ltmax is the database's longest profile).
extern "C" int cth_launch_dp_batch(const cth_dpset* q, const cth_dpset* db, const cth_bquery* bq,
                                   const cth_bpair* pairs, const int* list, const int* count, int n, int* next,
                                   int ltmax, int64_t ldsc, const float* scores, uint8_t* btck, cth_dpres* res,
                                   int grid, hipStream_t st) {
  if (ltmax > CTH_DP_LDS_ROW_MAX) { cth_set_error("dp batch: profile length %d over the LDS row limit", ltmax); return -1; }
  if (grid < 1) return 0;
  const size_t lds = sizeof(float) * 7 * (size_t)(ltmax > 0 ? ltmax : 1);
  if (btck)
    hipLaunchKernelGGL(k_dp_batch<true>, dim3(grid), dim3(64), lds, st, *q, *db, bq, pairs, list, count, n, next, ldsc,
                       scores, btck, res);
  else
    hipLaunchKernelGGL(k_dp_batch<false>, dim3(grid), dim3(64), lds, st, *q, *db, bq, pairs, list, count, n, next, ldsc,
                       scores, btck, res);
  CTH_HIP_CHECK(hipGetLastError());
  if (btck && list) {
    hipLaunchKernelGGL(k_dpfin_batch, dim3(grid), dim3(64), 0, st, *db, bq, pairs, list, count, ldsc, scores, btck,
                       res);
    CTH_HIP_CHECK(hipGetLastError());
  }
  return 0;
}
