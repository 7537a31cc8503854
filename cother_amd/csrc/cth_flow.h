// cth_flow.h: the reference's per-column buffer of maxima, shared by the
// phase-1 DP (dp.hip) and the MAP alignment DP (map.hip).
//
// The reference's 32x32 block DPs (cudp/CuBatchDP_init_btck.cu:246-263,328,
// 436-464; cumapdp/CuBatchMAPDP_map_btck.cu:232-244,286,341-370) keep each
// block thread's running maximum, seeded from a per-column buffer at the
// column before the thread's first cell and written back at the block's end;
// the finalisation kernels (cudp/CuBatchDP_final.cu:100-142,
// cumapdp/CuBatchMAPDP_map_final.cu:170-214) reduce the buffer's last 64
// columns.  The kernels here replay that flow in their stripe sweep (see
// dp_pair in dp.hip); these are its pieces that do not depend on the DP.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// MaxIterationsForBlock: the columns the reference's block starting at x0 in
// the block row ending at query row y iterates over
__device__ __forceinline__ int flow_block_iters(int x0, int y, int lq, int lt) {
  int x = lt - x0;
  if (x <= 0) return 0;
  const int yy = y - (lq - 1);
  if (0 < yy) x -= yy;
  if (x <= 0) return 0;
  return x < 32 ? x : 32;
}

// the buffer column a thread tx of the block at x0 writes at the block's
// end: its last column x0+tx+31, but its first for a block whose column
// x0+31 is past the end, and lt-32+tx for the block that covers the end
__device__ __forceinline__ int flow_write_column(int x0, int tx, int lt) {
  int ndx = 31;
  const int xtl = x0 + 31;
  if (lt <= xtl) ndx = 0;
  else if (lt <= xtl + ndx) ndx = lt - xtl - 1;
  return x0 + tx + ndx;
}

// Between the two half-wavefront phases of a block end (the lower half's
// buffer writes, then the upper half's compare-and-write): the phases run
// under exclusive lane conditions (hi == 0, hi != 0), so reasoning per
// thread, a compiler may sink the first phase's stores into the second's
// and issue both halves' stores as one instruction after the upper half's
// reads (hipcc did so with the DP's step loop unrolled).  An inline asm that
// may read and write any memory can neither be crossed by a store nor, being
// convergent, be duplicated into the branches.  The fence keeps the
// memory-model ordering; a wavefront's LDS accesses execute in issue order.
__device__ __forceinline__ void flow_phase() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  asm volatile("" ::: "memory");
}

// the finalisation's reduction of the last 64 buffer columns (bmv values,
// bmc cells): column lt-1-l against lt-33-l, then halves of 32 .. 1, a tie
// keeping the lower index (the later column); the result reaches every lane
__device__ __forceinline__ void flow_final(const float* bmv, const uint32_t* bmc, int lt, float& best,
                                           uint32_t& bestxy) {
  const int lane = threadIdx.x;
  float v = lane < lt ? bmv[lt - 1 - lane] : 0.0f;
  int ix = lane;
  {
    const float o = __shfl_down(v, 32);
    if (lane < 32 && v < o) { v = o; ix = lane + 32; }
  }
  for (int h = 16; h >= 1; h >>= 1) {
    const float o = __shfl_down(v, h);
    const int oi = __shfl_down(ix, h);
    if (lane < 32 && v < o) { v = o; ix = oi; }
  }
  best = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
  const int ib = __builtin_amdgcn_readfirstlane(ix);
  bestxy = bmc[lt - 1 - (ib < lt ? ib : lt - 1)];
}
