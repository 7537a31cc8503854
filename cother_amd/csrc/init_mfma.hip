// init_mfma.hip: the log-odds term of the score matrix as a product on the
// matrix cores (the reduced-precision contraction variant).
//
// The reference sums the term per cell in FP32 (CalcSM_Init_SMEMUnroll2x,
// cusco/CuBatchScoreMatrix_Init.cu:129-144,174-185):
//   init[i][j] = log( sum_a q_i(a) t_j(a) / (0.5 (bq(a) + bt_p(j)(a))) ) - 0.002
// which is a product of the query's target frequencies Q [Lq x 20] and the
// template's weighted frequencies T' [ncols x 20], T'[j][a] = t_j(a) /
// (0.5 (bq(a) + bt_p(j)(a))) (its weights depend on the query's background,
// so T' is formed per launch).  Here both operands are rounded to BF16 (or
// FP16) and multiplied on the MFMA units with FP32 accumulation, K = 20
// padded to 32 with zeros; the logarithm stays FP32.  The score kernel then
// adds the secondary-structure, context-vector and DDMS terms to it
// unchanged.  The FP32 path (the default) does not use this kernel.
//
// Tile: a workgroup of four wavefronts computes 64 query rows x 64 template
// columns; each wavefront one 32 x 32 block with two
// v_mfma_f32_32x32x16 (bf16 / f16) steps over K.  Operand layout of the 32x32x16
// instruction (lane l, 64 lanes): A holds row l%32, k = 8*(l/32) + 0..7; B
// holds column l%32, the same k; the 16 accumulators of lane l hold column
// l%32, rows 8*(r/4) + 4*(l/32) + r%4 for r = 0..15 (checked on the device by
// cth_selftest_mfma).
#include <hip/hip_runtime.h>

#include "cth_dev.h"
#include "cth_math.h"

namespace {

typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

constexpr int kTile = 64;
constexpr int kK = 32;

// FP32 -> BF16, round to nearest even (the operands are finite and >= 0)
__device__ __forceinline__ uint16_t to_bf16(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ uint16_t to_f16(float x) {
  const _Float16 h = (_Float16)x;
  return *reinterpret_cast<const uint16_t*>(&h);
}

template <int kPrec>  // 1: BF16, 2: FP16; a, b: 8 consecutive k of the lane's row / column
__device__ __forceinline__ v16f mfma(const uint16_t* a, const uint16_t* b, v16f acc) {
  const v8s av = *reinterpret_cast<const v8s*>(a);
  const v8s bv = *reinterpret_cast<const v8s*>(b);
  if constexpr (kPrec == 1)
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(v8bf, av), __builtin_bit_cast(v8bf, bv), acc,
                                                   0, 0, 0);
  else
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(v8h, av), __builtin_bit_cast(v8h, bv), acc, 0,
                                                  0, 0);
}

template <int kPrec>
__global__ __launch_bounds__(256) void k_init_mfma(cth_dpset q, int qp, cth_dpset db, int64_t col0, int ncols,
                                                   int64_t ldsc, int lq, float* __restrict__ out) {
  // +8: rows of 80 bytes, 16-byte aligned for the 8-element operand reads,
  // spread over the LDS banks
  __shared__ __align__(16) uint16_t A[kTile][kK + 8];
  __shared__ __align__(16) uint16_t B[kTile][kK + 8];
  __shared__ float bq[CTH_NAA];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int row0 = blockIdx.y * kTile, c0 = blockIdx.x * kTile;
  const int64_t q0 = q.dist[qp];
  if (tid < CTH_NAA) bq[tid] = q.bkg[qp * CTH_NAA + tid];
  __syncthreads();
  for (int e = tid; e < kTile * kK; e += 256) {
    const int r = e / kK, k = e % kK;
    float va = 0.0f, vb = 0.0f;
    if (k < CTH_NAA) {
      if (row0 + r < lq) va = q.trg[k * q.ld + q0 + row0 + r];
      if (c0 + r < ncols) {
        const int64_t tj = col0 + c0 + r;
        const int tp = db.pro[tj];
        vb = db.trg[k * db.ld + tj] / (0.5f * (bq[k] + db.bkg[tp * CTH_NAA + k]));
      }
    }
    A[r][k] = kPrec == 1 ? to_bf16(va) : to_f16(va);
    B[r][k] = kPrec == 1 ? to_bf16(vb) : to_f16(vb);
  }
  __syncthreads();
  const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
  // operands: row / column l32, k = kb + 8 (lane / 32) + 0..7; accumulators:
  // rows 8 (r / 4) + 4 (lane / 32) + r % 4
  const int l32 = lane & 31, kh = (lane >> 5) * 4, ko = (lane >> 5) * 8;
  v16f acc = {};
#pragma unroll
  for (int kb = 0; kb < kK; kb += 16) acc = mfma<kPrec>(&A[wr + l32][kb + ko], &B[wc + l32][kb + ko], acc);
  const int col = c0 + wc + l32;
  if (col >= ncols) return;
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const int row = row0 + wr + 8 * (r >> 2) + kh + (r & 3);
    if (row < lq) out[(int64_t)row * ldsc + col] = cth_logf(acc[r]) - CTH_INITSHIFT;
  }
}

// Lane layout check of the 32x32x16 instructions: A = I (32 x 16 block of
// the identity), B[k][c] = 32 (k - 8 h) + c for k in [8 h, 8 h + 8) and 0
// elsewhere (h = 0, 1; every value below 256, exact in BF16), so
// D[r][c] = B[r][c] for r < 16; out[lane*16 + r] receives accumulator r of
// each lane.
template <int kPrec>
__global__ void k_selftest_mfma(float* out, int h) {
  const int lane = threadIdx.x;
  const int l32 = lane & 31, ko = (lane >> 5) * 8;
  __align__(16) uint16_t a[8], b[8];
  for (int k = 0; k < 8; k++) {
    const int kk = ko + k;
    const float av = (l32 == kk) ? 1.0f : 0.0f;
    const float bv = (kk >> 3) == h ? (float)(32 * (kk - 8 * h) + l32) : 0.0f;
    a[k] = kPrec == 1 ? to_bf16(av) : to_f16(av);
    b[k] = kPrec == 1 ? to_bf16(bv) : to_f16(bv);
  }
  v16f acc = {};
  acc = mfma<kPrec>(a, b, acc);
  for (int r = 0; r < 16; r++) out[lane * 16 + r] = acc[r];
}

}  // namespace

extern "C" int cth_launch_init_mfma(const cth_dpset* q, int qp, const cth_dpset* db, int64_t col0, int ncols,
                                    int64_t ldsc, int lq, int prec, float* out, hipStream_t st) {
  if (ncols <= 0 || lq <= 0) return 0;
  if (prec != 1 && prec != 2) { cth_set_error("init_mfma: precision %d (1 BF16, 2 FP16)", prec); return -1; }
  if (ldsc < ncols) { cth_set_error("init_mfma: ldsc %lld < ncols %d", (long long)ldsc, ncols); return -1; }
  dim3 grid((ncols + kTile - 1) / kTile, (lq + kTile - 1) / kTile);
  if (prec == 1) hipLaunchKernelGGL(k_init_mfma<1>, grid, dim3(256), 0, st, *q, qp, *db, col0, ncols, ldsc, lq, out);
  else hipLaunchKernelGGL(k_init_mfma<2>, grid, dim3(256), 0, st, *q, qp, *db, col0, ncols, ldsc, lq, out);
  CTH_HIP_CHECK(hipGetLastError());
  return 0;
}

// Device check of the MFMA lane layout assumed above: 0 when every
// accumulator holds the expected element (for both precisions).
extern "C" int cth_selftest_mfma(int* bad) {
  float* d = nullptr;
  CTH_HIP_CHECK(hipMalloc((void**)&d, 64 * 16 * sizeof(float)));
  float h[64 * 16];
  *bad = 0;
  for (int prec = 1; prec <= 2; prec++)
    for (int half = 0; half < 2; half++) {
      if (prec == 1) hipLaunchKernelGGL(k_selftest_mfma<1>, dim3(1), dim3(64), 0, 0, d, half);
      else hipLaunchKernelGGL(k_selftest_mfma<2>, dim3(1), dim3(64), 0, 0, d, half);
      if (hipMemcpy(h, d, sizeof h, hipMemcpyDeviceToHost) != hipSuccess) {
        (void)hipFree(d);
        cth_set_error("selftest_mfma: copy failed");
        return -1;
      }
      for (int lane = 0; lane < 64; lane++)
        for (int r = 0; r < 16; r++) {
          const int row = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3), col = lane & 31;
          const float want = (row >> 3) == half ? (float)(32 * (row - 8 * half) + col) : 0.0f;
          if (h[lane * 16 + r] != want) (*bad)++;
        }
    }
  (void)hipFree(d);
  return 0;
}
