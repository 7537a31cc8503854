/* The log-odds quotient of the plain score kernel (score.hip k_score_plain)
 * in its division-free form: with y = RN(1/den), q0 = RN(num*y),
 * r = fma(-den, q0, num), q = fma(r, y, q0), the result must equal the IEEE
 * quotient RN(num/den) for every num = RN(a*b) with factors a, b of +0 or in
 * [2^-32, 2] and every den in [2^-30, 2^30] (the ranges the kernel admits,
 * compared on bit patterns as it does).  Checked on random bit patterns of
 * those ranges and on full significand sweeps against denominators whose
 * reciprocals round hardest (all-ones and near-one significands).
 * Usage: quotient_check [random pairs]; prints the mismatch count, exits 1
 * on any mismatch. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float bf(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t fb(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static uint64_t next(uint64_t* s) { *s ^= *s << 13; *s ^= *s >> 7; *s ^= *s << 17; return *s; }

static long bad = 0;
static void check(float a, float b, float den) {
  volatile float dv = den, nv = a * b;
  const float num = nv, y = 1.0f / dv;
  const float q0 = num * y;
  const float q = fmaf(fmaf(-den, q0, num), y, q0);
  const float ref = num / dv;
  if (fb(q) != fb(ref)) {
    if (bad < 5) printf("a %a b %a den %a: %a, division %a\n", a, b, den, q, ref);
    bad++;
  }
}

int main(int argc, char** argv) {
  const long n = argc > 1 ? atol(argv[1]) : 20000000L;
  const uint32_t f_lo = 0x2F800000u, f_span = 0x10800000u;  /* [2^-32, 2] */
  const uint32_t d_lo = 0x30800000u, d_span = 0x1E000000u;  /* [2^-30, 2^30] */
  uint64_t s = 0x243F6A8885A308D3ULL;
  for (long i = 0; i < n; i++) {
    const uint64_t r = next(&s), t = next(&s);
    const float a = (r & 0xff) == 0 ? 0.0f : bf(f_lo + (uint32_t)((r >> 8) % (f_span + 1)));
    const float b = bf(f_lo + (uint32_t)((r >> 36) % (f_span + 1)));
    check(a, b, bf(d_lo + (uint32_t)(t % (d_span + 1))));
  }
  const uint32_t hard[] = {0x3f7fffffu, 0x3f800001u, 0x3fffffffu, 0x3faaaaabu, 0x3f555555u};
  for (int k = 0; k < 5; k++)
    for (uint32_t m = 0; m < (1u << 23); m += 3) {
      check(bf(0x3f800000u | m), 1.0f, bf(hard[k]));                /* num in [1, 2) */
      check(bf(0x2f800000u | m), bf(0x2f800000u), bf(hard[k] + (28u << 23)));  /* tiny num, den near 2^28 */
    }
  printf("%ld mismatched\n", bad);
  return bad != 0;
}
