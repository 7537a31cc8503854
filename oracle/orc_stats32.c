/* orc_stats32.c: a single-precision twin of orc_stats.c (TEST
 * INFRASTRUCTURE ONLY; see oracle.h).
 *
 * The reference fits the score statistics (lambda, H, K) and computes the
 * E-values in single precision (cuss/CuBatchSS_dynplm.cu:134-580,
 * CuBatchSS_lambda.cuh, CuBatchSS_entropy.cuh, CuBatchSS_K.cuh,
 * CuBatchSS_SS18.cuh); the oracle and the GPU path run the same steps in
 * double precision.  This file compiles the same source with every double
 * quantity and math call in float (and, by its make rule, float constants)
 * under orc32_ names, so that tests can state how far the E-values of the
 * double-precision statistics lie from single-precision arithmetic.  It is
 * not the reference's float code: its exp/log are the C library's, not
 * CUDA's fast intrinsics. */
#include <math.h>
#include <stdint.h>
#include <string.h>

#define ORC_FPMIN 1e-30f
#define ORC_EPS 1e-7f
#define double float
#define exp expf
#define log logf
#define log1p log1pf
#define pow powf
#define fabs fabsf
#define lgamma lgammaf
#define rint rintf
#define orc_histogram orc32_histogram
#define orc_stat_params orc32_stat_params
#define orc_lbetainc orc32_lbetainc
#define orc_lgammaincq orc32_lgammaincq
#define orc_ss18_weights orc32_ss18_weights
#define orc_ss18_nn orc32_ss18_nn
#define orc_ss18_input_e orc32_ss18_input_e
#define orc_ss18_input_l orc32_ss18_input_l
#define orc_ss18_predict orc32_ss18_predict
#define orc_ss18_score_logp orc32_ss18_score_logp
#define orc_ss18_psts_logp orc32_ss18_psts_logp
#define orc_ss18_combine orc32_ss18_combine
#define orc_ss18_logevalue orc32_ss18_logevalue
#define orc_significance orc32_significance

#include "orc_stats.c"
