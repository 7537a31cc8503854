/* orc_stats.c: oracle restatement of the alignment statistics and E-values
 * (TEST INFRASTRUCTURE ONLY; see oracle.h).
 *
 * Per passed profile, in double precision (the statistics fit runs in FP64
 * in this build; the reference computes it in FP32):
 *   histogram   CalcScoreProbsXuninfDynPlmProfile   cuss/CuBatchSS_dynplm.cu:134-182
 *   parameters  CalcStatParamsDynPlmProfile         cuss/CuBatchSS_dynplm.cu:193-430
 *     expected  CalcExpectedScoreProfile            cuss/CuBatchSS_expected.cuh:22-56
 *     min/max   GetScoreMinMaxProfile               cuss/CuBatchSS_minmax.cuh:23-140
 *     lambda    FindLmbdRootByNRandBisection        cuss/CuBatchSS_lambda.cuh:132-330
 *     entropy   CalcEntropyProfile                  cuss/CuBatchSS_entropy.cuh:22-80
 *     K         CalcKProfile/CalcKSigma/
 *               CalcProbOfAlnLengthj                cuss/CuBatchSS_K.cuh:110-321
 *   E-values    CalcSignificanceSS18DynPlmProfile   cuss/CuBatchSS_dynplm.cu:447-580
 *               (EstimateGappedLambdaK, GetPairEvalue, PredictEVDParametersSS18,
 *                CalculateEvalueSS18: cuss/CuBatchSS_SS18.cuh:119-453;
 *                networks cuss/CuBatchSS_NN.cuh:37-202; weights
 *                cupro/CuDeviceMemory.cu:43-77)
 * The score grid keeps the reference's 21 x 32 layout (row y holds the
 * scaled scores around y-10, rows 0 and 20 clipped, row 10 one entry short)
 * because the K iteration reads it through that layout, including its row
 * wrap-around.  The reference's K iteration reads neighbouring cells while
 * other threads overwrite them (no barrier between the reads and the write
 * of P[j]); this restatement reads P[j-1] entirely before writing P[j].
 * Special functions: log of the regularized incomplete beta and of the
 * regularized upper incomplete gamma by the continued fractions of Lentz's
 * method (Press et al., Numerical Recipes, 3rd ed., 6.2 and 6.4) with the
 * series for the lower gamma, in place of the reference's single-precision
 * Cephes-derived routines (cusf/CuSF_betainc.cuh, cusf/CuSF_gammainc.cuh).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

/* the continued fractions' underflow guard and tolerance (the single-precision
 * twin, orc_stats32.c, sets its own) */
#ifndef ORC_FPMIN
#define ORC_FPMIN 1e-300
#define ORC_EPS 1e-15
#endif

#define NSC 21
#define SCALE 32
#define SSCMIN (-320)
#define NSSC 641
#define LOG_SP_MIN (-8.7336545e+01)
#define LOG_SP_MAX (8.8722839e+01)

/* NN weights and SS18 adjustment parameters (cupro/CuDeviceMemory.cu:43-77) */
static const float NNW[93] = {
    3.1179029f, 0.2732032f, -1.7865388f, -2.7370650f, -2.4113938f,
    2.8802990f, -0.0845794f, 0.7335855f, -0.4155596f, -3.5943765f,
    -2.3533863f, -0.1646748f, 2.6470928f, 1.6489798f, -0.7481717f,
    -0.1763256f, 0.3270687f, 0.5918245f, 0.4900585f, 0.0253580f,
    2.0547640f, 1.4051396f, -1.7874640f, 2.0048461f, 0.6947985f,
    -0.0605839f, 0.0533123f, -1.0046299f, 0.9682557f, 0.0f,
    -0.0173019f, -0.0051053f, 0.6215832f, -0.5038284f, 0.0f,
    0.0600411f, -0.0527767f, 1.0036328f, -0.9667071f, 0.0f,
    -5.9555713f, -3.1437676f, 1.8754493f, 3.4776044f, 0.0f,
    2.8086869f, -0.5250554f, 0.3085814f, 0.5244239f, 2.4166220f,
    -4.4935289f, -11.5416831f, 0.4601384f, 4.1707612f, 0.0f,
    1.4710020f, 0.7746900f, 0.5248285f, -2.2621661f, 0.0f,
    -1.4710009f, -0.7746948f, -0.5248293f, 2.2621667f, 0.0f,
    4.8137710f, 1.0000337f, 2.0837171f, -5.3104535f, 0.0f,
    2.5137293f, 4.2316344f, -0.8620832f, 0.8620837f, 3.1901578f,
    0.1f, 12.0f, 1.0f, 4.0f, 0.05f, 1.0f, 0.1f, 0.35f, 0.65f,
    0.6f, 16.0f, 1.0f, 12.0f, 0.3f, -1.7f, 0.1f, 0.45f, 0.65f};

/* ---- histogram ---------------------------------------------------------- */
void orc_histogram(const float* sc, int64_t ldsc, int64_t col, int lq, int lt, const char* qaa, const char* taa,
                   double counts[NSSC]) {
  memset(counts, 0, sizeof(double) * NSSC);
  for (int y = 0; y < lq; y++) {
    if (qaa[y] == 'X') continue;
    for (int x = 0; x < lt; x++) {
      if (taa[x] == 'X') continue;
      float s = sc[(int64_t)y * ldsc + col + x];
      if (s < -10.0f) s = -10.0f;
      else if (s > 10.0f) s = 10.0f;
      const int b = (int)rintf(s * 32.0f) - SSCMIN;
      counts[b] += 1.0;
    }
  }
}

/* grid layout of the scaled scores (CalcStatParamsDynPlmProfile :209-220) */
static void grid_bounds(int y, int* left, int* right) {
  const int u = y - 10;
  int l = u * SCALE - (SCALE >> 1) - SSCMIN;
  int r = u * SCALE + (SCALE >> 1) - SSCMIN;
  if (u <= 0) { l++; if (l < 0) l = 0; }
  if (0 <= u) { r--; if (NSSC - 1 < r) r = NSSC - 1; }
  *left = l; *right = r;
}

static void lambda_eq(const double G[NSC][SCALE], const int left[NSC], double x, double mn, double mxs, double* f,
                      double* df) {
  double lf = 0.0, ldf = 0.0;
  for (int y = 0; y < NSC; y++)
    for (int k = 0; k < SCALE; k++) {
      double sc = left[y] + k + SSCMIN;
      if (mn <= sc && sc <= mxs) {
        sc -= mxs;
        const double e = x * sc;
        if (LOG_SP_MIN < e) {
          const double t = exp(e) * G[y][k];
          lf += t;
          ldf += t * sc;
        }
      }
    }
  const double e = x * -mxs;
  if (LOG_SP_MIN < e) {
    const double t = exp(e);
    lf -= t;
    ldf += t * mxs;
  }
  *f = lf;
  *df = ldf;
}

/* FindLmbdRootByNRandBisection (CuBatchSS_lambda.cuh:132-226) */
static double find_lambda(const double G[NSC][SCALE], const int left[NSC], double mn, double mxs) {
  const double x1 = 1.0e-5, x2 = 10.0, acc = 1.0e-4;
  double f, df, fl, dx, dxold, xh, xl, rts;
  lambda_eq(G, left, x1, mn, mxs, &fl, &df);
  lambda_eq(G, left, x2, mn, mxs, &f, &df);
  if ((fl > 0.0 && f > 0.0) || (fl < 0.0 && f < 0.0)) return -1.0;
  if (fl < 0.0) { xl = x1; xh = x2; } else { xh = x1; xl = x2; }
  rts = 0.5 * (x1 + x2);
  dxold = fabs(x2 - x1);
  dx = dxold;
  lambda_eq(G, left, rts, mn, mxs, &f, &df);
  int n;
  for (n = 0; n < 32; n++) {
    int stop = 0;
    if ((((rts - xh) * df - f) * ((rts - xl) * df - f) > 0.0) || (fabs(2.0 * f) > fabs(dxold * df))) {
      dxold = dx;
      dx = 0.5 * (xh - xl);
      rts = xl + dx;
      if (xl == rts) stop = 1;
    } else {
      dxold = dx;
      dx = f / df;
      const double t = rts;
      rts -= dx;
      if (t == rts) stop = 1;
    }
    if (fabs(dx) < acc) stop = 1;
    if (stop) break;
    lambda_eq(G, left, rts, mn, mxs, &f, &df);
    if (f < 0.0) xl = rts; else xh = rts;
  }
  if (n == 32) return -1.0;
  return rts;
}

/* CalcStatParamsDynPlmProfile: out = [min, max, E, H, lambda, K] */
void orc_stat_params(const double counts[NSSC], double out[6]) {
  double G[NSC][SCALE], pu[NSC];
  int left[NSC], right[NSC];
  double total = 0.0;
  for (int b = 0; b < NSSC; b++) total += counts[b];
  for (int y = 0; y < NSC; y++) {
    grid_bounds(y, &left[y], &right[y]);
    double cnt = 0.0;
    for (int k = 0; k < SCALE; k++) {
      G[y][k] = (left[y] + k <= right[y]) ? counts[left[y] + k] : 0.0;
      cnt += G[y][k];
    }
    pu[y] = cnt / total;
  }
  for (int y = 0; y < NSC; y++)
    for (int k = 0; k < SCALE; k++) G[y][k] = G[y][k] / total;
  /* expected, min, max of scaled scores */
  double E = 0.0, mn = 0.0, mxs = 0.0;
  for (int y = 0; y < NSC; y++)
    for (int k = 0; k < SCALE; k++) {
      const double sc = left[y] + k + SSCMIN;
      E += sc * G[y][k];
      const double v = G[y][k] ? sc : 0.0;
      if (v < mn) mn = v;
      if (v > mxs) mxs = v;
    }
  double lambda = -1.0;
  if (!(0.0 <= E || mxs <= 0.0)) lambda = find_lambda(G, left, mn, mxs);
  double H = -1.0;
  if (lambda > 0.0) {
    H = 0.0;
    for (int y = 0; y < NSC; y++)
      for (int k = 0; k < SCALE; k++) {
        const double sc = left[y] + k + SSCMIN;
        if (mn <= sc && sc <= mxs) {
          const double yy = lambda * sc;
          if (LOG_SP_MIN < yy && yy < LOG_SP_MAX - 3.0) H += exp(yy) * yy * G[y][k];
        }
      }
  }
  for (int k = 0; k < SCALE; k++) G[10][k] = 1.0;
  lambda *= SCALE;
  E /= SCALE;
  /* min/max of unscaled scores */
  double umn = 0.0, umx = 0.0;
  for (int y = 0; y < NSC; y++) {
    const double v = pu[y] ? (double)(y - 10) : 0.0;
    if (v < umn) umn = v;
    if (v > umx) umx = v;
  }
  double K = -1.0;
  if (!(0.0 <= E || lambda <= 0.0 || H <= 0.0)) {
    if (umn == -1.0 && umx == 1.0) {
      const double a = pu[9], b = pu[11];
      K = (a - b) * (a - b) / a;
    } else if (umn == -1.0 || umx == 1.0) {
      if (umx != 1.0) K = lambda / H * (1.0 - exp(-lambda)) * (E * E);
      else K = H / lambda * (1.0 - exp(-lambda));
    } else {
      /* CalcKSigma / CalcProbOfAlnLengthj on the 21x32 grid */
      const int minu = (int)umn, maxu = (int)umx, range = (int)(umx - umn + 1.0);
      double sigma = 0.0, Pj = 1.0;
      int minscore = minu, maxscore = maxu;
      for (int j = 1; j <= 32 && 1.0e-4 < Pj; j++) {
        double NG[NSC][SCALE];
        Pj = 0.0;
        for (int y = 0; y < NSC; y++)
          for (int k = 0; k < SCALE; k++) {
            double pji = 0.0, pj = 0.0;
            const int sc = left[y] + k + SSCMIN;
            if (minscore <= sc && sc <= maxscore && -32 <= sc && sc <= 32) {
              const double etols = sc < 0 ? lambda * (double)sc : 0.0;
              if (LOG_SP_MIN < etols) {
                const int mnu = (maxscore - sc < range) ? maxu - (maxscore - sc) : minu;
                const int mxu = (sc - minscore < range) ? minu + (sc - minscore) : maxu;
                for (int i = mnu; i <= mxu; i++) {
                  int yy = y, xx = k - i;
                  if (SCALE <= xx) {
                    xx -= SCALE;
                    if (yy + 1 < NSC) yy++; else continue;
                  } else if (xx < 0) {
                    xx += SCALE;
                    if (0 < yy) yy--; else continue;
                  }
                  pji += G[yy][xx] * pu[i + 10];
                }
                pj = pji * exp(etols);
              }
            }
            NG[y][k] = pji;
            Pj += pj;
          }
        memcpy(G, NG, sizeof G);
        sigma += Pj / j;
        minscore += minu;
        maxscore += maxu;
      }
      K = lambda * exp(-2.0 * sigma) / (H * (1.0 - exp(-lambda)));
    }
  }
  out[0] = umn; out[1] = umx; out[2] = E; out[3] = H; out[4] = lambda; out[5] = K;
}

/* ---- special functions ------------------------------------------------------ */
/* continued fraction for the incomplete beta (Lentz) */
static double betacf(double a, double b, double x) {
  const double FPMIN = ORC_FPMIN, EPS = ORC_EPS;
  const double qab = a + b, qap = a + 1.0, qam = a - 1.0;
  double c = 1.0, d = 1.0 - qab * x / qap;
  if (fabs(d) < FPMIN) d = FPMIN;
  d = 1.0 / d;
  double h = d;
  for (int m = 1; m < 10000; m++) {
    const int m2 = 2 * m;
    double aa = m * (b - m) * x / ((qam + m2) * (a + m2));
    d = 1.0 + aa * d; if (fabs(d) < FPMIN) d = FPMIN;
    c = 1.0 + aa / c; if (fabs(c) < FPMIN) c = FPMIN;
    d = 1.0 / d;
    h *= d * c;
    aa = -(a + m) * (qab + m) * x / ((a + m2) * (qap + m2));
    d = 1.0 + aa * d; if (fabs(d) < FPMIN) d = FPMIN;
    c = 1.0 + aa / c; if (fabs(c) < FPMIN) c = FPMIN;
    d = 1.0 / d;
    const double del = d * c;
    h *= del;
    if (fabs(del - 1.0) < EPS) break;
  }
  return h;
}

/* log I_x(a,b), the log of the regularized incomplete beta function */
double orc_lbetainc(double a, double b, double x) {
  if (x <= 0.0) return -999.0;
  if (x >= 1.0) return 0.0;
  const double lbt = lgamma(a + b) - lgamma(a) - lgamma(b) + a * log(x) + b * log1p(-x);
  if (x < (a + 1.0) / (a + b + 2.0)) return lbt + log(betacf(a, b, x)) - log(a);
  const double c = exp(lbt + log(betacf(b, a, 1.0 - x)) - log(b));
  return log1p(-c);
}

/* log Q(a,x), the log of the regularized upper incomplete gamma function */
double orc_lgammaincq(double a, double x) {
  if (x <= 0.0) return 0.0;
  const double gln = lgamma(a);
  if (x < a + 1.0) {
    double ap = a, del = 1.0 / a, sum = del;
    for (int n = 0; n < 100000; n++) {
      ap += 1.0;
      del *= x / ap;
      sum += del;
      if (fabs(del) < fabs(sum) * 1e-16) break;
    }
    const double P = sum * exp(-x + a * log(x) - gln);
    return log1p(-P);
  }
  const double FPMIN = ORC_FPMIN;
  double b = x + 1.0 - a, c = 1.0 / FPMIN, d = 1.0 / b, h = d;
  for (int i = 1; i < 100000; i++) {
    const double an = -i * (i - a);
    b += 2.0;
    d = an * d + b; if (fabs(d) < FPMIN) d = FPMIN;
    c = b + an / c; if (fabs(c) < FPMIN) c = FPMIN;
    d = 1.0 / d;
    const double del = d * c;
    h *= del;
    if (fabs(del - 1.0) < 1e-16) break;
  }
  return -x + a * log(x) - gln + log(h);
}

/* ---- SS18 significance ------------------------------------------------------
 * Constants of cuss/CuBatchSS_SS18.cuh:17-41 and cuss/CuBatchSS_com.h:39; the
 * reference's values are single-precision literals, used here at their float
 * values. */
#define SS18_MINMU 3.0
#define SS18_MINSCALE ((double)0.1f)
#define SS18_MAXSCALE 5.0
#define SS18_LOG001 ((double)-4.605170186f)
#define SS18_LOGINVNPAIRS ((double)-9.210340372f)
#define SS18_PSTS_FACTOR 100000.0
#define SS18_PSTS_EXP 1.5
#define SS18_NBD_SHAPE ((double)0.1f)
#define SS18_NBD_PROB ((double)0.757213f)
#define SS18_EBM_DFO2 ((double)1.346206f)
#define SS18_EBM_C ((double)1.485656f)
#define SS18_PSTS_MIN_LENPRODUCT 625.0
#define SS18_LN1K 6.90775527898213681510242167860
#define REF_LAMBDA ((double)1.116276f)
#define REF_K ((double)0.115812f)
#define GAP_LAMBDA ((double)0.561132f)
#define GAP_K ((double)0.001488f)

void orc_ss18_weights(float out[93]) { memcpy(out, NNW, sizeof NNW); }

static double sig2(double v) { return 2.0 / (1.0 + exp(-v)) - 1.0; }

/* one network: hidden unit h has bias w[5h] and input weights w[5h+1..5h+4];
   the output has bias w[20] and unit weights w[21..24]; activation
   2/(1+e^-x) - 1 at both layers (cuss/CuBatchSS_NN.cuh:37-148) */
static double nn_out(const double in[4], const float* w, double factor) {
  double acc = w[20];
  for (int h = 0; h < 4; h++) {
    double pre = w[5 * h];
    for (int i = 0; i < 4; i++) pre += in[i] * w[5 * h + 1 + i];
    acc += sig2(pre) * w[21 + h];
  }
  return sig2(acc) * factor;
}

/* which 0: mu from profile attributes (E), 1: mu and 2: scale from the
   compositional similarity (L) */
double orc_ss18_nn(int which, const double in[4]) {
  return nn_out(in, NNW + 25 * which, which == 2 ? 5.0 : 30.0);
}

/* FormatInputConfigurationE (cuss/CuBatchSS_NN.cuh:150-176) */
void orc_ss18_input_e(int lq, int lt, double qeno, double teno, double in[4]) {
  double e1 = qeno, e2 = teno;
  int n1 = lq, n2 = lt;
  if (e1 < e2) { double t = e1; e1 = e2; e2 = t; int u = n1; n1 = n2; n2 = u; }
  else if (e1 == e2 && n1 < n2) { int u = n1; n1 = n2; n2 = u; }
  in[0] = (double)0.05f * e1; in[1] = log((double)n1) / SS18_LN1K;
  in[2] = (double)0.05f * e2; in[3] = log((double)n2) / SS18_LN1K;
}

/* FormatInputConfigurationL (cuss/CuBatchSS_NN.cuh:178-202) */
void orc_ss18_input_l(int lq, int lt, double lambda, double in[4]) {
  double lm = rint(lambda * 10.0) / 10.0;
  if (lm <= 0.0) lm = 0.1;
  if (10.0 < lm) lm = 10.0;
  int n1 = lq, n2 = lt;
  if (n1 < n2) { int u = n1; n1 = n2; n2 = u; }
  in[0] = (double)0.2f * lm; in[1] = log((double)n1) / SS18_LN1K; in[2] = log((double)n2) / SS18_LN1K; in[3] = 0.0;
}

/* PredictEVDParametersSS18 (cuss/CuBatchSS_SS18.cuh:119-233): the EVD's mu
   and scale from the three networks and the adjustment parameters of the
   model's variant (9 values each, CuBatchProcessing_com.cuh:61-72; variant 1
   from weight 75, variant 2 from weight 84) */
void orc_ss18_predict(int ssemodel, int lq, int lt, double qeno, double teno, double lambda, double* mu,
                      double* scale) {
  double in[4];
  orc_ss18_input_e(lq, lt, qeno, teno, in);
  double emu = orc_ss18_nn(0, in);
  orc_ss18_input_l(lq, lt, lambda, in);
  double lmu = orc_ss18_nn(1, in);
  double lsc = orc_ss18_nn(2, in);
  const float* adj = ssemodel == 2 ? NNW + 84 : NNW + 75;
  if (emu < SS18_MINMU) emu = SS18_MINMU;
  if (lmu < SS18_MINMU) lmu = SS18_MINMU;
  if (lsc < SS18_MINSCALE) lsc = SS18_MINSCALE;
  lsc = adj[6] * (exp(lsc) - 1.0);
  if (lsc < SS18_MINSCALE) lsc = SS18_MINSCALE;
  if (SS18_MAXSCALE < lsc) lsc = SS18_MAXSCALE;
  double scaleE = adj[4] * emu + adj[5];
  if (scaleE < SS18_MINSCALE) scaleE = SS18_MINSCALE;
  if (SS18_MAXSCALE < scaleE) scaleE = SS18_MAXSCALE;
  emu = adj[0] * emu + adj[1];
  lmu = adj[2] * lmu + adj[3];
  double alpha = adj[7], ibeta = adj[8];
  if (0.0 < lambda) {
    const double r = 1.0 / lambda;
    alpha -= alpha * (double)0.1f * (r < 10.0 ? r : 10.0);
    ibeta += (1.0 - ibeta) * (double)0.1f * (r < 10.0 ? r : 10.0);
  }
  *mu = alpha * emu + (1.0 - alpha) * lmu;
  *scale = (1.0 - ibeta) * scaleE + ibeta * lsc;
}

/* CalculateEvalueSS18 (cuss/CuBatchSS_SS18.cuh:255-351), step 1: log p-value
   of the score under the EVD (Gumbel) with location mu and scale,
   P(S >= s) = 1 - exp(-exp(-(s - mu')/scale)), mu' = mu + scale log(K/Kref)
   when K > Kref, for variant 2 with the expected count divided by 10^4; an
   expected count below 0.01 is taken as the p-value itself */
double orc_ss18_score_logp(int ssemodel, double score, double K, double mu, double scale) {
  double logK = mu / scale;
  if (K > REF_K && 0.0 < K && K < 1.0 && 0.0 < REF_K && REF_K < 1.0) logK += log(K / REF_K);
  double logexp = logK - score / scale;
  if (ssemodel == 2) logexp += SS18_LOGINVNPAIRS;
  double logp = logexp;
  if (SS18_LOG001 <= logexp) {
    if (LOG_SP_MAX - 3.0 < logexp) logp = 0.0;
    else {
      const double t = exp(logexp);
      if (LOG_SP_MAX - 3.0 < t) logp = 0.0;
      else logp = log(1.0 - exp(-t));
    }
  }
  return logp;
}

/* step 2: log p-value of the positive-score count, the negative binomial
   tail I_x(n + 1, 0.1) at x = 0.757213 of the normalised count
   n = rint(10^5 psts / max(Lq Lt, 625)^1.5) */
double orc_ss18_psts_logp(int lq, int lt, double psts) {
  const double l1l2 = (double)(float)((uint32_t)lq * (uint32_t)lt);
  const double lp = l1l2 < SS18_PSTS_MIN_LENPRODUCT ? SS18_PSTS_MIN_LENPRODUCT : l1l2;
  const double pn = rint(SS18_PSTS_FACTOR * psts / pow(lp, SS18_PSTS_EXP));
  return orc_lbetainc(pn + 1.0, SS18_NBD_SHAPE, SS18_NBD_PROB);
}

/* step 3: the two combined by Brown's method: chi-square-like statistic
   (-log p1 - log p2) / c with 2 k degrees of freedom, p = Q(k, .) */
double orc_ss18_combine(double logp_score, double logp_psts) {
  return orc_lgammaincq(SS18_EBM_DFO2, (-logp_score - logp_psts) / SS18_EBM_C);
}

/* step 4: p-value to log E-value (E = -log(1 - p); p itself below 0.01; 77
   for p = 1), scaled from the pair's Lq Lt to the search space */
double orc_ss18_logevalue(double logp, double searchspace, int lq, int lt) {
  const double l1l2 = (double)(float)((uint32_t)lq * (uint32_t)lt);
  double logexp = logp;
  if (SS18_LOG001 <= logp) {
    const double p = exp(logp);
    logexp = 77.0;
    if (p < 1.0) logexp = log(-log(1.0 - p));
  }
  return logexp + log(searchspace / l1l2);
}

/* out: [lmbdest, kest, logE, logEpA]
   EstimateGappedLambdaK and GetPairEvalue (cuss/CuBatchSS_SS18.cuh:384-453),
   then the log E-value of the model: SSEMODEL 0 CalcSignificanceObsDynPlmProfile
   (cuss/CuBatchSS_dynplm.cu:594-660, CalculateEvalueObs), 1 and 2
   CalcSignificanceSS18DynPlmProfile (cuss/CuBatchSS_dynplm.cu:447-580) */
void orc_significance(int ssemodel, int lq, int lt, double qeno, double teno, double searchspace, double score,
                      double psts, double lambda, double K, double out[4]) {
  out[0] = 0.0; out[1] = 0.0;
  if (0.0 < lambda) out[0] = lambda * (GAP_LAMBDA / REF_LAMBDA);
  if (0.0 < K) out[1] = K * (GAP_K / REF_K);
  const double l1l2 = (double)(float)((uint32_t)lq * (uint32_t)lt);
  if (lambda <= 0.0 || K <= 0.0) out[3] = log(REF_K * l1l2) - REF_LAMBDA * score;
  else out[3] = log(K * l1l2) - lambda * score;
  if (ssemodel == 0) {
    out[2] = log(GAP_K * searchspace) - GAP_LAMBDA * score;
    return;
  }
  double mu, scale;
  orc_ss18_predict(ssemodel, lq, lt, qeno, teno, lambda, &mu, &scale);
  if (mu <= 0.0 || scale <= 0.0) {
    out[2] = log(GAP_K * searchspace) - GAP_LAMBDA * score;
    return;
  }
  const double lps = orc_ss18_score_logp(ssemodel, score, K, mu, scale);
  const double lpp = orc_ss18_psts_logp(lq, lt, psts);
  out[2] = orc_ss18_logevalue(orc_ss18_combine(lps, lpp), searchspace, lq, lt);
}
