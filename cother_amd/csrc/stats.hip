// stats.hip: alignment statistics and E-values for profiles passed to phase 2.
//
// k_hist  : one workgroup per passed profile builds the 641-bin histogram of
//           rint(32*clamp(S,-10,10)) over the profile's Lq x Lt scores in LDS
//           (integer counts; masked 'X' positions skipped), the work of
//           CalcScoreProbsXuninfDynPlmProfile (cuss/CuBatchSS_dynplm.cu:134-182).
// k_stats : one wavefront per passed profile computes in double precision the
//           parameters of CalcStatParamsDynPlmProfile (cuss/CuBatchSS_dynplm.cu:
//           193-430): expected score, min/max, lambda by Newton-Raphson with
//           bisection, relative entropy, K by the sigma series on the
//           reference's 21x32 score grid (the 672 grid cells of one iteration
//           are spread over the 64 lanes; each iteration's P_j sum is reduced
//           in grid order), then the SS18 E-value of
//           CalcSignificanceSS18DynPlmProfile (cuss/CuBatchSS_dynplm.cu:447-580,
//           cuss/CuBatchSS_SS18.cuh:119-453) with the two small networks of
//           cuss/CuBatchSS_NN.cuh:37-202.  Special functions: log regularized
//           incomplete beta and upper incomplete gamma by Lentz continued
//           fractions and the lower-gamma series (Numerical Recipes 3rd ed.,
//           6.2, 6.4), in double precision.
#include <hip/hip_runtime.h>

#include "cth_batch.h"
#include "cth_dev.h"
#include "cth_dp.h"
#include "cth_stats.h"

namespace {

constexpr int kNSC = 21, kScale = 32, kSscMin = -320, kNssc = 641;
constexpr double kLogSpMin = -8.7336545e+01, kLogSpMax = 8.8722839e+01;

// NN weights and SS18 adjustment parameters (cupro/CuDeviceMemory.cu:43-77)
__constant__ float c_nnw[93] = {
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

// The 641-bin score histogram of query qp against profile p (scores from
// scores[colp], row stride ldsc) into the workgroup's LDS array h.
__device__ __forceinline__ void hist_fill(const cth_dpset& q, int qp, const cth_dpset& db, int p, int64_t colp,
                                          int64_t ldsc, const float* __restrict__ scores, int* h) {
  const int lq = q.len[qp], lt = db.len[p];
  const char* qaa = q.aa + q.dist[qp];
  const char* taa = db.aa + db.dist[p];
  for (int b = threadIdx.x; b < kNssc; b += blockDim.x) h[b] = 0;
  __syncthreads();
  const int64_t ncell = (int64_t)lq * lt;
  for (int64_t c = threadIdx.x; c < ncell; c += blockDim.x) {
    const int y = (int)(c / lt), x = (int)(c - (int64_t)y * lt);
    if (qaa[y] == 'X' || taa[x] == 'X') continue;
    float s = scores[(int64_t)y * ldsc + colp + x];
    if (s < -10.0f) s = -10.0f;
    else if (s > 10.0f) s = 10.0f;
    atomicAdd(&h[(int)rintf(s * 32.0f) - kSscMin], 1);
  }
  __syncthreads();
}

__global__ __launch_bounds__(256) void k_hist(const cth_dpset q, int qp, const cth_dpset db, int p0, int64_t col0,
                                              int64_t ldsc, const float* __restrict__ scores,
                                              const int* __restrict__ plist, int* __restrict__ hist) {
  __shared__ int h[kNssc];
  const int k = blockIdx.x;
  const int p = plist[k];
  hist_fill(q, qp, db, p, db.dist[p] - col0, ldsc, scores, h);
  for (int b = threadIdx.x; b < kNssc; b += blockDim.x) hist[(int64_t)k * kNssc + b] = h[b];
}

__device__ void grid_bounds(int y, int* left, int* right) {
  const int u = y - 10;
  int l = u * kScale - (kScale >> 1) - kSscMin;
  int r = u * kScale + (kScale >> 1) - kSscMin;
  if (u <= 0) { l++; if (l < 0) l = 0; }
  if (0 <= u) { r--; if (kNssc - 1 < r) r = kNssc - 1; }
  *left = l; *right = r;
}

__device__ void lambda_eq(const double* G, const int* left, double x, double mn, double mxs, double* f, double* df) {
  double lf = 0.0, ldf = 0.0;
  for (int y = 0; y < kNSC; y++)
    for (int k = 0; k < kScale; k++) {
      double sc = left[y] + k + kSscMin;
      if (mn <= sc && sc <= mxs) {
        sc -= mxs;
        const double e = x * sc;
        if (kLogSpMin < e) {
          const double t = exp(e) * G[y * kScale + k];
          lf += t;
          ldf += t * sc;
        }
      }
    }
  const double e = x * -mxs;
  if (kLogSpMin < e) {
    const double t = exp(e);
    lf -= t;
    ldf += t * mxs;
  }
  *f = lf;
  *df = ldf;
}

__device__ double find_lambda(const double* G, const int* left, double mn, double mxs) {
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
    bool stop = false;
    if ((((rts - xh) * df - f) * ((rts - xl) * df - f) > 0.0) || (fabs(2.0 * f) > fabs(dxold * df))) {
      dxold = dx;
      dx = 0.5 * (xh - xl);
      rts = xl + dx;
      if (xl == rts) stop = true;
    } else {
      dxold = dx;
      dx = f / df;
      const double t = rts;
      rts -= dx;
      if (t == rts) stop = true;
    }
    if (fabs(dx) < acc) stop = true;
    if (stop) break;
    lambda_eq(G, left, rts, mn, mxs, &f, &df);
    if (f < 0.0) xl = rts; else xh = rts;
  }
  return n == 32 ? -1.0 : rts;
}

__device__ double betacf(double a, double b, double x) {
  const double FPMIN = 1e-300, EPS = 1e-15;
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

__device__ double lbetainc(double a, double b, double x) {
  if (x <= 0.0) return -999.0;
  if (x >= 1.0) return 0.0;
  const double lbt = lgamma(a + b) - lgamma(a) - lgamma(b) + a * log(x) + b * log1p(-x);
  if (x < (a + 1.0) / (a + b + 2.0)) return lbt + log(betacf(a, b, x)) - log(a);
  const double c = exp(lbt + log(betacf(b, a, 1.0 - x)) - log(b));
  return log1p(-c);
}

__device__ double lgammaincq(double a, double x) {
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
    return log1p(-(sum * exp(-x + a * log(x) - gln)));
  }
  const double FPMIN = 1e-300;
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

__device__ double sig2(double v) { return 2.0 / (1.0 + exp(-v)) - 1.0; }

// one network (cuss/CuBatchSS_NN.cuh:37-148): hidden unit h has bias w[5h]
// and input weights w[5h+1..5h+4], the output bias w[20] and unit weights
// w[21..24]; activation 2/(1+e^-x) - 1
__device__ double nn_out(const double in[4], const float* w, double factor) {
  double acc = w[20];
  for (int h = 0; h < 4; h++) {
    double pre = w[5 * h];
    for (int i = 0; i < 4; i++) pre += in[i] * w[5 * h + 1 + i];
    acc += sig2(pre) * w[21 + h];
  }
  return sig2(acc) * factor;
}

// constants of cuss/CuBatchSS_SS18.cuh:17-41 at the values of the reference's
// single-precision literals
constexpr double kSS18MinScale = (double)0.1f, kSS18Log001 = (double)-4.605170186f,
                 kSS18LogInvNPairs = (double)-9.210340372f, kSS18NbdShape = (double)0.1f,
                 kSS18NbdProb = (double)0.757213f, kSS18EbmDfo2 = (double)1.346206f,
                 kSS18EbmC = (double)1.485656f;

// significance of one alignment, out = [lmbdest, kest, logE, logEpA]: SSEMODEL
// 0 observed statistics, 1 and 2 the SS18 model and its variant
__device__ void significance(int ssemodel, int lq, int lt, double qeno, double teno, double searchspace, double score,
                             double psts, double lambda, double K, double* out) {
  const double lref = CTH_UNGAP_LAMBDA, kref = CTH_UNGAP_K, lg = CTH_GAPPED_LAMBDA, kg = CTH_GAPPED_K;
  const double ln1k = 6.90775527898213681510242167860;
  out[0] = 0.0; out[1] = 0.0;
  if (0.0 < lambda) out[0] = lambda * (lg / lref);
  if (0.0 < K) out[1] = K * (kg / kref);
  const double l1l2 = (double)(float)((uint32_t)lq * (uint32_t)lt);
  if (lambda <= 0.0 || K <= 0.0) out[3] = log(kref * l1l2) - lref * score;
  else out[3] = log(K * l1l2) - lambda * score;
  if (ssemodel == 0) {
    // observed statistics: the gapped reference parameters alone
    // (CalcSignificanceObsDynPlmProfile -> CalculateEvalueObs,
    // cuss/CuBatchSS_dynplm.cu:594-660, cuss/CuBatchSS_SS18.cuh:355-381)
    out[2] = log(kg * searchspace) - lg * score;
    return;
  }
  // PredictEVDParametersSS18 (cuss/CuBatchSS_SS18.cuh:119-233)
  double in[4];
  {
    double e1 = qeno, e2 = teno;
    int n1 = lq, n2 = lt;
    if (e1 < e2) { double t = e1; e1 = e2; e2 = t; int u = n1; n1 = n2; n2 = u; }
    else if (e1 == e2 && n1 < n2) { int u = n1; n1 = n2; n2 = u; }
    in[0] = (double)0.05f * e1; in[1] = log((double)n1) / ln1k;
    in[2] = (double)0.05f * e2; in[3] = log((double)n2) / ln1k;
  }
  double emu = nn_out(in, c_nnw, 30.0);
  {
    double lm = rint(lambda * 10.0) / 10.0;
    if (lm <= 0.0) lm = 0.1;
    if (10.0 < lm) lm = 10.0;
    int n1 = lq, n2 = lt;
    if (n1 < n2) { int u = n1; n1 = n2; n2 = u; }
    in[0] = (double)0.2f * lm; in[1] = log((double)n1) / ln1k; in[2] = log((double)n2) / ln1k; in[3] = 0.0;
  }
  double lmu = nn_out(in, c_nnw + 25, 30.0);
  double lsc = nn_out(in, c_nnw + 50, 5.0);
  const float* adj = ssemodel == 2 ? c_nnw + 84 : c_nnw + 75;
  if (emu < 3.0) emu = 3.0;
  if (lmu < 3.0) lmu = 3.0;
  if (lsc < kSS18MinScale) lsc = kSS18MinScale;
  lsc = adj[6] * (exp(lsc) - 1.0);
  if (lsc < kSS18MinScale) lsc = kSS18MinScale;
  if (5.0 < lsc) lsc = 5.0;
  double scaleE = adj[4] * emu + adj[5];
  if (scaleE < kSS18MinScale) scaleE = kSS18MinScale;
  if (5.0 < scaleE) scaleE = 5.0;
  emu = adj[0] * emu + adj[1];
  lmu = adj[2] * lmu + adj[3];
  double alpha = adj[7], ibeta = adj[8];
  if (0.0 < lambda) {
    const double r = 1.0 / lambda;
    alpha -= alpha * (double)0.1f * (r < 10.0 ? r : 10.0);
    ibeta += (1.0 - ibeta) * (double)0.1f * (r < 10.0 ? r : 10.0);
  }
  const double mu = alpha * emu + (1.0 - alpha) * lmu;
  const double scale = (1.0 - ibeta) * scaleE + ibeta * lsc;
  // CalculateEvalueSS18 (cuss/CuBatchSS_SS18.cuh:255-351)
  if (mu <= 0.0 || scale <= 0.0) {
    out[2] = log(kg * searchspace) - lg * score;
    return;
  }
  double logK = mu / scale;
  if (K > kref && 0.0 < K && K < 1.0 && 0.0 < kref && kref < 1.0) logK += log(K / kref);
  double logexp = logK - score / scale;
  if (ssemodel == 2) logexp += kSS18LogInvNPairs;
  double logp = logexp;
  if (kSS18Log001 <= logexp) {
    if (kLogSpMax - 3.0 < logexp) logp = 0.0;
    else {
      const double t = exp(logexp);
      if (kLogSpMax - 3.0 < t) logp = 0.0;
      else logp = log(1.0 - exp(-t));
    }
  }
  const double lp = l1l2 < 625.0 ? 625.0 : l1l2;
  const double pn = rint(100000.0 * psts / pow(lp, 1.5));
  const double logpp = lbetainc(pn + 1.0, kSS18NbdShape, kSS18NbdProb);
  logp = lgammaincq(kSS18EbmDfo2, (-logp - logpp) / kSS18EbmC);
  logexp = logp;
  if (kSS18Log001 <= logp) {
    const double p = exp(logp);
    logexp = 77.0;
    if (p < 1.0) logexp = log(-log(1.0 - p));
  }
  logexp += log(searchspace / l1l2);
  out[2] = logexp;
}

// The statistics of one passed profile p (one wavefront): parameters from
// its histogram hc, then the significance of its phase-1 record r, into the
// CTH_STATS_NOUT doubles at o.
__device__ __forceinline__ void stats_one(const cth_dpset& q, int qp, const cth_dpset& db, int p, const int* hc,
                                          double searchspace, int ssemodel, const cth_dpres& r, double* o) {
  __shared__ double G[kNSC * kScale], NG[kNSC * kScale], PJ[kNSC * kScale];
  __shared__ double pu[kNSC];
  __shared__ int left[kNSC];
  __shared__ double sh[8];  // umn, umx, E, H, lambda, K, total
  const int lane = threadIdx.x;
  if (lane < kNSC) {
    int l, r;
    grid_bounds(lane, &l, &r);
    left[lane] = l;
    double cnt = 0.0;
    for (int x = 0; x < kScale; x++) {
      const double v = (l + x <= r) ? (double)hc[l + x] : 0.0;
      G[lane * kScale + x] = v;
      cnt += v;
    }
    pu[lane] = cnt;
  }
  __syncthreads();
  if (lane == 0) {
    double total = 0.0;
    for (int b = 0; b < kNssc; b++) total += (double)hc[b];
    for (int y = 0; y < kNSC; y++) pu[y] = pu[y] / total;
    for (int c = 0; c < kNSC * kScale; c++) G[c] = G[c] / total;
    double E = 0.0, mn = 0.0, mxs = 0.0;
    for (int y = 0; y < kNSC; y++)
      for (int x = 0; x < kScale; x++) {
        const double sc = left[y] + x + kSscMin;
        E += sc * G[y * kScale + x];
        const double v = G[y * kScale + x] ? sc : 0.0;
        if (v < mn) mn = v;
        if (v > mxs) mxs = v;
      }
    double lambda = -1.0;
    if (!(0.0 <= E || mxs <= 0.0)) lambda = find_lambda(G, left, mn, mxs);
    double H = -1.0;
    if (lambda > 0.0) {
      H = 0.0;
      for (int y = 0; y < kNSC; y++)
        for (int x = 0; x < kScale; x++) {
          const double sc = left[y] + x + kSscMin;
          if (mn <= sc && sc <= mxs) {
            const double yy = lambda * sc;
            if (kLogSpMin < yy && yy < kLogSpMax - 3.0) H += exp(yy) * yy * G[y * kScale + x];
          }
        }
    }
    for (int x = 0; x < kScale; x++) G[10 * kScale + x] = 1.0;
    lambda *= kScale;
    E /= kScale;
    double umn = 0.0, umx = 0.0;
    for (int y = 0; y < kNSC; y++) {
      const double v = pu[y] ? (double)(y - 10) : 0.0;
      if (v < umn) umn = v;
      if (v > umx) umx = v;
    }
    sh[0] = umn; sh[1] = umx; sh[2] = E; sh[3] = H; sh[4] = lambda; sh[5] = -1.0;
  }
  __syncthreads();
  const double umn = sh[0], umx = sh[1], E = sh[2], H = sh[3], lambda = sh[4];
  double K = -1.0;
  if (!(0.0 <= E || lambda <= 0.0 || H <= 0.0)) {
    if (umn == -1.0 && umx == 1.0) {
      const double a = pu[9], b = pu[11];
      K = (a - b) * (a - b) / a;
    } else if (umn == -1.0 || umx == 1.0) {
      if (umx != 1.0) K = lambda / H * (1.0 - exp(-lambda)) * (E * E);
      else K = H / lambda * (1.0 - exp(-lambda));
    } else {
      const int minu = (int)umn, maxu = (int)umx, range = (int)(umx - umn + 1.0);
      double sigma = 0.0, Pj = 1.0;
      int minscore = minu, maxscore = maxu;
      for (int j = 1; j <= 32 && 1.0e-4 < Pj; j++) {
        for (int c = lane; c < kNSC * kScale; c += 64) {
          const int y = c / kScale, x = c - y * kScale;
          double pji = 0.0, pj = 0.0;
          const int sc = left[y] + x + kSscMin;
          if (minscore <= sc && sc <= maxscore && -32 <= sc && sc <= 32) {
            const double etols = sc < 0 ? lambda * (double)sc : 0.0;
            if (kLogSpMin < etols) {
              const int mnu = (maxscore - sc < range) ? maxu - (maxscore - sc) : minu;
              const int mxu = (sc - minscore < range) ? minu + (sc - minscore) : maxu;
              for (int i = mnu; i <= mxu; i++) {
                int yy = y, xx = x - i;
                if (kScale <= xx) {
                  xx -= kScale;
                  if (yy + 1 < kNSC) yy++; else continue;
                } else if (xx < 0) {
                  xx += kScale;
                  if (0 < yy) yy--; else continue;
                }
                pji += G[yy * kScale + xx] * pu[i + 10];
              }
              pj = pji * exp(etols);
            }
          }
          NG[c] = pji;
          PJ[c] = pj;
        }
        __syncthreads();
        for (int c = lane; c < kNSC * kScale; c += 64) G[c] = NG[c];
        if (lane == 0) {
          double s = 0.0;
          for (int c = 0; c < kNSC * kScale; c++) s += PJ[c];
          sh[6] = s;
        }
        __syncthreads();
        Pj = sh[6];
        sigma += Pj / j;
        minscore += minu;
        maxscore += maxu;
      }
      K = lambda * exp(-2.0 * sigma) / (H * (1.0 - exp(-lambda)));
    }
  }
  if (lane == 0) {
    double sig[4];
    significance(ssemodel, q.len[qp], db.len[p], (double)q.eno[qp], (double)db.eno[p], searchspace,
                 (double)r.score, (double)r.psts, lambda, K, sig);
    o[CTH_ST_LMBDEST] = sig[0]; o[CTH_ST_KEST] = sig[1]; o[CTH_ST_LOGE] = sig[2]; o[CTH_ST_LOGEPA] = sig[3];
    o[CTH_ST_MIN] = umn; o[CTH_ST_MAX] = umx; o[CTH_ST_E] = E; o[CTH_ST_H] = H;
    o[CTH_ST_LMBD] = lambda; o[CTH_ST_K] = K;
  }
}

__global__ __launch_bounds__(64) void k_stats(const cth_dpset q, int qp, const cth_dpset db,
                                              const int* __restrict__ plist, const int* __restrict__ prel,
                                              const cth_dpres* __restrict__ dpres,
                                              const int* __restrict__ hist, double searchspace, int ssemodel,
                                              double* __restrict__ out) {
  const int k = blockIdx.x;
  stats_one(q, qp, db, plist[k], hist + (int64_t)k * kNssc, searchspace, ssemodel, dpres[prel[k]],
            out + (int64_t)k * CTH_STATS_NOUT);
}

// Batched statistics over the *count passed pairs listed in list
// (persistent, one wavefront per pair at a time): histogram in LDS, then the
// parameters and significance against the pair's query's search space; out
// and dpres are per pair.
__global__ __launch_bounds__(64) void k_stats_batch(const cth_dpset q, const cth_dpset db,
                                                    const cth_bquery* __restrict__ bq,
                                                    const cth_bpair* __restrict__ pairs, const int* __restrict__ list,
                                                    const int* __restrict__ count, int64_t ldsc,
                                                    const float* __restrict__ scores,
                                                    const cth_dpres* __restrict__ dpres, int ssemodel,
                                                    double* __restrict__ out) {
  __shared__ int h[kNssc];
  const int nj = *count;
  for (int k = blockIdx.x; k < nj; k += gridDim.x) {
    const int i = list[k];
    const cth_bpair pr = pairs[i];
    const cth_bquery Q = bq[pr.b];
    hist_fill(q, Q.qp, db, pr.p, db.dist[pr.p], ldsc, scores + Q.off, h);
    stats_one(q, Q.qp, db, pr.p, h, Q.sspace, ssemodel, dpres[i], out + (int64_t)i * CTH_STATS_NOUT);
    __syncthreads();
  }
}

}  // namespace

// Statistics for the np2 passed profiles listed in plist (indices into the
// launch's profile range, relative to p0 as the DP results are); hist is
// scratch of np2*641 ints, out receives np2 x CTH_STATS_NOUT doubles.
extern "C" int cth_launch_stats(const cth_dpset* q, int qp, const cth_dpset* db, int p0, int64_t col0, int64_t ldsc,
                                const float* scores, const int* plist_abs, const int* plist_rel,
                                const cth_dpres* dpres, int np2, double searchspace, int ssemodel, int* hist,
                                double* out, hipStream_t st) {
  if (np2 <= 0) return 0;
  (void)p0;
  hipLaunchKernelGGL(k_hist, dim3(np2), dim3(256), 0, st, *q, qp, *db, p0, col0, ldsc, scores, plist_abs, hist);
  CTH_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_stats, dim3(np2), dim3(64), 0, st, *q, qp, *db, plist_abs, plist_rel, dpres, hist,
                     searchspace, ssemodel, out);
  CTH_HIP_CHECK(hipGetLastError());
  return 0;
}

// Batched statistics (k_stats_batch) on a persistent grid of `grid`
// wavefronts over the *count pairs of list; out: CTH_STATS_NOUT doubles per pair.
extern "C" int cth_launch_stats_batch(const cth_dpset* q, const cth_dpset* db, const cth_bquery* bq,
                                      const cth_bpair* pairs, const int* list, const int* count, int64_t ldsc,
                                      const float* scores, const cth_dpres* dpres, int ssemodel, double* out,
                                      int grid, hipStream_t st) {
  if (grid < 1) return 0;
  hipLaunchKernelGGL(k_stats_batch, dim3(grid), dim3(64), 0, st, *q, *db, bq, pairs, list, count, ldsc, scores, dpres,
                     ssemodel, out);
  CTH_HIP_CHECK(hipGetLastError());
  return 0;
}
