// synth.cpp: seeded synthetic profile sets in the reference's value
// conventions (query sets and template databases).
//
// Every value is produced as the reference's text profile round trip would
// give it: written by TextWriteProfileCondensed (libpro/srcpro/PMProfileModel.cpp
// :1856-2082) with scale 10000 and read back by TextReadProfileData
// (cupro/IOProfileModelTxt.cu:296-1030).  Distance lists are built by a
// restatement of adddist's list transform (adddist/adddist.cpp:269-651):
// for each position the 99 partners nearest in sequence, the main-diagonal
// band |i-j|<=DSTFREE dropped, and gap penalties DSTSEP6/DSTGAPL inserted.
//
// Generators (seeded, counter-based so profiles are independent of thread
// scheduling):
//  - target frequencies ~ Dirichlet(0.5 * Robinson background), mixed with a
//    1% background pseudocount (a makepro profile always carries
//    pseudocounts; this keeps every frequency >= 1/10000 after quantisation);
//  - transitions: M->(M,I,D) ~ Dir(8,1,1), I->(M,I) ~ Dir(1,1),
//    D->(M,D) ~ Dir(1,1);
//  - SS probabilities ~ Dir(1,1,1), state = argmax;
//  - ENO ~ U[1,14], written with one decimal;
//  - query distances: per pair |i-j|>=3 a 17-bin distribution over 2-20 A
//    (Dir(1,..,1)); the mode's bin centre and probability go to adddist with a
//    probability threshold of 0.05;
//  - template distances: a 3.8 A-step self-avoiding C-alpha chain; pairs
//    |i-j|>=3 with a distance above 64 A are written as missing ('-').
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <unordered_map>
#include <utility>
#include <vector>

#include "cth_adddist.h"
#include "cth_pset.h"

void cth_set_error(const char* fmt, ...);

namespace {

// Robinson & Robinson background (libpro/srcpro/SUBSTABLE.cpp:406-412)
const float kRobinson[CTH_NAA] = {
    0.078050f, 0.051290f, 0.044870f, 0.053640f, 0.019250f, 0.042640f, 0.062950f,
    0.073770f, 0.021990f, 0.051420f, 0.090190f, 0.057440f, 0.022430f, 0.038560f,
    0.052030f, 0.071200f, 0.058410f, 0.013300f, 0.032160f, 0.064410f};
const char kAA[] = "ARNDCQEGHILKMFPSTWYV";  // liblib/alpha.h:31
const char kSS[] = " eh";                   // SS state letters as the device sees them (liblib/alpha.cpp:14)


struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed, uint64_t a, uint64_t b) {
    s = seed * 0x9E3779B97F4A7C15ull ^ (a + 0x632BE59BD9B4E019ull) * 0xBF58476D1CE4E5B9ull ^
        (b + 0x2545F4914F6CDD1Dull) * 0x94D049BB133111EBull;
    next(); next();
  }
  uint64_t next() {  // splitmix64
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uni() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }  // (0,1)
  double normal() {
    double u1 = uni(), u2 = uni();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
  // Marsaglia & Tsang (2000), with the a<1 boost g(a) = g(a+1) U^(1/a)
  double gamma(double a) {
    if (a < 1.0) {
      double u = uni();
      return gamma(a + 1.0) * std::pow(u, 1.0 / a);
    }
    double d = a - 1.0 / 3.0, c = 1.0 / std::sqrt(9.0 * d);
    for (;;) {
      double x, v;
      do { x = normal(); v = 1.0 + c * x; } while (v <= 0.0);
      v = v * v * v;
      double u = uni();
      if (u < 1.0 - 0.0331 * x * x * x * x) return d * v;
      if (std::log(u) < 0.5 * x * x + d * (1.0 - v + std::log(v))) return d * v;
    }
  }
  void dirichlet(const double* alpha, int n, double* out) {
    double s = 0;
    for (int i = 0; i < n; i++) { out[i] = gamma(alpha[i]); s += out[i]; }
    if (s <= 0) { for (int i = 0; i < n; i++) out[i] = 1.0 / n; return; }
    for (int i = 0; i < n; i++) out[i] /= s;
  }
};

// text round trip of a probability written as rint(scale*p) and read as k/scale
inline float q_rint(double p) { return (float)(int)rintf(CTH_PROSCALE * (float)p) / (float)CTH_PROSCALE; }
// SS probabilities are written with truncation (PMProfileModel.cpp:2050-2053)
inline float q_floor(double p) { return (float)(int)(CTH_PROSCALE * (float)p) / (float)CTH_PROSCALE; }
// transition read-back (IOProfileModelTxt.cu:501-518): log, x0.4 for states after MD
inline float trn_read(double p, int t9) {
  float v = q_rint(p);
  if (!v) return CTH_LOG0;
  return t9 > 2 ? logf(v) * CTH_TRNLOGFCT : logf(v);
}

// one profile's 9 transition probabilities in file order
// (MM,MI,MD,IM,II,ID,DM,DI,DD) -> 7 stored log values
void sample_transitions(Rng& r, float out7[CTH_NTRN]) {
  const double am[3] = {8, 1, 1}, a2[2] = {1, 1};
  double m[3], i[2], d[2];
  r.dirichlet(am, 3, m);
  r.dirichlet(a2, 2, i);
  r.dirichlet(a2, 2, d);
  const double p9[9] = {m[0], m[1], m[2], i[0], i[1], 0.0, d[0], 0.0, d[1]};
  int tt = 0;
  for (int t = 0; t < 9; t++) {
    if (t == 5 || t == 7) continue;  // P_ID, P_DI are not stored
    out7[tt++] = trn_read(p9[t], t);
  }
}

// drive the transform over a full list of pair lines (i<j, j-i>=3), as
// adddist's ReadDistances does (adddist.cpp:269-367); dst(i,j) < 0 marks a
// line whose distance is missing or below the probability threshold
template <typename DstFn>
void adddist_lists(int L, DstFn dst, std::vector<std::vector<float>>& dgram) {
  dgram.assign(L, std::vector<float>());
  std::vector<std::vector<std::pair<float, int>>> tri(L);
  std::vector<float> vdsts;
  std::vector<int> vposs;
  int prevpos = 1;
  bool any = false;
  for (int i = 0; i < L; i++) {
    for (int j = i + 3; j < L; j++) {
      any = true;
      int r0 = i + 1;
      if (prevpos < r0) {
        cth_adddist_position(prevpos - 1, L, vdsts, vposs, tri, dgram);
        vdsts.clear(); vposs.clear();
        prevpos = r0;
      }
      float d = dst(i, j);
      if (d >= 0.f) {
        vdsts.push_back(d);
        vposs.push_back(j);
        tri[j].push_back({d, i});
      }
    }
  }
  (void)any;
  if (L > 0) cth_adddist_position(prevpos - 1, L, vdsts, vposs, tri, dgram);
}

// approximately self-avoiding C-alpha chain with 3.8 A steps
void calpha_chain(Rng& r, int L, std::vector<double>& xyz) {
  const double step = 3.8, excl = 3.8, cell = 4.0;
  xyz.assign(3 * (size_t)L, 0.0);
  std::unordered_map<int64_t, std::vector<int>> grid;
  auto key = [&](double x, double y, double z) {
    int64_t a = (int64_t)std::floor(x / cell) + (1 << 20), b = (int64_t)std::floor(y / cell) + (1 << 20),
            c = (int64_t)std::floor(z / cell) + (1 << 20);
    return (a << 42) | (b << 21) | c;
  };
  for (int i = 0; i < L; i++) {
    double x = 0, y = 0, z = 0;
    if (i > 0) {
      for (int tries = 0; tries < 64; tries++) {
        double u = 2.0 * r.uni() - 1.0, ph = 6.283185307179586 * r.uni();
        double s = std::sqrt(std::max(0.0, 1.0 - u * u));
        x = xyz[3 * (i - 1)] + step * s * std::cos(ph);
        y = xyz[3 * (i - 1) + 1] + step * s * std::sin(ph);
        z = xyz[3 * (i - 1) + 2] + step * u;
        bool clash = false;
        int64_t cx = (int64_t)std::floor(x / cell), cy = (int64_t)std::floor(y / cell), cz = (int64_t)std::floor(z / cell);
        for (int dx = -1; dx <= 1 && !clash; dx++)
          for (int dy = -1; dy <= 1 && !clash; dy++)
            for (int dz = -1; dz <= 1 && !clash; dz++) {
              int64_t k = ((cx + dx + (1 << 20)) << 42) | ((cy + dy + (1 << 20)) << 21) | (cz + dz + (1 << 20));
              auto it = grid.find(k);
              if (it == grid.end()) continue;
              for (int j : it->second) {
                if (j >= i - 1) continue;
                double ex = xyz[3 * j] - x, ey = xyz[3 * j + 1] - y, ez = xyz[3 * j + 2] - z;
                if (ex * ex + ey * ey + ez * ez < excl * excl) { clash = true; break; }
              }
            }
        if (!clash) break;
      }
    }
    xyz[3 * i] = x; xyz[3 * i + 1] = y; xyz[3 * i + 2] = z;
    grid[key(x, y, z)].push_back(i);
  }
}

struct SynthCfg {
  int nprof;
  int len_mode;       // 0 uniform [lmin,lmax]; 1 lognormal(median,sigma) clipped; 2 fixed lmin
  double median, sigma;
  int lmin, lmax;
  int dist_mode;      // 0 predicted map (query style); 1 C-alpha chain (template style); 2 none
  int sort_desc;      // 1: order profiles by length, longest first (makedb order)
  uint64_t seed;
};

int sample_len(const SynthCfg& c, int idx) {
  Rng r(c.seed, 0x1e9, (uint64_t)idx);
  if (c.len_mode == 2) return c.lmin;
  if (c.len_mode == 0) return c.lmin + (int)(r.uni() * (c.lmax - c.lmin + 1));
  double l = c.median * std::exp(c.sigma * r.normal());
  int li = (int)std::lround(l);
  return std::min(c.lmax, std::max(c.lmin, li));
}

void gen_profile(const SynthCfg& c, int org, int L, cth_pset* ps, int p) {
  Rng r(c.seed, 0xbeef, (uint64_t)org);
  const int64_t npos = ps->npos, ntr = ps->npos + ps->nprof;
  const int64_t d0 = ps->dist[p];
  for (int a = 0; a < CTH_NAA; a++) ps->bkg[(size_t)p * CTH_NAA + a] = q_rint(kRobinson[a]);
  {
    char buf[32];
    snprintf(buf, sizeof buf, "%.1f", 1.0 + 13.0 * r.uni());
    ps->eno[p] = strtof(buf, nullptr);
  }
  float t7[CTH_NTRN];
  sample_transitions(r, t7);
  for (int t = 0; t < CTH_NTRN; t++) ps->trn[t * ntr + d0 + p] = t7[t];
  double alpha[CTH_NAA], f[CTH_NAA];
  for (int a = 0; a < CTH_NAA; a++) alpha[a] = 0.5 * kRobinson[a];
  for (int m = 0; m < L; m++) {
    const int64_t i = d0 + m;
    r.dirichlet(alpha, CTH_NAA, f);
    int amax = 0;
    for (int a = 0; a < CTH_NAA; a++) {
      f[a] = 0.99 * f[a] + 0.01 * kRobinson[a];
      if (f[a] > f[amax]) amax = a;
      ps->trg[a * npos + i] = q_rint(f[a]);
    }
    ps->aa[i] = kAA[amax];
    sample_transitions(r, t7);
    for (int t = 0; t < CTH_NTRN; t++) ps->trn[t * ntr + d0 + p + m + 1] = t7[t];
    const double a3[3] = {1, 1, 1};
    double s3[3];
    r.dirichlet(a3, 3, s3);
    int smax = 0;
    for (int t = 0; t < CTH_NSS; t++) {
      if (s3[t] > s3[smax]) smax = t;
      ps->ss[t * npos + i] = q_floor(s3[t]);
    }
    ps->sss[i] = kSS[smax];
    for (int t = 0; t < CTH_NCV; t++) ps->cv[t * npos + i] = 0.17302947f;
    ps->cvprior[i] = 0.0f;
    ps->cvnorm2[i] = 0.0f;
  }
  // distance lists
  std::vector<std::vector<float>> dg;
  if (c.dist_mode == 0) {
    Rng rd(c.seed, 0xd157, (uint64_t)org);
    // mode bin centre and its probability for every pair line (i<j, j-i>=3)
    std::vector<float> dmap((size_t)L * L, -1.f);
    for (int a = 0; a < L; a++)
      for (int b = a + 3; b < L; b++) {
        double e[17], s = 0;
        int km = 0;
        for (int k = 0; k < 17; k++) { e[k] = -std::log(rd.uni()); s += e[k]; if (e[k] > e[km]) km = k; }
        double prob = e[km] / s;
        float dst = (float)(2.0 + (km + 0.5) * 18.0 / 17.0);
        dmap[(size_t)a * L + b] = prob < 0.05 ? -1.f : dst;
      }
    adddist_lists(L, [&](int a, int b) { return dmap[(size_t)a * L + b]; }, dg);
  } else if (c.dist_mode == 1) {
    Rng rc(c.seed, 0xca1f, (uint64_t)org);
    std::vector<double> xyz;
    calpha_chain(rc, L, xyz);
    adddist_lists(L, [&](int a, int b) {
      double dx = xyz[3 * a] - xyz[3 * b], dy = xyz[3 * a + 1] - xyz[3 * b + 1], dz = xyz[3 * a + 2] - xyz[3 * b + 2];
      float d = (float)std::sqrt(dx * dx + dy * dy + dz * dz);
      return d > 64.f ? -1.f : d;
    }, dg);
  }
  for (int m = 0; m < L; m++) {
    uint8_t* rec = ps->dd + (size_t)(d0 + m) * CTH_DDSTRIDE;
    memset(rec, 0xff, CTH_DDSTRIDE);
    int n = 0;
    if (c.dist_mode != 2 && m < (int)dg.size()) n = (int)dg[m].size();
    rec[0] = (uint8_t)n;
    for (int k = 0; k < n; k++) {
      int v = (int)rintf(dg[m][k]);
      rec[1 + k] = (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v);
    }
  }
}

}  // namespace

extern "C" {

// Free a profile set allocated by cth_synth_make.
void cth_pset_free(cth_pset* ps) {
  if (!ps) return;
  free(ps->len); free(ps->dist); free(ps->eno); free(ps->bkg); free(ps->aa); free(ps->sss);
  free(ps->trg); free(ps->trn); free(ps->ss); free(ps->cv); free(ps->cvprior); free(ps->cvnorm2);
  free(ps->dd);
  for (int p = 0; p < ps->nprof; p++) {
    if (ps->name) free(ps->name[p]);
    if (ps->desc) free(ps->desc[p]);
  }
  free(ps->name); free(ps->desc);
  free(ps);
}

// Name and description of profile k (NULL when unset).
const char* cth_pset_name(const cth_pset* ps, int k) {
  return ps && ps->name && 0 <= k && k < ps->nprof ? ps->name[k] : nullptr;
}
const char* cth_pset_desc(const cth_pset* ps, int k) {
  return ps && ps->desc && 0 <= k && k < ps->nprof ? ps->desc[k] : nullptr;
}

// Set (copy) the name and description of profile k; a NULL argument clears it.
int cth_pset_set_text(cth_pset* ps, int k, const char* name, const char* desc) {
  if (!ps || k < 0 || k >= ps->nprof || !ps->name || !ps->desc) return -1;
  char* n = name ? strdup(name) : nullptr;
  char* d = desc ? strdup(desc) : nullptr;
  if ((name && !n) || (desc && !d)) { free(n); free(d); return -1; }
  free(ps->name[k]); ps->name[k] = n;
  free(ps->desc[k]); ps->desc[k] = d;
  return 0;
}

// Allocate an empty profile set with the given lengths.
cth_pset* cth_pset_alloc(int nprof, const int32_t* lens) {
  if (nprof < 0 || (nprof > 0 && !lens)) { cth_set_error("pset_alloc: %d profiles", nprof); return nullptr; }
  for (int p = 0; p < nprof; p++)
    if (lens[p] < 1 || lens[p] > CTH_MAXCOLUMNS) {
      // the readers accept 1..MAXCOLUMNS positions (liblib/mylimits.h:50; IOProfileModelBin.cu:166)
      cth_set_error("pset_alloc: profile %d has length %d (1..%d)", p, (int)lens[p], CTH_MAXCOLUMNS);
      return nullptr;
    }
  cth_pset* ps = (cth_pset*)calloc(1, sizeof(cth_pset));
  if (!ps) return nullptr;
  ps->nprof = nprof;
  ps->len = (int32_t*)malloc(sizeof(int32_t) * (nprof ? nprof : 1));
  ps->dist = (int64_t*)malloc(sizeof(int64_t) * (nprof ? nprof : 1));
  int64_t np = 0;
  for (int p = 0; p < nprof; p++) { ps->len[p] = lens[p]; ps->dist[p] = np; np += lens[p]; }
  ps->npos = np;
  const size_t n1 = (size_t)(np ? np : 1);
  ps->eno = (float*)calloc(nprof ? nprof : 1, sizeof(float));
  ps->bkg = (float*)calloc((size_t)(nprof ? nprof : 1) * CTH_NAA, sizeof(float));
  ps->aa = (char*)calloc(n1, 1);
  ps->sss = (char*)calloc(n1, 1);
  ps->trg = (float*)calloc(n1 * CTH_NAA, sizeof(float));
  ps->trn = (float*)calloc((size_t)(np + nprof + 1) * CTH_NTRN, sizeof(float));
  ps->ss = (float*)calloc(n1 * CTH_NSS, sizeof(float));
  ps->cv = (float*)calloc(n1 * CTH_NCV, sizeof(float));
  ps->cvprior = (float*)calloc(n1, sizeof(float));
  ps->cvnorm2 = (float*)calloc(n1, sizeof(float));
  ps->dd = (uint8_t*)malloc(n1 * CTH_DDSTRIDE);
  ps->name = (char**)calloc(nprof ? nprof : 1, sizeof(char*));
  ps->desc = (char**)calloc(nprof ? nprof : 1, sizeof(char*));
  if (!ps->eno || !ps->bkg || !ps->aa || !ps->sss || !ps->trg || !ps->trn || !ps->ss || !ps->cv ||
      !ps->cvprior || !ps->cvnorm2 || !ps->dd || !ps->name || !ps->desc) {
    cth_pset_free(ps);
    return nullptr;
  }
  memset(ps->dd, 0xff, n1 * CTH_DDSTRIDE);
  return ps;
}

// Generate a seeded synthetic profile set.
//   len_mode 0: lengths uniform in [lmin,lmax]; 1: lognormal(median, sigma)
//   clipped to [lmin,lmax]; 2: all lengths lmin.
//   dist_mode 0: predicted-map distances (query style); 1: C-alpha chain
//   (template style); 2: no distances.
//   sort_desc 1: profiles ordered by length, longest first, as makedb orders a database.
//   nthreads: host threads used (profiles are independent).
cth_pset* cth_synth_make(int nprof, int len_mode, double median, double sigma, int lmin, int lmax,
                         int dist_mode, int sort_desc, uint64_t seed, int nthreads) {
  SynthCfg c{nprof, len_mode, median, sigma, lmin, lmax, dist_mode, sort_desc, seed};
  std::vector<int> lens(nprof), order(nprof);
  for (int i = 0; i < nprof; i++) lens[i] = sample_len(c, i);
  std::iota(order.begin(), order.end(), 0);
  if (sort_desc)
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return lens[a] > lens[b]; });
  std::vector<int32_t> sl(nprof);
  for (int p = 0; p < nprof; p++) sl[p] = lens[order[p]];
  cth_pset* ps = cth_pset_alloc(nprof, sl.data());
  if (!ps) return nullptr;
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (int p = 0; p < nprof; p++) gen_profile(c, order[p], sl[p], ps, p);
  for (int p = 0; p < nprof; p++) {
    char name[64], desc[160];
    snprintf(name, sizeof name, "synth%d", order[p]);
    snprintf(desc, sizeof desc, "synthetic profile %d, length %d, seed %llu", order[p], sl[p],
             (unsigned long long)seed);
    if (cth_pset_set_text(ps, p, name, desc)) { cth_pset_free(ps); return nullptr; }
  }
  return ps;
}

}  // extern "C"
