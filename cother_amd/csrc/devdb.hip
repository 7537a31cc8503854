// devdb.hip: upload of a host profile set into its HBM-resident layout.
//
// The host set (cth_pset) follows the reference's field conventions; the
// device copy pads every position field to a multiple of 64 positions,
// stores the owning profile index of each position (the reference's
// pmv2DAddrPro field, cupro/PM2DVectorFields.h:79) and rewrites distance
// lists with distances saturated at 20 and 21 for the missing value 255 --
// the only form the distance-match kernel consumes
// (CacheIRDistanceValuesh2, cusco/CuBatchScoreMatrix_ddms2s_h2.cuh:55-108).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "cth_dev.h"
#include "cth_devdb.h"
#include "cth_pset.h"

namespace {

int64_t pad64(int64_t n) { return (n + 63) / 64 * 64; }

template <typename T>
int up(void** dst, const T* src, size_t n, size_t nalloc) {
  CTH_HIP_CHECK(hipMalloc(dst, nalloc * sizeof(T) + 256));
  CTH_HIP_CHECK(hipMemset(*dst, 0, nalloc * sizeof(T) + 256));
  if (n) CTH_HIP_CHECK(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

// field-major array [nf][n] -> [nf][ld]
int up_fields(void** dst, const float* src, int nf, int64_t n, int64_t ld) {
  CTH_HIP_CHECK(hipMalloc(dst, sizeof(float) * nf * ld + 256));
  CTH_HIP_CHECK(hipMemset(*dst, 0, sizeof(float) * nf * ld + 256));
  for (int f = 0; f < nf; f++)
    if (n) CTH_HIP_CHECK(hipMemcpy((float*)*dst + f * ld, src + f * n, sizeof(float) * n, hipMemcpyHostToDevice));
  return 0;
}

}  // namespace

int cth_devset_has_ddms(const cth_devset* s, int p0, int p1) {
  if (!s->h_maxdd) return 1;
  for (int p = p0; p < p1; p++)
    if (s->h_maxdd[p] >= CTH_DSTSEGM) return 1;
  return 0;
}

extern "C" void cth_devset_free(cth_devset* s) {
  if (!s) return;
  for (void* p : {(void*)s->d.len, (void*)s->d.dist, (void*)s->d.eno, (void*)s->d.bkg, (void*)s->d.aa, (void*)s->d.sss,
                  (void*)s->d.pro, (void*)s->d.trg, (void*)s->d.trn, (void*)s->d.ss, (void*)s->d.cv,
                  (void*)s->d.cvprior, (void*)s->d.cvnorm2, (void*)s->d.dd})
    if (p) (void)hipFree(p);
  if (s->ddoff) (void)hipFree(s->ddoff);
  free(s->h_len);
  free(s->h_dist);
  free(s->h_maxdd);
  memset(s, 0, sizeof *s);
}

extern "C" int cth_devset_upload(const cth_pset* h, cth_devset* s) {
  memset(s, 0, sizeof *s);
  const int64_t n = h->npos, ld = pad64(n > 0 ? n : 1);
  // distance lists hold at most MAX_N_IR_DISTANCE_VALUES = 99 values
  // (libpro/srcpro/MOptions.h:15); the score kernel's scan relies on it
  for (int64_t i = 0; i < n; i++)
    if (h->dd[(size_t)i * CTH_DDSTRIDE] > CTH_MAXDST) {
      cth_set_error("position %lld: distance list of %d values (at most %d)", (long long)i,
                    (int)h->dd[(size_t)i * CTH_DDSTRIDE], CTH_MAXDST);
      return -1;
    }
  s->h_maxdd = (int32_t*)calloc(h->nprof > 0 ? h->nprof : 1, sizeof(int32_t));
  if (!s->h_maxdd) { cth_set_error("out of host memory"); return -1; }
  for (int p = 0; p < h->nprof; p++) {
    int m = 0;
    for (int k = 0; k < h->len[p]; k++) m = std::max<int>(m, h->dd[(size_t)(h->dist[p] + k) * CTH_DDSTRIDE]);
    s->h_maxdd[p] = m;
  }
  const int64_t ntr = h->npos + h->nprof, ldtr = pad64(ntr > 0 ? ntr : 1);
  cth_dpset& d = s->d;
  d.nprof = h->nprof;
  d.npos = n;
  d.ld = ld;
  d.ldtr = ldtr;
  const int np = h->nprof > 0 ? h->nprof : 1;
  int rc = 0;
  rc |= up((void**)&d.len, h->len, h->nprof, np);
  rc |= up((void**)&d.dist, h->dist, h->nprof, np);
  rc |= up((void**)&d.eno, h->eno, h->nprof, np);
  rc |= up((void**)&d.bkg, h->bkg, (size_t)h->nprof * CTH_NAA, (size_t)np * CTH_NAA);
  rc |= up((void**)&d.aa, h->aa, n, ld);
  rc |= up((void**)&d.sss, h->sss, n, ld);
  {
    std::vector<int32_t> pro(ld, 0);
    for (int p = 0; p < h->nprof; p++)
      for (int m = 0; m < h->len[p]; m++) pro[h->dist[p] + m] = p;
    rc |= up((void**)&d.pro, pro.data(), ld, ld);
  }
  rc |= up_fields((void**)&d.trg, h->trg, CTH_NAA, n, ld);
  rc |= up_fields((void**)&d.trn, h->trn, CTH_NTRN, ntr, ldtr);
  rc |= up_fields((void**)&d.ss, h->ss, CTH_NSS, n, ld);
  rc |= up_fields((void**)&d.cv, h->cv, CTH_NCV, n, ld);
  rc |= up((void**)&d.cvprior, h->cvprior, n, ld);
  rc |= up((void**)&d.cvnorm2, h->cvnorm2, n, ld);
  if (!rc) {
    // distance lists in chunks of 1M positions
    CTH_HIP_CHECK(hipMalloc((void**)&d.dd, (size_t)ld * CTH_DDSTRIDE));
    const int64_t chunk = 1 << 20;
    std::vector<uint8_t> buf((size_t)std::min<int64_t>(chunk, ld) * CTH_DDSTRIDE);
    for (int64_t b = 0; b < ld; b += chunk) {
      const int64_t e = std::min<int64_t>(ld, b + chunk);
      for (int64_t i = b; i < e; i++) {
        uint8_t* o = &buf[(size_t)(i - b) * CTH_DDSTRIDE];
        if (i < n) {
          const uint8_t* r = h->dd + (size_t)i * CTH_DDSTRIDE;
          o[0] = r[0];
          for (int k = 1; k < CTH_DDSTRIDE; k++) {
            const uint8_t v = r[k];
            o[k] = v == 255 ? CTH_DD_MISSING : (v < CTH_DMS_MAXDST ? v : CTH_DMS_MAXDST);
          }
        } else {
          memset(o, CTH_DD_MISSING, CTH_DDSTRIDE);
          o[0] = 0;
        }
      }
      CTH_HIP_CHECK(hipMemcpy((void*)(d.dd + (size_t)b * CTH_DDSTRIDE), buf.data(), (size_t)(e - b) * CTH_DDSTRIDE,
                              hipMemcpyHostToDevice));
    }
  }
  s->h_len = (int32_t*)malloc(sizeof(int32_t) * np);
  s->h_dist = (int64_t*)malloc(sizeof(int64_t) * np);
  if (h->nprof) {
    memcpy(s->h_len, h->len, sizeof(int32_t) * h->nprof);
    memcpy(s->h_dist, h->dist, sizeof(int64_t) * h->nprof);
  }
  if (rc) { cth_devset_free(s); return -1; }
  return 0;
}

// The query set's distance lists as byte offsets into the 22 x 22 pair-score
// table (value v -> row v, 22 floats per row), one 128-entry record per
// position, so the score kernel adds the template value's column directly.
extern "C" int cth_devset_make_ddoff(const cth_pset* h, cth_devset* s) {
  const int64_t n = h->npos;
  std::vector<int32_t> buf((size_t)std::max<int64_t>(n, 1) * CTH_DDSTRIDE);
  for (int64_t i = 0; i < n; i++) {
    const uint8_t* r = h->dd + (size_t)i * CTH_DDSTRIDE;
    int32_t* o = &buf[(size_t)i * CTH_DDSTRIDE];
    o[0] = r[0];
    for (int k = 1; k < CTH_DDSTRIDE; k++) {
      const uint8_t v = r[k];
      const int c = v == 255 ? CTH_DD_MISSING : (v < CTH_DMS_MAXDST ? v : CTH_DMS_MAXDST);
      o[k] = c * 22 * (int)sizeof(float);
    }
  }
  if (s->ddoff) (void)hipFree(s->ddoff);
  s->ddoff = nullptr;
  CTH_HIP_CHECK(hipMalloc((void**)&s->ddoff, buf.size() * sizeof(int32_t)));
  CTH_HIP_CHECK(hipMemcpy(s->ddoff, buf.data(), buf.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  return 0;
}
