/* orc_search.c: the whole per-query search of one query against a range of
 * database profiles, composed from the oracle's stage restatements (TEST
 * INFRASTRUCTURE ONLY: the checker for cth_search and the timed CPU baseline).
 *
 * Pipeline as CuBatchProcessing::BatchProcessScoreMatrixDevice runs it
 * (src/libmycu/cupro/CuBatchProcessing.cu:230-575): score matrix, phase-1 DP
 * and finalisation, statistics of the passed profiles, the E-value
 * filter of FinalizeMAPDP_MAP (src/libmycu/cumapdp/CuBatchMAPDP_map_final.cu:
 * 117-123) and, with do_map, the MAP stage (hits without a MAP alignment are
 * dropped as the host finalizer drops them).  Each profile is processed on its own (its score matrix only),
 * so memory stays at one Lq x Lt block per thread.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define GAPPED_LAMBDA 0.561132f
#define GAPPED_K 0.001488f
#define GAPPED_ALPHA 0.5f
#define GAPPED_BETA (-1.0f)

static int hit_cmp(const void* a, const void* b) {
  const orc_hit* x = (const orc_hit*)a;
  const orc_hit* y = (const orc_hit*)b;
  if (x->logE < y->logE) return -1;
  if (y->logE < x->logE) return 1;
  return (x->profile > y->profile) - (x->profile < y->profile);
}

int orc_search_query(const orc_tables* t, const orc_pset* q, int qp, const orc_pset* db, int p0, int p1,
                     float eval, int ssemodel, int do_map, uint64_t dbsize, uint64_t ndb, orc_hit* hits, int maxhits,
                     int nthreads) {
  const int lq = q->len[qp];
  float scorethld, logethld;
  orc_score_threshold(eval, GAPPED_LAMBDA, GAPPED_K, (uint64_t)lq, dbsize, &scorethld, &logethld);
  uint64_t dl;
  const double sspace =
      (double)orc_length_adjustment(GAPPED_LAMBDA, GAPPED_K, GAPPED_ALPHA, GAPPED_BETA, (uint64_t)lq, dbsize, ndb, &dl);
  const int n = p1 - p0;
  orc_hit* all = (orc_hit*)calloc((size_t)(n > 0 ? n : 1), sizeof(orc_hit));
  char* keep = (char*)calloc((size_t)(n > 0 ? n : 1), 1);
  const char* qaa = q->aa + q->dist[qp];
  /* few long profiles: one at a time with the threads inside the score
   * matrix; many: the threads over profiles */
  const int inner = (n < nthreads) ? nthreads : 1;
#pragma omp parallel num_threads(inner > 1 ? 1 : (nthreads > 0 ? nthreads : 1))
  {
    float* sc = NULL;
    float* md = NULL;
    uint8_t* bt = NULL;
    size_t cap = 0;
#pragma omp for schedule(dynamic, 4)
    for (int k = 0; k < n; k++) {
      const int p = p0 + k;
      const int lt = db->len[p];
      const size_t need = (size_t)lq * (size_t)lt;
      if (need > cap) {
        free(sc); free(md); free(bt);
        sc = (float*)malloc(need * sizeof(float));
        md = (float*)malloc(need * sizeof(float));
        bt = (uint8_t*)malloc(need);
        cap = need;
      }
      orc_score_matrix(t, q, qp, db, p, p + 1, sc, md, inner);
      float maxsc, score;
      uint32_t maxxy, beg;
      int passed, psts;
      orc_dp_phase1(q, qp, db, p, p + 1, sc, bt, scorethld, &maxsc, &maxxy, &passed, &score, &psts, &beg, 1);
      if (!passed) continue;
      double counts[641];
      orc_histogram(sc, lt, 0, lq, lt, qaa, db->aa + db->dist[p], counts);
      double par[6], sig[4];
      orc_stat_params(counts, par);
      orc_significance(ssemodel, lq, lt, q->eno[qp], db->eno[p], sspace, score, psts, par[4], par[5], sig);
      const float loge = (float)sig[2];
      if (logethld < loge) continue;
      orc_mapres mr;
      memset(&mr, 0, sizeof mr);
      if (do_map) {
        orc_map_profile(q, qp, db, p, sc, md, lt, 0, (float)sig[3], 0.28f, &mr, NULL, 0);
        if (mr.alnlen < 1) continue; /* hits need an alignment (CuBatchProcessingFinalizerPlain.cpp:87-142) */
      }
      orc_hit* h = &all[k];
      h->query = qp;
      h->profile = p;
      h->score = score;
      h->logE = loge;
      h->logEpA = (float)sig[3];
      h->lambda_est = (float)sig[0];
      h->K_est = (float)sig[1];
      h->minsc = (float)par[0];
      h->maxsc = (float)par[1];
      h->expected = (float)par[2];
      h->H = (float)par[3];
      h->lambda = (float)par[4];
      h->K = (float)par[5];
      h->psts = psts;
      h->beg = beg;
      h->end = maxxy;
      if (do_map) {
        h->mapbeg = mr.beg;
        h->mapend = mr.end;
        h->idnts = mr.idnts;
        h->ngaps = mr.ngaps;
        h->alnlen = mr.alnlen;
        h->pstvs_map = mr.psts;
        h->mapscore = mr.maxsc;
      }
      keep[k] = 1;
    }
    free(sc); free(md); free(bt);
  }
  int nh = 0;
  for (int k = 0; k < n; k++)
    if (keep[k]) {
      if (nh >= maxhits) { nh = -1; break; }
      hits[nh++] = all[k];
    }
  free(all);
  free(keep);
  if (nh > 0) qsort(hits, (size_t)nh, sizeof(orc_hit), hit_cmp);
  return nh;
}
