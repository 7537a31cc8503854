// Test driver around the reference's length adjustment
// (AbstractScoreMatrix<float>::ComputeLengthAdjustment,
// src/libpro/srcsco/AbstractScoreMatrix.cpp:246-395, compiled from the
// reference's source by Makefile.ref).  Test infrastructure only.
//
// usage: lenadj_ref <lambda> <K> <alpha> <beta>, then lines "m n N" on stdin;
// prints "searchspace deltalength" per line.  The abstract class is
// instantiated through a derived class whose other members are never called.
#include <stdio.h>
#include <stdlib.h>

#include "libpro/srcpro/Configuration.h"
#include "libpro/srcsco/AbstractScoreMatrix.h"

namespace {
class Probe : public AbstractScoreMatrix<float> {
 public:
  explicit Probe(Configuration* cfg) : AbstractScoreMatrix<float>(ProfileSpecific, cfg, StatisticsGiven) {}
  float GetScore(int, int) const override { return 0.0f; }
  float GetModScore(int, int) const override { return 0.0f; }
  void ComputeScoreMatrix(bool) override {}
  float GetFinalScore(float v) const override { return v; }
  void ScaleScoreMatrix() override {}
  void ComputeStatisticalParameters(bool) override {}
  void ComputeScoreProbabilities(ScoresAttr<float>*) override {}
  void PrintParameterTable(TPrintFunction, void*) const override {}
  void SetScore(int, int, float) override {}
  void SetModScore(int, int, float) override {}
  bool GetAllNegatives() const override { return false; }
  void SetAllNegatives(bool) override {}
  const float** GetScores() const override { return nullptr; }
};
}  // namespace

int main(int argc, char** argv) {
  if (argc != 5) {
    fprintf(stderr, "usage: %s <lambda> <K> <alpha> <beta>  (lines m n N on stdin)\n", argv[0]);
    return 2;
  }
  Configuration cfg[NoCTypes];
  cfg[CTGapped].SetLambda(strtof(argv[1], nullptr));
  cfg[CTGapped].SetK(strtof(argv[2], nullptr));
  cfg[CTGapped].SetAlpha(strtof(argv[3], nullptr));
  cfg[CTGapped].SetBeta(strtof(argv[4], nullptr));
  unsigned long long m, n, N;
  while (scanf("%llu %llu %llu", &m, &n, &N) == 3) {
    Probe p(cfg);  // a fresh object: the public overload caches per query length
    p.ComputeLengthAdjustment((size_t)m, (uint64_mt)n, (size_t)N);
    printf("%llu %zu\n", (unsigned long long)p.GetSearchSpace(), p.GetDeltaLength());
  }
  return 0;
}
