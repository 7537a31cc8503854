// ref_ddms_wrap.cpp: C entry points over the reference's own CPU DDMS
// (cother-scores-optimizer/ddms.cxx, CSO_DDMS::CalculateDDMS) so the oracle's
// DDMS restatement can be checked against it.  TEST INFRASTRUCTURE ONLY: built
// from the reference sources in place by oracle/Makefile.ref into
// oracle/_ref/, never shipped or linked by the product.
#include "ddms.h"

extern "C" {
// lists are int[n+1+8]: n, d1..dn, then 255 padding (raw distances)
float ref_ddms_calculate(int* v1, int* v2) {
  static CSO_DDMS ddms;
  return ddms.CalculateDDMS(v1, v2);
}
}
