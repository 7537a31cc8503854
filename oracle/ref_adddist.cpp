// Test driver around the reference's distance-file transform (ReadDistances
// and ProcessDistancesForPosition of src/adddist/adddist.cpp, compiled from
// the reference's source by Makefile.ref with its main renamed).  Test
// infrastructure only.
//
// usage: adddist_ref <options file> <distance file> <profile length> <columns, e.g. 3,5> <probability threshold>
// prints one line per position: the list length, then the values as %a
#include <stdio.h>
#include <stdlib.h>

#include <string>
#include <vector>

#include "liblib/mybase.h"
#include "libpro/srcpro/MOptions.h"

void GetDstColumns(mystring, std::vector<int>&);
void ReadDistances(const char* filename, const int prolen, const std::vector<int>& colvec, float prbthrhld,
                   std::vector<std::vector<float>>& distogram);

int main(int argc, char** argv) {
  if (argc != 6) {
    fprintf(stderr, "usage: %s <options> <distances> <prolen> <columns> <prob threshold>\n", argv[0]);
    return 2;
  }
  try {
    MOptions::Read(argv[1]);
    std::vector<int> cols;
    GetDstColumns(argv[4], cols);
    const int L = atoi(argv[3]);
    std::vector<std::vector<float>> dg(L);
    ReadDistances(argv[2], L, cols, strtof(argv[5], nullptr), dg);
    for (const auto& l : dg) {
      printf("%zu", l.size());
      for (float v : l) printf(" %a", v);
      printf("\n");
    }
  } catch (myexception const& ex) {
    fprintf(stderr, "error: %s\n", ex.what());
    return 1;
  }
  return 0;
}
