// Test driver around the reference's results writer (class AlnWriter,
// src/libmycu/cualn/AlnWriter.cpp and AlnWriterPlain.cpp, compiled from the
// reference's sources by Makefile.ref).  Test infrastructure only.
//
// It reads per-hit records (annotation line and alignment record, as the
// reference's finalizer would hand them over) split into parts, passes each
// part to AlnWriter::PushPartOfResults exactly as CuBatchProcessingFinalizer::
// PassResultsToWriter does (cupro/CuBatchProcessingFinalizer.cpp:248-272,
// each part's indices sorted by log E-value as SortCompressedResults sorts
// them, :276-292), triggers the write as the job dispatcher does, and lets the
// reference write the output file for query 0.
//
// usage: alnwriter_ref [--json] <options file> <cother.par> <outdir> <input file>
//        (--json: the JSON format, CLOptions B_FMT = ofJSON)
//    or: alnwriter_ref --fmt <maxlen> <indent> <width> <lead char code> < text
//        (the reference's FormatDescription, liblib/fmtdescription.h:35-83,
//        applied to the text on stdin, result on stdout)
// input (text header lines, then raw record bytes):
//   <nqyposs> <deltalen> <sspace %a> <logevthld %a>
//   <query name>
//   <query description>
//   <dbsize> <ndbentries> <ndbnames>
//   <one database name per line>
//   <nparts>
//   per part: <nrec>, then per record "<logE %a> <annotation bytes> <record bytes>\n"
//   followed by the annotation and the record bytes
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <memory>
#include <string>
#include <vector>

#include "liblib/fmtdescription.h"
#include "liblib/msg.h"
#include "libpro/srcpro/Configuration.h"
#include "libpro/srcpro/CLOptions.h"
#include "libpro/srcpro/MOptions.h"
#include "libmycu/cualn/AlnWriter.h"

namespace {

bool read_line(FILE* fp, std::string& s) {
  s.clear();
  int c;
  while ((c = fgetc(fp)) != EOF && c != '\n') s += (char)c;
  return c != EOF || !s.empty();
}

char* read_bytes(FILE* fp, size_t n) {
  char* p = (char*)malloc(n + 1);
  if (!p || fread(p, 1, n, fp) != n) {
    fprintf(stderr, "truncated input\n");
    exit(2);
  }
  p[n] = 0;
  return p;
}

}  // namespace

int fmt_main(int argc, char** argv) {
  if (argc != 6) return 2;
  const int maxlen = atoi(argv[2]), indent = atoi(argv[3]), width = atoi(argv[4]);
  const char lead = (char)atoi(argv[5]);
  std::string text;
  int c;
  while ((c = getchar()) != EOF) text += (char)c;
  std::vector<char> out((size_t)maxlen + text.size() * (indent + 2) + 64);
  char* p = out.data();
  int outpos = 0, linepos = 0;
  FormatDescription(p, text.c_str(), maxlen, indent, width, outpos, linepos, lead);
  fwrite(out.data(), 1, p - out.data(), stdout);
  return 0;
}

int main(int argc, char** argv) {
  if (argc > 1 && !strcmp(argv[1], "--fmt")) return fmt_main(argc, argv);
  if (argc > 1 && !strcmp(argv[1], "--json")) {
    CLOptions::AssignCLOptB_FMT(CLOptions::ofJSON);
    argc--;
    argv++;
  }
  if (argc != 5) {
    fprintf(stderr, "usage: %s <options> <cother.par> <outdir> <input>\n", argv[0]);
    return 2;
  }
  SetProgramName("cother", "2.03.03");  // as cother/cother.cpp:60 with cother/cother.h:23
  MOptions::Read(argv[1]);
  Configuration config[NoCTypes];
  // JobDispatcher::ReadConfiguration (cualn/JobDispatcher.cpp:152-167)
  config[CTUngapped].SetFilename(argv[2]);
  config[CTGapped].SetFilename(argv[2]);
  config[CTUngapped].ReadUngapped();
  config[CTGapped].SetAutoGapOpenCost(true);
  config[CTGapped].SetGapOpenCost(4);
  config[CTGapped].SetGapExtendCost(1);
  config[CTGapped].Read();

  FILE* fp = fopen(argv[4], "rb");
  if (!fp) {
    fprintf(stderr, "cannot open %s\n", argv[4]);
    return 2;
  }
  std::string line, qname, qdesc;
  int nqyposs = 0;
  unsigned deltalen = 0;
  float sspace = 0, logevthld = 0;
  size_t dbsize = 0, ndbentries = 0;
  int ndb = 0, nparts = 0;
  read_line(fp, line);
  if (sscanf(line.c_str(), "%d %u %a %a", &nqyposs, &deltalen, &sspace, &logevthld) != 4) return 2;
  read_line(fp, qname);
  read_line(fp, qdesc);
  read_line(fp, line);
  if (sscanf(line.c_str(), "%zu %zu %d", &dbsize, &ndbentries, &ndb) != 3) return 2;
  std::vector<std::string> dbnames(ndb);
  for (auto& d : dbnames) read_line(fp, d);
  read_line(fp, line);
  nparts = atoi(line.c_str());

  try {
    AlnWriter writer(config, argv[3], dbnames, dbsize, ndbentries, 1);
    for (int part = 0; part < nparts; part++) {
      read_line(fp, line);
      const int nrec = atoi(line.c_str());
      std::vector<std::string> ann(nrec), aln(nrec);
      std::unique_ptr<std::vector<float>> loge(new std::vector<float>(nrec));
      for (int r = 0; r < nrec; r++) {
        size_t na = 0, nb = 0;
        read_line(fp, line);
        if (sscanf(line.c_str(), "%a %zu %zu", &(*loge)[r], &na, &nb) != 3) return 2;
        char* a = read_bytes(fp, na);
        char* b = read_bytes(fp, nb);
        ann[r].assign(a, na);
        aln[r].assign(b, nb);
        free(a);
        free(b);
      }
      // the finalizer's buffers: records one after another, each NUL-terminated
      size_t sa = 1, sb = 1;
      for (int r = 0; r < nrec; r++) {
        sa += ann[r].size() + 1;
        sb += aln[r].size() + 1;
      }
      std::unique_ptr<char, MyDataDestroyer> abuf((char*)malloc(sa)), bbuf((char*)malloc(sb));
      std::unique_ptr<std::vector<int>> srt(new std::vector<int>(nrec));
      std::unique_ptr<std::vector<char*>> aptrs(new std::vector<char*>(nrec)), bptrs(new std::vector<char*>(nrec));
      char* pa = abuf.get();
      char* pb = bbuf.get();
      for (int r = 0; r < nrec; r++) {
        (*srt)[r] = r;
        (*aptrs)[r] = pa;
        memcpy(pa, ann[r].c_str(), ann[r].size() + 1);
        pa += ann[r].size() + 1;
        (*bptrs)[r] = pb;
        memcpy(pb, aln[r].c_str(), aln[r].size() + 1);
        pb += aln[r].size() + 1;
      }
      const std::vector<float>& lv = *loge;
      std::sort(srt->begin(), srt->end(), [&lv](size_t a, size_t b) { return lv[a] < lv[b]; });
      writer.IncreaseQueryNParts(0);
      writer.PushPartOfResults(0, nqyposs, qname.c_str(), qdesc.c_str(), deltalen, sspace, logevthld,
                               std::move(abuf), std::move(bbuf), std::move(srt), std::move(loge), std::move(bptrs),
                               std::move(aptrs));
    }
    writer.DereaseNPartsAndTrigger(0);
    if (writer.WaitDone() == WRITERTHREAD_MSG_ERROR) {
      fprintf(stderr, "writer failed\n");
      return 1;
    }
    writer.Notify(AlnWriter::wrtthreadmsgTerminate);
  } catch (myexception const& ex) {
    fprintf(stderr, "%s\n", ex.what());
    return 1;
  }
  fclose(fp);
  return 0;
}
