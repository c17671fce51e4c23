// Test-only wrapper around the reference's Pannotia graph parsers
// (code/experiments/pannotia/graph_parser/parse.cpp), compiled from the
// reference tree by `make ref` into oracle/_ref/.  Gives the tests a C entry
// point that returns the parsed CSR arrays.
#include <cstdlib>
#include <cstring>
#include <string>
#include "parse.h"

extern "C" int ref_parse(const char* path, int kind, int directed, int* n, int* m, int** row, int** col,
                         int** w) {
  std::string p(path);
  csr_array* c = nullptr;
  if (kind == 0) c = parseMetis(&p[0], n, m, directed != 0);
  else if (kind == 1) c = parseCOO(&p[0], n, m, directed != 0);
  else c = parseCOO_transpose(&p[0], n, m, directed != 0);
  if (!c) return -1;
  *row = c->row_array;
  *col = c->col_array;
  *w = c->data_array;
  return 0;
}

extern "C" void ref_free(void* p) { free(p); }
