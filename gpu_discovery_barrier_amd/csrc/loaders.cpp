// Graph file loaders and writers (host C++).
//
// The edge multiset and in-row order of every loader follow the reference
// parser it replaces; the row offsets are always a correct CSR.  (The
// reference builders emit a row start only when the row id changes, so a
// vertex without out-edges shifts every later row: parse.cpp:152-162,
// 258-269, 776-787.  Graphs without empty rows load identically.)
//
//   gdb_load_metis      parseMetis         pannotia/graph_parser/parse.cpp:59-176
//   gdb_load_dimacs     parseCOO           parse.cpp:178-282 (directed = 0 lists
//                                          both directions, then a stable sort by row)
//   gdb_load_dimacs(transpose = 1)
//                       parseCOO_transpose parse.cpp:693-799 (each arc u->v becomes
//                                          v<-u; undirected adds the same tuple twice)
//   gdb_load_galois_gr  readFromGR         lonestar_gpu_opencl/include/graph.h:143-215
//   gdb_load_edges      readFromEdges      lonestar_gpu_opencl/include/graph.h:96-141
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include "../../include/gdb_graph.h"

namespace {

struct Tuple {
  int row, col, val;
};

int build_from_tuples(int n, std::vector<Tuple>& t, gdb_host_csr* out) {
  for (const Tuple& x : t)
    if (x.row < 0 || x.row >= n || x.col < 0 || x.col >= n) return GDB_GRAPH_E_FORMAT;
  std::stable_sort(t.begin(), t.end(), [](const Tuple& a, const Tuple& b) { return a.row < b.row; });
  const size_t m = t.size();
  if (m > 0x7fffffffu) return GDB_GRAPH_E_TOO_LARGE;
  out->n = n;
  out->m = (int)m;
  out->row = (int*)calloc((size_t)n + 1, sizeof(int));
  out->col = (int*)malloc(sizeof(int) * (m ? m : 1));
  out->w = (int*)malloc(sizeof(int) * (m ? m : 1));
  if (!out->row || !out->col || !out->w) return GDB_GRAPH_E_NOMEM;
  for (const Tuple& x : t) out->row[x.row + 1]++;
  for (int i = 0; i < n; ++i) out->row[i + 1] += out->row[i];
  for (size_t k = 0; k < m; ++k) {
    out->col[k] = t[k].col;
    out->w[k] = t[k].val;
  }
  return GDB_GRAPH_OK;
}

}  // namespace

extern "C" {

int gdb_load_metis(const char* path, int directed, gdb_host_csr* out) {
  if (!path || !out) return GDB_GRAPH_E_INVALID;
  *out = gdb_host_csr{};
  FILE* f = fopen(path, "r");
  if (!f) return GDB_GRAPH_E_IO;
  std::vector<char> line(1 << 20);
  std::vector<Tuple> t;
  long long lineno = 0;
  int n = 0, m_decl = 0;
  while (fgets(line.data(), (int)line.size(), f)) {
    if (line[0] == '%') continue;  // comment lines, parse.cpp:83
    if (lineno == 0) {
      if (sscanf(line.data(), "%d %d", &n, &m_decl) != 2 || n < 0) {
        fclose(f);
        return GDB_GRAPH_E_FORMAT;
      }
      t.reserve((size_t)m_decl * (directed ? 1 : 2));
    } else {
      // Neighbour list of vertex `lineno` (1-based), tokens split on " ,.-"
      // and ended by the first non-positive value (parse.cpp:107-125).
      char* save = nullptr;
      for (char* p = strtok_r(line.data(), " ,.-", &save); p; p = strtok_r(nullptr, " ,.-", &save)) {
        const int tail = atoi(p);
        if (tail <= 0) break;
        t.push_back(Tuple{(int)lineno - 1, tail - 1, 0});
      }
    }
    ++lineno;
  }
  fclose(f);
  if (lineno == 0) return GDB_GRAPH_E_FORMAT;
  return build_from_tuples(n, t, out);
}

int gdb_load_dimacs(const char* path, int directed, int transpose, gdb_host_csr* out) {
  if (!path || !out) return GDB_GRAPH_E_INVALID;
  *out = gdb_host_csr{};
  FILE* f = fopen(path, "r");
  if (!f) return GDB_GRAPH_E_IO;
  char line[512];
  std::vector<Tuple> t;
  int n = -1, m_decl = 0;
  while (fgets(line, sizeof(line), f)) {
    if (line[0] == 'p') {
      char sp[16];
      if (sscanf(line, "p %15s %d %d", sp, &n, &m_decl) != 3 || n < 0) {
        fclose(f);
        return GDB_GRAPH_E_FORMAT;
      }
      t.reserve((size_t)m_decl * (directed ? 1 : 2));
    } else if (line[0] == 'a') {
      int head, tail, weight;
      if (sscanf(line, "a %d %d %d", &head, &tail, &weight) != 3) {
        fclose(f);
        return GDB_GRAPH_E_FORMAT;
      }
      if (!transpose) {  // parse.cpp:221-233
        t.push_back(Tuple{head - 1, tail - 1, weight});
        if (!directed) t.push_back(Tuple{tail - 1, head - 1, weight});
      } else {  // parse.cpp:738-750
        t.push_back(Tuple{tail - 1, head - 1, weight});
        if (!directed) t.push_back(Tuple{tail - 1, head - 1, weight});
      }
    }
  }
  fclose(f);
  if (n < 0) return GDB_GRAPH_E_FORMAT;
  return build_from_tuples(n, t, out);
}

int gdb_load_galois_gr(const char* path, gdb_host_csr* out) {
  if (!path || !out) return GDB_GRAPH_E_INVALID;
  *out = gdb_host_csr{};
  FILE* f = fopen(path, "rb");
  if (!f) return GDB_GRAPH_E_IO;
  uint64_t hdr[4];
  if (fread(hdr, sizeof(uint64_t), 4, f) != 4 || hdr[0] != 1 || hdr[2] > 0x7fffffffull ||
      hdr[3] > 0x7fffffffull) {
    fclose(f);
    return GDB_GRAPH_E_FORMAT;
  }
  const uint64_t size_edge = hdr[1], n = hdr[2], m = hdr[3];
  std::vector<uint64_t> out_idx(n);
  std::vector<uint32_t> outs(m);
  bool ok = fread(out_idx.data(), sizeof(uint64_t), n, f) == n &&
            fread(outs.data(), sizeof(uint32_t), m, f) == m;
  if (ok && (m % 2)) {  // 4-byte pad to an 8-byte boundary (graph.h:168)
    uint32_t pad;
    ok = fread(&pad, sizeof(pad), 1, f) == 1 || m == 0;
  }
  std::vector<uint32_t> data(m, 1u);
  if (ok && size_edge >= 4 && m) {
    // Edge data is read as 32-bit values (graph.h:169, 200).
    if (size_edge == 4) {
      ok = fread(data.data(), sizeof(uint32_t), m, f) == m;
    } else {
      std::vector<char> raw(size_edge);
      for (uint64_t k = 0; k < m && ok; ++k) {
        ok = fread(raw.data(), 1, size_edge, f) == size_edge;
        memcpy(&data[k], raw.data(), 4);
      }
    }
  }
  fclose(f);
  if (!ok) return GDB_GRAPH_E_FORMAT;
  out->n = (int)n;
  out->m = (int)m;
  out->row = (int*)malloc(sizeof(int) * (n + 1));
  out->col = (int*)malloc(sizeof(int) * (m ? m : 1));
  out->w = (int*)malloc(sizeof(int) * (m ? m : 1));
  if (!out->row || !out->col || !out->w) return GDB_GRAPH_E_NOMEM;
  out->row[0] = 0;
  for (uint64_t i = 0; i < n; ++i) {
    if (out_idx[i] > m || (i && out_idx[i] < out_idx[i - 1])) return GDB_GRAPH_E_FORMAT;
    out->row[i + 1] = (int)out_idx[i];
  }
  if (n && (uint64_t)out->row[n] != m) return GDB_GRAPH_E_FORMAT;
  for (uint64_t k = 0; k < m; ++k) {
    if (outs[k] >= n) return GDB_GRAPH_E_FORMAT;
    out->col[k] = (int)outs[k];
    out->w[k] = (int)data[k];
  }
  return GDB_GRAPH_OK;
}

int gdb_write_galois_gr(const char* path, const gdb_host_csr* g) {
  if (!path || !g || g->n < 0 || g->m < 0) return GDB_GRAPH_E_INVALID;
  FILE* f = fopen(path, "wb");
  if (!f) return GDB_GRAPH_E_IO;
  const uint64_t hdr[4] = {1, 4, (uint64_t)g->n, (uint64_t)g->m};
  bool ok = fwrite(hdr, sizeof(uint64_t), 4, f) == 4;
  for (int i = 0; i < g->n && ok; ++i) {
    const uint64_t e = (uint64_t)g->row[i + 1];
    ok = fwrite(&e, sizeof(e), 1, f) == 1;
  }
  for (int k = 0; k < g->m && ok; ++k) {
    const uint32_t d = (uint32_t)g->col[k];
    ok = fwrite(&d, sizeof(d), 1, f) == 1;
  }
  if (ok && (g->m % 2)) {
    const uint32_t pad = 0;
    ok = fwrite(&pad, sizeof(pad), 1, f) == 1;
  }
  for (int k = 0; k < g->m && ok; ++k) {
    const uint32_t d = g->w ? (uint32_t)g->w[k] : 1u;
    ok = fwrite(&d, sizeof(d), 1, f) == 1;
  }
  if (fclose(f) != 0) ok = false;
  return ok ? GDB_GRAPH_OK : GDB_GRAPH_E_IO;
}

int gdb_load_edges(const char* path, gdb_host_csr* out) {
  if (!path || !out) return GDB_GRAPH_E_INVALID;
  *out = gdb_host_csr{};
  FILE* f = fopen(path, "r");
  if (!f) return GDB_GRAPH_E_IO;
  int n = 0, m = 0;
  if (fscanf(f, "%d %d", &n, &m) != 2 || n < 0 || m < 0) {
    fclose(f);
    return GDB_GRAPH_E_FORMAT;
  }
  // "src dst weight" lines, grouped by source (graph.h:113-137).
  std::vector<Tuple> t((size_t)m);
  for (int k = 0; k < m; ++k) {
    if (fscanf(f, "%d %d %d", &t[k].row, &t[k].col, &t[k].val) != 3) {
      fclose(f);
      return GDB_GRAPH_E_FORMAT;
    }
  }
  fclose(f);
  return build_from_tuples(n, t, out);
}

int gdb_csr_validate(const gdb_host_csr* g) {
  if (!g || g->n < 0 || g->m < 0 || !g->row) return GDB_GRAPH_E_INVALID;
  if (g->row[0] != 0 || g->row[g->n] != g->m) return GDB_GRAPH_E_FORMAT;
  for (int i = 0; i < g->n; ++i)
    if (g->row[i + 1] < g->row[i]) return GDB_GRAPH_E_FORMAT;
  for (int k = 0; k < g->m; ++k)
    if (g->col[k] < 0 || g->col[k] >= g->n) return GDB_GRAPH_E_FORMAT;
  return GDB_GRAPH_OK;
}

}  // extern "C"
