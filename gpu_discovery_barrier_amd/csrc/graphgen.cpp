// Seeded synthetic graph generators and CSR utilities (host C++).
//
// The reference ships no datasets (pannotia/datasets/README.txt only links
// an external download), so every workload is generated here from a
// counter-based PRNG (splitmix64 of seed and a draw index), which makes each
// graph a pure function of its parameters on any host and thread count.
//
// All generated graphs are symmetric, free of self-loops and duplicate edges,
// and stored as int32 CSR with each row sorted by destination.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>
#include "../../include/gdb_graph.h"

namespace {

inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
inline uint64_t draw(uint64_t seed, uint64_t stream, uint64_t i) {
  return splitmix64(splitmix64(seed * 0x100000001B3ull + stream) ^ (i * 0xD6E8FEB86659FD93ull));
}
inline double unit(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

int hw_threads() {
  unsigned t = std::thread::hardware_concurrency();
  if (t == 0) t = 4;
  if (t > 32) t = 32;
  return (int)t;
}

template <typename F>
void parallel_for(long long n, F f) {
  const int T = hw_threads();
  if (n < 1 << 16 || T == 1) {
    f(0LL, n);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < T; ++t) {
    const long long b = n * t / T, e = n * (t + 1) / T;
    th.emplace_back([=] { f(b, e); });
  }
  for (auto& x : th) x.join();
}

// Symmetric weight of undirected edge {u, v} in [lo, hi].
inline int edge_weight(uint64_t seed, int u, int v, int lo, int hi) {
  if (hi <= lo) return lo;
  const int a = std::min(u, v), b = std::max(u, v);
  const uint64_t r = draw(seed, 0x77ull, ((uint64_t)(uint32_t)a << 32) | (uint32_t)b);
  return lo + (int)(r % (uint64_t)(hi - lo + 1));
}

// Builds a symmetric, deduplicated, self-loop-free CSR from undirected pairs.
int build_symmetric(int n, const std::vector<uint64_t>& pairs, uint64_t wseed, int wlo, int whi,
                    gdb_host_csr* out) {
  std::vector<long long> deg(n + 1, 0);
  for (uint64_t p : pairs) {
    const int u = (int)(p >> 32), v = (int)(uint32_t)p;
    if (u == v) continue;
    ++deg[u];
    ++deg[v];
  }
  std::vector<long long> off(n + 1, 0);
  for (int i = 0; i < n; ++i) off[i + 1] = off[i] + deg[i];
  std::vector<int> tmp((size_t)off[n]);
  std::vector<long long> pos(off.begin(), off.end() - 1);
  for (uint64_t p : pairs) {
    const int u = (int)(p >> 32), v = (int)(uint32_t)p;
    if (u == v) continue;
    tmp[pos[u]++] = v;
    tmp[pos[v]++] = u;
  }
  std::vector<int> cnt(n, 0);
  parallel_for(n, [&](long long b, long long e) {
    for (long long i = b; i < e; ++i) {
      int* s = tmp.data() + off[i];
      int* t = tmp.data() + off[i + 1];
      std::sort(s, t);
      cnt[i] = (int)(std::unique(s, t) - s);
    }
  });
  long long m = 0;
  for (int i = 0; i < n; ++i) m += cnt[i];
  if (m > 0x7fffffffLL) return GDB_GRAPH_E_TOO_LARGE;
  out->n = n;
  out->m = (int)m;
  out->row = (int*)malloc(sizeof(int) * (size_t)(n + 1));
  out->col = (int*)malloc(sizeof(int) * (size_t)(m ? m : 1));
  out->w = (int*)malloc(sizeof(int) * (size_t)(m ? m : 1));
  if (!out->row || !out->col || !out->w) return GDB_GRAPH_E_NOMEM;
  out->row[0] = 0;
  for (int i = 0; i < n; ++i) out->row[i + 1] = out->row[i] + cnt[i];
  parallel_for(n, [&](long long b, long long e) {
    for (long long i = b; i < e; ++i) {
      const int r0 = out->row[i];
      for (int k = 0; k < cnt[i]; ++k) {
        const int v = tmp[off[i] + k];
        out->col[r0 + k] = v;
        out->w[r0 + k] = edge_weight(wseed, (int)i, v, wlo, whi);
      }
    }
  });
  return GDB_GRAPH_OK;
}

}  // namespace

extern "C" {

void gdb_csr_free(gdb_host_csr* g) {
  if (!g) return;
  free(g->row);
  free(g->col);
  free(g->w);
  g->row = g->col = g->w = nullptr;
  g->n = g->m = 0;
}

int gdb_gen_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed, int wlo,
                 int whi, gdb_host_csr* out) {
  if (!out || scale < 1 || scale > 30 || edge_factor < 1 || a < 0 || b < 0 || c < 0 ||
      a + b + c > 1.0)
    return GDB_GRAPH_E_INVALID;
  *out = gdb_host_csr{};
  const int n = 1 << scale;
  const long long pairs_n = (long long)edge_factor << scale;
  std::vector<uint64_t> pairs((size_t)pairs_n);
  const double ab = a + b, abc = a + b + c;
  parallel_for(pairs_n, [&](long long lo, long long hi) {
    for (long long i = lo; i < hi; ++i) {
      uint32_t u = 0, v = 0;
      for (int l = 0; l < scale; ++l) {
        const double x = unit(draw(seed, 1 + (uint64_t)l, (uint64_t)i));
        const uint32_t bit = 1u << (scale - 1 - l);
        if (x < a) {
        } else if (x < ab) {
          v |= bit;
        } else if (x < abc) {
          u |= bit;
        } else {
          u |= bit;
          v |= bit;
        }
      }
      pairs[(size_t)i] = ((uint64_t)u << 32) | v;
    }
  });
  return build_symmetric(n, pairs, seed, wlo, whi, out);
}

int gdb_gen_grid(int rows, int cols, double remove_frac, uint64_t seed, int wlo, int whi,
                 gdb_host_csr* out) {
  if (!out || rows < 1 || cols < 1 || remove_frac < 0 || remove_frac >= 1 ||
      (long long)rows * cols > 0x7fffffffLL)
    return GDB_GRAPH_E_INVALID;
  *out = gdb_host_csr{};
  const int n = rows * cols;
  const long long nh = (long long)rows * (cols - 1);
  const long long nv = (long long)(rows - 1) * cols;
  const long long ne = nh + nv;
  // Undirected edge e < nh: horizontal (r, c)-(r, c+1) with r = e / (cols-1);
  // otherwise vertical (r, c)-(r+1, c).
  auto endpoints = [&](long long e) -> uint64_t {
    int u, v;
    if (e < nh) {
      const int r = (int)(e / (cols - 1)), cc = (int)(e % (cols - 1));
      u = r * cols + cc;
      v = u + 1;
    } else {
      const long long f = e - nh;
      const int r = (int)(f / cols), cc = (int)(f % cols);
      u = r * cols + cc;
      v = u + cols;
    }
    return ((uint64_t)(uint32_t)u << 32) | (uint32_t)v;
  };
  const long long k = std::llround(remove_frac * (double)ne);
  std::vector<char> removed((size_t)ne, 0);
  if (k > 0) {
    // Partial Fisher-Yates: the first k slots of a seeded shuffle of 0..ne-1.
    std::vector<long long> idx((size_t)ne);
    for (long long i = 0; i < ne; ++i) idx[(size_t)i] = i;
    for (long long i = 0; i < k; ++i) {
      const long long j = i + (long long)(draw(seed, 2, (uint64_t)i) % (uint64_t)(ne - i));
      std::swap(idx[(size_t)i], idx[(size_t)j]);
      removed[(size_t)idx[(size_t)i]] = 1;
    }
  }
  std::vector<uint64_t> pairs;
  pairs.reserve((size_t)(ne - k));
  for (long long e = 0; e < ne; ++e)
    if (!removed[(size_t)e]) pairs.push_back(endpoints(e));
  return build_symmetric(n, pairs, seed, wlo, whi, out);
}

int gdb_gen_gnp(int n, double avg_degree, uint64_t seed, int wlo, int whi, gdb_host_csr* out) {
  if (!out || n < 1 || avg_degree < 0) return GDB_GRAPH_E_INVALID;
  *out = gdb_host_csr{};
  std::vector<uint64_t> pairs;
  if (n > 1 && avg_degree > 0) {
    const double p = std::min(1.0, avg_degree / (double)(n - 1));
    // Geometric skipping over the n(n-1)/2 pairs u < v in row-major order.
    const double lq = std::log(1.0 - p);
    long long u = 1, v = -1;
    uint64_t i = 0;
    while (u < n) {
      const double r = unit(draw(seed, 3, i++));
      const long long skip = p >= 1.0 ? 0 : (long long)std::floor(std::log(1.0 - r) / lq);
      v += 1 + skip;
      while (v >= u && u < n) {
        v -= u;
        ++u;
      }
      if (u < n) pairs.push_back(((uint64_t)(uint32_t)v << 32) | (uint32_t)u);
    }
  }
  return build_symmetric(n, pairs, seed, wlo, whi, out);
}

int gdb_gen_permutation(int n, uint64_t seed, int* out) {
  if (!out || n < 0) return GDB_GRAPH_E_INVALID;
  for (int i = 0; i < n; ++i) out[i] = i;
  for (int i = n - 1; i > 0; --i) {
    const int j = (int)(draw(seed, 5, (uint64_t)i) % (uint64_t)(i + 1));
    std::swap(out[i], out[j]);
  }
  return GDB_GRAPH_OK;
}

int gdb_gen_uniform_ints(long long count, int lo, int hi, uint64_t seed, int* out) {
  if (!out || count < 0 || hi < lo) return GDB_GRAPH_E_INVALID;
  const uint64_t span = (uint64_t)((long long)hi - lo + 1);
  parallel_for(count, [&](long long b, long long e) {
    for (long long i = b; i < e; ++i) out[i] = (int)((long long)lo + (long long)(draw(seed, 6, (uint64_t)i) % span));
  });
  return GDB_GRAPH_OK;
}

int gdb_csr_transpose(const gdb_host_csr* in, gdb_host_csr* out) {
  if (!in || !out || in->n < 0 || in->m < 0) return GDB_GRAPH_E_INVALID;
  *out = gdb_host_csr{};
  const int n = in->n, m = in->m;
  out->n = n;
  out->m = m;
  out->row = (int*)calloc((size_t)n + 1, sizeof(int));
  out->col = (int*)malloc(sizeof(int) * (size_t)(m ? m : 1));
  out->w = (int*)malloc(sizeof(int) * (size_t)(m ? m : 1));
  if (!out->row || !out->col || !out->w) return GDB_GRAPH_E_NOMEM;
  for (int k = 0; k < m; ++k) out->row[in->col[k] + 1]++;
  for (int i = 0; i < n; ++i) out->row[i + 1] += out->row[i];
  std::vector<int> pos(out->row, out->row + n);
  // Rows of the input are visited in order, so each output row lists its
  // sources in increasing order.
  for (int u = 0; u < n; ++u)
    for (int k = in->row[u]; k < in->row[u + 1]; ++k) {
      const int v = in->col[k];
      const int p = pos[v]++;
      out->col[p] = u;
      out->w[p] = in->w ? in->w[k] : 1;
    }
  return GDB_GRAPH_OK;
}

int gdb_csr_relabel(const gdb_host_csr* in, const int* new_id, gdb_host_csr* out) {
  if (!in || !new_id || !out || in->n < 0 || in->m < 0) return GDB_GRAPH_E_INVALID;
  *out = gdb_host_csr{};
  const int n = in->n, m = in->m;
  std::vector<int> old_of((size_t)n, -1);
  for (int v = 0; v < n; ++v) {
    const int x = new_id[v];
    if (x < 0 || x >= n || old_of[x] != -1) return GDB_GRAPH_E_INVALID;  // not a permutation
    old_of[x] = v;
  }
  out->n = n;
  out->m = m;
  out->row = (int*)calloc((size_t)n + 1, sizeof(int));
  out->col = (int*)malloc(sizeof(int) * (size_t)(m ? m : 1));
  out->w = (int*)malloc(sizeof(int) * (size_t)(m ? m : 1));
  if (!out->row || !out->col || !out->w) return GDB_GRAPH_E_NOMEM;
  for (int x = 0; x < n; ++x) {
    const int v = old_of[x];
    out->row[x + 1] = out->row[x] + (in->row[v + 1] - in->row[v]);
  }
  std::vector<std::pair<int, int>> nb;
  for (int x = 0; x < n; ++x) {
    const int v = old_of[x];
    nb.clear();
    for (int k = in->row[v]; k < in->row[v + 1]; ++k) {
      const int u = in->col[k];
      if (u < 0 || u >= n) return GDB_GRAPH_E_INVALID;
      nb.emplace_back(new_id[u], in->w ? in->w[k] : 1);
    }
    std::stable_sort(nb.begin(), nb.end(), [](const std::pair<int, int>& a, const std::pair<int, int>& b) {
      return a.first < b.first;
    });
    int p = out->row[x];
    for (const auto& e : nb) {
      out->col[p] = e.first;
      out->w[p] = e.second;
      ++p;
    }
  }
  return GDB_GRAPH_OK;
}

int gdb_order_by_degree(const gdb_host_csr* g, int* new_id) {
  if (!g || !new_id || g->n < 0) return GDB_GRAPH_E_INVALID;
  std::vector<int> order((size_t)g->n);
  for (int v = 0; v < g->n; ++v) order[v] = v;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return g->row[a + 1] - g->row[a] > g->row[b + 1] - g->row[b];
  });
  for (int x = 0; x < g->n; ++x) new_id[order[x]] = x;
  return GDB_GRAPH_OK;
}

}  // extern "C"
