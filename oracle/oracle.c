/* CPU oracle for the discovery-barrier graph path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker.
 *
 * Citations are to the reference tree (mc-imperial/gpu_discovery_barrier):
 *   PAN/ = code/experiments/pannotia/, LS/ = code/experiments/lonestar_gpu_opencl/,
 *   DP/  = code/discovery_protocol/.
 *
 * Pinning (which independent check fixes each restatement) is listed in the
 * docstrings of tests/test_oracle.py; every orc_*_ref function here has one.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <sched.h>
#include <semaphore.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------------ */
/* Colouring                                                                 */
/* ------------------------------------------------------------------------ */

/* PAN/graph_app/color/kernel/color_kernel.cl:180-251 (mega_kernel) executed
 * as sequential synchronous rounds; phase 1 reads only, phase 2 writes only
 * the vertex's own colour, so the order inside a phase does not matter. */
int orc_color_ref(int n, const int* row, const int* col, const float* nv, int* color) {
  float* max_d = (float*)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
  if (!max_d) return -2;
  for (int i = 0; i < n; ++i) color[i] = -1; /* color_gb.cpp:148-153 */
  int graph_color = 1;
  for (;;) {
    int stop = 0;
    for (int i = 0; i < n; ++i) { /* lines 185-213 */
      if (color[i] != -1) continue;
      const int start = row[i], end = row[i + 1];
      float maximum = -1;
      for (int e = start; e < end; ++e) {
        if (color[col[e]] == -1 && start != end - 1) { /* line 204: degree-1 quirk */
          stop = 1;
          if (nv[col[e]] > maximum) maximum = nv[col[e]];
        }
      }
      max_d[i] = maximum;
    }
    for (int i = 0; i < n; ++i) /* lines 229-238 */
      if (color[i] == -1 && nv[i] > max_d[i]) color[i] = graph_color;
    if (stop == 0) break; /* lines 240-242 */
    if (graph_color > n) {
      free(max_d);
      return -1;
    }
    ++graph_color;
  }
  free(max_d);
  return graph_color;
}

/* Max-min colouring (no reference counterpart: the reference colours with
 * the max rule only, color_kernel.cl:229-237; the max-min rule is the
 * two-sided variant of the same Jones-Plassmann-style scheme).  Round
 * r = 1, 2, ...: among the uncoloured vertices, v takes colour 2r-1 when
 * its value exceeds every uncoloured neighbour's (self-loops ignored) and
 * colour 2r when it is below every one of them; decisions use the state at
 * the start of the round.  Returns the number of rounds that coloured
 * something (0 for an empty graph), or -1 when n+1 rounds pass with
 * vertices left (adjacent equal values never resolve). */
int orc_color_maxmin_ref(int n, const int* row, const int* col, const float* nv, int* color) {
  for (int i = 0; i < n; ++i) color[i] = -1;
  int* next = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  if (!next) return -2;
  int r = 0, left = n;
  while (left > 0) {
    ++r;
    if (r > n + 1) {
      free(next);
      return -1;
    }
    for (int v = 0; v < n; ++v) {
      next[v] = color[v];
      if (color[v] != -1) continue;
      int is_max = 1, is_min = 1;
      for (int e = row[v]; e < row[v + 1]; ++e) {
        const int u = col[e];
        if (u == v || color[u] != -1) continue;
        if (!(nv[v] > nv[u])) is_max = 0;
        if (!(nv[v] < nv[u])) is_min = 0;
      }
      if (is_max) next[v] = 2 * r - 1;
      else if (is_min) next[v] = 2 * r;
    }
    for (int v = 0; v < n; ++v) {
      if (color[v] == -1 && next[v] != -1) --left;
      color[v] = next[v];
    }
  }
  free(next);
  return r;
}

static const float* g_sort_key;
static int cmp_desc(const void* a, const void* b) {
  const float x = g_sort_key[*(const int*)a], y = g_sort_key[*(const int*)b];
  return (x < y) - (x > y);
}
static int cmp_asc(const void* a, const void* b) { return -cmp_desc(a, b); }

int orc_color_tb(int n, const int* row, const int* col, const float* nv, int* color) {
  int* order = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  if (!order) return -2;
  for (int i = 0; i < n; ++i) order[i] = i;
  g_sort_key = nv;
  qsort(order, (size_t)n, sizeof(int), cmp_desc);
  int rmax = 0;
  for (int t = 0; t < n; ++t) {
    const int v = order[t];
    int r = 1;
    if (row[v + 1] - row[v] != 1) {
      int best = 0;
      for (int e = row[v]; e < row[v + 1]; ++e) {
        const int u = col[e];
        if (nv[u] > nv[v] && color[u] > best) best = color[u];
      }
      r = best + 1;
    }
    color[v] = r;
    if (r > rmax) rmax = r;
  }
  free(order);
  return rmax;
}

/* ------------------------------------------------------------------------ */
/* Maximal independent set                                                   */
/* ------------------------------------------------------------------------ */

/* PAN/graph_app/mis/kernel/mis_kernel.cl:220-337 (mega_kernel) as sequential
 * rounds; init per lines 78-93. */
int orc_mis_ref(int n, const int* row, const int* col, const float* nv, int* s) {
  const size_t sz = (size_t)(n > 0 ? n : 1);
  int* c = (int*)malloc(sizeof(int) * sz);
  int* cu = (int*)malloc(sizeof(int) * sz);
  float* min_array = (float*)malloc(sizeof(float) * sz);
  if (!c || !cu || !min_array) return -2;
  for (int i = 0; i < n; ++i) {
    c[i] = -1;
    cu[i] = -1;
    s[i] = 0;
  }
  int rounds = 0;
  for (;;) {
    ++rounds;
    int stop = 0;
    for (int t = 0; t < n; ++t) { /* mis1, lines 240-266 */
      if (c[t] != -1) continue;
      stop = 1;
      float mn = ORC_MIS_BIGNUM;
      for (int e = row[t]; e < row[t + 1]; ++e)
        if (c[col[e]] == -1 && nv[col[e]] < mn) mn = nv[col[e]];
      min_array[t] = mn;
    }
    for (int t = 0; t < n; ++t) { /* mis2, lines 276-313 */
      if (nv[t] < min_array[t] && c[t] == -1) {
        s[t] = 2;
        c[t] = -2;
        for (int e = row[t]; e < row[t + 1]; ++e)
          if (c[col[e]] == -1) cu[col[e]] = -2;
      }
    }
    if (stop == 0) break; /* lines 322-325 */
    for (int t = 0; t < n; ++t) /* mis3, lines 330-334 */
      if (cu[t] == -2) c[t] = cu[t];
    if (rounds > n + 1) {
      rounds = -1;
      break;
    }
  }
  free(c);
  free(cu);
  free(min_array);
  return rounds;
}

int orc_mis_tb(int n, const int* row, const int* col, const float* nv, int* s) {
  int* order = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  if (!order) return -2;
  for (int i = 0; i < n; ++i) {
    order[i] = i;
    s[i] = 0;
  }
  g_sort_key = nv;
  qsort(order, (size_t)n, sizeof(int), cmp_asc);
  int size = 0;
  for (int t = 0; t < n; ++t) {
    const int v = order[t];
    int ok = 1;
    for (int e = row[v]; e < row[v + 1] && ok; ++e)
      if (s[col[e]] == 2) ok = 0;
    if (ok) {
      s[v] = 2;
      ++size;
    }
  }
  free(order);
  return size;
}

/* ------------------------------------------------------------------------ */
/* Shortest paths                                                            */
/* ------------------------------------------------------------------------ */

/* PAN/graph_app/sssp/kernel/sssp_kernel.cl: vector_init (111-132) then
 * mega_kernel (170-244): for k = 1 .. n-1 { x = y; y = min(x, min(w + x[u]));
 * stop if x == y }.  Output is x (sssp_gb.cpp:308). */
int orc_sssp_pull_ref(int n, const int* row_t, const int* col_t, const int* w_t, int src, int* x) {
  int* y = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  if (!y) return -2;
  for (int i = 0; i < n; ++i) x[i] = y[i] = (i == src) ? 0 : ORC_PANNOTIA_INF;
  int k = 1;
  for (; k < n; ++k) {
    for (int i = 0; i < n; ++i) x[i] = y[i];
    for (int it = 0; it < n; ++it) {
      int mn = x[it];
      for (int j = row_t[it]; j < row_t[it + 1]; ++j)
        if (w_t[j] + x[col_t[j]] < mn) mn = w_t[j] + x[col_t[j]];
      y[it] = mn;
    }
    int stop = 0;
    for (int i = 0; i < n; ++i)
      if (y[i] != x[i]) stop = 1;
    if (stop == 0) break;
  }
  free(y);
  return k < n ? k : n - 1; /* rounds executed */
}

typedef struct {
  long long d;
  int v;
} heap_item;

static void heap_push(heap_item* h, int* sz, heap_item it) {
  int i = (*sz)++;
  h[i] = it;
  while (i > 0) {
    int p = (i - 1) / 2;
    if (h[p].d <= h[i].d) break;
    heap_item t = h[p];
    h[p] = h[i];
    h[i] = t;
    i = p;
  }
}
static heap_item heap_pop(heap_item* h, int* sz) {
  heap_item top = h[0];
  h[0] = h[--(*sz)];
  int i = 0;
  for (;;) {
    int l = 2 * i + 1, r = l + 1, m = i;
    if (l < *sz && h[l].d < h[m].d) m = l;
    if (r < *sz && h[r].d < h[m].d) m = r;
    if (m == i) break;
    heap_item t = h[m];
    h[m] = h[i];
    h[i] = t;
    i = m;
  }
  return top;
}

int orc_dijkstra_tb(int n, const int* row, const int* col, const int* w, int src, long long inf,
                    long long* dist) {
  for (int i = 0; i < n; ++i) dist[i] = inf;
  if (n == 0) return 0;
  const long long m = row[n];
  heap_item* h = (heap_item*)malloc(sizeof(heap_item) * (size_t)(m + 2));
  if (!h) return -2;
  int sz = 0;
  dist[src] = 0;
  heap_push(h, &sz, (heap_item){0, src});
  while (sz > 0) {
    heap_item it = heap_pop(h, &sz);
    if (it.d != dist[it.v]) continue;
    for (int e = row[it.v]; e < row[it.v + 1]; ++e) {
      const long long nd = it.d + (w ? w[e] : 1);
      if (nd < dist[col[e]]) {
        dist[col[e]] = nd;
        heap_push(h, &sz, (heap_item){nd, col[e]});
      }
    }
  }
  free(h);
  return 0;
}

int orc_bellman_ford_tb(int n, const int* row, const int* col, const int* w, int src, long long inf,
                        long long* dist) {
  for (int i = 0; i < n; ++i) dist[i] = inf;
  if (n == 0) return 0;
  dist[src] = 0;
  for (int it = 0; it < n; ++it) {
    int changed = 0;
    for (int u = 0; u < n; ++u) {
      if (dist[u] == inf) continue;
      for (int e = row[u]; e < row[u + 1]; ++e) {
        const long long nd = dist[u] + (w ? w[e] : 1);
        if (nd < dist[col[e]]) {
          dist[col[e]] = nd;
          changed = 1;
        }
      }
    }
    if (!changed) break;
  }
  return 0;
}

/* LS/apps/bfs: initialize (bfs_common.cl:64-71), dist[source] = 0
 * (bfs_port.cpp:112-121), iteration 0 pushes the source
 * (bfs_port_kernel.cl:130-137), then per iteration every popped vertex
 * relaxes its edges with processedge2 (bfs_common.cl:73-90): an unvisited
 * destination gets dist = iteration and is pushed. */
int orc_bfs_worklist_ref(int n, const int* row, const int* col, int src, uint32_t* dist) {
  for (int i = 0; i < n; ++i) dist[i] = ORC_LONESTAR_INF;
  if (n == 0) return 0;
  const long long m = row[n];
  int* in = (int*)malloc(sizeof(int) * (size_t)(m + n + 1));
  int* out = (int*)malloc(sizeof(int) * (size_t)(m + n + 1));
  if (!in || !out) return -2;
  dist[src] = 0;
  int nin = 1, iteration = 1;
  in[0] = src;
  while (nin > 0) { /* bfs_port_kernel.cl:173-189 */
    int nout = 0;
    for (int t = 0; t < nin; ++t) {
      const int nn = in[t];
      for (int e = row[nn]; e < row[nn + 1]; ++e) {
        const int dst = col[e];
        if (dst >= n) continue;
        if (dist[dst] == ORC_LONESTAR_INF) {
          dist[dst] = (uint32_t)iteration;
          out[nout++] = dst;
        }
      }
    }
    int* tmp = in;
    in = out;
    out = tmp;
    nin = nout;
    ++iteration;
  }
  free(in);
  free(out);
  return iteration - 1;
}

/* LS/apps/sssp: processedge2 (sssp_common.cl:58-81): alt = dist[src] + w;
 * if alt < dist[dst] { atomic_min(dist[dst], alt); push dst }, iterated over
 * worklists until empty (sssp_port_kernel.cl:152-197). */
int orc_sssp_worklist_ref(int n, const int* row, const int* col, const int* w, int src,
                          uint32_t* dist) {
  for (int i = 0; i < n; ++i) dist[i] = ORC_LONESTAR_INF;
  if (n == 0) return 0;
  const long long m = row[n];
  size_t cap = (size_t)(m + n + 1);
  int* in = (int*)malloc(sizeof(int) * cap);
  int* out = (int*)malloc(sizeof(int) * cap);
  char* queued = (char*)calloc((size_t)n, 1);
  if (!in || !out || !queued) return -2;
  dist[src] = 0;
  int nin = 1, iters = 0;
  in[0] = src;
  while (nin > 0) {
    int nout = 0;
    memset(queued, 0, (size_t)n);
    for (int t = 0; t < nin; ++t) {
      const int nn = in[t];
      for (int e = row[nn]; e < row[nn + 1]; ++e) {
        const int dst = col[e];
        if (dst >= n) continue;
        const uint32_t wt = (uint32_t)w[e];
        if (wt >= ORC_LONESTAR_INF) continue;
        const uint32_t alt = dist[nn] + wt;
        if (alt < dist[dst]) {
          dist[dst] = alt;
          if (!queued[dst]) {
            queued[dst] = 1;
            out[nout++] = dst;
          }
        }
      }
    }
    int* tmp = in;
    in = out;
    out = tmp;
    nin = nout;
    ++iters;
  }
  free(in);
  free(out);
  free(queued);
  return iters;
}

/* ---- multi-threaded worklist BFS / SSSP (the CPU baseline) ---------------
 * The same worklist iterations as orc_bfs_worklist_ref / orc_sssp_worklist_ref
 * (LS/apps/bfs/kernel/bfs_port_kernel.cl:146-192, sssp_port_kernel.cl:152-197)
 * run by `threads` persistent pthreads, one per core: each iteration the
 * threads take the input worklist with a grid stride (item t, t + T, ...),
 * relax edges with atomic compare-and-swap (BFS claims an unvisited vertex,
 * SSSP lowers the distance), buffer their pushes and reserve output space
 * with one fetch-add per 1024 items, then meet at a CPU restatement of the
 * XF barrier (DP/api/discovery.cl:129-190: the master waits for every
 * slave's flag, swaps the worklists inside the barrier, then clears the
 * flags).  SSSP dedupes pushes per iteration with a stamp. */
typedef struct {
  _Atomic int v;
  char pad[60];
} orc_pad_int;

typedef struct {
  int n, T;
  const int *row, *col, *w;
  _Atomic uint32_t* dist;
  _Atomic int* stamp;
  int* q[2];
  orc_pad_int out_size;
  orc_pad_int* flags;
  int nin, iter, cur; /* written by the master inside the barrier */
  int iterations;
} orc_wl_shared;

typedef struct {
  orc_wl_shared* S;
  int t;
} orc_wl_arg;

static void orc_spin_pause(unsigned* k) {
  if ((++*k & 1023u) == 0) sched_yield();
#if defined(__x86_64__) || defined(__i386__)
  else __builtin_ia32_pause();
#endif
}

/* XF barrier; `serial` runs on the master between the arrivals and the release. */
static void orc_xf_barrier(orc_wl_shared* S, int t, void (*serial)(orc_wl_shared*)) {
  if (t == 0) {
    for (int p = 1; p < S->T; ++p) {
      unsigned k = 0;
      while (atomic_load_explicit(&S->flags[p].v, memory_order_acquire) == 0) orc_spin_pause(&k);
    }
    serial(S);
    for (int p = 1; p < S->T; ++p) atomic_store_explicit(&S->flags[p].v, 0, memory_order_release);
  } else {
    atomic_store_explicit(&S->flags[t].v, 1, memory_order_release);
    unsigned k = 0;
    while (atomic_load_explicit(&S->flags[t].v, memory_order_acquire) == 1) orc_spin_pause(&k);
  }
}

static void orc_wl_swap(orc_wl_shared* S) {
  S->nin = atomic_load_explicit(&S->out_size.v, memory_order_relaxed);
  atomic_store_explicit(&S->out_size.v, 0, memory_order_relaxed);
  S->cur ^= 1;
  S->iter += 1;
  if (S->nin > 0) S->iterations += 1;
}

static void orc_wl_flush(orc_wl_shared* S, int* out, const int* buf, int cnt) {
  if (cnt == 0) return;
  const int at = atomic_fetch_add_explicit(&S->out_size.v, cnt, memory_order_relaxed);
  memcpy(out + at, buf, sizeof(int) * (size_t)cnt);
}

static void* orc_wl_worker(void* p) {
  orc_wl_arg* a = (orc_wl_arg*)p;
  orc_wl_shared* S = a->S;
  const int t = a->t, T = S->T;
  int buf[1024];
  for (;;) {
    const int nin = S->nin, it = S->iter;
    if (nin == 0) break;
    const int* in = S->q[S->cur];
    int* out = S->q[S->cur ^ 1];
    int cnt = 0;
    for (int i = t; i < nin; i += T) {
      const int v = in[i];
      const uint32_t dv = atomic_load_explicit(&S->dist[v], memory_order_relaxed);
      for (int e = S->row[v]; e < S->row[v + 1]; ++e) {
        const int dst = S->col[e];
        if (dst < 0 || dst >= S->n) continue;
        int push = 0;
        if (!S->w) { /* processedge2, bfs_common.cl:73-90 */
          uint32_t expect = ORC_LONESTAR_INF;
          if (atomic_load_explicit(&S->dist[dst], memory_order_relaxed) == ORC_LONESTAR_INF)
            push = atomic_compare_exchange_strong_explicit(&S->dist[dst], &expect, (uint32_t)it,
                                                           memory_order_relaxed, memory_order_relaxed);
        } else { /* processedge2, sssp_common.cl:58-81 */
          const uint32_t wt = (uint32_t)S->w[e];
          if (wt >= ORC_LONESTAR_INF) continue;
          const uint32_t alt = dv + wt;
          uint32_t cur = atomic_load_explicit(&S->dist[dst], memory_order_relaxed);
          while (alt < cur) {
            if (atomic_compare_exchange_weak_explicit(&S->dist[dst], &cur, alt, memory_order_relaxed,
                                                      memory_order_relaxed)) {
              push = atomic_exchange_explicit(&S->stamp[dst], it, memory_order_relaxed) != it;
              break;
            }
          }
        }
        if (push) {
          buf[cnt++] = dst;
          if (cnt == 1024) {
            orc_wl_flush(S, out, buf, cnt);
            cnt = 0;
          }
        }
      }
    }
    orc_wl_flush(S, out, buf, cnt);
    orc_xf_barrier(S, t, orc_wl_swap);
  }
  return NULL;
}

int orc_worklist_mt(int n, const int* row, const int* col, const int* w, int src, int threads, uint32_t* dist,
                    int* iterations) {
  if (threads < 1) threads = 1;
  if (n <= 0) {
    if (iterations) *iterations = 0;
    return 0;
  }
  orc_wl_shared S;
  memset(&S, 0, sizeof(S));
  S.n = n;
  S.T = threads;
  S.row = row;
  S.col = col;
  S.w = w;
  S.dist = (_Atomic uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
  S.stamp = (_Atomic int*)malloc(sizeof(int) * (size_t)n);
  S.q[0] = (int*)malloc(sizeof(int) * (size_t)n);
  S.q[1] = (int*)malloc(sizeof(int) * (size_t)n);
  S.flags = (orc_pad_int*)calloc((size_t)threads, sizeof(orc_pad_int));
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
  orc_wl_arg* args = (orc_wl_arg*)malloc(sizeof(orc_wl_arg) * (size_t)threads);
  if (!S.dist || !S.stamp || !S.q[0] || !S.q[1] || !S.flags || !th || !args) return -2;
  for (int i = 0; i < n; ++i) {
    atomic_init(&S.dist[i], ORC_LONESTAR_INF);
    atomic_init(&S.stamp[i], 0);
  }
  atomic_init(&S.dist[src], 0u);
  atomic_init(&S.out_size.v, 0);
  for (int p = 0; p < threads; ++p) atomic_init(&S.flags[p].v, 0);
  S.q[0][0] = src; /* iteration 0 pushes the source (bfs_port_kernel.cl:130-137) */
  S.nin = 1;
  S.iter = 1;
  S.cur = 0;
  S.iterations = 1;
  for (int t = 1; t < threads; ++t) {
    args[t].S = &S;
    args[t].t = t;
    pthread_create(&th[t], NULL, orc_wl_worker, &args[t]);
  }
  args[0].S = &S;
  args[0].t = 0;
  orc_wl_worker(&args[0]);
  for (int t = 1; t < threads; ++t) pthread_join(th[t], NULL);
  for (int i = 0; i < n; ++i) dist[i] = atomic_load_explicit(&S.dist[i], memory_order_relaxed);
  if (iterations) *iterations = S.iterations;
  free((void*)S.dist);
  free((void*)S.stamp);
  free(S.q[0]);
  free(S.q[1]);
  free(S.flags);
  free(th);
  free(args);
  return 0;
}

int orc_bfs_tb(int n, const int* row, const int* col, int src, uint32_t* dist) {
  for (int i = 0; i < n; ++i) dist[i] = ORC_LONESTAR_INF;
  if (n == 0) return 0;
  int* q = (int*)malloc(sizeof(int) * (size_t)n);
  if (!q) return -2;
  int head = 0, tail = 0;
  dist[src] = 0;
  q[tail++] = src;
  while (head < tail) {
    const int u = q[head++];
    for (int e = row[u]; e < row[u + 1]; ++e)
      if (dist[col[e]] == ORC_LONESTAR_INF) {
        dist[col[e]] = dist[u] + 1;
        q[tail++] = col[e];
      }
  }
  free(q);
  return 0;
}

/* dverifysolution: LS/apps/bfs/kernel/bfs_common.cl:43-60 (w = 1) and
 * LS/apps/sssp/kernel/sssp_common.cl:41-56. */
long long orc_verify_distances_ref(int n, const int* row, const int* col, const int* w,
                                   const uint32_t* dist) {
  long long nerr = 0;
  for (int u = 0; u < n; ++u)
    for (int e = row[u]; e < row[u + 1]; ++e) {
      const uint32_t wt = w ? (uint32_t)w[e] : 1u;
      if (wt > 0 && dist[u] + wt < dist[col[e]]) ++nerr;
    }
  return nerr;
}

/* ------------------------------------------------------------------------ */
/* Betweenness centrality                                                    */
/* ------------------------------------------------------------------------ */

int orc_bc_brandes_tb(int n, const int* row, const int* col, const int* sources, int nsrc,
                      double* bc) {
  const size_t sz = (size_t)(n > 0 ? n : 1);
  int* d = (int*)malloc(sizeof(int) * sz);
  int* order = (int*)malloc(sizeof(int) * sz);
  double* sigma = (double*)malloc(sizeof(double) * sz);
  double* delta = (double*)malloc(sizeof(double) * sz);
  if (!d || !order || !sigma || !delta) return -2;
  for (int i = 0; i < n; ++i) bc[i] = 0.0;
  for (int si = 0; si < nsrc; ++si) {
    const int s = sources[si];
    for (int i = 0; i < n; ++i) {
      d[i] = -1;
      sigma[i] = 0.0;
      delta[i] = 0.0;
    }
    int head = 0, tail = 0;
    d[s] = 0;
    sigma[s] = 1.0;
    order[tail++] = s;
    while (head < tail) {
      const int v = order[head++];
      for (int e = row[v]; e < row[v + 1]; ++e) {
        const int w = col[e];
        if (d[w] < 0) {
          d[w] = d[v] + 1;
          order[tail++] = w;
        }
        if (d[w] == d[v] + 1) sigma[w] += sigma[v];
      }
    }
    for (int t = tail - 1; t >= 0; --t) {
      const int v = order[t];
      double acc = 0.0;
      for (int e = row[v]; e < row[v + 1]; ++e) {
        const int w = col[e];
        if (d[w] == d[v] + 1) acc += sigma[v] / sigma[w] * (1.0 + delta[w]);
      }
      delta[v] = acc;
      if (v != s) bc[v] += delta[v];
    }
  }
  free(d);
  free(order);
  free(sigma);
  free(delta);
  return 0;
}

/* orc_bc_brandes_tb split over threads: thread t takes sources t, t + T, ...
 * into its own FP64 accumulator; the accumulators are summed in thread order,
 * so the result does not depend on scheduling. */
typedef struct {
  int n, nsrc, t, T, rc;
  const int *row, *col, *sources;
  double* acc;
} bc_mt_arg;

static void* bc_mt_worker(void* p) {
  bc_mt_arg* a = (bc_mt_arg*)p;
  int cnt = 0;
  for (int i = a->t; i < a->nsrc; i += a->T) ++cnt;
  int* mine = (int*)malloc(sizeof(int) * (size_t)(cnt > 0 ? cnt : 1));
  if (!mine) {
    a->rc = -2;
    return NULL;
  }
  cnt = 0;
  for (int i = a->t; i < a->nsrc; i += a->T) mine[cnt++] = a->sources[i];
  a->rc = orc_bc_brandes_tb(a->n, a->row, a->col, mine, cnt, a->acc);
  free(mine);
  return NULL;
}

int orc_bc_brandes_mt(int n, const int* row, const int* col, const int* sources, int nsrc, int threads,
                      double* bc) {
  if (threads < 1) threads = 1;
  if (threads > nsrc) threads = nsrc > 0 ? nsrc : 1;
  const size_t sz = (size_t)(n > 0 ? n : 1);
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
  bc_mt_arg* args = (bc_mt_arg*)calloc((size_t)threads, sizeof(bc_mt_arg));
  double* acc = (double*)malloc(sizeof(double) * sz * (size_t)threads);
  if (!th || !args || !acc) return -2;
  for (int t = 0; t < threads; ++t) {
    bc_mt_arg a = {n, nsrc, t, threads, 0, row, col, sources, acc + sz * (size_t)t};
    args[t] = a;
    pthread_create(&th[t], NULL, bc_mt_worker, &args[t]);
  }
  int rc = 0;
  for (int t = 0; t < threads; ++t) {
    pthread_join(th[t], NULL);
    if (args[t].rc) rc = args[t].rc;
  }
  for (int i = 0; i < n; ++i) {
    double v = 0.0;
    for (int t = 0; t < threads; ++t) v += acc[sz * (size_t)t + (size_t)i];
    bc[i] = v;
  }
  free(th);
  free(args);
  free(acc);
  return rc;
}

static void bfs_counts(int n, const int* row, const int* col, int s, int* d, double* sigma, int* q) {
  for (int i = 0; i < n; ++i) {
    d[i] = -1;
    sigma[i] = 0.0;
  }
  int head = 0, tail = 0;
  d[s] = 0;
  sigma[s] = 1.0;
  q[tail++] = s;
  while (head < tail) {
    const int v = q[head++];
    for (int e = row[v]; e < row[v + 1]; ++e) {
      const int w = col[e];
      if (d[w] < 0) {
        d[w] = d[v] + 1;
        q[tail++] = w;
      }
      if (d[w] == d[v] + 1) sigma[w] += sigma[v];
    }
  }
}

int orc_bc_pairs_tb(int n, const int* row, const int* col, const int* sources, int nsrc,
                    double* bc) {
  const size_t nn = (size_t)(n > 0 ? n : 1);
  int* dall = (int*)malloc(sizeof(int) * nn * nn);
  double* sall = (double*)malloc(sizeof(double) * nn * nn);
  int* q = (int*)malloc(sizeof(int) * nn);
  if (!dall || !sall || !q) return -2;
  for (int v = 0; v < n; ++v) bfs_counts(n, row, col, v, dall + (size_t)v * nn, sall + (size_t)v * nn, q);
  for (int i = 0; i < n; ++i) bc[i] = 0.0;
  for (int si = 0; si < nsrc; ++si) {
    const int s = sources[si];
    const int* ds = dall + (size_t)s * nn;
    const double* ss = sall + (size_t)s * nn;
    for (int t = 0; t < n; ++t) {
      if (t == s || ds[t] < 0) continue;
      for (int v = 0; v < n; ++v) {
        if (v == s || v == t || ds[v] < 0) continue;
        const int* dv = dall + (size_t)v * nn;
        const double* sv = sall + (size_t)v * nn;
        if (dv[t] >= 0 && ds[v] + dv[t] == ds[t]) bc[v] += ss[v] * sv[t] / ss[t];
      }
    }
  }
  free(dall);
  free(sall);
  free(q);
  return 0;
}

/* PAN/graph_app/bc/kernel/bc_kernel.cl: per source, clean_1d_array
 * (95-116), mega_bfs_kernel_func (295-371), mega_backtrack_kernel_func
 * (373-428), in FP32 with a sequential order inside each level. */
int orc_bc_ref_f32(int n, const int* row, const int* col, const int* row_t, const int* col_t,
                   const int* sources, int nsrc, float* bc) {
  const size_t sz = (size_t)(n > 0 ? n : 1);
  int* d = (int*)malloc(sizeof(int) * sz);
  float* rho = (float*)malloc(sizeof(float) * sz);
  float* sigma = (float*)malloc(sizeof(float) * sz);
  if (!d || !rho || !sigma) return -2;
  for (int i = 0; i < n; ++i) bc[i] = 0.f; /* clean_bc, lines 137-143 */
  for (int si = 0; si < nsrc; ++si) {
    const int s = sources[si];
    for (int i = 0; i < n; ++i) {
      sigma[i] = 0.f;
      rho[i] = (i == s) ? 1.f : 0.f;
      d[i] = (i == s) ? 0 : -1;
    }
    int local_dist = 0;
    for (;;) {
      int cont = 0;
      for (int i = 0; i < n; ++i) {
        if (d[i] != local_dist) continue;
        for (int e = row[i]; e < row[i + 1]; ++e) {
          const int w = col[e];
          if (d[w] < 0) {
            cont = 1;
            d[w] = local_dist + 1;
          }
          if (d[w] == local_dist + 1) rho[w] += rho[i];
        }
      }
      local_dist += 1;
      if (!cont) break;
    }
    for (; local_dist > 0; --local_dist) {
      for (int i = 0; i < n; ++i) {
        if (d[i] != local_dist - 1) continue;
        for (int e = row_t[i]; e < row_t[i + 1]; ++e) {
          const int w = col_t[e];
          if (d[w] == local_dist - 2) sigma[w] += rho[w] / rho[i] * (1 + sigma[i]);
        }
        if (i != s) bc[i] = bc[i] + sigma[i];
      }
    }
  }
  free(d);
  free(rho);
  free(sigma);
  return 0;
}

/* LS/include/block_scan_cl.h:7-57. */
int orc_block_scan_ref(int n_items, const int* in, int* out) {
  if (n_items <= 0 || (n_items & (n_items - 1))) return -1;
  int* b = (int*)malloc(sizeof(int) * (size_t)n_items);
  if (!b) return -2;
  memcpy(b, in, sizeof(int) * (size_t)n_items);
  int dp = 1;
  for (int s = n_items >> 1; s > 0; s >>= 1) {
    for (int lid = 0; lid < s; ++lid) {
      const int i = dp * (2 * lid + 1) - 1, j = dp * (2 * lid + 2) - 1;
      b[j] += b[i];
    }
    dp <<= 1;
  }
  b[n_items - 1] = 0;
  for (int s = 1; s < n_items; s <<= 1) {
    dp >>= 1;
    for (int lid = 0; lid < s; ++lid) {
      const int i = dp * (2 * lid + 1) - 1, j = dp * (2 * lid + 2) - 1;
      const int t = b[j];
      b[j] += b[i];
      b[i] = t;
    }
  }
  memcpy(out, b, sizeof(int) * (size_t)n_items);
  const int total = b[n_items - 1] + in[n_items - 1];
  free(b);
  return total;
}

/* ------------------------------------------------------------------------ */
/* CPU model of the discovery protocol and the barriers                       */
/* ------------------------------------------------------------------------ */

typedef struct {
  _Atomic int v;
  char pad[124];
} line_t;

#define CLOSED (1 << 30)
#define MASK (CLOSED - 1)

typedef struct model {
  int groups, resident, lock, bar, skip, barriers;
  line_t poll, prot_poll_open, prot_counter, kernel_counter, num_participating, closing_p;
  line_t lock_counter, now_serving, bar_count, bar_gen, cluster[64], rounds[3];
  line_t* flags;
  int* data; /* plain payload: 2 x groups ints, written before each barrier, read after */
  _Atomic long long progress;
  _Atomic int abort_flag;
  _Atomic long long violations;
  _Atomic int finished;
  sem_t slots;
  int* ids;
} model_t;

typedef struct {
  model_t* m;
  int gid;
} targ_t;

static double now_ms(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec * 1e3 + ts.tv_nsec / 1e6;
}

/* Memory orders follow the device code (include/gdb_discovery.hpp): polls
 * are relaxed, protocol and barrier read-modify-writes are acq_rel, the
 * publishing stores are release.  The device's "relaxed poll loop, then an
 * acquire fence" and "release fence, then a relaxed store" appear here as an
 * acquire poll and a release store on the same word: the same
 * synchronises-with edges, in a form ThreadSanitizer checks (it does not
 * model standalone fences).  g_weak plants a mistake for the tests: the
 * barrier's arrival and release operations become relaxed, and the plain
 * payload written before each barrier must then race with its readers. */
static int g_weak = 0;
void orc_dp_model_set_weak(int weak) { g_weak = weak; }
#define ORD_REL (g_weak ? memory_order_relaxed : memory_order_release)
#define ORD_ACQ (g_weak ? memory_order_relaxed : memory_order_acquire)
#define ORD_AR (g_weak ? memory_order_relaxed : memory_order_acq_rel)
#define LD(p, o) atomic_load_explicit((p), (o))
#define ST(p, v, o) atomic_store_explicit((p), (v), (o))

/* Spin helper: returns 0 if aborted. */
#define SPIN_WHILE(m, cond)                              \
  do {                                                   \
    while (cond) {                                       \
      if (atomic_load(&(m)->abort_flag)) return 0;       \
      sched_yield();                                     \
    }                                                    \
    atomic_fetch_add(&(m)->progress, 1);                 \
  } while (0)

/* DP/locks/locks.cl:11-34 with the device's orders (lock_acquire /
 * lock_release in gdb_discovery.hpp). */
static int m_lock(model_t* m) {
  if (m->lock == 2) {
    for (;;) {
      if (atomic_exchange_explicit(&m->lock_counter.v, 1, memory_order_acq_rel) == 0) break;
      if (atomic_load(&m->abort_flag)) return 0;
      sched_yield();
    }
    return 1;
  }
  const int ticket = atomic_fetch_add_explicit(&m->lock_counter.v, 1, memory_order_acq_rel);
  SPIN_WHILE(m, LD(&m->now_serving.v, memory_order_acquire) != ticket);
  return 1;
}
static void m_unlock(model_t* m) {
  if (m->lock == 2)
    ST(&m->lock_counter.v, 0, memory_order_release);
  else
    ST(&m->now_serving.v, LD(&m->now_serving.v, memory_order_relaxed) + 1, memory_order_release);
}

/* discovery_protocol_master, DP/api/discovery.cl:42-88, with a mutex.
 * Returns 1 on completion; sets *id / *P (id = -1: not participating). */
static int m_protocol_locked(model_t* m, int* id, int* P) {
  *id = -1;
  if (!m_lock(m)) return 0;
  if (LD(&m->prot_poll_open.v, memory_order_relaxed)) {
    *id = LD(&m->prot_counter.v, memory_order_relaxed);
    ST(&m->prot_counter.v, *id + 1, memory_order_relaxed);
  }
  m_unlock(m);
  if (!m_lock(m)) return 0;
  const int kc = LD(&m->kernel_counter.v, memory_order_relaxed) + 1;
  ST(&m->kernel_counter.v, kc, memory_order_relaxed);
  if (LD(&m->prot_poll_open.v, memory_order_relaxed)) {
    ST(&m->prot_poll_open.v, 0, memory_order_relaxed);
    ST(&m->num_participating.v, LD(&m->prot_counter.v, memory_order_relaxed), memory_order_relaxed);
  }
  *P = LD(&m->prot_counter.v, memory_order_relaxed);
  if (kc == m->groups) { /* reset_kernel_context, discovery.cl:31-37; the last store unlocks */
    ST(&m->prot_poll_open.v, 1, memory_order_relaxed);
    ST(&m->prot_counter.v, 0, memory_order_relaxed);
    ST(&m->kernel_counter.v, 0, memory_order_relaxed);
    ST(&m->now_serving.v, 0, memory_order_relaxed);
    ST(&m->lock_counter.v, 0, memory_order_release);
  } else {
    m_unlock(m);
  }
  return 1;
}

/* Lock-free variant (protocol_lockfree in gdb_discovery.hpp). */
static int m_protocol_lockfree(model_t* m, int* id, int* P) {
  const int old = atomic_fetch_add_explicit(&m->poll.v, 1, memory_order_acq_rel);
  const int part = (old & CLOSED) == 0;
  if (part) {
    int last = old + 1;
    double t_last = now_ms();
    for (;;) {
      const int cur = LD(&m->poll.v, memory_order_relaxed);
      if (cur & CLOSED) break;
      if ((cur & MASK) >= m->groups) break;
      const double t = now_ms();
      if (cur != last) {
        last = cur;
        t_last = t;
      } else if (t - t_last > 20.0) {
        break;
      }
      if (atomic_load(&m->abort_flag)) return 0;
      sched_yield();
    }
  }
  const int o2 = atomic_fetch_or_explicit(&m->poll.v, CLOSED, memory_order_acq_rel);
  int p = 0;
  if ((o2 & CLOSED) == 0) {
    p = o2 & MASK;
    ST(&m->num_participating.v, p, memory_order_relaxed);
    ST(&m->closing_p.v, p, memory_order_release);
  } else if (part) {
    SPIN_WHILE(m, LD(&m->closing_p.v, memory_order_acquire) == 0);
    p = LD(&m->closing_p.v, memory_order_acquire);
  }
  *id = part ? (old & MASK) : -1;
  *P = p;
  const int kc = atomic_fetch_add_explicit(&m->kernel_counter.v, 1, memory_order_acq_rel);
  if (kc == m->groups - 1) {
    ST(&m->poll.v, 0, memory_order_relaxed);
    ST(&m->closing_p.v, 0, memory_order_relaxed);
    ST(&m->kernel_counter.v, 0, memory_order_release);
  }
  return 1;
}

/* XF_barrier, DP/api/discovery.cl:129-190, one representative per group
 * (the master's per-thread loop over slaves is run by one thread here);
 * orders as xf_barrier in gdb_discovery.hpp. */
static int m_xf_barrier(model_t* m, int id, int P) {
  if (id == 0) {
    for (int peer = 1; peer < P; ++peer) SPIN_WHILE(m, LD(&m->flags[peer].v, ORD_ACQ) == 0);
    for (int peer = 1; peer < P; ++peer) ST(&m->flags[peer].v, 0, ORD_REL);
  } else {
    ST(&m->flags[id].v, 1, ORD_REL);
    SPIN_WHILE(m, LD(&m->flags[id].v, ORD_ACQ) == 1);
  }
  return 1;
}

/* counter_wait in gdb_discovery.hpp. */
static int m_counter_wait(model_t* m, _Atomic int* count, int members, int* gen_cache) {
  const int g = *gen_cache;
  const int old = atomic_fetch_add_explicit(count, 1, ORD_AR);
  if (old == members - 1) {
    ST(count, 0, memory_order_relaxed);
    ST(&m->bar_gen.v, g + 1, ORD_REL);
  } else {
    SPIN_WHILE(m, LD(&m->bar_gen.v, ORD_ACQ) == g);
  }
  *gen_cache = g + 1;
  return 1;
}

/* tree_barrier in gdb_discovery.hpp: clusters of consecutive ids. */
static int m_tree_barrier(model_t* m, int id, int P, int* gen_cache) {
  const int C = (P + 15) / 16;
  const int c = id / C;
  const int members = (P - c * C) < C ? (P - c * C) : C;
  const int nclus = (P + C - 1) / C;
  const int g = *gen_cache;
  const int old = atomic_fetch_add_explicit(&m->cluster[c].v, 1, ORD_AR);
  if (old == members - 1) {
    ST(&m->cluster[c].v, 0, memory_order_relaxed);
    const int o2 = atomic_fetch_add_explicit(&m->bar_count.v, 1, ORD_AR);
    if (o2 == nclus - 1) {
      ST(&m->bar_count.v, 0, memory_order_relaxed);
      ST(&m->bar_gen.v, g + 1, ORD_REL);
      *gen_cache = g + 1;
      return 1;
    }
  }
  SPIN_WHILE(m, LD(&m->bar_gen.v, ORD_ACQ) == g);
  *gen_cache = g + 1;
  return 1;
}

static void* m_group(void* arg) {
  targ_t* a = (targ_t*)arg;
  model_t* m = a->m;
  /* Wait for a residency slot (dispatch). */
  for (;;) {
    struct timespec ts;
    clock_gettime(CLOCK_REALTIME, &ts);
    ts.tv_nsec += 5 * 1000 * 1000;
    if (ts.tv_nsec >= 1000000000) {
      ts.tv_sec += 1;
      ts.tv_nsec -= 1000000000;
    }
    if (sem_timedwait(&m->slots, &ts) == 0) break;
    if (atomic_load(&m->abort_flag)) return NULL;
  }
  atomic_fetch_add(&m->progress, 1);
  int id = -1, P = 0, ok = 1;
  if (m->skip) { /* DISCOVERY_PROTOCOL skip path, discovery.cl:215-220 */
    id = a->gid;
    P = m->groups;
    atomic_store(&m->num_participating.v, P);
  } else if (m->lock == 0) {
    ok = m_protocol_lockfree(m, &id, &P);
  } else {
    ok = m_protocol_locked(m, &id, &P);
  }
  m->ids[a->gid] = ok ? id : -1;
  if (ok && id >= 0) {
    int gen = atomic_load_explicit(&m->bar_gen.v, memory_order_acquire);
    const int nb = (id + 1) % P;
    for (int k = 0; k < m->barriers && ok; ++k) {
      /* Before barrier k: a plain payload store and a relaxed count (as
       * gdb_k_barrier_check); after it both must be visible. */
      int* slot = m->data + (size_t)(k & 1) * (size_t)m->groups;
      slot[id] = k * 7919 + id;
      atomic_fetch_add_explicit(&m->rounds[k % 3].v, 1, memory_order_relaxed);
      if (P > 1) {
        if (m->bar == 0)
          ok = m_xf_barrier(m, id, P);
        else if (m->bar == 1)
          ok = m_counter_wait(m, &m->bar_count.v, P, &gen);
        else
          ok = m_tree_barrier(m, id, P, &gen);
      }
      if (ok && (atomic_load_explicit(&m->rounds[k % 3].v, memory_order_relaxed) != P * (k / 3 + 1) ||
                 slot[nb] != k * 7919 + nb))
        atomic_fetch_add(&m->violations, 1);
    }
  }
  atomic_fetch_add(&m->finished, 1);
  atomic_fetch_add(&m->progress, 1);
  sem_post(&m->slots);
  return NULL;
}

int orc_dp_model(int groups, int resident, int lock, int bar, int skip, int barriers,
                 int watchdog_ms, int* participating, int* ids, long long* violations) {
  if (groups < 1 || resident < 1 || groups > 4096) return -1;
  model_t* m = (model_t*)calloc(1, sizeof(model_t));
  if (!m) return -2;
  m->groups = groups;
  m->resident = resident;
  m->lock = lock;
  m->bar = bar;
  m->skip = skip;
  m->barriers = barriers;
  m->flags = (line_t*)calloc((size_t)groups, sizeof(line_t));
  m->data = (int*)calloc((size_t)groups * 2, sizeof(int));
  m->ids = ids;
  atomic_store(&m->prot_poll_open.v, 1);
  sem_init(&m->slots, 0, (unsigned)resident);
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)groups);
  targ_t* args = (targ_t*)malloc(sizeof(targ_t) * (size_t)groups);
  int deadlock = 0;
  for (int g = 0; g < groups; ++g) {
    args[g].m = m;
    args[g].gid = g;
    ids[g] = -1;
    pthread_attr_t attr;
    pthread_attr_init(&attr);
    pthread_attr_setstacksize(&attr, 256 * 1024);
    pthread_create(&th[g], &attr, m_group, &args[g]);
    pthread_attr_destroy(&attr);
  }
  long long last = -1;
  double t_last = now_ms();
  while (atomic_load(&m->finished) < groups) {
    struct timespec ts = {0, 2 * 1000 * 1000};
    nanosleep(&ts, NULL);
    const long long p = atomic_load(&m->progress);
    if (p != last) {
      last = p;
      t_last = now_ms();
    } else if (now_ms() - t_last > watchdog_ms) {
      deadlock = 1;
      atomic_store(&m->abort_flag, 1);
      break;
    }
  }
  for (int g = 0; g < groups; ++g) pthread_join(th[g], NULL);
  *participating = atomic_load(&m->num_participating.v);
  *violations = atomic_load(&m->violations);
  sem_destroy(&m->slots);
  free(th);
  free(args);
  free(m->flags);
  free(m->data);
  free(m);
  return deadlock;
}

/* ------------------------------------------------------------------------ */
/* Schedule selector MLP (no reference counterpart)                          */
/* ------------------------------------------------------------------------ */
enum { MIN = 16, MHID = 64, MOUT = 4, MSTRIDE = 20 };
enum {
  MW1 = 0,
  MB1 = MW1 + MIN * MHID,
  MW2 = MB1 + MHID,
  MB2 = MW2 + MHID * MHID,
  MW3 = MB2 + MHID,
  MB3 = MW3 + MHID * MOUT,
  MPARAMS = MB3 + MOUT
};

static void mlp_fwd(const double* w, const double* x, double* a1, double* h1, double* a2, double* h2,
                    double* y) {
  for (int j = 0; j < MHID; ++j) {
    double a = w[MB1 + j];
    for (int i = 0; i < MIN; ++i) a += w[MW1 + i * MHID + j] * x[i];
    a1[j] = a;
    h1[j] = a > 0 ? a : 0;
  }
  for (int j = 0; j < MHID; ++j) {
    double a = w[MB2 + j];
    for (int i = 0; i < MHID; ++i) a += w[MW2 + i * MHID + j] * h1[i];
    a2[j] = a;
    h2[j] = a > 0 ? a : 0;
  }
  for (int o = 0; o < MOUT; ++o) {
    double a = w[MB3 + o];
    for (int i = 0; i < MHID; ++i) a += w[MW3 + i * MOUT + o] * h2[i];
    y[o] = a;
  }
}

void orc_mlp_forward(const double* w, const double* x, double* y) {
  double a1[MHID], h1[MHID], a2[MHID], h2[MHID];
  mlp_fwd(w, x, a1, h1, a2, h2, y);
}

int orc_mlp_grad(const double* w, const float* records, int count, double* grad, double* loss) {
  double a1[MHID], h1[MHID], a2[MHID], h2[MHID], y[MOUT], x[MIN], dh2[MHID];
  memset(grad, 0, sizeof(double) * MPARAMS);
  double lsum = 0;
  int used = 0;
  for (int r = 0; r < count; ++r) {
    const float* R = records + (size_t)r * MSTRIDE;
    const double cost = R[MIN + 1];
    const int arm = (int)R[MIN];
    if (!(cost > 0) || arm < 0 || arm >= MOUT) continue;
    for (int i = 0; i < MIN; ++i) x[i] = R[i];
    mlp_fwd(w, x, a1, h1, a2, h2, y);
    const double err = y[arm] - log2(cost);
    lsum += err * err;
    ++used;
    const double dy = 2 * err;
    grad[MB3 + arm] += dy;
    for (int i = 0; i < MHID; ++i) {
      grad[MW3 + i * MOUT + arm] += h2[i] * dy;
      dh2[i] = a2[i] > 0 ? w[MW3 + i * MOUT + arm] * dy : 0;
    }
    for (int j = 0; j < MHID; ++j) {
      grad[MB2 + j] += dh2[j];
      for (int i = 0; i < MHID; ++i) grad[MW2 + i * MHID + j] += h1[i] * dh2[j];
    }
    for (int i = 0; i < MHID; ++i) {
      double d = 0;
      for (int j = 0; j < MHID; ++j) d += w[MW2 + i * MHID + j] * dh2[j];
      const double dh1 = a1[i] > 0 ? d : 0;
      grad[MB1 + i] += dh1;
      for (int k = 0; k < MIN; ++k) grad[MW1 + k * MHID + i] += x[k] * dh1;
    }
  }
  if (used > 0)
    for (int p = 0; p < MPARAMS; ++p) grad[p] /= used;
  *loss = used > 0 ? lsum / used : 0;
  return used;
}

/* ------------------------------------------------------------------------ */
/* Minimum spanning forest                                                   */
/* ------------------------------------------------------------------------ */
typedef struct {
  unsigned w;
  int a, b;
} orc_edge;

static int orc_edge_cmp(const void* x, const void* y) {
  const orc_edge* p = (const orc_edge*)x;
  const orc_edge* q = (const orc_edge*)y;
  if (p->w != q->w) return p->w < q->w ? -1 : 1;
  if (p->a != q->a) return p->a < q->a ? -1 : 1;
  return (p->b > q->b) - (p->b < q->b);
}

static int uf_find(int* par, int x) {
  while (par[x] != x) {
    par[x] = par[par[x]];
    x = par[x];
  }
  return x;
}

/* lonestar apps/mst/kernel/mst_common.cl (Boruvka); any algorithm gives the
 * same forest weight.  Each undirected edge is taken once (a < b). */
int orc_mst_kruskal(int n, const int* row, const int* col, const unsigned* w, unsigned long long* weight,
                    int* edges, int* components) {
  long long me = 0;
  for (int v = 0; v < n; ++v)
    for (int k = row[v]; k < row[v + 1]; ++k)
      if (v < col[k]) ++me;
  orc_edge* e = (orc_edge*)malloc(sizeof(orc_edge) * (size_t)(me > 0 ? me : 1));
  int* par = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  long long c = 0;
  for (int v = 0; v < n; ++v)
    for (int k = row[v]; k < row[v + 1]; ++k)
      if (v < col[k]) {
        e[c].w = w[k];
        e[c].a = v;
        e[c].b = col[k];
        ++c;
      }
  qsort(e, (size_t)me, sizeof(orc_edge), orc_edge_cmp);
  for (int v = 0; v < n; ++v) par[v] = v;
  unsigned long long total = 0;
  int taken = 0;
  for (long long i = 0; i < me; ++i) {
    const int ra = uf_find(par, e[i].a), rb = uf_find(par, e[i].b);
    if (ra == rb) continue;
    par[ra] = rb;
    total += e[i].w;
    ++taken;
  }
  *weight = total;
  *edges = taken;
  *components = n - taken;
  free(e);
  free(par);
  return 0;
}

/* Prim with a binary heap of (weight, vertex), restarted from every vertex
 * not yet in a tree. */
int orc_mst_prim_tb(int n, const int* row, const int* col, const unsigned* w, unsigned long long* weight,
                    int* edges, int* components) {
  const long long m = n > 0 ? row[n] : 0;
  unsigned long long* heap = (unsigned long long*)malloc(sizeof(unsigned long long) * (size_t)(m + n + 1));
  char* in = (char*)calloc((size_t)(n > 0 ? n : 1), 1);
  unsigned long long total = 0;
  int taken = 0, comps = 0;
  for (int s = 0; s < n; ++s) {
    if (in[s]) continue;
    ++comps;
    long long hs = 0;
    heap[hs++] = (unsigned long long)s;  /* weight 0, vertex s */
    while (hs > 0) {
      /* pop min */
      const unsigned long long top = heap[0];
      heap[0] = heap[--hs];
      for (long long i = 0;;) {
        long long l = 2 * i + 1, r = l + 1, sm = i;
        if (l < hs && heap[l] < heap[sm]) sm = l;
        if (r < hs && heap[r] < heap[sm]) sm = r;
        if (sm == i) break;
        const unsigned long long t = heap[i];
        heap[i] = heap[sm];
        heap[sm] = t;
        i = sm;
      }
      const int v = (int)(top & 0xffffffffu);
      if (in[v]) continue;
      in[v] = 1;
      if (v != s) {
        total += top >> 32;
        ++taken;
      }
      for (int k = row[v]; k < row[v + 1]; ++k) {
        const int u = col[k];
        if (u < 0 || u >= n || in[u]) continue;
        long long i = hs++;
        heap[i] = ((unsigned long long)w[k] << 32) | (unsigned)u;
        while (i > 0 && heap[(i - 1) / 2] > heap[i]) {
          const unsigned long long t = heap[i];
          heap[i] = heap[(i - 1) / 2];
          heap[(i - 1) / 2] = t;
          i = (i - 1) / 2;
        }
      }
    }
  }
  *weight = total;
  *edges = taken;
  *components = comps;
  free(heap);
  free(in);
  return 0;
}
