// Minimum spanning forest (Boruvka) on the discovery protocol.
//
// Reference: lonestar mst (code/experiments/lonestar_gpu_opencl/apps/mst):
// per round every component finds its lightest cross-component edge
// (dfindelemin, dfindelemin2, verify_min_elem in kernel/mst_common.cl) and
// the chosen edges unify components (dfindcompmintwo, mst_port_kernel.cl,
// with discovery barriers inside a uniform-trip loop); the host repeats
// rounds while the component count falls (mst_port.cpp:203-313) and prints
// weight, components and edges.
//
// Here one persistent kernel runs every round.  Edges are totally ordered by
// (weight, smaller endpoint, larger endpoint), so each component's choice is
// unique and the chosen edges form a forest except for pairs of components
// that chose the same edge; the smaller id of such a pair stays a root and
// the edge is counted once.  Per round:
//   A  reset the lightest-weight and best-key words of every root, parent = self
//   B  every vertex: lightest cross edge -> atomicMin(minw[comp])
//   C  every vertex: among its edges of that weight, the smallest key ->
//      atomicMin(best[comp])
//   D  every root with a choice points at the other side's component
//   E  mutual pairs keep one root; the others add the edge's weight
//   F  every vertex follows parents to its new root (path halving)
// The weight of a minimum spanning forest is unique, so it and the
// component and edge counts are compared exactly.
#include "graph_common.hpp"
#include "runtime.hpp"

namespace {

constexpr unsigned kNoW = 0xffffffffu;
constexpr unsigned long long kNoKey = ~0ull;

__device__ __forceinline__ unsigned long long edge_key(int a, int b) {
  const unsigned lo = (unsigned)(a < b ? a : b), hi = (unsigned)(a < b ? b : a);
  return ((unsigned long long)lo << 32) | hi;
}

// Root of component c: follows parent pointers, halving the path (a pointer
// is only ever replaced by one of its ancestors, so racing walkers agree).
__device__ __forceinline__ int find_root(int* parent, int c, int n) {
  for (int guard = 0; guard < n; ++guard) {
    const int p = parent[c];
    if (p == c || (unsigned)p >= (unsigned)n) return c;
    const int gp = parent[p];
    if ((unsigned)gp < (unsigned)n && gp != p) parent[c] = gp;
    c = p;
  }
  return c;
}

template <int WGS>
__global__ __launch_bounds__(WGS) void k_mst(gdb_kernel_ctx_t* ctx, const int* __restrict__ row,
                                             const int* __restrict__ col, const unsigned* __restrict__ w,
                                             int n, int* comp, int* parent, unsigned* minw,
                                             unsigned long long* best, int* flags,
                                             unsigned long long* total, int* out) {
  DISCOVERY_PROTOCOL(ctx);
  const int tid = p_get_global_id(ctx, &local_ctx);
  const int stride = p_get_global_size(ctx, &local_ctx);
  int rounds = 0, nbar = 0;
  unsigned long long wsum = 0;
  int added = 0;
  for (int v = tid; v < n; v += stride) comp[v] = v;
  for (;;) {
    // A
    for (int c = tid; c < n; c += stride) {
      if (comp[c] == c) {
        minw[c] = kNoW;
        best[c] = kNoKey;
        parent[c] = c;
      }
    }
    if (tid == 0) gdb::st_rlx(&flags[(rounds + 1) % 3], 0);
    if (!discovery_barrier(ctx, &local_ctx)) return;
    ++nbar;
    // B
    for (int v = tid; v < n; v += stride) {
      const int cv = comp[v];
      unsigned m = kNoW;
      for (int k = row[v]; k < row[v + 1]; ++k) {
        const int u = col[k];
        if ((unsigned)u >= (unsigned)n) continue;
        const unsigned wk = w[k];
        if (wk < m && comp[u] != cv) m = wk;
      }
      if (m != kNoW && (unsigned)cv < (unsigned)n) atomicMin(&minw[cv], m);
    }
    if (!discovery_barrier(ctx, &local_ctx)) return;
    ++nbar;
    // C
    for (int v = tid; v < n; v += stride) {
      const int cv = comp[v];
      if ((unsigned)cv >= (unsigned)n) continue;
      const unsigned m = minw[cv];
      if (m == kNoW) continue;
      unsigned long long key = kNoKey;
      for (int k = row[v]; k < row[v + 1]; ++k) {
        const int u = col[k];
        if ((unsigned)u >= (unsigned)n || w[k] != m || comp[u] == cv) continue;
        const unsigned long long kk = edge_key(v, u);
        if (kk < key) key = kk;
      }
      if (key != kNoKey) atomicMin(&best[cv], key);
    }
    if (!discovery_barrier(ctx, &local_ctx)) return;
    ++nbar;
    // D
    for (int c = tid; c < n; c += stride) {
      if (comp[c] != c || best[c] == kNoKey) continue;
      const int a = (int)(best[c] >> 32), b = (int)(best[c] & 0xffffffffu);
      if ((unsigned)a >= (unsigned)n || (unsigned)b >= (unsigned)n) continue;
      const int ca = comp[a], cb = comp[b];
      parent[c] = ca == c ? cb : ca;
    }
    if (!discovery_barrier(ctx, &local_ctx)) return;
    ++nbar;
    // E
    bool progress = false;
    for (int c = tid; c < n; c += stride) {
      if (comp[c] != c) continue;
      const int p = parent[c];
      if (p == c || (unsigned)p >= (unsigned)n) continue;
      if (parent[p] == c && c < p) {
        parent[c] = c;  // the pair's smaller id stays the root; p counts the edge
      } else {
        wsum += minw[c];
        ++added;
        progress = true;
      }
    }
    if (gdb::wave_any(progress) && gdb::lane_id() == 0) gdb::st_rlx(&flags[rounds % 3], 1);
    if (!discovery_barrier(ctx, &local_ctx)) return;
    ++nbar;
    // F
    for (int v = tid; v < n; v += stride) {
      const int c = comp[v];
      if ((unsigned)c < (unsigned)n) comp[v] = find_root(parent, c, n);
    }
    if (!discovery_barrier(ctx, &local_ctx)) return;
    ++nbar;
    ++rounds;
    if (gdb::ld_rlx(&flags[(rounds - 1) % 3]) == 0) break;
    if (rounds > 64) {  // each round at least halves the components with a cross edge
      gdb::set_error(ctx, GDB_ERR_QUEUE_OVERFLOW);
      break;
    }
  }
  // Totals: one atomic per wave.
  unsigned long long ws = wsum;
  int ad = added, roots = 0;
  for (int v = tid; v < n; v += stride) roots += comp[v] == v;
  for (int off = 32; off > 0; off >>= 1) {
    ws += __shfl_xor(ws, off, 64);
    ad += __shfl_xor(ad, off, 64);
    roots += __shfl_xor(roots, off, 64);
  }
  if (gdb::lane_id() == 0) {
    if (ws) atomicAdd(total, ws);
    if (ad) atomicAdd(&out[2], ad);
    if (roots) atomicAdd(&out[3], roots);
  }
  if (tid == 0) {
    out[0] = rounds;
    out[1] = nbar;
  }
}

}  // namespace

using namespace gdbrt;

extern "C" int gdb_mst(gdb_ctx* ctx, int n, int m, const int* row, const int* col, const unsigned* w,
                       unsigned long long* weight, int* edges, int* components, const gdb_opts* opts,
                       gdb_stats* stats) {
  if (!ctx || n < 0 || m < 0 || (n > 0 && !row) || (m > 0 && (!col || !w))) return GDB_E_INVALID;
  gdb_opts o{};
  if (opts) o = *opts;
  gdbrt::apply_tuning("mst", n, &o);
  const int wgs = default_wgs(&o);
  gdb_stats st{};
  if (n == 0) {
    if (weight) *weight = 0;
    if (edges) *edges = 0;
    if (components) *components = 0;
    if (stats) *stats = st;
    return GDB_SUCCESS;
  }
  const size_t bytes = carve_bytes<int>(n) * 2 + carve_bytes<unsigned>(n) +
                       carve_bytes<unsigned long long>(n) + carve_bytes<int>(32) +
                       carve_bytes<unsigned long long>(1);
  int rc = ensure_workspace(ctx, bytes);
  if (rc) return rc;
  Carve cv(ctx->ws);
  int* comp = cv.take<int>(n);
  int* parent = cv.take<int>(n);
  unsigned* minw = cv.take<unsigned>(n);
  unsigned long long* best = cv.take<unsigned long long>(n);
  int* flags = cv.take<int>(32);
  unsigned long long* total = cv.take<unsigned long long>(1);
  rc = fill_i32(ctx, flags, 0, 32);
  if (!rc) rc = fill_i32(ctx, (int*)total, 0, 2);
  if (rc) return rc;
  int* out = flags + 8;
  rc = prepare_launch(ctx);
  if (rc) return rc;
  rc = gdb_dispatch_wgs(wgs, [&](auto W) -> int {
    constexpr int WG = decltype(W)::value;
    int groups = 0;
    int r2 = launch_groups(ctx, (const void*)k_mst<WG>, &o, n, &groups);
    if (r2) return r2;
    r2 = timer_start(ctx);
    if (r2) return r2;
    hipLaunchKernelGGL(k_mst<WG>, dim3(groups), dim3(WG), 0, ctx->stream, ctx->d, row, col, w, n, comp, parent,
                       minw, best, flags, total, out);
    GDB_HIP(hipGetLastError());
    return timer_stop(ctx, &st.kernel_ms);
  });
  if (rc) return rc;
  int hout[4] = {0, 0, 0, 0};
  unsigned long long hw = 0;
  GDB_HIP(hipMemcpy(hout, out, sizeof(hout), hipMemcpyDeviceToHost));
  GDB_HIP(hipMemcpy(&hw, total, sizeof(hw), hipMemcpyDeviceToHost));
  st.rounds = hout[0];
  st.barriers = hout[1];
  st.launches = 1;
  st.work_edges = (long long)m * hout[0];
  st.work_items = (long long)n * hout[0];
  rc = read_participating(ctx, &st.participating);
  if (rc) return rc;
  rc = check_device_error(ctx);
  if (weight) *weight = hw;
  if (edges) *edges = hout[2];
  if (components) *components = hout[3];
  if (stats) *stats = st;
  return rc;
}
