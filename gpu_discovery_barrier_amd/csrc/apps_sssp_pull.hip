// Pannotia single-source shortest paths (pull Bellman-Ford) on the
// discovery protocol.
//
// Rule (code/experiments/pannotia/graph_app/sssp/kernel/sssp_kernel.cl):
// vector_init sets x = y = (0 at the source, BIG_NUM elsewhere) (lines
// 111-132); then for k = 1 .. n-1: x = y; y[v] = min(x[v], min over in-edges
// (w + x[u])); stop if no y[v] != x[v] (mega_kernel, lines 170-244).  The
// result is x (sssp_gb.cpp:308), which lags y by one round when the n-1
// round cap ends the loop.
//
// Schedules:
//   plain persistent - mega_kernel: assign / relax / diff with three
//                      barriers per round.
//   tuned persistent - one barrier per round: round k writes
//                      b_k = relax(b_{k-1}) into the buffer of parity k and
//                      raises a per-round change flag.  The loop ends at the
//                      first round K with no change or at K = n-1; the
//                      reference's x is then b_{K-1}.
//   relaunch         - vector_assign / spmv / vector_diff per round with a
//                      host loop (sssp.cpp:290-344).
#include "graph_common.hpp"
#include "runtime.hpp"

namespace {

constexpr int kBigNum = 99999999;  // sssp_kernel.cl:68
constexpr int kSerialLimit = 16;

__device__ __forceinline__ int relax_min(const int* __restrict__ row_t, const int* __restrict__ col_t,
                                         const int* __restrict__ w_t, const int* x, int v) {
  int mn = x[v];
  for (int j = row_t[v]; j < row_t[v + 1]; ++j) {
    const int cand = w_t[j] + x[col_t[j]];
    if (cand < mn) mn = cand;
  }
  return mn;
}

template <int WGS>
__global__ __launch_bounds__(WGS) void k_sssp_pull_plain(gdb_kernel_ctx_t* ctx, int n,
                                                         const int* __restrict__ row_t,
                                                         const int* __restrict__ col_t,
                                                         const int* __restrict__ w_t, int* x, int* y,
                                                         int* stop, int* out) {
  DISCOVERY_PROTOCOL(ctx);
  const int tid = p_get_global_id(ctx, &local_ctx);
  const int stride = p_get_global_size(ctx, &local_ctx);
  int k = 1, nbar = 0;
  for (; k < n; ++k) {
    for (int i = tid; i < n; i += stride) x[i] = y[i];  // vector_assign
    if (!discovery_barrier(ctx, &local_ctx)) return;
    ++nbar;
    if (tid == 0) gdb::st_rlx(stop, 0);
    for (int it = tid; it < n; it += stride) y[it] = relax_min(row_t, col_t, w_t, x, it);
    if (!discovery_barrier(ctx, &local_ctx)) return;
    ++nbar;
    bool ch = false;  // vector_diff
    for (int i = tid; i < n; i += stride)
      if (y[i] != x[i]) ch = true;
    if (ch) gdb::st_rlx(stop, 1);
    if (!discovery_barrier(ctx, &local_ctx)) return;
    ++nbar;
    if (gdb::ld_rlx(stop) == 0) break;
  }
  if (tid == 0) {
    out[0] = k < n ? k : n - 1;  // rounds executed
    out[1] = nbar;
  }
}

template <int WGS>
__global__ __launch_bounds__(WGS) void k_sssp_pull_tuned(gdb_kernel_ctx_t* ctx, int n,
                                                         const int* __restrict__ row_t,
                                                         const int* __restrict__ col_t,
                                                         const int* __restrict__ w_t, int* buf0,
                                                         int* buf1, int* changed, int* out) {
  DISCOVERY_PROTOCOL(ctx);
  const int tid = p_get_global_id(ctx, &local_ctx);
  const int stride = p_get_global_size(ctx, &local_ctx);
  int K = 0, nbar = 0;
  for (int k = 1; k < n; ++k) {
    K = k;
    const int* prev = (k & 1) ? buf0 : buf1;  // b_{k-1}
    int* cur = (k & 1) ? buf1 : buf0;         // b_k
    if (tid == 0) gdb::st_rlx(&changed[(k + 1) % 3], 0);
    bool ch = false;
    GDB_FOR_UNIFORM(v, valid, tid, stride, n) {
      int b = 0, e = 0;
      if (valid) {
        b = row_t[v];
        e = row_t[v + 1];
      }
      auto edge = [&](int j) -> int { return w_t[j] + prev[col_t[j]]; };
      auto red = [](int a, int c) { return a < c ? a : c; };
      const int mn = gdb::wave_coop_reduce(valid != 0, b, e, 0x7fffffff, edge, red, kSerialLimit);
      if (valid) {
        const int old = prev[v];
        const int nvl = mn < old ? mn : old;
        cur[v] = nvl;
        if (nvl != old) ch = true;
      }
    }
    if (gdb::wave_any(ch) && gdb::lane_id() == 0) gdb::st_rlx(&changed[k % 3], 1);
    if (!discovery_barrier(ctx, &local_ctx)) return;
    ++nbar;
    if (gdb::ld_rlx(&changed[k % 3]) == 0) break;
  }
  // The reference outputs b_{K-1}; buf0 is the output array, so copy when
  // b_{K-1} sits in buf1 (K-1 odd).
  if (K >= 1 && ((K - 1) & 1))
    for (int i = tid; i < n; i += stride) buf0[i] = buf1[i];
  if (tid == 0) {
    out[0] = K;
    out[1] = nbar;
  }
}

__global__ void k_vector_init(int* x, int* y, int src, int n, int* flags) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {  // vector_init, sssp_kernel.cl:111-132
    x[i] = i == src ? 0 : kBigNum;
    y[i] = x[i];
  }
  if (i < 16) flags[i] = 0;
}

// ---- relaunch: vector_assign (140-146), spmv (79-102), vector_diff (155-165).
// flags = {stop[0..2], rounds, done}: round r's diff raises stop[r % 3];
// round r's assign sees round r-1's stop and, when it is 0, raises the
// sticky done flag that turns every later kernel into a no-op.
__device__ __forceinline__ bool sssp_round_done(int* flags, int r) {
  return flags[4] != 0 || (r > 1 && flags[(r - 1) % 3] == 0);
}
__global__ void k_assign(int* x, const int* y, int n, int* flags, int r) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (sssp_round_done(flags, r)) {
    if (i == 0) flags[4] = 1;
    return;
  }
  if (i < n) x[i] = y[i];
  if (i == 0) flags[r % 3] = 0;
}
__global__ void k_spmv(int n, const int* __restrict__ row_t, const int* __restrict__ col_t,
                       const int* __restrict__ w_t, const int* x, int* y, int* flags, int r) {
  if (flags[4]) return;
  const int it = blockIdx.x * blockDim.x + threadIdx.x;
  if (it < n) y[it] = relax_min(row_t, col_t, w_t, x, it);
}
__global__ void k_diff(const int* x, const int* y, int n, int* flags, int r) {
  if (flags[4]) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && y[i] != x[i]) flags[r % 3] = 1;
  if (i == 0) atomicAdd(&flags[3], 1);
}

}  // namespace

using namespace gdbrt;

extern "C" int gdb_sssp_pull(gdb_ctx* ctx, int n, int m, const int* row_t, const int* col_t,
                             const int* w_t, int source, int* dist, const gdb_opts* opts,
                             gdb_stats* stats) {
  if (!ctx || n < 0 || m < 0 || (n > 0 && (!row_t || !dist)) || (m > 0 && (!col_t || !w_t)) ||
      (n > 0 && (source < 0 || source >= n)))
    return GDB_E_INVALID;
  gdb_opts o{};
  if (opts) o = *opts;
  if (o.variant == GDB_VARIANT_TUNED && o.schedule == GDB_SCHED_PERSISTENT) gdbrt::apply_tuning("sssp_pull", n, &o);
  const int wgs = default_wgs(&o);
  gdb_stats st{};
  if (n == 0) {
    if (stats) *stats = st;
    return GDB_SUCCESS;
  }
  const size_t bytes = carve_bytes<int>(n) + carve_bytes<int>(32);
  int rc = ensure_workspace(ctx, bytes);
  if (rc) return rc;
  Carve cv(ctx->ws);
  int* y = cv.take<int>(n);
  int* flags = cv.take<int>(32);
  int* x = dist;
  hipLaunchKernelGGL(k_vector_init, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, x, y, source, n,
                     flags);
  GDB_HIP(hipGetLastError());
  rc = prepare_launch(ctx);
  if (rc) return rc;

  if (o.schedule == GDB_SCHED_RELAUNCH) {
    const int grid = (n + wgs - 1) / wgs;
    const int check = o.relaunch_check > 1 ? o.relaunch_check : 1;
    rc = timer_start(ctx);
    if (rc) return rc;
    int r = 0;
    while (r < n - 1) {
      for (int k = 0; k < check && r < n - 1; ++k) {
        ++r;
        hipLaunchKernelGGL(k_assign, dim3(grid), dim3(wgs), 0, ctx->stream, x, y, n, flags, r);
        hipLaunchKernelGGL(k_spmv, dim3(grid), dim3(wgs), 0, ctx->stream, n, row_t, col_t, w_t, x, y,
                           flags, r);
        hipLaunchKernelGGL(k_diff, dim3(grid), dim3(wgs), 0, ctx->stream, x, y, n, flags, r);
        st.launches += 3;
      }
      GDB_HIP(hipGetLastError());
      int hf[5];
      GDB_HIP(hipMemcpyAsync(ctx->h_flag, flags, 4 * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
      GDB_HIP(hipStreamSynchronize(ctx->stream));
      for (int q = 0; q < 4; ++q) hf[q] = ctx->h_flag[q];
      GDB_HIP(hipMemcpy(&hf[4], flags + 4, sizeof(int), hipMemcpyDeviceToHost));
      if (hf[4] || hf[r % 3] == 0) break;
    }
    rc = timer_stop(ctx, &st.kernel_ms);
    if (rc) return rc;
    st.rounds = ctx->h_flag[3];
    st.participating = grid;
  } else {
    const bool tuned = o.variant == GDB_VARIANT_TUNED;
    int* out = flags + 8;
    rc = gdb_dispatch_wgs(wgs, [&](auto W) -> int {
      constexpr int WG = decltype(W)::value;
      const void* kern = tuned ? (const void*)k_sssp_pull_tuned<WG> : (const void*)k_sssp_pull_plain<WG>;
      int groups = 0;
      int r2 = launch_groups(ctx, kern, &o, n, &groups);
      if (r2) return r2;
      r2 = timer_start(ctx);
      if (r2) return r2;
      if (tuned)
        hipLaunchKernelGGL(k_sssp_pull_tuned<WG>, dim3(groups), dim3(WG), 0, ctx->stream, ctx->d, n,
                           row_t, col_t, w_t, x, y, flags, out);
      else
        hipLaunchKernelGGL(k_sssp_pull_plain<WG>, dim3(groups), dim3(WG), 0, ctx->stream, ctx->d, n,
                           row_t, col_t, w_t, x, y, flags, out);
      GDB_HIP(hipGetLastError());
      return timer_stop(ctx, &st.kernel_ms);
    });
    if (rc) return rc;
    st.launches = 1;
    int hout[2] = {0, 0};
    GDB_HIP(hipMemcpy(hout, out, sizeof(hout), hipMemcpyDeviceToHost));
    st.rounds = hout[0];
    st.barriers = hout[1];
    rc = read_participating(ctx, &st.participating);
    if (rc) return rc;
  }
  rc = check_device_error(ctx);
  st.work_edges = (long long)m * (st.rounds > 0 ? st.rounds : 1);
  st.work_items = (long long)n * (st.rounds > 0 ? st.rounds : 1);
  if (stats) *stats = st;
  return rc;
}
