// Lonestar data-driven worklist BFS and SSSP on the discovery protocol.
//
// BFS (code/experiments/lonestar_gpu_opencl/apps/bfs/): dist = MYINFINITY
// (1e9) everywhere, dist[source] = 0 (bfs_port.cpp:112-121); iteration 0
// pushes the source; iteration i pops every worklist item and, for each
// out-edge, sets an unvisited destination to dist = i and pushes it
// (processedge2, bfs_common.cl:73-90); the loop runs while the input
// worklist is non-empty (bfs_port_kernel.cl:173-189).
// SSSP (apps/sssp/): processedge2 relaxes alt = dist[src] + w with
// atomic_min and pushes the destination when alt improved it
// (sssp_common.cl:58-81).  Both converge to exact levels / distances.
//
// Queues hold each vertex at most once per iteration: BFS claims a vertex
// with a compare-and-swap on dist, SSSP stamps it with the iteration number;
// a queue therefore never exceeds n entries (the reference sizes them by the
// edge count and does not bounds-check pushes, worklist_cl.h:60-62).
//
// Schedules:
//   plain persistent - drelax2's loop: process, barrier, swap and reset the
//                      output index, barrier (two barriers per iteration);
//                      a thread walks its item's edges.
//   tuned persistent - one barrier per iteration (queue tails rotate over
//                      three slots so the reset of the next-next tail needs
//                      no barrier); edges of long rows are walked by a whole
//                      wave; pushes are compacted with ballot + one atomic per
//                      wave.
//   relaunch         - one launch per iteration with a host loop reading the
//                      output size (the non-persistent schedule).
#include <mutex>
#include "graph_common.hpp"
#include "runtime.hpp"

// Scratch of gdb_sssp_worklist_noalloc: the stamp array and both queues
// (3 x n ints) plus flags and work counters, for n up to 2^22 vertices.
// Each device that loads the library holds its own copy until the process
// ends.  Host code never passes `scratch` itself to a kernel: its device
// address comes from hipGetSymbolAddress (resolve_noalloc_scratch).
#define GDB_NOALLOC_SCRATCH_BYTES ((size_t)3 * (4u << 22) + 4096)
static __device__ __align__(256) unsigned char scratch[GDB_NOALLOC_SCRATCH_BYTES];

namespace gdbrt {

int resolve_noalloc_scratch(gdb_ctx* ctx) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static void* addr[kMaxDevices] = {};
  if (ctx->device < 0 || ctx->device >= kMaxDevices) return GDB_E_INVALID;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[ctx->device]) {
    int cur = -1;
    GDB_HIP(hipGetDevice(&cur));
    if (cur != ctx->device) GDB_HIP(hipSetDevice(ctx->device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != ctx->device) GDB_HIP(hipSetDevice(cur));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
    addr[ctx->device] = p;
  }
  ctx->noalloc_scratch = addr[ctx->device];
  return GDB_SUCCESS;
}

}  // namespace gdbrt

namespace {

constexpr unsigned kInf = 1000000000u;  // bfs_common.cl:35
constexpr int kSerialLimit = 16;
constexpr int kChunk = 4;  // edge slots a lane handles per step in the tuned kernel

// Relaxes edge k out of v (distance dv).  Returns whether dst must be pushed.
template <bool WEIGHTED>
__device__ __forceinline__ bool relax_edge(const int* __restrict__ col, const int* __restrict__ w,
                                           unsigned* dist, int* stamp, int k, unsigned dv, int iter,
                                           int n, int* dst_out) {
  const int dst = col[k];
  *dst_out = dst;
  if ((unsigned)dst >= (unsigned)n) return false;
  if (!WEIGHTED) {
    if (dist[dst] != kInf) return false;
    return atomicCAS(&dist[dst], kInf, (unsigned)iter) == kInf;
  } else {
    const unsigned wt = (unsigned)w[k];
    if (wt >= kInf) return false;
    const unsigned alt = dv + wt;
    if (alt >= dist[dst]) return false;
    const unsigned old = atomicMin(&dist[dst], alt);
    if (alt >= old) return false;
    return atomicExch(&stamp[dst], iter) != iter;
  }
}

template <int WGS, bool WEIGHTED>
__global__ __launch_bounds__(WGS) void k_wl_plain(gdb_kernel_ctx_t* ctx, const int* __restrict__ row,
                                                  const int* __restrict__ col,
                                                  const int* __restrict__ w, unsigned* dist,
                                                  int* stamp, int* q0, int* q1, int* idx, int n,
                                                  int* out) {
  DISCOVERY_PROTOCOL(ctx);
  const int tid = p_get_global_id(ctx, &local_ctx);
  const int threads = p_get_global_size(ctx, &local_ctx);
  int* in = q0;
  int* outq = q1;
  int* in_idx = &idx[0];
  int* out_idx = &idx[1];
  int iteration = 1, nbar = 0;
  while (gdb::ld_rlx(in_idx) > 0) {
    const int nin = gdb::ld_rlx(in_idx);
    if (nin > n) {  // a queue never holds more than n entries
      gdb::set_error(ctx, GDB_ERR_QUEUE_OVERFLOW);
      break;
    }
    GDB_FOR_UNIFORM(i, valid, tid, threads, nin) {
      int v = -1;
      if (valid) {
        v = in[i];
        if ((unsigned)v >= (unsigned)n) {
          gdb::set_error(ctx, GDB_ERR_QUEUE_OVERFLOW);
          valid = 0;
        }
      }
      const int b = valid ? row[v] : 0, e = valid ? row[v + 1] : 0;
      const unsigned dv = valid ? dist[v] : 0u;
      int maxlen = e - b;
      for (int off = 32; off > 0; off >>= 1) {
        const int o = __shfl_xor(maxlen, off, 64);
        maxlen = o > maxlen ? o : maxlen;
      }
      for (int t = 0; t < maxlen; ++t) {
        int dst = 0;
        const bool push =
            (b + t < e) && relax_edge<WEIGHTED>(col, w, dist, stamp, b + t, dv, iteration, n, &dst);
        gdb::wave_push(push, dst, outq, out_idx);
      }
    }
    if (!discovery_barrier(ctx, &local_ctx)) return;
    ++nbar;
    int* t = in;
    in = outq;
    outq = t;
    int* ti = in_idx;
    in_idx = out_idx;
    out_idx = ti;
    if (tid == 0) gdb::st_rlx(out_idx, 0);
    ++iteration;
    if (!discovery_barrier(ctx, &local_ctx)) return;  // bfs_port_kernel.cl:187-188
    ++nbar;
  }
  if (tid == 0) {
    out[0] = iteration - 1;
    out[1] = nbar;
  }
}

template <int WGS, bool WEIGHTED>
__global__ __launch_bounds__(WGS) void k_wl_tuned(gdb_kernel_ctx_t* ctx, const int* __restrict__ row,
                                                  const int* __restrict__ col,
                                                  const int* __restrict__ w, unsigned* dist,
                                                  int* stamp, int* q0, int* q1, int* tails, int n,
                                                  long long* work, int* out) {
  DISCOVERY_PROTOCOL(ctx);
  const int tid = p_get_global_id(ctx, &local_ctx);
  const int threads = p_get_global_size(ctx, &local_ctx);
  const int spare = n > 0 ? (int)((unsigned)tid % (unsigned)n) : 0;
  int iteration = 1, nbar = 0;
  long long items = 0, edges = 0;
  int nin = gdb::ld_rlx(&tails[1]);
  while (nin > 0) {
    if (nin > n) {
      gdb::set_error(ctx, GDB_ERR_QUEUE_OVERFLOW);
      break;
    }
    const int* in = (iteration & 1) ? q0 : q1;
    int* outq = (iteration & 1) ? q1 : q0;
    int* out_tail = &tails[(iteration + 1) % 3];
    if (tid == 0) gdb::st_rlx(&tails[(iteration + 2) % 3], 0);
    items += nin;
    GDB_FOR_UNIFORM(i, valid, tid, threads, nin) {
      int v = -1, b = 0, e = 0;
      unsigned dv = 0;
      if (valid) {
        v = in[i];
        if ((unsigned)v >= (unsigned)n) valid = 0;
      }
      if (valid) {
        b = row[v];
        e = row[v + 1];
        if (WEIGHTED) dv = dist[v];
        edges += e - b;
      }
      // kChunk edges per lane step (long rows: the whole wave, carrying the
      // row owner's distance): the column and weight loads, then the
      // targets' distances, then atomics only where a candidate improves
      // (lanes with none send no atomic), then the once-per-iteration stamps
      // of the improved targets, then one wave push.  Unused slots read edge
      // 0 and this thread's spare vertex, so that the slots stay branch-free.
      gdb::wave_walk_np<kChunk>(
          valid != 0, b, e, (int)dv, 0,
          [&](const bool (&ok)[kChunk], const int (&kk)[kChunk], int pd) -> int {
            const unsigned src = (unsigned)pd;
            int dst[kChunk];
            bool push[kChunk];
#pragma unroll
            for (int j = 0; j < kChunk; ++j) {
              dst[j] = -1;
              push[j] = false;
            }
            if (gdb::any_slot(ok)) {
              unsigned wt[kChunk], cur[kChunk];
              bool act[kChunk], need[kChunk];
#pragma unroll
              for (int j = 0; j < kChunk; ++j) {
                const int e0 = ok[j] ? kk[j] : 0;  // edge 0 exists here
                dst[j] = gdb::ldg(col + e0);
                wt[j] = WEIGHTED ? (unsigned)gdb::ldg(w + e0) : 1u;
              }
#pragma unroll
              for (int j = 0; j < kChunk; ++j) {
                act[j] = ok[j] && (unsigned)dst[j] < (unsigned)n && wt[j] < kInf;
                if (!ok[j]) dst[j] = -1;
                cur[j] = gdb::ldg(dist + (act[j] ? dst[j] : spare));
              }
              // A stale distance is never below the current one, so it can
              // only let a needless atomic through.
              bool any_need = false;
#pragma unroll
              for (int j = 0; j < kChunk; ++j) {
                need[j] = act[j] && (WEIGHTED ? src + wt[j] < cur[j] : cur[j] == kInf);
                any_need |= need[j];
              }
              if (any_need) {
                unsigned old[kChunk];
                bool imp[kChunk], any_imp = false;
#pragma unroll
                for (int j = 0; j < kChunk; ++j) {
                  unsigned* dp = &dist[need[j] ? dst[j] : spare];
                  old[j] = WEIGHTED ? gdb::atomic_min_g(dp, need[j] ? src + wt[j] : ~0u)
                                    : gdb::atomic_cas_g(dp, need[j] ? kInf : ~0u, need[j] ? (unsigned)iteration : ~0u);
                }
#pragma unroll
                for (int j = 0; j < kChunk; ++j) {
                  imp[j] = need[j] && (WEIGHTED ? src + wt[j] < old[j] : old[j] == kInf);
                  any_imp |= imp[j];
                }
                if (!WEIGHTED) {
#pragma unroll
                  for (int j = 0; j < kChunk; ++j) push[j] = imp[j];
                } else if (any_imp) {
                  // Stamps only grow (iteration numbers): the first to raise
                  // one to this iteration queues the target.
                  int sp[kChunk];
#pragma unroll
                  for (int j = 0; j < kChunk; ++j)
                    sp[j] = gdb::atomic_max_g(&stamp[imp[j] ? dst[j] : spare], imp[j] ? iteration : INT_MIN);
#pragma unroll
                  for (int j = 0; j < kChunk; ++j) push[j] = imp[j] && sp[j] < iteration;
                }
              }
            }
            gdb::wave_push_n<kChunk>(push, dst, outq, out_tail);
            return 0;
          },
          [](int a, int c) { return a | c; }, kSerialLimit);
    }
    if (!discovery_barrier(ctx, &local_ctx)) return;
    ++nbar;
    nin = gdb::ld_rlx(out_tail);
    ++iteration;
  }
  // Edges examined, summed over threads (one atomic per wave).
  for (int off = 32; off > 0; off >>= 1) edges += __shfl_xor(edges, off, 64);
  if (gdb::lane_id() == 0 && edges) atomicAdd((unsigned long long*)&work[1], (unsigned long long)edges);
  if (tid == 0) {
    out[0] = iteration - 1;
    out[1] = nbar;
    work[0] = items;
  }
}

// Relaunch schedule: one iteration per launch.  tails rotate as above.
template <bool WEIGHTED>
__global__ void k_wl_step(const int* __restrict__ row, const int* __restrict__ col,
                          const int* __restrict__ w, unsigned* dist, int* stamp, int* q0, int* q1,
                          int* tails, int n, int iteration, int* active) {
  // active[0] counts non-empty iterations; active[1] is a sticky done flag
  // raised by the first launch that finds an empty input.
  if (active[1]) return;
  int nin = tails[iteration % 3];
  if (nin > n) nin = n;
  if (nin == 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) active[1] = 1;
    return;
  }
  const int* in = (iteration & 1) ? q0 : q1;
  int* outq = (iteration & 1) ? q1 : q0;
  int* out_tail = &tails[(iteration + 1) % 3];
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid == 0) {
    tails[(iteration + 2) % 3] = 0;
    atomicAdd(active, 1);
  }
  const int stride = gridDim.x * blockDim.x;
  GDB_FOR_UNIFORM(i, valid, gid, stride, nin) {
    int v = -1, b = 0, e = 0;
    unsigned dv = 0;
    if (valid) {
      v = in[i];
      if ((unsigned)v >= (unsigned)n) valid = 0;
    }
    if (valid) {
      b = row[v];
      e = row[v + 1];
      if (WEIGHTED) dv = dist[v];
    }
    gdb::wave_coop_foreach(
        valid != 0, b, e, dv,
        [&](int k, bool ok, unsigned src_dist) {
          int dst = 0;
          const bool push =
              ok && relax_edge<WEIGHTED>(col, w, dist, stamp, k, src_dist, iteration, n, &dst);
          gdb::wave_push(push, dst, outq, out_tail);
        },
        kSerialLimit);
  }
}

__global__ void k_wl_init(unsigned* dist, int* stamp, int* q0, int* q1, int* flags, int n, int src) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {  // initialize (bfs_common.cl:64-71) + dist[source] = 0
    dist[i] = i == src ? 0u : kInf;
    stamp[i] = 0;
  }
  if (i < 16) flags[i] = 0;
  if (i == 0) {  // iteration 0: push the source (bfs_port_kernel.cl:130-137)
    q0[0] = src;
    q1[0] = src;
    flags[0] = 1;  // plain: input index
    flags[4 + 1] = 1;  // tuned / relaunch: tails[1] = input size of iteration 1
  }
}

// dverifysolution (bfs_common.cl:43-60, sssp_common.cl:41-56).
__global__ void k_verify(const int* __restrict__ row, const int* __restrict__ col,
                         const int* __restrict__ w, const unsigned* dist, int n,
                         unsigned long long* nerr) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n) return;
  unsigned long long bad = 0;
  for (int k = row[u]; k < row[u + 1]; ++k) {
    const int v = col[k];
    const unsigned wt = w ? (unsigned)w[k] : 1u;
    if (wt > 0 && dist[u] + wt < dist[v]) ++bad;
  }
  if (bad) atomicAdd(nerr, bad);
}

// `fixed_ws`: a caller-provided device buffer of `fixed_bytes` bytes that
// replaces the context's grow-only workspace (no allocation), or nullptr.
template <bool WEIGHTED>
int run_worklist(gdb_ctx* ctx, int n, int m, const int* row, const int* col, const int* w, int source,
                 unsigned* dist, const gdb_opts* opts, gdb_stats* stats, void* fixed_ws = nullptr,
                 size_t fixed_bytes = 0) {
  using namespace gdbrt;
  if (!ctx || n < 0 || m < 0 || (n > 0 && (!row || !dist)) || (m > 0 && !col) ||
      (WEIGHTED && m > 0 && !w) || (n > 0 && (source < 0 || source >= n)))
    return GDB_E_INVALID;
  gdb_opts o{};
  if (opts) o = *opts;
  if (o.variant == GDB_VARIANT_TUNED && o.schedule == GDB_SCHED_PERSISTENT) gdbrt::apply_tuning(WEIGHTED ? "sssp_wl" : "bfs_wl", n, &o);
  const int wgs = default_wgs(&o);
  gdb_stats st{};
  if (n == 0) {
    if (stats) *stats = st;
    return GDB_SUCCESS;
  }
  const size_t bytes = carve_bytes<int>(n) * 3 + carve_bytes<int>(32) + carve_bytes<long long>(2);
  void* base = fixed_ws;
  int rc = GDB_SUCCESS;
  if (base) {
    if (bytes > fixed_bytes) return GDB_E_SCRATCH;
  } else {
    rc = ensure_workspace(ctx, bytes);
    if (rc) return rc;
    base = ctx->ws;
  }
  Carve cv(base);
  int* stamp = cv.take<int>(n);
  int* q0 = cv.take<int>(n);
  int* q1 = cv.take<int>(n);
  int* flags = cv.take<int>(32);
  long long* work = cv.take<long long>(2);
  int* tails = flags + 4;
  int* out = flags + 8;
  rc = fill_i32(ctx, (int*)work, 0, 4);
  if (rc) return rc;
  hipLaunchKernelGGL(k_wl_init, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, dist, stamp, q0, q1,
                     flags, n, source);
  GDB_HIP(hipGetLastError());
  rc = prepare_launch(ctx);
  if (rc) return rc;

  if (o.schedule == GDB_SCHED_RELAUNCH) {
    const int check = o.relaunch_check > 1 ? o.relaunch_check : 1;
    int grid = (n + wgs - 1) / wgs;
    if (grid > 4096) grid = 4096;
    int* active = flags + 12;
    rc = timer_start(ctx);
    if (rc) return rc;
    int it = 0;
    for (;;) {
      for (int k = 0; k < check; ++k) {
        ++it;
        hipLaunchKernelGGL(k_wl_step<WEIGHTED>, dim3(grid), dim3(wgs), 0, ctx->stream, row, col, w,
                           dist, stamp, q0, q1, tails, n, it, active);
        st.launches += 1;
      }
      GDB_HIP(hipGetLastError());
      GDB_HIP(hipMemcpyAsync(ctx->h_flag, tails, 3 * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
      GDB_HIP(hipStreamSynchronize(ctx->stream));
      if (ctx->h_flag[(it + 1) % 3] == 0 || ctx->h_flag[it % 3] == 0) break;
      if (it > n + 2) return GDB_E_TIMEOUT;
    }
    rc = timer_stop(ctx, &st.kernel_ms);
    if (rc) return rc;
    int act = 0;
    GDB_HIP(hipMemcpy(&act, active, sizeof(int), hipMemcpyDeviceToHost));
    st.rounds = act;
    st.participating = grid;
  } else {
    const bool tuned = o.variant == GDB_VARIANT_TUNED;
    rc = gdb_dispatch_wgs(wgs, [&](auto W) -> int {
      constexpr int WG = decltype(W)::value;
      const void* kern = tuned ? (const void*)k_wl_tuned<WG, WEIGHTED> : (const void*)k_wl_plain<WG, WEIGHTED>;
      int groups = 0;
      int r2 = launch_groups(ctx, kern, &o, n, &groups);
      if (r2) return r2;
      r2 = timer_start(ctx);
      if (r2) return r2;
      if (tuned)
        hipLaunchKernelGGL((k_wl_tuned<WG, WEIGHTED>), dim3(groups), dim3(WG), 0, ctx->stream, ctx->d,
                           row, col, w, dist, stamp, q0, q1, tails, n, work, out);
      else
        hipLaunchKernelGGL((k_wl_plain<WG, WEIGHTED>), dim3(groups), dim3(WG), 0, ctx->stream, ctx->d,
                           row, col, w, dist, stamp, q0, q1, flags, n, out);
      GDB_HIP(hipGetLastError());
      return timer_stop(ctx, &st.kernel_ms);
    });
    if (rc) return rc;
    st.launches = 1;
    int hout[2] = {0, 0};
    GDB_HIP(hipMemcpy(hout, out, sizeof(hout), hipMemcpyDeviceToHost));
    st.rounds = hout[0];
    st.barriers = hout[1];
    if (tuned) {
      long long hw[2] = {0, 0};
      GDB_HIP(hipMemcpy(hw, work, sizeof(hw), hipMemcpyDeviceToHost));
      st.work_items = hw[0];
      st.work_edges = hw[1];
    }
    rc = read_participating(ctx, &st.participating);
    if (rc) return rc;
  }
  rc = check_device_error(ctx);
  if (stats) *stats = st;
  return rc;
}

}  // namespace

extern "C" int gdb_bfs_worklist(gdb_ctx* ctx, int n, int m, const int* row, const int* col, int source,
                                unsigned* dist, const gdb_opts* opts, gdb_stats* stats) {
  return run_worklist<false>(ctx, n, m, row, col, nullptr, source, dist, opts, stats);
}

extern "C" int gdb_sssp_worklist(gdb_ctx* ctx, int n, int m, const int* row, const int* col,
                                 const int* w, int source, unsigned* dist, const gdb_opts* opts,
                                 gdb_stats* stats) {
  return run_worklist<true>(ctx, n, m, row, col, w, source, dist, opts, stats);
}

extern "C" int gdb_sssp_worklist_noalloc(gdb_ctx* ctx, int n, int m, const int* row, const int* col,
                                         const int* w, int source, unsigned* dist, const gdb_opts* opts,
                                         gdb_stats* stats) {
  if (!ctx) return GDB_E_INVALID;
  if (!ctx->noalloc_scratch) return GDB_E_INVALID;  // resolved by gdb_ctx_create
  if (n > GDB_NOALLOC_MAX_N) return GDB_E_SCRATCH;
  return run_worklist<true>(ctx, n, m, row, col, w, source, dist, opts, stats, ctx->noalloc_scratch,
                            GDB_NOALLOC_SCRATCH_BYTES);
}

extern "C" int gdb_verify_distances(gdb_ctx* ctx, int n, int m, const int* row, const int* col,
                                    const int* w, const unsigned* dist, long long* nerr) {
  using namespace gdbrt;
  if (!ctx || !nerr || n < 0 || m < 0 || (n > 0 && (!row || !dist)) || (m > 0 && !col))
    return GDB_E_INVALID;
  *nerr = 0;
  if (n == 0) return GDB_SUCCESS;
  int rc = ensure_workspace(ctx, 256);
  if (rc) return rc;
  unsigned long long* d_err = (unsigned long long*)ctx->ws;
  rc = fill_i32(ctx, (int*)d_err, 0, 2);
  if (rc) return rc;
  hipLaunchKernelGGL(k_verify, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, row, col, w, dist, n,
                     d_err);
  GDB_HIP(hipGetLastError());
  unsigned long long h = 0;
  GDB_HIP(hipMemcpyAsync(&h, d_err, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  GDB_HIP(hipStreamSynchronize(ctx->stream));
  *nerr = (long long)h;
  return GDB_SUCCESS;
}
