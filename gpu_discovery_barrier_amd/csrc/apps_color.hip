// Pannotia graph colouring (max rule) on the discovery protocol.
//
// Rule (code/experiments/pannotia/graph_app/color/kernel/color_kernel.cl):
// round r = 1, 2, ...: for each uncoloured v, max_d[v] = max node_value[u]
// over neighbours u that are uncoloured, starting from -1, with every
// neighbour ignored when deg(v) == 1 (the `start != end - 1` test, lines
// 112/204); `stop` is raised when such a neighbour exists.  Then every
// uncoloured v with node_value[v] > max_d[v] takes colour r.  The loop ends
// after the first round whose stop flag stayed 0 (lines 240-242).
//
// Schedules:
//   plain persistent  - the mega_kernel structure (lines 159-252): two phases
//                       and two barriers per round, thread per vertex.
//   tuned persistent  - one barrier per round: phase r also applies round
//                       r-1's colouring (v coloured in r-1 iff uncoloured and
//                       node_value[v] > max_d_{r-1}[v]), so a neighbour's state
//                       is read as "uncoloured after r-1" from color[] and the
//                       previous max_d buffer.  Rows of more than 16 edges
//                       are cut into 64-edge chunks spread over every wave
//                       of the grid, their partial maxima combined by
//                       atomicMax on order-preserving int keys (exact: max
//                       does not depend on order).
//   relaunch          - color / color2 kernels per round with a host loop
//                       (color.cpp:294-334).
#include <algorithm>

#include "graph_common.hpp"
#include "runtime.hpp"

namespace {

constexpr int kSerialLimit = 16;

template <int WGS>
__global__ __launch_bounds__(WGS) void k_color_plain(gdb_kernel_ctx_t* ctx, const int* __restrict__ row,
                                                     const int* __restrict__ col,
                                                     const float* __restrict__ nv, int* color,
                                                     int* stop1, int* stop2, float* max_d, int n,
                                                     int max_rounds, int* out) {
  DISCOVERY_PROTOCOL(ctx);
  int* write_stop = stop1;
  int* read_stop = stop2;
  const int tid = p_get_global_id(ctx, &local_ctx);
  const int stride = p_get_global_size(ctx, &local_ctx);
  int graph_color = 1;
  int nbar = 0;
  for (;;) {
    bool st = false;
    for (int i = tid; i < n; i += stride) {
      if (color[i] == -1) {
        const int start = row[i], end = row[i + 1];
        float mx = -1.f;
        for (int e = start; e < end; ++e) {
          const int u = col[e];
          if (color[u] == -1 && start != end - 1) {
            st = true;
            const float x = nv[u];
            if (x > mx) mx = x;
          }
        }
        max_d[i] = mx;
      }
    }
    if (st) gdb::st_rlx(write_stop, 1);
    int* sw = read_stop;
    read_stop = write_stop;
    write_stop = sw;
    if (!discovery_barrier(ctx, &local_ctx)) return;
    ++nbar;
    for (int i = tid; i < n; i += stride)
      if (color[i] == -1 && nv[i] > max_d[i]) color[i] = graph_color;
    if (gdb::ld_rlx(read_stop) == 0) break;
    if (graph_color >= max_rounds) {
      gdb::set_error(ctx, GDB_ERR_QUEUE_OVERFLOW);
      break;
    }
    ++graph_color;
    gdb::st_rlx(write_stop, 0);
    if (!discovery_barrier(ctx, &local_ctx)) return;
    ++nbar;
  }
  if (tid == 0) {
    out[0] = graph_color;
    out[1] = nbar;
  }
}

// Max-min colouring (gdb_color_maxmin): round r gives colour 2r-1 to every
// uncoloured vertex whose value exceeds all its uncoloured neighbours' and
// colour 2r to every one below all of them (self-loops ignored).  Phase 1
// records, per uncoloured vertex, whether it is a local maximum / minimum
// (a 2-bit code, or 0); phase 2 applies the codes.  Two barriers per round:
// phase 1 reads colours that phase 2 writes.  Rows of at most kSerialLimit
// edges are walked by their lane; longer rows by the whole wave, 64 edges
// per step, with the owner's vertex and value broadcast.
template <int WGS>
__global__ __launch_bounds__(WGS) void k_color_maxmin(gdb_kernel_ctx_t* ctx, const int* __restrict__ row,
                                                      const int* __restrict__ col,
                                                      const float* __restrict__ nv, int* color, int* code,
                                                      int* stops, int n, int max_rounds, int* out) {
  DISCOVERY_PROTOCOL(ctx);
  const int tid = p_get_global_id(ctx, &local_ctx);
  const int stride = p_get_global_size(ctx, &local_ctx);
  const int lane = gdb::lane_id();
  int r = 1, nbar = 0;
  for (;;) {
    if (tid == 0) gdb::st_rlx(&stops[(r + 1) % 3], 0);
    bool any = false;
    GDB_FOR_UNIFORM(v, valid, tid, stride, n) {
      const bool act = valid && color[v] == -1;
      const int b = act ? row[v] : 0, e = act ? row[v + 1] : 0;
      const float x = act ? nv[v] : 0.f;
      any |= act;
      bool is_max = act, is_min = act;
      const bool heavy = act && (e - b) > kSerialLimit;
      if (act && !heavy) {
        for (int k = b; k < e; ++k) {
          const int u = col[k];
          if (u == v || (unsigned)u >= (unsigned)n || color[u] != -1) continue;
          const float y = nv[u];
          is_max &= x > y;
          is_min &= x < y;
        }
      }
      unsigned long long mask = __ballot(heavy);
      while (mask) {
        const int leader = __ffsll((long long)mask) - 1;
        const int lb = gdb::bcast(b, leader), le = gdb::bcast(e, leader), lv = gdb::bcast(v, leader);
        const float lx = __shfl(x, leader, 64);
        bool mx = true, mn = true;
        for (int k = lb + lane; k < le; k += 64) {
          const int u = col[k];
          if (u == lv || (unsigned)u >= (unsigned)n || color[u] != -1) continue;
          const float y = nv[u];
          mx &= lx > y;
          mn &= lx < y;
        }
        mx = !gdb::wave_any(!mx);
        mn = !gdb::wave_any(!mn);
        if (lane == leader) {
          is_max = mx;
          is_min = mn;
        }
        mask &= mask - 1;
      }
      if (act) code[v] = is_max ? 1 : (is_min ? 2 : 0);
    }
    if (gdb::wave_any(any) && lane == 0) gdb::st_rlx(&stops[r % 3], 1);
    if (!discovery_barrier(ctx, &local_ctx)) return;
    ++nbar;
    if (gdb::ld_rlx(&stops[r % 3]) == 0) break;
    if (r > max_rounds) {
      gdb::set_error(ctx, GDB_ERR_QUEUE_OVERFLOW);
      break;
    }
    for (int v = tid; v < n; v += stride)
      if (color[v] == -1 && code[v] != 0) color[v] = code[v] == 1 ? 2 * r - 1 : 2 * r;
    if (!discovery_barrier(ctx, &local_ctx)) return;
    ++nbar;
    ++r;
  }
  if (tid == 0) {
    out[0] = r - 1;
    out[1] = nbar;
  }
}

// Order-preserving int key of a float (IEEE total order, -0 taken as +0):
// key(a) > key(b) iff a > b for non-NaN a, b, so max and compares can run on
// ints and atomicMax combines partial maxima exactly.
__device__ __forceinline__ int fkey(float f) {
  const int b = __float_as_int(f + 0.f);
  return b < 0 ? b ^ 0x7fffffff : b;
}

constexpr int kHeavyDeg = 16;    // rows longer than this are split into chunks
constexpr int kChunkEdges = 64;  // edges of a heavy row per chunk (one per lane)
constexpr int kSlots = 8;        // edges a lane gathers per step on light rows
constexpr int kChunksPerStep = 4;  // heavy-row chunks a wave handles per step

template <int WGS>
__global__ __launch_bounds__(WGS) void k_color_tuned(gdb_kernel_ctx_t* ctx, const int* __restrict__ row,
                                                     const int* __restrict__ col,
                                                     const float* __restrict__ nv, int* color,
                                                     int* stops, int* mdk, const int* __restrict__ heavy,
                                                     int nheavy, const int4* __restrict__ chunks,
                                                     int nchunks, int4* clist, int n, int max_rounds,
                                                     int* out) {
  DISCOVERY_PROTOCOL(ctx);
  const int tid = p_get_global_id(ctx, &local_ctx);
  const int stride = p_get_global_size(ctx, &local_ctx);
  const int lane = gdb::lane_id();
  const int wave = tid >> 6, nwaves = stride >> 6;
  const int none = fkey(-1.f);
  // Unused gather slots read this thread's spare vertex, so that every
  // slot's loads are issued branch-free, before any of them is used.
  const int spare = n > 0 ? (int)((unsigned)tid % (unsigned)n) : 0;
  // Heavy-row chunks this wave still walks: its share of the chunk table in
  // round 1 (chunks wave + k * nwaves), then the ones whose rows were still
  // uncoloured, compacted in place into the same strided slots of clist.
  int wchunks = nchunks > wave ? (nchunks - wave + nwaves - 1) / nwaves : 0;
  // Light rows' blockers: the neighbour that held the row's maximum (-1
  // none).  max_d is the largest value among the uncoloured neighbours, and
  // neighbours only leave that set, so while the blocker is uncoloured the
  // maximum stands and the row is not walked; it is walked again once the
  // blocker is coloured.  The keys equal the full walk's every round.
  int* blk = mdk + 3 * (size_t)n;
  int r = 1;
  int nbar = 0;
  for (;;) {
    // Three key buffers: round r writes md_cur, reads md_prev, and resets
    // the heavy rows' entries of md_next (read by nobody this round).
    const int* md_prev = mdk + (size_t)((r + 2) % 3) * n;
    int* md_cur = mdk + (size_t)(r % 3) * n;
    int* md_next = mdk + (size_t)((r + 1) % 3) * n;
    const bool have_prev = r > 1;
    if (tid == 0) gdb::st_rlx(&stops[(r + 1) % 3], 0);
    for (int h = tid; h < nheavy; h += stride) md_next[heavy[h]] = none;
    // u counts as uncoloured in round r unless coloured before round r-1 or
    // by round r-1 itself (node_value[u] > max_d_{r-1}[u]): below, color[u]
    // == -1 && !(have_prev && fkey(nv[u]) > md_prev[u]).
    bool st = false;
    GDB_FOR_UNIFORM(v, valid, tid, stride, n) {
      bool active = false, blocked = false;
      int b = 0, e = 0, bk = -1, mp = 0;
      if (valid) {
        // All of v's words at once (one memory round trip), then decide.
        const int cv = color[v];
        const int kv = fkey(nv[v]);
        mp = md_prev[v];
        bk = blk[v];
        const int rb = row[v], re = row[v + 1];
        if (cv == -1) {
          if (have_prev && kv > mp) {
            color[v] = r - 1;
          } else {
            b = rb;
            e = re;
            active = (e - b) <= kHeavyDeg;  // longer rows: the chunk pass below
            if (active && bk >= 0)
              blocked = color[bk] == -1 && !(have_prev && fkey(nv[bk]) > md_prev[bk]);
          }
        }
      }
      const bool deg1 = (e - b) == 1;
      // kSlots edges per lane step: the column loads, then the three
      // gathers per neighbour, then the compares.
      auto step = [&](const bool (&ok)[kSlots], const int (&kk)[kSlots]) -> gdb::MaxAt {
        if (!gdb::any_slot(ok)) return gdb::MaxAt{-1.f, -1};  // no slot: no loads
        int u[kSlots], cu[kSlots], mu[kSlots];
        float xu[kSlots];
#pragma unroll
        for (int j = 0; j < kSlots; ++j) u[j] = gdb::ldg(col + (ok[j] ? kk[j] : 0));  // edge 0 exists here
#pragma unroll
        for (int j = 0; j < kSlots; ++j) {
          const int uu = ok[j] ? u[j] : spare;
          cu[j] = gdb::ldg(color + uu);
          xu[j] = gdb::ldg(nv + uu);
          mu[j] = gdb::ldg(md_prev + uu);
        }
        gdb::MaxAt a{-1.f, -1};
#pragma unroll
        for (int j = 0; j < kSlots; ++j) {
          const bool unc = ok[j] && cu[j] == -1 && !(have_prev && fkey(xu[j]) > mu[j]);
          if (unc && (a.at < 0 || xu[j] > a.mx)) a = gdb::MaxAt{xu[j], u[j]};
        }
        return a;
      };
      auto red = [](gdb::MaxAt a, gdb::MaxAt c) { return (c.at >= 0 && (a.at < 0 || c.mx > a.mx)) ? c : a; };
      const gdb::MaxAt res =
          gdb::wave_walk_n<kSlots>(active && !deg1 && !blocked, b, e, gdb::MaxAt{-1.f, -1}, step, red, kHeavyDeg);
      if (active) {
        if (blocked) {
          md_cur[v] = mp;  // the maximum of the previous round stands
          st = true;
        } else {
          md_cur[v] = deg1 ? none : fkey(res.mx > -1.f ? res.mx : -1.f);
          if (!deg1) {
            if (res.at != bk) blk[v] = res.at;
            if (res.at >= 0) st = true;
          }
        }
      }
    }
    // Long rows in chunks of kChunkEdges edges (one per lane), spread over
    // every wave of the grid, kChunksPerStep chunks per wave step with their
    // loads issued together; the chunks of one row meet in md_cur[v] by
    // atomicMax.  A wave keeps only the chunks of rows still uncoloured (a
    // row coloured by round r is dropped in round r+1), so the walk shrinks
    // as rows are coloured.  Compaction in place is safe: a step's loads
    // complete before its stores, which only reach slots already read.
    const int4* cin = r == 1 ? chunks : clist;
    int kept = 0;
    for (int k0 = 0; k0 < wchunks; k0 += kChunksPerStep) {
      int4 ch[kChunksPerStep];
      bool has[kChunksPerStep], keep[kChunksPerStep];
#pragma unroll
      for (int j = 0; j < kChunksPerStep; ++j) {
        has[j] = k0 + j < wchunks;
        ch[j] = cin[wave + (size_t)(has[j] ? k0 + j : k0) * nwaves];  // (row, first edge, end edge)
      }
      int u[kChunksPerStep], cv[kChunksPerStep], mv[kChunksPerStep];
      float xv[kChunksPerStep];
#pragma unroll
      for (int j = 0; j < kChunksPerStep; ++j) {
        const int k = ch[j].y + lane;
        u[j] = gdb::ldg(col + (k < ch[j].z ? k : ch[j].y));
        if (k >= ch[j].z) u[j] = -1;
        cv[j] = gdb::ldg(color + ch[j].x);
        xv[j] = gdb::ldg(nv + ch[j].x);
        mv[j] = gdb::ldg(md_prev + ch[j].x);
      }
      int cu[kChunksPerStep], mu[kChunksPerStep];
      float xu[kChunksPerStep];
#pragma unroll
      for (int j = 0; j < kChunksPerStep; ++j) {
        const int uu = u[j] >= 0 ? u[j] : spare;
        cu[j] = gdb::ldg(color + uu);
        xu[j] = gdb::ldg(nv + uu);
        mu[j] = gdb::ldg(md_prev + uu);
      }
#pragma unroll
      for (int j = 0; j < kChunksPerStep; ++j) {
        keep[j] = has[j] && cv[j] == -1 && !(have_prev && fkey(xv[j]) > mv[j]);  // wave-uniform
        if (keep[j]) {
          if (lane == 0) clist[wave + (size_t)kept * nwaves] = ch[j];
          ++kept;
        }
      }
      // Each chunk's maximum key over its uncoloured neighbours (INT_MIN:
      // none; only a NaN priority could have that key), the four wave reductions
      // independent of each other.
      int key[kChunksPerStep];
#pragma unroll
      for (int j = 0; j < kChunksPerStep; ++j) {
        const bool unc = u[j] >= 0 && cu[j] == -1 && !(have_prev && fkey(xu[j]) > mu[j]);
        key[j] = gdb::wave_max(unc ? fkey(xu[j]) : INT_MIN);
      }
#pragma unroll
      for (int j = 0; j < kChunksPerStep; ++j)
        if (keep[j] && key[j] != INT_MIN) {  // wave-uniform
          st = true;
          if (lane == 0) atomicMax(&md_cur[ch[j].x], key[j]);
        }
    }
    wchunks = kept;
    if (gdb::wave_any(st) && lane == 0) gdb::st_rlx(&stops[r % 3], 1);
    if (!discovery_barrier(ctx, &local_ctx)) return;
    ++nbar;
    if (gdb::ld_rlx(&stops[r % 3]) == 0) break;
    if (r >= max_rounds) {
      gdb::set_error(ctx, GDB_ERR_QUEUE_OVERFLOW);
      break;
    }
    ++r;
  }
  // Round r's own colouring (the reference's color2 of the last round).
  const int* md = mdk + (size_t)(r % 3) * n;
  for (int v = tid; v < n; v += stride)
    if (color[v] == -1 && fkey(nv[v]) > md[v]) color[v] = r;
  if (tid == 0) {
    out[0] = r;
    out[1] = nbar;
  }
}

// ---- relaunch schedule: color (color_kernel.cl:80-123) and color2 (134-154)
// `flags` = {stop[0..2], done, last}.  Round r writes stop[r % 3]; a kernel of
// round r is a no-op once round r-1 ended with stop 0 (done is sticky).
__global__ void k_color_r1(const int* __restrict__ row, const int* __restrict__ col,
                           const float* __restrict__ nv, const int* color, int* flags,
                           float* max_d, int n, int r) {
  if (flags[3] || (r > 1 && flags[(r - 1) % 3] == 0)) {
    if (blockIdx.x == 0 && threadIdx.x == 0 && !flags[3]) {
      flags[3] = 1;
      flags[4] = r - 1;
    }
    return;
  }
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && color[i] == -1) {
    const int start = row[i], end = row[i + 1];
    float mx = -1.f;
    bool st = false;
    for (int e = start; e < end; ++e) {
      const int u = col[e];
      if (color[u] == -1 && start != end - 1) {
        st = true;
        const float x = nv[u];
        if (x > mx) mx = x;
      }
    }
    max_d[i] = mx;
    if (st) flags[r % 3] = 1;
  }
}

__global__ void k_color_r2(const float* __restrict__ nv, int* color, int* flags,
                           const float* max_d, int n, int r) {
  if (flags[3] || (r > 1 && flags[(r - 1) % 3] == 0)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && color[i] == -1 && nv[i] > max_d[i]) color[i] = r;
  if (i == 0) flags[(r + 1) % 3] = 0;
}

__global__ void k_color_init(int* color, float* max_d, int* mdk, int* flags, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    color[i] = -1;
    max_d[i] = -1.f;
    max_d[i + n] = -1.f;
    if (mdk) {
      const int none = fkey(-1.f);
      mdk[i] = none;
      mdk[i + n] = none;
      mdk[i + 2 * (size_t)n] = none;
      mdk[i + 3 * (size_t)n] = -1;  // blockers: none yet
    }
  }
  if (i < 8) flags[i] = 0;
}

}  // namespace

using namespace gdbrt;

extern "C" int gdb_color(gdb_ctx* ctx, int n, int m, const int* row, const int* col,
                         const float* node_value, int* color, const gdb_opts* opts,
                         gdb_stats* stats) {
  if (!ctx || n < 0 || m < 0 || (n > 0 && (!row || !node_value || !color)) || (m > 0 && !col))
    return GDB_E_INVALID;
  gdb_opts o{};
  if (opts) o = *opts;
  if (o.variant == GDB_VARIANT_TUNED && o.schedule == GDB_SCHED_PERSISTENT) gdbrt::apply_tuning("color", n, &o);
  const int wgs = default_wgs(&o);
  gdb_stats s{};
  if (n == 0) {
    if (stats) *stats = s;
    return GDB_SUCCESS;
  }
  const bool tuned_persistent = o.variant == GDB_VARIANT_TUNED && o.schedule != GDB_SCHED_RELAUNCH;
  // Tuned schedule: rows longer than kHeavyDeg are split into chunks of
  // kChunkEdges edges; the rows and the (row, first edge) of every chunk
  // are listed here from the row offsets.
  std::vector<int> heavy;
  std::vector<int4> chunks;
  if (tuned_persistent) {
    std::vector<int> hrow((size_t)n + 1);
    GDB_HIP(hipMemcpy(hrow.data(), row, sizeof(int) * ((size_t)n + 1), hipMemcpyDeviceToHost));
    for (int v = 0; v < n; ++v) {
      const int d = hrow[v + 1] - hrow[v];
      if (d < 0 || hrow[v] < 0 || hrow[v + 1] > m) return GDB_E_INVALID;
      if (d > kHeavyDeg) {
        heavy.push_back(v);
        for (int eb = hrow[v]; eb < hrow[v + 1]; eb += kChunkEdges)
          chunks.push_back(make_int4(v, eb, std::min(eb + kChunkEdges, hrow[v + 1]), 0));
      }
    }
  }
  const int nheavy = (int)heavy.size();
  const int nchunks = (int)chunks.size();
  const size_t bytes = carve_bytes<float>((size_t)2 * n) + carve_bytes<int>(16) +
                       (tuned_persistent ? carve_bytes<int>((size_t)4 * n) + carve_bytes<int>(nheavy + 1) +
                                               carve_bytes<int4>(nchunks + 1) * 2
                                         : 0);
  int rc = ensure_workspace(ctx, bytes);
  if (rc) return rc;
  Carve cv(ctx->ws);
  float* maxd = cv.take<float>((size_t)2 * n);
  int* flags = cv.take<int>(16);
  int* mdk = tuned_persistent ? cv.take<int>((size_t)4 * n) : nullptr;  // 3 key buffers + blockers
  int* d_heavy = tuned_persistent ? cv.take<int>(nheavy + 1) : nullptr;
  int4* d_chunks = tuned_persistent ? cv.take<int4>(nchunks + 1) : nullptr;
  int4* d_clist = tuned_persistent ? cv.take<int4>(nchunks + 1) : nullptr;
  if (nheavy) {
    GDB_HIP(hipMemcpyAsync(d_heavy, heavy.data(), sizeof(int) * nheavy, hipMemcpyHostToDevice, ctx->stream));
    GDB_HIP(hipMemcpyAsync(d_chunks, chunks.data(), sizeof(int4) * nchunks, hipMemcpyHostToDevice,
                           ctx->stream));
  }
  const int max_rounds = n + 1;
  const int init_groups = (n + 255) / 256;
  hipLaunchKernelGGL(k_color_init, dim3(init_groups), dim3(256), 0, ctx->stream, color, maxd, mdk, flags, n);
  GDB_HIP(hipGetLastError());
  rc = prepare_launch(ctx);
  if (rc) return rc;

  if (o.schedule == GDB_SCHED_RELAUNCH) {
    const int grid = (n + wgs - 1) / wgs;
    const int check = o.relaunch_check > 1 ? o.relaunch_check : 1;
    rc = timer_start(ctx);
    if (rc) return rc;
    int r = 1, last = 0;
    for (;;) {
      for (int k = 0; k < check; ++k, ++r) {
        hipLaunchKernelGGL(k_color_r1, dim3(grid), dim3(wgs), 0, ctx->stream, row, col, node_value,
                           color, flags, maxd, n, r);
        hipLaunchKernelGGL(k_color_r2, dim3(grid), dim3(wgs), 0, ctx->stream, node_value, color,
                           flags, maxd, n, r);
        s.launches += 2;
      }
      GDB_HIP(hipGetLastError());
      GDB_HIP(hipMemcpyAsync(ctx->h_flag, flags, 4 * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
      GDB_HIP(hipStreamSynchronize(ctx->stream));
      const int lr = r - 1;  // last launched round
      if (ctx->h_flag[3]) {
        int tail[2];
        GDB_HIP(hipMemcpy(tail, flags + 3, 2 * sizeof(int), hipMemcpyDeviceToHost));
        last = tail[1];
        break;
      }
      if (ctx->h_flag[lr % 3] == 0) {
        last = lr;
        break;
      }
      if (lr >= max_rounds) return GDB_E_TIMEOUT;
    }
    rc = timer_stop(ctx, &s.kernel_ms);
    if (rc) return rc;
    s.rounds = last;
    s.participating = grid;
  } else {
    const bool tuned = o.variant == GDB_VARIANT_TUNED;
    int* out = flags + 8;
    rc = gdb_dispatch_wgs(wgs, [&](auto W) -> int {
      constexpr int WG = decltype(W)::value;
      const void* kern = tuned ? (const void*)k_color_tuned<WG> : (const void*)k_color_plain<WG>;
      int groups = 0;
      int r2 = launch_groups(ctx, kern, &o, n, &groups);
      if (r2) return r2;
      r2 = timer_start(ctx);
      if (r2) return r2;
      if (tuned)
        hipLaunchKernelGGL(k_color_tuned<WG>, dim3(groups), dim3(WG), 0, ctx->stream, ctx->d, row,
                           col, node_value, color, flags, mdk, d_heavy, nheavy, d_chunks, nchunks,
                           d_clist, n, max_rounds, out);
      else
        hipLaunchKernelGGL(k_color_plain<WG>, dim3(groups), dim3(WG), 0, ctx->stream, ctx->d, row,
                           col, node_value, color, flags, flags + 4, maxd, n, max_rounds, out);
      GDB_HIP(hipGetLastError());
      return timer_stop(ctx, &s.kernel_ms);
    });
    if (rc) return rc;
    s.launches = 1;
    int hout[2] = {0, 0};
    GDB_HIP(hipMemcpy(hout, out, sizeof(hout), hipMemcpyDeviceToHost));
    s.rounds = hout[0];
    s.barriers = hout[1];
    rc = read_participating(ctx, &s.participating);
    if (rc) return rc;
  }
  rc = check_device_error(ctx);
  if (rc == GDB_E_OVERFLOW) rc = GDB_E_TIMEOUT;  // round cap hit: equal priorities never resolve
  s.work_edges = m;
  s.work_items = n;
  if (stats) *stats = s;
  return rc;
}

extern "C" int gdb_color_maxmin(gdb_ctx* ctx, int n, int m, const int* row, const int* col,
                                const float* node_value, int* color, const gdb_opts* opts, gdb_stats* stats) {
  if (!ctx || n < 0 || m < 0 || (n > 0 && (!row || !node_value || !color)) || (m > 0 && !col))
    return GDB_E_INVALID;
  gdb_opts o{};
  if (opts) o = *opts;
  const int wgs = default_wgs(&o);
  gdb_stats s{};
  if (n == 0) {
    if (stats) *stats = s;
    return GDB_SUCCESS;
  }
  const size_t bytes = carve_bytes<int>(n) + carve_bytes<int>(16);
  int rc = ensure_workspace(ctx, bytes);
  if (rc) return rc;
  Carve cv(ctx->ws);
  int* code = cv.take<int>(n);
  int* flags = cv.take<int>(16);
  rc = fill_i32(ctx, color, -1, n);
  if (!rc) rc = fill_i32(ctx, flags, 0, 16);
  if (!rc) rc = prepare_launch(ctx);
  if (rc) return rc;
  int* out = flags + 8;
  rc = gdb_dispatch_wgs(wgs, [&](auto W) -> int {
    constexpr int WG = decltype(W)::value;
    int groups = 0;
    int r2 = launch_groups(ctx, (const void*)k_color_maxmin<WG>, &o, n, &groups);
    if (r2) return r2;
    r2 = timer_start(ctx);
    if (r2) return r2;
    hipLaunchKernelGGL(k_color_maxmin<WG>, dim3(groups), dim3(WG), 0, ctx->stream, ctx->d, row, col, node_value,
                       color, code, flags, n, n + 1, out);
    GDB_HIP(hipGetLastError());
    return timer_stop(ctx, &s.kernel_ms);
  });
  if (rc) return rc;
  s.launches = 1;
  int hout[2] = {0, 0};
  GDB_HIP(hipMemcpy(hout, out, sizeof(hout), hipMemcpyDeviceToHost));
  s.rounds = hout[0];
  s.barriers = hout[1];
  rc = read_participating(ctx, &s.participating);
  if (rc) return rc;
  rc = check_device_error(ctx);
  if (rc == GDB_E_OVERFLOW) rc = GDB_E_TIMEOUT;  // round cap hit: adjacent equal values never resolve
  s.work_edges = m;
  s.work_items = n;
  if (stats) *stats = s;
  return rc;
}
