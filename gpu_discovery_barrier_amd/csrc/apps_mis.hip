// Pannotia maximal independent set on the discovery protocol.
//
// Rule (code/experiments/pannotia/graph_app/mis/kernel/mis_kernel.cl): per
// round, every undecided vertex (c == -1) takes the minimum node_value over
// its undecided neighbours (start BIGNUM, lines 240-266); an undecided vertex
// whose own value is below that minimum joins the set (s = 2, c = -2) and
// marks its undecided neighbours inactive (lines 276-313, through cu, copied
// into c by mis3, lines 330-334).  The loop ends after the first round that
// found no undecided vertex.
//
// Schedules:
//   plain persistent - mega_kernel (lines 220-337): three phases and three
//                      barriers per round over all n vertices.
//   tuned persistent - two barriers per round over a queue of the undecided
//                      vertices: phase A computes the minimum over neighbours
//                      from a packed (status, value) gather and decides joins
//                      (s = 2); phase B retires joiners and vertices with a
//                      joining neighbour (c = -2) and compacts the survivors
//                      into the next queue.  Same joins as the reference: a
//                      join only reads statuses fixed in the previous round.
//   relaunch         - init / mis1 / mis2 / mis3 kernels with a host loop
//                      (mis.cpp:314-363).
#include "graph_common.hpp"
#include "runtime.hpp"

namespace {

constexpr float kBigNum = 99999999.0f;  // mis_kernel.cl:68
constexpr int kSerialLimit = 16;

template <int WGS>
__global__ __launch_bounds__(WGS) void k_mis_plain(gdb_kernel_ctx_t* ctx, const int* __restrict__ row,
                                                   const int* __restrict__ col,
                                                   const float* __restrict__ nv, int* s, int* c,
                                                   int* cu, float* min_array, int* stop, int n,
                                                   int max_rounds, int* out) {
  DISCOVERY_PROTOCOL(ctx);
  const int tid_start = p_get_global_id(ctx, &local_ctx);
  const int stride = p_get_global_size(ctx, &local_ctx);
  int rounds = 0, nbar = 0;
  for (;;) {
    ++rounds;
    bool st = false;
    for (int t = tid_start; t < n; t += stride) {  // mis1
      if (c[t] == -1) {
        st = true;
        float mn = kBigNum;
        for (int e = row[t]; e < row[t + 1]; ++e) {
          const int u = col[e];
          if (c[u] == -1 && nv[u] < mn) mn = nv[u];
        }
        min_array[t] = mn;
      }
    }
    if (st) gdb::st_rlx(stop, 1);
    if (!discovery_barrier(ctx, &local_ctx)) return;
    ++nbar;
    const int local_stop = gdb::ld_rlx(stop);
    for (int t = tid_start; t < n; t += stride) {  // mis2
      if (nv[t] < min_array[t] && c[t] == -1) {
        s[t] = 2;
        c[t] = -2;
        for (int e = row[t]; e < row[t + 1]; ++e)
          if (c[col[e]] == -1) cu[col[e]] = -2;
      }
    }
    if (!discovery_barrier(ctx, &local_ctx)) return;
    ++nbar;
    if (local_stop == 0) break;
    if (rounds > max_rounds) {
      gdb::set_error(ctx, GDB_ERR_QUEUE_OVERFLOW);
      break;
    }
    gdb::st_rlx(stop, 0);
    for (int t = tid_start; t < n; t += stride)  // mis3
      if (cu[t] == -2) c[t] = cu[t];
    if (!discovery_barrier(ctx, &local_ctx)) return;
    ++nbar;
  }
  if (tid_start == 0) {
    out[0] = rounds;
    out[1] = nbar;
  }
}

constexpr int kSlots = 4;  // edges a lane gathers per step in the tuned walks

// Packed per-vertex state for one 8-byte gather per edge: x = status
// (-1 undecided, -2 decided), y = node_value bits.
template <int WGS>
__global__ __launch_bounds__(WGS) void k_mis_tuned(gdb_kernel_ctx_t* ctx, const int* __restrict__ row,
                                                   const int* __restrict__ col, int2* cv, int* s,
                                                   int* q0, int* q1, int* tails, int n,
                                                   int max_rounds, int* out) {
  DISCOVERY_PROTOCOL(ctx);
  const int tid = p_get_global_id(ctx, &local_ctx);
  const int stride = p_get_global_size(ctx, &local_ctx);
  const int lane = gdb::lane_id();
  // Unused edge slots gather this thread's spare vertex, so that a step's
  // loads are issued branch-free, before any of them is used.
  const int spare = n > 0 ? (int)((unsigned)tid % (unsigned)n) : 0;
  int* qin = q0;
  int* qout = q1;
  int rounds = 0, nbar = 0;
  int qn = n;  // round 1: every vertex, queue holds 0..n-1
  for (;;) {
    ++rounds;
    if (qn == 0) break;
    if (qn > n) {  // the queue never holds more than n entries
      gdb::set_error(ctx, GDB_ERR_QUEUE_OVERFLOW);
      break;
    }
    if (rounds > max_rounds) {
      gdb::set_error(ctx, GDB_ERR_QUEUE_OVERFLOW);
      break;
    }
    // Phase A: minimum over undecided neighbours, join decision.
    GDB_FOR_UNIFORM(i, valid, tid, stride, qn) {
      int v = -1, b = 0, e = 0;
      if (valid) {
        v = rounds == 1 ? i : qin[i];
        if ((unsigned)v >= (unsigned)n) valid = 0;
      }
      if (valid) {
        b = row[v];
        e = row[v + 1];
      }
      // v joins iff its value is below kBigNum and below every undecided
      // neighbour's (the reference's minimum, lines 240-266, compared with
      // <): the walk stops at the first undecided neighbour whose value is
      // not above v's.  kSlots edges per lane step: the column loads, then
      // the 8-byte (status, value) gathers, then the compares.
      const float vv = valid ? __int_as_float(cv[v].y) : kBigNum;
      auto step = [&](const bool (&ok)[kSlots], const int (&kk)[kSlots], int own) -> bool {
        if (!gdb::any_slot(ok)) return false;  // no slot: no loads
        const float ov = __int_as_float(own);
        int u[kSlots];
        long long x[kSlots];
#pragma unroll
        for (int j = 0; j < kSlots; ++j) u[j] = gdb::ldg(col + (ok[j] ? kk[j] : 0));  // edge 0 exists here
#pragma unroll
        for (int j = 0; j < kSlots; ++j) {
          const bool in = ok[j] && (unsigned)u[j] < (unsigned)n;
          x[j] = gdb::ldg((const long long*)(cv + (in ? u[j] : spare)));
          if (!in) x[j] = 0;  // status 0: not undecided
        }
        bool hit = false;
#pragma unroll
        for (int j = 0; j < kSlots; ++j)
          hit |= (int)(x[j] & 0xffffffffll) == -1 && __int_as_float((int)(x[j] >> 32)) <= ov;
        return hit;
      };
      const bool blocked =
          gdb::wave_walk_any<kSlots>(valid != 0, b, e, __float_as_int(vv), step, kSerialLimit);
      if (valid && vv < kBigNum && !blocked) s[v] = 2;
    }
    if (tid == 0) gdb::st_rlx(&tails[(rounds + 1) % 3], 0);
    if (!discovery_barrier(ctx, &local_ctx)) return;
    ++nbar;
    // Phase B: retire joiners and their neighbours, compact survivors.
    int* tail = &tails[rounds % 3];
    GDB_FOR_UNIFORM(i, valid, tid, stride, qn) {
      int v = -1, b = 0, e = 0;
      bool joined = false;
      if (valid) {
        v = rounds == 1 ? i : qin[i];
        if ((unsigned)v >= (unsigned)n) valid = 0;
      }
      if (valid) {
        joined = s[v] == 2;
        if (!joined) {
          b = row[v];
          e = row[v + 1];
        }
      }
      // A joining neighbour retires v: the walk stops at the first one.
      auto step = [&](const bool (&ok)[kSlots], const int (&kk)[kSlots], int) -> bool {
        if (!gdb::any_slot(ok)) return false;  // no slot: no loads
        int u[kSlots], sj[kSlots];
#pragma unroll
        for (int j = 0; j < kSlots; ++j) u[j] = gdb::ldg(col + (ok[j] ? kk[j] : 0));
#pragma unroll
        for (int j = 0; j < kSlots; ++j) {
          const bool in = ok[j] && (unsigned)u[j] < (unsigned)n;
          sj[j] = gdb::ldg(s + (in ? u[j] : spare));
          if (!in) sj[j] = 0;
        }
        bool h = false;
#pragma unroll
        for (int j = 0; j < kSlots; ++j) h |= sj[j] == 2;
        return h;
      };
      const bool hit = gdb::wave_walk_any<kSlots>(valid && !joined, b, e, 0, step, kSerialLimit);
      const bool retire = valid && (joined || hit);
      if (retire) cv[v].x = -2;
      const bool keep = valid && !retire;
      const unsigned long long m = __ballot(keep);
      const int cnt = __popcll(m);
      int base = 0;
      if (lane == 0 && cnt) base = gdb::fadd_rlx(tail, cnt);
      base = __shfl(base, 0, 64);
      if (keep) qout[base + __popcll(m & ((1ull << lane) - 1ull))] = v;
    }
    if (!discovery_barrier(ctx, &local_ctx)) return;
    ++nbar;
    qn = gdb::ld_rlx(tail);
    int* t2 = qin;
    qin = qout;
    qout = t2;
  }
  if (tid == 0) {
    out[0] = rounds;
    out[1] = nbar;
  }
}

__global__ void k_mis_init(const float* __restrict__ nv, int* s, int* c, int* cu, int2* cv, int* flags,
                           int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {  // init kernel, mis_kernel.cl:78-93
    c[i] = -1;
    cu[i] = -1;
    s[i] = 0;
    cv[i] = make_int2(-1, __float_as_int(nv[i]));
  }
  if (i < 16) flags[i] = 0;
}

// ---- relaunch kernels: mis1 / mis2 / mis3 (mis_kernel.cl:107-216) ----------
// flags = {stop[0..2], active}: round r's mis1 raises stop[r % 3] when an
// undecided vertex exists; mis2 counts such rounds; mis3 clears the stop
// flag of round r + 1.  Rounds after the last active one change nothing.
__global__ void k_mis1(const int* __restrict__ row, const int* __restrict__ col,
                       const float* __restrict__ nv, const int* c, float* min_array, int* flags,
                       int n, int r) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n && c[t] == -1) {
    flags[r % 3] = 1;
    float mn = kBigNum;
    for (int e = row[t]; e < row[t + 1]; ++e)
      if (c[col[e]] == -1 && nv[col[e]] < mn) mn = nv[col[e]];
    min_array[t] = mn;
  }
}
__global__ void k_mis2(const int* __restrict__ row, const int* __restrict__ col,
                       const float* __restrict__ nv, int* s, int* c, int* cu,
                       const float* min_array, int* flags, int n, int r) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n && nv[t] < min_array[t] && c[t] == -1) {
    s[t] = 2;
    c[t] = -2;
    for (int e = row[t]; e < row[t + 1]; ++e)
      if (c[col[e]] == -1) cu[col[e]] = -2;
  }
  if (t == 0 && flags[r % 3]) flags[3] += 1;
}
__global__ void k_mis3(const int* cu, int* c, int* flags, int n, int r) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n && cu[t] == -2) c[t] = cu[t];
  if (t == 0) flags[(r + 1) % 3] = 0;
}

}  // namespace

using namespace gdbrt;

extern "C" int gdb_mis(gdb_ctx* ctx, int n, int m, const int* row, const int* col,
                       const float* node_value, int* s, const gdb_opts* opts, gdb_stats* stats) {
  if (!ctx || n < 0 || m < 0 || (n > 0 && (!row || !node_value || !s)) || (m > 0 && !col))
    return GDB_E_INVALID;
  gdb_opts o{};
  if (opts) o = *opts;
  if (o.variant == GDB_VARIANT_TUNED && o.schedule == GDB_SCHED_PERSISTENT) gdbrt::apply_tuning("mis", n, &o);
  const int wgs = default_wgs(&o);
  gdb_stats st{};
  if (n == 0) {
    if (stats) *stats = st;
    return GDB_SUCCESS;
  }
  const size_t bytes = carve_bytes<int>(n) * 4 + carve_bytes<int2>(n) + carve_bytes<float>(n) +
                       carve_bytes<int>(32);
  int rc = ensure_workspace(ctx, bytes);
  if (rc) return rc;
  Carve cvw(ctx->ws);
  int* c = cvw.take<int>(n);
  int* cu = cvw.take<int>(n);
  int* q0 = cvw.take<int>(n);
  int* q1 = cvw.take<int>(n);
  int2* cv = cvw.take<int2>(n);
  float* min_array = cvw.take<float>(n);
  int* flags = cvw.take<int>(32);
  const int max_rounds = n + 1;
  hipLaunchKernelGGL(k_mis_init, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, node_value, s, c,
                     cu, cv, flags, n);
  GDB_HIP(hipGetLastError());
  rc = prepare_launch(ctx);
  if (rc) return rc;

  if (o.schedule == GDB_SCHED_RELAUNCH) {
    const int grid = (n + wgs - 1) / wgs;
    const int check = o.relaunch_check > 1 ? o.relaunch_check : 1;
    rc = timer_start(ctx);
    if (rc) return rc;
    int r = 0;
    for (;;) {
      for (int k = 0; k < check; ++k) {
        ++r;
        hipLaunchKernelGGL(k_mis1, dim3(grid), dim3(wgs), 0, ctx->stream, row, col, node_value, c,
                           min_array, flags, n, r);
        hipLaunchKernelGGL(k_mis2, dim3(grid), dim3(wgs), 0, ctx->stream, row, col, node_value, s, c,
                           cu, min_array, flags, n, r);
        hipLaunchKernelGGL(k_mis3, dim3(grid), dim3(wgs), 0, ctx->stream, cu, c, flags, n, r);
        st.launches += 3;
      }
      GDB_HIP(hipGetLastError());
      GDB_HIP(hipMemcpyAsync(ctx->h_flag, flags, 4 * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
      GDB_HIP(hipStreamSynchronize(ctx->stream));
      if (ctx->h_flag[r % 3] == 0) break;
      if (r > max_rounds) return GDB_E_TIMEOUT;
    }
    rc = timer_stop(ctx, &st.kernel_ms);
    if (rc) return rc;
    st.rounds = ctx->h_flag[3] + 1;
    st.participating = grid;
  } else {
    const bool tuned = o.variant == GDB_VARIANT_TUNED;
    int* out = flags + 8;
    rc = gdb_dispatch_wgs(wgs, [&](auto W) -> int {
      constexpr int WG = decltype(W)::value;
      const void* kern = tuned ? (const void*)k_mis_tuned<WG> : (const void*)k_mis_plain<WG>;
      int groups = 0;
      int r2 = launch_groups(ctx, kern, &o, n, &groups);
      if (r2) return r2;
      r2 = timer_start(ctx);
      if (r2) return r2;
      if (tuned)
        hipLaunchKernelGGL(k_mis_tuned<WG>, dim3(groups), dim3(WG), 0, ctx->stream, ctx->d, row, col,
                           cv, s, q0, q1, flags + 4, n, max_rounds, out);
      else
        hipLaunchKernelGGL(k_mis_plain<WG>, dim3(groups), dim3(WG), 0, ctx->stream, ctx->d, row, col,
                           node_value, s, c, cu, min_array, flags, n, max_rounds, out);
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
  if (rc == GDB_E_OVERFLOW) rc = GDB_E_TIMEOUT;  // round cap: equal priorities never resolve
  st.work_edges = m;
  st.work_items = n;
  if (stats) *stats = st;
  return rc;
}
