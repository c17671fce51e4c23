// Pannotia betweenness centrality (Brandes) on the discovery protocol.
//
// Per source s (code/experiments/pannotia/graph_app/bc/kernel/bc_kernel.cl):
// clean_1d_array resets sigma (the dependency delta), rho (path counts) and
// the distance (lines 95-116); the forward BFS sets d and accumulates
// rho[w] += rho[v] over edges into the next level (mega_bfs_kernel_func,
// 295-371); the backward pass, level by level from the deepest, adds
// rho[w]/rho[v] * (1 + delta[v]) into delta[w] for predecessors w read from
// the transposed CSR and credits bc[v] += delta[v] for v != s
// (mega_backtrack_kernel_func, 373-428).  No 1/2 factor.
//
// The reference's persistent kernel loops over every vertex as a source and
// scales its reset stride by the workgroup size (line 454), so resets skip
// vertices; the n^2 array `p` is never read.  Here the caller passes the
// source list and every vertex is reset.
//
// Schedules:
//   plain persistent - the mega kernel's structure: every level scans all n
//                      vertices; rho and delta accumulate with float atomics.
//   tuned persistent - level queues, one barrier per level: the phase that
//                      expands level l (claiming level l+1 vertices with a
//                      compare-and-swap on the level) also pulls rho for the
//                      level-l vertices from their in-edges (one walk when the
//                      graph is symmetric); the backward pass pulls delta
//                      from out-edges.  Level and rho (later (1+delta)/rho)
//                      share an 8-byte word: one gather per edge.  No float
//                      atomics, so results do not depend on scheduling.
//   relaunch         - bfs_kernel / backtrack_kernel per level with a host
//                      loop (bc.cpp:322-419).
#include "graph_common.hpp"
#include "runtime.hpp"
#include <climits>

namespace {

constexpr int kSerialLimit = 16;
constexpr int kChunk = 4;  // edge slots a lane handles per step in the tuned walks

__device__ __forceinline__ void fatomic_add(float* p, float v) { unsafeAtomicAdd(p, v); }

// Level of queue slot i (a corrupt entry reads as unreachable).
__device__ __forceinline__ int level_of(const int* d, const int* q, int i, int n) {
  const int v = q[i];
  return (unsigned)v < (unsigned)n ? d[v] : -1;
}

template <int WGS>
__global__ __launch_bounds__(WGS) void k_bc_plain(gdb_kernel_ctx_t* ctx, const int* __restrict__ row,
                                                  const int* __restrict__ col,
                                                  const int* __restrict__ row_t,
                                                  const int* __restrict__ col_t, int* d, float* rho,
                                                  float* delta, float* bc, const int* sources,
                                                  int nsrc, int* stops, int n, int* out) {
  DISCOVERY_PROTOCOL(ctx);
  const int tid = p_get_global_id(ctx, &local_ctx);
  const int stride = p_get_global_size(ctx, &local_ctx);
  int nbar = 0, levels = 0;
  for (int i = tid; i < n; i += stride) bc[i] = 0.f;  // clean_bc
  for (int si = 0; si < nsrc; ++si) {
    const int s = sources[si];
    for (int i = tid; i < n; i += stride) {  // clean_1d_array
      delta[i] = 0.f;
      rho[i] = i == s ? 1.f : 0.f;
      d[i] = i == s ? 0 : -1;
    }
    if (tid == 0) {
      gdb::st_rlx(&stops[0], 0);
      gdb::st_rlx(&stops[1], 0);
    }
    if (!discovery_barrier(ctx, &local_ctx)) return;
    ++nbar;
    // Forward BFS with path counts (mega_bfs_kernel_func).
    int ld = 0;
    for (;;) {
      bool cont = false;
      for (int i = tid; i < n; i += stride) {
        if (d[i] != ld) continue;
        const float ri = rho[i];
        for (int e = row[i]; e < row[i + 1]; ++e) {
          const int w = col[e];
          if (d[w] < 0) {
            cont = true;
            d[w] = ld + 1;
          }
          if (d[w] == ld + 1) fatomic_add(&rho[w], ri);
        }
      }
      if (cont) gdb::st_rlx(&stops[ld % 3], 1);
      if (tid == 0) gdb::st_rlx(&stops[(ld + 1) % 3], 0);
      ++ld;
      if (!discovery_barrier(ctx, &local_ctx)) return;
      ++nbar;
      if (gdb::ld_rlx(&stops[(ld - 1) % 3]) == 0) break;
    }
    levels += ld;
    // Backward accumulation (mega_backtrack_kernel_func).
    while (ld > 0) {
      for (int i = tid; i < n; i += stride) {
        if (d[i] != ld - 1) continue;
        const float ri = rho[i], di = delta[i];
        for (int e = row_t[i]; e < row_t[i + 1]; ++e) {
          const int w = col_t[e];
          if (d[w] == ld - 2) fatomic_add(&delta[w], rho[w] / ri * (1.f + di));
        }
        if (i != s) bc[i] = bc[i] + di;
      }
      --ld;
      if (!discovery_barrier(ctx, &local_ctx)) return;
      ++nbar;
    }
  }
  if (tid == 0) {
    out[0] = levels;
    out[1] = nbar;
  }
}

// Level-queue BC.  q holds the reached vertices of the current source level
// by level; lvl_off[l] is the start of level l in q (lvl_off has n+2 slots).
// Per-vertex state of the tuned schedule in one 8-byte word, so that an edge
// costs one gather: x = level (-1 unreached); y = the bits of sigma in the
// forward pass, then of coef = (1 + delta) / sigma once the vertex's delta
// is final in the backward pass (delta itself is only needed there).
__device__ __forceinline__ int2 ld_state(const int2* ds, int v) {
  const long long x = gdb::ldg((const long long*)&ds[v]);
  return make_int2((int)(x & 0xffffffffll), (int)(x >> 32));
}

template <int WGS>
__global__ __launch_bounds__(WGS) void k_bc_tuned(gdb_kernel_ctx_t* ctx, const int* __restrict__ row,
                                                  const int* __restrict__ col,
                                                  const int* __restrict__ row_t,
                                                  const int* __restrict__ col_t, int2* ds, float* bc,
                                                  const int* sources, int nsrc, int* q, int* tail, int n,
                                                  int* out) {
  DISCOVERY_PROTOCOL(ctx);
  const int tid = p_get_global_id(ctx, &local_ctx);
  const int stride = p_get_global_size(ctx, &local_ctx);
  // A symmetric graph is its own transpose: one walk over a frontier
  // vertex's edges both sums sigma from level l-1 and claims level l+1.
  const bool sym = row == row_t && col == col_t;
  // Unused edge slots read this thread's spare vertex and send it a no-op
  // compare-and-swap (no level is INT_MIN), which keeps the slots
  // branch-free and spreads such accesses over many lines.
  const int spare = n > 0 ? (int)((unsigned)tid % (unsigned)n) : 0;
  int nbar = 0, levels = 0;
  for (int i = tid; i < n; i += stride) {
    bc[i] = 0.f;
    ds[i] = make_int2(-1, 0);
  }
  auto plus = [](float a, float c) { return a + c; };
  int prev_reached = 0;
  for (int si = 0; si < nsrc; ++si) {
    const int s = sources[si];
    // Reset what the previous source touched (its reached set is q[0..)).
    for (int i = tid; i < prev_reached; i += stride) {
      const int v = q[i];
      if ((unsigned)v < (unsigned)n) ds[v] = make_int2(-1, 0);
    }
    if (!discovery_barrier(ctx, &local_ctx)) return;
    ++nbar;
    if (tid == 0) {
      ds[s] = make_int2(0, __float_as_int(1.f));
      q[0] = s;
      gdb::st_rlx(tail, 1);
    }
    if (!discovery_barrier(ctx, &local_ctx)) return;
    ++nbar;
    // Forward: level l occupies q[lb, le).  Expanding level l claims level
    // l+1 vertices (compare-and-swap on the level word); the same phase
    // pulls sigma for level l (l >= 1) from in-edges whose source sits on
    // level l-1.  Level boundaries live in registers: every group reads the
    // same tail after each barrier.
    int lb = 0, le = 1, l = 0;
    for (;;) {
      GDB_FOR_UNIFORM(i, valid, tid, stride, le - lb) {
        int v = -1, b = 0, e = 0;
        if (valid) {
          v = q[lb + i];
          if ((unsigned)v >= (unsigned)n) valid = 0;
        }
        if (valid) {
          b = row[v];
          e = row[v + 1];
        }
        // Walks v's out-edges kChunk slots per lane step (light rows by their
        // own lane, long rows by the whole wave): claims unreached targets
        // for level l+1 and, when `pull_sigma` (a symmetric graph), sums
        // sigma over the targets on level l-1.  Per step the slots' column
        // loads, then their state gathers, then their claims are issued
        // together; one wave push serves the step.
        auto walk = [&](bool pull_sigma) -> float {
          return gdb::wave_walk_n<kChunk>(
              valid != 0, b, e, 0.f,
              [&](const bool (&ok)[kChunk], const int (&kk)[kChunk]) -> float {
                int w[kChunk];
                bool push[kChunk];
                float acc = 0.f;
                bool any = false;
#pragma unroll
                for (int j = 0; j < kChunk; ++j) {
                  w[j] = -1;
                  push[j] = false;
                  any |= ok[j];
                }
                if (any) {  // lanes without a slot this step stay out
                  int old[kChunk];
                  int2 x[kChunk];
                  bool in[kChunk], claim[kChunk];
#pragma unroll
                  for (int j = 0; j < kChunk; ++j) w[j] = gdb::ldg(col + (ok[j] ? kk[j] : 0));
#pragma unroll
                  for (int j = 0; j < kChunk; ++j) {
                    in[j] = ok[j] && (unsigned)w[j] < (unsigned)n;
                    x[j] = ld_state(ds, in[j] ? w[j] : spare);
                  }
                  bool any_claim = false;
#pragma unroll
                  for (int j = 0; j < kChunk; ++j) {
                    claim[j] = in[j] && x[j].x == -1;
                    any_claim |= claim[j];
                    if (pull_sigma && in[j] && x[j].x == l - 1) acc += __int_as_float(x[j].y);
                  }
                  // Most targets are claimed already: only lanes with a claim
                  // send compare-and-swaps (an atomic, even a failing one,
                  // dirties its line).  The results are consumed inside.
                  if (any_claim) {
#pragma unroll
                    for (int j = 0; j < kChunk; ++j)
                      old[j] = gdb::atomic_cas_g(&ds[claim[j] ? w[j] : spare].x, claim[j] ? -1 : INT_MIN,
                                                 claim[j] ? l + 1 : INT_MIN);
#pragma unroll
                    for (int j = 0; j < kChunk; ++j) push[j] = claim[j] && old[j] == -1;
                  }
                }
                gdb::wave_push_n<kChunk>(push, w, q, tail);
                return acc;
              },
              plus, kSerialLimit);
        };
        float sg = 0.f;
        if (sym) {
          sg = walk(l > 0);
        } else {
          int bt = 0, et = 0;
          if (valid && l > 0) {
            bt = row_t[v];
            et = row_t[v + 1];
          }
          sg = gdb::wave_walk_n<kChunk>(
              valid != 0 && l > 0, bt, et, 0.f,
              [&](const bool (&ok)[kChunk], const int (&kk)[kChunk]) -> float {
                if (!gdb::any_slot(ok)) return 0.f;  // no slot: no loads
                int u[kChunk];
                int2 x[kChunk];
#pragma unroll
                for (int j = 0; j < kChunk; ++j) u[j] = gdb::ldg(col_t + (ok[j] ? kk[j] : 0));
#pragma unroll
                for (int j = 0; j < kChunk; ++j)
                  x[j] = ld_state(ds, ok[j] && (unsigned)u[j] < (unsigned)n ? u[j] : spare);
                float acc = 0.f;
#pragma unroll
                for (int j = 0; j < kChunk; ++j)
                  if (ok[j] && (unsigned)u[j] < (unsigned)n && x[j].x == l - 1) acc += __int_as_float(x[j].y);
                return acc;
              },
              plus, kSerialLimit);
          walk(false);
        }
        if (valid && l > 0) ds[v].y = __float_as_int(sg);
      }
      if (!discovery_barrier(ctx, &local_ctx)) return;
      ++nbar;
      const int t = gdb::ld_rlx(tail);
      if (t > n || t < le) {  // each vertex is claimed once: never more than n
        gdb::set_error(ctx, GDB_ERR_QUEUE_OVERFLOW);
        return;
      }
      if (t == le) break;  // level l had no successors: l is the deepest level
      lb = le;
      le = t;
      ++l;
    }
    // q[lb, le) is the deepest level L = l (delta 0: coef = 1 / sigma); the
    // reached set is q[0, le).
    const int reached = le;
    levels += l + 1;
    for (int i = lb + tid; i < le; i += stride) {
      const int v = q[i];
      if ((unsigned)v < (unsigned)n) ds[v].y = __float_as_int(1.f / __int_as_float(ds[v].y));
    }
    if (!discovery_barrier(ctx, &local_ctx)) return;
    ++nbar;
    // Backward: delta[w] = sigma[w] * sum of coef over out-edges into level
    // +1, for w on level L-1 down to 0.  q is ordered by level, so each
    // level's range is found by binary search on the level words.
    for (int ll = l - 1; ll >= 0; --ll) {
      int lo = 0, hi = reached;
      while (lo < hi) {  // first position with level >= ll
        const int mid = (lo + hi) >> 1;
        const int v = q[mid];
        if (((unsigned)v < (unsigned)n ? ds[v].x : -1) < ll) lo = mid + 1; else hi = mid;
      }
      const int qb = lo;
      hi = reached;
      while (lo < hi) {  // first position with level >= ll + 1
        const int mid = (lo + hi) >> 1;
        const int v = q[mid];
        if (((unsigned)v < (unsigned)n ? ds[v].x : -1) < ll + 1) lo = mid + 1; else hi = mid;
      }
      const int qe = lo;
      GDB_FOR_UNIFORM(i, valid, tid, stride, qe - qb) {
        int w = -1, b = 0, e = 0;
        if (valid) {
          w = q[qb + i];
          if ((unsigned)w >= (unsigned)n) valid = 0;
        }
        if (valid) {
          b = row[w];
          e = row[w + 1];
        }
        const float acc = gdb::wave_walk_n<kChunk>(
            valid != 0, b, e, 0.f,
            [&](const bool (&ok)[kChunk], const int (&kk)[kChunk]) -> float {
              if (!gdb::any_slot(ok)) return 0.f;  // no slot: no loads
              int v[kChunk];
              int2 x[kChunk];
#pragma unroll
              for (int j = 0; j < kChunk; ++j) v[j] = gdb::ldg(col + (ok[j] ? kk[j] : 0));
#pragma unroll
              for (int j = 0; j < kChunk; ++j)
                x[j] = ld_state(ds, ok[j] && (unsigned)v[j] < (unsigned)n ? v[j] : spare);
              float a = 0.f;
#pragma unroll
              for (int j = 0; j < kChunk; ++j)
                if (ok[j] && (unsigned)v[j] < (unsigned)n && x[j].x == ll + 1) a += __int_as_float(x[j].y);
              return a;
            },
            plus, kSerialLimit);
        if (valid) {
          const float sw = __int_as_float(ds[w].y);
          const float dw = sw * acc;
          if (w != s) bc[w] += dw;
          ds[w].y = __float_as_int((1.f + dw) / sw);
        }
      }
      if (!discovery_barrier(ctx, &local_ctx)) return;
      ++nbar;
    }
    prev_reached = reached;
  }
  if (tid == 0) {
    out[0] = levels;
    out[1] = nbar;
  }
}

// ---- multi-source BC on a symmetric graph -----------------------------------
// kMS sources share one sweep, which reads the column array once for all of
// them.  Step t's queue range holds the vertices that sit on level t for at
// least one source; walking such a vertex pulls sigma from level t-1 and
// claims level t+1 for exactly those sources.  The backward pass walks the
// steps from the deepest down: a vertex on level t of source s takes
// delta_s = sigma_s * sum of coef_s over neighbours on level t+1 (none on its
// source's deepest level: delta 0), coef = (1 + delta) / sigma.
//
// Per vertex, three arrays:
//   lev[v]  kMS ints, the sources' levels (-1 unreached): read once per queue
//           entry, for the vertex's own source mask;
//   code[v] one word, a nibble per source: bit 3 reached, bits 0-1 the level
//           mod 4;
//   sg[v]   kMS floats (32 bytes): sigma in the forward pass, coef once the
//           vertex's delta is final in the backward pass.
// On a symmetric graph the levels of two adjacent vertices differ by at most
// one for every source, so for a vertex on level t a neighbour's nibble tells
// t-1, t, t+1 and unreached apart.  An edge therefore gathers the 4-byte code
// word, and the 32-byte sg record only when the neighbour sits on the level
// the walk pulls from (t-1 forward, t+1 backward) for one of the walking
// vertex's sources.  A claim of a neighbour for several sources is one
// fetch-or of their nibbles (every claimer in step t ors the same level-t+1
// nibble, and the winners are the sources whose reached bit it set); each
// winner stores the full level in lev.
constexpr int kMS = 8;
constexpr int kNR = kMS / 4;  // int4 per lev / sg record
// Edge slots per lane step: the gathers of a step are in flight together.
// Measured on R-MAT 22, 64 sources (same box): 4 slots 450-465 ms (4 waves
// per SIMD), 6: 395, 8: 355-363, 10: 340-346, 12: 331-340 (238 VGPRs, 2
// waves per SIMD, no spills): gathers in flight per lane beat occupancy.
constexpr int kSlotsM = 12;
constexpr unsigned kNibLo = 0x11111111u;

// Bit 4s of the result is set iff source s's nibble of `code` is reached and
// holds level `lv` mod 4.
__device__ __forceinline__ unsigned nib_on_level(unsigned code, unsigned lv) {
  const unsigned x = (code ^ (lv * kNibLo)) & 0x33333333u;
  return ~(x | (x >> 1)) & (code >> 3) & kNibLo;
}
// Bit s of `mask` moved to bit 4s.
__device__ __forceinline__ unsigned nib_spread(unsigned mask) {
  unsigned r = 0;
#pragma unroll
  for (int s = 0; s < kMS; ++s) r |= ((mask >> s) & 1u) << (4 * s);
  return r;
}

struct FV {  // per-source partial sums
  float v[kMS];
};
__device__ __forceinline__ FV shfl_xor(FV x, int o) {
  FV r;
#pragma unroll
  for (int s = 0; s < kMS; ++s) r.v[s] = __shfl_xor(x.v[s], o, 64);
  return r;
}
__device__ __forceinline__ FV fv_add(FV a, const FV& b) {
#pragma unroll
  for (int s = 0; s < kMS; ++s) a.v[s] += b.v[s];
  return a;
}
__device__ __forceinline__ FV fv_zero() {
  FV r;
#pragma unroll
  for (int s = 0; s < kMS; ++s) r.v[s] = 0.f;
  return r;
}
struct Half {  // kMS ints of a record
  int4 q[kNR];
};
__device__ __forceinline__ int at(const Half& h, int s) {
  const int4& x = h.q[s >> 2];
  const int k = s & 3;
  return k == 0 ? x.x : k == 1 ? x.y : k == 2 ? x.z : x.w;
}
__device__ __forceinline__ void ld_half(const int4* a, size_t v, Half& h) {
#pragma unroll
  for (int r = 0; r < kNR; ++r) h.q[r] = a[v * kNR + r];
}
__device__ __forceinline__ int* rec_int(int4* a, size_t v, int s) { return (int*)&a[v * kNR] + s; }

// Gathers in flight bound this kernel.
template <int WGS>
__global__ __launch_bounds__(WGS) void k_bc_multi(gdb_kernel_ctx_t* ctx, const int* __restrict__ row,
                                                  const int* __restrict__ col, int4* lev, int4* sg,
                                                  unsigned* code, float* bc, const int* sources, int nsrc,
                                                  int* q, int* stamp, int* loff, int* tail, int n, int* out) {
  DISCOVERY_PROTOCOL(ctx);
  const int tid = p_get_global_id(ctx, &local_ctx);
  const int stride = p_get_global_size(ctx, &local_ctx);
  // Unused edge slots read this thread's spare vertex (branch-free slots) and
  // send it no-op atomics there (an or of 0, a max with INT_MIN).
  const int spare = (int)((unsigned)tid % (unsigned)n);
  const int qcap = kMS * n;  // a vertex is queued at most once per source level
  int nbar = 0, levels = 0;
  auto reset = [&](int v) {
#pragma unroll
    for (int r = 0; r < kNR; ++r) {
      lev[(size_t)v * kNR + r] = make_int4(-1, -1, -1, -1);
      sg[(size_t)v * kNR + r] = make_int4(0, 0, 0, 0);
    }
    code[v] = 0u;
    stamp[v] = -1;
  };
  for (int i = tid; i < n; i += stride) {
    bc[i] = 0.f;
    reset(i);
  }
  int prev_reached = 0;
  for (int s0 = 0; s0 < nsrc; s0 += kMS) {
    const int ns = nsrc - s0 < kMS ? nsrc - s0 : kMS;
    for (int i = tid; i < prev_reached; i += stride) {  // what the previous sweep touched
      const int v = q[i];
      if ((unsigned)v < (unsigned)n) reset(v);
    }
    if (!discovery_barrier(ctx, &local_ctx)) return;
    ++nbar;
    if (tid == 0) {
      int cnt = 0;
      for (int j = 0; j < ns; ++j) {
        const int s = sources[s0 + j];
        *rec_int(lev, s, j) = 0;
        *rec_int(sg, s, j) = __float_as_int(1.f);
        code[s] |= 8u << (4 * j);  // reached, level 0
        if (stamp[s] != 0) {
          stamp[s] = 0;
          q[cnt++] = s;
        }
      }
      loff[0] = 0;
      loff[1] = cnt;
      gdb::st_rlx(tail, cnt);
    }
    if (!discovery_barrier(ctx, &local_ctx)) return;
    ++nbar;
    const unsigned all = ns >= 32 ? ~0u : (1u << ns) - 1u;
    // Forward.
    int lb = 0, le = gdb::ld_rlx(tail), t = 0;
    for (;;) {
      const unsigned par = (unsigned)(t - 1) & 3u;                    // parents' level mod 4
      const unsigned nxt_nib = 8u | ((unsigned)(t + 1) & 3u);         // a claimed nibble
      GDB_FOR_UNIFORM(i, valid, tid, stride, le - lb) {
        int v = -1, b = 0, e = 0;
        unsigned mask = 0;
        if (valid) {
          v = q[lb + i];
          if ((unsigned)v >= (unsigned)n) valid = 0;
        }
        if (valid) {
          Half lv;
          ld_half(lev, v, lv);
#pragma unroll
          for (int s = 0; s < kMS; ++s) mask |= (at(lv, s) == t ? 1u : 0u) << s;
          mask &= all;
          b = row[v];
          e = row[v + 1];
        }
        // On a long row every lane works for the row's owner: `om` is the
        // owner's source mask.
        const FV sgm = gdb::wave_walk_np<kSlotsM>(
            valid && mask != 0, b, e, (int)mask, fv_zero(),
            [&](const bool (&ok)[kSlotsM], const int (&kk)[kSlotsM], int om) -> FV {
              int w[kSlotsM];
              bool push[kSlotsM];
              FV acc = fv_zero();
              bool any = false;
#pragma unroll
              for (int j = 0; j < kSlotsM; ++j) {
                w[j] = -1;
                push[j] = false;
                any |= ok[j];
              }
              if (any) {  // lanes without a slot this step stay out
                const unsigned omn = nib_spread((unsigned)om);
                unsigned c[kSlotsM], pm[kSlotsM], cl[kSlotsM];
#pragma unroll
                for (int j = 0; j < kSlotsM; ++j) w[j] = gdb::ldg(col + (ok[j] ? kk[j] : 0));
#pragma unroll
                for (int j = 0; j < kSlotsM; ++j) {
                  const bool in = ok[j] && (unsigned)w[j] < (unsigned)n;
                  c[j] = gdb::ldg(code + (in ? w[j] : spare));
                  if (!in) w[j] = spare;
                  const unsigned sel = in ? omn : 0u;
                  pm[j] = nib_on_level(c[j], par) & sel;  // neighbour on level t-1
                  cl[j] = (~c[j] >> 3) & kNibLo & sel;    // neighbour unreached
                }
                // sigma of the neighbours on level t-1, where some source has one.
                Half sw[kSlotsM];
#pragma unroll
                for (int j = 0; j < kSlotsM; ++j)
                  if (pm[j]) ld_half(sg, w[j], sw[j]);
#pragma unroll
                for (int j = 0; j < kSlotsM; ++j)
                  if (pm[j])
#pragma unroll
                    for (int s = 0; s < kMS; ++s)
                      if ((pm[j] >> (4 * s)) & 1u) acc.v[s] += __int_as_float(at(sw[j], s));
                bool any_claim = false;
#pragma unroll
                for (int j = 0; j < kSlotsM; ++j) any_claim |= cl[j] != 0u;
                // Only lanes with an unreached target send claims: one fetch-or
                // of the claimed nibbles (winners are the sources whose reached
                // bit it set), then each winner stores the full level.
                if (any_claim) {
                  unsigned oldc[kSlotsM];
#pragma unroll
                  for (int j = 0; j < kSlotsM; ++j)
                    oldc[j] = (unsigned)gdb::atomic_or_g((int*)&code[cl[j] ? w[j] : spare],
                                                         (int)(cl[j] * nxt_nib));
                  bool won[kSlotsM];
                  bool any_won = false;
#pragma unroll
                  for (int j = 0; j < kSlotsM; ++j) {
                    const unsigned got = cl[j] & ~(oldc[j] >> 3);
                    won[j] = got != 0u;
                    any_won |= won[j];
#pragma unroll
                    for (int s = 0; s < kMS; ++s)
                      if ((got >> (4 * s)) & 1u) *rec_int(lev, w[j], s) = t + 1;
                  }
                  // A claimed vertex joins step t+1's queue once.
                  if (any_won) {
                    int sp[kSlotsM];
#pragma unroll
                    for (int j = 0; j < kSlotsM; ++j)
                      sp[j] = gdb::atomic_max_g(&stamp[won[j] ? w[j] : spare], won[j] ? t + 1 : INT_MIN);
#pragma unroll
                    for (int j = 0; j < kSlotsM; ++j) push[j] = won[j] && sp[j] < t + 1;
                  }
                }
              }
              gdb::wave_push_n<kSlotsM>(push, w, q, tail);
              return acc;
            },
            [](FV a, FV c) { return fv_add(a, c); }, kSerialLimit);
        if (valid && t > 0) {
#pragma unroll
          for (int s = 0; s < kMS; ++s)
            if ((mask >> s) & 1u) *rec_int(sg, v, s) = __float_as_int(sgm.v[s]);
        }
      }
      if (!discovery_barrier(ctx, &local_ctx)) return;
      ++nbar;
      const int tn = gdb::ld_rlx(tail);
      if (tn > qcap || tn < le) {  // stamps queue a vertex once per step
        gdb::set_error(ctx, GDB_ERR_QUEUE_OVERFLOW);
        return;
      }
      if (tn == le) break;  // step t claimed nothing: t is the deepest step
      lb = le;
      le = tn;
      ++t;
      if (tid == 0) loff[t + 1] = le;
    }
    const int reached = le;
    levels += t + 1;
    // Backward, deepest step first (loff is visible after the last barrier).
    for (int tt = t; tt >= 0; --tt) {
      const int qb = tt == t ? lb : loff[tt];
      const int qe = tt == t ? le : loff[tt + 1];
      GDB_FOR_UNIFORM(i, valid, tid, stride, qe - qb) {
        int v = -1, b = 0, e = 0;
        unsigned mask = 0;
        Half sv;
#pragma unroll
        for (int r = 0; r < kNR; ++r) sv.q[r] = make_int4(0, 0, 0, 0);
        if (valid) {
          v = q[qb + i];
          if ((unsigned)v >= (unsigned)n) valid = 0;
        }
        if (valid) {
          Half lv;
          ld_half(lev, v, lv);
          ld_half(sg, v, sv);
#pragma unroll
          for (int s = 0; s < kMS; ++s) mask |= (at(lv, s) == tt ? 1u : 0u) << s;
          mask &= all;
          b = row[v];
          e = row[v + 1];
        }
        const unsigned chi = (unsigned)(tt + 1) & 3u;  // children's level mod 4
        const FV acc = gdb::wave_walk_np<kSlotsM>(
            valid && mask != 0, b, e, (int)mask, fv_zero(),
            [&](const bool (&ok)[kSlotsM], const int (&kk)[kSlotsM], int om) -> FV {
              FV a = fv_zero();
              if (!gdb::any_slot(ok)) return a;  // no slot: no loads
              const unsigned omn = nib_spread((unsigned)om);
              int w[kSlotsM];
              unsigned cm[kSlotsM];
              Half sw[kSlotsM];
#pragma unroll
              for (int j = 0; j < kSlotsM; ++j) w[j] = gdb::ldg(col + (ok[j] ? kk[j] : 0));
#pragma unroll
              for (int j = 0; j < kSlotsM; ++j) {
                const bool in = ok[j] && (unsigned)w[j] < (unsigned)n;
                cm[j] = nib_on_level(gdb::ldg(code + (in ? w[j] : spare)), chi) & (in ? omn : 0u);
              }
              // coef of the neighbours on level tt+1, where some source has one.
#pragma unroll
              for (int j = 0; j < kSlotsM; ++j)
                if (cm[j]) ld_half(sg, w[j], sw[j]);
#pragma unroll
              for (int j = 0; j < kSlotsM; ++j)
                if (cm[j])
#pragma unroll
                  for (int s = 0; s < kMS; ++s)
                    if ((cm[j] >> (4 * s)) & 1u) a.v[s] += __int_as_float(at(sw[j], s));
              return a;
            },
            [](FV x, FV c) { return fv_add(x, c); }, kSerialLimit);
        if (valid && mask) {
          float add = 0.f;
#pragma unroll
          for (int s = 0; s < kMS; ++s)
            if ((mask >> s) & 1u) {
              const float sw = __int_as_float(at(sv, s));
              const float dw = sw * acc.v[s];
              if (v != sources[s0 + s]) add += dw;
              *rec_int(sg, v, s) = __float_as_int((1.f + dw) / sw);
            }
          bc[v] += add;
        }
      }
      if (!discovery_barrier(ctx, &local_ctx)) return;
      ++nbar;
    }
    prev_reached = reached;
  }
  if (tid == 0) {
    out[0] = levels;
    out[1] = nbar;
  }
}

__global__ void k_bc_clean(int* d, float* rho, float* delta, int s, int n, int* cont) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {  // clean_1d_array
    delta[i] = 0.f;
    rho[i] = i == s ? 1.f : 0.f;
    d[i] = i == s ? 0 : -1;
  }
  if (i == 0) *cont = 0;
}
__global__ void k_bc_zero(float* bc, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) bc[i] = 0.f;
}
__global__ void k_bc_bfs(const int* __restrict__ row, const int* __restrict__ col, int* d, float* rho,
                         int* cont, int n, int dist) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid < n && d[tid] == dist) {
    const float r = rho[tid];
    for (int e = row[tid]; e < row[tid + 1]; ++e) {
      const int w = col[e];
      if (d[w] < 0) {
        *cont = 1;
        d[w] = dist + 1;
      }
      if (d[w] == dist + 1) unsafeAtomicAdd(&rho[w], r);
    }
  }
}
__global__ void k_bc_back(const int* __restrict__ row_t, const int* __restrict__ col_t, const int* d,
                          const float* rho, float* delta, int n, int dist, int s, float* bc) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid < n && d[tid] == dist - 1) {
    const float r = rho[tid], di = delta[tid];
    for (int e = row_t[tid]; e < row_t[tid + 1]; ++e) {
      const int w = col_t[e];
      if (d[w] == dist - 2) unsafeAtomicAdd(&delta[w], rho[w] / r * (1.f + di));
    }
    if (tid != s) bc[tid] = bc[tid] + di;
  }
}

}  // namespace

using namespace gdbrt;

extern "C" int gdb_bc(gdb_ctx* ctx, int n, int m, const int* row, const int* col, const int* row_t,
                      const int* col_t, const int* sources, int nsrc, float* bc, const gdb_opts* opts,
                      gdb_stats* stats) {
  if (!ctx || n < 0 || m < 0 || nsrc < 0 || (n > 0 && (!row || !row_t || !bc)) ||
      (m > 0 && (!col || !col_t)) || (nsrc > 0 && !sources))
    return GDB_E_INVALID;
  gdb_opts o{};
  if (opts) o = *opts;
  if (o.variant == GDB_VARIANT_TUNED && o.schedule == GDB_SCHED_PERSISTENT) gdbrt::apply_tuning("bc", n, &o);
  const int wgs = default_wgs(&o);
  gdb_stats st{};
  if (n == 0) {
    if (stats) *stats = st;
    return GDB_SUCCESS;
  }
  // Host copy of the sources for validation and the relaunch loop.
  std::vector<int> hsrc((size_t)nsrc);
  if (nsrc) GDB_HIP(hipMemcpy(hsrc.data(), sources, sizeof(int) * nsrc, hipMemcpyDeviceToHost));
  for (int s : hsrc)
    if (s < 0 || s >= n) return GDB_E_INVALID;
  // The tuned persistent schedule sweeps kMS sources at once on a symmetric
  // graph (the CSR passed as its own transpose); a directed graph takes the
  // single-source kernel with separate in- and out-edge walks.
  const bool multi = o.variant == GDB_VARIANT_TUNED && o.schedule != GDB_SCHED_RELAUNCH && row == row_t &&
                     col == col_t && (long long)kMS * n + 2 <= INT_MAX;  // queue positions are ints
  const size_t qcap = (size_t)kMS * n;
  const size_t bytes = carve_bytes<int>(n) * 2 + carve_bytes<float>(n) * 2 + carve_bytes<int>(32) +
                       carve_bytes<int2>(n) +
                       (multi ? carve_bytes<int4>((size_t)2 * kNR * n) + carve_bytes<int>(qcap) +
                                    carve_bytes<int>(qcap + 2)
                              : 0);
  int rc = ensure_workspace(ctx, bytes);
  if (rc) return rc;
  Carve cv(ctx->ws);
  int* d = cv.take<int>(n);
  int* q = cv.take<int>(n);
  float* rho = cv.take<float>(n);
  float* delta = cv.take<float>(n);
  int* flags = cv.take<int>(32);
  int2* ds = cv.take<int2>(n);
  int4* mst = multi ? cv.take<int4>((size_t)2 * kNR * n) : nullptr;
  int* mq = multi ? cv.take<int>(qcap) : nullptr;
  int* mloff = multi ? cv.take<int>(qcap + 2) : nullptr;
  rc = fill_i32(ctx, flags, 0, 32);
  if (rc) return rc;
  rc = prepare_launch(ctx);
  if (rc) return rc;
  const int grid = (n + wgs - 1) / wgs;

  if (o.schedule == GDB_SCHED_RELAUNCH) {
    rc = timer_start(ctx);
    if (rc) return rc;
    hipLaunchKernelGGL(k_bc_zero, dim3(grid), dim3(wgs), 0, ctx->stream, bc, n);
    st.launches += 1;
    for (int s : hsrc) {
      hipLaunchKernelGGL(k_bc_clean, dim3(grid), dim3(wgs), 0, ctx->stream, d, rho, delta, s, n, flags);
      st.launches += 1;
      int dist = 0;
      for (;;) {  // bc.cpp:363-390: launch a level, read `cont`
        rc = fill_i32(ctx, flags, 0, 1);
        if (rc) return rc;
        hipLaunchKernelGGL(k_bc_bfs, dim3(grid), dim3(wgs), 0, ctx->stream, row, col, d, rho, flags, n,
                           dist);
        st.launches += 1;
        GDB_HIP(hipMemcpyAsync(ctx->h_flag, flags, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        GDB_HIP(hipStreamSynchronize(ctx->stream));
        ++dist;
        if (ctx->h_flag[0] == 0) break;
      }
      st.rounds += dist;
      for (; dist > 0; --dist) {  // bc.cpp:395-416
        hipLaunchKernelGGL(k_bc_back, dim3(grid), dim3(wgs), 0, ctx->stream, row_t, col_t, d, rho, delta,
                           n, dist, s, bc);
        st.launches += 1;
      }
      GDB_HIP(hipGetLastError());
    }
    rc = timer_stop(ctx, &st.kernel_ms);
    if (rc) return rc;
    st.participating = grid;
  } else {
    const bool tuned = o.variant == GDB_VARIANT_TUNED;
    int* out = flags + 8;
    rc = gdb_dispatch_wgs(wgs, [&](auto W) -> int {
      constexpr int WG = decltype(W)::value;
      const void* kern = multi   ? (const void*)k_bc_multi<WG>
                         : tuned ? (const void*)k_bc_tuned<WG>
                                 : (const void*)k_bc_plain<WG>;
      int groups = 0;
      int r2 = launch_groups(ctx, kern, &o, n, &groups);
      if (r2) return r2;
      r2 = timer_start(ctx);
      if (r2) return r2;
      if (multi)  // mst holds lev then sg; d the per-vertex queue stamps; q the level codes
        hipLaunchKernelGGL(k_bc_multi<WG>, dim3(groups), dim3(WG), 0, ctx->stream, ctx->d, row, col, mst,
                           mst + (size_t)kNR * n, (unsigned*)q, bc, sources, nsrc, mq, d, mloff, flags + 4, n,
                           out);
      else if (tuned)
        hipLaunchKernelGGL(k_bc_tuned<WG>, dim3(groups), dim3(WG), 0, ctx->stream, ctx->d, row, col,
                           row_t, col_t, ds, bc, sources, nsrc, q, flags + 4, n, out);
      else
        hipLaunchKernelGGL(k_bc_plain<WG>, dim3(groups), dim3(WG), 0, ctx->stream, ctx->d, row, col,
                           row_t, col_t, d, rho, delta, bc, sources, nsrc, flags, n, out);
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
  st.work_edges = (long long)m * nsrc;
  st.work_items = (long long)n * nsrc;
  if (stats) *stats = st;
  return rc;
}
