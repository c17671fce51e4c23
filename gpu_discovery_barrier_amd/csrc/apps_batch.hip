// Batched engine: many (graph, algorithm) tasks in one persistent launch.
//
// After discovery the P participating groups form floor(P / G) teams of G
// consecutive ids.  A team's rank-0 leader takes the next task index from a
// shared counter; the team solves the task with its own sub-barrier (a
// sense-reversing counter over its G groups, on lines of its own) and takes
// the next one.  A team of one group needs no inter-group barrier at all:
// it synchronises with workgroup barriers, which suits the thousands of
// small graphs of a batch; large graphs get large teams, and independent
// solves of one large graph (BFS and SSSP) run side by side.
//
// Algorithms per task (outputs identical to the single-task solvers):
//   BFS    worklist levels (lonestar bfs_port_kernel.cl semantics)
//   SSSP   worklist label-correcting relaxation (sssp_common.cl:58-81); with
//          delta > 0 the worklist is split near / far: relaxations below the
//          current threshold feed the next iteration, the rest wait in a far
//          pile that is re-split when the near queue runs dry, threshold +=
//          delta.  Same fixpoint (exact distances), far less re-relaxation.
//   COLOR  Pannotia max rule, one barrier per round (see apps_color.hip).
#include "graph_common.hpp"
#include "runtime.hpp"

namespace {

constexpr unsigned kInf = 1000000000u;  // bfs_common.cl:35
constexpr int kSerialLimit = 16;
constexpr int kChunk = 4;                  // edges a lane relaxes per step
constexpr int kLine = GDB_LINE_BYTES / 4;  // ints per 128 B line

// Team state, one 128 B line per word.
enum TeamWord { TW_COUNT = 0, TW_GEN, TW_TASK, TW_TAIL0, TW_TAIL1, TW_TAIL2, TW_FAR0, TW_FAR1, TW_FAR2,
                TW_STOP0, TW_STOP1, TW_STOP2, TW_WORDS };

struct DevTask {
  int algo, n, m, source, delta, hops;
  const int* row;
  const int* col;
  const int* w;
  const float* nv;
  void* out;
  int* scratch;
};

constexpr int kLocalQ = 2048;  // LDS queue capacity per buffer (ints)

struct Team {
  gdb_kernel_ctx_t* ctx;
  gdb_local_ctx* l;
  int* ts;        // this team's state lines
  int* gen_cache; // LDS
  int G, rank;
  int (*lq)[kLocalQ];  // LDS: two local queue buffers
  int* lcnt;           // LDS: three rotating local queue counters
  __device__ int word_off(int w) const { return w * kLine; }
  __device__ int* word(int w) const { return ts + w * kLine; }
  __device__ int tid() const { return rank * (int)blockDim.x + (int)threadIdx.x; }
  __device__ int stride() const { return G * (int)blockDim.x; }
};

__device__ inline int team_barrier_payload(Team& T, const int* src, int* lds_val);

// Plain team barrier (payload unused).
__device__ inline bool team_barrier(Team& T) {
  __shared__ int dummy;
  team_barrier_payload(T, T.word(TW_TASK), &dummy);
  return T.l->aborted == 0;
}

// Team barrier that also hands every member one int: the value at *src read
// by the last arriver, after every member's writes are visible to it (the
// size of the queue the iteration just filled).  The generation and the
// value travel in one 64-bit word, so a waiter gets both with one load.
__device__ inline int team_barrier_payload(Team& T, const int* src, int* lds_val) {
  gdb::drain_stores();
  __syncthreads();
  if (threadIdx.x == 0) {
    if (T.G == 1) {
      *lds_val = gdb::ld_rlx(src);
    } else {
      unsigned long long* word = (unsigned long long*)T.word(TW_GEN);
      int g = *T.gen_cache;
      if (g == GDB_GEN_UNKNOWN)
        g = (int)(unsigned)__hip_atomic_load(word, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
      const int old = gdb::fadd(T.word(TW_COUNT), 1);
      int val = 0;
      if (old == T.G - 1) {
        gdb::st_rlx(T.word(TW_COUNT), 0);
        val = gdb::ld_rlx(src);
        const unsigned long long w = ((unsigned long long)(unsigned)val << 32) | (unsigned)(g + 1);
        __hip_atomic_store(word, w, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        unsigned long long w = 0;
        const bool ok = gdb::wait_while(
            T.ctx,
            [&] {
              w = __hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              return (int)(unsigned)w == g;
            },
            GDB_ERR_BARRIER_TIMEOUT);
        gdb::fence_acq();
        if (!ok) T.l->aborted = 1;
        val = (int)(w >> 32);
      }
      *T.gen_cache = g + 1;
      *lds_val = val;
    }
  }
  __syncthreads();
  return *lds_val;
}

template <bool WEIGHTED>
__device__ __forceinline__ unsigned relax_target(const int* __restrict__ col, const int* __restrict__ w,
                                                 unsigned* dist, int k, unsigned dv, int n,
                                                 int* dst_out) {
  // Returns the new distance written to dst (kInf when nothing improved).
  const int dst = col[k];
  *dst_out = dst;
  if ((unsigned)dst >= (unsigned)n) return kInf;
  if (!WEIGHTED) {
    if (dist[dst] != kInf) return kInf;
    return atomicCAS(&dist[dst], kInf, dv) == kInf ? dv : kInf;
  }
  const unsigned wt = (unsigned)w[k];
  if (wt >= kInf) return kInf;
  const unsigned alt = dv + wt;
  if (alt >= dist[dst]) return kInf;
  const unsigned old = atomicMin(&dist[dst], alt);
  return alt < old ? alt : kInf;
}

// ---- BFS / SSSP on a team --------------------------------------------------
// scratch: stamp[n], fstamp[n], q0[n], q1[n], f0[n], f1[n], f2[n]
// Adds a per-thread count into *dst with one atomic per wave.
__device__ __forceinline__ void wave_add(unsigned long long* dst, long long v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if (gdb::lane_id() == 0 && v) atomicAdd(dst, (unsigned long long)v);
}

// Worklist BFS / SSSP on a team.
//   WEIGHTED = false : level-synchronous BFS, vertices claimed by CAS with
//                      dist = iteration (the reference's processedge2).
//   WEIGHTED = true  : label-correcting relaxation with a 64-bit
//                      (distance << 32 | mark) word per vertex; UNIT = true
//                      gives BFS levels (unit weights).
// With k.hops = H > 0 a group, after its share of the iteration's queue,
// keeps relaxing the vertices it improved through an LDS queue for up to H
// local steps (workgroup barriers only), so one team barrier covers several
// hops; what is left spills to the team's queue.  Any relaxation order
// reaches the same fixpoint (exact levels / distances).
template <bool WEIGHTED, bool UNIT>
__device__ int team_worklist(Team& T, const DevTask& k, unsigned long long* work) {
  // scratch: BFS (CAS)  q0[n], q1[n]
  //          otherwise  d64[n] (8 B each), fstamp[n], gstamp[n], q0[n], q1[n], far0..2[n]
  // Marks in the low word of d64 (level-synchronous iterations only):
  // 2*it for a near push of iteration it, 2*it+1 for a far push; on equal
  // distances the older mark stays, so "already queued in this iteration"
  // is one compare.  Iterations with local hops leave the mark at 0 and
  // dedupe team-queue pushes with gstamp instead, because a vertex pushed
  // to an LDS queue may be relaxed again before the iteration ends.
  __shared__ int lds_n;
  int (*lq)[kLocalQ] = T.lq;
  int* lcnt = T.lcnt;
  const int n = k.n;
  unsigned* dist = (unsigned*)k.out;
  unsigned long long* d64 = (unsigned long long*)k.scratch;
  int* base = WEIGHTED ? k.scratch + 2 * (size_t)n : k.scratch;
  int* fstamp = base;
  int* gstamp = base + n;
  int* q[2] = {WEIGHTED ? base + 2 * (size_t)n : base, WEIGHTED ? base + 3 * (size_t)n : base + n};
  int* far[3] = {base + 4 * (size_t)n, base + 5 * (size_t)n, base + 6 * (size_t)n};
  const int tid = T.tid(), stride = T.stride();
  const bool near_far = WEIGHTED && !UNIT && k.delta > 0;
  const int hops = WEIGHTED ? k.hops : 0;
  for (int v = tid; v < n; v += stride) {
    if (WEIGHTED) {
      d64[v] = (unsigned long long)(v == k.source ? 0u : kInf) << 32;
      fstamp[v] = -1;
      gstamp[v] = 0;
    } else {
      dist[v] = v == k.source ? 0u : kInf;
    }
  }
  if (tid == 0) {
    q[0][0] = k.source;
    gdb::st_rlx(T.word(TW_TAIL1), 1);
    gdb::st_rlx(T.word(TW_TAIL2), 0);
    gdb::st_rlx(T.word(TW_TAIL0), 0);
    gdb::st_rlx(T.word(TW_FAR0), 0);
    gdb::st_rlx(T.word(TW_FAR1), 0);
    gdb::st_rlx(T.word(TW_FAR2), 0);
  }
  if (!team_barrier(T)) return -1;
  unsigned thresh = near_far ? (unsigned)k.delta : kInf;
  int gen = 0;  // far pile generation being filled
  int it = 1;
  int nin = 1;
  long long items = 0, edges = 0;
  for (;;) {
    const int* in = q[(it - 1) & 1];
    int* outq = q[it & 1];
    int* out_tail = T.word(TW_TAIL0 + (it + 1) % 3);
    if (tid == 0) gdb::st_rlx(T.word(TW_TAIL0 + (it + 2) % 3), 0);
    const unsigned near_mark = 2u * (unsigned)it, far_mark = near_mark + 1u;
    if (nin > 0) {
      if (nin > n) {
        gdb::set_error(T.ctx, GDB_ERR_QUEUE_OVERFLOW);
        return -1;
      }
      int* fpile = far[gen % 3];
      int* ftail = T.word(TW_FAR0 + gen % 3);
      if (hops > 0 && threadIdx.x < 3) lcnt[threadIdx.x] = 0;
      if (hops > 0) __syncthreads();
      // Relaxes edge kk from a vertex at distance src_d.  Returns the push
      // class: 0 none, 1 team queue, 2 far pile, 3 local (LDS) queue.
      auto relax = [&](int kk, unsigned src_d, int* dst_out) -> int {
        const int dst = k.col[kk];
        *dst_out = dst;
        if ((unsigned)dst >= (unsigned)n) return 0;
        if (!WEIGHTED) return atomicCAS(&dist[dst], kInf, src_d) == kInf ? 1 : 0;
        const unsigned wt = UNIT ? 1u : (unsigned)k.w[kk];
        if (wt >= kInf) return 0;
        const unsigned cand = src_d + wt;
        const bool nearp = cand < thresh;
        if (hops > 0) {
          const unsigned long long old = atomicMin(&d64[dst], (unsigned long long)cand << 32);
          if (cand >= (unsigned)(old >> 32)) return 0;
          if (nearp) return 3;
          return atomicExch(&fstamp[dst], gen) != gen ? 2 : 0;
        }
        const unsigned long long nw = ((unsigned long long)cand << 32) | (nearp ? near_mark : far_mark);
        const unsigned long long old = atomicMin(&d64[dst], nw);
        if (cand >= (unsigned)(old >> 32)) return 0;
        if (nearp) return (unsigned)old == near_mark ? 0 : 1;
        return atomicExch(&fstamp[dst], gen) != gen ? 2 : 0;
      };
      // Pushes the results of one relaxation step of this wave.  Class 3
      // goes to LDS buffer `lb` when there is room, else to the team queue.
      auto push_n = [&](const int (&dst)[kChunk], const int (&cls)[kChunk], int lb, int* lc) {
        bool to_team[kChunk], to_far[kChunk];
#pragma unroll
        for (int j = 0; j < kChunk; ++j) {
          bool team_push = cls[j] == 1;
          if (cls[j] == 3) {
            const int pos = atomicAdd(lc, 1);
            if (pos < kLocalQ) lq[lb][pos] = dst[j];
            else team_push = atomicExch(&gstamp[dst[j]], it) != it;
          }
          to_team[j] = team_push;
          to_far[j] = cls[j] == 2;
        }
        gdb::wave_push_n<kChunk>(to_team, dst, outq, out_tail);
        if (near_far) gdb::wave_push_n<kChunk>(to_far, dst, fpile, ftail);
      };
      // Relaxes the out-edges of items[start, count) with stride `step`.
      auto process = [&](const int* src_items, int count, int start, int step, int lb, int* lc) {
        GDB_FOR_UNIFORM(i, valid, start, step, count) {
          int v = -1, b = 0, e = 0;
          unsigned dv = 0;
          if (valid) {
            v = src_items[i];
            if ((unsigned)v >= (unsigned)n) valid = 0;
          }
          if (valid) {
            b = k.row[v];
            e = k.row[v + 1];
            dv = WEIGHTED ? (unsigned)(d64[v] >> 32) : (unsigned)it;
            ++items;
            edges += e - b;
          }
          // Light rows: each lane relaxes its own edges four at a time, the
          // four atomics of a chunk issued back to back.
          const bool heavy = valid && (e - b) > kSerialLimit;
          const int len = (valid && !heavy) ? (e - b) : 0;
          const int maxlen = gdb::wave_max(len);
          for (int c = 0; c < maxlen; c += kChunk) {
            int dst[kChunk], cls[kChunk];
#pragma unroll
            for (int j = 0; j < kChunk; ++j) {
              dst[j] = -1;
              cls[j] = (c + j < len) ? relax(b + c + j, dv, &dst[j]) : 0;
            }
            push_n(dst, cls, lb, lc);
          }
          // Heavy rows: the whole wave walks one row, 64 edges per step.
          gdb::wave_coop_foreach(
              heavy, b, e, dv,
              [&](int kk, bool ok, unsigned src_d) {
                int dst[kChunk] = {-1, -1, -1, -1}, cls[kChunk] = {0, 0, 0, 0};
                cls[0] = ok ? relax(kk, src_d, &dst[0]) : 0;
                push_n(dst, cls, lb, lc);
              },
              -1);
        }
      };
      // The team's share of the iteration; improved vertices go to LDS
      // buffer 1 (local step 1 reads it).
      process(in, nin, tid, stride, 1, &lcnt[1]);
      // Local steps: this group alone relaxes what it just improved.
      int s = 1;
      for (; s <= hops; ++s) {
        gdb::drain_stores();
        __syncthreads();
        const int ln = min(lcnt[s % 3], kLocalQ);
        if (ln == 0) break;
        if (threadIdx.x == 0) lcnt[(s + 2) % 3] = 0;
        process(lq[s & 1], ln, (int)threadIdx.x, (int)blockDim.x, (s + 1) & 1, &lcnt[(s + 1) % 3]);
      }
      if (hops > 0 && s > hops) {
        // Budget spent: flush the last local buffer to the team queue.
        __syncthreads();
        const int ln = min(lcnt[s % 3], kLocalQ);
        GDB_FOR_UNIFORM(i, valid, (int)threadIdx.x, (int)blockDim.x, ln) {
          int v = 0;
          bool p = false;
          if (valid) {
            v = lq[s & 1][i];
            p = atomicExch(&gstamp[v], it) != it;
          }
          gdb::wave_push(p, v, outq, out_tail);
        }
      }
    } else {
      // Near queue empty: split the far pile of generation `gen` at the next
      // threshold into the near queue and generation gen+1.
      const int fsz = gdb::ld_rlx(T.word(TW_FAR0 + gen % 3));
      if (!near_far || fsz == 0) break;
      if (fsz > n) {
        gdb::set_error(T.ctx, GDB_ERR_QUEUE_OVERFLOW);
        return -1;
      }
      thresh += (unsigned)k.delta;
      const int* src_pile = far[gen % 3];
      int* next_pile = far[(gen + 1) % 3];
      int* next_tail = T.word(TW_FAR0 + (gen + 1) % 3);
      if (tid == 0) gdb::st_rlx(T.word(TW_FAR0 + (gen + 2) % 3), 0);
      GDB_FOR_UNIFORM(i, valid, tid, stride, fsz) {
        int v = -1;
        bool to_near = false, to_far = false;
        if (valid) {
          v = src_pile[i];  // each vertex at most once per generation
          if ((unsigned)v < (unsigned)n) {
            if ((unsigned)(d64[v] >> 32) < thresh) to_near = atomicExch(&gstamp[v], it) != it;
            else to_far = atomicExch(&fstamp[v], gen + 1) != gen + 1;
          }
        }
        gdb::wave_push(to_near, v, outq, out_tail);
        gdb::wave_push(to_far, v, next_pile, next_tail);
      }
      ++gen;
    }
    nin = team_barrier_payload(T, out_tail, &lds_n);
    if (T.l->aborted) return -1;
    ++it;
  }
  if (WEIGHTED)
    for (int v = tid; v < n; v += stride) dist[v] = (unsigned)(d64[v] >> 32);
  wave_add(&work[0], items);
  wave_add(&work[1], edges);
  return it - 1;
}

// ---- colour on a team (one barrier per round, as k_color_tuned) ------------
// scratch: maxd[2n] (floats)
__device__ int team_color(Team& T, const DevTask& k) {
  const int n = k.n;
  int* color = (int*)k.out;
  float* maxd = (float*)k.scratch;
  const float* nv = k.nv;
  const int tid = T.tid(), stride = T.stride();
  for (int v = tid; v < n; v += stride) color[v] = -1;
  if (tid == 0) {
    gdb::st_rlx(T.word(TW_STOP0), 0);
    gdb::st_rlx(T.word(TW_STOP1), 0);
    gdb::st_rlx(T.word(TW_STOP2), 0);
  }
  if (!team_barrier(T)) return -1;
  int r = 1;
  for (;;) {
    const float* md_prev = maxd + (size_t)((r - 1) & 1) * n;
    float* md_cur = maxd + (size_t)(r & 1) * n;
    const bool have_prev = r > 1;
    if (tid == 0) gdb::st_rlx(T.word(TW_STOP0 + (r + 1) % 3), 0);
    bool st = false;
    GDB_FOR_UNIFORM(v, valid, tid, stride, n) {
      bool active = false;
      int b = 0, e = 0;
      if (valid && color[v] == -1) {
        if (have_prev && nv[v] > md_prev[v]) {
          color[v] = r - 1;
        } else {
          active = true;
          b = k.row[v];
          e = k.row[v + 1];
        }
      }
      const bool deg1 = (e - b) == 1;
      auto edge = [&](int kk) -> gdb::MaxFound {
        const int u = k.col[kk];
        const bool unc = color[u] == -1 && !(have_prev && nv[u] > md_prev[u]);
        return unc ? gdb::MaxFound{nv[u], 1} : gdb::MaxFound{-1.f, 0};
      };
      auto red = [](gdb::MaxFound a, gdb::MaxFound c) {
        return gdb::MaxFound{a.mx > c.mx ? a.mx : c.mx, a.found | c.found};
      };
      const gdb::MaxFound res =
          gdb::wave_coop_reduce(active && !deg1, b, e, gdb::MaxFound{-1.f, 0}, edge, red, kSerialLimit);
      if (active) {
        md_cur[v] = deg1 ? -1.f : res.mx;
        if (!deg1 && res.found) st = true;
      }
    }
    if (gdb::wave_any(st) && gdb::lane_id() == 0) gdb::st_rlx(T.word(TW_STOP0 + r % 3), 1);
    if (!team_barrier(T)) return -1;
    if (gdb::ld_rlx(T.word(TW_STOP0 + r % 3)) == 0) break;
    if (r > n) {
      gdb::set_error(T.ctx, GDB_ERR_QUEUE_OVERFLOW);
      return -1;
    }
    ++r;
  }
  const float* md = maxd + (size_t)(r & 1) * n;
  for (int v = tid; v < n; v += stride)
    if (color[v] == -1 && nv[v] > md[v]) color[v] = r;
  return r;
}

template <int WGS>
__global__ __launch_bounds__(WGS) void k_batch(gdb_kernel_ctx_t* ctx, const DevTask* __restrict__ tasks,
                                               int ntasks, int G, int* team_state, int* next_task,
                                               int* results, int* teams_used,
                                               unsigned long long* work) {
  DISCOVERY_PROTOCOL(ctx);
  __shared__ int gen_cache;
  __shared__ int lq[2][kLocalQ];
  __shared__ int lcnt[3];
  const int P = p_get_num_groups(ctx, &local_ctx);
  const int id = p_get_group_id(ctx, &local_ctx);
  if (G > P) G = P;
  if (G < 1) G = 1;
  const int nteams = P / G;
  const int team = id / G;
  if (team >= nteams) return;
  if (threadIdx.x == 0) gen_cache = GDB_GEN_UNKNOWN;
  Team T{ctx, &local_ctx, team_state + (size_t)team * TW_WORDS * kLine, &gen_cache, G, id % G, lq, lcnt};
  if (id == 0 && threadIdx.x == 0) *teams_used = nteams;
  for (;;) {
    if (T.rank == 0 && threadIdx.x == 0) gdb::st_rlx(T.word(TW_TASK), gdb::fadd_rlx(next_task, 1));
    if (!team_barrier(T)) return;
    const int t = gdb::ld_rlx(T.word(TW_TASK));
    if (t >= ntasks) break;
    const DevTask k = tasks[t];
    if (T.tid() == 0) work[4 * (size_t)t + 2] = gdb::now();
    int rounds = 0;
    if (k.n > 0) {
      unsigned long long* wk = work + 4 * (size_t)t;
      if (k.algo == GDB_TASK_BFS)
        rounds = k.hops > 0 ? team_worklist<true, true>(T, k, wk) : team_worklist<false, false>(T, k, wk);
      else if (k.algo == GDB_TASK_SSSP) rounds = team_worklist<true, false>(T, k, wk);
      else rounds = team_color(T, k);
    }
    if (rounds < 0) return;
    if (T.tid() == 0) {
      results[t] = rounds;
      work[4 * (size_t)t + 3] = gdb::now();
    }
    if (!team_barrier(T)) return;
  }
}

}  // namespace

using namespace gdbrt;

extern "C" int gdb_batch_solve(gdb_ctx* ctx, gdb_task* tasks, int ntasks, int team_groups,
                               const gdb_opts* opts, gdb_stats* stats) {
  if (!ctx || ntasks < 0 || (ntasks > 0 && !tasks) || team_groups < 0) return GDB_E_INVALID;
  gdb_opts o{};
  if (opts) o = *opts;
  const int wgs = default_wgs(&o);
  gdb_stats st{};
  if (ntasks == 0) {
    if (stats) *stats = st;
    return GDB_SUCCESS;
  }
  // Per-task scratch sizes (ints) and the device task table.
  std::vector<DevTask> dt((size_t)ntasks);
  size_t scratch_total = 0;
  long long max_n = 0;
  for (int t = 0; t < ntasks; ++t) {
    const gdb_task& a = tasks[t];
    if (a.n < 0 || a.m < 0 || a.algo < GDB_TASK_BFS || a.algo > GDB_TASK_COLOR || !a.out ||
        (a.n > 0 && !a.row) || (a.m > 0 && !a.col) || (a.algo == GDB_TASK_SSSP && a.m > 0 && !a.w) ||
        (a.algo == GDB_TASK_COLOR && a.n > 0 && !a.node_value) ||
        (a.algo != GDB_TASK_COLOR && a.n > 0 && (a.source < 0 || a.source >= a.n)) || a.delta < 0 ||
        a.hops < 0 || a.hops > 64)
      return GDB_E_INVALID;
    const size_t ints = (size_t)(a.algo == GDB_TASK_SSSP || a.hops > 0 ? 9 : 2) * a.n;
    dt[t] = DevTask{a.algo, a.n, a.m, a.source, a.delta, a.hops, a.row, a.col, a.w, a.node_value, a.out,
                    nullptr};
    scratch_total += (ints + 63) & ~(size_t)63;
    if (a.n > max_n) max_n = a.n;
  }
  int groups = 0;
  int rc = gdb_dispatch_wgs(wgs, [&](auto W) -> int {
    constexpr int WG = decltype(W)::value;
    return launch_groups(ctx, (const void*)k_batch<WG>, &o, (long long)GDB_MAX_GROUPS * WG, &groups);
  });
  if (rc) return rc;
  const size_t team_ints = (size_t)groups * TW_WORDS * kLine;
  const size_t bytes = carve_bytes<DevTask>(ntasks) + carve_bytes<int>(team_ints) +
                       carve_bytes<int>(ntasks + 64) + carve_bytes<unsigned long long>(4 * (size_t)ntasks) +
                       carve_bytes<int>(scratch_total);
  rc = ensure_workspace(ctx, bytes);
  if (rc) return rc;
  Carve cv(ctx->ws);
  DevTask* d_tasks = cv.take<DevTask>(ntasks);
  int* team_state = cv.take<int>(team_ints);
  int* misc = cv.take<int>(ntasks + 64);
  unsigned long long* work = cv.take<unsigned long long>(4 * (size_t)ntasks);
  int* scratch = cv.take<int>(scratch_total);
  size_t off = 0;
  for (int t = 0; t < ntasks; ++t) {
    const size_t ints = (size_t)(dt[t].algo == GDB_TASK_SSSP || dt[t].hops > 0 ? 9 : 2) * dt[t].n;
    dt[t].scratch = scratch + off;
    off += (ints + 63) & ~(size_t)63;
  }
  GDB_HIP(hipMemcpyAsync(d_tasks, dt.data(), sizeof(DevTask) * ntasks, hipMemcpyHostToDevice, ctx->stream));
  rc = fill_i32(ctx, team_state, 0, (long long)team_ints);
  if (!rc) rc = fill_i32(ctx, misc, 0, ntasks + 64);
  if (!rc) rc = fill_i32(ctx, (int*)work, 0, 8 * (long long)ntasks);
  if (rc) return rc;
  int* next_task = misc;
  int* teams_used = misc + 32;
  int* results = misc + 64;
  rc = prepare_launch(ctx);
  if (rc) return rc;
  const int G = team_groups > 0 ? team_groups : 1;
  rc = gdb_dispatch_wgs(wgs, [&](auto W) -> int {
    constexpr int WG = decltype(W)::value;
    int r2 = timer_start(ctx);
    if (r2) return r2;
    hipLaunchKernelGGL(k_batch<WG>, dim3(groups), dim3(WG), 0, ctx->stream, ctx->d, d_tasks, ntasks, G,
                       team_state, next_task, results, teams_used, work);
    GDB_HIP(hipGetLastError());
    return timer_stop(ctx, &st.kernel_ms);
  });
  if (rc) return rc;
  std::vector<int> hres((size_t)ntasks);
  std::vector<unsigned long long> hwork((size_t)4 * ntasks);
  GDB_HIP(hipMemcpy(hres.data(), results, sizeof(int) * ntasks, hipMemcpyDeviceToHost));
  GDB_HIP(hipMemcpy(hwork.data(), work, sizeof(unsigned long long) * 4 * ntasks, hipMemcpyDeviceToHost));
  for (int t = 0; t < ntasks; ++t) {
    tasks[t].rounds = hres[t];
    st.rounds += hres[t];
    // Worklist tasks report items / edges processed; colour tasks count the
    // graph once.
    const bool wl = tasks[t].algo != GDB_TASK_COLOR;
    st.work_items += wl ? (long long)hwork[4 * t] : tasks[t].n;
    st.work_edges += wl ? (long long)hwork[4 * t + 1] : tasks[t].m;
    // Device wall clock (100 MHz) from the team's start to end of the task.
    tasks[t].time_us = (double)(hwork[4 * t + 3] - hwork[4 * t + 2]) / 100.0;
  }
  st.launches = 1;
  rc = read_participating(ctx, &st.participating);
  if (rc) return rc;
  rc = check_device_error(ctx);
  if (stats) *stats = st;
  return rc;
}
