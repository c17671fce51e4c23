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
//
// BFS / SSSP iterations may also run a schedule other than "push on the whole
// team" (pull over in-edges, or on a narrow subset of the team's groups): the
// group that completes each team barrier picks the next iteration's arm
// (selector.hpp) and publishes it with the release.  All arms reach the same
// fixpoint, so outputs do not depend on the choices.
#include <algorithm>

#include "graph_common.hpp"
#include "runtime.hpp"
#include "selector.hpp"

namespace {

constexpr unsigned kInf = 1000000000u;  // bfs_common.cl:35
constexpr int kSerialLimit = 16;
constexpr int kChunk = 4;                  // edges a lane relaxes per step
constexpr int kLine = GDB_LINE_BYTES / 4;  // ints per 128 B line

// Team state, one 128 B line per word.  TW_RESUME and TW_SEL are lines of
// several words (see below).
enum TeamWord { TW_COUNT = 0, TW_GEN, TW_TASK, TW_TAIL0, TW_TAIL1, TW_TAIL2, TW_FAR0, TW_FAR1, TW_FAR2,
                TW_STOP0, TW_STOP1, TW_STOP2, TW_PARK, TW_RESUME, TW_SEL,
                TW_OARR0, TW_OARR1, TW_OARR2, TW_OFIN0, TW_OFIN1, TW_OFIN2, TW_OMIN0, TW_OMIN1, TW_OMIN2,
                TW_WORDS };
// TW_OARR* / TW_OFIN* / TW_OMIN*: arrival count, report count and smallest
// deferred distance of the owned solve's epochs (three, rotating).
// TW_RESUME line: what a parked group needs to rejoin (written before the
// release of TW_PARK).
enum ResumeWord {
  RS_DONE = 0, RS_IT, RS_PAYLOAD, RS_FGEN, RS_THRESH, RS_BARGEN, RS_TOTAL, RS_EVAL_NIN, RS_EVAL_IT, RS_WORDS
};
// TW_SEL line (ints, trace mode only): [0,1] time of the last release, [2]
// trace record awaiting its cost.
constexpr int kNinBits = 30;  // barrier payload = frontier size | arm << 30
constexpr int kNinMask = (1 << kNinBits) - 1;

struct DevTask {
  int algo, n, m, source, delta, hops, owned;
  const int* row;
  const int* col;
  const int* w;
  const float* nv;
  void* out;
  int* scratch;
  const int* trow;
  const int* tcol;
  const int* tw;
};

// Selector inputs of one task, identical in every group of the team.
struct SelRun {
  gdbsel::Cfg cfg;
  const unsigned short* lw;  // LDS copy of the weights (MLP mode) or nullptr
  int task;
  int Gn;          // groups of a narrow iteration
  unsigned valid;  // arms this task can run
};

// Loop state of the active groups (identical in each of them): what the
// releasing group decides from and publishes to parked groups.  `total` is
// the sum of the frontier sizes so far (for BFS, the vertices reached);
// eval_nin / eval_it are the frontier size and iteration of the last MLP
// evaluation.
struct IterState {
  int it, arm, nin, fgen;
  unsigned thresh;
  int wakes, total, eval_nin, eval_it;
};

// An MLP decision is reused while the frontier stays within 25% of the size
// it had when the network last ran, for at most 32 iterations, so that long
// runs of similar iterations (a grid's wavefront) do not pay the inference
// every time.  Every active group evaluates this on the same values.
__device__ __forceinline__ bool mlp_reuse(const IterState& S, int nin) {
  return S.eval_it > 0 && S.it + 1 - S.eval_it < 32 && nin > 0 &&
         (long long)nin * 4 <= (long long)S.eval_nin * 5 && (long long)nin * 5 >= (long long)S.eval_nin * 4;
}

// LDS queue capacity per buffer (ints): gdb_opts.lds_queue, in dynamic LDS.
constexpr int kLocalQDefault = 2048;
constexpr int kLocalQMin = 256, kLocalQMax = 4096;
constexpr int kMaxHops = 4096;  // local-step budget limit per team iteration

struct Team {
  gdb_kernel_ctx_t* ctx;
  gdb_local_ctx* l;
  int* ts;        // this team's state lines
  int* gen_cache; // LDS
  int G, rank;
  int* lq;             // LDS: two local queue buffers of qcap ints each
  int qcap;
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

__device__ __forceinline__ int sat_add(int a, int b) {
  const long long s = (long long)a + b;
  return s > 0x7fffffff ? 0x7fffffff : (int)s;
}
__device__ __forceinline__ unsigned long long ld64(const void* p) {
  return __hip_atomic_load((const unsigned long long*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st64(void* p, unsigned long long v) {
  __hip_atomic_store((unsigned long long*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Run by wave 0 of the group that completes a team barrier: decides the
// next iteration's arm and returns the payload (frontier size | arm << 30).
// The features come from the frontier size just read and the loop state
// every active group carries, so the decision costs no extra memory round
// trip; only the trace mode reads and writes the TW_SEL line (the time of
// the last release and the record awaiting its cost).
__device__ int sel_release(Team& T, const int* tail, const SelRun& R, const DevTask& k,
                           const IterState& S) {
  const int lane = (int)__lane_id();
  const gdbsel::Cfg& c = R.cfg;
  const bool tracing = c.trace && c.trace_count;
  int* sl = T.word(TW_SEL);
  // One load per lane, issued together.
  unsigned long long x = 0;
  if (lane == 0) x = (unsigned)gdb::ld_rlx(tail);
  else if (tracing && lane == 1) x = ld64(sl);
  else if (tracing && lane == 2) x = (unsigned)gdb::ld_rlx(sl + 2);
  auto get = [&](int l) -> unsigned long long {
    const unsigned lo = (unsigned)__shfl((int)(unsigned)x, l, 64);
    const unsigned hi = (unsigned)__shfl((int)(unsigned)(x >> 32), l, 64);
    return ((unsigned long long)hi << 32) | lo;
  };
  const int nin = (int)(unsigned)get(0);
  int arm, rec = -1;
  unsigned long long t = 0;
  if (tracing) {
    t = gdb::now();
    const unsigned long long t_rel = get(1);
    const int pending = (int)(unsigned)get(2);
    if (lane == 0 && pending >= 0 && pending < c.trace_capacity)
      c.trace[(size_t)pending * gdbsel::kStride + gdbsel::kIn + 1] = (float)(t - t_rel) * 0.01f;
  }
  if (nin == 0) {
    // A far-pile split (or the end): the whole team splits the pile, which
    // can hold most of the graph; nothing to learn.
    arm = GDB_ARM_PUSH;
  } else {
    gdbsel::FeatIn a;
    a.n = (float)k.n;
    a.m = (float)k.m;
    a.nin = (float)nin;
    a.total = (float)S.total + (float)nin;
    a.prev_nin = (float)S.nin;
    a.it = S.it + 1;
    a.algo = k.algo;
    a.prev_arm = S.arm;
    a.team_threads = (float)(T.G * (int)blockDim.x);
    a.narrow_threads = (float)(R.Gn * (int)blockDim.x);
    // The MLP input (lane i holds feature i); only the MLP and the trace use it.
    const bool use_mlp = c.mode == GDB_SEL_MLP && R.lw;
    const bool reuse = use_mlp && mlp_reuse(S, nin);
    const bool need_x = (use_mlp && !reuse) || tracing;
    const float xi = (need_x && lane < gdbsel::kIn) ? gdbsel::feature(a, lane) : 0.f;
    arm = reuse ? S.arm : gdbsel::decide(c, R.lw, a, R.valid, (unsigned)R.task * 0x10001u + (unsigned)S.it, xi);
    if (tracing) {
      if (lane == 0) rec = atomicAdd(c.trace_count, 1);
      rec = __shfl(rec, 0, 64);
      if (rec < c.trace_capacity) {
        float* r = c.trace + (size_t)rec * gdbsel::kStride;
        if (lane < gdbsel::kIn) r[lane] = xi;
        if (lane == 0) {
          r[gdbsel::kIn] = (float)arm;
          r[gdbsel::kIn + 1] = -1.f;
          r[gdbsel::kIn + 2] = (float)R.task;
          r[gdbsel::kIn + 3] = (float)(S.it + 1);
        }
      } else {
        rec = -1;
      }
    }
  }
  if (tracing && lane == 0) {
    st64(sl, t);
    gdb::st_rlx(sl + 2, rec);
  }
  return nin | (arm << kNinBits);
}

// Team barrier over the `members` active groups that also picks the next
// iteration's arm (in the group that completes it) and, when a narrow team
// widens again, wakes the parked groups.  Returns the payload.  Only wave 0
// takes part (lane 0 arrives or waits, the whole wave runs the selector), so
// the group pays the same two workgroup barriers as the plain team barrier.
__device__ int team_barrier_sel(Team& T, int members, const int* tail, const SelRun& R,
                                const DevTask& k, const IterState& S) {
  __shared__ int s_val;
  gdb::drain_stores();
  __syncthreads();
  if (threadIdx.x < 64) {
    const int lane = (int)threadIdx.x;
    const bool cross = members > 1;
    unsigned long long* word = (unsigned long long*)T.word(TW_GEN);
    int g = 0, last = 1;
    if (lane == 0) {
      g = *T.gen_cache;
      if (g == GDB_GEN_UNKNOWN && T.G > 1)
        g = (int)(unsigned)__hip_atomic_load(word, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
      if (cross) {
        const int old = gdb::fadd(T.word(TW_COUNT), 1);
        last = old == members - 1;
        if (last) gdb::st_rlx(T.word(TW_COUNT), 0);
      }
    }
    last = __shfl(last, 0, 64);
    if (last) {
      const int val = sel_release(T, tail, R, k, S);
      if (lane == 0) {
        const int new_arm = (int)((unsigned)val >> kNinBits);
        const bool wake = members < T.G && !gdbsel::arm_narrow(new_arm);
        const int ng = cross ? g + 1 : g;
        if (wake) {
          const int nn = val & kNinMask;
          const bool ev = R.cfg.mode == GDB_SEL_MLP && R.lw && nn > 0 && !mlp_reuse(S, nn);
          int* rs = T.word(TW_RESUME);
          gdb::st_rlx(rs + RS_DONE, 0);
          gdb::st_rlx(rs + RS_IT, S.it + 1);
          gdb::st_rlx(rs + RS_PAYLOAD, val);
          gdb::st_rlx(rs + RS_FGEN, S.fgen);
          gdb::st_rlx(rs + RS_THRESH, (int)S.thresh);
          gdb::st_rlx(rs + RS_BARGEN, ng);
          gdb::st_rlx(rs + RS_TOTAL, sat_add(S.total, nn));
          gdb::st_rlx(rs + RS_EVAL_NIN, ev ? nn : S.eval_nin);
          gdb::st_rlx(rs + RS_EVAL_IT, ev ? S.it + 1 : S.eval_it);
        }
        if (cross) {
          const unsigned long long w = ((unsigned long long)(unsigned)val << 32) | (unsigned)ng;
          __hip_atomic_store(word, w, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (wake) gdb::st_rel(T.word(TW_PARK), S.wakes + 1);
        *T.gen_cache = ng;
        s_val = val;
      }
    } else if (lane == 0) {
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
      s_val = (int)(w >> 32);
      *T.gen_cache = g + 1;
    }
  }
  __syncthreads();
  return __builtin_amdgcn_readfirstlane(s_val);
}

// A parked group waits until the team's wake count passes `wakes`, then
// reads the resume line.  Returns false when the wait gave up.
__device__ bool team_park(Team& T, int wakes, int (&rs)[RS_WORDS]) {
  __shared__ int s_rs[RS_WORDS];
  __shared__ int s_ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int* pw = T.word(TW_PARK);
    const bool ok = gdb::wait_while_idle(T.ctx, [&] { return gdb::ld_rlx(pw) <= wakes; },
                                         GDB_ERR_BARRIER_TIMEOUT);
    gdb::fence_acq();
    s_ok = ok;
    if (!ok) T.l->aborted = 1;
    const int* r = T.word(TW_RESUME);
    for (int i = 0; i < RS_WORDS; ++i) s_rs[i] = gdb::ld_rlx(r + i);
    *T.gen_cache = s_rs[RS_BARGEN];
  }
  __syncthreads();
  for (int i = 0; i < RS_WORDS; ++i) rs[i] = s_rs[i];
  return s_ok != 0;
}

// ---- BFS / SSSP on a team --------------------------------------------------
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
template <bool WEIGHTED, bool UNIT, bool SEL>
__device__ __forceinline__ int team_worklist(Team& T, const DevTask& k, unsigned long long* work,
                                             const SelRun& R) {
  // scratch: BFS (CAS)  q0[n], q1[n]
  //          otherwise  d64[n] (8 B each), fstamp[n], gstamp[n], q0[n], q1[n], far0..2[n]
  // Marks in the low word of d64 (level-synchronous iterations only):
  // 2*it for a near push of iteration it, 2*it+1 for a far push; on equal
  // distances the older mark stays, so "already queued in this iteration"
  // is one compare.  Iterations with local hops leave the mark at 0 and
  // dedupe team-queue pushes with gstamp instead, because a vertex pushed
  // to an LDS queue may be relaxed again before the iteration ends.
  __shared__ int lds_n;
  int* const lq = T.lq;
  const int qcap = T.qcap;
  int* lcnt = T.lcnt;
  const int n = k.n;
  const int G = T.G;
  unsigned* dist = (unsigned*)k.out;
  unsigned long long* d64 = (unsigned long long*)k.scratch;
  int* base = WEIGHTED ? k.scratch + 2 * (size_t)n : k.scratch;
  int* fstamp = base;
  int* gstamp = base + n;
  int* q[2] = {WEIGHTED ? base + 2 * (size_t)n : base, WEIGHTED ? base + 3 * (size_t)n : base + n};
  int* far[3] = {base + 4 * (size_t)n, base + 5 * (size_t)n, base + 6 * (size_t)n};
  const bool near_far = WEIGHTED && !UNIT && k.delta > 0;
  const int hops = WEIGHTED ? k.hops : 0;
  // The selector runs on level-synchronous tasks only (no local hops).  The
  // SEL = false instantiation carries none of its code (fewer registers).
  const bool sel = SEL && !UNIT && R.cfg.mode != GDB_SEL_NONE && k.hops == 0;
  const int Gn = sel ? R.Gn : G;
  {
    const int tid = T.tid(), stride = T.stride();
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
      if (sel) {
        int* sl = T.word(TW_SEL);
        st64(sl, gdb::now());
        gdb::st_rlx(sl + 2, -1);
      }
    }
  }
  unsigned thresh = near_far ? (unsigned)k.delta : kInf;
  int gen = 0;  // far pile generation being filled
  int it = 1;
  int nin = 1;
  int arm = GDB_ARM_PUSH;
  int wakes = 0;  // wake-ups of parked groups seen so far
  int total = 0;  // sum of frontier sizes so far
  int eval_nin = 0, eval_it = 0;
  const bool use_mlp = R.cfg.mode == GDB_SEL_MLP && R.lw;
  if (SEL && sel) {
    const IterState S0{0, GDB_ARM_PUSH, 0, 0, thresh, 0, 0, 0, 0};
    const int pl = team_barrier_sel(T, G, T.word(TW_TAIL1), R, k, S0);
    if (T.l->aborted) return -1;
    nin = pl & kNinMask;
    arm = (int)((unsigned)pl >> kNinBits);
    total = nin;
    if (use_mlp) {
      eval_nin = nin;
      eval_it = 1;
    }
  } else if (!team_barrier(T)) {
    return -1;
  }
  long long items = 0, edges = 0;
  int ai[4] = {0, 0, 0, 0};
  for (;;) {
    const int width = gdbsel::arm_narrow(arm) ? Gn : G;
    if (SEL && T.rank >= width) {
      // Parked until the active groups widen the team again or finish.
      int rs[RS_WORDS];
      if (!team_park(T, wakes, rs)) return -1;
      ++wakes;
      it = rs[RS_IT];
      if (rs[RS_DONE]) break;
      nin = rs[RS_PAYLOAD] & kNinMask;
      arm = (int)((unsigned)rs[RS_PAYLOAD] >> kNinBits);
      gen = rs[RS_FGEN];
      thresh = (unsigned)rs[RS_THRESH];
      total = rs[RS_TOTAL];
      eval_nin = rs[RS_EVAL_NIN];
      eval_it = rs[RS_EVAL_IT];
      continue;
    }
    const int tid = T.rank * (int)blockDim.x + (int)threadIdx.x, stride = width * (int)blockDim.x;
    if (nin > 0) {  // relaxing iterations per arm (kept in scalar registers)
      ai[0] += arm == 0;
      ai[1] += arm == 1;
      ai[2] += arm == 2;
      ai[3] += arm == 3;
    }
    const int* in = q[(it - 1) & 1];
    int* outq = q[it & 1];
    int* out_tail = T.word(TW_TAIL0 + (it + 1) % 3);
    if (tid == 0) gdb::st_rlx(T.word(TW_TAIL0 + (it + 2) % 3), 0);
    const unsigned near_mark = 2u * (unsigned)it, far_mark = near_mark + 1u;
    int* fpile = far[gen % 3];
    int* ftail = T.word(TW_FAR0 + gen % 3);
    if (SEL && !UNIT && nin > 0 && gdbsel::arm_pull(arm)) {
      if (nin > n) {
        gdb::set_error(T.ctx, GDB_ERR_QUEUE_OVERFLOW);
        return -1;
      }
      // Pull: every vertex scans its in-edges.  BFS: an unvisited vertex
      // with an in-neighbour on level it-1 takes level it (bottom-up, first
      // hit ends the scan).  SSSP: d[v] = min(d[v], min d[u] + w(u, v)).
      // Each vertex is written by its own thread only.
      if constexpr (SEL && !UNIT) {
        GDB_FOR_UNIFORM(v, valid, tid, stride, n) {
          int b = 0, e = 0;
          bool cand = false;
          unsigned cur = kInf;
          if (valid) {
            cur = WEIGHTED ? (unsigned)(d64[v] >> 32) : dist[v];
            cand = WEIGHTED || cur == kInf;
            if (cand) {
              b = k.trow[v];
              e = k.trow[v + 1];
              ++items;
              edges += e - b;
            }
          }
          unsigned best = kInf;
          if constexpr (!WEIGHTED) {
            const unsigned lvl = (unsigned)(it - 1);
            const bool found = gdb::wave_coop_any(
                cand, b, e,
                [&](int kk) {
                  const int u = k.tcol[kk];
                  return (unsigned)u < (unsigned)n && dist[u] == lvl;
                },
                kSerialLimit);
            if (found) best = (unsigned)it;
          } else {
            best = gdb::wave_coop_reduce(
                cand, b, e, kInf,
                [&](int kk) -> unsigned {
                  const int u = k.tcol[kk];
                  if ((unsigned)u >= (unsigned)n) return kInf;
                  const unsigned wt = (unsigned)k.tw[kk];
                  const unsigned du = (unsigned)(d64[u] >> 32);
                  return (du < kInf && wt < kInf) ? du + wt : kInf;
                },
                [](unsigned x, unsigned y) { return x < y ? x : y; }, kSerialLimit);
          }
          int cls = 0;
          if (valid && best < cur) {
            if (!WEIGHTED) {
              dist[v] = best;
              cls = 1;
            } else {
              const bool nearp = best < thresh;
              d64[v] = ((unsigned long long)best << 32) | (nearp ? near_mark : far_mark);
              cls = nearp ? 1 : (atomicExch(&fstamp[v], gen) != gen ? 2 : 0);
            }
          }
          gdb::wave_push(cls == 1, v, outq, out_tail);
          if (near_far) gdb::wave_push(cls == 2, v, fpile, ftail);
        }
      }
    } else if (nin > 0) {
      if (nin > n) {
        gdb::set_error(T.ctx, GDB_ERR_QUEUE_OVERFLOW);
        return -1;
      }
      if (hops > 0 && threadIdx.x < 3) lcnt[threadIdx.x] = 0;
      if (hops > 0) __syncthreads();
      // Relaxes up to kChunk edges per lane (slot j is edge kk[j] when
      // ok[j]) from a vertex at distance src_d.  Fills the destinations and
      // their push classes: 0 none, 1 team queue, 2 far pile, 3 local (LDS)
      // queue.  The slots are branch-free, so every load is issued before
      // any atomic and every atomic before any compare: the slots cost one
      // chain of memory round trips, not kChunk chains.  An unused slot
      // loads edge 0 (present whenever a slot is used) and sends a no-op
      // atomic (min with all ones, max with INT_MIN) to this thread's spare
      // word, a valid vertex that spreads such atomics over many lines.
      const int spare = (int)(((unsigned)blockIdx.x * blockDim.x + threadIdx.x) % (unsigned)n);
      auto relax_n = [&](const bool (&ok)[kChunk], const int (&kk)[kChunk], unsigned src_d,
                         int (&dst)[kChunk], int (&cls)[kChunk]) {
#pragma unroll
        for (int j = 0; j < kChunk; ++j) {
          dst[j] = -1;
          cls[j] = 0;
        }
        bool any = false;
#pragma unroll
        for (int j = 0; j < kChunk; ++j) any |= ok[j];
        // Lanes without a slot this step stay out (no memory traffic); inside
        // the branch every result is consumed before it closes.
        if (!any) return;
        unsigned wt[kChunk];
        bool act[kChunk];
        unsigned cand[kChunk];
        int at[kChunk];
#pragma unroll
        for (int j = 0; j < kChunk; ++j) {
          const int e = ok[j] ? kk[j] : 0;
          dst[j] = gdb::ldg(k.col + e);
          wt[j] = (!WEIGHTED || UNIT) ? 1u : (unsigned)gdb::ldg(k.w + e);
        }
#pragma unroll
        for (int j = 0; j < kChunk; ++j) {
          act[j] = ok[j] && (unsigned)dst[j] < (unsigned)n && wt[j] < kInf;
          if (!ok[j]) dst[j] = -1;
          at[j] = act[j] ? dst[j] : spare;
          cand[j] = src_d + wt[j];
          cls[j] = 0;
        }
        if constexpr (!WEIGHTED) {
          // An unused slot compares against all ones, which no distance is.
          unsigned old[kChunk];
#pragma unroll
          for (int j = 0; j < kChunk; ++j)
            old[j] = gdb::atomic_cas_g(&dist[at[j]], act[j] ? kInf : ~0u, act[j] ? src_d : ~0u);
#pragma unroll
          for (int j = 0; j < kChunk; ++j) cls[j] = (act[j] && old[j] == kInf) ? 1 : 0;
        } else {
          bool imp[kChunk], nearp[kChunk];
          if (hops > 0) {
            // The mark word stays 0 with local hops, so a 32-bit atomic on
            // the distance word does (a 64-bit one has a third of the
            // throughput).
            unsigned old[kChunk];
#pragma unroll
            for (int j = 0; j < kChunk; ++j)
              old[j] = gdb::atomic_min_g(reinterpret_cast<unsigned*>(&d64[at[j]]) + 1, act[j] ? cand[j] : ~0u);
#pragma unroll
            for (int j = 0; j < kChunk; ++j) {
              imp[j] = act[j] && cand[j] < old[j];
              nearp[j] = cand[j] < thresh;
              if (imp[j] && nearp[j]) cls[j] = 3;
            }
          } else {
            unsigned long long old[kChunk];
#pragma unroll
            for (int j = 0; j < kChunk; ++j) {
              nearp[j] = cand[j] < thresh;
              const unsigned long long nw =
                  ((unsigned long long)cand[j] << 32) | (nearp[j] ? near_mark : far_mark);
              old[j] = gdb::atomic_min_g(&d64[at[j]], act[j] ? nw : ~0ull);
            }
#pragma unroll
            for (int j = 0; j < kChunk; ++j) {
              imp[j] = act[j] && cand[j] < (unsigned)(old[j] >> 32);
              if (imp[j] && nearp[j]) cls[j] = (unsigned)old[j] == near_mark ? 0 : 1;
            }
          }
          // An improvement beyond the threshold joins the far pile once per
          // generation: the first to raise the vertex's stamp to `gen` pushes
          // it (stamps only grow, and none exceeds gen while relaxing).
          bool far_any = false;
#pragma unroll
          for (int j = 0; j < kChunk; ++j) far_any |= imp[j] && !nearp[j];
          if (far_any) {
            int fs[kChunk];
#pragma unroll
            for (int j = 0; j < kChunk; ++j) {
              const bool f = imp[j] && !nearp[j];
              fs[j] = gdb::atomic_max_g(&fstamp[f ? dst[j] : spare], f ? gen : INT_MIN);
            }
#pragma unroll
            for (int j = 0; j < kChunk; ++j)
              if (imp[j] && !nearp[j]) cls[j] = fs[j] < gen ? 2 : 0;
          }
        }
      };
      // Pushes the results of one relaxation step of this wave.  Class 3
      // goes to LDS buffer `lb` when there is room, else to the team queue.
      auto push_n = [&](const int (&dst)[kChunk], const int (&cls)[kChunk], int lb, int* lc) {
        bool to_team[kChunk], to_far[kChunk];
        unsigned long long m3[kChunk];
        int n3 = 0;
#pragma unroll
        for (int j = 0; j < kChunk; ++j) {
          to_team[j] = cls[j] == 1;
          to_far[j] = cls[j] == 2;
          m3[j] = __ballot(cls[j] == 3);
          n3 += __popcll(m3[j]);
        }
        if (n3) {  // one LDS atomic per wave for the step's local pushes
          int pos = 0;
          if (gdb::lane_id() == 0) pos = atomicAdd(lc, n3);
          pos = gdb::bcast(pos, 0);
          const unsigned long long lt = (1ull << gdb::lane_id()) - 1ull;
#pragma unroll
          for (int j = 0; j < kChunk; ++j) {
            if (cls[j] == 3) {
              const int at = pos + __popcll(m3[j] & lt);
              if (at < qcap) lq[lb * qcap + at] = dst[j];
              else to_team[j] = atomicExch(&gstamp[dst[j]], it) != it;
            }
            pos += __popcll(m3[j]);
          }
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
            bool ok[kChunk];
            int kk[kChunk], dst[kChunk], cls[kChunk];
#pragma unroll
            for (int j = 0; j < kChunk; ++j) {
              ok[j] = c + j < len;
              kk[j] = b + c + j;
            }
            relax_n(ok, kk, dv, dst, cls);
            push_n(dst, cls, lb, lc);
          }
          // Heavy rows: the whole wave walks one row, 64 edges per step.
          gdb::wave_coop_foreach(
              heavy, b, e, dv,
              [&](int e, bool ok, unsigned src_d) {
                bool oks[kChunk];
                int kk[kChunk], dst[kChunk], cls[kChunk];
#pragma unroll
                for (int j = 0; j < kChunk; ++j) {
                  oks[j] = j == 0 && ok;
                  kk[j] = e;
                }
                relax_n(oks, kk, src_d, dst, cls);
                push_n(dst, cls, lb, lc);
              },
              -1);
        }
      };
      // The team's share of the iteration; improved vertices go to LDS
      // buffer 1 (local step 1 reads it).
      process(in, nin, tid, stride, 1, &lcnt[1]);
      // Local steps: this group alone relaxes what it just improved.  Only
      // LDS is exchanged between its waves here (the relaxations are
      // returning atomics; queue stores are read after the team barrier,
      // which drains them), so a workgroup barrier suffices.
      int s = 1;
      for (; s <= hops; ++s) {
        __syncthreads();
        const int ln = min(lcnt[s % 3], qcap);
        if (ln == 0) break;
        if (threadIdx.x == 0) lcnt[(s + 2) % 3] = 0;
        process(lq + (s & 1) * qcap, ln, (int)threadIdx.x, (int)blockDim.x, (s + 1) & 1, &lcnt[(s + 1) % 3]);
      }
      if (hops > 0 && s > hops) {
        // Budget spent: flush the last local buffer to the team queue.
        __syncthreads();
        const int ln = min(lcnt[s % 3], qcap);
        GDB_FOR_UNIFORM(i, valid, (int)threadIdx.x, (int)blockDim.x, ln) {
          int v = 0;
          bool p = false;
          if (valid) {
            v = lq[(s & 1) * qcap + i];
            p = atomicExch(&gstamp[v], it) != it;
          }
          gdb::wave_push(p, v, outq, out_tail);
        }
      }
    } else {
      // Near queue empty: split the far pile of generation `gen` at the next
      // threshold into the near queue and generation gen+1.
      const int fsz = gdb::ld_rlx(T.word(TW_FAR0 + gen % 3));
      if (!near_far || fsz == 0) {
        if (width < G && T.rank == 0 && threadIdx.x == 0) {
          // Finished on a narrow team: release the parked groups.
          int* rs = T.word(TW_RESUME);
          gdb::st_rlx(rs + RS_DONE, 1);
          gdb::st_rlx(rs + RS_IT, it);
          gdb::st_rlx(rs + RS_BARGEN, *T.gen_cache);
          gdb::st_rel(T.word(TW_PARK), wakes + 1);
        }
        break;
      }
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
    if (SEL && sel) {
      const IterState S{it, arm, nin, gen, thresh, wakes, total, eval_nin, eval_it};
      const int pl = team_barrier_sel(T, width, out_tail, R, k, S);
      if (T.l->aborted) return -1;
      const int new_arm = (int)((unsigned)pl >> kNinBits);
      if (width < G && !gdbsel::arm_narrow(new_arm)) ++wakes;  // the releaser woke the parked groups
      nin = pl & kNinMask;
      arm = new_arm;
      total = sat_add(total, nin);
      if (use_mlp && nin > 0 && !mlp_reuse(S, nin)) {  // the releaser ran the network
        eval_nin = nin;
        eval_it = it + 1;
      }
    } else {
      nin = team_barrier_payload(T, out_tail, &lds_n);
      if (T.l->aborted) return -1;
    }
    ++it;
  }
  if (WEIGHTED)
    for (int v = T.tid(); v < n; v += T.stride()) dist[v] = (unsigned)(d64[v] >> 32);
  wave_add(&work[0], items);
  wave_add(&work[1], edges);
  if (T.rank == 0 && threadIdx.x == 0)
    for (int a = 0; a < 4; ++a) work[4 + a] = (unsigned long long)ai[a];
  return it - 1;
}

constexpr int kColorVerts = 2;  // vertices per lane per colour pass (4: slower, spills)

// ---- colour on a team (one barrier per round, as k_color_tuned) ------------
// scratch: pk[n] (int4: colour, value bits, max_d of the even and odd rounds),
// then blk[n] (the neighbour that set max_d, -1 none).
// A vertex's words travel together, so a neighbour costs one 16-byte load.
// A racing read of a neighbour's record sees its colour as -1 or r-1, which
// give the same answer, and round r never reads the max_d word it writes.
// The colours are copied to `out` at the end.
//
// Blockers: a vertex's max_d is the largest value among its uncoloured
// neighbours, and neighbours only ever leave that set.  While the neighbour
// that held the maximum (the blocker) is still uncoloured, the maximum is
// unchanged, so the vertex checks that one record instead of its row; it
// walks its row again only once the blocker is coloured.  Colours, rounds
// and max_d values equal those of the full walk every round.
__device__ __forceinline__ int team_color(Team& T, const DevTask& k) {
  const int n = k.n;
  int* color = (int*)k.out;
  int4* pk = (int4*)k.scratch;
  int* blk = (int*)(pk + n);
  const float* nv = k.nv;
  const int tid = T.tid(), stride = T.stride();
  for (int v = tid; v < n; v += stride) {
    pk[v] = make_int4(-1, __float_as_int(nv[v]), 0, 0);
    blk[v] = -1;
  }
  if (tid == 0) {
    gdb::st_rlx(T.word(TW_STOP0), 0);
    gdb::st_rlx(T.word(TW_STOP1), 0);
    gdb::st_rlx(T.word(TW_STOP2), 0);
  }
  if (!team_barrier(T)) return -1;
  int r = 1;
  for (;;) {
    const int cur = r & 1;  // max_d word written this round (z even, w odd)
    const bool have_prev = r > 1;
    if (tid == 0) gdb::st_rlx(T.word(TW_STOP0 + (r + 1) % 3), 0);
    // Uncoloured as seen in round r: not coloured before, and not colouring
    // now (value above the max_d of the previous round).
    auto uncoloured = [&](const int4 p) {
      return p.x == -1 && !(have_prev && __int_as_float(p.y) > __int_as_float(cur ? p.z : p.w));
    };
    bool st = false;
    auto edge = [&](int kk) -> gdb::MaxAt {
      const int u = k.col[kk];
      const int4 q = pk[u];
      return uncoloured(q) ? gdb::MaxAt{__int_as_float(q.y), u} : gdb::MaxAt{-1.f, -1};
    };
    auto red = [](gdb::MaxAt a, gdb::MaxAt c) { return (c.at >= 0 && (a.at < 0 || c.mx > a.mx)) ? c : a; };
    // kColorVerts vertices per lane per pass (v, v + stride, ...), their loads
    // issued together: a blocked vertex is a short chain of dependent loads (its
    // record and blocker id, then the blocker's record), so several chains
    // in flight hide more latency.  Loads are branch-free (vertex 0 stands in
    // for a missing vertex or blocker); the row walks follow one by one.
    for (long long base = (long long)tid - threadIdx.x; base < (long long)n; base += (long long)kColorVerts * stride) {
      int vv[kColorVerts], bk[kColorVerts], b[kColorVerts], e[kColorVerts];
      bool val[kColorVerts], active[kColorVerts], blocked[kColorVerts];
      int4 p[kColorVerts];
#pragma unroll
      for (int j = 0; j < kColorVerts; ++j) {
        const long long vi = base + threadIdx.x + (long long)j * stride;
        val[j] = vi < (long long)n;
        vv[j] = val[j] ? (int)vi : 0;
        p[j] = pk[vv[j]];
        bk[j] = blk[vv[j]];
        b[j] = k.row[vv[j]];
        e[j] = k.row[vv[j] + 1];
      }
      int4 q[kColorVerts];
#pragma unroll
      for (int j = 0; j < kColorVerts; ++j) {
        active[j] = false;
        if (val[j] && p[j].x == -1) {
          if (have_prev && __int_as_float(p[j].y) > __int_as_float(cur ? p[j].z : p[j].w))
            ((int*)&pk[vv[j]])[0] = r - 1;
          else
            active[j] = true;
        }
        q[j] = pk[active[j] && bk[j] >= 0 ? bk[j] : 0];
      }
#pragma unroll
      for (int j = 0; j < kColorVerts; ++j) {
        blocked[j] = active[j] && bk[j] >= 0 && uncoloured(q[j]);
        if (!active[j]) b[j] = e[j] = 0;
      }
      // Light rows of all kColorVerts vertices walked together, two edges
      // of each per step, every load branch-free (edge 0 and vertex 0 stand
      // in for unused slots; a slot is used only when the graph has edges).
      bool walk[kColorVerts], heavy[kColorVerts];
      int len[kColorVerts];
      gdb::MaxAt acc[kColorVerts];
      int maxlen = 0;
#pragma unroll
      for (int j = 0; j < kColorVerts; ++j) {
        walk[j] = active[j] && (e[j] - b[j]) != 1 && !blocked[j];
        heavy[j] = walk[j] && (e[j] - b[j]) > kSerialLimit;
        len[j] = walk[j] && !heavy[j] ? e[j] - b[j] : 0;
        maxlen = max(maxlen, len[j]);
        acc[j] = gdb::MaxAt{-1.f, -1};
      }
      for (int t = 0; t < maxlen; t += 2) {
        int u[kColorVerts][2];
        int4 qq[kColorVerts][2];
#pragma unroll
        for (int j = 0; j < kColorVerts; ++j)
#pragma unroll
          for (int h = 0; h < 2; ++h) u[j][h] = k.col[t + h < len[j] ? b[j] + t + h : 0];
#pragma unroll
        for (int j = 0; j < kColorVerts; ++j)
#pragma unroll
          for (int h = 0; h < 2; ++h) qq[j][h] = pk[t + h < len[j] ? u[j][h] : 0];
#pragma unroll
        for (int j = 0; j < kColorVerts; ++j)
#pragma unroll
          for (int h = 0; h < 2; ++h)
            if (t + h < len[j] && uncoloured(qq[j][h]))
              acc[j] = red(acc[j], gdb::MaxAt{__int_as_float(qq[j][h].y), u[j][h]});
      }
#pragma unroll
      for (int j = 0; j < kColorVerts; ++j) {
        const bool deg1 = (e[j] - b[j]) == 1;
        // Rows longer than kSerialLimit: the whole wave walks them.
        const gdb::MaxAt hres =
            gdb::wave_coop_reduce(heavy[j], b[j], e[j], gdb::MaxAt{-1.f, -1}, edge, red, kSerialLimit);
        const gdb::MaxAt res = red(acc[j], hres);
        if (active[j]) {
          const int v = vv[j];
          if (blocked[j]) {
            // The maximum of the previous round stands; copy it to this
            // round's word when they differ.
            const int prev = cur ? p[j].z : p[j].w;
            if ((cur ? p[j].w : p[j].z) != prev) ((int*)&pk[v])[cur ? 3 : 2] = prev;
            st = true;
          } else {
            const float md = deg1 ? -1.f : (res.mx > -1.f ? res.mx : -1.f);
            ((int*)&pk[v])[cur ? 3 : 2] = __float_as_int(md);
            if (!deg1) {
              if (res.at != bk[j]) blk[v] = res.at;
              if (res.at >= 0) st = true;
            }
          }
        }
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
  const int cur = r & 1;
  for (int v = tid; v < n; v += stride) {
    const int4 p = pk[v];
    const float md = __int_as_float(cur ? p.w : p.z);
    color[v] = (p.x == -1 && __int_as_float(p.y) > md) ? r : p.x;
  }
  return r;
}

// ---- Owner-partitioned BFS / SSSP on a team (gdb_task.owned) --------------
// Group r of a team of G owns the vertex range [r*B, r*B + B), B = ceil(n/G)
// rounded up to 32, and keeps the range's distances, queue marks and queues
// (and row offsets, when they fit) in LDS.  A relaxation whose destination
// lies in the range is an LDS atomic, and the group's rounds are separated by
// workgroup barriers only, so a hop along a path inside a range costs one
// memory round trip (the vertex's slot record, below) instead of three.  A
// relaxation into another range is a fire-and-forget device-scope atomic min
// on the destination's output word; an owner reads the output words of its
// boundary vertices (the ones with an in-edge from another range, marked
// once per task) every few rounds (one device-scope load per boundary vertex,
// up to kScanBatch in flight per thread: 3.3-3.7 us per scan on the grid)
// and while it is idle.  (A per-owner counter of such updates, to skip scans that find
// nothing, measured slower on a 2048^2 grid: every scan there finds
// something, and the extra atomic per edge costs.)
//
// Distance bands: the team relaxes only below a common threshold T (SSSP:
// delta; BFS: kOwnedBfsBand levels); an improvement at or above T waits in
// the owner's far set.  Without the common band, ranges that run ahead
// spread distances that later prove too long and relax them again (an
// order of magnitude more work on a grid).
//
// Epochs e = 0, 1, ...: a group with nothing left below T drains its
// atomics and arrives on the epoch's counter, then keeps polling its
// boundary; an improvement below T found while fewer than G have arrived
// un-arrives it (compare-and-swap, so the count never drops from G).  Once
// all G have arrived no relaxation is in flight: every group scans its
// boundary once more and reports whether anything below T improved, and the
// smallest distance in its far set.  If something improved the band stays;
// otherwise T moves to that smallest distance plus the band width, or the
// solve ends when every far set is empty.  Every order of relaxations
// reaches the same fixpoint, so outputs are the exact levels / distances of
// the other engines.
//
// LDS (dynamic, `region` bytes): dist[B] u32 | inq[B/32] | far[B/32] |
// queue u16[2][qcap] | row offsets[B + 1] when they fit | boundary list u16.
// Rounds between boundary scans of a busy group, and BFS levels per band,
// as measured on the 2048^2 grid (128 ranges, four lanes per queued vertex,
// scans beside the rounds when the spare waves cover the boundary in one
// batch): with four lanes per vertex, BFS+SSSP took 11.2 ms at 8 rounds, 11.4
// at 6 and 11.8 at 4 (every 2 to 16 measured); with two lanes per SSSP
// vertex, 10.80-10.83 ms at 10 rounds against 10.88-10.97 at 8, 10.87-10.93
// at 12 and 10.95-10.96 at 6 (same box).  BFS bands of 1024 levels (2048: no
// better).
constexpr int kOwnedScanEvery = 10;
constexpr int kOwnedBfsBand = 1024;
// Queue entries per buffer at most: a larger tile would take the LDS the
// boundary list needs (a range whose list does not fit scans all of its
// vertices instead: 23.8 against 19.5 ms at a 4096-entry tile).
constexpr int kOwnedQueueMax = 2048;
// Dynamic LDS of a batch with an owned task: a CU's 160 KiB less the kernel's
// static LDS (a few hundred bytes).
constexpr long long kOwnedLdsTotal = 163840 - 1024;
// Slot records (device memory, task scratch after the boundary marks): each
// range writes one record per vertex at the start of the solve, and a round
// reads it in place of the row offsets and the CSR, one dependent load
// instead of two.  A record holds up to kChunk out-edges (BFS: ids, 16 B;
// SSSP: (id, weight) pairs, 32 B) padded with kSlotNone; a vertex with more
// marks slot 0 with kSlotCsr and is walked from the CSR.
constexpr int kSlotNone = -2147483647 - 1;
constexpr int kSlotCsr = -2147483647;
static_assert(kChunk == 4, "a slot record is one or two 16-byte vectors");
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) i32x4 g_i32x4;
// Scratch ints of an owned task: boundary marks, then the slot records.
__host__ __device__ inline size_t owned_scratch_ints(int n, bool unit) {
  return ((((size_t)n + 31) / 32 + 63) & ~(size_t)63) + (size_t)n * kChunk * (unit ? 1 : 2);
}

// Bytes of LDS the owned solve needs without the optional parts.
__host__ __device__ constexpr long long owned_lds_core(long long B, int qcap) {
  return 4 * B + 2 * (B / 32) * 4 + 4LL * qcap;
}
__host__ __device__ inline int owned_range(int n, int G) { return (((n + G - 1) / G) + 31) & ~31; }

// LDS pointers of the owned solve carry their address space, so that its
// accesses are ds_* instructions (a generic pointer would give flat ones,
// whose waits also cover the memory loads in flight).
typedef __attribute__((address_space(3))) unsigned lds_u32;
typedef __attribute__((address_space(3))) unsigned short lds_u16;
typedef __attribute__((address_space(3))) int lds_i32;
typedef __attribute__((address_space(3))) char lds_char;
typedef __attribute__((address_space(1))) unsigned g_u32;
typedef __attribute__((address_space(1))) const int g_ci32;
__device__ __forceinline__ unsigned lds_or(lds_u32* p, unsigned v) {
  return __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ unsigned lds_and(lds_u32* p, unsigned v) {
  return __hip_atomic_fetch_and(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <bool UNIT>
__device__ int team_owned(Team& T, const DevTask& k, unsigned long long* work, int region) {
  __shared__ int s_cnt[3];   // rotating local queue counters
  __shared__ int s_nbl;      // boundary list length
  __shared__ int s_flag;     // an improvement was found (scan) / far set non-empty
  __shared__ unsigned s_min;
  __shared__ int s_c;
  __shared__ unsigned s_gmin;
  const int n = k.n, G = T.G, r = T.rank;
  const int B = owned_range(n, G);
  const int lo = min(n, r * B), hi = min(n, lo + B), nb = hi - lo;
  const int W = B >> 5;
  const int qc = min(T.qcap, kOwnedQueueMax);
  lds_char* const lb = (lds_char*)T.lq;
  lds_u32* const dist = (lds_u32*)lb;
  lds_u32* const inq = dist + B;
  lds_u32* const farb = inq + W;
  lds_u16* const q = (lds_u16*)(farb + W);
  long long used = owned_lds_core(B, qc);
  lds_i32* rowl = nullptr;
  if (used + 4LL * (B + 1) <= region) {
    rowl = (lds_i32*)(lb + used);
    used += 4LL * (B + 1);
  }
  lds_u16* const bl = (lds_u16*)(lb + used);
  const int blcap = (int)((region - used) / 2);
  // Device memory through global-address-space pointers (global_*, not
  // flat_* instructions, whose waits would also cover the LDS operations);
  // task fields in registers, since k lives in memory these stores could
  // alias.
  g_u32* const out = (g_u32*)k.out;
  g_u32* const bmark = (g_u32*)k.scratch;
  g_i32x4* const slots = (g_i32x4*)(k.scratch + ((((size_t)n + 31) / 32 + 63) & ~(size_t)63));
  g_ci32* const row = (g_ci32*)k.row;
  g_ci32* const col = (g_ci32*)k.col;
  g_ci32* const wgt = (g_ci32*)k.w;
  const int delta = k.delta, source = k.source;
  const int nw = (n + 31) >> 5;
  const int lane = gdb::lane_id();
  const unsigned long long lt = (1ull << lane) - 1ull;

  // Team-wide set-up: output words (remote minima land there), boundary
  // marks and the termination counters.
  {
    const int tid = T.tid(), stride = T.stride();
    for (int v = tid; v < n; v += stride) out[v] = kInf;
    for (int w = tid; w < nw; w += stride) bmark[w] = 0;
    if (tid == 0)
      for (int e = 0; e < 3; ++e) {
        gdb::st_rlx(T.word(TW_OARR0 + e), 0);
        gdb::st_rlx(T.word(TW_OFIN0 + e), 0);
        gdb::st_rlx(T.word(TW_OMIN0 + e), (int)kInf);
      }
  }
  for (int i = threadIdx.x; i < B; i += blockDim.x) dist[i] = kInf;
  for (int i = threadIdx.x; i < W; i += blockDim.x) {
    inq[i] = 0;
    farb[i] = 0;
  }
  if (rowl)
    for (int i = threadIdx.x; i <= nb; i += blockDim.x) rowl[i] = row[lo + i];
  if (threadIdx.x == 0) {
    s_cnt[0] = s_cnt[1] = s_cnt[2] = 0;
    s_nbl = 0;
  }
  if (!team_barrier(T)) return -1;
  // Boundary marks: every edge of this range into another range marks its
  // destination.
  GDB_FOR_UNIFORM(i, valid, (int)threadIdx.x, (int)blockDim.x, nb) {
    int b = 0, e = 0;
    if (valid) {
      b = row[lo + i];
      e = row[lo + i + 1];
    }
    gdb::wave_coop_foreach(
        valid, b, e, 0u,
        [&](int j, bool ok, unsigned) {
          const int v = ok ? col[j] : lo;
          if (ok && (unsigned)v < (unsigned)n && (unsigned)(v - lo) >= (unsigned)nb)
            __hip_atomic_fetch_or(&bmark[v >> 5], 1u << (v & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        },
        kSerialLimit);
  }
  // Slot records of this range (only this group reads them; the barrier
  // below orders the stores before the first round).
  {
    constexpr int kU = 4;  // vertices per thread per pass, their loads in flight together
    for (int base = 0; base < nb; base += kU * (int)blockDim.x) {
      int b[kU], d[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int i = base + u * (int)blockDim.x + (int)threadIdx.x;
        b[u] = 0;
        d[u] = 0;
        if (i < nb) {
          b[u] = row[lo + i];
          d[u] = row[lo + i + 1] - b[u];
        }
      }
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int i = base + u * (int)blockDim.x + (int)threadIdx.x;
        int c[kChunk], x[kChunk];
#pragma unroll
        for (int j = 0; j < kChunk; ++j) {
          const bool in = j < d[u] && d[u] <= kChunk;
          c[j] = in ? col[b[u] + j] : kSlotNone;
          x[j] = (in && !UNIT) ? wgt[b[u] + j] : 0;
        }
        if (d[u] > kChunk) c[0] = kSlotCsr;
        if (i < nb) {
          if (UNIT) {
            slots[lo + i] = i32x4{c[0], c[1], c[2], c[3]};
          } else {
            slots[2 * (size_t)(lo + i)] = i32x4{c[0], x[0], c[1], x[1]};
            slots[2 * (size_t)(lo + i) + 1] = i32x4{c[2], x[2], c[3], x[3]};
          }
        }
      }
    }
  }
  if ((unsigned)(source - lo) < (unsigned)nb && threadIdx.x == 0) {
    const int s0 = source - lo;
    dist[s0] = 0;
    inq[s0 >> 5] |= 1u << (s0 & 31);
    q[0] = (unsigned short)s0;
    s_cnt[0] = 1;
  }
  if (!team_barrier(T)) return -1;
  // Boundary list of this range.
  for (int w = threadIdx.x; w < W; w += blockDim.x) {
    unsigned bits = (lo + 32 * w < hi) ? __hip_atomic_load(&bmark[(lo >> 5) + w], __ATOMIC_RELAXED,
                                                           __HIP_MEMORY_SCOPE_AGENT)
                                       : 0u;
    if (bits) {
      int pos = atomicAdd(&s_nbl, __popc(bits));
      while (bits) {
        const int j = __ffs(bits) - 1;
        bits &= bits - 1;
        if (pos < blcap) bl[pos] = (unsigned short)(32 * w + j);
        ++pos;
      }
    }
  }
  __syncthreads();
  const int nbl = s_nbl;
  const bool scan_all = nbl > blcap;  // list did not fit: scan the whole range
  const int nscan = scan_all ? nb : nbl;

  const unsigned band = UNIT ? (unsigned)kOwnedBfsBand : delta > 0 ? (unsigned)delta : kInf;
  constexpr int kScanEvery = kOwnedScanEvery;
  unsigned thresh = band;
  long long items = 0, edges = 0;
  // Relaxes local vertices li[j] to cand[j] on lanes with act[j], N targets
  // per lane together: the N LDS minima are issued back to back, then the
  // marks, then one queue reservation for the wave (three LDS round trips
  // instead of 3N).  An improvement below the threshold joins queue buffer
  // qb (once: its inq bit); one above it, or one that finds the buffer full,
  // joins the far set.  All lanes of the wave call it.  imp[j]: target j of
  // this lane improved.
  auto relax_many = [&](const auto& act, const auto& li, const auto& cand, auto& imp, int qb, int cb) {
    constexpr int N = (int)(sizeof(act) / sizeof(act[0]));
    unsigned old[N];
#pragma unroll
    for (int j = 0; j < N; ++j)
      old[j] = act[j] ? __hip_atomic_fetch_min(&dist[li[j]], cand[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)
                      : 0u;
    unsigned prev[N];  // the inq word before this lane's mark (the bit itself when nothing is to be queued)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      imp[j] = act[j] && cand[j] < old[j];
      const unsigned bit = 1u << (li[j] & 31);
      prev[j] = bit;
      if (imp[j]) {
        if (cand[j] < thresh) prev[j] = lds_or(&inq[li[j] >> 5], bit);
        else lds_or(&farb[li[j] >> 5], bit);
      }
    }
    unsigned long long m[N], all = 0;
    int total = 0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
      m[j] = __ballot(!(prev[j] & (1u << (li[j] & 31))));
      all |= m[j];
      total += __popcll(m[j]);
    }
    if (all) {
      const int leader = __ffsll((long long)all) - 1;
      int base = 0;
      if (lane == leader) base = atomicAdd(&s_cnt[cb], total);
      base = gdb::bcast(base, leader);
#pragma unroll
      for (int j = 0; j < N; ++j) {
        if ((m[j] >> lane) & 1ull) {
          const int pos = base + __popcll(m[j] & lt);
          const unsigned bit = 1u << (li[j] & 31);
          if (pos < qc) {
            q[qb * qc + pos] = (unsigned short)li[j];
          } else {
            lds_and(&inq[li[j] >> 5], ~bit);
            lds_or(&farb[li[j] >> 5], bit);
          }
        }
        base += __popcll(m[j]);
      }
    }
  };
  // Relaxes every boundary vertex from its output word, kScanBatch loads per
  // thread in flight; sets s_flag when an improvement joined the queue
  // (below the threshold).
  // The scan runs on nt threads (t = 0 .. nt-1 among them); thread t's
  // boundary entries are base + j * nt + t.  scan_load loads kScanGroup of
  // them, scan_relax relaxes them (and says whether one improved below the
  // threshold).
  constexpr int kScanBatch = 8, kScanGroup = 4;  // 16 in flight spilled registers into the rounds
  auto scan_load = [&](int base, int t, int nt, int (&li)[kScanGroup], unsigned (&x)[kScanGroup]) {
#pragma unroll
    for (int j = 0; j < kScanGroup; ++j) {
      const int i = base + j * nt + t;
      li[j] = i < nscan ? (scan_all ? i : (int)bl[i]) : -1;
      x[j] = li[j] >= 0 ? __hip_atomic_load(&out[lo + li[j]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : kInf;
    }
  };
  auto scan_relax = [&](const int (&li)[kScanGroup], const unsigned (&x)[kScanGroup], int qb, int cb) -> bool {
    bool act[kScanGroup], imp[kScanGroup];
    int lg[kScanGroup];
#pragma unroll
    for (int j = 0; j < kScanGroup; ++j) {
      act[j] = x[j] < kInf;
      lg[j] = li[j] < 0 ? 0 : li[j];
    }
    relax_many(act, lg, x, imp, qb, cb);
    bool any = false;
#pragma unroll
    for (int j = 0; j < kScanGroup; ++j) any |= imp[j] && x[j] < thresh;
    return any;
  };
  // Every boundary entry, kScanBatch loads per thread in flight (relaxed
  // kScanGroup at a time: register pressure).
  auto scan = [&](int qb, int cb, int t, int nt) {
    bool any = false;
    const int step = kScanBatch * nt;
    for (int base = 0; base < nscan; base += step) {
      int li[kScanBatch / kScanGroup][kScanGroup];
      unsigned x[kScanBatch / kScanGroup][kScanGroup];
#pragma unroll
      for (int h = 0; h < kScanBatch / kScanGroup; ++h) scan_load(base + h * kScanGroup * nt, t, nt, li[h], x[h]);
#pragma unroll
      for (int h = 0; h < kScanBatch / kScanGroup; ++h) any |= scan_relax(li[h], x[h], qb, cb);
    }
    if (__ballot(any) && lane == 0) s_flag = 1;
  };
  // Relaxes the out-edges of queue buffer (qa, count ca) into (qb, cb).
  // Several lanes per queued vertex, each relaxing kPer of its slot-record
  // edges: a round is usually run by a few waves, so its length is their
  // instruction count.  BFS: four lanes of one edge; SSSP: two lanes of two
  // (one 16-byte load each; 10.9 against 11.2 ms with four lanes on the
  // 2048^2 grid).  The first lane of a vertex marked kSlotCsr walks its CSR
  // row.
  constexpr int kLanes = UNIT ? 4 : 2, kPer = kChunk / kLanes;
  auto round = [&](int qa, int ca, int qb, int cb) {
    const int ln = min(s_cnt[ca], qc);
    GDB_FOR_UNIFORM(t, valid, (int)threadIdx.x, (int)blockDim.x, ln * kLanes) {
      // A wave without an entry skips the body: its instructions would take
      // issue slots from the waves on its SIMD that have entries.
      if (!__ballot(valid)) continue;
      const int sl = t % kLanes;
      int li = 0, b = 0, e = 0;
      int v[kPer];
      unsigned wt[kPer];
#pragma unroll
      for (int j = 0; j < kPer; ++j) {
        v[j] = kSlotNone;
        wt[j] = 1u;
      }
      unsigned du = kInf;
      if (valid) {
        li = q[qa * qc + t / kLanes];
        // Clear the mark before reading the distance: a later improvement
        // queues the vertex again.  (A vertex's lanes share a wave, whose
        // LDS operations complete in order; the signal fence keeps the
        // compiler from swapping the two.)
        if (sl == 0) lds_and(&inq[li >> 5], ~(1u << (li & 31)));
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        du = __hip_atomic_load(&dist[li], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if constexpr (UNIT) {
          v[0] = ((g_ci32*)slots)[kChunk * (size_t)(lo + li) + sl];
        } else {
          const i32x4 p = slots[2 * (size_t)(lo + li) + sl];  // two (id, weight) pairs
          v[0] = p.x;
          wt[0] = (unsigned)p.y;
          v[1] = p.z;
          wt[1] = (unsigned)p.w;
        }
        if (sl == 0) ++items;
      }
      const bool csr = v[0] == kSlotCsr;  // slot 0 of a vertex with more than kChunk edges
      {
        bool act[kPer], imp[kPer];
        int lq[kPer];
        unsigned cd[kPer];
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
          const bool edge = v[j] != kSlotNone && v[j] != kSlotCsr;
          const bool ok = edge && wt[j] < kInf;
          const bool loc = (unsigned)(v[j] - lo) < (unsigned)nb;
          cd[j] = du + wt[j];
          if (ok && !loc && (unsigned)v[j] < (unsigned)n)
            __hip_atomic_fetch_min(&out[v[j]], cd[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          act[j] = ok && loc;
          lq[j] = loc ? v[j] - lo : 0;
          edges += edge;
        }
        relax_many(act, lq, cd, imp, qb, cb);
      }
      if (!__ballot(csr)) continue;
      if (csr) {
        b = rowl ? rowl[li] : row[lo + li];
        e = rowl ? rowl[li + 1] : row[lo + li + 1];
      }
      // Relaxes up to kChunk targets v (weights wt) of a vertex at src_d.
      auto relax_targets = [&](const bool (&ok)[kChunk], const int (&v)[kChunk], const unsigned (&wt)[kChunk],
                               unsigned src_d) {
        bool act[kChunk], imp[kChunk];
        int li[kChunk];
        unsigned cand[kChunk];
#pragma unroll
        for (int j = 0; j < kChunk; ++j) {
          const bool loc = (unsigned)(v[j] - lo) < (unsigned)nb;
          const bool rem = !loc && (unsigned)v[j] < (unsigned)n;
          cand[j] = src_d + wt[j];
          if (ok[j] && rem && wt[j] < kInf)
            __hip_atomic_fetch_min(&out[v[j]], cand[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          act[j] = ok[j] && loc && wt[j] < kInf;
          li[j] = loc ? v[j] - lo : 0;
        }
        relax_many(act, li, cand, imp, qb, cb);
      };
      auto relax_edges = [&](const bool (&ok)[kChunk], const int (&kk)[kChunk], unsigned src_d) {
        int v[kChunk];
        unsigned wt[kChunk];
#pragma unroll
        for (int j = 0; j < kChunk; ++j) {
          const int ee = ok[j] ? kk[j] : 0;
          v[j] = ok[j] ? col[ee] : -1;
          wt[j] = UNIT ? 1u : (ok[j] ? (unsigned)wgt[ee] : kInf);
        }
        relax_targets(ok, v, wt, src_d);
      };
      if (csr) edges += e - b;
      const bool heavy = csr && (e - b) > kSerialLimit;
      const int len = (csr && !heavy) ? (e - b) : 0;
      const int maxlen = gdb::wave_max(len);
      for (int c = 0; c < maxlen; c += kChunk) {
        bool ok[kChunk];
        int kk[kChunk];
#pragma unroll
        for (int j = 0; j < kChunk; ++j) {
          ok[j] = c + j < len;
          kk[j] = b + c + j;
        }
        relax_edges(ok, kk, du);
      }
      gdb::wave_coop_foreach(
          heavy, b, e, du,
          [&](int j, bool ok, unsigned src_d) {
            bool oks[kChunk];
            int kk[kChunk];
#pragma unroll
            for (int t = 0; t < kChunk; ++t) {
              oks[t] = t == 0 && ok;
              kk[t] = j;
            }
            relax_edges(oks, kk, src_d);
          },
          -1);
    }
  };
  // Smallest distance in the far set (kInf when empty), in s_min.
  auto far_min = [&]() {
    if (threadIdx.x == 0) s_min = kInf;
    __syncthreads();
    for (int w = threadIdx.x; w < W; w += blockDim.x) {
      unsigned bits = farb[w];
      while (bits) {
        const int j = __ffs(bits) - 1;
        bits &= bits - 1;
        atomicMin(&s_min, (unsigned)dist[32 * w + j]);
      }
    }
    __syncthreads();
  };
  // Far set: the vertices improved at or beyond the threshold, or that found
  // the queue full.  Queues every one now below the threshold; returns
  // whether it queued any.
  auto drain_far = [&](int qb, int cb) -> bool {
    far_min();
    if (s_min >= thresh) return false;
    for (int w = threadIdx.x; w < W; w += blockDim.x) {
      unsigned bits = farb[w];
      unsigned keep = bits;
      while (bits) {
        const int j = __ffs(bits) - 1;
        bits &= bits - 1;
        const int li = 32 * w + j;
        if (dist[li] < thresh) {
          keep &= ~(1u << j);
          const unsigned bit = 1u << j;
          if (!(lds_or(&inq[w], bit) & bit)) {
            const int pos = atomicAdd(&s_cnt[cb], 1);
            if (pos < qc) {
              q[qb * qc + pos] = (unsigned short)li;
            } else {
              lds_and(&inq[w], ~bit);
              keep |= bit;
            }
          }
        }
      }
      farb[w] = keep;
    }
    __syncthreads();
    return s_cnt[cb] > 0;
  };

  const unsigned long long limit = (unsigned long long)(unsigned)gdb::ld_rlx(&T.ctx->timeout_ticks.v);
  int s = 0;  // round number: queue buffer s & 1, counter s % 3
  int epoch = 0;
  bool done = false;
  __syncthreads();
  while (!done) {
    // Local rounds until the queue runs dry.
    for (;;) {
      const int ca = s % 3;
      if (s_cnt[ca] == 0) break;
      const int cb = (s + 1) % 3, qa = s & 1, qb = qa ^ 1;
      if (threadIdx.x == 0) s_cnt[(s + 2) % 3] = 0;
      // A boundary scan every kScanEvery rounds.  When the waves without a
      // queue entry can load every boundary word in one batch, they scan
      // while the others run the round (the two overlap); otherwise the
      // whole group scans after it (a scan costs a far-memory latency per
      // batch).
      const int split = (min(s_cnt[ca], qc) * kLanes + 63) & ~63;  // threads with an entry, in whole 64-lane waves
      const bool due = s % kScanEvery == kScanEvery - 1;
      const bool beside = ((int)blockDim.x - split) * kScanBatch >= nscan;
      if (due && beside && (int)threadIdx.x >= split) {
        scan(qb, cb, (int)threadIdx.x - split, (int)blockDim.x - split);
      } else {
        round(qa, ca, qb, cb);
        if (due && !beside) scan(qb, cb, (int)threadIdx.x, (int)blockDim.x);
      }
      ++s;
      __syncthreads();
    }
    if (threadIdx.x == 0) s_cnt[(s + 1) % 3] = 0;
    if (drain_far(s & 1, s % 3)) continue;
    // Idle: arrive on this epoch's counter and poll the boundary.
    int* const arr = T.word(TW_OARR0 + epoch % 3);
    int* const fin = T.word(TW_OFIN0 + epoch % 3);
    gdb::drain_stores();
    __syncthreads();
    if (threadIdx.x == 0) {
      if (r == 0) {  // counters of epoch + 1 (last used by epoch - 2, which every group has left)
        gdb::st_rlx(T.word(TW_OARR0 + (epoch + 1) % 3), 0);
        gdb::st_rlx(T.word(TW_OFIN0 + (epoch + 1) % 3), 0);
        gdb::st_rlx(T.word(TW_OMIN0 + (epoch + 1) % 3), (int)kInf);
      }
      gdb::fadd(arr, 1);
      s_flag = 0;
    }
    const unsigned long long t0 = gdb::now();
    bool active = false, pending = false;
    for (unsigned it = 0;; ++it) {
      __syncthreads();
      if (threadIdx.x == 0) {
        s_c = gdb::ld_acq(arr);
        if (s_c < G && (it & 15u) == 15u) {
          if (gdb::ld_rlx(&T.ctx->error.v) != GDB_OK || gdb::now() - t0 > limit) {
            gdb::set_error(T.ctx, GDB_ERR_BARRIER_TIMEOUT);
            T.l->aborted = 1;
          }
        }
      }
      __syncthreads();
      if (T.l->aborted) return -1;
      if (s_c == G) break;
      scan(s & 1, s % 3, (int)threadIdx.x, (int)blockDim.x);
      __syncthreads();
      if (s_flag) {
        if (threadIdx.x == 0) {
          int c = gdb::ld_rlx(arr);
          s_c = c;
          while (c < G) {
            if (__hip_atomic_compare_exchange_strong(arr, &c, c - 1, __ATOMIC_ACQ_REL, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT)) {
              s_c = -1;  // un-arrived
              break;
            }
          }
        }
        __syncthreads();
        if (s_c == -1) active = true;
        else pending = true;
        break;
      }
      if (threadIdx.x == 0) __builtin_amdgcn_s_sleep(8);
    }
    if (active) continue;
    // Everyone arrived: final scan, then report (improved below the
    // threshold?, smallest far distance).
    if (!pending) {
      if (threadIdx.x == 0) s_flag = 0;
      __syncthreads();
      scan(s & 1, s % 3, (int)threadIdx.x, (int)blockDim.x);
      __syncthreads();
      pending = s_flag != 0;
    }
    far_min();
    if (threadIdx.x == 0) {
      int* const mnw = T.word(TW_OMIN0 + epoch % 3);
      if (s_min < kInf)
        __hip_atomic_fetch_min((unsigned*)mnw, s_min, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      gdb::fadd(fin, 1 + (pending ? (1 << 16) : 0));
      int f = 0;
      const bool ok = gdb::wait_while(
          T.ctx, [&] { return ((f = gdb::ld_acq(fin)) & 0xffff) < G; }, GDB_ERR_BARRIER_TIMEOUT);
      if (!ok) T.l->aborted = 1;
      s_c = f >> 16;
      s_gmin = (unsigned)gdb::ld_rlx(mnw);
    }
    __syncthreads();
    if (T.l->aborted) return -1;
    ++epoch;
    if (s_c == 0) {
      // Nothing moved below the threshold: the next band, or the end.
      const unsigned gmin = s_gmin;
      if (gmin >= kInf) {
        done = true;
      } else {
        thresh = (band >= kInf - gmin) ? kInf : gmin + band;
        drain_far(s & 1, s % 3);
      }
    }
  }
  for (int i = threadIdx.x; i < nb; i += blockDim.x) out[lo + i] = dist[i];
  wave_add(&work[0], items);
  wave_add(&work[1], edges);
  return s;
}

// 512-thread groups serve many small tasks at once (throughput-bound): they
// are held to 80 VGPRs, three groups per CU instead of two, at the price of a
// few spilled long-lived values.
template <int WGS, bool SEL, bool OWN = false>
__global__ __launch_bounds__(WGS) __attribute__((amdgpu_waves_per_eu(WGS == 512 && !SEL && !OWN ? 6 : 1))) void k_batch(gdb_kernel_ctx_t* ctx, const DevTask* __restrict__ tasks,
                                               int ntasks, int G, int* team_state, int* next_task,
                                               int* results, int* teams_used,
                                               unsigned long long* work, gdbsel::Cfg sel, int qcap,
                                               int region) {
  DISCOVERY_PROTOCOL(ctx);
  __shared__ int gen_cache;
  __shared__ int lcnt[3];
  // Dynamic LDS: `region` bytes for the local queues (two buffers of qcap
  // ints) or an owned solve's range, then the MLP weights (MLP mode only).
  extern __shared__ int dyn_lds[];
  int* const lq = dyn_lds;
  unsigned short* const sel_w = (unsigned short*)((char*)dyn_lds + region);
  const int P = p_get_num_groups(ctx, &local_ctx);
  const int id = p_get_group_id(ctx, &local_ctx);
  if (G > P) G = P;
  if (G < 1) G = 1;
  const int nteams = P / G;
  const int team = id / G;
  if (team >= nteams) return;
  if (threadIdx.x == 0) gen_cache = GDB_GEN_UNKNOWN;
  const bool mlp = SEL && sel.mode == GDB_SEL_MLP && sel.w;
  if (mlp)
    for (int i = threadIdx.x; i < gdbsel::kParams; i += blockDim.x) sel_w[i] = sel.w[i];
  Team T{ctx, &local_ctx, team_state + (size_t)team * TW_WORDS * kLine, &gen_cache, G, id % G, lq, qcap, lcnt};
  if (id == 0 && threadIdx.x == 0) *teams_used = nteams;
  int Gn = sel.narrow_groups > 0 ? sel.narrow_groups : 1;
  if (Gn > G) Gn = G;
  for (;;) {
    if (T.rank == 0 && threadIdx.x == 0) gdb::st_rlx(T.word(TW_TASK), gdb::fadd_rlx(next_task, 1));
    if (!team_barrier(T)) return;
    const int t = gdb::ld_rlx(T.word(TW_TASK));
    if (t >= ntasks) break;
    const DevTask k = tasks[t];
    if (T.tid() == 0) work[8 * (size_t)t + 2] = gdb::now();
    int rounds = 0;
    if (k.n > 0) {
      unsigned long long* wk = work + 8 * (size_t)t;
      // Pull arms need the in-edge CSR.  The modes that choose (heuristic,
      // MLP, random) offer them to BFS only: a BFS iteration reaches the
      // same next level whichever arm runs, so the cheapest arm per
      // iteration is the cheapest schedule; an SSSP pull iteration makes
      // different progress than a push one, which a per-iteration cost
      // cannot weigh.  GDB_SEL_FIXED may still pull SSSP.
      const bool pull_ok = k.trow && k.tcol && (k.algo == GDB_TASK_BFS || (k.tw && sel.mode == GDB_SEL_FIXED));
      unsigned valid = 1u << GDB_ARM_PUSH;
      if (pull_ok) valid |= 1u << GDB_ARM_PULL;
      if (Gn < G) valid |= 1u << GDB_ARM_PUSH_NARROW;
      if (pull_ok && Gn < G) valid |= 1u << GDB_ARM_PULL_NARROW;
      const SelRun R{sel, mlp ? sel_w : nullptr, t, Gn, valid};
      // Owned solves run in the OWN instantiation only (a batch with an
      // owned task and no selector), which keeps their registers out of the
      // other kernels.
      const bool owned = OWN && k.owned && k.algo != GDB_TASK_COLOR && owned_range(k.n, G) <= 65536 &&
                         owned_lds_core(owned_range(k.n, G), min(qcap, kOwnedQueueMax)) <= region;
      if (owned)
        rounds = k.algo == GDB_TASK_BFS ? team_owned<true>(T, k, wk, region) : team_owned<false>(T, k, wk, region);
      else if (k.algo == GDB_TASK_BFS)
        rounds = k.hops > 0 ? team_worklist<true, true, false>(T, k, wk, R)
                            : team_worklist<false, false, SEL>(T, k, wk, R);
      else if (k.algo == GDB_TASK_SSSP) rounds = team_worklist<true, false, SEL>(T, k, wk, R);
      else rounds = team_color(T, k);
    }
    if (rounds < 0) return;
    if (T.tid() == 0) {
      results[t] = rounds;
      work[8 * (size_t)t + 3] = gdb::now();
    }
    if (!team_barrier(T)) return;
  }
}

}  // namespace

using namespace gdbrt;

namespace {

// A batch laid out on the device: the task table (longest first), the team
// state, counters, per-task work words and scratch, in `mem` (the context's
// workspace, or an allocation of its own for a captured graph).
struct BatchPlan {
  gdb_ctx* ctx = nullptr;
  int ntasks = 0, G = 1, groups = 0, wgs = 256, qcap = kLocalQDefault;
  size_t dyn_lds = 0;
  int region = 0;  // bytes of dynamic LDS before the MLP weights
  bool owned_kernel = false;  // launch the instantiation that runs owned solves
  gdbsel::Cfg cfg{GDB_SEL_NONE, 0, 1, 0u, nullptr, nullptr, 0, nullptr};
  std::vector<int> order;  // order[i] = caller's index of table entry i
  DevTask* d_tasks = nullptr;
  int* team_state = nullptr;
  size_t team_ints = 0;
  int* misc = nullptr;
  unsigned long long* work = nullptr;
  void* own = nullptr;  // allocation owned by a graph handle
};

int plan_batch(gdb_ctx* ctx, const gdb_task* tasks, int ntasks, int team_groups, const gdb_opts* opts,
               const gdb_selector* sel, bool own_memory, BatchPlan* P) {
  if (!ctx || ntasks <= 0 || !tasks || team_groups < 0) return GDB_E_INVALID;
  gdb_opts o{};
  if (opts) o = *opts;
  {
    // A batch with an owner-partitioned task takes the "batch_owned" row
    // when the table has one (one group per CU), else the "batch" row.
    long long nmax = 0;
    bool any_owned = false;
    for (int t = 0; t < ntasks; ++t) {
      nmax = std::max<long long>(nmax, tasks[t].n);
      any_owned |= tasks[t].owned != 0;
    }
    if (!any_owned || !apply_tuning("batch_owned", nmax, &o, &team_groups))
      apply_tuning("batch", nmax, &o, &team_groups);
  }
  P->ctx = ctx;
  P->ntasks = ntasks;
  P->wgs = default_wgs(&o);
  P->G = team_groups > 0 ? team_groups : 1;
  P->qcap = o.lds_queue > 0 ? o.lds_queue : kLocalQDefault;
  if (P->qcap < kLocalQMin || P->qcap > kLocalQMax) return GDB_E_INVALID;
  if (sel) {
    if (sel->mode < GDB_SEL_NONE || sel->mode > GDB_SEL_MLP || sel->narrow_groups < 0 ||
        sel->trace_capacity < 0 || (sel->trace && !sel->trace_count))
      return GDB_E_INVALID;
    P->cfg = gdbsel::Cfg{sel->mode, sel->fixed_arm, sel->narrow_groups, sel->seed, sel->weights,
                         sel->trace, sel->trace ? sel->trace_capacity : 0, sel->trace_count};
  }
  const gdbsel::Cfg& cfg = P->cfg;
  // The device task table lists the tasks longest first (teams take them in
  // table order, so the large ones do not start last and stretch the tail):
  // estimated cost (n + m) x {SSSP 3, BFS 2, colour 1}.
  std::vector<int>& order = P->order;
  order.resize((size_t)ntasks);
  for (int t = 0; t < ntasks; ++t) order[t] = t;
  auto cost = [&](int t) {
    const gdb_task& a = tasks[t];
    const long long f = a.algo == GDB_TASK_SSSP ? 3 : a.algo == GDB_TASK_BFS ? 2 : 1;
    return ((long long)a.n + a.m) * f;
  };
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost(x) > cost(y); });
  // Per-task scratch sizes (ints) and the device task table.
  std::vector<DevTask> dt((size_t)ntasks);
  std::vector<size_t> sints((size_t)ntasks);
  size_t scratch_total = 0;
  for (int t = 0; t < ntasks; ++t) {
    const gdb_task& a = tasks[order[t]];
    if (a.n < 0 || a.m < 0 || a.algo < GDB_TASK_BFS || a.algo > GDB_TASK_COLOR || !a.out ||
        (a.n > 0 && !a.row) || (a.m > 0 && !a.col) || (a.algo == GDB_TASK_SSSP && a.m > 0 && !a.w) ||
        (a.algo == GDB_TASK_COLOR && a.n > 0 && !a.node_value) ||
        (a.algo != GDB_TASK_COLOR && a.n > 0 && (a.source < 0 || a.source >= a.n)) || a.delta < 0 ||
        a.hops < 0 || a.hops > kMaxHops || (a.trow && !a.tcol) || a.owned < 0 || a.owned > 1 ||
        (cfg.mode != GDB_SEL_NONE && a.algo != GDB_TASK_COLOR && a.n > kNinMask))
      return GDB_E_INVALID;
    size_t ints = (size_t)(a.algo == GDB_TASK_SSSP || a.hops > 0 ? 9 : a.algo == GDB_TASK_COLOR ? 5 : 2) * a.n;
    // An owned solve keeps its boundary marks (a bit per vertex) and slot
    // records there (or runs the shared-queue engine, which needs `ints`).
    if (a.owned && a.algo != GDB_TASK_COLOR) ints = std::max(ints, owned_scratch_ints(a.n, a.algo == GDB_TASK_BFS));
    dt[t] = DevTask{a.algo, a.n, a.m, a.source, a.delta, a.hops, a.owned, a.row, a.col, a.w, a.node_value, a.out,
                    nullptr, a.trow, a.tcol, a.tw};
    sints[t] = (ints + 63) & ~(size_t)63;
    scratch_total += sints[t];
  }
  // Only local hops and owned solves use the LDS queue: without them it
  // takes no LDS (more groups fit a CU).  An owned solve takes what a CU
  // has left (one group per CU).
  bool any_hops = false, any_owned = false;
  for (int t = 0; t < ntasks; ++t) {
    any_hops |= tasks[t].algo != GDB_TASK_COLOR && tasks[t].hops > 0;
    any_owned |= tasks[t].algo != GDB_TASK_COLOR && tasks[t].owned;
  }
  // With a selector, owned tasks run the shared-queue engine.
  P->owned_kernel = any_owned && cfg.mode == GDB_SEL_NONE;
  if (!any_hops && !P->owned_kernel) P->qcap = 0;
  const size_t sel_bytes = (cfg.mode == GDB_SEL_MLP && cfg.w) ? sizeof(unsigned short) * gdbsel::kParams : 0;
  P->region = (int)(sizeof(int) * 2 * (size_t)P->qcap);
  if (P->owned_kernel) P->region = std::max<int>(P->region, (int)kOwnedLdsTotal);
  P->dyn_lds = (size_t)P->region + sel_bytes;
  int rc = gdb_dispatch_wgs(P->wgs, [&](auto W) -> int {
    constexpr int WG = decltype(W)::value;
    const void* kern = cfg.mode != GDB_SEL_NONE ? (const void*)k_batch<WG, true>
                       : P->owned_kernel        ? (const void*)k_batch<WG, false, true>
                                                : (const void*)k_batch<WG, false>;
    // Above 64 KiB of dynamic LDS a kernel has to opt in (a CU has 160 KiB).
    if (P->dyn_lds > 65536)
      GDB_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)P->dyn_lds));
    return launch_groups(ctx, kern, &o, (long long)GDB_MAX_GROUPS * WG, &P->groups, P->dyn_lds);
  });
  if (rc) return rc;
  P->team_ints = (size_t)P->groups * TW_WORDS * kLine;
  const size_t bytes = carve_bytes<DevTask>(ntasks) + carve_bytes<int>(P->team_ints) +
                       carve_bytes<int>(ntasks + 64) + carve_bytes<unsigned long long>(8 * (size_t)ntasks) +
                       carve_bytes<int>(scratch_total);
  void* base = nullptr;
  if (own_memory) {
    GDB_HIP(hipMalloc(&P->own, bytes));
    base = P->own;
  } else {
    rc = ensure_workspace(ctx, bytes);
    if (rc) return rc;
    base = ctx->ws;
  }
  Carve cv(base);
  P->d_tasks = cv.take<DevTask>(ntasks);
  P->team_state = cv.take<int>(P->team_ints);
  P->misc = cv.take<int>(ntasks + 64);
  P->work = cv.take<unsigned long long>(8 * (size_t)ntasks);
  int* scratch = cv.take<int>(scratch_total);
  size_t off = 0;
  for (int t = 0; t < ntasks; ++t) {
    dt[t].scratch = scratch + off;
    off += sints[t];
  }
  GDB_HIP(hipMemcpyAsync(P->d_tasks, dt.data(), sizeof(DevTask) * ntasks, hipMemcpyHostToDevice, ctx->stream));
  GDB_HIP(hipStreamSynchronize(ctx->stream));  // dt is a local copy
  return GDB_SUCCESS;
}

// Clears the team state, counters and work words, and the context's error
// word (kernels only, so that the sequence can be captured in a graph).
int enqueue_prep(const BatchPlan& P) {
  int rc = fill_i32(P.ctx, P.team_state, 0, (long long)P.team_ints);
  if (!rc) rc = fill_i32(P.ctx, P.misc, 0, P.ntasks + 64);
  if (!rc) rc = fill_i32(P.ctx, (int*)P.work, 0, 16 * (long long)P.ntasks);
  if (!rc) rc = prepare_launch(P.ctx);
  return rc;
}

int enqueue_kernel(const BatchPlan& P) {
  gdb_ctx* ctx = P.ctx;
  int* next_task = P.misc;
  int* teams_used = P.misc + 32;
  int* results = P.misc + 64;
  return gdb_dispatch_wgs(P.wgs, [&](auto W) -> int {
    constexpr int WG = decltype(W)::value;
    if (P.cfg.mode != GDB_SEL_NONE)
      hipLaunchKernelGGL((k_batch<WG, true>), dim3(P.groups), dim3(WG), P.dyn_lds, ctx->stream, ctx->d,
                         P.d_tasks, P.ntasks, P.G, P.team_state, next_task, results, teams_used, P.work, P.cfg,
                         P.qcap, P.region);
    else if (P.owned_kernel)
      hipLaunchKernelGGL((k_batch<WG, false, true>), dim3(P.groups), dim3(WG), P.dyn_lds, ctx->stream, ctx->d,
                         P.d_tasks, P.ntasks, P.G, P.team_state, next_task, results, teams_used, P.work, P.cfg,
                         P.qcap, P.region);
    else
      hipLaunchKernelGGL((k_batch<WG, false>), dim3(P.groups), dim3(WG), P.dyn_lds, ctx->stream, ctx->d,
                         P.d_tasks, P.ntasks, P.G, P.team_state, next_task, results, teams_used, P.work, P.cfg,
                         P.qcap, P.region);
    GDB_HIP(hipGetLastError());
    return GDB_SUCCESS;
  });
}

// Copies rounds, work and per-task times back into the caller's tasks.
int collect_batch(const BatchPlan& P, gdb_task* tasks, gdb_stats* st) {
  gdb_ctx* ctx = P.ctx;
  const int ntasks = P.ntasks;
  std::vector<int> hres((size_t)ntasks);
  std::vector<unsigned long long> hwork((size_t)8 * ntasks);
  GDB_HIP(hipMemcpy(hres.data(), P.misc + 64, sizeof(int) * ntasks, hipMemcpyDeviceToHost));
  GDB_HIP(hipMemcpy(hwork.data(), P.work, sizeof(unsigned long long) * 8 * ntasks, hipMemcpyDeviceToHost));
  for (int i = 0; i < ntasks; ++i) {
    const unsigned long long* wk = &hwork[(size_t)8 * i];
    gdb_task& tk = tasks[P.order[i]];
    tk.rounds = hres[i];
    st->rounds += hres[i];
    // Worklist tasks report items / edges processed; colour tasks count the
    // graph once.
    const bool wl = tk.algo != GDB_TASK_COLOR;
    tk.work_items = wl ? (long long)wk[0] : tk.n;
    tk.work_edges = wl ? (long long)wk[1] : tk.m;
    st->work_items += tk.work_items;
    st->work_edges += tk.work_edges;
    // Device wall clock (100 MHz) from the team's start to end of the task.
    tk.time_us = (double)(wk[3] - wk[2]) / 100.0;
    for (int a = 0; a < 4; ++a) tk.arm_iters[a] = (int)wk[4 + a];
  }
  st->launches = 1;
  int rc = read_participating(ctx, &st->participating);
  if (rc) return rc;
  return check_device_error(ctx);
}

}  // namespace

extern "C" int gdb_batch_solve_sel(gdb_ctx* ctx, gdb_task* tasks, int ntasks, int team_groups,
                                   const gdb_opts* opts, const gdb_selector* sel, gdb_stats* stats) {
  if (!ctx || ntasks < 0 || (ntasks > 0 && !tasks) || team_groups < 0) return GDB_E_INVALID;
  gdb_stats st{};
  if (ntasks == 0) {
    if (stats) *stats = st;
    return GDB_SUCCESS;
  }
  BatchPlan P;
  int rc = plan_batch(ctx, tasks, ntasks, team_groups, opts, sel, false, &P);
  if (!rc) rc = enqueue_prep(P);
  if (!rc) rc = timer_start(ctx);
  if (!rc) rc = enqueue_kernel(P);
  if (!rc) rc = timer_stop(ctx, &st.kernel_ms);
  if (rc) return rc;
  rc = collect_batch(P, tasks, &st);
  if (stats) *stats = st;
  return rc;
}

// ---- the batch as a captured HIP graph ------------------------------------
struct gdb_batch_graph {
  BatchPlan plan;
  gdb_task* tasks = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

extern "C" int gdb_batch_graph_create(gdb_ctx* ctx, gdb_task* tasks, int ntasks, int team_groups,
                                      const gdb_opts* opts, const gdb_selector* sel, gdb_batch_graph** out) {
  if (!out || !ctx || ntasks <= 0 || !tasks) return GDB_E_INVALID;
  *out = nullptr;
  gdb_batch_graph* g = new gdb_batch_graph();
  g->tasks = tasks;
  int rc = plan_batch(ctx, tasks, ntasks, team_groups, opts, sel, true, &g->plan);
  if (!rc) {
    hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
      rc = hip_fail(e, "hipStreamBeginCapture", __FILE__, __LINE__);
    } else {
      rc = enqueue_prep(g->plan);
      if (!rc) rc = enqueue_kernel(g->plan);
      hipGraph_t graph = nullptr;
      e = hipStreamEndCapture(ctx->stream, &graph);  // always end the capture
      if (!rc && e != hipSuccess) rc = hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
      g->graph = graph;
      if (!rc) {
        e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) rc = hip_fail(e, "hipGraphInstantiate", __FILE__, __LINE__);
      }
    }
  }
  if (rc) {
    gdb_batch_graph_destroy(g);
    return rc;
  }
  *out = g;
  return GDB_SUCCESS;
}

extern "C" int gdb_batch_graph_launch(gdb_batch_graph* g, gdb_stats* stats) {
  if (!g || !g->exec) return GDB_E_INVALID;
  gdb_ctx* ctx = g->plan.ctx;
  gdb_stats st{};
  int rc = timer_start(ctx);
  if (rc) return rc;
  GDB_HIP(hipGraphLaunch(g->exec, ctx->stream));
  rc = timer_stop(ctx, &st.kernel_ms);
  if (rc) return rc;
  rc = collect_batch(g->plan, g->tasks, &st);
  if (stats) *stats = st;
  return rc;
}

extern "C" int gdb_batch_graph_destroy(gdb_batch_graph* g) {
  if (!g) return GDB_SUCCESS;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->plan.own) (void)hipFree(g->plan.own);
  delete g;
  return GDB_SUCCESS;
}

extern "C" int gdb_batch_solve(gdb_ctx* ctx, gdb_task* tasks, int ntasks, int team_groups,
                               const gdb_opts* opts, gdb_stats* stats) {
  return gdb_batch_solve_sel(ctx, tasks, ntasks, team_groups, opts, nullptr, stats);
}
