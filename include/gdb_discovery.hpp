// Device-side occupancy-discovery protocol, execution-environment queries
// and inter-workgroup barriers for gfx950 (HIP).  Public header: a caller's
// persistent kernel includes it (as the reference's kernels include
// discovery.cl), takes the context's device pointer (gdb_ctx_device_ptr in
// gdb_amd.h) as an argument, and is launched with gdb_launch_on_ctx /
// gdb_time_protocol_kernel or on the context's stream by the caller:
//
//   #include "gdb_discovery.hpp"
//   __global__ void my_kernel(gdb_kernel_ctx_t* ctx, int* data) {
//     DISCOVERY_PROTOCOL(ctx);                 // non-participants return here
//     for (int i = p_get_global_id(ctx, &local_ctx); ...;
//          i += p_get_global_size(ctx, &local_ctx)) ...
//     if (!discovery_barrier(ctx, &local_ctx)) return;  // false: a wait gave up
//   }
//
// Every participating group must call discovery_barrier the same number of
// times with all its threads (it contains workgroup barriers), as in the
// reference.  The header only needs hip_runtime.h; code built from it with
// hipcc --offload-arch=gfx950 shares the context with libgdb_amd.so.
//
// Reference API kept (code/discovery_protocol/api/discovery.cl):
//   DISCOVERY_PROTOCOL(ctx)       discovery.cl:213-226 (non-participants return)
//   discovery_barrier(ctx, &lctx) discovery.cl:209
//   p_get_num_groups / p_get_group_id / p_get_global_id / p_get_global_size
//                                 discovery.cl:112-126
//   is_participating              discovery.cl:106-108
//
// What changes for gfx950:
//   * All cross-group words are agent-scope __hip_atomic operations; release
//     and acquire are agent-scope fences, which the compiler lowers to the
//     L2 write-back / invalidate that makes data written on one XCD visible
//     on another (each XCD has its own L2).
//   * The default protocol needs no mutex: one atomic add on a ticket word
//     whose bit 30 means "poll closed".  A group that adds while the bit is
//     clear participates, with id = the old arrival count; the first group to
//     set the bit publishes P.  Before closing, a participant waits until the
//     expected number of groups (the host's occupancy bound, clamped to the
//     launch) has arrived or no group has arrived for a quiet period, so the
//     close does not race the dispatcher.  The reference's ticket and spin
//     locks (locks.cl) are kept as selectable variants.
//   * Every wait sleeps between polls (s_sleep) and is bounded by a wall-clock
//     deadline; on expiry the context's sticky error word is set and the wait
//     gives up, so a mis-sized launch ends instead of hanging the device.
#pragma once
#include <hip/hip_runtime.h>
#include <climits>
#include "gdb_ctx_layout.h"

#define GDB_POLL_CLOSED (1 << 30)
#define GDB_POLL_MASK (GDB_POLL_CLOSED - 1)
#define GDB_GEN_UNKNOWN INT_MIN

// Per-group execution environment, kept in LDS.  Counterpart of
// discovery_local_ctx (discovery.cl:19-23) plus the barrier flavour, the
// barrier generation the group expects next, and a flag raised when a wait
// gave up.
struct gdb_local_ctx {
  int participating_group_id;
  int participating_group_size;
  int is_participating;
  int bar_kind;   // resolved barrier flavour (never GDB_BAR_AUTO)
  int bar_gen;    // generation of the global counter this group expects next
  int aborted;
  int xcd;        // XCD this group runs on (hardware register)
};

namespace gdb {

__device__ __forceinline__ int ld_rlx(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int ld_acq(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rlx(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rel(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int fadd(int* p, int v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int fadd_rlx(int* p, int v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int f_or(int* p, int v) {
  return __hip_atomic_fetch_or(p, v, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int xchg(int* p, int v) {
  return __hip_atomic_exchange(p, v, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void fence_acq() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }
__device__ __forceinline__ void fence_rel() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }
// Waits until this wave's earlier stores have completed at L2.  A
// workgroup-scope fence or __syncthreads() does not wait for them on gfx950,
// so every wave runs this before a barrier's arrival is signalled.
__device__ __forceinline__ void drain_stores() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }
__device__ __forceinline__ int xcc_id() {
  unsigned x;
  asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(x));
  return (int)(x & (GDB_MAX_XCDS - 1));
}
// 100 MHz constant-rate wall clock.
__device__ __forceinline__ unsigned long long now() { return __builtin_amdgcn_s_memrealtime(); }

__device__ __forceinline__ void set_error(gdb_kernel_ctx_t* ctx, int code) {
  int expected = GDB_OK;
  __hip_atomic_compare_exchange_strong(&ctx->error.v, &expected, code, __ATOMIC_RELAXED,
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Bounded wait: spins while pred() holds.  Returns true once pred() is false;
// false when the deadline passes (raising err_code) or when another group has
// already raised an error.
template <typename Pred>
__device__ __forceinline__ bool wait_while(gdb_kernel_ctx_t* ctx, Pred pred, int err_code) {
  if (!pred()) return true;
  const unsigned long long limit = (unsigned long long)(unsigned)ld_rlx(&ctx->timeout_ticks.v);
  if (limit == 0) {  // zero bound: any wait that would block fails at once
    set_error(ctx, err_code);
    return false;
  }
  const unsigned long long t0 = now();
  unsigned it = 0;
  while (pred()) {
    __builtin_amdgcn_s_sleep(1);
    if ((++it & 31u) == 0) {
      if (ld_rlx(&ctx->error.v) != GDB_OK) return false;
      if (now() - t0 > limit) {
        set_error(ctx, err_code);
        return false;
      }
    }
  }
  return true;
}

// Bounded wait for long idle periods (a group parked while the rest of its
// team works): after 32 quick polls it polls about once a microsecond, so
// that many parked groups do not load the memory system the active groups
// use, while a wake-up still takes about a microsecond.
template <typename Pred>
__device__ __forceinline__ bool wait_while_idle(gdb_kernel_ctx_t* ctx, Pred pred, int err_code) {
  if (!pred()) return true;
  const unsigned long long limit = (unsigned long long)(unsigned)ld_rlx(&ctx->timeout_ticks.v);
  if (limit == 0) {
    set_error(ctx, err_code);
    return false;
  }
  const unsigned long long t0 = now();
  unsigned it = 0;
  while (pred()) {
    if (it < 32) __builtin_amdgcn_s_sleep(2);
    else __builtin_amdgcn_s_sleep(32);  // 32 x 64 cycles
    if ((++it & 7u) == 0) {
      if (ld_rlx(&ctx->error.v) != GDB_OK) return false;
      if (now() - t0 > limit) {
        set_error(ctx, err_code);
        return false;
      }
    }
  }
  return true;
}

// ---- mutex variants of the reference protocol (locks.cl) -----------------
__device__ inline bool lock_acquire(gdb_kernel_ctx_t* ctx, int kind) {
  if (kind == GDB_LOCK_SPIN)  // locks.cl:11-15
    return wait_while(ctx, [&] { return xchg(&ctx->lock_counter.v, 1) == 1; },
                      GDB_ERR_DISCOVERY_TIMEOUT);
  const int ticket = fadd(&ctx->lock_counter.v, 1);  // locks.cl:25-28
  return wait_while(ctx, [&] { return ld_acq(&ctx->lock_now_serving.v) != ticket; },
                    GDB_ERR_DISCOVERY_TIMEOUT);
}

__device__ inline void lock_release(gdb_kernel_ctx_t* ctx, int kind) {
  if (kind == GDB_LOCK_SPIN)
    st_rel(&ctx->lock_counter.v, 0);  // locks.cl:17-20
  else
    st_rel(&ctx->lock_now_serving.v, ld_rlx(&ctx->lock_now_serving.v) + 1);  // locks.cl:30-34
}

// reset_kernel_context (discovery.cl:31-37).  Run by the last launched
// group through the closing phase; the release store of the lock word is
// also the unlock.
__device__ inline void reset_locked_protocol(gdb_kernel_ctx_t* ctx) {
  st_rlx(&ctx->prot_poll_open.v, 1);
  st_rlx(&ctx->prot_counter.v, 0);
  st_rlx(&ctx->kernel_counter.v, 0);
  st_rlx(&ctx->lock_now_serving.v, 0);
  st_rel(&ctx->lock_counter.v, 0);
}

// discovery_protocol_master (discovery.cl:42-88) with a mutex.
__device__ inline void protocol_locked(gdb_kernel_ctx_t* ctx, gdb_local_ctx* l, int kind) {
  const int total = (int)gridDim.x;
  l->is_participating = 0;
  // Polling phase.
  if (!lock_acquire(ctx, kind)) return;
  if (ld_rlx(&ctx->prot_poll_open.v) && ld_rlx(&ctx->prot_counter.v) < GDB_MAX_GROUPS) {
    const int id = ld_rlx(&ctx->prot_counter.v);
    l->is_participating = 1;
    l->participating_group_id = id;
    st_rlx(&ctx->prot_counter.v, id + 1);
  }
  lock_release(ctx, kind);
  // Closing phase.
  if (!lock_acquire(ctx, kind)) {
    l->is_participating = 0;
    return;
  }
  const int kc = ld_rlx(&ctx->kernel_counter.v) + 1;
  st_rlx(&ctx->kernel_counter.v, kc);
  if (ld_rlx(&ctx->prot_poll_open.v)) {
    st_rlx(&ctx->prot_poll_open.v, 0);
    st_rlx(&ctx->num_participating.v, ld_rlx(&ctx->prot_counter.v));
  }
  l->participating_group_size = ld_rlx(&ctx->prot_counter.v);
  if (kc == total)
    reset_locked_protocol(ctx);
  else
    lock_release(ctx, kind);
}

// Lock-free protocol: same two phases, one atomic each.  With xcd_only >= 0
// groups on other XCDs do not poll and never participate.
__device__ inline void protocol_lockfree(gdb_kernel_ctx_t* ctx, gdb_local_ctx* l) {
  const int total = (int)gridDim.x;
  const int xo = ld_rlx(&ctx->xcd_only.v);
  const bool eligible = xo < 0 || l->xcd == xo;
  int old = GDB_POLL_CLOSED;
  if (eligible) old = fadd(&ctx->poll.v, 1);
  const bool part = (old & GDB_POLL_CLOSED) == 0;
  if (part) {
    // Let co-resident groups that are still being dispatched arrive before
    // closing: stop at the expected count or after a quiet period.
    int expect = ld_rlx(&ctx->expect.v);
    if (expect <= 0 || expect > total) expect = total;
    if (xo >= 0) expect = (expect + 7 - xo) / 8;  // groups are dealt round-robin to 8 XCDs
    if (expect < 1) expect = 1;
    if (expect > GDB_MAX_GROUPS) expect = GDB_MAX_GROUPS;
    const unsigned long long quiet = (unsigned)ld_rlx(&ctx->quiet_ticks.v);
    const unsigned long long limit = (unsigned)ld_rlx(&ctx->timeout_ticks.v);
    int last = old + 1;
    unsigned long long t_last = now(), t_start = t_last;
    for (;;) {
      const int cur = ld_rlx(&ctx->poll.v);
      if (cur & GDB_POLL_CLOSED) break;
      if ((cur & GDB_POLL_MASK) >= expect) break;
      const unsigned long long t = now();
      if (cur != last) {
        last = cur;
        t_last = t;
      } else if (t - t_last > quiet || t - t_start > limit) {
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
  }
  // Closing phase: the first group to set the closed bit publishes P.
  int P = 0;
  if (eligible) {
    const int o2 = f_or(&ctx->poll.v, GDB_POLL_CLOSED);
    if ((o2 & GDB_POLL_CLOSED) == 0) {
      P = o2 & GDB_POLL_MASK;
      if (P > GDB_MAX_GROUPS) P = GDB_MAX_GROUPS;
      st_rlx(&ctx->num_participating.v, P);
      st_rel(&ctx->closing_p.v, P);
    } else if (part) {
      if (wait_while(ctx, [&] { return ld_acq(&ctx->closing_p.v) == 0; }, GDB_ERR_DISCOVERY_TIMEOUT))
        P = ld_acq(&ctx->closing_p.v);
    }
  }
  l->participating_group_id = old & GDB_POLL_MASK;
  l->is_participating = part && P > 0 && l->participating_group_id < P;
  l->participating_group_size = P;
  // The last launched group resets the words for the next launch.
  const int kc = fadd(&ctx->kernel_counter.v, 1);
  if (kc == total - 1) {
    st_rlx(&ctx->poll.v, 0);
    st_rlx(&ctx->closing_p.v, 0);
    st_rel(&ctx->kernel_counter.v, 0);
  }
}

// discovery_protocol (discovery.cl:91-102) plus the skip path of the
// DISCOVERY_PROTOCOL macro (discovery.cl:215-220).  Returns whether this
// group participates.
__device__ inline bool discovery_protocol(gdb_kernel_ctx_t* ctx, gdb_local_ctx* l) {
  if (threadIdx.x == 0) {
    l->aborted = 0;
    l->bar_gen = GDB_GEN_UNKNOWN;
    l->xcd = xcc_id();
    int kind = ld_rlx(&ctx->bar_kind.v);
    const bool skip = ld_rlx(&ctx->skip.v) != 0;
    if (skip) {
      l->participating_group_size = (int)gridDim.x;
      l->is_participating = 1;
      l->participating_group_id = (int)blockIdx.x;
      st_rlx(&ctx->num_participating.v, (int)gridDim.x);
    } else {
      const int lk = ld_rlx(&ctx->lock_kind.v);
      if (lk == GDB_LOCK_FREE)
        protocol_lockfree(ctx, l);
      else
        protocol_locked(ctx, l, lk);
    }
    // Resolve the barrier flavour.
    if (kind == GDB_BAR_AUTO)
      kind = l->participating_group_size >= GDB_AUTO_TREE_MIN_P ? GDB_BAR_TREE : GDB_BAR_COUNTER;
    if (kind != GDB_BAR_XF && kind != GDB_BAR_TREE) kind = GDB_BAR_COUNTER;
    l->bar_kind = kind;
  }
  __syncthreads();
  return l->is_participating != 0;
}

// ---- barriers ---------------------------------------------------------------
// Each returns false when a wait gave up (the caller should leave the kernel).

// XF barrier (discovery.cl:129-190): group 0 is the master; master thread t
// waits for the flags of slaves t+1, t+1+L, ...; then releases them all.
// Each flag lives on its own 128 B line.
__device__ inline bool xf_barrier(gdb_kernel_ctx_t* ctx, gdb_local_ctx* l) {
  const int id = l->participating_group_id;
  const int P = l->participating_group_size;
  drain_stores();
  __syncthreads();
  if (id == 0) {
    bool ok = true;
    for (int peer = threadIdx.x + 1; peer < P && ok; peer += blockDim.x) {
      const int* f = &ctx->bar_flags[peer].v;
      ok = wait_while(ctx, [&] { return ld_rlx(f) == 0; }, GDB_ERR_BARRIER_TIMEOUT);
    }
    if (!ok) l->aborted = 1;
    __syncthreads();
    fence_acq();
    fence_rel();
    for (int peer = threadIdx.x + 1; peer < P; peer += blockDim.x)
      st_rlx(&ctx->bar_flags[peer].v, 0);
  } else {
    if (threadIdx.x == 0) {
      int* f = &ctx->bar_flags[id].v;
      st_rel(f, 1);
      if (!wait_while(ctx, [&] { return ld_rlx(f) == 1; }, GDB_ERR_BARRIER_TIMEOUT))
        l->aborted = 1;
      fence_acq();
    }
  }
  __syncthreads();
  return l->aborted == 0;
}

// Sense-reversing arrival counter over `members` groups, run by one thread
// of each group: add 1 (release+acquire); the last arriver zeroes the count
// and bumps the generation the others poll.  `gen_cache` holds the
// generation this group expects, or GDB_GEN_UNKNOWN before its first use.
__device__ inline bool counter_wait(gdb_kernel_ctx_t* ctx, int* count, int* gen, int members,
                                    int* gen_cache, int* aborted) {
  int g = *gen_cache;
  if (g == GDB_GEN_UNKNOWN) g = ld_acq(gen);
  bool ok = true;
  const int old = fadd(count, 1);
  if (old == members - 1) {
    st_rlx(count, 0);
    st_rel(gen, g + 1);
  } else {
    ok = wait_while(ctx, [&] { return ld_rlx(gen) == g; }, GDB_ERR_BARRIER_TIMEOUT);
    fence_acq();
  }
  *gen_cache = g + 1;
  if (!ok) *aborted = 1;
  return ok;
}

// One group leader's agent-scope fences serve its whole group: the waves of
// a group share one CU and so one L1, which the acquire invalidates.
__device__ inline bool counter_barrier(gdb_kernel_ctx_t* ctx, gdb_local_ctx* l) {
  drain_stores();
  __syncthreads();
  if (threadIdx.x == 0)
    counter_wait(ctx, &ctx->bar_count.v, &ctx->bar_gen.v, l->participating_group_size, &l->bar_gen,
                 &l->aborted);
  __syncthreads();
  return l->aborted == 0;
}

// Two-level counter: groups arrive on their cluster's counter (clusters of
// consecutive participating ids, at most GDB_TREE_CLUSTERS of them, each counter
// on its own line); a cluster's last arriver arrives on the global counter;
// the last global arriver bumps the generation that every leader polls.
// Every arrival is an agent-scope acq_rel read-modify-write, so each
// group's writes reach the generation bump through a chain of
// release/acquire pairs.  Contention on any one counter drops from P to
// about P / GDB_TREE_CLUSTERS.
__device__ inline bool tree_barrier(gdb_kernel_ctx_t* ctx, gdb_local_ctx* l) {
  drain_stores();
  __syncthreads();
  if (threadIdx.x == 0) {
    const int id = l->participating_group_id;
    const int P = l->participating_group_size;
    const int C = (P + GDB_TREE_CLUSTERS - 1) / GDB_TREE_CLUSTERS;
    const int c = id / C;
    const int members = min(C, P - c * C);
    const int nclus = (P + C - 1) / C;
    int g = l->bar_gen;
    if (g == GDB_GEN_UNKNOWN) g = ld_acq(&ctx->bar_gen.v);
    bool ok = true, waiter = true;
    const int old = fadd(&ctx->cluster_count[c].v, 1);
    if (old == members - 1) {
      st_rlx(&ctx->cluster_count[c].v, 0);
      const int o2 = fadd(&ctx->bar_count.v, 1);
      if (o2 == nclus - 1) {
        st_rlx(&ctx->bar_count.v, 0);
        st_rel(&ctx->bar_gen.v, g + 1);
        waiter = false;
      }
    }
    if (waiter) {
      ok = wait_while(ctx, [&] { return ld_rlx(&ctx->bar_gen.v) == g; }, GDB_ERR_BARRIER_TIMEOUT);
      fence_acq();
    }
    l->bar_gen = g + 1;
    if (!ok) l->aborted = 1;
  }
  __syncthreads();
  return l->aborted == 0;
}

__device__ inline bool barrier(gdb_kernel_ctx_t* ctx, gdb_local_ctx* l) {
  if (l->participating_group_size <= 1) {
    drain_stores();
    __syncthreads();
    return l->aborted == 0;
  }
  switch (l->bar_kind) {
    case GDB_BAR_XF: return xf_barrier(ctx, l);
    case GDB_BAR_TREE: return tree_barrier(ctx, l);
    default: return counter_barrier(ctx, l);
  }
}

}  // namespace gdb

// ---- reference-named execution-environment API (discovery.cl:106-126) -----
__device__ __forceinline__ int is_participating(gdb_kernel_ctx_t*, gdb_local_ctx* l) {
  return l->is_participating;
}
__device__ __forceinline__ int p_get_num_groups(gdb_kernel_ctx_t*, gdb_local_ctx* l) {
  return l->participating_group_size;
}
__device__ __forceinline__ int p_get_group_id(gdb_kernel_ctx_t*, gdb_local_ctx* l) {
  return l->participating_group_id;
}
__device__ __forceinline__ int p_get_global_id(gdb_kernel_ctx_t*, gdb_local_ctx* l) {
  return (int)threadIdx.x + (int)blockDim.x * l->participating_group_id;
}
__device__ __forceinline__ int p_get_global_size(gdb_kernel_ctx_t*, gdb_local_ctx* l) {
  return l->participating_group_size * (int)blockDim.x;
}

// discovery_barrier (discovery.cl:209).  Returns false when the barrier gave
// up (peers never arrived); the caller should then leave the kernel.
#define discovery_barrier(gl_ctx, lctx) gdb::barrier((gl_ctx), (lctx))

// DISCOVERY_PROTOCOL (discovery.cl:213-226): declares `local_ctx` in LDS and
// makes non-participating groups return.
#define DISCOVERY_PROTOCOL(gl_ctx)          \
  __shared__ gdb_local_ctx local_ctx;       \
  if (!gdb::discovery_protocol((gl_ctx), &local_ctx)) return;
