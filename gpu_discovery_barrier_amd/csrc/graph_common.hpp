// Wave64 helpers shared by the graph kernels.
#pragma once
#include <hip/hip_runtime.h>
#include "gdb_discovery.hpp"

namespace gdb {

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Value of `v` in lane `src` (wave-uniform) for every lane: a scalar read of
// the register, no LDS round trip.
__device__ __forceinline__ int bcast(int v, int src) { return __builtin_amdgcn_readlane(v, src); }

// Wave-wide max over the active lanes, by DPP row shifts (no LDS round
// trips), from the device library.
extern "C" __device__ int __ockl_wfred_max_i32(int);
__device__ __forceinline__ int wave_max(int v) { return __ockl_wfred_max_i32(v); }

__device__ __forceinline__ float shfl_xor(float x, int o) { return __shfl_xor(x, o, 64); }
__device__ __forceinline__ int shfl_xor(int x, int o) { return __shfl_xor(x, o, 64); }
__device__ __forceinline__ unsigned shfl_xor(unsigned x, int o) {
  return (unsigned)__shfl_xor((int)x, o, 64);
}

// Global-address-space accessors.  Graph arrays reach the kernels through
// structs, so the compiler sees generic pointers and emits flat instructions,
// whose waits also cover LDS traffic; these keep the accesses global_*.
template <typename T>
__device__ __forceinline__ T ldg(const T* p) {
  return *(const __attribute__((address_space(1))) T*)p;
}
__device__ __forceinline__ unsigned atomic_min_g(unsigned* p, unsigned v) {
  return __hip_atomic_fetch_min((__attribute__((address_space(1))) unsigned*)p, v, __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long atomic_min_g(unsigned long long* p, unsigned long long v) {
  return __hip_atomic_fetch_min((__attribute__((address_space(1))) unsigned long long*)p, v, __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int atomic_or_g(int* p, int v) {
  return __hip_atomic_fetch_or((__attribute__((address_space(1))) int*)p, v, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int atomic_max_g(int* p, int v) {
  return __hip_atomic_fetch_max((__attribute__((address_space(1))) int*)p, v, __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned atomic_cas_g(unsigned* p, unsigned expected, unsigned desired) {
  __hip_atomic_compare_exchange_strong((__attribute__((address_space(1))) unsigned*)p, &expected, desired,
                                       __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return expected;
}
__device__ __forceinline__ int atomic_cas_g(int* p, int expected, int desired) {
  __hip_atomic_compare_exchange_strong((__attribute__((address_space(1))) int*)p, &expected, desired,
                                       __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return expected;
}

struct MaxFound {  // running max plus "at least one candidate seen"
  float mx;
  int found;
};
__device__ __forceinline__ MaxFound shfl_xor(MaxFound x, int o) {
  return MaxFound{__shfl_xor(x.mx, o, 64), __shfl_xor(x.found, o, 64)};
}
struct MaxAt {  // running max and the candidate holding it (at = -1: none seen)
  float mx;
  int at;
};
__device__ __forceinline__ MaxAt shfl_xor(MaxAt x, int o) {
  return MaxAt{__shfl_xor(x.mx, o, 64), __shfl_xor(x.at, o, 64)};
}

// Thread / wave cooperative reduction over the edge ranges [b, e) held by
// the lanes of one wave.  Lanes whose range is at most `serial_limit` edges
// walk it alone; longer ranges are walked by the whole wave (64 consecutive
// edges per step, coalesced) and reduced with xor-shuffles.  `edge(k)` maps
// an edge index to a value, `red` combines two values.  All 64 lanes must
// call this (wave-uniform control flow).
template <typename T, typename EdgeFn, typename RedFn>
__device__ __forceinline__ T wave_coop_reduce(bool active, int b, int e, T identity, EdgeFn edge,
                                              RedFn red, int serial_limit) {
  T acc = identity;
  const bool heavy = active && (e - b) > serial_limit;
  // Unrolled so that the loads of several edges are in flight together.
  if (active && !heavy) {
#pragma unroll 4
    for (int k = b; k < e; ++k) acc = red(acc, edge(k));
  }
  unsigned long long mask = __ballot(heavy);
  const int lane = lane_id();
  while (mask) {
    const int leader = __ffsll((long long)mask) - 1;
    const int lb = bcast(b, leader);
    const int le = bcast(e, leader);
    T part = identity;
#pragma unroll 4
    for (int k = lb + lane; k < le; k += 64) part = red(part, edge(k));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) part = red(part, shfl_xor(part, off));
    if (lane == leader) acc = part;
    mask &= mask - 1;
  }
  return acc;
}

// Thread / wave cooperative walk over the edge ranges [b, e) of the lanes of
// one wave, calling fn(k, valid, payload) in wave-uniform steps so that fn
// may use wave collectives (ballot-based queue pushes).  `payload` is the
// value the range's owner lane passed in (for example its vertex's distance):
// a lane walking another lane's range sees the owner's payload.  Short ranges
// are walked by their own lane (trip count = the longest short range in the
// wave); long ranges by all 64 lanes, 64 consecutive edges per step.  Lanes
// without an edge in a step get valid = false.
template <typename F>
__device__ __forceinline__ void wave_coop_foreach(bool active, int b, int e, unsigned payload, F fn,
                                                  int serial_limit) {
  const bool heavy = active && (e - b) > serial_limit;
  int len = (active && !heavy) ? (e - b) : 0;
  const int maxlen = wave_max(len);
  for (int t = 0; t < maxlen; ++t) fn(b + t, t < len, payload);
  unsigned long long mask = __ballot(heavy);
  const int lane = lane_id();
  while (mask) {
    const int leader = __ffsll((long long)mask) - 1;
    const int lb = bcast(b, leader);
    const int le = bcast(e, leader);
    const unsigned lp = (unsigned)bcast((int)payload, leader);
    for (int base = lb; base < le; base += 64) fn(base + lane, base + lane < le, lp);
    mask &= mask - 1;
  }
}

// Wave cooperative "some edge k in [b, e) has pred(k)" with early exit:
// short ranges are scanned by their own lane up to the first hit, long ones
// by the whole wave, 64 consecutive edges per step, up to the first step
// with a hit.  All 64 lanes must call it.
template <typename P>
__device__ __forceinline__ bool wave_coop_any(bool active, int b, int e, P pred, int serial_limit) {
  const bool heavy = active && (e - b) > serial_limit;
  bool found = false;
  if (active && !heavy)
    for (int k = b; k < e; ++k)
      if (pred(k)) {
        found = true;
        break;
      }
  unsigned long long mask = __ballot(heavy);
  const int lane = lane_id();
  while (mask) {
    const int leader = __ffsll((long long)mask) - 1;
    const int lb = bcast(b, leader);
    const int le = bcast(e, leader);
    bool hit = false;
    for (int base = lb; base < le; base += 64) {
      const bool h = base + lane < le && pred(base + lane);
      if (__ballot(h) != 0ull) {
        hit = true;
        break;
      }
    }
    if (lane == leader) found = hit;
    mask &= mask - 1;
  }
  return found;
}

// Appends `item` for every lane with push == true to queue[*tail ...]:
// one atomic per wave.  Returns nothing; all 64 lanes must call it.
__device__ __forceinline__ void wave_push(bool push, int item, int* queue, int* tail) {
  const unsigned long long m = __ballot(push);
  if (m == 0ull) return;
  const int lane = lane_id();
  int base = 0;
  if (lane == __ffsll((long long)m) - 1) base = fadd_rlx(tail, __popcll(m));
  base = bcast(base, __ffsll((long long)m) - 1);
  if (push) queue[base + __popcll(m & ((1ull << lane) - 1ull))] = item;
}

// Appends the items of N push slots of every lane with one atomic per wave.
template <int N>
__device__ __forceinline__ void wave_push_n(const bool (&push)[N], const int (&item)[N], int* queue,
                                            int* tail) {
  unsigned long long m[N];
  int cnt = 0;
#pragma unroll
  for (int j = 0; j < N; ++j) {
    m[j] = __ballot(push[j]);
    cnt += __popcll(m[j]);
  }
  if (cnt == 0) return;
  const int lane = lane_id();
  int base = 0;
  if (lane == 0) base = fadd_rlx(tail, cnt);
  base = bcast(base, 0);
  const unsigned long long lt = (1ull << lane) - 1ull;
#pragma unroll
  for (int j = 0; j < N; ++j) {
    if (push[j]) queue[base + __popcll(m[j] & lt)] = item[j];
    base += __popcll(m[j]);
  }
}


// Wave-wide OR of a predicate.
__device__ __forceinline__ bool wave_any(bool p) { return __ballot(p) != 0ull; }

// Whether a lane has any edge slot in a walk step (lanes without one can
// skip the step's loads).
template <int N>
__device__ __forceinline__ bool any_slot(const bool (&ok)[N]) {
  bool a = false;
#pragma unroll
  for (int j = 0; j < N; ++j) a |= ok[j];
  return a;
}

// Chunked cooperative walk over the edge ranges [b, e) of the lanes of one
// wave: step(ok, kk) handles N edge slots of the calling lane at once (slot j
// is edge kk[j] when ok[j]) and returns the lane's contribution; it is called
// in wave-uniform steps, so it may use wave collectives, and it should issue
// the slots' loads and atomics branch-free, all loads before any use, so that
// the N slots cost one chain of memory round trips.  Short ranges are walked
// by their own lane N edges per step; long ranges by the whole wave, N x 64
// consecutive edges per step (slot j of lane i takes edge base + i + 64 j, so
// each slot's loads are coalesced), and reduced into the owner lane.
// As wave_walk_n, but step(ok, kk, owner) also gets the `payload` of the
// range's owner lane (for a long row walked by the whole wave, the leader's
// payload on every lane).
template <int N, typename T, typename Step, typename R>
__device__ __forceinline__ T wave_walk_np(bool active, int b, int e, int payload, T identity, Step step, R red,
                                          int serial_limit) {
  const bool heavy = active && (e - b) > serial_limit;
  const int len = (active && !heavy) ? (e - b) : 0;
  const int maxlen = wave_max(len);
  T acc = identity;
  for (int c = 0; c < maxlen; c += N) {
    bool ok[N];
    int kk[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
      ok[j] = c + j < len;
      kk[j] = b + c + j;
    }
    acc = red(acc, step(ok, kk, payload));
  }
  unsigned long long mask = __ballot(heavy);
  const int lane = lane_id();
  while (mask) {
    const int leader = __ffsll((long long)mask) - 1;
    const int lb = bcast(b, leader);
    const int le = bcast(e, leader);
    const int lp = bcast(payload, leader);
    T part = identity;
    for (int base = lb; base < le; base += 64 * N) {
      bool ok[N];
      int kk[N];
#pragma unroll
      for (int j = 0; j < N; ++j) {
        kk[j] = base + lane + 64 * j;
        ok[j] = kk[j] < le;
      }
      part = red(part, step(ok, kk, lp));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) part = red(part, shfl_xor(part, off));
    if (lane == leader) acc = part;
    mask &= mask - 1;
  }
  return acc;
}

// Chunked cooperative "some edge of [b, e) has pred" with early exit:
// step(ok, kk, owner) tests N edge slots of the calling lane at once (the
// same slot layout and payload as wave_walk_np) and returns whether one of
// them hit.  A lane walking its own short range stops at its first step with
// a hit; a long range walked by the whole wave stops at the first step in
// which some lane hit.  All 64 lanes must call it.
template <int N, typename Step>
__device__ __forceinline__ bool wave_walk_any(bool active, int b, int e, int payload, Step step,
                                              int serial_limit) {
  const bool heavy = active && (e - b) > serial_limit;
  const int len = (active && !heavy) ? (e - b) : 0;
  bool found = false;
  for (int c = 0; __ballot(c < len && !found) != 0ull; c += N) {
    bool ok[N];
    int kk[N];
    const bool live = c < len && !found;
#pragma unroll
    for (int j = 0; j < N; ++j) {
      ok[j] = live && c + j < len;
      kk[j] = b + c + j;
    }
    found |= step(ok, kk, payload);
  }
  unsigned long long mask = __ballot(heavy);
  const int lane = lane_id();
  while (mask) {
    const int leader = __ffsll((long long)mask) - 1;
    const int lb = bcast(b, leader);
    const int le = bcast(e, leader);
    const int lp = bcast(payload, leader);
    bool hit = false;
    for (int base = lb; base < le; base += 64 * N) {
      bool ok[N];
      int kk[N];
#pragma unroll
      for (int j = 0; j < N; ++j) {
        kk[j] = base + lane + 64 * j;
        ok[j] = kk[j] < le;
      }
      if (__ballot(step(ok, kk, lp)) != 0ull) {
        hit = true;
        break;
      }
    }
    if (lane == leader) found = hit;
    mask &= mask - 1;
  }
  return found;
}

template <int N, typename T, typename Step, typename R>
__device__ __forceinline__ T wave_walk_n(bool active, int b, int e, T identity, Step step, R red,
                                         int serial_limit) {
  return wave_walk_np<N>(
      active, b, e, 0, identity, [&](const bool (&ok)[N], const int (&kk)[N], int) { return step(ok, kk); }, red,
      serial_limit);
}

// Uniform-trip grid-stride loop over [0, n): every thread of a group runs
// the same number of iterations (i may be >= n in the last one), so the body
// may use wave- and group-wide collectives.
#define GDB_FOR_UNIFORM(i, valid, start, stride, n)                             \
  for (long long _base = (long long)(start) - (long long)threadIdx.x;           \
       _base < (long long)(n); _base += (long long)(stride))                      \
    for (int _once = 1, i = (int)(_base + threadIdx.x),                          \
             valid = (_base + (long long)threadIdx.x) < (long long)(n);          \
         _once; _once = 0)

}  // namespace gdb
