// Per-iteration schedule selector of the batched BFS / SSSP engine: the
// iteration features, the 16-64-64-4 MLP evaluated by one wave (bf16
// weights, fp32 accumulation), the plain heuristic it falls back to, the
// random arm used to collect traces, and the trace records.
//
// One wave of the group that completes a team barrier runs it; inputs are
// uniform across that wave, so every lane returns the same arm.
#pragma once
#include <hip/hip_runtime.h>
#include "../../include/gdb_amd.h"

namespace gdbsel {

constexpr int kIn = GDB_MLP_IN, kHid = GDB_MLP_HIDDEN, kOut = GDB_MLP_OUT;
// Input-major parameter layout: layer output j = b[j] + sum_i W[i][j] x[i],
// so lane j of a wave reads W[i][j] at consecutive addresses.
constexpr int kW1 = 0, kB1 = kW1 + kIn * kHid, kW2 = kB1 + kHid, kB2 = kW2 + kHid * kHid,
              kW3 = kB2 + kHid, kB3 = kW3 + kHid * kOut, kParams = kB3 + kOut;
static_assert(kParams == GDB_MLP_PARAMS, "parameter count");
constexpr int kStride = GDB_TRACE_STRIDE;
constexpr float kSwitchMargin = 0.15f;  // log2 units

struct Cfg {
  int mode, fixed_arm, narrow_groups;
  unsigned seed;
  const unsigned short* w;
  float* trace;
  int trace_capacity;
  int* trace_count;
};

__device__ __forceinline__ float bf16(unsigned short b) { return __uint_as_float((unsigned)b << 16); }

// Round to nearest even (NaN stays NaN).
__device__ __forceinline__ unsigned short to_bf16(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (unsigned short)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

__device__ __forceinline__ bool arm_pull(int a) { return a == GDB_ARM_PULL || a == GDB_ARM_PULL_NARROW; }
__device__ __forceinline__ bool arm_narrow(int a) { return a >= GDB_ARM_PUSH_NARROW; }

// Iteration features (all lanes compute the same values).  They use only
// counts every group of the team already holds: the next frontier's size,
// the running sum of frontier sizes (for BFS, the vertices reached), the
// graph's size and the team's shape; frontier edges are estimated with the
// average degree.  They leave out the previous arm: the network predicts
// the steady-state cost of each arm (it is trained on iterations that kept
// the arm of the one before).
struct FeatIn {
  float n, m;           // graph size
  float nin;            // next frontier size
  float total;          // sum of frontier sizes including the next one
  float prev_nin;       // frontier of the iteration that just ended
  int it;               // index of the next iteration
  int algo, prev_arm;
  float team_threads, narrow_threads;
};

__device__ __forceinline__ float lg(float x) { return __log2f(1.f + fmaxf(x, 0.f)) * (1.f / 32.f); }

__device__ __forceinline__ float feature(const FeatIn& a, int i) {
  const float deg = a.m / fmaxf(a.n, 1.f);
  const float unv = fmaxf(a.n - a.total, 0.f);
  switch (i) {
    case 0: return lg(a.nin);
    case 1: return lg(a.nin * deg);
    case 2: return lg(unv);
    case 3: return lg(unv * deg);
    case 4: return lg(a.n);
    case 5: return lg(a.m);
    case 6: return a.nin / fmaxf(a.n, 1.f);
    case 7: return fminf(a.total / fmaxf(a.n, 1.f), 4.f) * 0.25f;
    case 8: return unv / fmaxf(a.n, 1.f);
    case 9: return (float)a.it / ((float)a.it + 64.f);
    case 10: return (float)a.algo;
    case 11: return lg(a.nin * deg / fmaxf(a.team_threads, 1.f) * 64.f);
    case 12: return lg(a.nin * deg / fmaxf(a.narrow_threads, 1.f) * 64.f);
    case 13: return lg(a.prev_nin);
    case 14: return lg(a.team_threads);
    default: return lg(a.narrow_threads);
  }
}

// Broadcast of lane `i` (a compile-time constant after unrolling) to the
// wave through a scalar register.
__device__ __forceinline__ float bcast(float v, int i) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), i));
}

// y = MLP(x) for the features held by lanes 0..15 (`xi`); every lane of the
// wave must call it and gets all four outputs.  Lane j computes hidden unit j
// of both layers (its weights W[i][j] are consecutive across lanes), the
// activations travel by readlane, the output layer is a wave reduction.
__device__ inline void mlp_wave(const unsigned short* __restrict__ w, float xi, float (&y)[kOut]) {
  const int lane = (int)__lane_id();
  float h1 = bf16(w[kB1 + lane]);
#pragma unroll
  for (int i = 0; i < kIn; ++i) h1 += bf16(w[kW1 + i * kHid + lane]) * bcast(xi, i);
  h1 = fmaxf(h1, 0.f);
  float h2 = bf16(w[kB2 + lane]);
#pragma unroll
  for (int i = 0; i < kHid; ++i) h2 += bf16(w[kW2 + i * kHid + lane]) * bcast(h1, i);
  h2 = fmaxf(h2, 0.f);
#pragma unroll
  for (int o = 0; o < kOut; ++o) {
    float p = h2 * bf16(w[kW3 + lane * kOut + o]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) p += __shfl_xor(p, off, 64);
    y[o] = p + bf16(w[kB3 + o]);
  }
}

__device__ __forceinline__ unsigned hash32(unsigned x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// Direction-optimising BFS rule on estimated frontier edges (frontier size
// x average degree): top-down -> bottom-up when they exceed the unexplored
// edges / 14, back when the frontier falls below n / 24; SSSP pulls while
// the frontier holds over a quarter of the vertices.  Narrow when the
// frontier's estimated edges are fewer than the narrow groups' threads.
__device__ inline int heuristic(const FeatIn& a, unsigned valid) {
  const float deg = a.m / fmaxf(a.n, 1.f);
  const float fe = a.nin * deg, unv_e = fmaxf(a.n - a.total, 0.f) * deg;
  bool pull = false;
  if (valid & (1u << GDB_ARM_PULL)) {
    if (a.algo == GDB_TASK_BFS)
      pull = arm_pull(a.prev_arm) ? a.nin >= a.n / 24.f : fe > unv_e / 14.f;
    else
      pull = a.nin > a.n * 0.25f;
  }
  if (pull) return GDB_ARM_PULL;
  if ((valid & (1u << GDB_ARM_PUSH_NARROW)) && fe < a.narrow_threads) return GDB_ARM_PUSH_NARROW;
  return GDB_ARM_PUSH;
}

// The arm of the next iteration.  `lw` is the weight copy in LDS (NULL when
// the MLP is not in use).  Uniform across the wave.
__device__ inline int decide(const Cfg& c, const unsigned short* lw, const FeatIn& a, unsigned valid,
                             unsigned salt, float xi) {
  switch (c.mode) {
    case GDB_SEL_FIXED:
      return (c.fixed_arm >= 0 && c.fixed_arm < kOut && (valid & (1u << c.fixed_arm))) ? c.fixed_arm
                                                                                     : GDB_ARM_PUSH;
    case GDB_SEL_RANDOM: {
      // Sticky random arms: a new draw one iteration in eight, so most
      // records are taken in steady state (the same arm as the iteration
      // before, no wake-up of parked groups in the cost).
      const unsigned h = hash32(c.seed * 0x9e3779b9u ^ salt);
      if ((h & 7u) != 0 && a.prev_arm >= 0 && a.prev_arm < kOut && (valid & (1u << a.prev_arm)))
        return a.prev_arm;
      const int cnt = __popc(valid);
      int pick = (int)((h >> 3) % (unsigned)cnt);
      for (int a2 = 0; a2 < kOut; ++a2)
        if (valid & (1u << a2)) {
          if (pick == 0) return a2;
          --pick;
        }
      return GDB_ARM_PUSH;
    }
    case GDB_SEL_HEURISTIC: return heuristic(a, valid);
    case GDB_SEL_MLP: {
      if (!lw) return heuristic(a, valid);
      float y[kOut];
      mlp_wave(lw, xi, y);
      // Smallest predicted cost; leaving the current arm needs a predicted
      // gain of over 10% (0.15 in log2), so that model noise does not
      // flip the schedule back and forth.
      const int cur = (a.prev_arm >= 0 && a.prev_arm < kOut && (valid & (1u << a.prev_arm))) ? a.prev_arm
                                                                                            : GDB_ARM_PUSH;
      int best = cur;
      float by = 0.f;
#pragma unroll
      for (int a2 = 0; a2 < kOut; ++a2)
        if (a2 == cur) by = y[a2] - kSwitchMargin;
#pragma unroll
      for (int a2 = 0; a2 < kOut; ++a2)
        if (a2 != cur && (valid & (1u << a2)) && y[a2] < by) {
          best = a2;
          by = y[a2];
        }
      return best;
    }
    default: return GDB_ARM_PUSH;
  }
}

}  // namespace gdbsel
