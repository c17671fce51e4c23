// Memory-latency probes behind the barrier and worklist designs: what one
// dependent global load, one dependent agent-scope atomic and a wave step of
// returning atomics cost on this device.  No reference counterpart (the
// reference measures only the protocol, discovery.h:71-105); these replace
// one-off measurements with a reproducible entry point (bench.py --config
// probe).
//
//   chain   one lane follows a random cycle through a buffer, one access per
//           128 B line: next = buf[cur] (kind 0: plain load; kind 2:
//           agent-scope relaxed atomic load) or
//           next = atomic_add(&buf[cur], 0) (kind 1: returning, agent scope).  The
//           buffer size picks the level that serves it (L2 4 MB per XCD,
//           Infinity Cache 256 MB, HBM beyond); fine = 1 puts it in
//           fine-grained memory (as the discovery context).
//   atomics every lane of `groups` groups does `iters` dependent returning
//           atomic-min operations (32- or 64-bit) on its own random lines.
#include <chrono>
#include <random>
#include <vector>
#include "gdb_discovery.hpp"
#include "runtime.hpp"

namespace {

template <bool SCOPED>
__global__ void k_chain_load(const unsigned* __restrict__ buf, int steps, unsigned start, unsigned* sink) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  unsigned cur = start;
  for (int i = 0; i < steps; ++i)
    cur = SCOPED ? __hip_atomic_load(&buf[cur], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : buf[cur];
  *sink = cur;
}

__global__ void k_chain_atomic(unsigned* buf, int steps, unsigned start, unsigned* sink) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  unsigned cur = start;
  for (int i = 0; i < steps; ++i)
    cur = __hip_atomic_fetch_add(&buf[cur], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  *sink = cur;
}

template <typename T>
__global__ void k_atomic_tput(T* buf, unsigned lines, int iters, unsigned long long* sink) {
  const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned x = t * 2654435761u + 12345u;
  T acc = 0;
  const unsigned stride = 128 / sizeof(T);
  for (int i = 0; i < iters; ++i) {
    x = x * 1664525u + 1013904223u + (unsigned)acc;  // the next address depends on the last result
    const unsigned line = x % lines;
    acc += __hip_atomic_fetch_min(&buf[(size_t)line * stride], (T)x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (acc == (T)0x12345) sink[0] = (unsigned long long)acc;
}

// Two groups hand a counter back and forth `iters` times each: group A waits
// for 2k and writes 2k+1, group B waits for 2k+1 and writes 2k+2.  mode 0:
// relaxed agent-scope store / relaxed poll; mode 1: release store / relaxed
// poll + acquire fence (the barrier's message pattern); mode 2: acq_rel
// fetch-add / relaxed poll + acquire fence (a counter arrival).  Every wait
// is bounded by the context's deadline.  xcc[0..1] receive the groups' XCDs.
__global__ void k_pingpong(gdb_kernel_ctx_t* ctx, int* flag, int iters, int mode, int partner, int* xcc) {
  const int b = (int)blockIdx.x;
  if ((b != 0 && b != partner) || threadIdx.x != 0) return;
  const int side = b == 0 ? 0 : 1;
  xcc[side] = gdb::xcc_id();
  for (int k = 0; k < iters; ++k) {
    const int want = 2 * k + side;
    if (!gdb::wait_while(ctx, [&] { return gdb::ld_rlx(flag) != want; }, GDB_ERR_BARRIER_TIMEOUT)) return;
    if (mode == 0) {
      gdb::st_rlx(flag, want + 1);
    } else if (mode == 1) {
      gdb::fence_acq();
      gdb::st_rel(flag, want + 1);
    } else {
      gdb::fence_acq();
      gdb::fadd(flag, 1);
    }
  }
}

}  // namespace

using namespace gdbrt;

extern "C" int gdb_probe_pingpong(gdb_ctx* ctx, int mode, int fine, int same_xcd, int iters, double* ns_per_handoff,
                                  int* xcds) {
  if (!ctx || mode < 0 || mode > 2 || iters < 1 || !ns_per_handoff) return GDB_E_INVALID;
  void* d = nullptr;
  if (fine) GDB_HIP(hipExtMallocWithFlags(&d, 256, hipDeviceMallocFinegrained));
  else GDB_HIP(hipMalloc(&d, 256));
  int* xcc = (int*)((char*)d + 128);
  // Groups are dealt to the XCDs round-robin: group 8 shares group 0's XCD,
  // group 1 sits on the next one.
  const int partner = same_xcd ? 8 : 1;
  int rc = prepare_launch(ctx);
  if (!rc) rc = fill_i32(ctx, (int*)d, 0, 64);
  hipError_t e = hipSuccess;
  if (!rc) e = hipEventRecord(ctx->ev0, ctx->stream);
  if (!rc && e == hipSuccess) {
    hipLaunchKernelGGL(k_pingpong, dim3(partner + 1), dim3(64), 0, ctx->stream, ctx->d, (int*)d, iters, mode,
                       partner, xcc);
    e = hipGetLastError();
  }
  if (!rc && e == hipSuccess) e = hipEventRecord(ctx->ev1, ctx->stream);
  if (!rc && e == hipSuccess) e = hipEventSynchronize(ctx->ev1);
  float ms = 0.f;
  if (!rc && e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
  int hx[2] = {-1, -1};
  if (!rc && e == hipSuccess) e = hipMemcpy(hx, xcc, sizeof(hx), hipMemcpyDeviceToHost);
  if (e != hipSuccess) rc = hip_fail(e, "gdb_probe_pingpong", __FILE__, __LINE__);
  if (!rc) rc = check_device_error(ctx);
  if (!rc) {
    *ns_per_handoff = (double)ms * 1e6 / (2.0 * iters);
    if (xcds) {
      xcds[0] = hx[0];
      xcds[1] = hx[1];
    }
  }
  (void)hipFree(d);
  return rc;
}

extern "C" int gdb_probe_chain(gdb_ctx* ctx, int kind, size_t bytes, int fine, int steps, double* ns_per_step) {
  if (!ctx || kind < 0 || kind > 2 || bytes < 4096 || bytes > ((size_t)1 << 34) || steps < 1 || !ns_per_step)
    return GDB_E_INVALID;
  const size_t words = bytes / 4;
  const size_t lines = bytes / GDB_LINE_BYTES;
  // A random cyclic order of the lines; word 0 of each line holds the index
  // of word 0 of the next line in the cycle.
  std::vector<unsigned> order(lines);
  for (size_t i = 0; i < lines; ++i) order[i] = (unsigned)i;
  std::mt19937_64 rng(12345);
  for (size_t i = lines - 1; i > 0; --i) std::swap(order[i], order[rng() % (i + 1)]);
  std::vector<unsigned> h(words, 0u);
  const unsigned per_line = GDB_LINE_BYTES / 4;
  for (size_t i = 0; i < lines; ++i) h[(size_t)order[i] * per_line] = order[(i + 1) % lines] * per_line;
  void* d = nullptr;
  if (fine) GDB_HIP(hipExtMallocWithFlags(&d, bytes, hipDeviceMallocFinegrained));
  else GDB_HIP(hipMalloc(&d, bytes));
  unsigned* sink = nullptr;
  int rc = GDB_SUCCESS;
  hipError_t e = hipMalloc(&sink, 256);
  if (e == hipSuccess) e = hipMemcpy(d, h.data(), bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    unsigned start = order[0] * per_line;
    // Warm-up pass over the whole cycle (at most 4M lines, 512 MB), then the
    // timed pass continues from where it stopped: a buffer that fits a cache
    // level is served by it; a larger one is timed on lines the warm-up never
    // touched.
    for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) {
      const int n = rep == 0 ? (int)std::min<size_t>(lines, (size_t)1 << 22) : steps;
      if (rep == 1) {
        e = hipMemcpy(&start, sink, sizeof(unsigned), hipMemcpyDeviceToHost);
        if (e == hipSuccess && start >= words) e = hipErrorInvalidValue;
        if (e == hipSuccess) e = hipEventRecord(ctx->ev0, ctx->stream);
        if (e != hipSuccess) break;
      }
      if (kind == 0)
        hipLaunchKernelGGL(k_chain_load<false>, dim3(1), dim3(64), 0, ctx->stream, (const unsigned*)d, n, start, sink);
      else if (kind == 2)
        hipLaunchKernelGGL(k_chain_load<true>, dim3(1), dim3(64), 0, ctx->stream, (const unsigned*)d, n, start, sink);
      else
        hipLaunchKernelGGL(k_chain_atomic, dim3(1), dim3(64), 0, ctx->stream, (unsigned*)d, n, start, sink);
      if (e == hipSuccess) e = hipGetLastError();
      if (rep == 1 && e == hipSuccess) e = hipEventRecord(ctx->ev1, ctx->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    }
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    if (e == hipSuccess) *ns_per_step = (double)ms * 1e6 / steps;
  }
  if (e != hipSuccess) rc = hip_fail(e, "gdb_probe_chain", __FILE__, __LINE__);
  (void)hipFree(sink);
  (void)hipFree(d);
  return rc;
}

extern "C" int gdb_probe_atomics(gdb_ctx* ctx, int bits, int groups, int wgs, int iters, double* ns_per_step) {
  if (!ctx || (bits != 32 && bits != 64) || groups < 1 || groups > 65536 || wgs < 64 || wgs > 1024 || iters < 1 ||
      !ns_per_step)
    return GDB_E_INVALID;
  const unsigned lines = 1u << 20;  // 128 MB of lines: spread over every channel
  const size_t bytes = (size_t)lines * GDB_LINE_BYTES;
  void* d = nullptr;
  GDB_HIP(hipMalloc(&d, bytes));
  unsigned long long* sink = nullptr;
  hipError_t e = hipMalloc(&sink, 256);
  if (e == hipSuccess) e = hipMemset(d, 0xff, bytes);
  int rc = GDB_SUCCESS;
  for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) {
    if (rep == 1) e = hipEventRecord(ctx->ev0, ctx->stream);
    if (bits == 32)
      hipLaunchKernelGGL(k_atomic_tput<unsigned>, dim3(groups), dim3(wgs), 0, ctx->stream, (unsigned*)d, lines, iters,
                         sink);
    else
      hipLaunchKernelGGL(k_atomic_tput<unsigned long long>, dim3(groups), dim3(wgs), 0, ctx->stream,
                         (unsigned long long*)d, lines, iters, sink);
    if (e == hipSuccess) e = hipGetLastError();
    if (rep == 1 && e == hipSuccess) e = hipEventRecord(ctx->ev1, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  }
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
  if (e == hipSuccess) *ns_per_step = (double)ms * 1e6 / iters;
  if (e != hipSuccess) rc = hip_fail(e, "gdb_probe_atomics", __FILE__, __LINE__);
  (void)hipFree(sink);
  (void)hipFree(d);
  return rc;
}
