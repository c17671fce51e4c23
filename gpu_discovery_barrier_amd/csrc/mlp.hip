// Forward, backward and Adam training of the schedule selector's MLP
// (16-64-64-4, ReLU, bf16 weights, fp32 accumulation).  No reference
// counterpart: the reference has no learned component.
//
// Loss over trace records with a measured cost: mean of
// (y[arm] - log2(cost_us))^2, so each output learns log2 of the iteration
// time of its arm.  One wave evaluates one record at a time: lane j holds
// hidden unit j of both layers; inputs and activations travel by lane
// readlane broadcasts; each wave keeps its share of the gradient in registers and adds
// it to the global sum once.
#include <algorithm>

#include "runtime.hpp"
#include "selector.hpp"

namespace {

using namespace gdbsel;
constexpr int kPad = kHid + 1;  // padded row stride of the W2 copy read row-wise

// y[count][4] = MLP(x[count][16]); one wave per row, the engine's own
// inference routine.
__global__ __launch_bounds__(64) void k_mlp_forward(const unsigned short* __restrict__ w,
                                                    const float* __restrict__ x, int count,
                                                    float* __restrict__ y) {
  const int r = blockIdx.x;
  if (r >= count) return;
  const int lane = (int)threadIdx.x;
  const float xi = lane < kIn ? x[(size_t)r * kIn + lane] : 0.f;
  float out[kOut];
  mlp_wave(w, xi, out);
  if (lane < kOut) y[(size_t)r * kOut + lane] = out[lane];
}

__device__ __forceinline__ float wave_sum(float p) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) p += __shfl_xor(p, off, 64);
  return p;
}

// Sum over the selected records of d/dparams (y[arm] - t)^2, plus the loss
// sum and the number of records used.  nsel records are taken: record s
// itself when `sample` is 0, else a record drawn by hashing (seed, step, s).
__global__ __launch_bounds__(256) void k_mlp_grad(const unsigned short* __restrict__ w,
                                                  const float* __restrict__ rec, int count, int nsel,
                                                  int sample, unsigned seed, unsigned step,
                                                  float* __restrict__ grad, float* loss_sum, int* used) {
  __shared__ float sw[kParams];
  __shared__ float sw2[kHid * kPad];  // W2[i][j] at i * kPad + j
  for (int i = threadIdx.x; i < kParams; i += blockDim.x) sw[i] = bf16(w[i]);
  for (int i = threadIdx.x; i < kHid * kHid; i += blockDim.x)
    sw2[(i / kHid) * kPad + (i % kHid)] = bf16(w[kW2 + i]);
  __syncthreads();
  const int lane = (int)(threadIdx.x & 63);
  const int wave = (int)(blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64);
  const int nwaves = (int)(gridDim.x * (blockDim.x / 64));
  float gW1[kIn], gW2[kHid], gW3[kOut], gb3[kOut];
#pragma unroll
  for (int i = 0; i < kIn; ++i) gW1[i] = 0.f;
#pragma unroll
  for (int i = 0; i < kHid; ++i) gW2[i] = 0.f;
#pragma unroll
  for (int o = 0; o < kOut; ++o) gW3[o] = gb3[o] = 0.f;
  float gb1 = 0.f, gb2 = 0.f, lsum = 0.f;
  int nused = 0;
  for (int s = wave; s < nsel; s += nwaves) {
    const int r = sample ? (int)(hash32(seed ^ hash32(step * 0x9e3779b9u + (unsigned)s)) % (unsigned)count) : s;
    const float* R = rec + (size_t)r * kStride;
    const float cost = R[kIn + 1];
    const int arm = (int)R[kIn];
    if (!(cost > 0.f) || arm < 0 || arm >= kOut) continue;  // uniform across the wave
    const float xi = lane < kIn ? R[lane] : 0.f;
    float a1 = sw[kB1 + lane];
#pragma unroll
    for (int i = 0; i < kIn; ++i) a1 += sw[kW1 + i * kHid + lane] * bcast(xi, i);
    const float h1 = fmaxf(a1, 0.f);
    float a2 = sw[kB2 + lane];
#pragma unroll
    for (int i = 0; i < kHid; ++i) a2 += sw[kW2 + i * kHid + lane] * bcast(h1, i);
    const float h2 = fmaxf(a2, 0.f);
    const float w3 = sw[kW3 + lane * kOut + arm];
    const float y = wave_sum(h2 * w3) + sw[kB3 + arm];
    const float err = y - __log2f(cost);
    lsum += err * err;
    ++nused;
    const float dy = 2.f * err;
#pragma unroll
    for (int o = 0; o < kOut; ++o)
      if (o == arm) {
        gW3[o] += h2 * dy;
        gb3[o] += dy;
      }
    const float dh2 = a2 > 0.f ? w3 * dy : 0.f;
    gb2 += dh2;
#pragma unroll
    for (int i = 0; i < kHid; ++i) gW2[i] += bcast(h1, i) * dh2;
    float d1 = 0.f;
#pragma unroll
    for (int j = 0; j < kHid; ++j) d1 += sw2[lane * kPad + j] * bcast(dh2, j);
    const float dh1 = a1 > 0.f ? d1 : 0.f;
    gb1 += dh1;
#pragma unroll
    for (int i = 0; i < kIn; ++i) gW1[i] += bcast(xi, i) * dh1;
  }
  if (nused == 0) return;  // uniform across the wave
#pragma unroll
  for (int i = 0; i < kIn; ++i) atomicAdd(&grad[kW1 + i * kHid + lane], gW1[i]);
  atomicAdd(&grad[kB1 + lane], gb1);
#pragma unroll
  for (int i = 0; i < kHid; ++i) atomicAdd(&grad[kW2 + i * kHid + lane], gW2[i]);
  atomicAdd(&grad[kB2 + lane], gb2);
#pragma unroll
  for (int o = 0; o < kOut; ++o) atomicAdd(&grad[kW3 + lane * kOut + o], gW3[o]);
  if (lane < kOut) {
    float b = 0.f;
#pragma unroll
    for (int o = 0; o < kOut; ++o)
      if (o == lane) b = gb3[o];
    atomicAdd(&grad[kB3 + lane], b);
  }
  if (lane == 0) {
    atomicAdd(loss_sum, lsum);
    atomicAdd(used, nused);
  }
}

// grad /= used (the mean-loss gradient).
__global__ void k_mlp_scale(float* grad, const int* used) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int u = *used;
  if (p < kParams && u > 0) grad[p] /= (float)u;
}

__global__ void k_to_bf16(const float* __restrict__ p, unsigned short* __restrict__ w, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) w[i] = to_bf16(p[i]);
}

// One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the mean gradient;
// refreshes the bf16 copy and clears the gradient sum for the next step.
__global__ void k_adam(float* __restrict__ params, float* __restrict__ m, float* __restrict__ v,
                       float* __restrict__ grad, const int* used, unsigned short* __restrict__ w,
                       float lr, float bc1, float bc2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kParams) return;
  const int u = *used;
  const float g = u > 0 ? grad[p] / (float)u : 0.f;
  const float mm = 0.9f * m[p] + 0.1f * g;
  const float vv = 0.999f * v[p] + 0.001f * g * g;
  m[p] = mm;
  v[p] = vv;
  const float x = params[p] - lr * (mm / bc1) / (sqrtf(vv / bc2) + 1e-8f);
  params[p] = x;
  w[p] = to_bf16(x);
  grad[p] = 0.f;
}

constexpr int kGradBlocks = 128;

}  // namespace

using namespace gdbrt;

extern "C" int gdb_mlp_forward(gdb_ctx* ctx, const unsigned short* w_bf16, const float* x, int count,
                               float* y) {
  if (!ctx || !w_bf16 || count < 0 || (count > 0 && (!x || !y))) return GDB_E_INVALID;
  if (count == 0) return GDB_SUCCESS;
  hipLaunchKernelGGL(k_mlp_forward, dim3(count), dim3(64), 0, ctx->stream, w_bf16, x, count, y);
  GDB_HIP(hipGetLastError());
  GDB_HIP(hipStreamSynchronize(ctx->stream));
  return GDB_SUCCESS;
}

extern "C" int gdb_mlp_to_bf16(gdb_ctx* ctx, const float* params, unsigned short* w_bf16, int count) {
  if (!ctx || count < 0 || (count > 0 && (!params || !w_bf16))) return GDB_E_INVALID;
  if (count == 0) return GDB_SUCCESS;
  hipLaunchKernelGGL(k_to_bf16, dim3((count + 255) / 256), dim3(256), 0, ctx->stream, params, w_bf16, count);
  GDB_HIP(hipGetLastError());
  GDB_HIP(hipStreamSynchronize(ctx->stream));
  return GDB_SUCCESS;
}

namespace {
// Scratch for the loss sum and record count (workspace head).
struct LossSlot {
  float* loss;
  int* used;
};
int loss_slot(gdb_ctx* ctx, size_t extra_floats, LossSlot* s, float** extra) {
  const size_t bytes = carve_bytes<float>(64) + carve_bytes<float>(extra_floats);
  int rc = ensure_workspace(ctx, bytes);
  if (rc) return rc;
  Carve cv(ctx->ws);
  float* head = cv.take<float>(64);
  s->loss = head;
  s->used = (int*)(head + 32);
  if (extra) *extra = cv.take<float>(extra_floats);
  return GDB_SUCCESS;
}
int read_loss(gdb_ctx* ctx, const LossSlot& s, float* loss, int* used) {
  float ls = 0.f;
  int u = 0;
  GDB_HIP(hipMemcpyAsync(&ls, s.loss, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  GDB_HIP(hipMemcpyAsync(&u, s.used, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  GDB_HIP(hipStreamSynchronize(ctx->stream));
  if (loss) *loss = u > 0 ? ls / (float)u : 0.f;
  if (used) *used = u;
  return GDB_SUCCESS;
}
}  // namespace

extern "C" int gdb_mlp_grad(gdb_ctx* ctx, const unsigned short* w_bf16, const float* records, int count,
                            float* grad, float* loss, int* used) {
  if (!ctx || !w_bf16 || !grad || count < 0 || (count > 0 && !records)) return GDB_E_INVALID;
  LossSlot s;
  int rc = loss_slot(ctx, 0, &s, nullptr);
  if (rc) return rc;
  rc = fill_i32(ctx, (int*)s.loss, 0, 64);
  if (!rc) rc = fill_i32(ctx, (int*)grad, 0, kParams);
  if (rc) return rc;
  if (count > 0) {
    hipLaunchKernelGGL(k_mlp_grad, dim3(kGradBlocks), dim3(256), 0, ctx->stream, w_bf16, records, count,
                       count, 0, 0u, 0u, grad, s.loss, s.used);
    GDB_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_mlp_scale, dim3((kParams + 255) / 256), dim3(256), 0, ctx->stream, grad,
                       (const int*)s.used);
    GDB_HIP(hipGetLastError());
  }
  return read_loss(ctx, s, loss, used);
}

extern "C" int gdb_mlp_train(gdb_ctx* ctx, float* params, float* adam_m, float* adam_v,
                             unsigned short* w_bf16, const float* records, int count, int steps, int batch,
                             float lr, unsigned seed, float* loss) {
  if (!ctx || !params || !adam_m || !adam_v || !w_bf16 || count <= 0 || !records || steps < 0 ||
      batch <= 0 || !(lr > 0.f))
    return GDB_E_INVALID;
  LossSlot s;
  float* grad = nullptr;
  int rc = loss_slot(ctx, kParams, &s, &grad);
  if (rc) return rc;
  rc = fill_i32(ctx, (int*)grad, 0, kParams);
  if (rc) return rc;
  hipLaunchKernelGGL(k_to_bf16, dim3((kParams + 255) / 256), dim3(256), 0, ctx->stream, params, w_bf16,
                     kParams);
  GDB_HIP(hipGetLastError());
  const int blocks = std::min(kGradBlocks, (batch + 3) / 4);
  double bc1 = 1.0, bc2 = 1.0;
  for (int t = 1; t <= steps; ++t) {
    bc1 *= 0.9;
    bc2 *= 0.999;
    rc = fill_i32(ctx, (int*)s.loss, 0, 64);
    if (rc) return rc;
    hipLaunchKernelGGL(k_mlp_grad, dim3(blocks), dim3(256), 0, ctx->stream, w_bf16, records, count, batch, 1,
                       seed, (unsigned)t, grad, s.loss, s.used);
    GDB_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_adam, dim3((kParams + 255) / 256), dim3(256), 0, ctx->stream, params, adam_m, adam_v,
                       grad, (const int*)s.used, w_bf16, lr, (float)(1.0 - bc1), (float)(1.0 - bc2));
    GDB_HIP(hipGetLastError());
  }
  // Full-set loss with the final weights.
  rc = fill_i32(ctx, (int*)s.loss, 0, 64);
  if (rc) return rc;
  hipLaunchKernelGGL(k_mlp_grad, dim3(kGradBlocks), dim3(256), 0, ctx->stream, w_bf16, records, count, count,
                     0, 0u, 0u, grad, s.loss, s.used);
  GDB_HIP(hipGetLastError());
  return read_loss(ctx, s, loss, nullptr);
}
