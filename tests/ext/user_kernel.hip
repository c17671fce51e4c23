// A caller's persistent kernel built OUTSIDE libgdb_amd.so, against the
// public device header only (include/gdb_discovery.hpp), as a reference
// app's kernel is built against discovery.cl.  The tests launch it on a
// library context (gdb_ctx_device_ptr + gdb_launch_on_ctx) and time it with
// gdb_time_protocol_kernel.
//
// Per barrier k: every group adds 1 to count[k % 3] and every thread tags
// its slot of data[k % 2]; after the barrier each group checks the count
// (P * (k / 3 + 1)) and its neighbour group's tags.  Each participant also
// marks ids[p_get_group_id], so the host can check that the ids form
// 0..P-1.  Reference: the run_test kernel (occupancy_tests/kernels/
// occupancy_test.cl:8-11), a protocol and a barrier.
#include "gdb_discovery.hpp"

__global__ void user_barrier_kernel(gdb_kernel_ctx_t* ctx, int iters, int* count, int* data, int stride,
                                    unsigned long long* bad, int* ids) {
  DISCOVERY_PROTOCOL(ctx);
  const int id = p_get_group_id(ctx, &local_ctx);
  const int P = p_get_num_groups(ctx, &local_ctx);
  if (threadIdx.x == 0 && id < stride) atomicAdd(&ids[id], 1);
  const int nb = (id + 1) % P;
  const int gid = p_get_global_id(ctx, &local_ctx);
  const int nb_gid = nb * (int)blockDim.x + (int)threadIdx.x;
  unsigned long long wrong = 0;
  for (int k = 0; k < iters; ++k) {
    int* slot = data + (size_t)(k & 1) * stride;
    if (gid < stride) slot[gid] = k * 7919 + id;
    if (threadIdx.x == 0) atomicAdd(&count[k % 3], 1);
    if (!discovery_barrier(ctx, &local_ctx)) return;
    if (nb_gid < stride && __hip_atomic_load(&slot[nb_gid], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) !=
                               k * 7919 + nb)
      ++wrong;
    if (threadIdx.x == 0 &&
        __hip_atomic_load(&count[k % 3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != P * (k / 3 + 1))
      ++wrong;
  }
  if (wrong) atomicAdd(bad, wrong);
}

extern "C" const void* user_barrier_kernel_symbol(void) { return (const void*)&user_barrier_kernel; }
