// Host-side runtime shared by every translation unit of libgdb_amd.so.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdio>
#include <vector>
#include "gdb_ctx_layout.h"
#include "../../include/gdb_amd.h"

struct gdb_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  gdb_kernel_ctx_t* d = nullptr;  // device context (fine-grained memory)
  bool fine_grained = false;
  int lock_kind = GDB_LOCK_FREE;
  int bar_kind = GDB_BAR_AUTO;
  int xcd_only = -1;
  int skip = 0;
  unsigned timeout_ticks = 1000000000u;  // 10 s at 100 MHz
  unsigned quiet_ticks = 2000u;          // 20 us
  int num_cus = 0;
  int xcds = 8;
  // Grow-only device workspace reused across calls (allocated outside the
  // timed region; never freed while the context lives).
  void* ws = nullptr;
  size_t ws_bytes = 0;
  // Device address, on this context's device, of the library's static
  // scratch array (gdb_sssp_worklist_noalloc); resolved at context creation.
  void* noalloc_scratch = nullptr;
  // Pinned host word for stop-flag readbacks of the relaunch schedule.
  int* h_flag = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace gdbrt {

// Maps a HIP error to the C-ABI code, printing the HIP message once.
int hip_fail(hipError_t e, const char* what, const char* file, int line);

#define GDB_HIP(call)                                                        \
  do {                                                                       \
    hipError_t _e = (call);                                                  \
    if (_e != hipSuccess) return gdbrt::hip_fail(_e, #call, __FILE__, __LINE__); \
  } while (0)

// Ensures ctx->ws holds at least `bytes`; returns GDB_SUCCESS or an error.
int ensure_workspace(gdb_ctx* ctx, size_t bytes);

// Bump allocator over the workspace: 256-byte aligned slices.
struct Carve {
  char* base;
  size_t off = 0;
  explicit Carve(void* b) : base((char*)b) {}
  template <typename T>
  T* take(size_t count) {
    T* p = (T*)(base + off);
    off += (count * sizeof(T) + 255) & ~(size_t)255;
    return p;
  }
};
template <typename T>
inline size_t carve_bytes(size_t count) { return (count * sizeof(T) + 255) & ~(size_t)255; }

// Fills count ints with value on the context stream (a kernel, not a memset).
int fill_i32(gdb_ctx* ctx, int* p, int value, long long count);

// Writes the per-launch protocol configuration and clears the sticky error.
int prepare_launch(gdb_ctx* ctx);

// Reads the device sticky error after a launch (blocking).  Returns
// GDB_SUCCESS, GDB_E_TIMEOUT or GDB_E_OVERFLOW.
int check_device_error(gdb_ctx* ctx);

// Occupancy bound (co-resident groups on the whole device) for `kernel`.
int occupancy_bound(gdb_ctx* ctx, const void* kernel, int wgs, size_t lds, int* out);

// Resolves the number of groups to launch: opts->groups if positive, else the
// occupancy bound (with `dyn_lds` bytes of dynamic LDS per group), clamped to [1, min(work_groups, GDB_MAX_GROUPS)].
int launch_groups(gdb_ctx* ctx, const void* kernel, const gdb_opts* opts, long long work_items,
                  int* out, size_t dyn_lds = 0);

inline int default_wgs(const gdb_opts* o) { return (o && o->wgs > 0) ? o->wgs : 256; }

// Reads the participating count of the last launch (blocking).
int read_participating(gdb_ctx* ctx, int* out);

// Fills a solver's launch shape from the tuning table when the caller left
// it open (o->wgs == 0): wgs, groups (if 0) and lds_queue (if 0) from the row
// for `solver` with the smallest n_max >= n.  `team_groups` (may be NULL)
// is filled when it points to 0.  Solvers other than the batched engine
// call it for the tuned persistent variant only.  Returns whether a row was
// applied (false when the shape was given or the table has no such row).
bool apply_tuning(const char* solver, long long n, gdb_opts* o, int* team_groups = nullptr);

// Resolves, once per device, the device address of the static scratch array
// of the allocation-free worklist entry point and stores it in the context
// (defined beside that array, in apps_worklist.hip).
int resolve_noalloc_scratch(gdb_ctx* ctx);

// Start / stop the solve timer on the context stream.
int timer_start(gdb_ctx* ctx);
int timer_stop(gdb_ctx* ctx, double* ms);

}  // namespace gdbrt

// Dispatches a kernel template instantiated for workgroup sizes 64..1024 on
// a runtime wgs value.  F must be a generic lambda taking an integral
// constant.
template <typename F>
inline int gdb_dispatch_wgs(int wgs, F&& f) {
  switch (wgs) {
    case 64: return f(std::integral_constant<int, 64>{});
    case 128: return f(std::integral_constant<int, 128>{});
    case 256: return f(std::integral_constant<int, 256>{});
    case 512: return f(std::integral_constant<int, 512>{});
    case 1024: return f(std::integral_constant<int, 1024>{});
    default: return GDB_E_INVALID;
  }
}
