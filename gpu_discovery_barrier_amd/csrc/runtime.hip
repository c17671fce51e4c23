// Discovery context lifecycle, protocol timing, occupancy tests and the
// barrier / relaunch micro-benchmarks.
//
// Host API counterpart of code/discovery_protocol/api/discovery.h and of the
// occupancy harness in code/experiments/occupancy_tests/src/.
#include <dlfcn.h>
#include <chrono>
#include <cstddef>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>
#include "gdb_discovery.hpp"
#include "runtime.hpp"

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// init_discovery_kernel_ctx (discovery.cl:194-202): one thread per 128 B line
// writes that line's initial value, so the whole context is reset by one
// launch and no two threads write the same word.
__global__ void gdb_k_init_ctx(gdb_kernel_ctx_t* ctx, int skip, int lock_kind, int bar_kind,
                               int quiet_ticks, int timeout_ticks, int xcd_only) {
  const int nlines = (int)(sizeof(gdb_kernel_ctx_t) / GDB_LINE_BYTES);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nlines) return;
  gdb_line_t* lines = (gdb_line_t*)ctx;
  const int off = i * GDB_LINE_BYTES;
  int v = 0;
  if (off == (int)offsetof(gdb_kernel_ctx_t, prot_poll_open)) v = 1;
  else if (off == (int)offsetof(gdb_kernel_ctx_t, skip)) v = skip;
  else if (off == (int)offsetof(gdb_kernel_ctx_t, lock_kind)) v = lock_kind;
  else if (off == (int)offsetof(gdb_kernel_ctx_t, bar_kind)) v = bar_kind;
  else if (off == (int)offsetof(gdb_kernel_ctx_t, quiet_ticks)) v = quiet_ticks;
  else if (off == (int)offsetof(gdb_kernel_ctx_t, timeout_ticks)) v = timeout_ticks;
  else if (off == (int)offsetof(gdb_kernel_ctx_t, xcd_only)) v = xcd_only;
  gdb::st_rlx(&lines[i].v, v);
}

// Fills `count` ints with `value` (used instead of hipMemsetAsync so that a
// captured stream never records memset nodes).
__global__ void gdb_k_fill_i32(int* p, int value, long long count) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = value;
}

// Clears the sticky error word before a solve.
__global__ void gdb_k_clear_error(gdb_kernel_ctx_t* ctx) { gdb::st_rlx(&ctx->error.v, GDB_OK); }

// Rewrites the wait bounds of a context without resetting its other words.
__global__ void gdb_k_set_timeouts(gdb_kernel_ctx_t* ctx, int quiet_ticks, int timeout_ticks) {
  gdb::st_rlx(&ctx->quiet_ticks.v, quiet_ticks);
  gdb::st_rlx(&ctx->timeout_ticks.v, timeout_ticks);
}

// run_prot (occupancy_test.cl:15-17): protocol only.
__global__ void gdb_k_run_prot(gdb_kernel_ctx_t* ctx) {
  extern __shared__ int loc_mem[];
  DISCOVERY_PROTOCOL(ctx);
  if (threadIdx.x == 0) loc_mem[0] = local_ctx.participating_group_id;
}

// run_test (occupancy_test.cl:8-11): protocol, then one barrier.
__global__ void gdb_k_run_test(gdb_kernel_ctx_t* ctx) {
  extern __shared__ int loc_mem[];
  DISCOVERY_PROTOCOL(ctx);
  if (threadIdx.x == 0) loc_mem[0] = local_ctx.participating_group_id;
  discovery_barrier(ctx, &local_ctx);
}

// Barrier cost kernel: `iters` back-to-back barriers.
__global__ void gdb_k_barrier_loop(gdb_kernel_ctx_t* ctx, int iters) {
  DISCOVERY_PROTOCOL(ctx);
  for (int i = 0; i < iters; ++i)
    if (!discovery_barrier(ctx, &local_ctx)) return;
}

// Barrier invariant kernel.  Before barrier k every group adds 1 to slot
// k % 3 of `count` and every thread stores a tag into its slot of data[k % 2];
// after barrier k every group checks that slot k % 3 holds P * (k / 3 + 1) and
// that the next group's tags are visible.  Violations are counted.
__global__ void gdb_k_barrier_check(gdb_kernel_ctx_t* ctx, int iters, int* count, int* data,
                                    int stride, unsigned long long* violations) {
  DISCOVERY_PROTOCOL(ctx);
  const int id = p_get_group_id(ctx, &local_ctx);
  const int P = p_get_num_groups(ctx, &local_ctx);
  const int nb = (id + 1) % P;
  unsigned long long bad = 0;
  for (int k = 0; k < iters; ++k) {
    int* slot = data + (size_t)(k & 1) * stride;
    slot[id * blockDim.x + threadIdx.x] = k * 7919 + id;
    if (threadIdx.x == 0) gdb::fadd_rlx(&count[k % 3], 1);
    if (!discovery_barrier(ctx, &local_ctx)) return;
    if (slot[nb * blockDim.x + threadIdx.x] != k * 7919 + nb) ++bad;
    if (threadIdx.x == 0 && gdb::ld_rlx(&count[k % 3]) != P * (k / 3 + 1)) ++bad;
  }
  if (bad) atomicAdd(violations, bad);
}

__global__ void gdb_k_empty(int* sink) {
  if (sink && threadIdx.x == 0 && blockIdx.x == 0x7fffffff) sink[0] = 0;
}

// ---------------------------------------------------------------------------
// Runtime helpers
// ---------------------------------------------------------------------------
namespace gdbrt {

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
  fprintf(stderr, "gdb_amd: HIP error %d (%s) in %s at %s:%d\n", (int)e, hipGetErrorString(e), what,
          file, line);
  return e == hipErrorOutOfMemory ? GDB_E_NOMEM : GDB_E_HIP;
}

int ensure_workspace(gdb_ctx* ctx, size_t bytes) {
  if (bytes <= ctx->ws_bytes) return GDB_SUCCESS;
  if (ctx->ws) GDB_HIP(hipFree(ctx->ws));
  ctx->ws = nullptr;
  ctx->ws_bytes = 0;
  size_t want = bytes + bytes / 4 + 4096;
  GDB_HIP(hipMalloc(&ctx->ws, want));
  ctx->ws_bytes = want;
  return GDB_SUCCESS;
}

int prepare_launch(gdb_ctx* ctx) {
  hipLaunchKernelGGL(gdb_k_clear_error, dim3(1), dim3(1), 0, ctx->stream, ctx->d);
  GDB_HIP(hipGetLastError());
  return GDB_SUCCESS;
}

int fill_i32(gdb_ctx* ctx, int* p, int value, long long count) {
  if (count <= 0) return GDB_SUCCESS;
  const long long blocks = (count + 255) / 256;
  hipLaunchKernelGGL(gdb_k_fill_i32, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, p, value, count);
  GDB_HIP(hipGetLastError());
  return GDB_SUCCESS;
}

static int init_ctx(gdb_ctx* ctx, int skip);

int check_device_error(gdb_ctx* ctx) {
  int err = 0;
  GDB_HIP(hipMemcpyAsync(&err, &ctx->d->error.v, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  GDB_HIP(hipStreamSynchronize(ctx->stream));
  if (err == GDB_ERR_BARRIER_TIMEOUT || err == GDB_ERR_DISCOVERY_TIMEOUT) {
    // A wait that gave up leaves arrival counters, XF flags or protocol
    // words part-way through an episode; the next launch would start from
    // them.  Reset the whole context (same configuration) before reporting.
    int rc = init_ctx(ctx, ctx->skip);
    return rc ? rc : GDB_E_TIMEOUT;
  }
  if (err == GDB_ERR_QUEUE_OVERFLOW) return GDB_E_OVERFLOW;
  return GDB_SUCCESS;
}

int occupancy_bound(gdb_ctx* ctx, const void* kernel, int wgs, size_t lds, int* out) {
  int per_cu = 0;
  GDB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, wgs, lds));
  *out = per_cu * ctx->num_cus;
  return GDB_SUCCESS;
}

int launch_groups(gdb_ctx* ctx, const void* kernel, const gdb_opts* opts, long long work_items,
                  int* out, size_t dyn_lds) {
  const int wgs = default_wgs(opts);
  int g = (opts && opts->groups > 0) ? opts->groups : 0;
  if (g == 0) {
    int rc = occupancy_bound(ctx, kernel, wgs, dyn_lds, &g);
    if (rc) return rc;
  }
  long long need = (work_items + wgs - 1) / wgs;
  if (need < 1) need = 1;
  if (g > need) g = (int)need;
  if (g > GDB_MAX_GROUPS) g = GDB_MAX_GROUPS;
  if (g < 1) g = 1;
  *out = g;
  return GDB_SUCCESS;
}

int read_participating(gdb_ctx* ctx, int* out) {
  GDB_HIP(hipMemcpyAsync(out, &ctx->d->num_participating.v, sizeof(int), hipMemcpyDeviceToHost,
                         ctx->stream));
  GDB_HIP(hipStreamSynchronize(ctx->stream));
  return GDB_SUCCESS;
}

int timer_start(gdb_ctx* ctx) {
  GDB_HIP(hipEventRecord(ctx->ev0, ctx->stream));
  return GDB_SUCCESS;
}

int timer_stop(gdb_ctx* ctx, double* ms) {
  GDB_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  GDB_HIP(hipEventSynchronize(ctx->ev1));
  float f = 0.f;
  GDB_HIP(hipEventElapsedTime(&f, ctx->ev0, ctx->ev1));
  if (ms) *ms = f;
  return GDB_SUCCESS;
}

// ---- tuning table ------------------------------------------------------------
namespace {
std::mutex g_tune_mu;
std::vector<gdb_tune_entry> g_tune;
bool g_tune_loaded = false;

std::string default_tune_path() {
  if (const char* env = getenv("GDB_TUNE_FILE")) return env;
  Dl_info info;
  if (dladdr((void*)&gdb_tune_load, &info) && info.dli_fname) {
    std::string p = info.dli_fname;
    const size_t slash = p.rfind('/');
    return (slash == std::string::npos ? std::string(".") : p.substr(0, slash)) + "/tune_table.txt";
  }
  return "tune_table.txt";
}

// Parses the table; false on a malformed row.
bool parse_tune(FILE* f, std::vector<gdb_tune_entry>* rows) {
  char line[512];
  while (fgets(line, sizeof(line), f)) {
    char* hash = strchr(line, '#');
    if (hash) *hash = 0;
    gdb_tune_entry e{};
    char name[64];
    const int got = sscanf(line, "%63s %lld %d %d %d %d %d %d", name, &e.n_max, &e.wgs, &e.groups, &e.team_groups,
                           &e.delta, &e.hops, &e.lds_queue);
    if (got <= 0) continue;  // blank or comment line
    if (got != 8 || strlen(name) >= sizeof(e.solver) || e.n_max < 0 || e.groups < 0 || e.team_groups < 0 ||
        e.delta < 0 || e.hops < 0 || e.lds_queue < 0 ||
        (e.wgs != 64 && e.wgs != 128 && e.wgs != 256 && e.wgs != 512 && e.wgs != 1024))
      return false;
    memcpy(e.solver, name, strlen(name) + 1);
    rows->push_back(e);
  }
  return true;
}

int load_tune_locked(const char* path) {
  const std::string p = path ? std::string(path) : default_tune_path();
  g_tune.clear();
  g_tune_loaded = true;
  FILE* f = fopen(p.c_str(), "r");
  if (!f) return path ? GDB_E_INVALID : 0;  // no default table: defaults stand
  std::vector<gdb_tune_entry> rows;
  const bool ok = parse_tune(f, &rows);
  fclose(f);
  if (!ok) return GDB_E_INVALID;
  g_tune = rows;
  return (int)g_tune.size();
}

bool lookup_tune_locked(const char* solver, long long n, gdb_tune_entry* out) {
  const gdb_tune_entry* best = nullptr;
  for (const gdb_tune_entry& e : g_tune)
    if (!strcmp(e.solver, solver) && e.n_max >= n && (!best || e.n_max < best->n_max)) best = &e;
  if (best) *out = *best;
  return best != nullptr;
}
}  // namespace

bool apply_tuning(const char* solver, long long n, gdb_opts* o, int* team_groups) {
  if (o->wgs != 0) return false;
  gdb_tune_entry e{};
  {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    if (!g_tune_loaded) load_tune_locked(nullptr);
    if (!lookup_tune_locked(solver, n, &e)) return false;
  }
  o->wgs = e.wgs;
  if (o->groups == 0) o->groups = e.groups;
  if (o->lds_queue == 0) o->lds_queue = e.lds_queue;
  if (team_groups && *team_groups == 0) *team_groups = e.team_groups;
  return true;
}

static int init_ctx(gdb_ctx* ctx, int skip) {
  ctx->skip = skip;
  const int nlines = (int)(sizeof(gdb_kernel_ctx_t) / GDB_LINE_BYTES);
  hipLaunchKernelGGL(gdb_k_init_ctx, dim3((nlines + 255) / 256), dim3(256), 0, ctx->stream, ctx->d,
                     skip, ctx->lock_kind, ctx->bar_kind, (int)ctx->quiet_ticks,
                     (int)ctx->timeout_ticks, ctx->xcd_only);
  GDB_HIP(hipGetLastError());
  GDB_HIP(hipStreamSynchronize(ctx->stream));  // discovery.h:28 blocks on clFinish
  return GDB_SUCCESS;
}

}  // namespace gdbrt

using namespace gdbrt;

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int gdb_malloc(void** ptr, size_t bytes) {
  if (!ptr) return GDB_E_INVALID;
  *ptr = nullptr;
  if (bytes == 0) bytes = 4;
  GDB_HIP(hipMalloc(ptr, bytes));
  return GDB_SUCCESS;
}

int gdb_free(void* ptr) {
  if (ptr) GDB_HIP(hipFree(ptr));
  return GDB_SUCCESS;
}

int gdb_memcpy_htod(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return GDB_SUCCESS;
  if (!dst || !src) return GDB_E_INVALID;
  GDB_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return GDB_SUCCESS;
}

int gdb_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return GDB_SUCCESS;
  if (!dst || !src) return GDB_E_INVALID;
  GDB_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return GDB_SUCCESS;
}

int gdb_memcpy_dtod(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return GDB_SUCCESS;
  if (!dst || !src) return GDB_E_INVALID;
  GDB_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  return GDB_SUCCESS;
}

int gdb_device_synchronize(void) {
  GDB_HIP(hipDeviceSynchronize());
  return GDB_SUCCESS;
}

int gdb_device_count(int* out) {
  if (!out) return GDB_E_INVALID;
  *out = 0;
  hipError_t e = hipGetDeviceCount(out);
  if (e != hipSuccess) {
    *out = 0;
    return GDB_SUCCESS;  // no device is not an error for the query
  }
  return GDB_SUCCESS;
}

int gdb_tune_load(const char* path) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  return load_tune_locked(path);
}

int gdb_tune_lookup(const char* solver, long long n, gdb_tune_entry* out) {
  if (!solver || !out) return GDB_E_INVALID;
  std::lock_guard<std::mutex> lock(g_tune_mu);
  if (!g_tune_loaded) load_tune_locked(nullptr);
  return lookup_tune_locked(solver, n, out) ? GDB_SUCCESS : GDB_E_INVALID;
}

int gdb_ctx_create(gdb_ctx** out) {
  if (!out) return GDB_E_INVALID;
  *out = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    if (!g_tune_loaded) load_tune_locked(nullptr);
  }
  gdb_ctx* ctx = new gdb_ctx();
  GDB_HIP(hipGetDevice(&ctx->device));
  hipDeviceProp_t prop;
  GDB_HIP(hipGetDeviceProperties(&prop, ctx->device));
  ctx->num_cus = prop.multiProcessorCount;
  GDB_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  // Fine-grained device memory: flag words are coherent across XCDs without
  // relying on L2 write-back of the polled lines.
  void* p = nullptr;
  hipError_t e = hipExtMallocWithFlags(&p, sizeof(gdb_kernel_ctx_t), hipDeviceMallocFinegrained);
  if (e == hipSuccess) {
    ctx->fine_grained = true;
  } else {
    (void)hipGetLastError();
    GDB_HIP(hipMalloc(&p, sizeof(gdb_kernel_ctx_t)));
  }
  ctx->d = (gdb_kernel_ctx_t*)p;
  GDB_HIP(hipHostMalloc((void**)&ctx->h_flag, sizeof(int) * 4, hipHostMallocDefault));
  GDB_HIP(hipEventCreate(&ctx->ev0));
  GDB_HIP(hipEventCreate(&ctx->ev1));
  int rc = resolve_noalloc_scratch(ctx);
  if (rc) return rc;
  rc = init_ctx(ctx, 0);
  if (rc) return rc;
  *out = ctx;
  return GDB_SUCCESS;
}

int gdb_ctx_destroy(gdb_ctx* ctx) {
  if (!ctx) return GDB_SUCCESS;
  if (ctx->stream) GDB_HIP(hipStreamSynchronize(ctx->stream));
  if (ctx->ws) GDB_HIP(hipFree(ctx->ws));
  if (ctx->d) GDB_HIP(hipFree(ctx->d));
  if (ctx->h_flag) GDB_HIP(hipHostFree(ctx->h_flag));
  if (ctx->ev0) GDB_HIP(hipEventDestroy(ctx->ev0));
  if (ctx->ev1) GDB_HIP(hipEventDestroy(ctx->ev1));
  if (ctx->stream) GDB_HIP(hipStreamDestroy(ctx->stream));
  delete ctx;
  return GDB_SUCCESS;
}

int gdb_ctx_configure(gdb_ctx* ctx, int lock_kind, int bar_kind) {
  if (!ctx) return GDB_E_INVALID;
  if (lock_kind < GDB_LOCK_FREE || lock_kind > GDB_LOCK_SPIN) return GDB_E_INVALID;
  if (bar_kind < GDB_BAR_XF || bar_kind > GDB_BAR_AUTO) return GDB_E_INVALID;
  ctx->lock_kind = lock_kind;
  ctx->bar_kind = bar_kind;
  return GDB_SUCCESS;
}

int gdb_ctx_set_xcd_only(gdb_ctx* ctx, int xcd) {
  if (!ctx || xcd < -1 || xcd >= GDB_MAX_XCDS) return GDB_E_INVALID;
  ctx->xcd_only = xcd;
  return GDB_SUCCESS;
}

int gdb_ctx_set_timeouts(gdb_ctx* ctx, double wait_us, double quiet_us) {
  if (!ctx || wait_us < 0 || quiet_us < 0) return GDB_E_INVALID;
  // 100 MHz wall clock: 100 ticks per microsecond.  A zero wait bound makes
  // every wait that would block fail immediately (used to force the timeout
  // path from the host).
  double w = wait_us * 100.0, q = quiet_us * 100.0;
  if (w > 2.0e9 || q > 2.0e9) return GDB_E_INVALID;
  ctx->timeout_ticks = (unsigned)(wait_us == 0 ? 0 : (w < 1 ? 1 : w));
  ctx->quiet_ticks = (unsigned)q;
  // The bounds also apply to the next launch without a re-init.
  hipLaunchKernelGGL(gdb_k_set_timeouts, dim3(1), dim3(1), 0, ctx->stream, ctx->d, (int)ctx->quiet_ticks,
                     (int)ctx->timeout_ticks);
  GDB_HIP(hipGetLastError());
  GDB_HIP(hipStreamSynchronize(ctx->stream));
  return GDB_SUCCESS;
}

int gdb_ctx_init(gdb_ctx* ctx) {
  if (!ctx) return GDB_E_INVALID;
  return init_ctx(ctx, 0);
}

int gdb_ctx_init_skip(gdb_ctx* ctx, int skip) {
  if (!ctx) return GDB_E_INVALID;
  return init_ctx(ctx, skip ? 1 : 0);
}

int gdb_number_of_participating_groups(gdb_ctx* ctx) {
  if (!ctx) return GDB_E_INVALID;
  int p = 0;
  int rc = read_participating(ctx, &p);
  return rc ? rc : p;
}

int gdb_ctx_device_error(gdb_ctx* ctx) {
  if (!ctx) return GDB_E_INVALID;
  return check_device_error(ctx);
}

void* gdb_ctx_stream(gdb_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int gdb_ctx_device_ptr(gdb_ctx* ctx, void** out) {
  if (!ctx || !out) return GDB_E_INVALID;
  *out = (void*)ctx->d;
  return GDB_SUCCESS;
}

int gdb_launch_on_ctx(gdb_ctx* ctx, const void* kernel, void** args, int groups, int wgs, size_t dyn_lds) {
  if (!ctx || !kernel || groups < 1 || wgs < 1 || wgs > 1024) return GDB_E_INVALID;
  int rc = prepare_launch(ctx);
  if (rc) return rc;
  GDB_HIP(hipLaunchKernel(kernel, dim3(groups), dim3(wgs), args, dyn_lds, ctx->stream));
  GDB_HIP(hipStreamSynchronize(ctx->stream));
  return check_device_error(ctx);
}

int gdb_time_protocol_kernel(gdb_ctx* ctx, const void* kernel, void** args, int upper_bound, int wgs,
                             size_t dyn_lds, double* ns) {
  if (!ctx || !kernel || upper_bound < 1 || wgs < 1 || wgs > 1024) return GDB_E_INVALID;
  int rc = init_ctx(ctx, 0);  // discovery.h:74
  if (rc) return rc;
  GDB_HIP(hipEventRecord(ctx->ev0, ctx->stream));
  GDB_HIP(hipLaunchKernel(kernel, dim3(upper_bound), dim3(wgs), args, dyn_lds, ctx->stream));
  GDB_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  GDB_HIP(hipEventSynchronize(ctx->ev1));
  float f = 0.f;
  GDB_HIP(hipEventElapsedTime(&f, ctx->ev0, ctx->ev1));
  if (ns) *ns = (double)f * 1.0e6;
  return check_device_error(ctx);
}

int gdb_max_resident_groups(int wgs, int lms, int* out) {
  if (!out || wgs < 1 || wgs > 1024 || lms < 0) return GDB_E_INVALID;
  int dev = 0, per_cu = 0;
  GDB_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  GDB_HIP(hipGetDeviceProperties(&prop, dev));
  GDB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)gdb_k_run_test, wgs,
                                                       (size_t)lms));
  *out = per_cu * prop.multiProcessorCount;
  return GDB_SUCCESS;
}

int gdb_time_protocol(gdb_ctx* ctx, int upper_bound, int wgs, int lms, double* ns) {
  if (!ctx || upper_bound < 1 || wgs < 1 || wgs > 1024 || lms < 0) return GDB_E_INVALID;
  int rc = init_ctx(ctx, 0);  // discovery.h:74
  if (rc) return rc;
  GDB_HIP(hipEventRecord(ctx->ev0, ctx->stream));
  hipLaunchKernelGGL(gdb_k_run_prot, dim3(upper_bound), dim3(wgs), (size_t)lms + 4, ctx->stream,
                     ctx->d);
  GDB_HIP(hipGetLastError());
  GDB_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  GDB_HIP(hipEventSynchronize(ctx->ev1));
  float f = 0.f;
  GDB_HIP(hipEventElapsedTime(&f, ctx->ev0, ctx->ev1));
  if (ns) *ns = (double)f * 1.0e6;
  return check_device_error(ctx);
}

int gdb_occupancy_test(gdb_ctx* ctx, int wgc, int wgs, int lms, int prot, int* participating) {
  if (!ctx || wgc < 1 || wgs < 1 || wgs > 1024 || lms < 0) return GDB_E_INVALID;
  if (!prot) {
    // Skip mode makes every launched group wait in the barrier; if not all
    // of them can be resident at once that wait can never finish.
    int bound = 0;
    int rc = gdb_max_resident_groups(wgs, lms + 4, &bound);
    if (rc) return rc;
    if (wgc > bound) return GDB_E_UNSAFE;
  }
  if (wgc > GDB_MAX_GROUPS && !prot) return GDB_E_INVALID;
  int rc = init_ctx(ctx, prot ? 0 : 1);  // occupancy_test.c:62
  if (rc) return rc;
  hipLaunchKernelGGL(gdb_k_run_test, dim3(wgc), dim3(wgs), (size_t)lms + 4, ctx->stream, ctx->d);
  GDB_HIP(hipGetLastError());
  GDB_HIP(hipStreamSynchronize(ctx->stream));
  rc = check_device_error(ctx);
  if (rc) return rc;
  if (participating) {
    rc = read_participating(ctx, participating);
    if (rc) return rc;
  }
  return GDB_SUCCESS;
}

int gdb_device_query(char* buf, size_t len) {
  if (!buf || len == 0) return GDB_E_INVALID;
  int dev = 0;
  GDB_HIP(hipGetDevice(&dev));
  hipDeviceProp_t p;
  GDB_HIP(hipGetDeviceProperties(&p, dev));
  int wall = 0;
  (void)hipDeviceGetAttribute(&wall, hipDeviceAttributeWallClockRate, dev);
  snprintf(buf, len,
           "device name: %s\narch: %s\ncompute units: %d\nmax workgroup size: %d\n"
           "max threads per CU: %d\nLDS per workgroup: %zu bytes\nL2 cache: %d bytes\n"
           "global memory: %zu bytes\nclock: %d kHz\nmemory clock: %d kHz\nbus width: %d bits\n"
           "wall clock rate: %d kHz\nwarp size: %d\n",
           p.name, p.gcnArchName, p.multiProcessorCount, p.maxThreadsPerBlock,
           p.maxThreadsPerMultiProcessor, p.sharedMemPerBlock, p.l2CacheSize, p.totalGlobalMem,
           p.clockRate, p.memoryClockRate, p.memoryBusWidth, wall, p.warpSize);
  return GDB_SUCCESS;
}

int gdb_barrier_bench(gdb_ctx* ctx, int wgs, int groups, int iters, double* ms, int* participating) {
  if (!ctx || wgs < 1 || wgs > 1024 || groups < 1 || groups > GDB_MAX_GROUPS || iters < 0)
    return GDB_E_INVALID;
  int rc = init_ctx(ctx, 0);
  if (rc) return rc;
  rc = timer_start(ctx);
  if (rc) return rc;
  hipLaunchKernelGGL(gdb_k_barrier_loop, dim3(groups), dim3(wgs), 0, ctx->stream, ctx->d, iters);
  GDB_HIP(hipGetLastError());
  rc = timer_stop(ctx, ms);
  if (rc) return rc;
  rc = check_device_error(ctx);
  if (rc) return rc;
  if (participating) return read_participating(ctx, participating);
  return GDB_SUCCESS;
}

int gdb_barrier_check(gdb_ctx* ctx, int wgs, int groups, int iters, long long* violations,
                      int* participating) {
  if (!ctx || !violations || wgs < 1 || wgs > 1024 || groups < 1 || groups > GDB_MAX_GROUPS ||
      iters < 0)
    return GDB_E_INVALID;
  int rc = init_ctx(ctx, 0);
  if (rc) return rc;
  const int stride = groups * wgs;
  const size_t bytes = carve_bytes<int>(4) + carve_bytes<int>((size_t)2 * stride) +
                       carve_bytes<unsigned long long>(1);
  rc = ensure_workspace(ctx, bytes);
  if (rc) return rc;
  Carve cv(ctx->ws);
  int* count = cv.take<int>(4);
  int* data = cv.take<int>((size_t)2 * stride);
  unsigned long long* bad = cv.take<unsigned long long>(1);
  rc = fill_i32(ctx, count, 0, 4);
  if (!rc) rc = fill_i32(ctx, data, -1, (long long)2 * stride);
  if (!rc) rc = fill_i32(ctx, (int*)bad, 0, 2);
  if (rc) return rc;
  hipLaunchKernelGGL(gdb_k_barrier_check, dim3(groups), dim3(wgs), 0, ctx->stream, ctx->d, iters,
                     count, data, stride, bad);
  GDB_HIP(hipGetLastError());
  unsigned long long hbad = 0;
  GDB_HIP(hipMemcpyAsync(&hbad, bad, sizeof(hbad), hipMemcpyDeviceToHost, ctx->stream));
  GDB_HIP(hipStreamSynchronize(ctx->stream));
  *violations = (long long)hbad;
  rc = check_device_error(ctx);
  if (rc) return rc;
  if (participating) return read_participating(ctx, participating);
  return GDB_SUCCESS;
}

int gdb_relaunch_bench(gdb_ctx* ctx, int wgs, int groups, int iters, int mode, double* ms) {
  if (!ctx || wgs < 1 || wgs > 1024 || groups < 1 || iters < 1 || (mode != 0 && mode != 1))
    return GDB_E_INVALID;
  int rc = ensure_workspace(ctx, 256);
  if (rc) return rc;
  int* sink = (int*)ctx->ws;
  // Warm the launch path once.
  hipLaunchKernelGGL(gdb_k_empty, dim3(groups), dim3(wgs), 0, ctx->stream, sink);
  GDB_HIP(hipStreamSynchronize(ctx->stream));
  auto t0 = std::chrono::steady_clock::now();
  GDB_HIP(hipEventRecord(ctx->ev0, ctx->stream));
  for (int i = 0; i < iters; ++i) {
    hipLaunchKernelGGL(gdb_k_empty, dim3(groups), dim3(wgs), 0, ctx->stream, sink);
    if (mode == 1) {
      GDB_HIP(hipMemcpyAsync(ctx->h_flag, sink, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
      GDB_HIP(hipStreamSynchronize(ctx->stream));
    }
  }
  GDB_HIP(hipGetLastError());
  GDB_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  GDB_HIP(hipEventSynchronize(ctx->ev1));
  auto t1 = std::chrono::steady_clock::now();
  float f = 0.f;
  GDB_HIP(hipEventElapsedTime(&f, ctx->ev0, ctx->ev1));
  double host_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  if (ms) *ms = mode == 0 ? (double)f : host_ms;
  return GDB_SUCCESS;
}

}  // extern "C"
