/* C ABI of libgdb_amd.so: occupancy discovery, inter-workgroup barriers and
 * the persistent graph kernels built on them, for AMD Instinct MI355X
 * (gfx950).
 *
 * Every entry point returns an int: GDB_SUCCESS (0) or a negative GDB_E_*
 * code.  Graph arrays passed to the solvers are DEVICE pointers (use
 * gdb_malloc / gdb_memcpy_htod, or any HIP allocation); all work runs on the
 * context's stream and the call returns after it finished.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   gdb_ctx_create + gdb_ctx_init
 *        clCreateBuffer(sizeof(discovery_kernel_ctx)) + init_discovery_kernel_ctx
 *        (code/discovery_protocol/api/discovery.h:10-32; e.g.
 *         code/experiments/pannotia/graph_app/color/color_gb.cpp:182-185)
 *   gdb_ctx_init_skip          init_discovery_kernel_ctx_skip  (discovery.h:36-59)
 *   gdb_number_of_participating_groups
 *                              number_of_participating_groups  (discovery.h:63-68)
 *   gdb_time_protocol_kernel   time_protocol                   (discovery.h:71-105)
 *   gdb_time_protocol          time_protocol on run_prot (occupancy_test.cl:15-17)
 *   gdb_ctx_device_ptr         the context cl_mem handed to kernels (color_gb.cpp:182-185,248)
 *   gdb_launch_on_ctx          clEnqueueNDRangeKernel of a user kernel on the context
 *                              (e.g. lonestar bfs_port.cpp:139-191)
 *   gdb_occupancy_test         occupancy_test main + run_test kernel
 *        (code/experiments/occupancy_tests/src/occupancy_test.c:21-89,
 *         kernels/occupancy_test.cl:8-11)
 *   gdb_device_query           device_query / print_device_info
 *        (occupancy_tests/src/device_query.c; common/include/OpenCL/my_opencl.h:193-217)
 *   gdb_color                  color-gb mega_kernel / color relaunch loop
 *        (pannotia/graph_app/color/kernel/color_kernel.cl:80-252; color.cpp:294-334)
 *   gdb_mis                    mis-gb mega_kernel / mis relaunch loop
 *        (pannotia/graph_app/mis/kernel/mis_kernel.cl:78-337; mis.cpp:314-363)
 *   gdb_sssp_pull              sssp-gb mega_kernel / sssp relaunch loop
 *        (pannotia/graph_app/sssp/kernel/sssp_kernel.cl:79-244; sssp.cpp:290-344)
 *   gdb_bc                     bc-gb mega_bc_kernel / bc relaunch loop
 *        (pannotia/graph_app/bc/kernel/bc_kernel.cl:95-513; bc.cpp:322-419)
 *   gdb_bfs_worklist           bfs-port drelax2 + dverifysolution
 *        (lonestar_gpu_opencl/apps/bfs/kernel/bfs_port_kernel.cl:35-192,
 *         bfs_common.cl:43-90; bfs_port.cpp:62-215)
 *   gdb_sssp_worklist          sssp-port drelax2 + dverifysolution
 *        (lonestar_gpu_opencl/apps/sssp/kernel/sssp_port_kernel.cl:42-197,
 *         sssp_common.cl:41-81; sssp_port.cpp:65-178)
 * Device side (kernels that include gdb_discovery.hpp): DISCOVERY_PROTOCOL,
 * discovery_barrier, p_get_*, is_participating (discovery.cl:106-226).
 */
#ifndef GDB_AMD_H
#define GDB_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes ---------------------------------------------------- */
#define GDB_SUCCESS 0
#define GDB_E_INVALID (-1)   /* bad argument                                   */
#define GDB_E_HIP (-2)       /* HIP runtime error (message printed to stderr)  */
#define GDB_E_TIMEOUT (-3)   /* a device-side wait hit its deadline            */
#define GDB_E_UNSAFE (-4)    /* skip-mode launch larger than the residency bound */
#define GDB_E_OVERFLOW (-5)  /* a device queue overflowed                      */
#define GDB_E_NOMEM (-6)     /* device allocation failed                       */
#define GDB_E_SCRATCH (-7)   /* input larger than the library's static scratch
                                (gdb_sssp_worklist_noalloc: n > GDB_NOALLOC_MAX_N) */

/* ---- protocol / barrier / schedule selectors ---------------------------- */
#define GDB_PROTOCOL_LOCKFREE 0 /* single atomic ticket word (default)        */
#define GDB_PROTOCOL_TICKET 1   /* fair ticket lock, locks.cl:25-34           */
#define GDB_PROTOCOL_SPIN 2     /* unfair spin lock, locks.cl:11-20           */

#define GDB_BARRIER_XF 0       /* master/slave flags, discovery.cl:129-190    */
#define GDB_BARRIER_COUNTER 1  /* central sense-reversing counter             */
#define GDB_BARRIER_TREE 2     /* two-level counter: 16 clusters, then their leaders */
#define GDB_BARRIER_AUTO 3     /* COUNTER below 64 groups, TREE above (default) */

#define GDB_SCHED_PERSISTENT 0 /* one launch, inter-workgroup barriers        */
#define GDB_SCHED_RELAUNCH 1   /* one launch per phase, host loop             */

#define GDB_VARIANT_PLAIN 0    /* phase structure of the reference kernels    */
#define GDB_VARIANT_TUNED 1    /* gfx950 kernels: frontier queues, wave-level work split */

typedef struct gdb_ctx gdb_ctx;

typedef struct {
  int wgs;            /* threads per workgroup: 64,128,256,512,1024 (0 = 256) */
  int groups;         /* groups launched (0 = device residency bound)          */
  int schedule;       /* GDB_SCHED_*                                           */
  int variant;        /* GDB_VARIANT_*                                         */
  int barrier;        /* GDB_BARRIER_* or -1 for the context default           */
  int relaunch_check; /* relaunch schedule: rounds between host stop checks (0/1 = every round) */
  int lds_queue;      /* batched engine: ints per LDS queue buffer of a group's local hops,
                         256..4096 (0 = the tuning table's value, else 2048) */
} gdb_opts;

typedef struct {
  int participating;     /* groups that took part (discovered P)              */
  int rounds;            /* rounds / levels / iterations executed             */
  long long barriers;    /* inter-workgroup barriers executed (per group)     */
  long long launches;    /* kernel launches in the solve                      */
  double kernel_ms;      /* solve time, device events around the kernels      */
  long long work_edges;  /* edges examined (for algorithmic-byte accounting)  */
  long long work_items;  /* vertices / worklist items processed               */
} gdb_stats;

/* ---- device memory helpers ------------------------------------------------ */
int gdb_malloc(void** ptr, size_t bytes);
int gdb_free(void* ptr);
int gdb_memcpy_htod(void* dst, const void* src, size_t bytes);
int gdb_memcpy_dtoh(void* dst, const void* src, size_t bytes);
int gdb_memcpy_dtod(void* dst, const void* src, size_t bytes);
int gdb_device_synchronize(void);
int gdb_device_count(int* out);

/* ---- discovery context ------------------------------------------------------ */
int gdb_ctx_create(gdb_ctx** out);
int gdb_ctx_destroy(gdb_ctx* ctx);
/* lock_kind: GDB_PROTOCOL_*, bar_kind: GDB_BARRIER_*; takes effect at next init. */
int gdb_ctx_configure(gdb_ctx* ctx, int lock_kind, int bar_kind);
/* Restricts participation to the groups running on one XCD (0..7), or
 * lifts the restriction (-1, default); takes effect at next init.  Such
 * launches synchronise through that XCD's shared L2 (cheapest barrier). */
int gdb_ctx_set_xcd_only(gdb_ctx* ctx, int xcd);
/* Bounds every device wait (default 10 s) and the lock-free close's quiet
 * period (default 20 us); both in microseconds.  wait_us = 0 makes any wait
 * that would block fail at once with GDB_E_TIMEOUT. */
int gdb_ctx_set_timeouts(gdb_ctx* ctx, double wait_us, double quiet_us);
int gdb_ctx_init(gdb_ctx* ctx);
int gdb_ctx_init_skip(gdb_ctx* ctx, int skip);
/* Returns P of the last launch (>= 0) or a negative error. */
int gdb_number_of_participating_groups(gdb_ctx* ctx);
/* Device-side sticky error of the last launch as a GDB_E_* code. */
int gdb_ctx_device_error(gdb_ctx* ctx);
void* gdb_ctx_stream(gdb_ctx* ctx);

/* The context's device buffer (a gdb_kernel_ctx_t*, include/gdb_ctx_layout.h),
 * to pass to a caller's own persistent kernel that includes
 * gdb_discovery.hpp and opens with DISCOVERY_PROTOCOL(ctx).  It is the
 * cl_mem the reference's hosts create and hand to their kernels
 * (pannotia/graph_app/color/color_gb.cpp:182-185, clSetKernelArg at :248).
 * The buffer lives until gdb_ctx_destroy. */
int gdb_ctx_device_ptr(gdb_ctx* ctx, void** out);

/* Launches a caller's kernel (a __global__ function's host symbol, as taken
 * by hipLaunchKernel; `args` as for hipLaunchKernel) with `groups` groups of
 * `wgs` threads on the context's stream, waits, and reports the device
 * error word (GDB_E_TIMEOUT when a barrier wait gave up).  Clears the error
 * word first; does not re-initialise the context, which resets itself after
 * every launch (discovery.cl:67-84), as the reference's apps rely on
 * (lonestar bfs_port.cpp:139-191 launches twice after one init). */
int gdb_launch_on_ctx(gdb_ctx* ctx, const void* kernel, void** args, int groups, int wgs, size_t dyn_lds);

/* time_protocol (discovery.h:71-105): initialises the context, then times
 * one launch of the CALLER's kernel with `upper_bound` groups of `wgs`
 * threads (`dyn_lds` bytes of dynamic LDS) on the context's stream; *ns is
 * the device time.  The kernel's arguments must include the context's
 * device pointer (gdb_ctx_device_ptr), as the reference's caller sets
 * argument 0 to the context buffer (occupancy_tests/src/time_prot.c:57-62). */
int gdb_time_protocol_kernel(gdb_ctx* ctx, const void* kernel, void** args, int upper_bound, int wgs,
                             size_t dyn_lds, double* ns);

/* time_protocol with the library's own protocol-only kernel (run_prot,
 * occupancy_tests/kernels/occupancy_test.cl:15-17) as the caller's kernel:
 * `upper_bound` groups of `wgs` threads with `lms` bytes of LDS; *ns is its
 * device time. */
int gdb_time_protocol(gdb_ctx* ctx, int upper_bound, int wgs, int lms, double* ns);
/* Protocol (prot=1) or skip mode (prot=0) followed by one barrier, with wgc
 * groups of wgs threads and lms bytes of LDS.  Skip-mode launches beyond the
 * residency bound would deadlock on any GPU and are refused (GDB_E_UNSAFE). */
int gdb_occupancy_test(gdb_ctx* ctx, int wgc, int wgs, int lms, int prot, int* participating);
/* Co-resident groups of the occupancy-test kernel for (wgs, lms). */
int gdb_max_resident_groups(int wgs, int lms, int* out);
/* Writes a human-readable device description. */
int gdb_device_query(char* buf, size_t len);

/* Barrier cost: a persistent kernel running `iters` barriers (or none, for
 * iters = 0) on `groups` groups; returns device ms and the discovered P. */
int gdb_barrier_bench(gdb_ctx* ctx, int wgs, int groups, int iters, double* ms, int* participating);
/* Barrier invariants over `iters` barriers: per-round arrival counts equal
 * P * k and plain stores made before a barrier are visible after it;
 * returns the number of violations seen. */
int gdb_barrier_check(gdb_ctx* ctx, int wgs, int groups, int iters, long long* violations,
                      int* participating);
/* Relaunch cost: `iters` launches of an empty kernel of `groups` groups.
 * mode 0: back-to-back async launches, one final sync.
 * mode 1: reference host loop, blocking 4-byte readback after every launch. */
int gdb_relaunch_bench(gdb_ctx* ctx, int wgs, int groups, int iters, int mode, double* ms);

/* Memory-latency probes (no reference counterpart; they measure what the
 * barrier and worklist designs are built on).  gdb_probe_chain: one lane
 * follows a random cycle of `steps` dependent accesses, one per 128 B line,
 * over a buffer of `bytes` bytes (fine != 0: fine-grained memory); kind 0 a
 * plain load, 1 a returning agent-scope atomic add, 2 an agent-scope relaxed
 * atomic load.  *ns_per_step = device time / steps after a warm-up pass.
 * gdb_probe_atomics: every lane of `groups` x `wgs` does `iters` dependent
 * returning atomic-min operations (bits 32 or 64) on random lines of a
 * 128 MB buffer; *ns_per_step = device time / iters. */
int gdb_probe_chain(gdb_ctx* ctx, int kind, size_t bytes, int fine, int steps, double* ns_per_step);
int gdb_probe_atomics(gdb_ctx* ctx, int bits, int groups, int wgs, int iters, double* ns_per_step);
/* Two groups (on one XCD when same_xcd != 0, else on two) hand a counter back
 * and forth `iters` times each; *ns_per_handoff = device time / (2 iters).
 * mode 0 relaxed store / relaxed poll, 1 release store / poll + acquire,
 * 2 acq_rel fetch-add / poll + acquire; fine != 0 puts the counter in
 * fine-grained memory.  xcds (2 ints, may be NULL) receive the groups' XCDs.
 * The floor under any inter-group barrier: a barrier is at least one
 * arrival and one release handoff. */
int gdb_probe_pingpong(gdb_ctx* ctx, int mode, int fine, int same_xcd, int iters, double* ns_per_handoff,
                       int* xcds);

/* ---- launch-shape tuning table -----------------------------------------------
 * Counterpart of the reference's tuned work-group-size table
 * (pannotia/scripts/wg_size_data.py:7-79), written by its tuner
 * (tune_wg_size.py:106-124) and read by its suite runner.  Here the tuner
 * (python -m gpu_discovery_barrier_amd.autotune) times launch shapes on the
 * device and writes a text table, one row per line:
 *     solver n_max wgs groups team_groups delta hops lds_queue
 * (solver: color, mis, bc, sssp_pull, bfs_wl, sssp_wl, mst, batch; '#'
 * starts a comment).  The library loads it once per process, at the first
 * gdb_ctx_create, from $GDB_TUNE_FILE or else tune_table.txt beside
 * libgdb_amd.so.  A solver called with opts == NULL or opts->wgs == 0 takes
 * wgs, groups (when opts->groups == 0) and lds_queue from the row for its
 * solver with the smallest n_max >= n (the tuned persistent variant only;
 * the plain variant keeps the reference's default launch); gdb_batch_solve
 * also takes team_groups when it is 0, keyed by the largest task's n.
 * Without a row the defaults stand (256 threads, the residency bound). */
typedef struct {
  char solver[16];
  long long n_max;
  int wgs, groups, team_groups, delta, hops, lds_queue;
} gdb_tune_entry;
/* Replaces the table with the file's rows (path NULL: the default path);
 * returns the number of rows, or GDB_E_INVALID for an unreadable or
 * malformed file (the table is then empty). */
int gdb_tune_load(const char* path);
/* The row a solver of graph size n would use: GDB_SUCCESS, or GDB_E_INVALID
 * when there is none. */
int gdb_tune_lookup(const char* solver, long long n, gdb_tune_entry* out);

/* ---- solvers (device pointers; CSR int32, row has n+1 entries) ------------- */
/* Pannotia colouring (max rule): color[v] = round in which v was coloured. */
int gdb_color(gdb_ctx* ctx, int n, int m, const int* row, const int* col, const float* node_value,
              int* color, const gdb_opts* opts, gdb_stats* stats);
/* Max-min colouring: the two-sided variant of the max rule above, which the
 * reference does not implement (it colours with the max rule only,
 * color_kernel.cl:229-237).  Round r gives colour 2r-1 to every uncoloured
 * vertex whose value exceeds all its uncoloured neighbours' and 2r to every
 * one below all of them (self-loops ignored); a proper colouring for
 * distinct values, in about half the max rule's rounds.  One persistent
 * launch, two barriers per round; stats->rounds = rounds.  Adjacent equal
 * values never resolve: GDB_E_TIMEOUT after n+1 rounds. */
int gdb_color_maxmin(gdb_ctx* ctx, int n, int m, const int* row, const int* col, const float* node_value,
                     int* color, const gdb_opts* opts, gdb_stats* stats);
/* Pannotia MIS: s[v] = 2 for set members, 0 otherwise. */
int gdb_mis(gdb_ctx* ctx, int n, int m, const int* row, const int* col, const float* node_value,
            int* s, const gdb_opts* opts, gdb_stats* stats);
/* Pannotia pull Bellman-Ford over the transposed CSR (in-edges), source given;
 * unreachable = 99999999. */
int gdb_sssp_pull(gdb_ctx* ctx, int n, int m, const int* row_t, const int* col_t, const int* w_t,
                  int source, int* dist, const gdb_opts* opts, gdb_stats* stats);
/* Brandes betweenness over `nsrc` sources (FP32 sigma / delta / bc).  With the
 * tuned persistent schedule a symmetric graph (pass the CSR as its own
 * transpose: row_t == row, col_t == col) is swept eight sources at a time;
 * otherwise one source at a time with separate in- and out-edge walks. */
int gdb_bc(gdb_ctx* ctx, int n, int m, const int* row, const int* col, const int* row_t,
           const int* col_t, const int* sources, int nsrc, float* bc, const gdb_opts* opts,
           gdb_stats* stats);
/* Lonestar worklist BFS: levels from `source`, unreachable = 1000000000. */
int gdb_bfs_worklist(gdb_ctx* ctx, int n, int m, const int* row, const int* col, int source,
                     unsigned* dist, const gdb_opts* opts, gdb_stats* stats);
/* Lonestar worklist SSSP (label-correcting, atomic min). */
int gdb_sssp_worklist(gdb_ctx* ctx, int n, int m, const int* row, const int* col, const int* w,
                      int source, unsigned* dist, const gdb_opts* opts, gdb_stats* stats);
/* gdb_sssp_worklist without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot call hipMalloc (code under HIP
 * graph capture, tight serving loops, harnesses that forbid allocation).
 * Same arguments, schedules, results and stats as gdb_sssp_worklist; its
 * queues and counters live in a file-scope device array of the library
 * (48 MiB + 4 KiB per GPU, held until the process ends) instead of the
 * context's workspace, so n is limited to GDB_NOALLOC_MAX_N (larger inputs
 * return GDB_E_SCRATCH).  Every call on one GPU shares that GPU's array and
 * re-initialises what it reads; calls that overlap on one GPU (from
 * several threads or contexts) race and must be serialised by the caller. */
#define GDB_NOALLOC_MAX_N (1 << 22)
int gdb_sssp_worklist_noalloc(gdb_ctx* ctx, int n, int m, const int* row, const int* col, const int* w,
                              int source, unsigned* dist, const gdb_opts* opts, gdb_stats* stats);

/* dverifysolution: counts edges (u,v,w) with dist[u] + w < dist[v]; w = 1
 * when `w` is NULL (BFS). */
int gdb_verify_distances(gdb_ctx* ctx, int n, int m, const int* row, const int* col, const int* w,
                         const unsigned* dist, long long* nerr);

/* Minimum spanning forest (Boruvka) in one persistent launch.  Replaces the
 * lonestar mst-port host loop and kernels (lonestar_gpu_opencl/apps/mst/
 * mst_port.cpp:203-313; kernel/mst_common.cl, mst_port_kernel.cl).  The graph
 * must be symmetric (each undirected edge stored in both directions, as the
 * reference's .sym.gr inputs).  Outputs: total weight of the forest, its
 * edge count and the number of components (isolated vertices included), as
 * mst-port prints them ("weight: %u, components: %u, edges: %u"). */
int gdb_mst(gdb_ctx* ctx, int n, int m, const int* row, const int* col, const unsigned* w,
            unsigned long long* weight, int* edges, int* components, const gdb_opts* opts,
            gdb_stats* stats);

/* ---- batched engine -------------------------------------------------------
 * One persistent launch solves many (graph, algorithm) tasks: discovered
 * groups form teams of `team_groups` groups; each team takes tasks from a
 * shared counter and synchronises on its own sub-barrier (a team of one group
 * uses workgroup barriers only).  Outputs equal the single-task solvers'. */
#define GDB_TASK_BFS 0    /* Lonestar worklist BFS levels (u32, 1e9 unreachable) */
#define GDB_TASK_SSSP 1   /* worklist SSSP distances (u32); near-far when delta > 0 */
#define GDB_TASK_COLOR 2  /* Pannotia max-rule colouring (i32) */

typedef struct {
  int algo;                 /* GDB_TASK_* */
  int n, m;
  const int* row;           /* device CSR */
  const int* col;
  const int* w;             /* SSSP weights */
  int source;               /* BFS / SSSP */
  const float* node_value;  /* colour priorities */
  void* out;                /* device output: u32 distances or i32 colours */
  int delta;                /* SSSP near-far band width (0 = plain worklist) */
  int hops;                 /* BFS / SSSP: local hops a group may run past its share of
                               an iteration before the team barrier, 0..4096
                               (0 = level-synchronous) */
  int owned;                /* BFS / SSSP, 0 or 1: owner-partitioned solve.  Group r of the
                               team owns vertices [r*B, r*B + B), B = ceil(n / team groups)
                               rounded up to 32, and keeps their distances and row offsets
                               in LDS; only relaxations between ranges go through device
                               memory.  Suits graphs whose ids follow locality (grids,
                               meshes, BFS-ordered graphs).  Needs B <= 65536 and the LDS
                               of one group per CU; otherwise the task runs as owned = 0.
                               hops is ignored. */
  const int* trow;         /* in-edge CSR (the transpose; the CSR itself for a symmetric
                               graph) for pull iterations; NULL = push only */
  const int* tcol;
  const int* tw;            /* in-edge weights (SSSP) */
  int rounds;               /* out: levels / iterations / colour rounds */
  double time_us;           /* out: device time from task start to task end */
  int arm_iters[4];         /* out: relaxing iterations (non-empty frontier) run with each
                               schedule arm (GDB_ARM_*) */
  long long work_items;     /* out: BFS / SSSP worklist items processed (duplicates included);
                               colour: n */
  long long work_edges;     /* out: BFS / SSSP edges examined; colour: m */
} gdb_task;

/* team_groups: groups per team (0 = 1).  opts->wgs and opts->groups pick the
 * launch shape.  Teams take the tasks longest first (estimated (n + m),
 * SSSP before BFS before colour).  stats->rounds is the sum of task rounds. */
int gdb_batch_solve(gdb_ctx* ctx, gdb_task* tasks, int ntasks, int team_groups, const gdb_opts* opts,
                    gdb_stats* stats);

/* ---- per-iteration schedule selector ---------------------------------------
 * No reference counterpart: the reference fixes one schedule per binary
 * (push worklists in lonestar bfs_port_kernel.cl / sssp_port_kernel.cl, pull
 * in pannotia sssp_kernel.cl:79-165).  Here every BFS / SSSP iteration of a
 * batched task picks one of four arms; all arms reach the same fixpoint, so
 * outputs are identical whatever is chosen:
 *   PUSH          relax the out-edges of the frontier queue, whole team
 *   PULL          every vertex scans its in-edges (bottom-up BFS: unvisited
 *                 vertices look for a parent on the previous level; SSSP: a
 *                 Jacobi/Gauss-Seidel min over in-edges), whole team
 *   *_NARROW      the same with only the team's first `narrow_groups` groups;
 *                 the rest park on a team word until the schedule widens.
 * The group that completes each team barrier evaluates the selector and
 * publishes the arm with the barrier release.  MLP mode evaluates a
 * 16-64-64-4 network (bf16 weights, fp32 accumulation) whose outputs predict
 * log2 of the iteration's time in microseconds per arm, and takes the valid
 * arm with the smallest prediction; it reruns the network only when the
 * frontier has moved by over 25% (or after 32 iterations).  The choosing
 * modes offer pull arms to BFS only (a BFS iteration makes the same progress
 * under every arm, an SSSP pull iteration does not); GDB_SEL_FIXED may pull
 * SSSP too. */
#define GDB_SEL_NONE 0       /* PUSH on the whole team every iteration (the plain schedule) */
#define GDB_SEL_FIXED 1      /* fixed_arm every iteration (when valid) */
#define GDB_SEL_HEURISTIC 2  /* direction-optimising rule (alpha 14, beta 24) + narrow when the
                                frontier has fewer edges than the narrow groups have threads */
#define GDB_SEL_RANDOM 3     /* uniformly random valid arm (trace collection) */
#define GDB_SEL_MLP 4        /* smallest predicted cost; the heuristic when weights == NULL */

#define GDB_ARM_PUSH 0
#define GDB_ARM_PULL 1
#define GDB_ARM_PUSH_NARROW 2
#define GDB_ARM_PULL_NARROW 3

#define GDB_MLP_IN 16
#define GDB_MLP_HIDDEN 64
#define GDB_MLP_OUT 4
/* Parameter vector: W1[16][64], b1[64], W2[64][64], b2[64], W3[64][4], b3[4]
 * (input-major: layer output j = b[j] + sum_i W[i][j] * x[i]). */
#define GDB_MLP_PARAMS 5508
/* Trace record (floats): features[16], arm, cost_us (-1 until the iteration
 * ends), task index, iteration. */
#define GDB_TRACE_STRIDE 20

typedef struct {
  int mode;                       /* GDB_SEL_* */
  int fixed_arm;                  /* GDB_SEL_FIXED */
  int narrow_groups;              /* groups of a narrow iteration (0 = 1) */
  unsigned seed;                  /* GDB_SEL_RANDOM */
  const unsigned short* weights;  /* device bf16 parameters (GDB_MLP_PARAMS) or NULL */
  float* trace;                   /* device trace records or NULL */
  int trace_capacity;             /* records that fit in `trace` */
  int* trace_count;               /* device counter; may end above the capacity */
} gdb_selector;

/* gdb_batch_solve with a selector for the BFS / SSSP tasks (sel == NULL is
 * GDB_SEL_NONE).  Tasks with hops > 0 always push.  Selector modes other
 * than NONE need n < 2^30 per task. */
int gdb_batch_solve_sel(gdb_ctx* ctx, gdb_task* tasks, int ntasks, int team_groups,
                        const gdb_opts* opts, const gdb_selector* sel, gdb_stats* stats);

/* The same batch recorded once as a HIP graph (sel may be NULL; the clearing kernels and the
 * persistent launch) and replayed by gdb_batch_graph_launch; the task array
 * must stay alive (each replay writes rounds, time_us and arm_iters into
 * it), and so must the graphs and outputs it points to.  The handle owns its
 * device memory.  Replays on one context must not overlap. */
typedef struct gdb_batch_graph gdb_batch_graph;
int gdb_batch_graph_create(gdb_ctx* ctx, gdb_task* tasks, int ntasks, int team_groups, const gdb_opts* opts,
                           const gdb_selector* sel, gdb_batch_graph** out);
int gdb_batch_graph_launch(gdb_batch_graph* g, gdb_stats* stats);
int gdb_batch_graph_destroy(gdb_batch_graph* g);


/* ---- MLP forward / backward / training (device buffers) ----------------- */
/* y[count][4] = MLP(x[count][16]) with bf16 weights. */
int gdb_mlp_forward(gdb_ctx* ctx, const unsigned short* w_bf16, const float* x, int count, float* y);
/* Loss and gradient of L = mean over records with cost > 0 of
 * (y[arm] - log2(cost_us))^2, evaluated with bf16 weights; grad has
 * GDB_MLP_PARAMS floats.  *used = records with cost > 0 (host pointers for
 * loss and used). */
int gdb_mlp_grad(gdb_ctx* ctx, const unsigned short* w_bf16, const float* records, int count,
                 float* grad, float* loss, int* used);
/* Rounds fp32 parameters to bf16 (round to nearest even). */
int gdb_mlp_to_bf16(gdb_ctx* ctx, const float* params, unsigned short* w_bf16, int count);
/* Adam on the fp32 master parameters (m, v: GDB_MLP_PARAMS floats each, zero
 * at the start): `steps` steps over minibatches of `batch` records drawn by
 * a counter-based generator from `seed`; forward / backward use the bf16
 * copy, which is refreshed after every step.  *loss = full-set loss at the
 * end (host pointer). */
int gdb_mlp_train(gdb_ctx* ctx, float* params, float* adam_m, float* adam_v, unsigned short* w_bf16,
                  const float* records, int count, int steps, int batch, float lr, unsigned seed,
                  float* loss);

#ifdef __cplusplus
}
#endif
#endif /* GDB_AMD_H */
