// Layout of the device-resident discovery/barrier context, shared by host
// and device code.
//
// Counterpart of `discovery_kernel_ctx` / `discovery_mutex`
// (code/discovery_protocol/api/common.h:27-48).  Differences, all driven by
// gfx950 (wave64, 256 CUs in 8 XCDs, one L2 per XCD, 128 B L2 lines):
//   * every protocol word and every barrier flag sits on its own 128 B line,
//     so a poller never shares a line with a word another group writes;
//   * the flag array holds GDB_MAX_GROUPS entries (the reference caps
//     participation at 1024, common.h:20; a gfx950 part co-schedules up to
//     256 CUs x 32 groups);
//   * a sticky error word and a wall-clock deadline bound every wait, so a
//     barrier whose peers never arrive ends the kernel instead of hanging it.
#pragma once

#define GDB_LINE_BYTES 128
#define GDB_MAX_GROUPS 8192

// Lock used by the discovery protocol (reference: locks.cl).
#define GDB_LOCK_FREE 0    // single atomic ticket word, no mutex (gfx950 default)
#define GDB_LOCK_TICKET 1  // fair ticket lock, locks.cl:25-34
#define GDB_LOCK_SPIN 2    // unfair exchange lock, locks.cl:11-20

// Inter-workgroup barrier flavours.
#define GDB_BAR_XF 0       // master/slave flags, discovery.cl:129-190
#define GDB_BAR_COUNTER 1  // central sense-reversing arrival counter
#define GDB_BAR_TREE 2     // two-level counter: clusters of groups, then cluster leaders
#define GDB_BAR_AUTO 3     // COUNTER for small P, TREE otherwise

#define GDB_MAX_XCDS 16        // XCD ids read from the hardware are masked to this
#define GDB_TREE_CLUSTERS 16   // clusters of the two-level counter
#define GDB_AUTO_TREE_MIN_P 64

// Error codes written to the context's error word and returned by the C ABI.
#define GDB_OK 0
#define GDB_ERR_BARRIER_TIMEOUT 1
#define GDB_ERR_DISCOVERY_TIMEOUT 2
#define GDB_ERR_QUEUE_OVERFLOW 3

typedef struct {
  int v;
  int pad[GDB_LINE_BYTES / 4 - 1];
} gdb_line_t;

typedef struct {
  // ---- discovery protocol ------------------------------------------------
  gdb_line_t poll;            // lock-free: bit 30 = poll closed, low bits = arrivals
  gdb_line_t prot_poll_open;  // lock variants: 1 while the poll is open
  gdb_line_t prot_counter;    // lock variants: next participating id
  gdb_line_t kernel_counter;  // groups through the closing phase (self-reset)
  gdb_line_t num_participating;  // P of the last launch (read by the host)
  gdb_line_t closing_p;       // lock-free: P published by the closer (0 = not yet)
  gdb_line_t lock_counter;    // ticket / spin lock word
  gdb_line_t lock_now_serving;
  // ---- configuration written by the host at init --------------------------
  gdb_line_t skip;            // 1: every launched group participates (no discovery)
  gdb_line_t lock_kind;       // GDB_LOCK_*
  gdb_line_t bar_kind;        // GDB_BAR_*
  gdb_line_t expect;          // lock-free close: stop waiting once this many arrived
  gdb_line_t quiet_ticks;     // lock-free close: ...or after this long with no arrival
  gdb_line_t timeout_ticks;   // bound on every wait (wall clock, 100 MHz ticks)
  gdb_line_t error;           // sticky GDB_ERR_* (0 = fine)
  gdb_line_t xcd_only;        // >= 0: only groups on this XCD participate
  // ---- counter / tree barrier state (self-resetting) -----------------------
  gdb_line_t bar_count;
  gdb_line_t bar_gen;
  gdb_line_t cluster_count[GDB_TREE_CLUSTERS];
  // ---- XF barrier flags ----------------------------------------------------
  gdb_line_t bar_flags[GDB_MAX_GROUPS];
} gdb_kernel_ctx_t;
