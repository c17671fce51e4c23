// Occupancy harness tools: occupancy_test, time_prot, device_query
// (occupancy_tests/src/*.c).
//   occupancy_test <wgc> <wgs> <lms> <prot> <ticket>
//   time_prot <wgc> <wgs> <lms> <ticket>
// `ticket`: 1 ticket lock, 0 spin lock (as the reference), 2 lock-free.
#include <cstring>
#include "cli_common.h"

using namespace cli;

static int lock_of(int ticket) {
  return ticket == 1 ? GDB_PROTOCOL_TICKET : ticket == 0 ? GDB_PROTOCOL_SPIN : GDB_PROTOCOL_LOCKFREE;
}

int main(int argc, char** argv) {
#if defined(GDB_TOOL_DEVICE_QUERY)
  (void)argc;
  (void)argv;
  char buf[4096];
  check(gdb_device_query(buf, sizeof(buf)), "gdb_device_query");
  fputs(buf, stdout);
  return 0;
#elif defined(GDB_TOOL_OCCUPANCY)
  if (argc != 6) {
    printf("please provide number of workgroups, workgroup size, local memory size, flag for protocol, flag for ticket lock\n");
    return 0;
  }
  const int wgc = atoi(argv[1]), wgs = atoi(argv[2]), lms = atoi(argv[3]), prot = atoi(argv[4]),
            ticket = atoi(argv[5]);
  printf("running with\nworkgroup count: %d\nworkgroup size: %d\nlocal memory size: %d\nprotocol enabled: %d\nusing ticket lock: %d\n",
         wgc, wgs, lms, prot, ticket);
  gdb_ctx* ctx = nullptr;
  check(gdb_ctx_create(&ctx), "gdb_ctx_create");
  check(gdb_ctx_configure(ctx, lock_of(ticket), GDB_BARRIER_XF), "gdb_ctx_configure");
  int p = 0;
  const int rc = gdb_occupancy_test(ctx, wgc, wgs, lms, prot, &p);
  if (rc == GDB_E_UNSAFE) {
    printf("refused: %d groups exceed the residency bound; a skip-mode barrier would never complete\n", wgc);
    return 2;
  }
  check(rc, "occupancy_test");
  printf("kernel ran with a total of %d workgroups\n", p);
  gdb_ctx_destroy(ctx);
  return 0;
#else
  if (argc != 5) {
    printf("please provide number of workgroups, workgroup size, local memory size, flag for ticket lock\n");
    return 0;
  }
  const int wgc = atoi(argv[1]), wgs = atoi(argv[2]), lms = atoi(argv[3]), ticket = atoi(argv[4]);
  printf("running with\nworkgroup count: %d\nworkgroup size: %d\nlocal memory size: %d\nusing ticket lock: %d\n",
         wgc, wgs, lms, ticket);
  gdb_ctx* ctx = nullptr;
  check(gdb_ctx_create(&ctx), "gdb_ctx_create");
  check(gdb_ctx_configure(ctx, lock_of(ticket), GDB_BARRIER_XF), "gdb_ctx_configure");
  double ns = 0;
  // time_prot.c:62 always launches 1000 groups; wgc is only echoed.
  check(gdb_time_protocol(ctx, 1000, wgs, lms, &ns), "time_protocol");
  printf("kernel ran with a total of %d workgroups\n", gdb_number_of_participating_groups(ctx));
  printf("kernel time: %f ms\n", ns / 1000000.0);
  gdb_ctx_destroy(ctx);
  return 0;
#endif
}
