// Shared helpers of the command-line tools (host C++ over the C ABI).
#pragma once
#include <cstdio>
#include <cstdlib>
#include <string>
#include "gdb_amd.h"
#include "gdb_graph.h"

namespace cli {

[[noreturn]] void die(const char* fmt, ...);
double now_ms();
// Device copy of a host CSR (row, col, w).
struct DevCSR {
  int n = 0, m = 0;
  int* row = nullptr;
  int* col = nullptr;
  int* w = nullptr;
};
DevCSR upload(const gdb_host_csr& g);
void release(DevCSR& d);
void* dalloc(size_t bytes);
void to_host(void* dst, const void* src, size_t bytes);
// Launch options from the environment: GDB_VARIANT (plain|tuned),
// GDB_GROUPS (groups to launch, 0 = residency bound).
gdb_opts opts_from_env(int wgs, int schedule);
void check(int rc, const char* what);

}  // namespace cli
