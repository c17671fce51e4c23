#include "cli_common.h"

#include <chrono>
#include <cstdarg>
#include <cstring>

namespace cli {

void die(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vfprintf(stderr, fmt, ap);
  va_end(ap);
  fputc('\n', stderr);
  exit(1);
}

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

void check(int rc, const char* what) {
  if (rc != GDB_SUCCESS) die("%s failed with code %d", what, rc);
}

void* dalloc(size_t bytes) {
  void* p = nullptr;
  check(gdb_malloc(&p, bytes), "gdb_malloc");
  return p;
}

void to_host(void* dst, const void* src, size_t bytes) { check(gdb_memcpy_dtoh(dst, src, bytes), "gdb_memcpy_dtoh"); }

DevCSR upload(const gdb_host_csr& g) {
  if (gdb_csr_validate(&g) != GDB_GRAPH_OK) die("malformed graph");
  DevCSR d;
  d.n = g.n;
  d.m = g.m;
  d.row = (int*)dalloc(sizeof(int) * ((size_t)g.n + 1));
  d.col = (int*)dalloc(sizeof(int) * (size_t)(g.m ? g.m : 1));
  d.w = (int*)dalloc(sizeof(int) * (size_t)(g.m ? g.m : 1));
  check(gdb_memcpy_htod(d.row, g.row, sizeof(int) * ((size_t)g.n + 1)), "gdb_memcpy_htod");
  check(gdb_memcpy_htod(d.col, g.col, sizeof(int) * (size_t)g.m), "gdb_memcpy_htod");
  check(gdb_memcpy_htod(d.w, g.w, sizeof(int) * (size_t)g.m), "gdb_memcpy_htod");
  return d;
}

void release(DevCSR& d) {
  gdb_free(d.row);
  gdb_free(d.col);
  gdb_free(d.w);
  d = DevCSR{};
}

gdb_opts opts_from_env(int wgs, int schedule) {
  gdb_opts o{};
  o.wgs = wgs;
  o.schedule = schedule;
  o.barrier = -1;
  o.relaunch_check = 1;
  const char* v = getenv("GDB_VARIANT");
  o.variant = (v && !strcmp(v, "tuned")) ? GDB_VARIANT_TUNED : GDB_VARIANT_PLAIN;
  const char* g = getenv("GDB_GROUPS");
  o.groups = g ? atoi(g) : 0;
  return o;
}

}  // namespace cli
