// Pannotia command-line tools: color, color-gb, mis, mis-gb, sssp, sssp-gb,
// bc, bc-gb.  Same usage, stdout strings and output files as the reference
// hosts (pannotia/graph_app/*/): `<exe> GRAPH_FILE GRAPH_FORMAT WORKGROUP_SIZE`
// with GRAPH_FORMAT 1 = METIS, 0 = DIMACS.  The -gb tools run the persistent
// (discovery + barrier) schedule, the others the relaunch schedule.
// Environment: GDB_VARIANT=plain|tuned, GDB_GROUPS=<n>, and for bc
// GDB_BC_SOURCES=<k> to use the first k vertices as sources (default: all,
// as bc_kernel.cl:456 does).
#include <cstring>
#include <vector>
#include "cli_common.h"

#if defined(GDB_APP_COLOR)
#define APP "color"
#elif defined(GDB_APP_MIS)
#define APP "mis"
#elif defined(GDB_APP_SSSP)
#define APP "sssp"
#else
#define APP "bc"
#endif
#ifdef GDB_APP_RELAUNCH
#define EXE APP
#define SCHED GDB_SCHED_RELAUNCH
#define OUTFILE APP ".out"
#else
#define EXE APP "-gb"
#define SCHED GDB_SCHED_PERSISTENT
#define OUTFILE APP "_gb.out"
#endif

using namespace cli;

static void usage() {
  printf("invalid usage:\n");
  printf("./" EXE " GRAPH_FILE GRAPH_FORMAT WORKGROUP_SIZE\n\n");
  exit(1);
}

int main(int argc, char** argv) {
  if (argc != 4) usage();
  const char* file = argv[1];
  const int format = atoi(argv[2]);
  const int wgs = atoi(argv[3]);
  if (format != 0 && format != 1) {
    printf("reserve for future");
    return 1;
  }
  gdb_host_csr g{};
  int rc;
#if defined(GDB_APP_SSSP)
  // sssp_gb.cpp:96,124-133: directed; METIS as is, DIMACS transposed.
  rc = format == 1 ? gdb_load_metis(file, 1, &g) : gdb_load_dimacs(file, 1, 1, &g);
#elif defined(GDB_APP_BC)
  // bc_gb.cpp:105,131: directed DIMACS.
  rc = format == 1 ? gdb_load_metis(file, 1, &g) : gdb_load_dimacs(file, 1, 0, &g);
#else
  // color_gb.cpp:101,128-135 and mis_gb.cpp: undirected.
  rc = format == 1 ? gdb_load_metis(file, 0, &g) : gdb_load_dimacs(file, 0, 0, &g);
#endif
  if (rc != GDB_GRAPH_OK) die("cannot read graph %s (code %d)", file, rc);
  const int n = g.n, m = g.m;
  printf("Read from file: num_nodes = %d, num_edges = %d\n", n, m);

  gdb_ctx* ctx = nullptr;
  check(gdb_ctx_create(&ctx), "gdb_ctx_create");
  check(gdb_ctx_init(ctx), "init_discovery_kernel_ctx");
  gdb_opts o = opts_from_env(wgs, SCHED);
  gdb_stats st{};
  const double t1 = now_ms();
  DevCSR d = upload(g);
  FILE* fp = fopen(OUTFILE, "w");
  if (!fp) printf("ERROR: unable to open result.txt\n");

#if defined(GDB_APP_COLOR) || defined(GDB_APP_MIS)
  // node_value[i] = i / (float)(n + 1) (color_gb.cpp:148-153, mis_gb.cpp:148-152).
  std::vector<float> nv(n);
  for (int i = 0; i < n; ++i) nv[i] = i / (float)(n + 1);
  float* dnv = (float*)dalloc(sizeof(float) * (n ? n : 1));
  int* dout = (int*)dalloc(sizeof(int) * (n ? n : 1));
  check(gdb_memcpy_htod(dnv, nv.data(), sizeof(float) * n), "gdb_memcpy_htod");
#if defined(GDB_APP_COLOR)
  check(gdb_color(ctx, n, m, d.row, d.col, dnv, dout, &o, &st), "gdb_color");
#else
  check(gdb_mis(ctx, n, m, d.row, d.col, dnv, dout, &o, &st), "gdb_mis");
#endif
  std::vector<int> out(n);
  to_host(out.data(), dout, sizeof(int) * n);
  const double t2 = now_ms();
  if (SCHED == GDB_SCHED_PERSISTENT) printf("kernel ran with a total of %d workgroups\n", st.participating);
#if defined(GDB_APP_COLOR)
  printf("total number of colors used: %d\n", st.rounds);
#endif
  printf("kernel time = %lf ms\n", st.kernel_ms);
  printf("kernel + memcpy time = %lf ms\n", t2 - t1);
  if (fp) {
    for (int i = 0; i < n; ++i) {
#if defined(GDB_APP_COLOR)
      fprintf(fp, "%d: %d\n", i + 1, out[i]);  // color_gb.cpp:394
#else
      fprintf(fp, "%d\n", out[i]);  // mis_gb.cpp:374
#endif
    }
  }
  gdb_free(dnv);
  gdb_free(dout);
#elif defined(GDB_APP_SSSP)
  int* dout = (int*)dalloc(sizeof(int) * (n ? n : 1));
  check(gdb_sssp_pull(ctx, n, m, d.row, d.col, d.w, 0, dout, &o, &st), "gdb_sssp_pull");  // source 0
  std::vector<int> out(n);
  to_host(out.data(), dout, sizeof(int) * n);
  const double t2 = now_ms();
  if (SCHED == GDB_SCHED_PERSISTENT) printf("kernel ran with a total of %d workgroups\n", st.participating);
  printf("kernel + memcpy time = %lf ms\n", t2 - t1);
  printf("kernel time = %lf ms\n", st.kernel_ms);
  if (fp)
    for (int i = 0; i < n; ++i) fprintf(fp, "%d: %d\n", i + 1, out[i]);  // sssp_gb.cpp:351
  gdb_free(dout);
#else
  gdb_host_csr gt{};
  check(gdb_csr_transpose(&g, &gt), "gdb_csr_transpose");
  DevCSR dt = upload(gt);
  int nsrc = n;
  if (const char* k = getenv("GDB_BC_SOURCES")) nsrc = std::min(n, atoi(k));
  std::vector<int> src(nsrc);
  for (int i = 0; i < nsrc; ++i) src[i] = i;
  int* dsrc = (int*)dalloc(sizeof(int) * (nsrc ? nsrc : 1));
  float* dout = (float*)dalloc(sizeof(float) * (n ? n : 1));
  check(gdb_memcpy_htod(dsrc, src.data(), sizeof(int) * nsrc), "gdb_memcpy_htod");
  check(gdb_bc(ctx, n, m, d.row, d.col, dt.row, dt.col, dsrc, nsrc, dout, &o, &st), "gdb_bc");
  std::vector<float> out(n);
  to_host(out.data(), dout, sizeof(float) * n);
  const double t2 = now_ms();
  if (SCHED == GDB_SCHED_PERSISTENT) printf("kernel ran with a total of %d workgroups\n", st.participating);
  printf("kernel time = %lf ms\n", st.kernel_ms);
  printf("kernel + memcpy time = %lf ms\n", t2 - t1);
  if (fp)
    for (int i = 0; i < n; ++i) fprintf(fp, "%f\n", out[i]);  // bc_gb.cpp:433
  gdb_free(dsrc);
  gdb_free(dout);
  release(dt);
  gdb_csr_free(&gt);
#endif
  if (fp) fclose(fp);
  release(d);
  gdb_csr_free(&g);
  gdb_ctx_destroy(ctx);
  return 0;
}
