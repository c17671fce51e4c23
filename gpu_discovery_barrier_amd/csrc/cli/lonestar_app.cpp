// Lonestar command-line tools: bfs-port, sssp-port, mst-port.  Usage and
// stdout strings of lonestar_gpu_opencl/apps/{bfs,sssp,mst}/*_port.cpp:
// `<exe> <graph.gr|graph.edges> <wgs>` (wgs a power of two); source 0.
// Environment: GDB_VARIANT=plain|tuned, GDB_GROUPS=<n>.
#include <cstring>
#include <vector>
#include "cli_common.h"

#if defined(GDB_APP_BFS)
#define EXE "bfs-port"
#elif defined(GDB_APP_MST)
#define EXE "mst-port"
#else
#define EXE "sssp-port"
#endif

using namespace cli;

int main(int argc, char** argv) {
  if (argc != 3) {
    printf("Usage: %s <graph> <wgs>\n", EXE);
    return 1;
  }
  const char* file = argv[1];
  const int wgs = atoi(argv[2]);
  if (wgs <= 0 || (wgs & (wgs - 1))) die("workgroup size must be a power of two");  // bfs_port.cpp:235-238
  gdb_host_csr g{};
  int rc = GDB_GRAPH_E_FORMAT;
  if (strstr(file, ".edges")) rc = gdb_load_edges(file, &g);
  else if (strstr(file, ".gr")) rc = gdb_load_galois_gr(file, &g);
  if (rc != GDB_GRAPH_OK) die("cannot read graph %s (code %d)", file, rc);
  printf("nnodes=%d, nedges=%d.\n", g.n, g.m);
  gdb_ctx* ctx = nullptr;
  check(gdb_ctx_create(&ctx), "gdb_ctx_create");
  check(gdb_ctx_init(ctx), "init_discovery_kernel_ctx");
  DevCSR d = upload(g);
  gdb_opts o = opts_from_env(wgs, GDB_SCHED_PERSISTENT);
  gdb_stats st{};
#ifdef GDB_APP_MST
  // mst_port.cpp:202-313: rounds until the component count stops falling;
  // prints weight (u32, as the reference's cl_uint), components and edges.
  printf("finding mst.\n");
  unsigned long long weight = 0;
  int edges = 0, comps = 0;
  check(gdb_mst(ctx, g.n, g.m, d.row, d.col, (const unsigned*)d.w, &weight, &edges, &comps, &o, &st), "gdb_mst");
  printf("\t%s result: weight: %u, components: %u, edges: %u\n", file, (unsigned)weight, (unsigned)comps,
         (unsigned)edges);
  printf("\tapp runtime = %f ms.\n", st.kernel_ms);
  printf("\tnumber of participating groups = %d\n", st.participating);
  release(d);
  gdb_csr_free(&g);
  gdb_ctx_destroy(ctx);
  return 0;
#endif
  unsigned* dist = (unsigned*)dalloc(sizeof(unsigned) * (g.n ? g.n : 1));
  printf("solving.\n");
  printf("starting...\n");
#ifdef GDB_APP_BFS
  check(gdb_bfs_worklist(ctx, g.n, g.m, d.row, d.col, 0, dist, &o, &st), "gdb_bfs_worklist");
#else
  check(gdb_sssp_worklist(ctx, g.n, g.m, d.row, d.col, d.w, 0, dist, &o, &st), "gdb_sssp_worklist");
#endif
  printf("\tapp runtime = %f ms.\n", st.kernel_ms);
  printf("\tnumber of participating groups = %d\n", st.participating);
  long long nerr = 0;
#ifdef GDB_APP_BFS
  check(gdb_verify_distances(ctx, g.n, g.m, d.row, d.col, nullptr, dist, &nerr), "dverifysolution");
#else
  check(gdb_verify_distances(ctx, g.n, g.m, d.row, d.col, d.w, dist, &nerr), "dverifysolution");
#endif
  printf("\tno of errors = %lld.\n", nerr);
  gdb_free(dist);
  release(d);
  gdb_csr_free(&g);
  gdb_ctx_destroy(ctx);
  return nerr == 0 ? 0 : 1;  // bfs_port.cpp:328-329 asserts no errors
}
