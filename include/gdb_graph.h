/* Host-side graph generation, conversion and file loading for libgdb_amd.so.
 *
 * Graphs are int32 CSR in host memory: row[n+1], col[m], w[m] (weights; 1
 * when a format has none).  Buffers are malloc'ed by the library and released
 * with gdb_csr_free.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   gdb_load_metis          parseMetis        pannotia/graph_parser/parse.cpp:59-176
 *   gdb_load_dimacs         parseCOO          pannotia/graph_parser/parse.cpp:178-282
 *   gdb_csr_transpose       parseCOO_transpose / parseMetis_transpose
 *                                             pannotia/graph_parser/parse.cpp:582-799
 *   gdb_load_galois_gr      readFromGR        lonestar_gpu_opencl/include/graph.h:143-215
 *   gdb_load_edges          readFromEdges     lonestar_gpu_opencl/include/graph.h:96-141
 * The generators have no counterpart: the reference only links external
 * datasets (pannotia/datasets/README.txt).
 */
#ifndef GDB_GRAPH_H
#define GDB_GRAPH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GDB_GRAPH_OK 0
#define GDB_GRAPH_E_INVALID (-1)
#define GDB_GRAPH_E_NOMEM (-6)
#define GDB_GRAPH_E_TOO_LARGE (-8)
#define GDB_GRAPH_E_IO (-9)
#define GDB_GRAPH_E_FORMAT (-10)

typedef struct {
  int n;    /* vertices */
  int m;    /* directed edges */
  int* row; /* n + 1 offsets */
  int* col; /* m destinations, sorted within each row for generated graphs */
  int* w;   /* m weights */
} gdb_host_csr;

void gdb_csr_free(gdb_host_csr* g);

/* R-MAT with 2^scale vertices and edge_factor * 2^scale sampled pairs,
 * quadrant probabilities (a, b, c, 1-a-b-c); symmetrised, self-loops and
 * duplicates removed; weights uniform in [wlo, whi] per undirected edge. */
int gdb_gen_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed, int wlo,
                 int whi, gdb_host_csr* out);
/* rows x cols 4-neighbour grid, vertex r*cols+c; exactly round(frac * E)
 * undirected edges removed by a seeded partial Fisher-Yates shuffle. */
int gdb_gen_grid(int rows, int cols, double remove_frac, uint64_t seed, int wlo, int whi,
                 gdb_host_csr* out);
/* Erdos-Renyi G(n, p) with p = avg_degree / (n - 1), by geometric skipping. */
int gdb_gen_gnp(int n, double avg_degree, uint64_t seed, int wlo, int whi, gdb_host_csr* out);
/* Seeded permutation of 0..n-1 (Fisher-Yates). */
int gdb_gen_permutation(int n, uint64_t seed, int* out);
/* `count` seeded integers uniform in [lo, hi]. */
int gdb_gen_uniform_ints(long long count, int lo, int hi, uint64_t seed, int* out);
/* In-edge CSR: row u of the output lists the sources of edges into u. */
int gdb_csr_transpose(const gdb_host_csr* in, gdb_host_csr* out);
/* Relabels a graph for locality: vertex v becomes new_id[v] (a permutation
 * of 0..n-1); rows are laid out in new-id order and each row's neighbours
 * sorted by new id.  Per-vertex results of a solve on the relabelled graph
 * map back as result_old[v] = result_new[new_id[v]]. */
int gdb_csr_relabel(const gdb_host_csr* in, const int* new_id, gdb_host_csr* out);
/* new_id ordering vertices by decreasing degree (ties by id): the hubs of a
 * power-law graph, which every traversal touches, share the first lines of
 * every per-vertex array. */
int gdb_order_by_degree(const gdb_host_csr* g, int* new_id);
/* GDB_GRAPH_OK if offsets are monotone from 0 to m and every destination is
 * in [0, n). */
int gdb_csr_validate(const gdb_host_csr* g);

/* ---- files (edge multiset and in-row order follow the reference parser;
 * row offsets are always a correct CSR, also for vertices without edges) -- */
/* METIS .graph: "n m" then one 1-based neighbour list per vertex.  With
 * directed = 0 the declared m counts undirected edges. */
int gdb_load_metis(const char* path, int directed, gdb_host_csr* out);
/* DIMACS "p sp n m" / "a u v w" (1-based).  directed = 0 adds v->u for every
 * arc; transpose = 1 stores in-edges the way parseCOO_transpose does. */
int gdb_load_dimacs(const char* path, int directed, int transpose, gdb_host_csr* out);
/* Galois binary .gr version 1 (u64 header, u64 end offsets, u32 targets,
 * 32-bit edge data). */
int gdb_load_galois_gr(const char* path, gdb_host_csr* out);
int gdb_write_galois_gr(const char* path, const gdb_host_csr* g);
/* Lonestar .edges text: "n m" then m lines "src dst weight" (0-based). */
int gdb_load_edges(const char* path, gdb_host_csr* out);

#ifdef __cplusplus
}
#endif
#endif /* GDB_GRAPH_H */
