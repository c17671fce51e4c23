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

#ifdef __cplusplus
}
#endif
#endif /* GDB_GRAPH_H */
