/* CPU oracle for the discovery-barrier graph path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, always as the checker, never as the product.
 *
 * Two kinds of functions:
 *   orc_*_ref   restatements of the reference's rules, phase by phase, each
 *               citing the reference file:line it follows;
 *   orc_*_tb    independent textbook forms (greedy MIS, Dijkstra, queue BFS,
 *               pair-dependency betweenness, a longest-path recurrence for the
 *               colouring) used by the tests to pin the restatements.
 * Graphs are int32 CSR: row[n+1], col[m] (and w[m] where weighted).
 */
#ifndef GDB_ORACLE_H
#define GDB_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_PANNOTIA_INF 99999999    /* sssp_kernel.cl:68 */
#define ORC_LONESTAR_INF 1000000000u /* bfs_common.cl:35 */
#define ORC_MIS_BIGNUM 99999999.0f   /* mis_kernel.cl:68 */

/* Colouring, max rule.  Returns the number of rounds, or -1 if no round
 * limit of n+1 was reached without termination (ties livelock). */
int orc_color_ref(int n, const int* row, const int* col, const float* nv, int* color);
/* r(v) = 1 if deg(v) == 1, else 1 + max{r(u) : u in N(v), nv[u] > nv[v]}
 * (0 when empty); valid for distinct values >= 0. */
int orc_color_tb(int n, const int* row, const int* col, const float* nv, int* color);

/* Max-min colouring (the two-sided variant; no reference counterpart):
 * round r gives colour 2r-1 to uncoloured local maxima and 2r to uncoloured
 * local minima among uncoloured neighbours (self-loops ignored).  Returns
 * the rounds, or -1 if adjacent equal values livelock it. */
int orc_color_maxmin_ref(int n, const int* row, const int* col, const float* nv, int* color);

/* MIS (s = 2 in the set, 0 otherwise).  Returns rounds or -1 (livelock). */
int orc_mis_ref(int n, const int* row, const int* col, const float* nv, int* s);
/* Greedy sequential MIS in increasing value order (distinct values). */
int orc_mis_tb(int n, const int* row, const int* col, const float* nv, int* s);

/* Pannotia Jacobi Bellman-Ford over in-edges, at most n-1 rounds; output x.
 * Returns rounds executed. */
int orc_sssp_pull_ref(int n, const int* row_t, const int* col_t, const int* w_t, int src, int* x);
/* Dijkstra along out-edges; unreachable = inf. */
int orc_dijkstra_tb(int n, const int* row, const int* col, const int* w, int src, long long inf,
                    long long* dist);
/* Bellman-Ford along out-edges (second textbook form). */
int orc_bellman_ford_tb(int n, const int* row, const int* col, const int* w, int src, long long inf,
                        long long* dist);

/* Lonestar worklist BFS (levels; bfs_port_kernel.cl) and SSSP
 * (sssp_port_kernel.cl, sssp_common.cl), unreachable = 1e9.  Return the
 * number of worklist iterations. */
int orc_bfs_worklist_ref(int n, const int* row, const int* col, int src, uint32_t* dist);
int orc_sssp_worklist_ref(int n, const int* row, const int* col, const int* w, int src,
                          uint32_t* dist);
/* The same worklist iterations on `threads` persistent pthreads with a CPU
 * XF barrier (the CPU baseline): w == NULL gives BFS levels, else SSSP
 * distances; *iterations = worklist iterations with a non-empty input. */
int orc_worklist_mt(int n, const int* row, const int* col, const int* w, int src, int threads, uint32_t* dist,
                    int* iterations);
/* Queue BFS. */
int orc_bfs_tb(int n, const int* row, const int* col, int src, uint32_t* dist);
/* dverifysolution: edges with dist[u] + w < dist[v] (w = 1 if w == NULL). */
long long orc_verify_distances_ref(int n, const int* row, const int* col, const int* w,
                                   const uint32_t* dist);

/* Brandes over the given sources (FP64), bc_kernel.cl semantics: ordered
 * pairs, no 1/2 factor, bc[s] not credited for its own source. */
int orc_bc_brandes_tb(int n, const int* row, const int* col, const int* sources, int nsrc,
                      double* bc);
/* orc_bc_brandes_tb with the sources dealt to `threads` pthreads (each with
 * its own FP64 accumulator, summed in thread order). */
int orc_bc_brandes_mt(int n, const int* row, const int* col, const int* sources, int nsrc, int threads,
                      double* bc);
/* Pair-dependency definition: sum over s in sources, t != s, v != s,t of
 * sigma_sv * sigma_vt / sigma_st when d(s,v) + d(v,t) = d(s,t).  O(n^2 m). */
int orc_bc_pairs_tb(int n, const int* row, const int* col, const int* sources, int nsrc,
                    double* bc);
/* Level-synchronous FP32 restatement of mega_bc_kernel (bc_kernel.cl:295-513)
 * with sequential accumulation order. */
int orc_bc_ref_f32(int n, const int* row, const int* col, const int* row_t, const int* col_t,
                   const int* sources, int nsrc, float* bc);

/* Work-group exclusive scan as in block_scan_cl.h:7-57 (Blelloch up-sweep /
 * down-sweep over a power-of-two array); returns the total. */
int orc_block_scan_ref(int n_items, const int* in, int* out);

/* ---- CPU model of the discovery protocol and barriers ------------------
 * `groups` threads model launched workgroups; at most `resident` of them run
 * at once (a non-preemptive scheduler: a group leaves its slot only when it
 * finishes).  Each runs the protocol (lock: 0 lock-free, 1 ticket, 2 spin;
 * skip: every group participates with its launch id) and, if participating,
 * `barriers` barriers (bar: 0 XF, 1 counter, 2 tree).  Per-round counts are
 * checked after every barrier.  Results: P, ids[groups] (-1 for
 * non-participants), violations.  Returns 0, or 1 if the watchdog saw no
 * progress for `watchdog_ms` (deadlock). */
int orc_dp_model(int groups, int resident, int lock, int bar, int skip, int barriers,
                 int watchdog_ms, int* participating, int* ids, long long* violations);
/* The model uses the device's memory orders (relaxed polls, acq_rel
 * read-modify-writes, release publication), with each device fence carried
 * by the atomic it pairs with, and a plain payload written before every
 * barrier and read after it, so a ThreadSanitizer build reports any missing
 * synchronisation.  weak != 0 plants a mistake for the tests: the barrier's
 * arrival and release operations become relaxed. */
void orc_dp_model_set_weak(int weak);

/* ---- schedule selector MLP (FP64) ----------------------------------------
 * No reference counterpart.  16-64-64-4 ReLU network, parameters in the
 * library's input-major layout: W1[16][64], b1[64], W2[64][64], b2[64],
 * W3[64][4], b3[4]; layer output j = b[j] + sum_i W[i][j] x[i]. */
void orc_mlp_forward(const double* w, const double* x, double* y);
/* Loss L = mean over trace records (20 floats each: features[16], arm,
 * cost_us, ...) with cost > 0 of (y[arm] - log2(cost))^2, and dL/dw
 * (5508 doubles) by backpropagation.  Returns the number of records used. */
int orc_mlp_grad(const double* w, const float* records, int count, double* grad, double* loss);

/* ---- minimum spanning forest ---------------------------------------------
 * Kruskal (edges sorted by weight, union-find) over the undirected edges of
 * a symmetric CSR; the forest weight is unique.  Restates what lonestar mst
 * computes (apps/mst/kernel/mst_common.cl; mst_port.cpp:300-313 prints
 * weight, components, edges).  Returns 0. */
int orc_mst_kruskal(int n, const int* row, const int* col, const unsigned* w, unsigned long long* weight,
                    int* edges, int* components);
/* Prim from every unvisited vertex (binary heap): the second textbook form. */
int orc_mst_prim_tb(int n, const int* row, const int* col, const unsigned* w, unsigned long long* weight,
                    int* edges, int* components);

#ifdef __cplusplus
}
#endif
#endif
