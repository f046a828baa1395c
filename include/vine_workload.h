/*
 * vine_workload.h -- seeded synthetic graphs and query sets for the exact
 * constrained-shortest-path (CSP) solver.
 *
 * The reference ships no usable graphs (its road graphs ny/nw/ne.graph are
 * absent), so every workload here is generated from a seed.  Generation is
 * deterministic: the same parameters and seed give bit-identical arrays on
 * every machine and for any thread count.
 *
 * Graph layout produced here is the adjacency-insertion-order CSR that the
 * reference's Graph::adj_to_csr builds (include/graph.hpp:136-167): for every
 * undirected edge {u,v} in generation order the arc u->v is appended to u's
 * list and v->u to v's list, both arcs carrying the same (len, delay), as the
 * reference's undirected loader does (include/io.hpp:50-60).  Weights are
 * integer-valued and stored as float32, so all path sums below 2^24 are exact.
 *
 * Query bounds are integers floor(mindelay * num / den) computed in integer
 * arithmetic, so a bound has the same meaning on every device.
 */
#ifndef VINE_WORKLOAD_H
#define VINE_WORKLOAD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vine_host_graph vine_host_graph;

/* 0 on success, negative on bad arguments / allocation failure. */

/* rows x cols 4-neighbour grid, vertex id = r*cols + c.  Undirected edges are
 * enumerated row-major (right neighbour, then down neighbour); exactly
 * floor(delete_permille * E / 1000) of them are removed by a seeded
 * Fisher-Yates shuffle.  len ~ U{len_lo..len_hi}, delay ~ U{dly_lo..dly_hi}. */
int vine_gen_grid(int32_t rows, int32_t cols, int32_t delete_permille,
                  int32_t len_lo, int32_t len_hi, int32_t dly_lo, int32_t dly_hi,
                  uint64_t seed, vine_host_graph** out);

/* n points uniform in [0,1)^2, each linked to its k nearest neighbours
 * (self excluded, ties by index), union-symmetrised.
 * len = max(1, round(dist * len_scale)), delay = max(1, round(len * U[0.5,2.0])). */
int vine_gen_knn(int64_t n, int32_t k, double len_scale, uint64_t seed,
                 vine_host_graph** out);

/* R-MAT graph, 2^scale vertices, edge_factor * 2^scale generated edges with
 * quadrant probabilities (a, b, c, 1-a-b-c), symmetrised, self loops and
 * duplicates removed.  len, delay ~ U{lo..hi} independent. */
int vine_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                  int32_t lo, int32_t hi, uint64_t seed, vine_host_graph** out);

/* Build from caller CSR arrays (copied). */
int vine_host_graph_from_csr(int32_t n, int64_t m, const int32_t* row_ptr,
                             const int32_t* dst, const float* len, const float* delay,
                             vine_host_graph** out);

void vine_host_graph_free(vine_host_graph* g);
int32_t vine_host_graph_n(const vine_host_graph* g);
int64_t vine_host_graph_m(const vine_host_graph* g);
/* Borrowed pointers, valid until vine_host_graph_free. */
const int32_t* vine_host_graph_row_ptr(const vine_host_graph* g);
const int32_t* vine_host_graph_dst(const vine_host_graph* g);
const float* vine_host_graph_len(const vine_host_graph* g);
const float* vine_host_graph_delay(const vine_host_graph* g);

/* Minimum total delay from s to t (Dijkstra on delay), -1 if unreachable. */
int64_t vine_min_delay(const vine_host_graph* g, int32_t s, int32_t t);

/* nq random pairs (s != t); if lcc_only, both ends are drawn from the largest
 * weakly connected component.  bound = floor(mindelay*num/den), or 0 when t
 * is unreachable (out_mindelay = -1).  Uses up to nthreads threads. */
int vine_gen_queries_random(const vine_host_graph* g, int32_t nq, int32_t lcc_only,
                            int32_t num, int32_t den, uint64_t seed, int32_t nthreads,
                            int32_t* out_src, int32_t* out_dst, float* out_bound,
                            int64_t* out_mindelay);

/* nq queries whose hop distance is log-uniform in [hop_lo, hop_hi]: draw s and
 * h, BFS from s, pick t uniformly at hop distance exactly h (redraw if none).
 * bound = floor(mindelay * f / 1000) with f ~ U{f_lo_permille..f_hi_permille}.
 * out_hops receives the achieved hop distance. */
int vine_gen_queries_hops(const vine_host_graph* g, int32_t nq, int32_t hop_lo,
                          int32_t hop_hi, int32_t f_lo_permille, int32_t f_hi_permille,
                          uint64_t seed, int32_t nthreads, int32_t* out_src,
                          int32_t* out_dst, float* out_bound, int64_t* out_mindelay,
                          int32_t* out_hops);

/* Standard workload presets.  id: 1 = 64x64 grid (10% deleted), 2 = 264,346
 * vertex 3-NN geometric graph, 3 = 1,524,453 vertex 3-NN, 4 = R-MAT scale 17. */
int vine_gen_preset_graph(int32_t id, uint64_t seed, vine_host_graph** out);

/* ---- Graph and query files ------------------------------------------------
 * The input formats of the reference's read_graph(file, g, graph_type)
 * (include/io.hpp:22-91) and task::read_query(file, graph_type) /
 * task::load_query(file) (include/task.hpp:27-76); the format numbers are the
 * reference's graph_type values.  Vertex ids are 0-based, weights and bounds
 * integers in [0, 2^24).  Errors return a negative vine_status (VINE_E_ARG for
 * an unreadable file or bad content) with file:line in vine_last_error(). */
#define VINE_FILE_SNAP 1 /* SNAP edge list, undirected */
#define VINE_FILE_COLA 2 /* COLA road network, directed arcs */

/* Graph.  VINE_FILE_COLA: the vertex count, then lines "u: v w c v w c ... -1",
 *   each triple an arc u->v with len w and delay c, kept directed as listed.
 * VINE_FILE_SNAP: each row is an undirected edge, stored as both arcs with the
 *   same (len, delay).  Either the reference's layout ("N?M?": two counts each
 *   followed by one separator character, then M rows "u v w c") or the
 *   MatrixMarket one of the shipped inputs/tech-*.mtx.txt ('%' comment lines,
 *   "N N M", then M rows "id u v w c"), told apart by a leading '%'.
 * The CSR keeps the file's arc order per vertex (adjacency insertion order). */
int vine_read_graph(const char* path, int32_t format, vine_host_graph** out);

/* Queries.  VINE_FILE_COLA: rows "u v lower upper limit", bound = limit,
 *   aux = (lower, upper).  VINE_FILE_SNAP: a header line, then rows
 *   "id,src,snk,delay,hop,cost", bound = delay, aux = (hop, cost).
 * Returns the number of queries in the file and writes the first min(count,
 * cap) of them (out_aux: 2 per query, may be NULL; all outputs may be NULL when
 * cap = 0).  Vertex ids are checked against a graph only when solved. */
int64_t vine_read_queries(const char* path, int32_t format, int64_t cap, int32_t* out_src,
                          int32_t* out_dst, float* out_bound, int32_t* out_aux);

#ifdef __cplusplus
}
#endif

#endif /* VINE_WORKLOAD_H */
