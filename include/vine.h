/*
 * vine.h -- C ABI of the MI355X-native exact constrained-shortest-path (CSP)
 * solver.  Plain pointers and sizes only; no C++ or torch types cross it.
 *
 * Reference interfaces each entry point replaces (the reference exposes a C++
 * template API, not a C ABI):
 *   vine_graph_create   Graph<Edge_t>::adj_to_csr + make_gpu_graph
 *                       (include/graph.hpp:136-195) and create_edge_list
 *                       (main.cu:28-36): host CSR -> device CSR.
 *   vine_solver_create  the allocation half of RUN_CSPP
 *                       (algorithm/cspp.cuh:82-123), done once per handle
 *                       instead of once per call.
 *   vine_solve          cspp<c_label,c_edge>::solve() -> RUN_CSPP
 *                       (algorithm/cspp.hpp:36-53, algorithm/cspp.cuh:69-325)
 *                       for the c_label / c_constraint / c_edge instantiation
 *                       (include/label.hpp:170-247, include/constraint.hpp:81-105,
 *                       include/edge.hpp:16-26): one query per (source, sink,
 *                       delay bound), answers written to caller buffers.
 *   vine_solve_device   the same with query and answer arrays already in HBM.
 *
 * Semantics.  For each query (s, t, bound) the solver returns the
 * lexicographically smallest (len, delay) over s->t walks whose total delay
 * is <= bound (c_constraint::pass_constraints_check, include/constraint.hpp:99-101),
 * i.e. the min-d label of t's Pareto set under the non-strict dominance of
 * c_label::dominate (include/label.hpp:234-236), together with one path that
 * realises it.  cost(s, s) = 0 (the source sits in its own Pareto set, as
 * ck_init does, algorithm/cspp.cuh:63-67).  Weights are integer-valued float32;
 * results are exact while every path sum stays below 2^24.
 *
 * Errors: every entry point returns a vine_status; vine_last_error() gives a
 * thread-local message.  Threading: separate solver handles are independent;
 * one handle must not be used by two threads at once.
 */
#ifndef VINE_H
#define VINE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  VINE_OK = 0,
  VINE_E_ARG = -1,      /* invalid argument (bad id, size, pointer, option) */
  VINE_E_OOM = -2,      /* device or host allocation failed */
  VINE_E_HIP = -3,      /* HIP runtime error */
  VINE_E_OVERFLOW = -4, /* a device capacity (labels, candidates, pool) was exceeded */
  VINE_E_RANGE = -5     /* a path sum reached 2^24, exactness no longer guaranteed */
} vine_status;

/* Pruning modes.  All return the same (cost, delay, feasible). */
enum {
  VINE_PRUNE_NONE = 0,   /* reference one-to-all search: full Pareto sets at every vertex */
  VINE_PRUNE_TARGET = 1, /* drop labels lexicographically no better than t's best */
  VINE_PRUNE_BOUNDS = 2  /* lower bounds on remaining len/delay to t (reverse SSSP on the GPU); exact for
                            integer-valued weights below 2^24, so a graph with other weights runs
                            VINE_PRUNE_TARGET instead (same answers) */
};

/* Bucket-width (frontier window) policies.  Scheduling only; never change answers. */
enum {
  VINE_DELTA_INF = 0,   /* expand the whole frontier every round */
  VINE_DELTA_FIXED = 1, /* expand labels with key <= min key + delta_fixed */
  VINE_DELTA_AR = 2,    /* the reference's DOP rule driven by an AR(3) model of log2 frontier size */
  VINE_DELTA_MLP = 3,   /* per-query window predicted by the on-device MLP (vine_solver_set_mlp) */
  VINE_DELTA_DOP = 4    /* the reference's DOP rule with its own count semantics (algorithm/cspp.cuh:179-216):
                           each round expands the last DOP entries of the batch's frontier (all of them when
                           DOP >= frontier) and carries the rest; no key window */
};

/* How the label search is scheduled. */
enum {
  VINE_ENGINE_AUTO = 0,      /* the workgroup engine where it applies, else grid-wide rounds */
  VINE_ENGINE_ROUNDS = 1,    /* grid-wide rounds: every query advances one round per kernel sequence */
  VINE_ENGINE_WORKGROUP = 2  /* one persistent workgroup per query runs its whole search, rounds
                                separated by workgroup barriers (needs VINE_PRUNE_BOUNDS, a fixed or
                                learned window and out-degrees <= 512) */
};

typedef struct vine_graph vine_graph;
typedef struct vine_solver vine_solver;

typedef struct {
  int32_t max_batch;        /* query slots solved together; 1 = plain single-query solver */
  int32_t prune;            /* VINE_PRUNE_* */
  int32_t delta_policy;     /* VINE_DELTA_* */
  float delta_fixed;        /* window for VINE_DELTA_FIXED */
  int32_t use_graph;        /* 1: the rounds_per_sync rounds between two host checks are captured once
                               into a HIP graph and replayed (one graph launch per check) with the
                               full grid_blocks grid; with time_kernels the first round of each check
                               is launched directly and timed, the graph holds the rest.
                               VINE_DELTA_AR launches directly (it sets the window from the host after
                               every round).  2: the same, but only once the frontier holds at most 32,768
                               entries (the batch's tail).  0: direct launches */
  int32_t rounds_per_sync;  /* rounds between host checks of the frontier size */
  int32_t block_size;       /* threads per workgroup of the round kernels (64, 128, 192 or 256; anything
                               else is refused with VINE_E_ARG) */
  int32_t grid_blocks;      /* workgroups per round kernel at most (0 = 384) */
  int32_t hash_slots;       /* first-level candidate table in LDS, slots per workgroup of expand (0 = none;
                               else a power of two in [16, 1024]): candidates of one workgroup for the same
                               (query, vertex) meet in a slot of 4 and those another one dominates never
                               reach HBM; no slot or a full slot sends a candidate to HBM directly.  With
                               VINE_DELTA_MLP the network's expansion forecast sizes the active part each
                               round (about two slots per predicted candidate of a workgroup) */
  int64_t label_capacity;   /* label log entries (0 = auto from free memory) */
  int64_t cand_capacity;    /* candidate slots per round (0 = auto); half as many list chunks are allocated */
  int64_t pool_capacity;    /* Pareto-set pool entries (0 = auto) */
  int64_t frontier_capacity;/* carried-frontier entries (0 = auto) */
  double ar_max_labels;     /* MAX for the DOP rule (0 = free device memory / 12 B, as the reference) */
  int32_t record_trace;     /* keep per-round frontier statistics (for vine_solver_trace) */
  int32_t time_kernels;     /* HIP-event timing (vine_stats.ms_*): 0 off, k >= 1 brackets the kernels of
                               every k-th round (and every bound search); timed_rounds counts them */
  float bound_delta;        /* distance window of the bound search (0 = 2 x mean arc weight) */
  int32_t bound_ball;       /* 0: a forward min-delay search from s first, then the reverse bound
                               searches only inside fwd + lbd <= bound; 1: reverse searches over
                               the whole ball lbd <= bound (no forward search) */
  int32_t engine;           /* VINE_ENGINE_* */
  int32_t labels_per_wave;  /* frontier entries a wave takes per step in expand (power of two <= 64;
                               0 = 64); fewer spread a heavy round's arcs over more waves */
  int32_t concurrency;      /* 0 or 1: one stream.  k in 2..16: k internal solvers on k HIP streams,
                               each with ceil(max_batch / k) slots and 1/k of every explicit
                               capacity (automatic ones are sized for the whole batch); a solve
                               gives query i to part i % k and runs the parts from k host threads,
                               so one part's long tail overlaps the others' rounds.  With k > 1,
                               vine_solver_pareto's q is the query index (nq <= max_batch) */
  float near_windows;       /* the near tier (carried labels scanned again every round) holds keys up
                               to this many windows past the query's front; later keys wait in the far
                               tier's key buckets until the front comes within that reach (0 = 1) */
  int32_t lagrange_bound;   /* with VINE_PRUNE_BOUNDS, a Lagrangian bound on the remaining length: a reverse
                               search per query under the integer weights 256 len + k delay gives lbw, and a
                               label (d, delay) at v gets the bound d + (lbw(v) + k (delay - bound)) / 256,
                               which keys and prunes it together with the plain length bound (same answers).
                               k / 256 is a Lagrange multiplier found from a minimum-delay and a shortest path:
                               0: no Lagrangian bound;
                               1: one search, at the slope between those two paths;
                               2 (vine_options_default): one search under four multipliers at once (0.25,
                                  0.4, 0.55 and 0.75 x that slope), then one hull step; the best of the five;
                               3..8: that many searches of the sequential hull iteration, the best kept */
} vine_options;

typedef struct {
  int64_t rounds;           /* CSP rounds, summed over batches */
  int64_t sssp_rounds;      /* bound-search rounds of the longest query, summed over batches (VINE_PRUNE_BOUNDS) */
  int64_t expanded;         /* labels expanded (X) */
  int64_t relaxed;          /* arcs relaxed (E) */
  int64_t candidates;       /* candidates that survived the read-only dominance check */
  int64_t inserted;         /* labels inserted into Pareto sets (I) */
  int64_t sssp_relaxed;     /* arcs relaxed by the bound computation */
  int64_t max_frontier;     /* largest frontier seen */
  int64_t batches;
  double ms_total;          /* wall time of the solve (host clock) */
  double ms_kernels;        /* summed HIP-event time of the round kernels */
  double ms_expand;         /* HIP-event time of the expand kernels */
  double ms_merge;          /* HIP-event time of the merge kernels */
  double ms_sssp;           /* HIP-event time of the bound kernels */
  int64_t expand_launches;
  int64_t merge_launches;
  int64_t sssp_launches;
  int64_t touched;          /* (query, vertex) Pareto sets merged into */
  int64_t merge_scanned;    /* Pareto entries compared by the merge kernel */
  int64_t merge_chain_max;  /* longest candidate list folded by one lane */
  int64_t far_pulled;       /* frontier entries moved from far tiers to the near tier */
  double ms_pull;           /* HIP-event time of the far-tier pull kernels */
  int64_t near_carried;     /* frontier entries re-queued in the near tier (summed over rounds) */
  int64_t far_appended;     /* frontier entries parked in far tiers */
  int64_t scanned;          /* frontier entries read by the expand kernels */
  int64_t timed_rounds;     /* rounds whose kernels the ms_* figures cover (time_kernels) */
  int64_t far_units;        /* far-tier units (16 entries each) allocated, summed over batches */
  int64_t cand_chunks;      /* candidate-list chunks folded by the merge kernel (records / chunks = fill) */
  int64_t hash_dropped;     /* candidates the LDS table dropped (dominated by a candidate of the workgroup) */
  int64_t hash_overflow;    /* candidates that found no slot or a full one (appended to HBM directly) */
} vine_stats;

void vine_options_default(vine_options* opt);
const char* vine_last_error(void);
/* Number of devices visible; negative vine_status on runtime error. */
int32_t vine_device_count(void);

vine_status vine_graph_create(int32_t n, int64_t m, const int32_t* row_ptr, const int32_t* dst,
                              const float* len, const float* delay, vine_graph** out);
vine_status vine_graph_destroy(vine_graph* g);

vine_status vine_solver_create(const vine_graph* g, const vine_options* opt, vine_solver** out);
vine_status vine_solver_destroy(vine_solver* s);

/* Host buffers.  out_path_off has nq+1 entries; path i is
 * out_path[out_path_off[i] .. out_path_off[i+1]) (vertex ids s..t; empty if
 * infeasible).  Paths are skipped when out_path is NULL; if path_cap is too
 * small VINE_E_OVERFLOW is returned after costs are written. */
vine_status vine_solve(vine_solver* s, int32_t nq, const int32_t* src, const int32_t* dst,
                       const float* bound, float* out_cost, float* out_delay,
                       int32_t* out_feasible, int64_t* out_path_off, int32_t* out_path,
                       int64_t path_cap, vine_stats* out_stats);

/* Paths of the last solve again, without solving: for a caller whose path
 * buffer was too small (vine_solve returned VINE_E_OVERFLOW with the needed
 * size in out_path_off[nq]).  The last solve must have had exactly these nq
 * queries in one batch (nq <= max_batch, so its label log is still resident);
 * otherwise VINE_E_ARG.  Same layout and overflow rule as vine_solve.  No
 * counterpart in the reference, which never extracts a path. */
vine_status vine_solver_paths(vine_solver* s, int32_t nq, int64_t* out_path_off, int32_t* out_path,
                              int64_t path_cap);

/* Device buffers (HBM), no paths.  Costs/delays are +inf for infeasible queries. */
vine_status vine_solve_device(vine_solver* s, int32_t nq, const int32_t* d_src, const int32_t* d_dst,
                              const float* d_bound, float* d_cost, float* d_delay,
                              int32_t* d_feasible, vine_stats* out_stats);

/* No-allocation variant of vine_solver_create + vine_solve_device, for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  It calls no hipMalloc, hipFree or other allocating runtime
 * routine: every buffer of the solve (per-query and per-(query, vertex) state,
 * label log, Pareto pool, candidate lists, frontier tiers, counters, the MLP
 * parameters) is cut from a static device array of VINE_STATIC_SCRATCH_BYTES
 * that each GPU holds once the library is loaded on it, and the automatic
 * capacities are sized to fill it.  Options as vine_solver_create, except
 * that concurrency > 1, use_graph, time_kernels and record_trace are refused
 * (VINE_E_ARG), as is a configuration whose buffers do not fit the array (the
 * message names the bytes needed).  mlp_params (host, mlp_nparams floats, the
 * vine_solver_set_mlp layout) only for VINE_DELTA_MLP.  The graph must live on
 * the current device.  stream: a hipStream_t, or NULL for the default stream;
 * the call returns when the answers are in d_cost / d_delay / d_feasible.
 * Every call on one GPU shares that GPU's array (each call clears what it needs
 * cleared), so calls that overlap on one GPU race: serialise them. */
#define VINE_STATIC_SCRATCH_BYTES (64ull << 20)
vine_status vine_solve_device_static(const vine_graph* g, const vine_options* opt, const float* mlp_params,
                                     int64_t mlp_nparams, int32_t nq, const int32_t* d_src, const int32_t* d_dst,
                                     const float* d_bound, float* d_cost, float* d_delay, int32_t* d_feasible,
                                     void* stream, vine_stats* out_stats);

/* Per-query finish time in microseconds from the start of its batch (last
 * solve), measured with the device wall clock. */
vine_status vine_solver_latency(const vine_solver* s, int32_t nq, double* out_us);

/* Per-(query, round) scheduler features of the last solve (record_trace = k
 * records every k-th round of every live query, up to 8M rows): row i is
 * rows[i*10 .. i*10+10) = {query, round, x0..x7} with the MLP inputs x0..x7
 * (see vine_amd/csrc/solver_kernels.hpp, policy_query).  With rows = NULL returns
 * the row count; else copies up to max_rows and returns the number copied.
 * The query column is the index into the solve's query arrays. */
int64_t vine_solver_trace(const vine_solver* s, float* rows, int64_t max_rows);

/* MLP Delta policy weights (8 -> 32 -> 32 -> 2, fp32, row-major W[out][in]):
 * w1[32*8], b1[32], w2[32*32], b2[32], w3[2*32], b3[2], plus input mean/std
 * normalisation xmu[8], xsd[8].  Copied to the device. */
vine_status vine_solver_set_mlp(vine_solver* s, const float* params, int64_t nparams);

/* Device-resident Pareto set of vertex v for batch slot q after the last
 * batch of the last solve (VINE_PRUNE_NONE gives the reference's full sets).
 * Writes up to cap (d, delay) pairs, returns the set size or a negative status. */
int64_t vine_solver_pareto(const vine_solver* s, int32_t q, int32_t v, float* out_pairs, int64_t cap);

/* Bucket-width MLP (8 -> 32 -> 32 -> 2, fp32, ReLU), trained and evaluated on
 * the GPU.  No counterpart in the reference: it replaces the AR(3) model of
 * include/AR.cpp as the window scheduler.  Parameter layout: W1[32*8] b1[32]
 * W2[32*32] b2[32] W3[2*32] b3[2] xmu[8] xsd[8] (row-major W[out][in]); the
 * last 16 entries normalise the inputs and are not trained.  Loss per sample:
 * 0.5 * sum_k mask_k (y_k - target_k)^2.  Arrays are host buffers: x[n*8],
 * target[n*2], mask[n*2], y[n*2], grad[1410]. */
typedef struct vine_mlp vine_mlp;
int32_t vine_mlp_param_count(void);
const char* vine_mlp_last_error(void);
vine_status vine_mlp_create(vine_mlp** out);
vine_status vine_mlp_destroy(vine_mlp* m);
vine_status vine_mlp_set_params(vine_mlp* m, const float* params);
vine_status vine_mlp_get_params(const vine_mlp* m, float* params);
vine_status vine_mlp_forward(const vine_mlp* m, const float* x, int64_t n, float* y);
vine_status vine_mlp_gradient(const vine_mlp* m, const float* x, const float* target, const float* mask,
                              int64_t n, float* grad, double* loss);
vine_status vine_mlp_train(vine_mlp* m, const float* x, const float* target, const float* mask, int64_t n,
                           int32_t epochs, int32_t batch, float lr, uint64_t seed, double* loss_curve);

/* Per-query work of the last solve: labels expanded and rounds in which the
 * query had live frontier labels. */
vine_status vine_solver_query_stats(const vine_solver* s, int32_t nq, int64_t* expanded, int32_t* rounds);

/* Host scheduling helpers (also used by the solver's VINE_DELTA_AR policy).
 * vine_ar_burg: Burg AR fit, the reference's AR::AutoRegression with the
 *   default MAXENTROPY method (include/AR.cpp:6-147); 0 on success.
 * vine_dop_sequence: the reference's DOP rule (algorithm/cspp.cuh:254-284 with
 *   calculate_p :33-41), fed one frontier size per round; writes the DOP after
 *   each round and returns n. */
int vine_ar_burg(const double* x, int n, int degree, double* coef);
int64_t vine_dop_sequence(const int64_t* sizes, int n, double max_labels, int64_t* out_dop);

/* The internal vertex order vine_graph_create uses (no counterpart in the
 * reference, whose CSR keeps the input ids, include/graph.hpp:136-167):
 * out_order[new id] = caller's id, per connected component along a Hilbert
 * curve over hop distances from two far-apart vertices, so that neighbours
 * share cache lines of the per-(query, vertex) arrays.  Host only. */
int32_t vine_locality_order(int32_t n, const int32_t* row_ptr, const int32_t* dst, int32_t* out_order);

/* The landmark delays vine_graph_create keeps for the forward bound search
 * (no counterpart in the reference): out[v * 4 + j] = min delay from landmark
 * j to v (-1: unreachable), four landmarks chosen farthest-first (the vertex
 * farthest from vertex 0, then each time the vertex farthest from its nearest
 * landmark).  The search bounds the delay from v to t below by
 * d(L, t) - d(L, v), and on graphs with twin arcs also by d(L, v) - d(L, t).
 * Host only. */
int32_t vine_landmark_delays(int32_t n, const int32_t* row_ptr, const int32_t* dst, const float* delay, float* out);

/* ---- Other label families ------------------------------------------------
 * The reference compiles one label family at a time (CSP in include/label.hpp:
 * 46-169, 248-461 and include/constraint.hpp:21-80, 106-186); here each is a
 * family id of a generic label-correcting GPU search (functors in
 * include/vine/families.hpp).  Per query: limits[3*i .. 3*i+3) =
 * (delay_L, hop_L, maxcost_L) (hop_L and maxcost_L are read by the hop
 * families only; the reference's defaults are 100 and 10000), and the answer
 * out_label[4*i .. 4*i+4) = (d, delay, hop, maxcost) of the best label in t's
 * Pareto set (+inf when t's set is empty; unused fields 0):
 *   VINE_FAMILY_CSP, VINE_FAMILY_2OSP (CSP == 1, 4): min (d, delay); delay <= delay_L
 *   VINE_FAMILY_HOP (CSP == 2): + hop count, hop <= hop_L; min (d, delay, hop)
 *   VINE_FAMILY_HOP_MAXCOST (CSP == 3): + max arc len on the path <= maxcost_L;
 *     min (d, delay, hop, maxcost)
 *   VINE_FAMILY_SOCIAL (CSP == 5): products from (1, 1), larger dominates,
 *     delay <= delay_L; max (d, delay); arc weights must lie in [0, 1]
 *   VINE_FAMILY_QOS (default branch): d additive, delay = bottleneck (min arc
 *     delay, 21474836 at the source), bottleneck >= delay_L; min d, then max
 *     bottleneck.
 * No counterpart in the reference for the answer order: it never extracts an
 * answer for these families. */
typedef enum {
  VINE_FAMILY_CSP = 1,
  VINE_FAMILY_HOP = 2,
  VINE_FAMILY_HOP_MAXCOST = 3,
  VINE_FAMILY_2OSP = 4,
  VINE_FAMILY_SOCIAL = 5,
  VINE_FAMILY_QOS = 6
} vine_family;

typedef struct vine_family_solver vine_family_solver;

/* Device memory for max_batch queries per batch and label_capacity labels
 * (0 = 2^25); about 37 B per label plus 8 B per (query slot, vertex).
 * prune = 0: the reference's one-to-all search (every vertex's Pareto set is
 * complete, vine_family_front returns it); prune = 1: labels that cannot beat
 * the best answer found at t so far, or (integer weights) cannot complete
 * within the limits by their shortest delay / arc count to t, are dropped
 * (the answer is the same, Pareto sets other than the answer's may be
 * partial).  window > 0 (families other than social): each round expands only
 * the labels whose length plus shortest length to t is within `window` of
 * the query's smallest, as the batched solver's fixed window does; 0 expands
 * every frontier label each round. */
vine_status vine_family_solver_create(const vine_graph* g, int32_t family, int32_t max_batch,
                                      int64_t label_capacity, int32_t prune, float window,
                                      vine_family_solver** out);
vine_status vine_family_solver_destroy(vine_family_solver* s);

/* Host buffers; queries run in batches of max_batch.  Paths as in vine_solve
 * (skipped when out_path is NULL; out_path_off may then still receive the
 * offsets).  out_counters (optional, 4 entries): rounds, labels expanded,
 * arcs relaxed, labels inserted, summed over batches. */
vine_status vine_family_solve(vine_family_solver* s, int32_t nq, const int32_t* src, const int32_t* dst,
                              const float* limits, float* out_label, int32_t* out_feasible,
                              int64_t* out_path_off, int32_t* out_path, int64_t path_cap,
                              int64_t* out_counters);

/* Pareto set of vertex v for query q of the last batch: up to cap labels as
 * (d, delay, hop, maxcost); returns the set size or a negative status. */
int64_t vine_family_front(const vine_family_solver* s, int32_t q, int32_t v, float* out_labels, int64_t cap);

#ifdef __cplusplus
}
#endif

#endif /* VINE_H */
