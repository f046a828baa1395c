// Host orchestration and C ABI of the exact CSP solver (libvine.so).
//
// Counterpart of the reference's RUN_CSPP host driver (algorithm/cspp.cuh:69-325)
// and solver facade (algorithm/cspp.hpp:11-69), re-designed for one MI355X:
//  - device memory is owned by the solver handle and sized once, not
//    cudaMallocManaged per call (algorithm/cspp.cuh:82-114);
//  - per-batch reset touches only per-(query, vertex) headers, not O(N*128)
//    label arrays (algorithm/cspp.cuh:142-154);
//  - many queries share every round launch (max_batch query slots), with
//    per-query bounds (the reference uses C_l[0] for all, cspp.cuh:196,234);
//  - the frontier size stays on the device; the host polls it every
//    rounds_per_sync rounds (the reference syncs 5-6 times per round).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <mutex>
#include <thread>
#include <vector>

#include "mlp.hpp"
#include "device_graph.hpp"
#include "errors.hpp"
#include "layout.hpp"
#include "sched.hpp"
#include "solver_kernels.hpp"
#include "solver_wg.hpp"
#include "vine.h"

using namespace vine;
using vine::dev::Counters;
using vine::dev::Ctx;

namespace {

using vine::set_error;

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) {                                                        \
      set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
      return VINE_E_HIP;                                                           \
    }                                                                              \
  } while (0)

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct vine_solver {
  const vine_graph* g = nullptr;
  vine_options opt{};
  int B = 1;
  int grid = 1024;
  int block = 256;
  hipStream_t stream = nullptr;
  std::vector<void*> allocs;
  // no-allocation mode (vine_solve_device_static): every device buffer is cut
  // from this arena instead of hipMalloc'd; count_only sizes a layout
  char* arena = nullptr;
  size_t arena_cap = 0, arena_used = 0;
  bool count_only = false;
  Counters h_ct_local{};  // host copy of the counters when no pinned buffer is allocated

  // per query slot
  int32_t *d_qsrc = nullptr, *d_qdst = nullptr;
  float *d_qbound = nullptr, *d_qdelta = nullptr;
  float2* d_qbest = nullptr;
  int32_t* d_qbest_id = nullptr;
  uint32_t* d_qfmin = nullptr;
  int32_t* d_qlast = nullptr;
  int32_t *d_qexp = nullptr, *d_qins = nullptr, *d_qnear = nullptr;
  long long* d_qexptot = nullptr;
  float* d_qfcast = nullptr;
  float* d_hash_pred = nullptr;
  float* d_trace = nullptr;
  long long trace_cap = 0;
  // per (query, vertex)
  int2* d_pset = nullptr;
  unsigned long long* d_cand_head = nullptr;
  float* d_lbl = nullptr;  // lbl per (query, vertex)
  float* d_lbw = nullptr;  // lbw per (query, vertex)
  float* d_qlk = nullptr;
  unsigned long long* d_lbd = nullptr;
  float* d_fwd = nullptr;
  uint32_t* d_ell = nullptr;
  int32_t ell_words = 0;
  uint32_t* d_inq = nullptr;
  int32_t* d_wq = nullptr;
  int32_t* d_qorder = nullptr;
  int4* d_wnear = nullptr;
  unsigned long long* d_qfin = nullptr;
  bool wg = false;  // workgroup engine
  int lpw = 64;     // frontier entries per wave step in expand
  // pools
  float2* d_pool_cd = nullptr;
  int32_t* d_pool_id = nullptr;
  int4* d_labs = nullptr;
  int2* d_lmeta = nullptr;
  uint8_t* d_dead = nullptr;
  int2* d_chd = nullptr;
  int4* d_crec = nullptr;
  long long* d_touched = nullptr;
  int4* d_carry[2] = {nullptr, nullptr};
  int32_t* d_far_head = nullptr;
  unsigned long long* d_far_occ = nullptr;
  int32_t* d_far_next = nullptr;
  int32_t* d_far_cnt = nullptr;
  float* d_far_base = nullptr;
  float* d_far_bw = nullptr;
  int2* d_fhd = nullptr;
  int2* d_dl = nullptr;
  int4* d_frec = nullptr;
  unsigned long long* d_ts = nullptr;
  Counters* d_ct = nullptr;
  Counters* h_ct = nullptr;  // pinned
  // outputs for the host-buffer path
  float *d_out_cost = nullptr, *d_out_delay = nullptr;
  int32_t *d_out_feas = nullptr, *d_out_plen = nullptr;
  int32_t* d_in_q = nullptr;  // staging for host query arrays (src, dst)
  float* d_in_bound = nullptr;
  int32_t in_cap = 0;
  // MLP parameters (device)
  float* d_mlp = nullptr;
  int64_t mlp_nparams = 0;

  Ctx ctx{};
  // use_graph: the rounds between two host checks captured once as a HIP
  // graph and replayed; re-captured when the batch size or chunk changes
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int graph_nb = -1, graph_k = 0, graph_gsz = 0;
  // HIP-event brackets of timed launches, resolved at the end of each batch
  struct EvRec {
    hipEvent_t a, b;
    int kind;  // 0 expand, 1 merge, 2 bound search
  };
  std::vector<hipEvent_t> evpool;
  size_t ev_used = 0;
  long long round_seq = 0;  // round launches of the current solve (for sampled timing)
  std::vector<EvRec> evrec;

  // last-solve records
  std::vector<double> latency_us;
  std::vector<float> trace;  // rows of 10 floats
  std::vector<int64_t> q_expanded;
  std::vector<int32_t> q_rounds;
  int last_batch_q = 0;
  int nb_active = 0;  // queries in the current batch
  // the last solve: its query count and whether it ran as one batch (then
  // its paths are still in the label log, for vine_solver_paths)
  int32_t last_nq = -1;
  bool last_one_batch = false;
  // path extraction buffers, grown on demand and kept (no allocation per solve)
  long long* d_path_off = nullptr;
  int32_t* d_path = nullptr;
  long long path_off_cap = 0, path_cap = 0;

  // concurrency > 1: the parts do all the work, query i going to part
  // i % parts.size() as its (i / parts.size())-th query
  std::vector<vine_solver*> parts;
  std::vector<void*> split_bufs;  // gathered device query/answer arrays, per part
  int32_t split_cap = 0;          // queries each split_bufs entry holds
};

namespace {
vine_status create_single(const vine_graph* g, const vine_options& o, double mem_frac, long long want_slots,
                          vine_solver** out);
vine_status create_split(const vine_graph* g, const vine_options& o, vine_solver** out);
}  // namespace

static vine_status dev_alloc(vine_solver* s, void** p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  if (s->arena || s->count_only) {
    // bump allocation, 256-byte aligned
    const size_t at = (s->arena_used + 255) & ~(size_t)255;
    if (!s->count_only && (at > s->arena_cap || bytes > s->arena_cap - at)) {
      set_error("the static scratch of %zu bytes is too small for this configuration (needs more than %zu)",
                s->arena_cap, at + bytes);
      return VINE_E_ARG;
    }
    *p = s->count_only ? nullptr : s->arena + at;
    s->arena_used = at + bytes;
    return VINE_OK;
  }
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    set_error("hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
    (void)hipGetLastError();
    return VINE_E_OOM;
  }
  s->allocs.push_back(*p);
  return VINE_OK;
}

#define ALLOC(ptr, bytes)                                                 \
  do {                                                                    \
    vine_status st_ = dev_alloc(s, (void**)&(ptr), (size_t)(bytes));       \
    if (st_ != VINE_OK) {                                                 \
      if (!s->arena) vine_solver_destroy(s);                              \
      return st_;                                                         \
    }                                                                     \
  } while (0)

extern "C" {

void vine_options_default(vine_options* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->max_batch = 1;
  o->prune = VINE_PRUNE_BOUNDS;
  o->delta_policy = VINE_DELTA_INF;
  o->delta_fixed = 0.f;
  o->use_graph = 0;
  o->rounds_per_sync = 1;
  o->block_size = 128;
  o->grid_blocks = 0;
  o->hash_slots = 0;
  o->ar_max_labels = 0.0;
  o->lagrange_bound = 2;
}

int32_t vine_device_count(void) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) {
    set_error("hipGetDeviceCount: %s", hipGetErrorString(e));
    return VINE_E_HIP;
  }
  return n;
}

vine_status vine_graph_create(int32_t n, int64_t m, const int32_t* row_ptr, const int32_t* dst,
                              const float* len, const float* delay, vine_graph** out) {
  if (!out || n <= 0 || m < 0 || !row_ptr || (m > 0 && (!dst || !len || !delay))) {
    set_error("vine_graph_create: invalid arguments");
    return VINE_E_ARG;
  }
  if (m >= (int64_t)INT32_MAX) {
    set_error("vine_graph_create: %lld arcs exceed the 32-bit CSR offsets", (long long)m);
    return VINE_E_ARG;
  }
  if (row_ptr[0] != 0 || row_ptr[n] != m) {
    set_error("vine_graph_create: row_ptr must start at 0 and end at m");
    return VINE_E_ARG;
  }
  for (int32_t i = 0; i < n; i++)
    if (row_ptr[i + 1] < row_ptr[i]) {
      set_error("vine_graph_create: row_ptr not monotone at %d", i);
      return VINE_E_ARG;
    }
  for (int64_t p = 0; p < m; p++) {
    if (dst[p] < 0 || dst[p] >= n) {
      set_error("vine_graph_create: arc %lld has invalid head %d", (long long)p, dst[p]);
      return VINE_E_ARG;
    }
    if (!(len[p] >= 0.f) || !(delay[p] >= 0.f)) {
      set_error("vine_graph_create: arc %lld has a negative or NaN weight", (long long)p);
      return VINE_E_ARG;
    }
  }
  // renumber for locality, keeping each vertex's arcs in their order
  const std::vector<int32_t> order = vine::locality_order(n, row_ptr, dst);
  std::vector<int32_t> perm(n);
  for (int32_t i = 0; i < n; i++) perm[order[i]] = i;
  std::vector<int32_t> prow(n + 1, 0), pdst(std::max<int64_t>(m, 1));
  std::vector<float> plen(std::max<int64_t>(m, 1)), pdelay(std::max<int64_t>(m, 1));
  for (int32_t i = 0; i < n; i++) {
    const int32_t u = order[i];
    int32_t w = prow[i];
    for (int32_t p = row_ptr[u]; p < row_ptr[u + 1]; p++, w++) {
      pdst[w] = perm[dst[p]];
      plen[w] = len[p];
      pdelay[w] = delay[p];
    }
    prow[i + 1] = w;
  }
  row_ptr = prow.data();
  dst = pdst.data();
  len = plen.data();
  delay = pdelay.data();
  // packed records {dst, len, delay, dst's arc range}; transposed copy for
  // the bound search (its records carry the ranges of the transposed CSR)
  std::vector<int4> e(std::max<int64_t>(m, 1)), re(std::max<int64_t>(m, 1));
  std::vector<int32_t> rrow(n + 1, 0);
  for (int64_t p = 0; p < m; p++) rrow[dst[p] + 1]++;
  for (int32_t v = 0; v < n; v++) rrow[v + 1] += rrow[v];
  std::vector<int32_t> fill(rrow.begin(), rrow.end() - 1);
  auto row_info = [m](const int32_t* rp, int32_t v) -> int32_t {
    const int64_t first = rp[v], deg = rp[v + 1] - rp[v];
    if (m >= (1LL << 24) || deg >= 255) return (int32_t)dev::ROW_UNKNOWN;
    return (int32_t)(((uint32_t)first << 8) | (uint32_t)deg);
  };
  for (int32_t u = 0; u < n; u++)
    for (int32_t p = row_ptr[u]; p < row_ptr[u + 1]; p++) {
      int4 rec;
      rec.x = dst[p];
      std::memcpy(&rec.y, &len[p], 4);
      std::memcpy(&rec.z, &delay[p], 4);
      rec.w = row_info(row_ptr, dst[p]);
      e[p] = rec;
      int4 rr = rec;
      rr.x = u;
      rr.w = row_info(rrow.data(), u);
      re[fill[dst[p]]++] = rr;
    }
  auto* g = new vine_graph();
  g->n = n;
  g->m = m;
  g->perm = perm;
  if (m > 0) {
    double sl = 0, sd = 0;
    for (int64_t p = 0; p < m; p++) {
      sl += len[p];
      sd += delay[p];
    }
    g->mean_len = (float)std::max(sl / m, 1e-3);
    g->mean_delay = (float)std::max(sd / m, 1e-3);
  }
  for (int32_t u = 0; u < n; u++) g->max_deg = std::max<int32_t>(g->max_deg, row_ptr[u + 1] - row_ptr[u]);
  for (int64_t p = 0; p < m; p++) {
    g->max_len = std::max(g->max_len, len[p]);
    g->max_delay = std::max(g->max_delay, delay[p]);
    if (len[p] != std::floor(len[p]) || delay[p] != std::floor(delay[p]) || len[p] >= 16777216.f ||
        delay[p] >= 16777216.f)
      g->integer_weights = false;
  }
  {
    // twins: the smallest delay of each (tail, head) pair, looked up reversed
    std::vector<std::pair<uint64_t, float>> arcs((size_t)m);
    for (int32_t u = 0; u < n; u++)
      for (int32_t p = row_ptr[u]; p < row_ptr[u + 1]; p++)
        arcs[p] = {((uint64_t)(uint32_t)u << 32) | (uint32_t)dst[p], delay[p]};
    std::sort(arcs.begin(), arcs.end());
    g->twin_delay = true;
    for (int64_t p = 0; p < m && g->twin_delay; p++) {
      const uint64_t rk = (arcs[p].first << 32) | (arcs[p].first >> 32);
      auto it = std::lower_bound(arcs.begin(), arcs.end(), std::make_pair(rk, -1.f));
      g->twin_delay = it != arcs.end() && it->first == rk && it->second <= arcs[p].second;
    }
  }
  std::vector<float> lmk;
  if (g->integer_weights && m > 0) {
    vine::landmark_distances(n, row_ptr, dst, delay, 4, lmk);
    for (float x : lmk)
      if (!(x < 16777216.f)) {
        lmk.clear();
        break;
      }
  }
  hipError_t err = hipGetDevice(&g->device);
  size_t eb = sizeof(int4) * std::max<int64_t>(m, 1);
  if (err == hipSuccess) err = hipMalloc(&g->d_row, sizeof(int32_t) * (n + 1));
  if (err == hipSuccess) err = hipMalloc(&g->d_rrow, sizeof(int32_t) * (n + 1));
  if (err == hipSuccess) err = hipMalloc(&g->d_edges, eb);
  if (err == hipSuccess) err = hipMalloc(&g->d_redges, eb);
  if (err == hipSuccess) err = hipMemcpy(g->d_row, row_ptr, sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(g->d_rrow, rrow.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(g->d_edges, e.data(), eb, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(g->d_redges, re.data(), eb, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMalloc(&g->d_perm, sizeof(int32_t) * n);
  if (err == hipSuccess) err = hipMalloc(&g->d_inv, sizeof(int32_t) * n);
  if (err == hipSuccess) err = hipMemcpy(g->d_perm, perm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(g->d_inv, order.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice);
  if (err == hipSuccess && !lmk.empty()) err = hipMalloc(&g->d_lmk, sizeof(float4) * n);
  if (err == hipSuccess && !lmk.empty())
    err = hipMemcpy(g->d_lmk, lmk.data(), sizeof(float4) * n, hipMemcpyHostToDevice);
  if (err != hipSuccess) {
    set_error("vine_graph_create: %s", hipGetErrorString(err));
    vine_graph_destroy(g);
    return err == hipErrorOutOfMemory ? VINE_E_OOM : VINE_E_HIP;
  }
  *out = g;
  return VINE_OK;
}

vine_status vine_graph_destroy(vine_graph* g) {
  if (!g) return VINE_OK;
  if (g->d_row) (void)hipFree(g->d_row);
  if (g->d_rrow) (void)hipFree(g->d_rrow);
  if (g->d_edges) (void)hipFree(g->d_edges);
  if (g->d_redges) (void)hipFree(g->d_redges);
  if (g->d_lmk) (void)hipFree(g->d_lmk);
  if (g->d_perm) (void)hipFree(g->d_perm);
  if (g->d_inv) (void)hipFree(g->d_inv);
  delete g;
  return VINE_OK;
}

vine_status vine_solver_destroy(vine_solver* s) {
  if (!s) return VINE_OK;
  for (vine_solver* p : s->parts) vine_solver_destroy(p);
  for (void* p : s->split_bufs) (void)hipFree(p);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (void* p : s->allocs) (void)hipFree(p);
  if (s->d_path_off) (void)hipFree(s->d_path_off);
  if (s->d_path) (void)hipFree(s->d_path);
  if (s->h_ct) (void)hipHostFree(s->h_ct);
  if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  for (auto& e : s->evpool) (void)hipEventDestroy(e);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return VINE_OK;
}

}  // extern "C"

namespace {
// Option checks shared by vine_solver_create and vine_solve_device_static;
// may switch the pruning mode (see below).
vine_status validate_options(const vine_graph* g, vine_options& o) {
  if (o.max_batch < 1 || o.prune < 0 || o.prune > 2 || o.delta_policy < 0 || o.delta_policy > 4 ||
      o.rounds_per_sync < 1 || o.grid_blocks < 0 || o.hash_slots < 0 || !(o.delta_fixed >= 0.f) ||
      o.bound_ball < 0 || o.bound_ball > 1) {
    set_error("vine_solver_create: invalid options");
    return VINE_E_ARG;
  }
  // the round kernels' per-wave LDS tables hold dev::MAX_WAVES waves: a larger
  // workgroup would index past them, so it is refused before any launch
  if (o.block_size < dev::WAVE || o.block_size > dev::MAX_BLOCK || o.block_size % dev::WAVE) {
    set_error("vine_solver_create: block_size %d must be a multiple of %d in [%d, %d]", o.block_size, dev::WAVE,
              dev::WAVE, dev::MAX_BLOCK);
    return VINE_E_ARG;
  }
  // The bound pruning (lower bounds from reverse searches, the upper bound
  // from a minimum-delay path) adds a label's sums in another order than the
  // label itself; that is exact only while every sum is an integer below 2^24.
  // Other weights fall back to target pruning, which compares a label's own
  // sums only and is exact for any non-negative weights.
  if (o.prune == VINE_PRUNE_BOUNDS && !g->integer_weights) o.prune = VINE_PRUNE_TARGET;
  if (o.labels_per_wave < 0 || o.labels_per_wave > 64 || (o.labels_per_wave & (o.labels_per_wave - 1))) {
    set_error("vine_solver_create: labels_per_wave must be 0 or a power of two <= 64");
    return VINE_E_ARG;
  }
  if (o.engine < 0 || o.engine > 2) {
    set_error("vine_solver_create: invalid engine");
    return VINE_E_ARG;
  }
  // the workgroup engine needs the bounds (windows keyed by d + lbl), a
  // finite window and every label's arcs fitting the LDS candidate arena
  const bool wg_ok = o.prune == VINE_PRUNE_BOUNDS &&
                     (o.delta_policy == VINE_DELTA_FIXED || o.delta_policy == VINE_DELTA_MLP) &&
                     g->max_deg <= dev::ARENA / 2;
  if (o.engine == VINE_ENGINE_WORKGROUP && !wg_ok) {
    set_error("vine_solver_create: the workgroup engine needs VINE_PRUNE_BOUNDS (integer weights), a fixed or "
            "learned window and out-degrees <= %d (graph: %d)", dev::ARENA / 2, g->max_deg);
    return VINE_E_ARG;
  }
  if (o.use_graph < 0 || o.use_graph > 2) {
    set_error("vine_solver_create: use_graph must be 0, 1 or 2");
    return VINE_E_ARG;
  }
  if (o.hash_slots && (o.hash_slots < 16 || o.hash_slots > dev::HASH_MAX_SLOTS ||
                       (o.hash_slots & (o.hash_slots - 1)))) {
    set_error("vine_solver_create: hash_slots must be 0 or a power of two in [16, %d]", dev::HASH_MAX_SLOTS);
    return VINE_E_ARG;
  }
  if (o.lagrange_bound < 0 || o.lagrange_bound > 8) {
    set_error("vine_solver_create: lagrange_bound must lie in [0, 8]");
    return VINE_E_ARG;
  }
  if (!(o.near_windows >= 0.f) || o.near_windows > 64.f) {
    set_error("vine_solver_create: near_windows must lie in [0, 64]");
    return VINE_E_ARG;
  }
  if (o.concurrency < 0 || o.concurrency > 16) {
    set_error("vine_solver_create: concurrency must lie in [0, 16]");
    return VINE_E_ARG;
  }
  return VINE_OK;
}
}  // namespace

extern "C" {

vine_status vine_solver_create(const vine_graph* g, const vine_options* opt, vine_solver** out) {
  if (!g || !out) {
    set_error("vine_solver_create: null argument");
    return VINE_E_ARG;
  }
  vine_options o;
  if (opt) o = *opt;
  else vine_options_default(&o);
  vine_status v = validate_options(g, o);
  if (v != VINE_OK) return v;
  if (o.concurrency > 1) return create_split(g, o, out);
  return create_single(g, o, 0.6, o.max_batch, out);
}

}  // extern "C"

namespace {

// Fields of a solver taken from its options (no device work).
void init_fields(vine_solver* s, const vine_graph* g, const vine_options& o) {
  s->g = g;
  s->opt = o;
  s->B = o.max_batch;
  s->block = o.block_size;
  // about 768 waves: the round kernels are latency-bound, and measured on
  // MI355X fewer, fuller waves beat a grid that fills every CU
  s->grid = o.grid_blocks ? o.grid_blocks : 384;
  const bool wg_ok = o.prune == VINE_PRUNE_BOUNDS &&
                     (o.delta_policy == VINE_DELTA_FIXED || o.delta_policy == VINE_DELTA_MLP) &&
                     g->max_deg <= dev::ARENA / 2;
  s->wg = wg_ok && (o.engine == VINE_ENGINE_WORKGROUP || o.engine == VINE_ENGINE_AUTO);
  s->lpw = o.labels_per_wave ? o.labels_per_wave : 64;
}

// Per-(query, vertex) state: pset, cand_head; with the bounds lbl, lbd, fwd,
// the bound searches' queue stamp and the HBM part of their two queues (and
// a bit of the region bitmap, counted as a byte).
size_t per_qv_bytes(const vine_options& o) {
  return sizeof(int2) + sizeof(unsigned long long) +
         (o.prune == 2 ? sizeof(float2) + sizeof(unsigned long long) + sizeof(float) + sizeof(uint32_t) +
                             2 * sizeof(int32_t) + 1
                       : 0);
}

struct Caps {
  long long lab = 0, pool = 0, cand = 0, carry = 0;
};

// The other capacities follow the label capacity unless the options set them.
vine_status derive_caps(const vine_options& o, long long B, long long lab_cap, Caps* k) {
  if (lab_cap > INT32_MAX - 1 || lab_cap < B + 1) {
    set_error("vine_solver_create: label_capacity must be in [max_batch+1, 2^31-1)");
    return VINE_E_ARG;
  }
  k->lab = lab_cap;
  k->pool = o.pool_capacity ? o.pool_capacity : std::min<long long>(2 * lab_cap + 4 * B, 0xFFFFFFFFll);
  k->cand = o.cand_capacity ? o.cand_capacity : std::min<long long>(lab_cap / 2 + 1024, INT32_MAX - 1);
  k->carry = o.frontier_capacity ? o.frontier_capacity : std::min<long long>(lab_cap / 2 + 1024, INT32_MAX - 1);
  if (k->pool < 4 * B || k->pool > 0xFFFFFFFFll || k->cand < 64 || k->cand > INT32_MAX - 1 || k->carry < 64 ||
      k->carry > INT32_MAX - 1) {
    set_error("vine_solver_create: capacities out of range");
    return VINE_E_ARG;
  }
  return VINE_OK;
}

// Every device buffer of one solver, through dev_alloc (hipMalloc, or the
// arena of the no-allocation mode), then the kernels' context.
vine_status alloc_buffers(vine_solver* s, const Caps& k) {
  const vine_graph* g = s->g;
  const vine_options& o = s->opt;
  const long long N = g->n, B = s->B, BN = N * B;
  const long long lab_cap = k.lab, pool_cap = k.pool, cand_cap = k.cand, carry_cap = k.carry;
  ALLOC(s->d_qsrc, sizeof(int32_t) * B);
  ALLOC(s->d_qdst, sizeof(int32_t) * B);
  ALLOC(s->d_qbound, sizeof(float) * B);
  ALLOC(s->d_qdelta, sizeof(float) * B);
  ALLOC(s->d_qbest, sizeof(float2) * B);
  ALLOC(s->d_qbest_id, sizeof(int32_t) * B);
  ALLOC(s->d_qfmin, sizeof(uint32_t) * 2 * B);
  ALLOC(s->d_qlast, sizeof(int32_t) * B);
  ALLOC(s->d_qexp, sizeof(int32_t) * B);
  ALLOC(s->d_qins, sizeof(int32_t) * B);
  ALLOC(s->d_qnear, sizeof(int32_t) * B);
  ALLOC(s->d_qexptot, sizeof(long long) * B);
  ALLOC(s->d_qfcast, sizeof(float) * B);
  ALLOC(s->d_hash_pred, sizeof(float) * 2);
  if (o.record_trace > 0) {
    s->trace_cap = 8LL << 20;
    ALLOC(s->d_trace, sizeof(float) * 10 * s->trace_cap);
  }
  ALLOC(s->d_pset, sizeof(int2) * BN);
  ALLOC(s->d_cand_head, sizeof(unsigned long long) * BN);
  if (o.prune == 2) {
    ALLOC(s->d_lbl, sizeof(float) * BN);
    ALLOC(s->d_lbw, sizeof(float) * BN);
    ALLOC(s->d_qlk, sizeof(float) * B);
    ALLOC(s->d_lbd, sizeof(unsigned long long) * BN);
    ALLOC(s->d_fwd, sizeof(float) * BN);
    s->ell_words = (int32_t)((g->n + 31) / 32);
    ALLOC(s->d_ell, sizeof(uint32_t) * s->ell_words * B);
    ALLOC(s->d_inq, sizeof(uint32_t) * BN);
    ALLOC(s->d_wq, sizeof(int32_t) * 2 * BN);
    ALLOC(s->d_qorder, sizeof(int32_t) * B);
  }
  long long wn_cap = 0;
  if (s->wg) {
    // per query two near lists; each holds at least the largest far chunk so
    // that a pull into an empty list always makes progress
    wn_cap = std::max<long long>((long long)dev::FU << dev::FCLS_MAX, std::min<long long>(carry_cap / B, 1LL << 22));
    ALLOC(s->d_wnear, sizeof(int4) * 2 * wn_cap * B);
    ALLOC(s->d_qfin, sizeof(unsigned long long) * B);
  }
  ALLOC(s->d_pool_cd, sizeof(float2) * pool_cap);
  ALLOC(s->d_pool_id, sizeof(int32_t) * pool_cap);
  ALLOC(s->d_labs, sizeof(int4) * lab_cap);
  ALLOC(s->d_lmeta, sizeof(int2) * lab_cap);
  ALLOC(s->d_dead, lab_cap);
  // candidates travel in chunks of dev::CCHUNK; a round needs one chunk per
  // touched (query, vertex) plus one per CCHUNK further candidates
  long long chunk_cap = std::max<long long>(cand_cap / 2, 1024);
  ALLOC(s->d_chd, sizeof(int2) * chunk_cap);
  ALLOC(s->d_crec, sizeof(int4) * dev::CCHUNK * chunk_cap);
  ALLOC(s->d_touched, sizeof(long long) * chunk_cap);
  ALLOC(s->d_carry[0], sizeof(int4) * carry_cap);
  ALLOC(s->d_carry[1], sizeof(int4) * carry_cap);
  // far tier: units of dev::FU records from a bump allocator (reset per
  // batch), sized for four times the near tier's entries; one drain-list
  // segment at most per unit
  long long far_cap = std::max<long long>(4 * carry_cap / dev::FU, 16LL * B + (1LL << dev::FCLS_MAX) + 4096);
  if (far_cap > (1LL << 27) - 1) far_cap = (1LL << 27) - 1;  // chunk pointers keep 27 bits of unit
  ALLOC(s->d_far_head, sizeof(int32_t) * B * dev::NBK);
  ALLOC(s->d_far_occ, sizeof(unsigned long long) * B * dev::NBW);
  ALLOC(s->d_far_next, sizeof(int32_t) * B);
  ALLOC(s->d_far_cnt, sizeof(int32_t) * B);
  ALLOC(s->d_far_base, sizeof(float) * B);
  ALLOC(s->d_far_bw, sizeof(float) * B);
  ALLOC(s->d_fhd, sizeof(int2) * far_cap);
  ALLOC(s->d_dl, sizeof(int2) * far_cap);
  ALLOC(s->d_frec, sizeof(int4) * dev::FU * far_cap);
  ALLOC(s->d_ts, sizeof(unsigned long long) * dev::TS_CAP);
  ALLOC(s->d_ct, sizeof(Counters));
  if (s->arena || s->count_only) {
    s->h_ct = &s->h_ct_local;
  } else if (hipHostMalloc((void**)&s->h_ct, sizeof(Counters), hipHostMallocDefault) != hipSuccess) {
    set_error("hipHostMalloc failed");
    vine_solver_destroy(s);
    return VINE_E_OOM;
  }
  Ctx& c = s->ctx;
  c.n = g->n;
  c.B = s->B;
  c.row_ptr = g->d_row;
  c.edges = g->d_edges;
  c.rrow_ptr = g->d_rrow;
  c.redges = g->d_redges;
  c.vinv = g->d_inv;
  c.q_src = s->d_qsrc;
  c.q_dst = s->d_qdst;
  c.q_bound = s->d_qbound;
  c.q_delta = s->d_qdelta;
  c.q_best = s->d_qbest;
  c.q_best_id = s->d_qbest_id;
  c.q_fmin = s->d_qfmin;
  c.q_last = s->d_qlast;
  c.q_exp = s->d_qexp;
  c.q_ins = s->d_qins;
  c.q_near = s->d_qnear;
  c.q_exptot = s->d_qexptot;
  c.q_fcast = s->d_qfcast;
  c.hash_slots = o.hash_slots;
  c.hash_mlp = (o.hash_slots > 0 && o.delta_policy == VINE_DELTA_MLP) ? 1 : 0;
  c.hash_pred = s->d_hash_pred;
  c.avg_deg = g->n > 0 ? (float)g->m / (float)g->n : 0.f;
  c.policy = o.delta_policy;
  c.dop_count = -1;
  c.stats_on = (o.delta_policy == VINE_DELTA_MLP || o.record_trace > 0) ? 1 : 0;
  c.mlp = nullptr;
  c.trace = s->d_trace;
  c.trace_cap = s->trace_cap;
  c.trace_every = o.record_trace;
  c.pset = s->d_pset;
  c.cand_head = s->d_cand_head;
  c.lbl = s->d_lbl;
  c.lbw = s->d_lbw;
  c.q_lk = s->d_qlk;
  c.lagrange = o.prune == 2 ? o.lagrange_bound : 0;
  c.lbd = s->d_lbd;
  c.fwd = s->d_fwd;
  c.ell = s->d_ell;
  c.ell_words = s->ell_words;
  c.inq = s->d_inq;
  c.wq = s->d_wq;
  c.q_order = nullptr;  // set per batch when the slots are sorted
  c.wnear = s->d_wnear;
  c.wn_cap = wn_cap;
  c.q_fin = s->d_qfin;
  c.pool_cd = s->d_pool_cd;
  c.pool_id = s->d_pool_id;
  c.pool_cap = (unsigned long long)pool_cap;
  c.labs = s->d_labs;
  c.lmeta = s->d_lmeta;
  c.dead = s->d_dead;
  c.label_cap = lab_cap;
  c.chd = s->d_chd;
  c.crec = s->d_crec;
  c.touched = s->d_touched;
  c.chunk_cap = (int32_t)chunk_cap;
  c.carry[0] = s->d_carry[0];
  c.carry[1] = s->d_carry[1];
  c.carry_cap = (int32_t)carry_cap;
  c.far_head = s->d_far_head;
  c.far_occ = s->d_far_occ;
  c.far_next = s->d_far_next;
  c.far_cnt = s->d_far_cnt;
  c.far_base = s->d_far_base;
  c.far_bw = s->d_far_bw;
  c.fhd = s->d_fhd;
  c.dl = s->d_dl;
  c.frec = s->d_frec;
  c.far_cap = (int32_t)far_cap;
  c.far_floor = 0.0625f * g->mean_len;
  c.far_mult = o.near_windows > 0.f ? o.near_windows : 1.0f;
  c.ts = s->d_ts;
  c.ct = s->d_ct;
  c.ellipse = o.bound_ball ? 0 : 1;
  c.fwd_cut = g->twin_delay && g->integer_weights ? 1 : 0;
  c.lmk = g->d_lmk;
  return VINE_OK;
}

// One solver on one stream.  Automatic capacities are sized for want_slots
// query slots, capped by mem_frac of the free memory.
vine_status create_single(const vine_graph* g, const vine_options& o, double mem_frac, long long want_slots,
                          vine_solver** out) {
  auto* s = new vine_solver();
  init_fields(s, g, o);
  hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    set_error("hipStreamCreate: %s", hipGetErrorString(e));
    delete s;
    return VINE_E_HIP;
  }
  const long long N = g->n, B = s->B, BN = N * B;
  size_t freeb = 0, totalb = 0;
  if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) {
    set_error("hipMemGetInfo failed");
    vine_solver_destroy(s);
    return VINE_E_HIP;
  }
  const size_t per_qv = per_qv_bytes(o);
  double remaining = (double)freeb - (double)per_qv * BN - 64.0 * (1 << 20);
  if (remaining <= 0) {
    set_error("vine_solver_create: %lld query slots x %lld vertices need %.1f GB, %.1f GB free",
            B, N, (double)per_qv * BN / 1e9, (double)freeb / 1e9);
    vine_solver_destroy(s);
    return VINE_E_OOM;
  }
  // auto capacities: share the remaining memory (labels 25 B, pool 12 B x 2,
  // candidate chunks 144 B per two candidate slots, carry 32 B, bound frontier 16 B)
  long long lab_cap = o.label_capacity;
  if (!lab_cap) {
    // sized to the problem (64 labels per (query, vertex) pair plus slack),
    // capped by mem_frac of the free memory
    double per = 25.0 + 2 * 12.0 + 0.25 * 144.0 + 0.5 * 32.0 + 0.25 * 16.0;
    double want = 64.0 * (double)(N * want_slots) + (double)(1 << 22);
    lab_cap = (long long)std::min(std::min(remaining * mem_frac / per, want), (double)INT32_MAX - 1);
    lab_cap = std::max<long long>(lab_cap, 4 * B + 1024);
  }
  Caps k;
  vine_status r = derive_caps(o, B, lab_cap, &k);
  if (r != VINE_OK) {
    vine_solver_destroy(s);
    return r;
  }
  r = alloc_buffers(s, k);
  if (r != VINE_OK) return r;  // alloc_buffers destroyed s
  *out = s;
  return VINE_OK;
}

// `concurrency` solvers, each with its share of the query slots, on their
// own streams.  Explicit capacities are divided among them; automatic ones
// are sized as for the whole batch (the heaviest queries may meet in one
// part), each part capped by an equal share of 60% of the free memory.
vine_status create_split(const vine_graph* g, const vine_options& o, vine_solver** out) {
  const int k = o.concurrency;
  auto* s = new vine_solver();
  s->g = g;
  s->opt = o;
  s->B = o.max_batch;
  vine_options oc = o;
  oc.concurrency = 0;
  oc.max_batch = (o.max_batch + k - 1) / k;
  auto share = [k](int64_t v, int64_t lo) { return v ? std::max<int64_t>((v + k - 1) / k, lo) : 0; };
  oc.label_capacity = share(o.label_capacity, oc.max_batch + 1);
  oc.pool_capacity = share(o.pool_capacity, 4 * (int64_t)oc.max_batch);
  oc.cand_capacity = share(o.cand_capacity, 64);
  oc.frontier_capacity = share(o.frontier_capacity, 64);
  for (int i = 0; i < k; i++) {
    vine_solver* p = nullptr;
    vine_status r = create_single(g, oc, 0.6 / (k - 0.6 * i), o.max_batch, &p);
    if (r != VINE_OK) {
      vine_solver_destroy(s);
      return r;
    }
    s->parts.push_back(p);
  }
  *out = s;
  return VINE_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
namespace {

// Brackets a launch with HIP events on the solver stream (time_kernels = 1);
// no synchronisation here, the pairs are resolved by resolve_events().
struct Timer {
  vine_solver* s;
  bool on;
  hipEvent_t a = nullptr;
  int kind;
  // sample = true: this launch belongs to a round picked by time_kernels
  Timer(vine_solver* s_, int kind_, bool sample = true)
      : s(s_), on(s_->opt.time_kernels != 0 && sample), kind(kind_) {
    if (on) {
      a = next_event(s);
      if (a) (void)hipEventRecord(a, s->stream);
      else on = false;
    }
  }
  static hipEvent_t next_event(vine_solver* s) {
    if (s->ev_used == s->evpool.size()) {
      hipEvent_t e = nullptr;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      s->evpool.push_back(e);
    }
    return s->evpool[s->ev_used++];
  }
  hipError_t stop() {
    if (!on) return hipSuccess;
    hipEvent_t b = next_event(s);
    if (!b) return hipErrorOutOfMemory;
    hipError_t e = hipEventRecord(b, s->stream);
    s->evrec.push_back({a, b, kind});
    return e;
  }
};

vine_status resolve_events(vine_solver* s, vine_stats* st) {
  if (s->evrec.empty()) return VINE_OK;
  HIP_TRY(hipStreamSynchronize(s->stream));
  for (const auto& r : s->evrec) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, r.a, r.b));
    if (r.kind == 0) st->ms_expand += ms;
    else if (r.kind == 1) st->ms_merge += ms;
    else if (r.kind == 2) st->ms_sssp += ms;
    else st->ms_pull += ms;
  }
  s->evrec.clear();
  s->ev_used = 0;
  return VINE_OK;
}

vine_status poll(vine_solver* s) {
  HIP_TRY(hipMemcpyAsync(s->h_ct, s->d_ct, sizeof(Counters), hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (s->h_ct->overflow) {
    set_error("device capacity exceeded (flags 0x%x: 1 labels, 2 candidates, 4 touched, 8 pool, "
            "16 carry, 32 bound frontier, 64 far tier (frontier_capacity)); raise the capacities in vine_options",
            s->h_ct->overflow);
    return VINE_E_OVERFLOW;
  }
  return VINE_OK;
}

// Dynamic LDS of k_expand: the first-level candidate table.
size_t expand_smem(const vine_solver* s) { return (size_t)s->opt.hash_slots * sizeof(dev::HashSlot); }

// sample: 1 brackets this round's kernels with HIP events, 0 never, -1 every
// time_kernels-th round
template <int PRUNE>
vine_status launch_round(vine_solver* s, int gsz, vine_stats* st, int sample_mode) {
  Ctx& c = s->ctx;
  const int tk = std::max(1, s->opt.time_kernels);
  const bool sample = s->opt.time_kernels != 0 &&
                      (sample_mode < 0 ? (s->round_seq++ % tk) == 0 : sample_mode > 0);
  if (sample) st->timed_rounds++;
  {
    Timer t(s, 0, sample);
    hipLaunchKernelGGL((dev::k_expand<c_label, c_constraint<c_label>, c_edge, PRUNE>), dim3(gsz),
                       dim3(s->block), expand_smem(s), s->stream, c, s->lpw);
    HIP_TRY(hipGetLastError());
    HIP_TRY(t.stop());
  }
  {
    Timer t(s, 1, sample);
    hipLaunchKernelGGL((dev::k_merge<c_label, PRUNE>), dim3(gsz), dim3(s->block), 0, s->stream, c);
    HIP_TRY(hipGetLastError());
    HIP_TRY(t.stop());
  }
  if (s->opt.delta_policy != VINE_DELTA_INF || c.stats_on) {
    Timer t(s, 3, sample);
    hipLaunchKernelGGL((dev::k_pull_policy<PRUNE>), dim3((s->B + 3) / 4), dim3(256), 0, s->stream, c,
                       s->nb_active);
    HIP_TRY(hipGetLastError());
    HIP_TRY(t.stop());
  }
  st->expand_launches++;
  st->merge_launches++;
  return VINE_OK;
}

vine_status launch_round_any(vine_solver* s, int gsz, vine_stats* st, int sample_mode = -1) {
  switch (s->opt.prune) {
    case 0:
      return launch_round<0>(s, gsz, st, sample_mode);
    case 1:
      return launch_round<1>(s, gsz, st, sample_mode);
    default:
      return launch_round<2>(s, gsz, st, sample_mode);
  }
}

// use_graph 2 replays rounds once the frontier (near tier, new labels, far
// segments and far tier) holds at most this many entries: the tail of a batch,
// where a few queries' rounds are short and launch gaps dominate
constexpr long long GRAPH_TAIL_FRONTIER = 1 << 15;

// use_graph: K rounds (expand, merge, pull/policy each) as one HIP graph.
// Every round kernel reads its round number and frontier sizes from the
// device counters, so the same K launches are valid for any K consecutive
// rounds: the graph is captured once per (batch size, K) and replayed.  The
// grid is the full s->grid (grid-stride loops cover any frontier size).
// With time_kernels the first round of each chunk is launched directly,
// bracketed by HIP events, and the graph holds the other K - 1.
vine_status capture_rounds(vine_solver* s, int K) {
  if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  s->gexec = nullptr;
  s->graph = nullptr;
  const int gsz = s->grid;
  Ctx& c = s->ctx;
  HIP_TRY(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  hipError_t err = hipSuccess;
  for (int k = 0; k < K && err == hipSuccess; k++) {
    switch (s->opt.prune) {
      case 0:
        hipLaunchKernelGGL((dev::k_expand<c_label, c_constraint<c_label>, c_edge, 0>), dim3(gsz), dim3(s->block),
                           expand_smem(s), s->stream, c, s->lpw);
        break;
      case 1:
        hipLaunchKernelGGL((dev::k_expand<c_label, c_constraint<c_label>, c_edge, 1>), dim3(gsz), dim3(s->block),
                           expand_smem(s), s->stream, c, s->lpw);
        break;
      default:
        hipLaunchKernelGGL((dev::k_expand<c_label, c_constraint<c_label>, c_edge, 2>), dim3(gsz), dim3(s->block),
                           expand_smem(s), s->stream, c, s->lpw);
    }
    if (err == hipSuccess) err = hipGetLastError();
    switch (s->opt.prune) {
      case 0:
        hipLaunchKernelGGL((dev::k_merge<c_label, 0>), dim3(gsz), dim3(s->block), 0, s->stream, c);
        break;
      case 1:
        hipLaunchKernelGGL((dev::k_merge<c_label, 1>), dim3(gsz), dim3(s->block), 0, s->stream, c);
        break;
      default:
        hipLaunchKernelGGL((dev::k_merge<c_label, 2>), dim3(gsz), dim3(s->block), 0, s->stream, c);
    }
    if (err == hipSuccess) err = hipGetLastError();
    if (s->opt.delta_policy != VINE_DELTA_INF || c.stats_on) {
      switch (s->opt.prune) {
        case 0:
          hipLaunchKernelGGL((dev::k_pull_policy<0>), dim3((s->B + 3) / 4), dim3(256), 0, s->stream, c, s->nb_active);
          break;
        case 1:
          hipLaunchKernelGGL((dev::k_pull_policy<1>), dim3((s->B + 3) / 4), dim3(256), 0, s->stream, c, s->nb_active);
          break;
        default:
          hipLaunchKernelGGL((dev::k_pull_policy<2>), dim3((s->B + 3) / 4), dim3(256), 0, s->stream, c, s->nb_active);
      }
      if (err == hipSuccess) err = hipGetLastError();
    }
  }
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(s->stream, &g);
  if (err == hipSuccess) err = ec;
  if (err == hipSuccess) err = hipGraphInstantiate(&s->gexec, g, nullptr, nullptr, 0);
  s->graph = g;
  if (err != hipSuccess) {
    set_error("capturing %d rounds into a HIP graph: %s", K, hipGetErrorString(err));
    (void)hipGetLastError();
    return VINE_E_HIP;
  }
  s->graph_nb = s->nb_active;
  s->graph_k = K;
  s->graph_gsz = gsz;
  return VINE_OK;
}

vine_status launch_rounds_graph(vine_solver* s, int K, vine_stats* st) {
  int kg = K;
  if (s->opt.time_kernels) {
    vine_status ls = launch_round_any(s, s->grid, st, 1);
    if (ls != VINE_OK) return ls;
    kg = K - 1;
  }
  if (kg == 0) return VINE_OK;
  if (!s->gexec || s->graph_nb != s->nb_active || s->graph_k != kg || s->graph_gsz != s->grid) {
    vine_status cs = capture_rounds(s, kg);
    if (cs != VINE_OK) return cs;
  }
  HIP_TRY(hipGraphLaunch(s->gexec, s->stream));
  st->expand_launches += kg;
  st->merge_launches += kg;
  return VINE_OK;
}

vine_status fill_f32(vine_solver* s, float* p, float v, long long n) {
  uint32_t bits;
  std::memcpy(&bits, &v, 4);
  int blocks = (int)std::min<long long>((n + 255) / 256, 65536);
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(dev::k_fill_u32, dim3(blocks), dim3(256), 0, s->stream, (uint32_t*)p, bits, n);
  HIP_TRY(hipGetLastError());
  return VINE_OK;
}

// Bound computation: delta-stepping SSSP on the transposed graph.  The host
// launches every round, so the round number travels as a kernel argument.
int grid_for(const vine_solver* s, long long items) {
  long long want = 2 * items / s->block + 4;
  return (int)std::max<long long>(4, std::min<long long>(want, s->grid));
}

// Bound searches: one workgroup per query slot runs its searches to the end
// (dev::k_bounds), so the whole computation is a single launch.
vine_status run_bounds(vine_solver* s, int nb, vine_stats* st) {
  Ctx& c = s->ctx;
  const float dd = s->opt.bound_delta > 0 ? s->opt.bound_delta : 2.0f * s->g->mean_delay;
  const float dl = s->opt.bound_delta > 0 ? s->opt.bound_delta : 2.0f * s->g->mean_len;
  c.q_order = nullptr;
  if (nb <= dev::ORDER_MAX && nb > 1) {
    c.q_order = s->d_qorder;
    hipLaunchKernelGGL(dev::k_bound_order, dim3(1), dim3(1024), 0, s->stream, c, nb);
    HIP_TRY(hipGetLastError());
  }
  Timer t(s, 2);
  hipLaunchKernelGGL(dev::k_bounds, dim3(nb), dim3(dev::BOUND_BLOCK), 0, s->stream, c, nb, dd, dl);
  HIP_TRY(hipGetLastError());
  HIP_TRY(t.stop());
  st->sssp_launches++;
  return VINE_OK;
}

// One batch of nb <= B queries whose arrays are already in the slot buffers.
vine_status solve_batch(vine_solver* s, int nb, int q0, vine_stats* st) {
  Ctx& c = s->ctx;
  const long long BN = (long long)s->B * c.n;
  // reset per-(query, vertex) headers and counters
  HIP_TRY(hipMemsetAsync(s->d_pset, 0, sizeof(int2) * BN, s->stream));
  // the workgroup engine never reads candidate lists (the four-multiplier
  // pass sets the words it uses itself)
  if (!s->wg) HIP_TRY(hipMemsetAsync(s->d_cand_head, 0xFF, sizeof(unsigned long long) * BN, s->stream));
  HIP_TRY(hipMemsetAsync(s->d_far_head, 0xFF, sizeof(int32_t) * s->B * dev::NBK, s->stream));
  HIP_TRY(hipMemsetAsync(s->d_far_occ, 0, sizeof(unsigned long long) * s->B * dev::NBW, s->stream));
  std::memset(s->h_ct, 0, sizeof(Counters));
  s->h_ct->label_ctr = (unsigned long long)s->B;
  s->h_ct->pool_ctr = (unsigned long long)s->B * dev::INIT_CAP;
  HIP_TRY(hipMemcpyAsync(s->d_ct, s->h_ct, sizeof(Counters), hipMemcpyHostToDevice, s->stream));
  // window per query
  // initial window: fixed policy uses delta_fixed, the learned one starts
  // from delta_fixed (or a quarter arc) and the AR rule from one arc length
  float delta_ar = s->g->mean_len;
  {
    float dv = dev::FINF;
    if (s->opt.delta_policy == VINE_DELTA_FIXED) dv = s->opt.delta_fixed;
    if (s->opt.delta_policy == VINE_DELTA_MLP) dv = s->opt.delta_fixed > 0 ? s->opt.delta_fixed : 0.25f * s->g->mean_len;
    if (s->opt.delta_policy == VINE_DELTA_AR) dv = delta_ar;
    // VINE_DELTA_DOP: no key window, a count of frontier entries per round
    vine_status fs = fill_f32(s, s->d_qdelta, dv, s->B);
    if (fs != VINE_OK) return fs;
  }
  s->nb_active = nb;
  if (s->opt.delta_policy == VINE_DELTA_MLP && !c.mlp) {
    set_error("VINE_DELTA_MLP needs parameters: call vine_solver_set_mlp first");
    return VINE_E_ARG;
  }
  if (s->opt.prune == 2) {
    vine_status fs = fill_f32(s, (float*)s->d_lbd, dev::FINF, 2 * BN);
    if (fs == VINE_OK) fs = fill_f32(s, s->d_lbl, dev::FINF, BN);
    if (fs == VINE_OK) fs = fill_f32(s, s->d_lbw, dev::FINF, BN);
    if (fs == VINE_OK && c.ellipse) fs = fill_f32(s, s->d_fwd, dev::FINF, BN);
    if (fs != VINE_OK) return fs;
    HIP_TRY(hipMemsetAsync(s->d_ell, 0, sizeof(uint32_t) * s->ell_words * s->B, s->stream));
    HIP_TRY(hipMemsetAsync(s->d_inq, 0, sizeof(uint32_t) * BN, s->stream));
    vine_status ss = run_bounds(s, nb, st);
    if (ss != VINE_OK) return ss;
  }
  switch (s->opt.prune) {
    case 0:
      hipLaunchKernelGGL(dev::k_init_batch<0>, dim3((s->B + 255) / 256), dim3(256), 0, s->stream, c, nb);
      break;
    case 1:
      hipLaunchKernelGGL(dev::k_init_batch<1>, dim3((s->B + 255) / 256), dim3(256), 0, s->stream, c, nb);
      break;
    default:
      hipLaunchKernelGGL(dev::k_init_batch<2>, dim3((s->B + 255) / 256), dim3(256), 0, s->stream, c, nb);
  }
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(dev::k_stamp_start, dim3(1), dim3(1), 0, s->stream, c);
  HIP_TRY(hipGetLastError());

  int64_t rounds = 0;
  if (s->wg) {
    // workgroup engine: one launch runs every query's search to the end, in
    // the bound searches' order (largest bound first); its time is reported
    // as ms_expand
    {
      Timer t(s, 0);
      hipLaunchKernelGGL((dev::k_search<c_label, c_constraint<c_label>, c_edge>), dim3(nb), dim3(dev::WG_BLOCK), 0,
                         s->stream, c, nb);
      HIP_TRY(hipGetLastError());
      HIP_TRY(t.stop());
    }
    st->expand_launches++;
    vine_status ps = poll(s);
    if (ps != VINE_OK) return ps;
    rounds = s->h_ct->round_e;
  } else {
    // scheduling
    const bool count = s->opt.delta_policy == VINE_DELTA_DOP;
    const bool ar = s->opt.delta_policy == VINE_DELTA_AR || count;
    DopRule rule;
    c.dop_count = -1;
    if (ar) {
      double mx = s->opt.ar_max_labels;
      if (mx <= 0) {
        size_t fb = 0, tb = 0;
        HIP_TRY(hipMemGetInfo(&fb, &tb));
        mx = (double)fb / 12.0;  // free memory / sizeof(c_label), algorithm/cspp.cuh:43-47
      }
      rule.reset(mx);
      if (count) c.dop_count = std::max<int64_t>(rule.dop(), 1);
    }
    const int K = ar ? 1 : std::max(1, s->opt.rounds_per_sync);
    // the AR rule sets the window from the host after every round, so it
    // launches directly; the other policies may replay a captured chunk
    const bool graph = s->opt.use_graph && !ar;
    long long fn_last = nb;
    unsigned long long exp_prev = 0;
    for (;;) {
      // use_graph 2: replay only in the tail, once the frontier is small
      if (graph && (s->opt.use_graph == 1 || fn_last <= GRAPH_TAIL_FRONTIER)) {
        vine_status ls = launch_rounds_graph(s, K, st);
        if (ls != VINE_OK) return ls;
        rounds += K;
      } else {
        for (int k = 0; k < K; k++) {
          // the frontier may grow between polls: size the grid for 4x growth
          // (grid-stride loops keep any size correct)
          vine_status ls = launch_round_any(s, grid_for(s, fn_last * 4), st);
          if (ls != VINE_OK) return ls;
          rounds++;
        }
      }
      vine_status ps = poll(s);
      if (ps != VINE_OK) return ps;
      const Counters& h = *s->h_ct;
      int par = h.round_e & 1;
      // near tier, new labels, far segments queued for the next expand, far tier
      long long fn = (long long)h.carry_n[par] + ((long long)h.label_ctr - h.lstart[par]) +
                     (long long)h.drain_n * dev::FSEG + h.far_live;
      st->max_frontier = std::max<int64_t>(st->max_frontier, fn);
      fn_last = std::max<long long>(fn, 1);
      if (fn == 0) break;
      // every round expands at least the minimum-key label of each live query
      // (or re-keys carried labels once), so rounds are bounded by twice the
      // number of labels ever created
      if (rounds > 2 * (int64_t)h.label_ctr + 64) {
        set_error("solver made no progress after %lld rounds", (long long)rounds);
        return VINE_E_OVERFLOW;
      }
      if (count) {
        // the reference's count semantics (algorithm/cspp.cuh:179-216): the
        // next round expands the last DOP entries of the frontier (all of them
        // when DOP >= frontier) and carries the rest
        rule.update((int64_t)fn);
        c.dop_count = std::max<int64_t>(rule.dop(), 1);
      } else if (ar) {
        // the reference's DOP rule picks how many labels the next round should
        // expand; a multiplicative controller turns that count into the key
        // window (DOP >= frontier means the whole frontier, as in the reference)
        rule.update((int64_t)fn);
        const double dop = (double)rule.dop();
        const unsigned long long x = h.expanded - exp_prev;
        exp_prev = h.expanded;
        if (dop >= (double)fn) {
          delta_ar = dev::FINF;
        } else {
          if (!(delta_ar < dev::FINF)) delta_ar = s->g->mean_len;
          double ratio = x > 0 ? dop / (double)x : 4.0;
          delta_ar = (float)std::min(1e7, std::max(0.25, delta_ar * std::min(4.0, std::max(0.25, std::sqrt(ratio)))));
        }
        vine_status fs = fill_f32(s, s->d_qdelta, delta_ar, s->B);
        if (fs != VINE_OK) return fs;
      }
    }
  }
  hipLaunchKernelGGL(dev::k_stamp_round, dim3(1), dim3(1), 0, s->stream, c, (int)rounds);
  HIP_TRY(hipGetLastError());
  if (s->h_ct->range_flag) {
    set_error("a label sum reached 2^24; float32 sums are no longer exact");
    return VINE_E_RANGE;
  }
  vine_status es = resolve_events(s, st);
  if (es != VINE_OK) return es;
  if (s->d_trace) {
    long long nrow = std::min<long long>(s->h_ct->trace_n, s->trace_cap);
    size_t at = s->trace.size();
    s->trace.resize(at + 10 * (size_t)nrow);
    if (nrow > 0)
      HIP_TRY(hipMemcpy(s->trace.data() + at, s->d_trace, sizeof(float) * 10 * nrow, hipMemcpyDeviceToHost));
    // slot -> query index of the whole solve
    for (long long i = 0; i < nrow; i++) s->trace[at + 10 * i] += (float)q0;
  }
  st->rounds += rounds;
  st->expanded += (int64_t)s->h_ct->expanded;
  st->relaxed += (int64_t)s->h_ct->relaxed;
  st->candidates += (int64_t)s->h_ct->cands;
  st->inserted += (int64_t)s->h_ct->inserted;
  st->sssp_relaxed += (int64_t)s->h_ct->sssp_relaxed;
  st->sssp_rounds += s->h_ct->sround;
  st->touched += (int64_t)s->h_ct->touched_total;
  st->merge_scanned += (int64_t)s->h_ct->merge_scanned;
  st->merge_chain_max = std::max<int64_t>(st->merge_chain_max, s->h_ct->merge_chain_max);
  st->far_pulled += (int64_t)s->h_ct->far_pulled;
  st->near_carried += (int64_t)s->h_ct->near_carried;
  st->far_appended += (int64_t)s->h_ct->far_appended;
  st->far_units += (int64_t)s->h_ct->far_chunks;
  st->cand_chunks += (int64_t)s->h_ct->cand_chunks;
  st->scanned += (int64_t)s->h_ct->scanned;
  st->hash_dropped += (int64_t)s->h_ct->hash_dropped;
  st->hash_overflow += (int64_t)s->h_ct->hash_overflow;
  st->batches++;
  return VINE_OK;
}

vine_status check_queries_host(const vine_solver* s, int32_t nq, const int32_t* src, const int32_t* dst,
                               const float* bound) {
  for (int32_t i = 0; i < nq; i++) {
    if (src[i] < 0 || src[i] >= s->g->n || dst[i] < 0 || dst[i] >= s->g->n) {
      set_error("query %d: vertex id out of range [0, %d)", i, s->g->n);
      return VINE_E_ARG;
    }
    if (!(bound[i] >= 0.f) || !(bound[i] < dev::EXACT_LIMIT)) {
      set_error("query %d: bound %g must be in [0, 2^24)", i, (double)bound[i]);
      return VINE_E_ARG;
    }
  }
  return VINE_OK;
}

__global__ void k_check_queries(const int32_t* src, const int32_t* dst, const float* bound, int nq, int n,
                                int* bad) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  if (src[i] < 0 || src[i] >= n || dst[i] < 0 || dst[i] >= n || !(bound[i] >= 0.f) ||
      !(bound[i] < dev::EXACT_LIMIT))
    atomicMin(bad, i);
}

// Solve nq queries whose arrays are in device memory; answers to device memory.
// Per-query records (latency, work, trace rows) go to index rec_base + i of
// records sized rec_total; rec_base = 0 starts a new solve's records.
vine_status solve_impl(vine_solver* s, int32_t nq, const int32_t* d_src, const int32_t* d_dst,
                       const float* d_bound, float* d_cost, float* d_delay, int32_t* d_feas,
                       int32_t* d_plen, vine_stats* st, int32_t rec_base, int32_t rec_total) {
  Ctx& c = s->ctx;
  if (rec_base == 0) {
    s->latency_us.assign(rec_total, 0.0);
    s->round_seq = 0;
    s->q_expanded.assign(rec_total, 0);
    s->q_rounds.assign(rec_total, 0);
    s->trace.clear();
  }
  for (int32_t q0 = 0; q0 < nq; q0 += s->B) {
    int nb = std::min<int32_t>(s->B, nq - q0);
    // endpoints in internal vertex ids (ids were validated by the caller)
    hipLaunchKernelGGL(dev::k_map_ids, dim3((nb + 255) / 256), dim3(256), 0, s->stream, d_src + q0, s->g->d_perm,
                       s->d_qsrc, nb);
    hipLaunchKernelGGL(dev::k_map_ids, dim3((nb + 255) / 256), dim3(256), 0, s->stream, d_dst + q0, s->g->d_perm,
                       s->d_qdst, nb);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(s->d_qbound, d_bound + q0, sizeof(float) * nb, hipMemcpyDeviceToDevice, s->stream));
    if (nb < s->B) {
      // idle slots: harmless query 0 -> 0 with bound 0 (never initialised as live)
      HIP_TRY(hipMemsetAsync(s->d_qsrc + nb, 0, sizeof(int32_t) * (s->B - nb), s->stream));
      HIP_TRY(hipMemsetAsync(s->d_qdst + nb, 0, sizeof(int32_t) * (s->B - nb), s->stream));
      HIP_TRY(hipMemsetAsync(s->d_qbound + nb, 0, sizeof(float) * (s->B - nb), s->stream));
    }
    vine_status bs = solve_batch(s, nb, rec_base + q0, st);
    if (bs != VINE_OK) return bs;
    // per-query finish times: merge-end clock of the last round with live labels
    std::vector<int32_t> last(nb);
    HIP_TRY(hipMemcpyAsync(last.data(), s->d_qlast, sizeof(int32_t) * nb, hipMemcpyDeviceToHost, s->stream));
    std::vector<unsigned long long> ts(dev::TS_CAP);
    int64_t nr = std::min<int64_t>(s->h_ct->round_e + 2, dev::TS_CAP);
    HIP_TRY(hipMemcpyAsync(ts.data(), s->d_ts, sizeof(unsigned long long) * nr, hipMemcpyDeviceToHost, s->stream));
    unsigned long long tstart = 0;
    HIP_TRY(hipMemcpyAsync(&tstart, &s->d_ct->t_start, sizeof(tstart), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    int rate_khz = 100000;
    if (hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, s->g->device) != hipSuccess ||
        rate_khz <= 0)
      rate_khz = 100000;
    {
      std::vector<long long> ex(nb);
      HIP_TRY(hipMemcpy(ex.data(), s->d_qexptot, sizeof(long long) * nb, hipMemcpyDeviceToHost));
      for (int i = 0; i < nb; i++) {
        const int32_t at = rec_base + q0 + i;
        if (at < (int32_t)s->q_expanded.size()) {
          s->q_expanded[at] = ex[i];
          s->q_rounds[at] = last[i] + 1;
        }
      }
    }
    std::vector<unsigned long long> fin;
    if (s->wg) {
      // workgroup engine: each query's workgroup stamps its own end
      fin.resize(nb);
      HIP_TRY(hipMemcpy(fin.data(), s->d_qfin, sizeof(unsigned long long) * nb, hipMemcpyDeviceToHost));
    }
    for (int i = 0; i < nb; i++) {
      int r = last[i] + 1;  // the query's last active round ends when the next one starts
      double us = 0.0;
      if (s->wg) {
        if (fin[i] >= tstart) us = (double)(fin[i] - tstart) * 1000.0 / rate_khz;
      } else if (r >= 1 && r < nr && ts[r] >= tstart) {
        us = (double)(ts[r] - tstart) * 1000.0 / rate_khz;
      }
      if (rec_base + q0 + i < (int32_t)s->latency_us.size()) s->latency_us[rec_base + q0 + i] = us;
    }
    hipLaunchKernelGGL(dev::k_answers, dim3((nb + 255) / 256), dim3(256), 0, s->stream, c, nb, d_cost + q0,
                       d_delay + q0, d_feas + q0, d_plen ? d_plen + q0 : nullptr);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipStreamSynchronize(s->stream));
  return VINE_OK;
}

vine_status ensure_io(vine_solver* s, int32_t nq) {
  if (nq <= s->in_cap) return VINE_OK;
  if (s->arena && s->in_cap > 0) {
    set_error("the static scratch holds the query staging buffers of one call only");
    return VINE_E_ARG;
  }
  for (void* p : {(void*)s->d_in_q, (void*)s->d_in_bound, (void*)s->d_out_cost, (void*)s->d_out_delay,
                  (void*)s->d_out_feas, (void*)s->d_out_plen}) {
    if (!p) continue;
    auto it = std::find(s->allocs.begin(), s->allocs.end(), p);
    if (it != s->allocs.end()) s->allocs.erase(it);
    (void)hipFree(p);
  }
  s->in_cap = 0;
  vine_status st;
  if ((st = dev_alloc(s, (void**)&s->d_in_q, sizeof(int32_t) * 2 * nq)) != VINE_OK) return st;
  if ((st = dev_alloc(s, (void**)&s->d_in_bound, sizeof(float) * nq)) != VINE_OK) return st;
  if ((st = dev_alloc(s, (void**)&s->d_out_cost, sizeof(float) * nq)) != VINE_OK) return st;
  if ((st = dev_alloc(s, (void**)&s->d_out_delay, sizeof(float) * nq)) != VINE_OK) return st;
  if ((st = dev_alloc(s, (void**)&s->d_out_feas, sizeof(int32_t) * nq)) != VINE_OK) return st;
  if ((st = dev_alloc(s, (void**)&s->d_out_plen, sizeof(int32_t) * nq)) != VINE_OK) return st;
  s->in_cap = nq;
  return VINE_OK;
}

// ---- concurrency > 1: query i is part (i % k)'s query (i / k)
__global__ void k_part_gather(int nq, int k, int p, const int32_t* src, const int32_t* dst, const float* bound,
                              int32_t* psrc, int32_t* pdst, float* pbound) {
  const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long i = j * k + p;
  if (i >= nq) return;
  psrc[j] = src[i];
  pdst[j] = dst[i];
  pbound[j] = bound[i];
}

__global__ void k_part_scatter(int nq, int k, int p, const float* pcost, const float* pdelay, const int32_t* pfeas,
                               float* cost, float* delay, int32_t* feas) {
  const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long i = j * k + p;
  if (i >= nq) return;
  cost[i] = pcost[j];
  delay[i] = pdelay[j];
  feas[i] = pfeas[j];
}

int32_t part_size(int32_t nq, int k, int p) { return p < nq ? (nq - p + k - 1) / k : 0; }

void merge_stats(vine_stats& a, const vine_stats& b) {
  a.rounds += b.rounds;
  a.sssp_rounds += b.sssp_rounds;
  a.expanded += b.expanded;
  a.relaxed += b.relaxed;
  a.candidates += b.candidates;
  a.inserted += b.inserted;
  a.sssp_relaxed += b.sssp_relaxed;
  a.max_frontier = std::max(a.max_frontier, b.max_frontier);
  a.batches += b.batches;
  a.ms_kernels += b.ms_kernels;
  a.ms_expand += b.ms_expand;
  a.ms_merge += b.ms_merge;
  a.ms_sssp += b.ms_sssp;
  a.expand_launches += b.expand_launches;
  a.merge_launches += b.merge_launches;
  a.sssp_launches += b.sssp_launches;
  a.touched += b.touched;
  a.merge_scanned += b.merge_scanned;
  a.merge_chain_max = std::max(a.merge_chain_max, b.merge_chain_max);
  a.far_pulled += b.far_pulled;
  a.ms_pull += b.ms_pull;
  a.near_carried += b.near_carried;
  a.far_appended += b.far_appended;
  a.scanned += b.scanned;
  a.timed_rounds += b.timed_rounds;
  a.far_units += b.far_units;
  a.cand_chunks += b.cand_chunks;
  a.hash_dropped += b.hash_dropped;
  a.hash_overflow += b.hash_overflow;
}

// Per-query records of the parts, in the caller's query order.
void gather_records(vine_solver* s, int32_t nq) {
  const int k = (int)s->parts.size();
  s->latency_us.assign(nq, 0.0);
  s->q_expanded.assign(nq, 0);
  s->q_rounds.assign(nq, 0);
  s->trace.clear();
  for (int p = 0; p < k; p++) {
    const vine_solver* c = s->parts[p];
    const int32_t np = part_size(nq, k, p);
    for (int32_t j = 0; j < np; j++) {
      const int32_t i = j * k + p;
      if (j < (int32_t)c->latency_us.size()) s->latency_us[i] = c->latency_us[j];
      if (j < (int32_t)c->q_expanded.size()) {
        s->q_expanded[i] = c->q_expanded[j];
        s->q_rounds[i] = c->q_rounds[j];
      }
    }
    const size_t at = s->trace.size();
    s->trace.insert(s->trace.end(), c->trace.begin(), c->trace.end());
    for (size_t r = at; r < s->trace.size(); r += 10) s->trace[r] = s->trace[r] * (float)k + (float)p;
  }
}

// Runs fn(p) for every part with queries on its own host thread; the first
// failing part's status and message become the call's.
template <typename Fn>
vine_status run_parts(vine_solver* s, int32_t nq, Fn fn) {
  const int k = (int)s->parts.size();
  std::vector<vine_status> rc(k, VINE_OK);
  std::vector<std::string> msg(k);
  std::vector<std::thread> th;
  for (int p = 0; p < k; p++) {
    if (part_size(nq, k, p) == 0) continue;
    th.emplace_back([&, p] {
      rc[p] = fn(p);
      if (rc[p] != VINE_OK) msg[p] = vine_last_error();
    });
  }
  for (auto& t : th) t.join();
  for (int p = 0; p < k; p++)
    if (rc[p] != VINE_OK && rc[p] != VINE_E_OVERFLOW) {
      set_error("part %d of %d (queries %d, %d + %d, ...): %s", p, k, p, p, k, msg[p].c_str());
      return rc[p];
    }
  for (int p = 0; p < k; p++)
    if (rc[p] == VINE_E_OVERFLOW) {
      set_error("part %d of %d: %s", p, k, msg[p].c_str());
      return rc[p];
    }
  return VINE_OK;
}

vine_status solve_split_device(vine_solver* s, int32_t nq, const int32_t* d_src, const int32_t* d_dst,
                               const float* d_bound, float* d_cost, float* d_delay, int32_t* d_feas,
                               vine_stats* out_stats) {
  const int k = (int)s->parts.size();
  const int32_t need = (nq + k - 1) / k;
  if (need > s->split_cap) {
    for (void* p : s->split_bufs) (void)hipFree(p);
    s->split_bufs.assign(k, nullptr);
    s->split_cap = 0;
    for (int p = 0; p < k; p++) {
      hipError_t e = hipMalloc(&s->split_bufs[p], (size_t)24 * need);
      if (e != hipSuccess) {
        set_error("hipMalloc of %zu bytes for the query parts: %s", (size_t)24 * need, hipGetErrorString(e));
        return VINE_E_OOM;
      }
    }
    s->split_cap = need;
  }
  std::vector<vine_stats> st(k);
  const double t0 = now_ms();
  vine_status r = run_parts(s, nq, [&](int p) -> vine_status {
    vine_solver* c = s->parts[p];
    const int32_t np = part_size(nq, k, p);
    char* base = (char*)s->split_bufs[p];
    int32_t* psrc = (int32_t*)base;
    int32_t* pdst = psrc + s->split_cap;
    float* pbound = (float*)(pdst + s->split_cap);
    float* pcost = pbound + s->split_cap;
    float* pdelay = pcost + s->split_cap;
    int32_t* pfeas = (int32_t*)(pdelay + s->split_cap);
    const dim3 grid((np + 255) / 256);
    hipLaunchKernelGGL(k_part_gather, grid, dim3(256), 0, c->stream, nq, k, p, d_src, d_dst, d_bound, psrc, pdst,
                       pbound);
    HIP_TRY(hipGetLastError());
    vine_status v = vine_solve_device(c, np, psrc, pdst, pbound, pcost, pdelay, pfeas, &st[p]);
    if (v != VINE_OK) return v;
    hipLaunchKernelGGL(k_part_scatter, grid, dim3(256), 0, c->stream, nq, k, p, pcost, pdelay, pfeas, d_cost,
                       d_delay, d_feas);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    return VINE_OK;
  });
  if (r != VINE_OK) return r;
  vine_stats acc{};
  for (int p = 0; p < k; p++) merge_stats(acc, st[p]);
  acc.ms_total = now_ms() - t0;
  if (out_stats) *out_stats = acc;
  gather_records(s, nq);
  return VINE_OK;
}

vine_status solve_split_host(vine_solver* s, int32_t nq, const int32_t* src, const int32_t* dst, const float* bound,
                             float* out_cost, float* out_delay, int32_t* out_feasible, int64_t* out_path_off,
                             int32_t* out_path, int64_t path_cap, vine_stats* out_stats) {
  const int k = (int)s->parts.size();
  struct Part {
    std::vector<int32_t> src, dst, feas, path;
    std::vector<float> bound, cost, delay;
    std::vector<int64_t> off;
    vine_stats st{};
    bool path_overflow = false;
  };
  std::vector<Part> part(k);
  const bool want_off = out_path_off != nullptr;
  for (int p = 0; p < k; p++) {
    Part& P = part[p];
    const int32_t np = part_size(nq, k, p);
    for (int32_t j = 0; j < np; j++) {
      const int32_t i = j * k + p;
      P.src.push_back(src[i]);
      P.dst.push_back(dst[i]);
      P.bound.push_back(bound[i]);
    }
    P.cost.resize(np);
    P.delay.resize(np);
    P.feas.resize(np);
    if (want_off) P.off.resize(np + 1);
    if (out_path) P.path.resize(std::max<int64_t>(path_cap, 1));
  }
  const double t0 = now_ms();
  vine_status r = run_parts(s, nq, [&](int p) -> vine_status {
    Part& P = part[p];
    const int32_t np = (int32_t)P.src.size();
    vine_status v = vine_solve(s->parts[p], np, P.src.data(), P.dst.data(), P.bound.data(), P.cost.data(),
                               P.delay.data(), P.feas.data(), want_off ? P.off.data() : nullptr,
                               out_path ? P.path.data() : nullptr, out_path ? path_cap : 0, &P.st);
    // a part whose paths alone overflow has its answers and offsets; the
    // total is checked below
    if (v == VINE_E_OVERFLOW && out_path && P.off[np] > path_cap) {
      P.path_overflow = true;
      return VINE_OK;
    }
    return v;
  });
  if (r != VINE_OK) return r;
  vine_stats acc{};
  for (int p = 0; p < k; p++) {
    const Part& P = part[p];
    merge_stats(acc, P.st);
    for (size_t j = 0; j < P.src.size(); j++) {
      const size_t i = j * k + p;
      out_cost[i] = P.cost[j];
      out_delay[i] = P.delay[j];
      out_feasible[i] = P.feas[j];
    }
  }
  if (want_off) {
    out_path_off[0] = 0;
    for (int32_t i = 0; i < nq; i++) {
      const Part& P = part[i % k];
      const int32_t j = i / k;
      out_path_off[i + 1] = out_path_off[i] + (P.off[j + 1] - P.off[j]);
    }
  }
  acc.ms_total = now_ms() - t0;
  if (out_stats) *out_stats = acc;
  gather_records(s, nq);
  if (out_path) {
    bool overflow = out_path_off[nq] > path_cap;
    for (const Part& P : part) overflow = overflow || P.path_overflow;
    if (overflow) {
      set_error("vine_solve: path_cap %lld too small (need %lld)", (long long)path_cap, (long long)out_path_off[nq]);
      return VINE_E_OVERFLOW;
    }
    for (int32_t i = 0; i < nq; i++) {
      const Part& P = part[i % k];
      const int32_t j = i / k;
      std::copy(P.path.begin() + P.off[j], P.path.begin() + P.off[j + 1], out_path + out_path_off[i]);
    }
  }
  return VINE_OK;
}

// Paths of the nb queries of the batch whose label log is resident: lengths
// plen[0..nb) (host), written to out_path in the caller's vertex ids.  The
// device buffers grow on demand and stay with the handle.
vine_status extract_paths(vine_solver* s, int nb, const int32_t* plen, int32_t* out_path) {
  long long need = 0;
  for (int i = 0; i < nb; i++) need += plen[i];
  if (need == 0) return VINE_OK;
  if (s->path_off_cap < nb + 1) {
    if (s->d_path_off) (void)hipFree(s->d_path_off);
    s->d_path_off = nullptr;
    s->path_off_cap = 0;
    hipError_t e = hipMalloc(&s->d_path_off, sizeof(long long) * (nb + 1));
    if (e != hipSuccess) {
      set_error("hipMalloc for path offsets: %s", hipGetErrorString(e));
      return VINE_E_OOM;
    }
    s->path_off_cap = nb + 1;
  }
  if (s->path_cap < need) {
    if (s->d_path) (void)hipFree(s->d_path);
    s->d_path = nullptr;
    s->path_cap = 0;
    const long long cap = std::max<long long>(need, s->path_cap * 2);
    hipError_t e = hipMalloc(&s->d_path, sizeof(int32_t) * cap);
    if (e != hipSuccess) {
      set_error("hipMalloc for paths: %s", hipGetErrorString(e));
      return VINE_E_OOM;
    }
    s->path_cap = cap;
  }
  std::vector<long long> off(nb + 1, 0);
  for (int i = 0; i < nb; i++) off[i + 1] = off[i] + plen[i];
  HIP_TRY(hipMemcpyAsync(s->d_path_off, off.data(), sizeof(long long) * (nb + 1), hipMemcpyHostToDevice, s->stream));
  hipLaunchKernelGGL(dev::k_paths, dim3((nb + 255) / 256), dim3(256), 0, s->stream, s->ctx, nb, s->d_path_off,
                     s->d_path);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out_path, s->d_path, sizeof(int32_t) * need, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return VINE_OK;
}

}  // namespace

extern "C" {

vine_status vine_solve_device(vine_solver* s, int32_t nq, const int32_t* d_src, const int32_t* d_dst,
                              const float* d_bound, float* d_cost, float* d_delay, int32_t* d_feasible,
                              vine_stats* out_stats) {
  if (!s || nq < 0 || (nq > 0 && (!d_src || !d_dst || !d_bound || !d_cost || !d_delay || !d_feasible))) {
    set_error("vine_solve_device: invalid arguments");
    return VINE_E_ARG;
  }
  if (!s->parts.empty()) return solve_split_device(s, nq, d_src, d_dst, d_bound, d_cost, d_delay, d_feasible,
                                                   out_stats);
  vine_stats st{};
  double t0 = now_ms();
  if (nq > 0) {
    // validate ids and bounds on the device (no host copy of the query set)
    vine_status es = ensure_io(s, 1);
    if (es != VINE_OK) return es;
    int big = INT32_MAX;
    HIP_TRY(hipMemcpyAsync(s->d_out_plen, &big, sizeof(int), hipMemcpyHostToDevice, s->stream));
    hipLaunchKernelGGL(k_check_queries, dim3((nq + 255) / 256), dim3(256), 0, s->stream, d_src, d_dst, d_bound,
                       nq, s->g->n, s->d_out_plen);
    int bad = 0;
    HIP_TRY(hipMemcpyAsync(&bad, s->d_out_plen, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (bad != INT32_MAX) {
      set_error("query %d: vertex id out of range or bound not in [0, 2^24)", bad);
      return VINE_E_ARG;
    }
  }
  vine_status r = solve_impl(s, nq, d_src, d_dst, d_bound, d_cost, d_delay, d_feasible, nullptr, &st, 0, nq);
  s->last_nq = r == VINE_OK ? nq : -1;
  s->last_one_batch = nq <= s->B;
  st.ms_total = now_ms() - t0;
  st.ms_kernels = st.ms_expand + st.ms_merge + st.ms_sssp + st.ms_pull;
  if (out_stats) *out_stats = st;
  return r;
}

vine_status vine_solve(vine_solver* s, int32_t nq, const int32_t* src, const int32_t* dst, const float* bound,
                       float* out_cost, float* out_delay, int32_t* out_feasible, int64_t* out_path_off,
                       int32_t* out_path, int64_t path_cap, vine_stats* out_stats) {
  if (!s || nq < 0 || (nq > 0 && (!src || !dst || !bound || !out_cost || !out_delay || !out_feasible)) ||
      (out_path && (!out_path_off || path_cap < 0))) {
    set_error("vine_solve: invalid arguments");
    return VINE_E_ARG;
  }
  vine_status v = check_queries_host(s, nq, src, dst, bound);
  if (v != VINE_OK) return v;
  if (!s->parts.empty())
    return solve_split_host(s, nq, src, dst, bound, out_cost, out_delay, out_feasible, out_path_off, out_path,
                            path_cap, out_stats);
  vine_stats st{};
  double t0 = now_ms();
  if (nq == 0) {
    if (out_path_off) out_path_off[0] = 0;
    if (out_stats) *out_stats = st;
    return VINE_OK;
  }
  v = ensure_io(s, nq);
  if (v != VINE_OK) return v;
  HIP_TRY(hipMemcpyAsync(s->d_in_q, src, sizeof(int32_t) * nq, hipMemcpyHostToDevice, s->stream));
  HIP_TRY(hipMemcpyAsync(s->d_in_q + nq, dst, sizeof(int32_t) * nq, hipMemcpyHostToDevice, s->stream));
  HIP_TRY(hipMemcpyAsync(s->d_in_bound, bound, sizeof(float) * nq, hipMemcpyHostToDevice, s->stream));
  const bool want_path = out_path != nullptr;
  // paths must be read before the next batch overwrites the label log: solve
  // batch by batch and extract paths per batch
  std::vector<int32_t> plen(nq, 0);
  if (out_path_off) out_path_off[0] = 0;
  bool path_overflow = false;
  for (int32_t q0 = 0; q0 < nq; q0 += s->B) {
    int nb = std::min<int32_t>(s->B, nq - q0);
    v = solve_impl(s, nb, s->d_in_q + q0, s->d_in_q + nq + q0, s->d_in_bound + q0, s->d_out_cost + q0,
                   s->d_out_delay + q0, s->d_out_feas + q0, s->d_out_plen + q0, &st, q0, nq);
    if (v != VINE_OK) return v;
    if (want_path || out_path_off) {
      HIP_TRY(hipMemcpy(plen.data() + q0, s->d_out_plen + q0, sizeof(int32_t) * nb, hipMemcpyDeviceToHost));
      for (int i = 0; i < nb; i++) out_path_off[q0 + i + 1] = out_path_off[q0 + i] + plen[q0 + i];
    }
    if (want_path) {
      if (out_path_off[q0 + nb] > path_cap) {
        path_overflow = true;
      } else {
        v = extract_paths(s, nb, plen.data() + q0, out_path + out_path_off[q0]);
        if (v != VINE_OK) return v;
      }
    }
  }
  s->last_nq = nq;
  s->last_one_batch = nq <= s->B;
  HIP_TRY(hipMemcpy(out_cost, s->d_out_cost, sizeof(float) * nq, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(out_delay, s->d_out_delay, sizeof(float) * nq, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(out_feasible, s->d_out_feas, sizeof(int32_t) * nq, hipMemcpyDeviceToHost));
  st.ms_total = now_ms() - t0;
  st.ms_kernels = st.ms_expand + st.ms_merge + st.ms_sssp + st.ms_pull;
  if (out_stats) *out_stats = st;
  if (path_overflow) {
    set_error("vine_solve: path_cap %lld too small (need %lld)", (long long)path_cap,
            (long long)out_path_off[nq]);
    return VINE_E_OVERFLOW;
  }
  return VINE_OK;
}

vine_status vine_solver_paths(vine_solver* s, int32_t nq, int64_t* out_path_off, int32_t* out_path,
                              int64_t path_cap) {
  if (!s || nq < 0 || !out_path_off || (path_cap > 0 && !out_path) || path_cap < 0) {
    set_error("vine_solver_paths: invalid arguments");
    return VINE_E_ARG;
  }
  if (!s->parts.empty()) {
    // query i is part (i % k)'s query (i / k): each part's paths, interleaved back
    const int k = (int)s->parts.size();
    std::vector<std::vector<int64_t>> off(k);
    std::vector<std::vector<int32_t>> path(k);
    for (int p = 0; p < k; p++) {
      const int32_t np = part_size(nq, k, p);
      off[p].assign(np + 1, 0);
      if (np == 0) continue;
      vine_status r = vine_solver_paths(s->parts[p], np, off[p].data(), nullptr, 0);
      if (r != VINE_OK && r != VINE_E_OVERFLOW) return r;
      path[p].resize(std::max<int64_t>(off[p][np], 1));
      r = vine_solver_paths(s->parts[p], np, off[p].data(), path[p].data(), (int64_t)path[p].size());
      if (r != VINE_OK) return r;
    }
    out_path_off[0] = 0;
    for (int32_t i = 0; i < nq; i++) {
      const auto& o = off[i % k];
      out_path_off[i + 1] = out_path_off[i] + (o[i / k + 1] - o[i / k]);
    }
    if (out_path_off[nq] > path_cap) {
      set_error("vine_solver_paths: path_cap %lld too small (need %lld)", (long long)path_cap,
                (long long)out_path_off[nq]);
      return VINE_E_OVERFLOW;
    }
    for (int32_t i = 0; i < nq; i++) {
      const auto& o = off[i % k];
      const auto& pv = path[i % k];
      std::copy(pv.begin() + o[i / k], pv.begin() + o[i / k + 1], out_path + out_path_off[i]);
    }
    return VINE_OK;
  }
  if (s->last_nq != nq || !s->last_one_batch) {
    set_error("vine_solver_paths: the last solve must have had exactly these %d queries in one batch "
              "(at most max_batch = %d); it had %d", nq, s->B, s->last_nq);
    return VINE_E_ARG;
  }
  out_path_off[0] = 0;
  if (nq == 0) return VINE_OK;
  vine_status es = ensure_io(s, nq);
  if (es != VINE_OK) return es;
  std::vector<int32_t> plen(nq);
  hipLaunchKernelGGL(dev::k_answers, dim3((nq + 255) / 256), dim3(256), 0, s->stream, s->ctx, nq, (float*)nullptr,
                     (float*)nullptr, (int32_t*)nullptr, s->d_out_plen);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(plen.data(), s->d_out_plen, sizeof(int32_t) * nq, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  for (int32_t i = 0; i < nq; i++) out_path_off[i + 1] = out_path_off[i] + plen[i];
  if (out_path_off[nq] > path_cap) {
    set_error("vine_solver_paths: path_cap %lld too small (need %lld)", (long long)path_cap,
              (long long)out_path_off[nq]);
    return VINE_E_OVERFLOW;
  }
  return extract_paths(s, nq, plen.data(), out_path);
}

vine_status vine_solver_latency(const vine_solver* s, int32_t nq, double* out_us) {
  if (!s || !out_us || nq < 0) return VINE_E_ARG;
  for (int32_t i = 0; i < nq; i++) out_us[i] = i < (int32_t)s->latency_us.size() ? s->latency_us[i] : 0.0;
  return VINE_OK;
}

int64_t vine_solver_trace(const vine_solver* s, float* rows, int64_t max_rows) {
  if (!s) return VINE_E_ARG;
  int64_t n = (int64_t)s->trace.size() / 10;
  if (!rows) return n;
  int64_t k = std::min(n, max_rows);
  std::memcpy(rows, s->trace.data(), sizeof(float) * 10 * k);
  return k;
}

vine_status vine_solver_query_stats(const vine_solver* s, int32_t nq, int64_t* expanded, int32_t* rounds) {
  if (!s || nq < 0 || (nq > 0 && (!expanded || !rounds))) return VINE_E_ARG;
  for (int32_t i = 0; i < nq; i++) {
    bool have = i < (int32_t)s->q_expanded.size();
    expanded[i] = have ? s->q_expanded[i] : 0;
    rounds[i] = have ? s->q_rounds[i] : 0;
  }
  return VINE_OK;
}

vine_status vine_solver_set_mlp(vine_solver* s, const float* params, int64_t nparams) {
  if (!s || !params || nparams != vine::mlp::NPARAM) {
    set_error("vine_solver_set_mlp: invalid arguments");
    return VINE_E_ARG;
  }
  for (vine_solver* p : s->parts) {
    vine_status r = vine_solver_set_mlp(p, params, nparams);
    if (r != VINE_OK) return r;
  }
  if (!s->parts.empty()) return VINE_OK;
  if (!s->d_mlp || s->mlp_nparams != nparams) {
    if (s->d_mlp) {
      auto it = std::find(s->allocs.begin(), s->allocs.end(), (void*)s->d_mlp);
      if (it != s->allocs.end()) s->allocs.erase(it);
      (void)hipFree(s->d_mlp);
      s->d_mlp = nullptr;
    }
    vine_status st = dev_alloc(s, (void**)&s->d_mlp, sizeof(float) * nparams);
    if (st != VINE_OK) return st;
    s->mlp_nparams = nparams;
  }
  HIP_TRY(hipMemcpy(s->d_mlp, params, sizeof(float) * nparams, hipMemcpyHostToDevice));
  s->ctx.mlp = s->d_mlp;
  return VINE_OK;
}

int64_t vine_solver_pareto(const vine_solver* s, int32_t q, int32_t v, float* out_pairs, int64_t cap) {
  if (!s || q < 0 || q >= s->B || v < 0 || v >= s->g->n || (cap > 0 && !out_pairs)) {
    set_error("vine_solver_pareto: invalid arguments");
    return VINE_E_ARG;
  }
  if (!s->parts.empty()) {
    const int k = (int)s->parts.size();
    return vine_solver_pareto(s->parts[q % k], q / k, v, out_pairs, cap);
  }
  int2 ps;
  hipError_t e = hipMemcpy(&ps, s->d_pset + (long long)q * s->g->n + s->g->perm[v], sizeof(int2),
                           hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    set_error("vine_solver_pareto: %s", hipGetErrorString(e));
    return VINE_E_HIP;
  }
  // the stored front holds exactly its live entries (sorted by d)
  int64_t slots = ps.y & ((1 << 26) - 1);
  std::vector<float2> tmp(slots);
  if (slots > 0) {
    e = hipMemcpy(tmp.data(), s->d_pool_cd + (uint32_t)ps.x, sizeof(float2) * slots, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      set_error("vine_solver_pareto: %s", hipGetErrorString(e));
      return VINE_E_HIP;
    }
  }
  int64_t live = 0;
  for (const float2& p : tmp) {
    if (std::isinf(p.x)) continue;
    if (live < cap) {
      out_pairs[2 * live] = p.x;
      out_pairs[2 * live + 1] = p.y;
    }
    live++;
  }
  return live;
}


// ---------------------------------------------------------------------------
// No-allocation entry point.  Every device buffer of the solve is cut from
// `scratch`, a file-scope device array of VINE_STATIC_SCRATCH_BYTES that each
// GPU holds from the time the library is loaded on it until the process ends.
// Host code never passes the array's name to the device: it keeps the address
// hipGetSymbolAddress returned for each device (scratch_ptr) and uses that.
}  // extern "C"

static __device__ __align__(256) unsigned char scratch[VINE_STATIC_SCRATCH_BYTES];

namespace {
constexpr int SCRATCH_DEVICES = 64;
std::mutex scratch_mu;
void* scratch_ptr[SCRATCH_DEVICES] = {};

// The scratch address on the current device, looked up once per device.
vine_status scratch_on_current_device(int* dev_out, char** out) {
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= SCRATCH_DEVICES) {
    set_error("vine_solve_device_static: device %d out of range", dev);
    return VINE_E_ARG;
  }
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
      return VINE_E_HIP;
    }
    scratch_ptr[dev] = p;
  }
  *dev_out = dev;
  *out = (char*)scratch_ptr[dev];
  return VINE_OK;
}
}  // namespace

extern "C" {

vine_status vine_solve_device_static(const vine_graph* g, const vine_options* opt, const float* mlp_params,
                                     int64_t mlp_nparams, int32_t nq, const int32_t* d_src, const int32_t* d_dst,
                                     const float* d_bound, float* d_cost, float* d_delay, int32_t* d_feasible,
                                     void* stream, vine_stats* out_stats) {
  if (!g || !opt || nq < 0 || (nq > 0 && (!d_src || !d_dst || !d_bound || !d_cost || !d_delay || !d_feasible))) {
    set_error("vine_solve_device_static: invalid arguments");
    return VINE_E_ARG;
  }
  vine_options o = *opt;
  vine_status v = validate_options(g, o);
  if (v != VINE_OK) return v;
  if (o.concurrency > 1 || o.use_graph || o.time_kernels || o.record_trace) {
    set_error("vine_solve_device_static: concurrency > 1, use_graph, time_kernels and record_trace need "
              "allocations and are not available here");
    return VINE_E_ARG;
  }
  if (o.delta_policy == VINE_DELTA_MLP && (!mlp_params || mlp_nparams != vine::mlp::NPARAM)) {
    set_error("vine_solve_device_static: VINE_DELTA_MLP needs the %d network parameters", vine::mlp::NPARAM);
    return VINE_E_ARG;
  }
  int dev = -1;
  char* base = nullptr;
  v = scratch_on_current_device(&dev, &base);
  if (v != VINE_OK) return v;
  if (dev != g->device) {
    set_error("vine_solve_device_static: the graph lives on device %d, the current device is %d", g->device, dev);
    return VINE_E_ARG;
  }
  const long long B = o.max_batch;
  // the largest label capacity whose buffers fit the scratch (the other
  // capacities follow it), by bisection over sizing passes that only add up
  // the buffers; room is left for the MLP parameters and query staging
  const double room = 0.98 * (double)VINE_STATIC_SCRATCH_BYTES - 64.0 * 1024;
  auto bytes_for = [&](long long lc, size_t* out) -> vine_status {
    vine_solver probe;
    init_fields(&probe, g, o);
    probe.count_only = true;
    Caps k;
    vine_status r = derive_caps(o, B, lc, &k);
    if (r == VINE_OK) r = alloc_buffers(&probe, k);
    *out = probe.arena_used;
    return r;
  };
  long long lab_cap = o.label_capacity;
  size_t need = 0;
  if (!lab_cap) {
    long long lo = B + 1024, hi = INT32_MAX - 1;
    v = bytes_for(lo, &need);
    if (v != VINE_OK) return v;
    if ((double)need > room) {
      set_error("vine_solve_device_static: %lld query slots x %d vertices need %zu bytes, more than the %zu-byte "
                "static scratch holds", B, g->n, need, (size_t)VINE_STATIC_SCRATCH_BYTES);
      return VINE_E_ARG;
    }
    while (lo < hi) {
      const long long mid = lo + (hi - lo + 1) / 2;
      v = bytes_for(mid, &need);
      if (v == VINE_OK && (double)need <= room) lo = mid;
      else hi = mid - 1;
    }
    lab_cap = lo;
  } else {
    v = bytes_for(lab_cap, &need);
    if (v != VINE_OK) return v;
    if ((double)need > room) {
      set_error("vine_solve_device_static: label_capacity %lld needs %zu bytes, more than the %zu-byte static "
                "scratch holds", lab_cap, need, (size_t)VINE_STATIC_SCRATCH_BYTES);
      return VINE_E_ARG;
    }
  }
  vine_solver s;
  init_fields(&s, g, o);
  s.arena = base;
  s.arena_cap = VINE_STATIC_SCRATCH_BYTES;
  s.stream = (hipStream_t)stream;
  Caps k;
  v = derive_caps(o, B, lab_cap, &k);
  if (v == VINE_OK) v = alloc_buffers(&s, k);
  if (v == VINE_OK && o.delta_policy == VINE_DELTA_MLP) v = vine_solver_set_mlp(&s, mlp_params, mlp_nparams);
  if (v == VINE_OK) v = vine_solve_device(&s, nq, d_src, d_dst, d_bound, d_cost, d_delay, d_feasible, out_stats);
  // nothing to free: the buffers belong to the scratch, the stream to the caller
  s.stream = nullptr;
  s.h_ct = nullptr;
  s.allocs.clear();
  return v;
}

}  // extern "C"
