// Generic label-correcting search for the reference's other label families
// (include/vine/families.hpp): hop-bounded, hop-bounded with a bottleneck
// arc length, 2-OSP, social (multiplicative) and network QoS, plus c_label
// itself.  Every kernel is a template over Label_t (its constraint type and
// answer order come from vine::family_traits), so a user-defined label with
// the same members and a family_traits specialisation drops in.
//
// The reference's search (RUN_CSPP, algorithm/cspp.cuh:69-325) runs one
// query at a time over an LDS hash and 128-entry per-vertex lists guarded by
// spin locks (algorithm/dominance.cuh:15-153).  Here many queries share
// every round and no lock is taken:
//   k_fam_expand  one thread per frontier label: extend along every out-arc,
//                 check the constraint, drop the son if an entry of its
//                 (query, vertex) Pareto list dominates it (the list is
//                 read-only during this kernel), else log it and push it on
//                 the vertex's candidate chain with one atomic exchange; the
//                 first push of a round records the (query, vertex) pair;
//   k_fam_merge   one thread per recorded pair: merge the chain into the
//                 Pareto list in order (drop a dominated candidate; unlink and
//                 mark dead the entries a candidate dominates), survivors form
//                 the next frontier.  Only this thread touches the pair's list.
// Labels marked dead after they joined a frontier are skipped when expanded.
// With dominance that is preserved by expand and by the constraint (all
// families here but social with delay_L < 1), the final Pareto set of every
// vertex is the set of non-dominated label values over all walks, whatever
// the processing order, so the answer is exactly the oracle's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "device_graph.hpp"
#include "errors.hpp"
#include "vine.h"
#include "vine/families.hpp"

using vine::set_error;

namespace vine {
namespace bounds {
// Distance to t for the bound searches of the pruned search: the generic
// search itself, run from t on the transposed graph with a one-criterion
// label, leaves at every vertex the single label of its shortest distance.
//   W = 0 length, 1 delay, 2 arcs;
//   W = 3 length over the arcs whose delay is at least the query's limit
//         (QoS: a feasible path uses only those arcs, so this is the length
//         still needed, not just a lower bound of it)
template <int W>
struct dist_label {
  int vid;
  float d;
  float gate;  // W = 3: delay of the arc that reached vid, tested by the constraint
  __host__ __device__ dist_label() : vid(-1), d(0.f), gate(INFINITY) {}
  __host__ __device__ explicit dist_label(int v) : vid(v), d(0.f), gate(INFINITY) {}
  template <typename Edge_t>
  __host__ __device__ void expand(int next_v, const Edge_t& e, dist_label* son) const {
    son->vid = next_v;
    son->d = d + (W == 0 || W == 3 ? e.len : (W == 1 ? e.delay : 1.f));
    son->gate = e.delay;
  }
  __host__ __device__ static bool dominate(const dist_label& a, const dist_label& b) { return a.d <= b.d; }
};
struct no_constraint {
  template <typename L>
  __host__ __device__ bool pass_constraints_check(const L&) const { return true; }
};
// the QoS arc filter: an arc whose delay is below the limit is never used
struct min_gate {
  float L;
  template <typename Lab>
  __host__ __device__ bool pass_constraints_check(const Lab& l) const { return !(l.gate < L); }
};
template <int W>
struct bound_constraint {
  using type = no_constraint;
  __host__ __device__ static type make(const float*) { return type(); }
};
// delay to t: vertices beyond the query's delay limit are pruned anyway
struct max_d {
  float L;
  template <typename Lab>
  __host__ __device__ bool pass_constraints_check(const Lab& l) const { return !(l.d > L); }
};
template <>
struct bound_constraint<1> {
  using type = max_d;
  __host__ __device__ static type make(const float* lim) { return type{lim[0]}; }
};
template <>
struct bound_constraint<3> {
  using type = min_gate;
  __host__ __device__ static type make(const float* lim) { return type{lim[0]}; }
};
}  // namespace bounds

template <int W>
struct family_traits<bounds::dist_label<W>> {
  using constraint = typename bounds::bound_constraint<W>::type;
  __host__ __device__ static void fields(const bounds::dist_label<W>& l, float* f) {
    f[0] = l.d;
    f[1] = f[2] = f[3] = 0.f;
  }
  __host__ __device__ static bool better(const bounds::dist_label<W>& a, const bounds::dist_label<W>& b) {
    return a.d < b.d;
  }
  __host__ __device__ static constraint make(const float* lim) { return bounds::bound_constraint<W>::make(lim); }
};
}  // namespace vine

namespace {

constexpr int kBlock = 256;

struct FamCounters {
  int32_t label_ctr;  // next free label id
  int32_t fn[2];      // frontier sizes (current, next)
  int32_t touched_n[2];  // (query, vertex) pairs with candidates, per round parity
  int32_t overflow;   // 1 = label capacity exceeded, 2 = path walk too long
  int32_t pad;
  long long expanded, relaxed, inserted, rounds;
};

struct FamCtx {
  int32_t n;
  int32_t B;    // query slots
  int32_t cap;  // label capacity
  const int32_t* row;
  const int4* edges;  // {dst, len bits, delay bits, 0}
  void* lab;          // Label_t[cap]
  int32_t* parent;    // [cap]
  int32_t* lq;        // [cap] batch slot
  int32_t* nxt;       // [cap] next entry of the Pareto list
  int32_t* cnext;     // [cap] next candidate of the chain
  uint8_t* dead;      // [cap]
  int32_t* head;      // [B*n] Pareto list heads
  int32_t* chead;     // [B*n] candidate chain heads
  int32_t* front[2];  // [cap] frontier label ids
  long long* touched; // [cap] (slot * n + vertex)
  const float* lim;   // [B*3] per-slot limits
  FamCounters* ct;
  // target pruning (prune = 1): the best label at t so far, per slot
  int32_t prune;
  const int32_t* qdst;  // [B] target (internal id)
  void* best;           // Label_t[B]
  int32_t* has_best;    // [B]
  // distance bounds to t per (slot, vertex), prune = 1 only (null: unused):
  // length (answer orders led by min d), delay (additive delay), arcs (hop limit)
  const float* blen;
  const float* bdelay;
  const float* bhop;
  // key-ordered windows (window > 0, families ordered by min d first): a
  // round expands the labels whose key d + blen is within `window` of the
  // query's smallest key and carries the rest to the next round
  float window;
  uint32_t* qkey;  // [2][B] smallest key of the query's frontier (float bits), per round parity
  // bound searches only: per (slot, vertex) delay to t; a vertex farther than
  // the query's delay limit is not entered (null: no gate)
  const float* vgate;
};

// key of label L of slot q: length plus the length bound to t (if computed)
template <typename Label_t>
__device__ __forceinline__ float fam_key(const FamCtx& c, const Label_t& L, int q) {
  float f[4];
  vine::family_traits<Label_t>::fields(L, f);
  return f[0] + (c.blen ? c.blen[(long long)q * c.n + L.vid] : 0.f);
}

// Sums a and b over the workgroup and adds them to *da / *db with one atomic
// each; every thread of the workgroup must call it.
__device__ void block_add2(long long a, long long b, long long* da, long long* db) {
  __shared__ long long sa[kBlock / 64], sb[kBlock / 64];
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o);
    b += __shfl_xor(b, o);
  }
  if ((threadIdx.x & 63) == 0) {
    sa[threadIdx.x >> 6] = a;
    sb[threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    long long x = 0, y = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
      x += sa[w];
      y += sb[w];
    }
    if (x) atomicAdd((unsigned long long*)da, (unsigned long long)x);
    if (y && db) atomicAdd((unsigned long long*)db, (unsigned long long)y);
  }
}

__global__ void k_fill(int32_t* p, long long count, int32_t v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

template <typename Label_t>
__global__ void k_fam_init(FamCtx c, int nb, const int32_t* src) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b == 0) {
    c.ct->label_ctr = nb;
    c.ct->fn[0] = nb;
    c.ct->fn[1] = 0;
    c.ct->touched_n[0] = c.ct->touched_n[1] = 0;
    c.ct->overflow = 0;
    c.ct->expanded = c.ct->relaxed = c.ct->inserted = c.ct->rounds = 0;
  }
  if (b >= nb) return;
  Label_t* lab = (Label_t*)c.lab;
  lab[b] = Label_t(src[b]);  // the source sits in its own Pareto set (algorithm/cspp.cuh:63-67)
  c.parent[b] = -1;
  c.lq[b] = b;
  c.nxt[b] = -1;
  c.dead[b] = 0;
  c.head[(long long)b * c.n + src[b]] = b;
  c.front[0][b] = b;
  // s == t: the source is the answer already
  c.has_best[b] = src[b] == c.qdst[b];
  if (src[b] == c.qdst[b]) ((Label_t*)c.best)[b] = lab[b];
  if (c.window > 0.f) {
    c.qkey[b] = __float_as_uint(fam_key(c, lab[b], b));
    c.qkey[c.B + b] = __float_as_uint(INFINITY);
  }
}

template <typename Label_t>
__global__ __launch_bounds__(kBlock) void k_fam_expand(FamCtx c, int cur, int first) {
  using Traits = vine::family_traits<Label_t>;
  Label_t* lab = (Label_t*)c.lab;
  const int F = c.ct->fn[cur];
  // rounds counts the rounds that left a frontier behind
  if (!first && F > 0 && blockIdx.x == 0 && threadIdx.x == 0) c.ct->rounds++;
  long long n_exp = 0, n_rel = 0;
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < F; f += gridDim.x * blockDim.x) {
    const int i = c.front[cur][f];
    if (c.dead[i]) continue;
    const Label_t L = lab[i];
    const int q = c.lq[i];
    if (c.window > 0.f) {
      const float k = fam_key(c, L, q);
      if (!(k < INFINITY)) continue;  // t is unreachable from here (length bound +inf)
      if (c.prune && c.has_best[q]) {
        float bf[4];
        Traits::fields(((const Label_t*)c.best)[q], bf);
        if (k > bf[0]) continue;  // no completion can beat t's best length
      }
      if (k > __uint_as_float(c.qkey[cur * c.B + q]) + c.window) {
        // outside this round's window: carried to the next frontier
        c.front[cur ^ 1][atomicAdd(&c.ct->fn[cur ^ 1], 1)] = i;
        atomicMin(&c.qkey[(cur ^ 1) * c.B + q], __float_as_uint(k));
        continue;
      }
    }
    const typename Traits::constraint C = Traits::make(c.lim + 3 * q);
    n_exp++;
    for (int p = c.row[L.vid]; p < c.row[L.vid + 1]; p++) {
      n_rel++;
      const int4 e = c.edges[p];
      Label_t son;
      L.expand(e.x, vine::c_edge(__int_as_float(e.y), __int_as_float(e.z)), &son);
      if (!C.pass_constraints_check(son)) continue;  // the reference breaks here (algorithm/bucket.cuh:76-78)
      if (c.vgate && !(c.vgate[(long long)q * c.n + son.vid] <= c.lim[3 * q])) continue;
      // target pruning: extending never makes a label better in the answer
      // order (every family's expand is monotone), so a son that is not
      // better than t's best so far cannot lead to the answer
      if (c.prune && c.has_best[q] && !Traits::better(son, ((const Label_t*)c.best)[q])) continue;
      const long long key = (long long)q * c.n + son.vid;
      if (c.prune && (c.blen || c.bdelay || c.bhop)) {
        // no completion through son.vid can be feasible (delay, arcs) or beat
        // t's best length; integer weights keep these sums exact
        float f[4];
        Traits::fields(son, f);
        const float* lv = c.lim + 3 * q;
        if (c.blen && !(c.blen[key] < INFINITY)) continue;  // t is unreachable from son.vid
        if (c.bdelay && f[1] + c.bdelay[key] > lv[0]) continue;
        if (c.bhop && f[2] + c.bhop[key] > (float)(int)lv[1]) continue;
        if (c.blen && c.has_best[q]) {
          float bf[4];
          Traits::fields(((const Label_t*)c.best)[q], bf);
          if (f[0] + c.blen[key] > bf[0]) continue;
        }
      }
      bool dominated = false;
      for (int x = c.head[key]; x >= 0; x = c.nxt[x])
        if (Label_t::dominate(lab[x], son)) {
          dominated = true;
          break;
        }
      if (dominated) continue;
      const int id = atomicAdd(&c.ct->label_ctr, 1);
      if (id < 0 || id >= c.cap) {  // the counter keeps counting (and may wrap) once full
        atomicOr(&c.ct->overflow, 1);
        continue;
      }
      lab[id] = son;
      c.parent[id] = i;
      c.lq[id] = q;
      c.nxt[id] = -1;
      c.dead[id] = 0;
      const int old = atomicExch(&c.chead[key], id);
      c.cnext[id] = old;
      if (old < 0) c.touched[atomicAdd(&c.ct->touched_n[cur], 1)] = key;
    }
  }
  block_add2(n_exp, n_rel, &c.ct->expanded, &c.ct->relaxed);
}

template <typename Label_t>
__global__ __launch_bounds__(kBlock) void k_fam_merge(FamCtx c, int cur) {
  using Traits = vine::family_traits<Label_t>;
  Label_t* lab = (Label_t*)c.lab;
  const int T = c.ct->touched_n[cur];
  const int nx = cur ^ 1;
  long long n_ins = 0;
  // the frontier expand consumed and the pair count of the round after this
  // one: read by no kernel until the next expand
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    c.ct->fn[cur] = 0;
    c.ct->touched_n[nx] = 0;
  }
  if (c.window > 0.f)  // this round's keys are consumed: start the round after next from +inf
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < c.B; b += gridDim.x * blockDim.x)
      c.qkey[cur * c.B + b] = __float_as_uint(INFINITY);
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < T; t += gridDim.x * blockDim.x) {
    const long long key = c.touched[t];
    int cand = c.chead[key];
    c.chead[key] = -1;
    while (cand >= 0) {
      const int next_cand = c.cnext[cand];
      const Label_t L = lab[cand];
      // one pass: a dominated candidate cannot dominate another entry (the
      // list is mutually non-dominated), so rejection and removal exclude
      // each other
      bool rejected = false;
      int prev = -1;
      for (int x = c.head[key]; x >= 0;) {
        const int xn = c.nxt[x];
        const Label_t E = lab[x];
        if (Label_t::dominate(E, L)) {
          rejected = true;
          break;
        }
        if (Label_t::dominate(L, E)) {
          c.dead[x] = 1;
          if (prev < 0)
            c.head[key] = xn;
          else
            c.nxt[prev] = xn;
        } else {
          prev = x;
        }
        x = xn;
      }
      if (rejected) {
        c.dead[cand] = 1;
      } else {
        c.nxt[cand] = c.head[key];
        c.head[key] = cand;
        c.front[nx][atomicAdd(&c.ct->fn[nx], 1)] = cand;
        if (c.window > 0.f) atomicMin(&c.qkey[nx * c.B + c.lq[cand]], __float_as_uint(fam_key(c, L, c.lq[cand])));
        n_ins++;
      }
      cand = next_cand;
    }
    if (c.prune) {
      const int q = (int)(key / c.n);
      if (key - (long long)q * c.n == c.qdst[q]) {
        // only this thread writes slot q's best in this kernel
        Label_t* best = (Label_t*)c.best;
        for (int x = c.head[key]; x >= 0; x = c.nxt[x])
          if (!c.has_best[q] || Traits::better(lab[x], best[q])) {
            best[q] = lab[x];
            c.has_best[q] = 1;
          }
      }
    }
  }
  block_add2(n_ins, 0, &c.ct->inserted, nullptr);
}

// Best label of t's list per slot; path length by the parent chain.
template <typename Label_t>
__global__ void k_fam_answer(FamCtx c, int nb, const int32_t* dst, float* out_f, int32_t* out_feas,
                             int32_t* out_best, int32_t* out_hops) {
  using Traits = vine::family_traits<Label_t>;
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const Label_t* lab = (const Label_t*)c.lab;
  int best = -1;
  for (int x = c.head[(long long)b * c.n + dst[b]]; x >= 0; x = c.nxt[x])
    if (best < 0 || Traits::better(lab[x], lab[best])) best = x;
  float f[4] = {INFINITY, INFINITY, INFINITY, INFINITY};
  int hops = 0;
  if (best >= 0) {
    Traits::fields(lab[best], f);
    for (int x = best; x >= 0; x = c.parent[x]) {
      if (++hops > c.n) {  // a simple path has at most n vertices
        atomicOr(&c.ct->overflow, 2);
        break;
      }
    }
  }
  for (int k = 0; k < 4; k++) out_f[4 * b + k] = f[k];
  out_feas[b] = best >= 0;
  out_best[b] = best;
  out_hops[b] = hops;
}

template <typename Label_t>
__global__ void k_fam_paths(FamCtx c, int nb, const int32_t* best, const long long* off, const int32_t* inv,
                            int32_t* out) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb || best[b] < 0) return;
  const Label_t* lab = (const Label_t*)c.lab;
  long long k = off[b + 1] - 1;
  for (int x = best[b]; x >= 0 && k >= off[b]; x = c.parent[x]) out[k--] = inv[lab[x].vid];
}

template <typename Label_t>
__global__ void k_fam_front(FamCtx c, long long key, float* out, int cap, int32_t* count) {
  using Traits = vine::family_traits<Label_t>;
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const Label_t* lab = (const Label_t*)c.lab;
  int k = 0;
  for (int x = c.head[key]; x >= 0; x = c.nxt[x], k++)
    if (k < cap) Traits::fields(lab[x], out + 4 * k);
  *count = k;
}

// bound[slot * n + v] = the distance label's value at v (+inf where none)
template <typename Label_t>
__global__ void k_fam_dist(FamCtx c, long long cells, float* bound) {
  const Label_t* lab = (const Label_t*)c.lab;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < cells; i += (long long)gridDim.x * blockDim.x) {
    const int x = c.head[i];
    bound[i] = x >= 0 ? lab[x].d : INFINITY;
  }
}

__global__ void k_map(int nb, const int32_t* perm, const int32_t* in, int32_t* out) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb) out[b] = perm[in[b]];
}

// the bound searches run in the family's label arena
static_assert(sizeof(vine::bounds::dist_label<3>) <= sizeof(vine::c_label) &&
                  sizeof(vine::bounds::dist_label<3>) <= sizeof(vine::qos::c_label) &&
                  sizeof(vine::bounds::dist_label<3>) <= sizeof(vine::hop::c_label) &&
                  sizeof(vine::bounds::dist_label<3>) <= sizeof(vine::hop_maxcost::c_label),
              "bound labels must fit the family's label slots");

size_t label_size(int family) {
  switch (family) {
    case VINE_FAMILY_HOP: return sizeof(vine::hop::c_label);
    case VINE_FAMILY_HOP_MAXCOST: return sizeof(vine::hop_maxcost::c_label);
    case VINE_FAMILY_SOCIAL: return sizeof(vine::social::c_label);
    case VINE_FAMILY_QOS: return sizeof(vine::qos::c_label);
    default: return sizeof(vine::c_label);
  }
}

}  // namespace

struct vine_family_solver {
  const vine_graph* g = nullptr;
  int family = 0;
  int B = 1;
  int32_t cap = 0;
  hipStream_t stream = nullptr;
  std::vector<void*> allocs;
  FamCtx c{};
  int32_t *d_src = nullptr, *d_dst = nullptr, *d_isrc = nullptr, *d_idst = nullptr, *d_has_best = nullptr;
  void* d_bestlab = nullptr;
  float *d_blen = nullptr, *d_bdelay = nullptr, *d_bhop = nullptr;
  uint32_t* d_qkey = nullptr;
  int32_t *d_feas = nullptr, *d_best = nullptr, *d_hops = nullptr, *d_count = nullptr;
  float *d_out = nullptr, *d_lim = nullptr, *d_front = nullptr;
  long long* d_off = nullptr;
  int32_t* d_path = nullptr;
  long long path_cap = 0;
  int last_nb = 0;
  std::vector<int32_t> last_slot_query;  // slot -> query index of the last batch
};

namespace {

vine_status fam_alloc(vine_family_solver* s, void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 16));
  if (e != hipSuccess) {
    *p = nullptr;
    set_error("vine_family_solver: hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
    return VINE_E_OOM;
  }
  s->allocs.push_back(*p);
  return VINE_OK;
}

#define FAM_TRY(expr)                                                                            \
  do {                                                                                           \
    hipError_t e_ = (expr);                                                                      \
    if (e_ != hipSuccess) {                                                                      \
      set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__);      \
      return VINE_E_HIP;                                                                         \
    }                                                                                            \
  } while (0)

int grid_for(long long work) { return (int)std::max<long long>(1, std::min<long long>((work + kBlock - 1) / kBlock, 4096)); }

// One batch: nb queries already in d_src/d_dst (caller ids) and d_lim.
// Rounds of the generic search for nb slots from the sources `from`
// (internal ids), on the graph c.row / c.edges, until the frontier is empty.
template <typename Label_t>
vine_status run_search(vine_family_solver* s, FamCtx& c, int nb, const int32_t* from) {
  const int n = s->g->n;
  hipStream_t st = s->stream;
  const long long cells = (long long)nb * n;
  hipLaunchKernelGGL(k_fill, dim3(grid_for(cells)), dim3(kBlock), 0, st, c.head, cells, -1);
  hipLaunchKernelGGL(k_fill, dim3(grid_for(cells)), dim3(kBlock), 0, st, c.chead, cells, -1);
  hipLaunchKernelGGL(k_fam_init<Label_t>, dim3(grid_for(nb)), dim3(kBlock), 0, st, c, nb, from);
  FAM_TRY(hipGetLastError());
  // rounds until the frontier is empty; the host looks every `sync` rounds
  int cur = 0;
  const int sync = 8;
  // workgroups per round kernel: sized to the frontier the host last read
  // (up to 16x growth until the next read; the kernels loop over any size),
  // so that the many small rounds of a windowed search launch small grids
  int grid = 2048;
  // every family here converges (dominance discards cycles); the cap only
  // turns a runaway search into an error
  const long long max_rounds = 64LL * n + 1024;
  for (long long r = 0;; r++) {
    if (r > max_rounds) {
      set_error("vine_family_solve: no convergence after %lld rounds", r);
      return VINE_E_HIP;
    }
    hipLaunchKernelGGL(k_fam_expand<Label_t>, dim3(grid), dim3(kBlock), 0, st, c, cur, r == 0 ? 1 : 0);
    hipLaunchKernelGGL(k_fam_merge<Label_t>, dim3(grid), dim3(kBlock), 0, st, c, cur);
    cur ^= 1;
    if ((r + 1) % sync == 0) {
      FAM_TRY(hipGetLastError());
      FamCounters h;
      FAM_TRY(hipMemcpyAsync(&h, c.ct, sizeof(h), hipMemcpyDeviceToHost, st));
      FAM_TRY(hipStreamSynchronize(st));
      if (h.overflow & 1) {
        set_error("vine_family_solve: label capacity (%d) exceeded; raise label_capacity", s->cap);
        return VINE_E_OVERFLOW;
      }
      if (h.fn[cur] == 0) break;
      grid = (int)std::max<long long>(16, std::min<long long>(2048, (16LL * h.fn[cur] + kBlock - 1) / kBlock));
    }
  }
  return VINE_OK;
}

constexpr float BOUND_WINDOW = 2.f;

// Distances to t for the pruned search: the generic search from each slot's
// target on the transposed graph, then one value per (slot, vertex).
template <int W>
vine_status run_bound(vine_family_solver* s, int nb, float* out) {
  using L = vine::bounds::dist_label<W>;
  FamCtx c = s->c;
  c.row = s->g->d_rrow;
  c.edges = s->g->d_redges;
  c.prune = 0;
  // length and delay: windows of BOUND_WINDOW mean arcs order the search like
  // delta-stepping (fewer re-relaxations; every label is still expanded, so
  // the distances are exact).  Arc counts settle in BFS order without one.
  c.window = (W == 0 || W == 3) ? BOUND_WINDOW * s->g->mean_len : (W == 1 ? BOUND_WINDOW * s->g->mean_delay : 0.f);
  c.blen = c.bdelay = c.bhop = nullptr;
  // length and arc counts over the vertices within the query's delay limit of
  // t (the delay search runs first): every feasible completion stays there,
  // so the restricted distances are still lower bounds, and tighter ones
  c.vgate = ((W == 0 || W == 2) && s->c.bdelay) ? s->d_bdelay : nullptr;
  vine_status r = run_search<L>(s, c, nb, s->d_idst);
  if (r != VINE_OK) return r;
  const long long cells = (long long)nb * s->g->n;
  hipLaunchKernelGGL(k_fam_dist<L>, dim3(grid_for(cells)), dim3(kBlock), 0, s->stream, c, cells, out);
  FAM_TRY(hipGetLastError());
  return VINE_OK;
}

template <typename Label_t>
vine_status run_batch(vine_family_solver* s, int nb, int64_t* counters) {
  FamCtx& c = s->c;
  hipStream_t st = s->stream;
  hipLaunchKernelGGL(k_map, dim3(grid_for(nb)), dim3(kBlock), 0, st, nb, s->g->d_perm, s->d_src, s->d_isrc);
  hipLaunchKernelGGL(k_map, dim3(grid_for(nb)), dim3(kBlock), 0, st, nb, s->g->d_perm, s->d_dst, s->d_idst);
  FAM_TRY(hipGetLastError());
  if (c.bdelay) {  // first: it gates the length and arc-count searches
    vine_status r = run_bound<1>(s, nb, s->d_bdelay);
    if (r != VINE_OK) return r;
  }
  if (c.blen) {
    // QoS: length over the arcs the query may use
    vine_status r = s->family == VINE_FAMILY_QOS ? run_bound<3>(s, nb, s->d_blen) : run_bound<0>(s, nb, s->d_blen);
    if (r != VINE_OK) return r;
  }
  if (c.bhop) {
    vine_status r = run_bound<2>(s, nb, s->d_bhop);
    if (r != VINE_OK) return r;
  }
  vine_status r = run_search<Label_t>(s, c, nb, s->d_isrc);
  if (r != VINE_OK) return r;
  hipLaunchKernelGGL(k_fam_answer<Label_t>, dim3(grid_for(nb)), dim3(kBlock), 0, st, c, nb, s->d_idst, s->d_out,
                     s->d_feas, s->d_best, s->d_hops);
  FAM_TRY(hipGetLastError());
  FamCounters h;
  FAM_TRY(hipMemcpyAsync(&h, c.ct, sizeof(h), hipMemcpyDeviceToHost, st));
  FAM_TRY(hipStreamSynchronize(st));
  if (h.overflow & 2) {
    set_error("vine_family_solve: a parent chain is longer than the vertex count");
    return VINE_E_HIP;
  }
  if (counters) {
    counters[0] += h.rounds;
    counters[1] += h.expanded;
    counters[2] += h.relaxed;
    counters[3] += h.inserted;
  }
  return VINE_OK;
}

template <typename Label_t>
vine_status run_paths(vine_family_solver* s, int nb, std::vector<long long>& off_host) {
  // offsets from the hop counts, then the walk
  std::vector<int32_t> hops(nb);
  FAM_TRY(hipMemcpyAsync(hops.data(), s->d_hops, sizeof(int32_t) * nb, hipMemcpyDeviceToHost, s->stream));
  FAM_TRY(hipStreamSynchronize(s->stream));
  off_host.assign(nb + 1, 0);
  for (int b = 0; b < nb; b++) off_host[b + 1] = off_host[b] + hops[b];
  const long long total = off_host[nb];
  if (total > s->path_cap) {
    if (s->d_path) {
      FAM_TRY(hipStreamSynchronize(s->stream));
      (void)hipFree(s->d_path);
      s->allocs.erase(std::find(s->allocs.begin(), s->allocs.end(), (void*)s->d_path));
      s->d_path = nullptr;
    }
    vine_status r = fam_alloc(s, (void**)&s->d_path, sizeof(int32_t) * total);
    if (r != VINE_OK) return r;
    s->path_cap = total;
  }
  FAM_TRY(hipMemcpyAsync(s->d_off, off_host.data(), sizeof(long long) * (nb + 1), hipMemcpyHostToDevice, s->stream));
  hipLaunchKernelGGL(k_fam_paths<Label_t>, dim3(grid_for(nb)), dim3(kBlock), 0, s->stream, s->c, nb, s->d_best,
                     s->d_off, s->g->d_inv, s->d_path);
  FAM_TRY(hipGetLastError());
  return VINE_OK;
}

template <typename Label_t>
vine_status solve_family(vine_family_solver* s, int32_t nq, const int32_t* src, const int32_t* dst,
                         const float* limits, float* out_label, int32_t* out_feasible, int64_t* out_path_off,
                         int32_t* out_path, int64_t path_cap, int64_t* out_counters) {
  int64_t counters[4] = {0, 0, 0, 0};
  int64_t path_pos = 0;
  bool path_overflow = false;
  if (out_path_off) out_path_off[0] = 0;
  for (int32_t q0 = 0; q0 < nq; q0 += s->B) {
    const int nb = std::min<int32_t>(s->B, nq - q0);
    FAM_TRY(hipMemcpyAsync(s->d_src, src + q0, sizeof(int32_t) * nb, hipMemcpyHostToDevice, s->stream));
    FAM_TRY(hipMemcpyAsync(s->d_dst, dst + q0, sizeof(int32_t) * nb, hipMemcpyHostToDevice, s->stream));
    FAM_TRY(hipMemcpyAsync(s->d_lim, limits + 3 * (size_t)q0, sizeof(float) * 3 * nb, hipMemcpyHostToDevice,
                           s->stream));
    vine_status r = run_batch<Label_t>(s, nb, counters);
    if (r != VINE_OK) return r;
    std::vector<int32_t> feas(nb);
    FAM_TRY(hipMemcpyAsync(out_label + 4 * (size_t)q0, s->d_out, sizeof(float) * 4 * nb, hipMemcpyDeviceToHost,
                           s->stream));
    FAM_TRY(hipMemcpyAsync(feas.data(), s->d_feas, sizeof(int32_t) * nb, hipMemcpyDeviceToHost, s->stream));
    FAM_TRY(hipStreamSynchronize(s->stream));
    if (out_feasible) std::copy(feas.begin(), feas.end(), out_feasible + q0);
    if (out_path_off) {
      std::vector<long long> off;
      if (out_path) {
        r = run_paths<Label_t>(s, nb, off);
        if (r != VINE_OK) return r;
      } else {
        std::vector<int32_t> hops(nb);
        FAM_TRY(hipMemcpyAsync(hops.data(), s->d_hops, sizeof(int32_t) * nb, hipMemcpyDeviceToHost, s->stream));
        FAM_TRY(hipStreamSynchronize(s->stream));
        off.assign(nb + 1, 0);
        for (int b = 0; b < nb; b++) off[b + 1] = off[b] + hops[b];
      }
      for (int b = 0; b < nb; b++) out_path_off[q0 + b + 1] = path_pos + off[b + 1];
      if (out_path) {
        if (path_pos + off[nb] <= path_cap) {
          if (off[nb] > 0)
            FAM_TRY(hipMemcpyAsync(out_path + path_pos, s->d_path, sizeof(int32_t) * off[nb], hipMemcpyDeviceToHost,
                                   s->stream));
          FAM_TRY(hipStreamSynchronize(s->stream));
        } else {
          path_overflow = true;
        }
      }
      path_pos += off[nb];
    }
    s->last_nb = nb;
    s->last_slot_query.assign(nb, 0);
    for (int b = 0; b < nb; b++) s->last_slot_query[b] = q0 + b;
  }
  if (out_counters)
    for (int k = 0; k < 4; k++) out_counters[k] = counters[k];
  if (path_overflow) {
    set_error("vine_family_solve: paths need %lld entries, path_cap is %lld", (long long)path_pos,
              (long long)path_cap);
    return VINE_E_OVERFLOW;
  }
  return VINE_OK;
}

template <typename Label_t>
int64_t front_family(const vine_family_solver* s, int32_t slot, int32_t v, float* out, int64_t cap) {
  const long long key = (long long)slot * s->g->n + s->g->perm[v];
  hipLaunchKernelGGL(k_fam_front<Label_t>, dim3(1), dim3(64), 0, s->stream, s->c, key, s->d_front,
                     (int)std::min<int64_t>(cap, 4096), s->d_count);
  int32_t count = 0;
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(&count, s->d_count, sizeof(int32_t), hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
      hipStreamSynchronize(s->stream) != hipSuccess) {
    set_error("vine_family_front: HIP error");
    return VINE_E_HIP;
  }
  const int64_t k = std::min<int64_t>(std::min<int64_t>(count, cap), 4096);
  if (k > 0 && (hipMemcpy(out, s->d_front, sizeof(float) * 4 * k, hipMemcpyDeviceToHost) != hipSuccess)) {
    set_error("vine_family_front: HIP error");
    return VINE_E_HIP;
  }
  return count;
}

}  // namespace

extern "C" {

vine_status vine_family_solver_create(const vine_graph* g, int32_t family, int32_t max_batch, int64_t label_capacity,
                                      int32_t prune, float window, vine_family_solver** out) {
  if (!g || !out || family < VINE_FAMILY_CSP || family > VINE_FAMILY_QOS || max_batch < 1 || prune < 0 ||
      prune > 1 || !(window >= 0.f)) {
    set_error("vine_family_solver_create: invalid arguments");
    return VINE_E_ARG;
  }
  if (label_capacity <= 0) label_capacity = 1ll << 25;
  if (label_capacity < max_batch || label_capacity > (1ll << 30)) {
    set_error("vine_family_solver_create: label_capacity must lie in [max_batch, 2^30]");
    return VINE_E_ARG;
  }
  if (family == VINE_FAMILY_SOCIAL && (g->max_len > 1.f || g->max_delay > 1.f)) {
    set_error("vine_family_solver_create: the social family multiplies weights, which must lie in [0, 1] "
              "(larger weights grow without bound around cycles)");
    return VINE_E_ARG;
  }
  auto* s = new vine_family_solver();
  s->g = g;
  s->family = family;
  s->B = max_batch;
  s->cap = (int32_t)label_capacity;
  vine_status r = VINE_OK;
  if (hipSetDevice(g->device) != hipSuccess || hipStreamCreate(&s->stream) != hipSuccess) {
    set_error("vine_family_solver_create: cannot create a stream");
    delete s;
    return VINE_E_HIP;
  }
  FamCtx& c = s->c;
  c.n = g->n;
  c.B = s->B;
  c.cap = s->cap;
  c.row = g->d_row;
  c.edges = g->d_edges;
  const size_t cap = (size_t)s->cap, cells = (size_t)s->B * g->n, B = (size_t)s->B;
  void** slots[] = {&c.lab, (void**)&c.parent, (void**)&c.lq, (void**)&c.nxt, (void**)&c.cnext, (void**)&c.dead,
                    (void**)&c.head, (void**)&c.chead, (void**)&c.front[0], (void**)&c.front[1], (void**)&c.touched,
                    (void**)&s->d_lim, (void**)&c.ct, (void**)&s->d_src, (void**)&s->d_dst, (void**)&s->d_isrc,
                    (void**)&s->d_idst, (void**)&s->d_feas, (void**)&s->d_best, (void**)&s->d_hops,
                    (void**)&s->d_count, (void**)&s->d_out, (void**)&s->d_front, (void**)&s->d_off,
                    (void**)&s->d_bestlab, (void**)&s->d_has_best, (void**)&s->d_qkey};
  const size_t bytes[] = {cap * label_size(family), cap * 4, cap * 4, cap * 4, cap * 4, cap, cells * 4, cells * 4,
                          cap * 4, cap * 4, cap * 8, B * 12, sizeof(FamCounters), B * 4, B * 4, B * 4,
                          B * 4, B * 4, B * 4, B * 4, 4, B * 16, 4096 * 16, (B + 1) * 8, B * label_size(family),
                          B * 4, B * 8};
  for (size_t k = 0; k < sizeof(bytes) / sizeof(bytes[0]) && r == VINE_OK; k++) r = fam_alloc(s, slots[k], bytes[k]);
  c.lim = s->d_lim;
  c.prune = prune;
  c.qdst = s->d_idst;
  c.best = s->d_bestlab;
  c.has_best = s->d_has_best;
  c.qkey = s->d_qkey;
  c.vgate = nullptr;
  // windows need an order led by min d (all families but social)
  c.window = family == VINE_FAMILY_SOCIAL ? 0.f : window;
  // distance bounds for the pruned search, where the family's order and
  // constraints are additive in them and sums of integer weights are exact
  if (r == VINE_OK && prune && g->integer_weights && family != VINE_FAMILY_SOCIAL) {
    const bool additive_delay = family != VINE_FAMILY_QOS;
    const bool hops = family == VINE_FAMILY_HOP || family == VINE_FAMILY_HOP_MAXCOST;
    r = fam_alloc(s, (void**)&s->d_blen, cells * 4);
    if (r == VINE_OK && additive_delay) r = fam_alloc(s, (void**)&s->d_bdelay, cells * 4);
    if (r == VINE_OK && hops) r = fam_alloc(s, (void**)&s->d_bhop, cells * 4);
    c.blen = s->d_blen;
    c.bdelay = s->d_bdelay;
    c.bhop = s->d_bhop;
  }
  if (r != VINE_OK) {
    vine_family_solver_destroy(s);
    return r;
  }
  *out = s;
  return VINE_OK;
}

vine_status vine_family_solver_destroy(vine_family_solver* s) {
  if (!s) return VINE_OK;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (void* p : s->allocs) (void)hipFree(p);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return VINE_OK;
}

vine_status vine_family_solve(vine_family_solver* s, int32_t nq, const int32_t* src, const int32_t* dst,
                              const float* limits, float* out_label, int32_t* out_feasible, int64_t* out_path_off,
                              int32_t* out_path, int64_t path_cap, int64_t* out_counters) {
  if (!s || nq < 0 || (nq > 0 && (!src || !dst || !limits || !out_label)) || (out_path && !out_path_off) ||
      path_cap < 0) {
    set_error("vine_family_solve: invalid arguments");
    return VINE_E_ARG;
  }
  for (int32_t i = 0; i < nq; i++) {
    if (src[i] < 0 || src[i] >= s->g->n || dst[i] < 0 || dst[i] >= s->g->n) {
      set_error("vine_family_solve: query %d has a vertex id outside [0, %d)", i, s->g->n);
      return VINE_E_ARG;
    }
    for (int k = 0; k < 3; k++)
      if (std::isnan(limits[3 * (size_t)i + k])) {
        set_error("vine_family_solve: query %d has a NaN limit", i);
        return VINE_E_ARG;
      }
  }
  if (nq == 0) {
    if (out_path_off) out_path_off[0] = 0;
    return VINE_OK;
  }
  if (hipSetDevice(s->g->device) != hipSuccess) {
    set_error("vine_family_solve: hipSetDevice failed");
    return VINE_E_HIP;
  }
  switch (s->family) {
    case VINE_FAMILY_HOP:
      return solve_family<vine::hop::c_label>(s, nq, src, dst, limits, out_label, out_feasible, out_path_off, out_path,
                                              path_cap, out_counters);
    case VINE_FAMILY_HOP_MAXCOST:
      return solve_family<vine::hop_maxcost::c_label>(s, nq, src, dst, limits, out_label, out_feasible, out_path_off,
                                                      out_path, path_cap, out_counters);
    case VINE_FAMILY_SOCIAL:
      return solve_family<vine::social::c_label>(s, nq, src, dst, limits, out_label, out_feasible, out_path_off,
                                                 out_path, path_cap, out_counters);
    case VINE_FAMILY_QOS:
      return solve_family<vine::qos::c_label>(s, nq, src, dst, limits, out_label, out_feasible, out_path_off, out_path,
                                              path_cap, out_counters);
    default:  // VINE_FAMILY_CSP and VINE_FAMILY_2OSP share c_label
      return solve_family<vine::c_label>(s, nq, src, dst, limits, out_label, out_feasible, out_path_off, out_path,
                                         path_cap, out_counters);
  }
}

int64_t vine_family_front(const vine_family_solver* s, int32_t q, int32_t v, float* out_labels, int64_t cap) {
  if (!s || q < 0 || q >= s->last_nb || v < 0 || v >= s->g->n || cap < 0 || (cap > 0 && !out_labels)) {
    set_error("vine_family_front: invalid arguments");
    return VINE_E_ARG;
  }
  switch (s->family) {
    case VINE_FAMILY_HOP: return front_family<vine::hop::c_label>(s, q, v, out_labels, cap);
    case VINE_FAMILY_HOP_MAXCOST: return front_family<vine::hop_maxcost::c_label>(s, q, v, out_labels, cap);
    case VINE_FAMILY_SOCIAL: return front_family<vine::social::c_label>(s, q, v, out_labels, cap);
    case VINE_FAMILY_QOS: return front_family<vine::qos::c_label>(s, q, v, out_labels, cap);
    default: return front_family<vine::c_label>(s, q, v, out_labels, cap);
  }
}

}  // extern "C"
