// Device side of the exact CSP solver for gfx950 (CDNA4, wave64).
//
// A round is two main kernels (plus the far-tier pull and the window policy):
//   expand  (k_expand)  -- counterpart of ck_expand_bucket_with_filter
//                          (algorithm/bucket.cuh:15-195) and ck_recheck_bucket
//                          (:217-245): every live frontier label inside the
//                          window is extended along its out-arcs; a candidate is
//                          filtered by the constraint, by the pruning bounds and
//                          by a read-only dominance check against the target
//                          vertex's Pareto front, and pushed onto a
//                          per-(query, vertex) candidate list.  Labels outside
//                          the window are carried.
//   merge   (k_merge)   -- counterpart of second_level_pruning
//                          (algorithm/dominance.cuh:15-153): a group of lanes
//                          owns each touched (query, vertex) and folds its
//                          candidates into the sorted Pareto front (drop
//                          dominated, evict dominated entries, append survivors
//                          to the label log, which is the next frontier).  No
//                          locks: the two kernels never run concurrently and
//                          each front has a single writer.
// The thrust scan + D2H size + ck_fill_buckets sequence of the reference
// (algorithm/cspp.cuh:245-292, bucket.cuh:197-215) is replaced by wave-level
// ballot/mbcnt compaction into the label log; the frontier size never leaves
// the device unless the host polls it.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mlp.hpp"
#include "vine/label.hpp"

namespace vine {
namespace dev {

constexpr int WAVE = 64;
// Threads per workgroup of the round kernels (k_expand, k_merge, k_pull_policy)
// at most: their per-wave LDS tables have MAX_BLOCK / WAVE rows, and the host
// refuses a larger block_size before any launch.
constexpr int MAX_BLOCK = 256;
constexpr int MAX_WAVES = MAX_BLOCK / WAVE;
static_assert(MAX_BLOCK % WAVE == 0 && MAX_WAVES >= 1, "round kernels run whole waves");
constexpr float FINF = __builtin_huge_valf();
// The fourth word of an arc record: the target's own arc range in the same
// CSR, first arc << 8 | out-degree, where that fits (fewer than 2^24 arcs and
// an out-degree below 255); ROW_UNKNOWN otherwise (read the row pointers).
constexpr uint32_t ROW_UNKNOWN = 0xFFFFFFFFu;
__device__ __forceinline__ bool row_known(uint32_t ri) { return ri != ROW_UNKNOWN; }
constexpr float EXACT_LIMIT = 16777216.0f;  // 2^24
constexpr int INIT_CAP = 4;                 // first Pareto-set allocation (entries)
constexpr int TS_CAP = 1 << 20;             // per-round timestamps kept
constexpr int CCHUNK = 8;                   // candidate records per list chunk
constexpr int FU = 16;                      // far-tier allocation unit (records)
constexpr int FSEG = 64;                    // records per drain segment (one per lane)
constexpr int FCLS_MAX = 12;                // far chunks grow 16, 32 .. 65,536 records along a bucket
constexpr int NBK = 32768;                  // far-tier key buckets per query
constexpr int NBW = NBK / 64;               // occupancy words per query
constexpr float MLP_LOG2_MIN = 0.f;         // learned window range: 1 ..
constexpr float MLP_LOG2_MAX = 7.f;         //   128 key units

enum : int { OVF_LABEL = 1, OVF_CAND = 2, OVF_TOUCH = 4, OVF_POOL = 8, OVF_CARRY = 16, OVF_SSSP = 32, OVF_FAR = 64 };

struct Counters {
  unsigned long long label_ctr;  // labels in the log
  unsigned long long pool_ctr;   // Pareto pool entries handed out
  int cand_ctr[2];   // candidate chunks handed out this round
  int touched_n[2];
  int carry_n[2];
  long long lstart[2];   // first label of the frontier's "new" range, per parity
  int round_e;           // round number read by expand
  int round_m;           // round number read by merge (written by expand)
  int sround;            // most bound-search rounds of one query (all searches)
  int overflow;
  int range_flag;
  unsigned long long t_start;  // device wall clock at batch start
  unsigned long long expanded, relaxed, cands, inserted, sssp_relaxed;
  unsigned long long touched_total;  // (query, vertex) merges
  unsigned long long merge_scanned;  // Pareto entries compared in merge
  unsigned long long cand_chunks;    // candidate chunks folded by merge
  int merge_chain_max;               // longest candidate list merged by one lane
  int far_chunks;                    // far-tier units handed out (bump)
  int drain_n;                       // drain-list segments queued for the next expand
  long long far_live;                // records currently in far tiers
  unsigned long long far_pulled;     // records moved far -> near
  unsigned long long near_carried, far_appended, scanned;
  unsigned long long hash_dropped;   // candidates dropped by the LDS table (dominated by another candidate)
  unsigned long long hash_overflow;  // candidates that found no room in the LDS table (appended directly)
  long long trace_n;  // scheduler feature rows recorded
};

struct Ctx {
  int32_t n;
  int32_t B;
  // graph
  const int32_t* row_ptr;
  const int4* edges;  // {dst, len bits, delay bits, 0}
  const int32_t* rrow_ptr;
  const int4* redges;
  const int32_t* vinv;  // internal vertex id -> caller's id
  // per query slot
  const int32_t* q_src;
  const int32_t* q_dst;
  const float* q_bound;
  float* q_delta;
  float2* q_best;
  int32_t* q_best_id;
  uint32_t* q_fmin;  // [2][B] float bits
  int32_t* q_last;   // last round with live frontier labels
  // per (query, vertex)
  int2* pset;        // {pool offset, cnt | capcode << 26}
  // per (query, vertex) candidate list head: chunk index << 32 | records
  // claimed in that chunk (may exceed CCHUNK; ~0 = empty)
  unsigned long long* cand_head;
  // lbl = length to t (restricted to lbd <= bound); lbw = the Lagrangian
  // distance to t under weights LSCALE * len + q_lk[q] * delay (same
  // restriction; +inf where not searched or q_lk[q] = 0), as integer bits.
  // Two 4-B arrays: a bound search's line holds 32 vertices' values.
  float* lbl;
  float* lbw;
  float* q_lk;       // [B] Lagrange multiplier x LSCALE of the query (0 = none)
  int32_t lagrange;  // multipliers tried by the Lagrangian bound searches (0: no lbw)
  unsigned long long* lbd;  // packed (delay bits << 32 | len bits): lexicographic
                            // (delay, len) distance to t; the len part is the
                            // length of a minimum-delay path
  // pools
  float2* pool_cd;
  int32_t* pool_id;
  unsigned long long pool_cap;
  int4* labs;        // {vid, d bits, delay bits, parent}
  int2* lmeta;       // {query slot, key bits}
  uint8_t* dead;
  long long label_cap;
  // per-round candidates: per (query, vertex) a list of chunks of CCHUNK
  // records {d bits, delay bits, parent label, 0}; chunk header {count, next}
  int2* chd;
  int4* crec;
  long long* touched;
  int32_t chunk_cap;
  int4* carry[2];    // near tier: carried entries {label id, query slot, key bits, pruning delay bits}
  int32_t carry_cap;
  // far tier: per query NBK key buckets, bucket b holding the carried entries
  // with key in [base + b*bw, base + (b+1)*bw) (the last one also everything
  // beyond) as a list of chunks.  A chunk is 2^cls units of FU records
  // (cls grows by one along a bucket's list up to FCLS_MAX); a head or next
  // pointer is (first unit << 4) | cls, and the chunk header lives at
  // fhd[first unit].  Buckets below far_next are "pulled": their keys belong
  // to the near tier.
  int32_t* far_head;  // [B * NBK]
  unsigned long long* far_occ;  // [B * NBW] bit b: bucket b holds a chunk
  int32_t* far_next;  // [B] first bucket not yet pulled
  int32_t* far_cnt;   // [B] entries in the query's far tier
  float* far_base;    // [B] smallest possible key of the query
  float* far_bw;      // [B] bucket width
  int2* fhd;          // [far_cap] chunk header {count, next pointer}, at the chunk's first unit
  int4* frec;         // [far_cap * FU] records
  int32_t far_cap;    // units
  int2* dl;           // [far_cap] drain list: segments {first unit, records <= FSEG} for the next expand
  // first-level candidate table in LDS (hash_slots > 0): per workgroup,
  // hash_slots slots keyed by (query, vertex), each a small Pareto set of
  // HASH_K candidates; hash_pred[2] (by round parity) is the learned policy's
  // predicted expansions of the next round, from which the active capacity
  // follows (hash_mlp)
  int32_t hash_slots;
  int32_t hash_mlp;
  float* hash_pred;
  float avg_deg;
  float far_floor;    // smallest near-tier reach (key units)
  float far_mult;     // the near tier reaches far_mult windows past the front
  uint32_t* inq;     // [B][n] bound searches: stamp of the round whose queue last took the vertex (0 = none)
  int32_t* wq;       // [B][2][n] bound-search frontier queues
  int32_t* q_order;  // [B] bound-search dispatch order (null: slot order)
  int4* wnear;       // workgroup engine: [B][2][wn_cap] near-tier records per query
  long long wn_cap;
  unsigned long long* q_fin;  // [B] device clock when the query's search ended (workgroup engine)
  unsigned long long* ts;  // [TS_CAP] per-round merge-end wall clock
  Counters* ct;
  int32_t ellipse;  // reverse bound searches restricted to fwd + dist <= bound
  float* fwd;           // [B][n] forward search's min delay from s (+inf: not reached)
  uint32_t* ell;        // [B][ell_words] bit v: lbd(v) is finite (v lies in the bound region)
  int32_t ell_words;
  int32_t fwd_cut;      // twin arcs and integer weights: fwd stops at (bound + fwd(t)) / 2
  const float4* lmk;    // [n] min delays from four landmarks (-1: unreachable); null: no landmark test
  // scheduler: per-query round statistics, MLP, feature trace
  int32_t stats_on;
  int32_t policy;     // VINE_DELTA_*
  long long dop_count;  // VINE_DELTA_DOP: frontier entries the round expands (the last ones), -1 = off
  int32_t* q_exp;     // [B] labels expanded this round
  int32_t* q_ins;     // [B] labels inserted this round
  int32_t* q_near;    // [B] near-tier entries queued this round
  long long* q_exptot;  // [B] labels expanded in the batch
  float* q_fcast;     // [B] MLP forecast of log2(1 + next round's expansions)
  const float* mlp;   // MLP parameters (VINE_DELTA_MLP)
  float* trace;       // [trace_cap][10] rows {q, round, x0..x7}
  long long trace_cap;
  int32_t trace_every;
};

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

__device__ __forceinline__ int mbcnt64(unsigned long long m) {
  return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                        __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Wave-aggregated slot allocation: one atomic per wave, ranks by mbcnt.
__device__ __forceinline__ int wave_alloc32(int* ctr, bool pred) {
  unsigned long long m = __ballot(pred);
  if (m == 0ull) return -1;
  int leader = __ffsll((long long)m) - 1;
  int base = 0;
  if (lane_id() == leader) base = atomicAdd(ctr, __popcll(m));
  base = __shfl(base, leader, WAVE);
  return pred ? base + mbcnt64(m) : -1;
}

__device__ __forceinline__ long long wave_alloc64(unsigned long long* ctr, bool pred) {
  unsigned long long m = __ballot(pred);
  if (m == 0ull) return -1;
  int leader = __ffsll((long long)m) - 1;
  unsigned long long base = 0;
  if (lane_id() == leader) base = atomicAdd(ctr, (unsigned long long)__popcll(m));
  base = __shfl(base, leader, WAVE);
  return pred ? (long long)base + mbcnt64(m) : -1;
}

// Variable-size wave allocation: returns this lane's offset for `size` items.
__device__ __forceinline__ long long wave_alloc64_n(unsigned long long* ctr, int size) {
  int lane = lane_id();
  int incl = size;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    int t = __shfl_up(incl, o, WAVE);
    if (lane >= o) incl += t;
  }
  int total = __shfl(incl, WAVE - 1, WAVE);
  if (total == 0) return -1;
  unsigned long long base = 0;
  if (lane == 0) base = atomicAdd(ctr, (unsigned long long)total);
  base = __shfl(base, 0, WAVE);
  return (long long)base + (incl - size);
}

// Workgroup sums of per-thread counts, one atomic per counter and workgroup
// (the statistics counters are shared by every workgroup of the grid, so
// per-wave atomics on them serialise).  Every thread must call it: it
// synchronises the workgroup.  Blocks have at most MAX_BLOCK threads.
template <int N>
__device__ __forceinline__ void block_add(unsigned long long* const (&ctr)[N], const long long (&v)[N]) {
  static_assert(N <= MAX_BLOCK, "one thread per counter");
  __shared__ long long red[N][MAX_WAVES];
  const int lane = (int)__lane_id();
  const int wid = threadIdx.x / 64;
#pragma unroll
  for (int k = 0; k < N; k++) {
    long long x = v[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    if (lane == 0) red[k][wid] = x;
  }
  __syncthreads();
  if ((int)threadIdx.x < N) {
    long long sum = 0;
    for (int w = 0; w < (int)(blockDim.x + 63) / 64; w++) sum += red[threadIdx.x][w];
    if (sum) atomicAdd(ctr[threadIdx.x], (unsigned long long)sum);
  }
}

// Orders the calling wave's LDS accesses around it (a wave executes its LDS
// instructions in order; this keeps the compiler from moving them across).
__device__ __forceinline__ void wave_lds_order() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Groups the wave's lanes with `act` set by a 64-bit key: a lane that joins
// gets its group's lanes in `gm` and one of them as `lead`.  Matching goes
// through a table of GROUP_SLOTS entries in LDS owned by the calling wave
// (slot = hash of the key); a lane whose slot a different key took stays
// alone (lead = lane, gm = its own bit, as on entry).  Called by every lane
// of the wave; the mask table holds zeros on entry and again on return.
constexpr int GROUP_SLOTS = 128;

__device__ __forceinline__ void wave_group_by_key(int* s_lead, unsigned long long* s_mask, bool act,
                                                  long long key, int lane, int& lead, unsigned long long& gm) {
  const unsigned long long many = __ballot(act);
  if ((many & (many - 1ull)) == 0ull) return;  // wave-uniform: at most one lane
  const int hs = (int)(((unsigned long long)key * 0x9E3779B97F4A7C15ull) >> 57);
  if (act) s_lead[hs] = lane;
  wave_lds_order();
  const int l = act ? s_lead[hs] : lane;
  const int klo = __shfl((int)(key & 0xffffffffll), l, WAVE);
  const int khi = __shfl((int)(key >> 32), l, WAVE);
  const bool join = act && klo == (int)(key & 0xffffffffll) && khi == (int)(key >> 32);
  if (join) atomicOr(&s_mask[hs], 1ull << lane);
  wave_lds_order();
  if (join) {
    gm = s_mask[hs];
    lead = l;
  }
  wave_lds_order();
  if (join && lane == l) s_mask[hs] = 0ull;
}

__device__ __forceinline__ void wave_add(unsigned long long* ctr, int v) {
  int lane = lane_id();
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  if (lane == 0 && v) atomicAdd(ctr, (unsigned long long)v);
}

// Per-key counter increments aggregated over the wave: lanes holding the same
// key add together (one atomic per distinct key).  Wave-uniform call.
__device__ __forceinline__ void wave_add_keyed(int* arr, int key, bool pred) {
  unsigned long long todo = __ballot(pred);
  while (todo) {
    const int leader = __ffsll((long long)todo) - 1;
    const int k = __shfl(key, leader, WAVE);
    const unsigned long long same = __ballot(pred && key == k) & todo;
    if (lane_id() == leader) atomicAdd(&arr[k], __popcll(same));
    todo &= ~same;
  }
}

// arr[key] += n summed over the lanes holding each key (one atomic per key).
// Wave-uniform call.
__device__ __forceinline__ void wave_add_keyed_n(int* arr, int key, int n, bool pred) {
  unsigned long long todo = __ballot(pred);
  while (todo) {
    const int leader = __ffsll((long long)todo) - 1;
    const int k = __shfl(key, leader, WAVE);
    const bool mine = pred && key == k;
    const unsigned long long same = __ballot(mine) & todo;
    int v = mine ? n : 0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    if (lane_id() == leader && v) atomicAdd(&arr[k], v);
    todo &= ~same;
  }
}

// arr[key] = min(arr[key], val) over the lanes holding each key (one atomic
// per key; vals are float bits of non-negative keys).  Wave-uniform call.
__device__ __forceinline__ void wave_min_keyed(uint32_t* arr, int key, uint32_t val, bool pred) {
  unsigned long long todo = __ballot(pred);
  while (todo) {
    const int leader = __ffsll((long long)todo) - 1;
    const int k = __shfl(key, leader, WAVE);
    const bool mine = pred && key == k;
    const unsigned long long same = __ballot(mine) & todo;
    uint32_t v = mine ? val : 0xFFFFFFFFu;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o, WAVE));
    if (lane_id() == leader) atomicMin(&arr[k], v);
    todo &= ~same;
  }
}

// Visit every (item, sub-item) pair of the wave, 64 pairs per step: lane i
// holds an item with `deg` sub-items starting at `row`.  The wave scans the
// degrees, then each lane finds its owner by binary search over the exclusive
// offsets (all through cross-lane permutes, no LDS).  body(owner, pos, active)
// runs with wave-uniform control flow.
template <class F>
__device__ __forceinline__ void wave_flat(int deg, int row, F&& body) {
  int lane = lane_id();
  int incl = deg;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    int t = __shfl_up(incl, o, WAVE);
    if (lane >= o) incl += t;
  }
  int excl = incl - deg;
  int total = __shfl(incl, WAVE - 1, WAVE);
  for (int k = 0; k < total; k += WAVE) {
    int e = k + lane;
    bool act = e < total;
    int ec = act ? e : total - 1;
    int lo = 0;
#pragma unroll
    for (int step = 32; step > 0; step >>= 1) {
      int v = __shfl(excl, lo + step, WAVE);
      if (v <= ec) lo += step;
    }
    int pos = __shfl(row, lo, WAVE) + (ec - __shfl(excl, lo, WAVE));
    body(lo, pos, act);
  }
}

// Loads that bypass the (non-coherent) vector L1, for values other CUs update
// with atomics during the same kernel.
__device__ __forceinline__ float load_agent(const float* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t load_agent_u32(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ unsigned long long pack2(float hi, float lo) {
  return ((unsigned long long)__float_as_uint(hi) << 32) | (unsigned long long)__float_as_uint(lo);
}
__device__ __forceinline__ float hi_of(unsigned long long p) { return __uint_as_float((uint32_t)(p >> 32)); }
__device__ __forceinline__ float lo_of(unsigned long long p) { return __uint_as_float((uint32_t)p); }
__device__ __forceinline__ float lbd_of(const Ctx& c, long long qv) { return hi_of(c.lbd[qv]); }
// {lbl, lbw} at (query, vertex) qv, as the label keys take them
__device__ __forceinline__ float2 lbl_pair(const Ctx& c, long long qv) { return make_float2(c.lbl[qv], c.lbw[qv]); }
// lbd before its search reaches the vertex: both words +inf
constexpr unsigned long long LBD_NONE = 0x7F8000007F800000ull;

// Key and lower bound on the final cost of any feasible completion of a label
// (d, delay) at a vertex with bounds lw = {lbl, lbw}, for a query with
// multiplier k / LSCALE and delay bound bnd.  Two bounds, the larger is kept:
//   d + lbl                                  (shortest remaining length)
//   d + (lbw + k * (delay - bnd)) / LSCALE   (Lagrangian relaxation: a completion
//     p with delay(p) <= bnd - delay has LSCALE*len(p) + k*delay(p) >= lbw, so
//     LSCALE*len(p) >= lbw - k*(bnd - delay))
// lbw is an integer held in the bits of lw.y (LBW_NONE: not computed).  With
// integer weights the final cost is an integer, so the second bound is rounded
// up; it is evaluated in double, where every term is exact.  The maximum of two
// consistent bounds is consistent: a child's key is never below its parent's.
constexpr float LSCALE = 256.f;
constexpr uint32_t LBW_NONE = 0x7F800000u;  // the bits of +inf, which the batch reset writes
__device__ __forceinline__ float lower_key(float d, float delay, float2 lw, float k, float bnd) {
  float key = d + lw.x;
  const uint32_t wb = __float_as_uint(lw.y);
  if (wb < LBW_NONE) {
    const double r = ceil(((double)wb + (double)k * ((double)delay - (double)bnd)) * (1.0 / (double)LSCALE));
    key = fmaxf(key, __double2float_rd((double)d + r));
  }
  return key;
}

// Lagrangian arc weight LSCALE * len + k * delay of integer-valued len and delay
__device__ __forceinline__ unsigned long long lagr_weight(int4 e, float k) {
  return (unsigned long long)(LSCALE * __int_as_float(e.y)) + (unsigned long long)k * (unsigned long long)__int_as_float(e.z);
}

__device__ __forceinline__ int far_bucket(const Ctx& c, int q, float key) {
  float t = (key - c.far_base[q]) / c.far_bw[q];
  if (!(t < (float)(NBK - 1))) return NBK - 1;
  return t < 0.f ? 0 : (int)t;
}

__device__ __forceinline__ int pset_cnt(int2 p) { return p.y & ((1 << 26) - 1); }
__device__ __forceinline__ int pset_cap(int2 p) {
  int code = (int)((uint32_t)p.y >> 26);
  return code ? (1 << (code - 1)) : 0;
}
__device__ __forceinline__ int2 pset_make(uint32_t off, int cnt, int cap) {
  int code = 0;
  if (cap) code = 32 - __clz(cap);  // cap is a power of two: log2(cap) + 1
  return make_int2((int)off, cnt | (code << 26));
}

__device__ __forceinline__ bool lex_ge(float a, float b, float c, float d) {
  // (a, b) >= (c, d) lexicographically
  return a > c || (a == c && b >= d);
}
// Pareto fronts in the pool are sorted by d ascending (so delay strictly
// descends; label_traits::pack gives the two criteria, and the store needs
// dominate(a, b) to be their componentwise <=).  Is `cand` dominated by one of
// the cnt entries at ent?  Only the last entry with d <= cand.d can dominate
// it: usually the set's last entry, else found by binary search.
template <typename Label_t>
__device__ __forceinline__ bool front_dominates(const float2* ent, int cnt, int vid, const Label_t& cand) {
  using LT = label_traits<Label_t>;
  if (cnt <= 0) return false;
  const float cx = LT::pack(cand).x;
  float2 e = ent[cnt - 1];
  if (!(e.x <= cx)) {
    int lo = 0, hi = cnt - 1;  // entries >= hi have d > cand.d
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (ent[mid].x <= cx) lo = mid + 1;
      else hi = mid;
    }
    if (lo == 0) return false;
    e = ent[lo - 1];
  }
  return Label_t::dominate(LT::unpack(vid, e), cand);
}

template <int PRUNE>
__global__ void k_init_batch(Ctx c, int nq) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= c.B) return;
  if (q < nq) {
    int s = c.q_src[q];
    long long qs = (long long)q * c.n + s;
    float key = 0.f;
    if (PRUNE == 2) key = lower_key(0.f, 0.f, lbl_pair(c, qs), c.q_lk[q], c.q_bound[q]);
    c.labs[q] = make_int4(s, __float_as_int(0.f), __float_as_int(0.f), -1);
    c.lmeta[q] = make_int2(q, __float_as_int(key));
    c.dead[q] = 0;
    // source in its own Pareto set, storage [q*INIT_CAP, +INIT_CAP)
    c.pool_cd[(long long)q * INIT_CAP] = make_float2(0.f, 0.f);
    c.pool_id[(long long)q * INIT_CAP] = q;
    c.pset[qs] = pset_make((uint32_t)(q * INIT_CAP), 1, INIT_CAP);
    bool st = (s == c.q_dst[q]);
    float2 best0 = make_float2(FINF, FINF);
    if (PRUNE == 2) {
      // a minimum-delay s->t path is feasible, so its length UB bounds the
      // answer: start t's best at (UB + 1, 0), which prunes keys > UB and no
      // label that could still reach an answer (sums are integers)
      unsigned long long pk = c.lbd[qs];
      if (hi_of(pk) <= c.q_bound[q]) best0 = make_float2(lo_of(pk) + 1.0f, 0.f);
    }
    c.q_best[q] = st ? make_float2(0.f, 0.f) : best0;
    c.q_best_id[q] = st ? q : -1;
    c.q_fmin[q] = __float_as_uint(key);
    // far-tier buckets: from the source key (the smallest any label can have)
    // up to the upper bound when there is one
    float reach = fmaxf(c.far_mult * c.q_delta[q], c.far_floor);
    float bw = reach;
    if (PRUNE == 2 && best0.x < FINF) bw = fmaxf(bw, (best0.x - key) / (float)(NBK - 2));
    c.far_base[q] = key;
    c.far_bw[q] = bw;
    c.far_cnt[q] = 0;
    // no far tier without a finite window or for an unreachable source
    c.far_next[q] = (reach < FINF && key < FINF) ? 1 + (int)(reach / bw) : NBK;
  } else {
    // unused slot: a dead label so that the log stays dense
    c.labs[q] = make_int4(0, 0, 0, -1);
    c.lmeta[q] = make_int2(q, __float_as_int(FINF));
    c.dead[q] = 1;
    c.q_best[q] = make_float2(FINF, FINF);
    c.q_best_id[q] = -1;
    c.q_fmin[q] = __float_as_uint(FINF);
    c.far_base[q] = 0.f;
    c.far_bw[q] = 1.f;
    c.far_next[q] = NBK;
    c.far_cnt[q] = 0;
  }
  c.q_fmin[c.B + q] = __float_as_uint(FINF);
  if (q == 0 && c.hash_pred) {
    c.hash_pred[0] = -1.f;
    c.hash_pred[1] = -1.f;
  }
  c.q_last[q] = -1;
  c.q_exp[q] = 0;
  c.q_ins[q] = 0;
  c.q_near[q] = 0;
  c.q_exptot[q] = 0;
}

// ---------------------------------------------------------------------------
// Lower bounds to t (VINE_PRUNE_BOUNDS).  One workgroup per query slot runs
// its searches to completion with its own frontier queues: delta-stepping
// rounds separated by workgroup barriers, no grid-wide rounds, so a short
// query's workgroup retires early and a long one is never held up by others.
// Each round relaxes the frontier vertices whose distance is within `delta`
// of the round's smallest frontier distance and carries the rest, which keeps
// re-relaxation close to Dijkstra's.
//   WHICH = 2: forward min delay fwd from s (on the graph itself, into fwd),
//   limited to radius `bound`.  Run first when c.ellipse is set.  With
//   c.fwd_cut the radius shrinks to (bound + D) / 2, D = fwd(t): a twin arc
//   of no larger delay for every arc makes the min delay from t to v at most
//   lbd(v), so fwd(v) <= D + lbd(v), and a vertex with fwd + lbd <= bound has
//   fwd <= (bound + D) / 2.  The current fwd(t) is an upper bound on D, so
//   the radius from it is safe at every round, and integer delays (below
//   2^24, path sums exact) make the floor of the half exact.  A vertex left
//   with a larger tentative fwd is outside the region all the same.
//   WHICH = 0: delay to t on the transposed graph, limited to radius `bound`,
//   or with c.ellipse to fwd + lbd <= bound.  A label at u has delay >=
//   fwd(u), so outside that region every label fails the delay test and
//   lbd = +inf prunes it the same.  The packed (delay, len) pair makes the
//   distance lexicographic; its len part is the length of a min-delay path.
//   WHICH = 1: len to t, restricted to vertices with lbd <= bound: only those
//   can lie on a feasible completion, so the restricted distance is a valid
//   lower bound for every label that passes the delay test.  The region is
//   the set of vertices the search of lbd reached (it only lowers lbd to
//   values within the bound), kept as a bitmap (in_region).
//   WHICH = 3: the Lagrangian distance lbw to t under the integer weights
//   LSCALE * len + lk * delay, with the same restriction, as unsigned integers
//   in the bits of lbw (sums saturate at LBW_NONE, which is never stored);
//   the window works on their float values.
constexpr int BOUND_BLOCK = 512;
// Frontier of a bound search: the first BQ_LDS entries of each round's queue
// in LDS, the rest in the query's HBM queue.  A vertex enters the next
// round's queue once: an LDS set of BQ_HASH slots, cleared every round,
// records who is in it; a vertex whose BQ_PROBES probes find neither itself
// nor a free slot (the set only fills up during a round, so that happens on
// its every push of the round) is checked against its HBM stamp instead,
// the round's number in the workgroup's count of rounds (`stamp`, never 0).
// An LDS entry carries the vertex's arc range, from the arc that reached it,
// and its key when queued, which decides whether the round takes it (a later
// decrease found the vertex queued already, so the key can be stale high:
// such an entry is carried once and queued again with its current key); the
// distance it is relaxed with is read when it is taken, so a vertex lowered
// after it was queued is relaxed with its newest distance.  Those reads and
// the arc reads are issued together.  HBM entries hold only the vertex and
// are taken or carried by their distance (within four windows).
constexpr int BQ_LDS = 1024;
constexpr int BQ_HASH = 2048;
constexpr int BQ_PROBES = 8;

struct BoundLds {
  int n[2];          // entries in each queue
  uint32_t kmin[2];  // smallest key queued in each (float bits)
  uint32_t stamp;    // rounds of this workgroup so far, + 1
  int4 q[2][BQ_LDS];  // {vertex, its arc range (row_info), key bits when queued, 0}
  int hkey[BQ_HASH];  // vertices already in the next queue (-1: free)
};

__device__ __forceinline__ void bq_push(const Ctx& c, BoundLds& L, long long qb, int nxt, int32_t* hq, uint32_t stamp,
                                        bool pred, int w, uint32_t ri, uint32_t keybits) {
  bool ins = pred;
  if (pred) {
    int h = (int)(((uint32_t)w * 2654435761u) >> 21) & (BQ_HASH - 1);
    bool set = false;
    for (int p = 0; p < BQ_PROBES; p++) {
      const int k = atomicCAS(&L.hkey[h], -1, w);
      if (k == -1 || k == w) {  // claimed (new in the queue) or found (queued already)
        set = true;
        ins = k == -1;
        break;
      }
      h = (h + 1) & (BQ_HASH - 1);
    }
    if (!set) ins = atomicExch(&c.inq[qb + w], stamp) != stamp;
  }
  const int slot = wave_alloc32(&L.n[nxt], ins);
  if (ins) {
    if (slot < BQ_LDS) L.q[nxt][slot] = make_int4(w, (int)ri, (int)keybits, 0);
    else if (slot - BQ_LDS < c.n) hq[slot - BQ_LDS] = w;
    else atomicOr(&c.ct->overflow, OVF_SSSP);
  }
}

// distance field of search WHICH at (query, vertex) qv
// The forward search's distance fwd, 4 B per (query, vertex): a 128-B line
// holds 32 vertices' values.
__device__ __forceinline__ float* fwd_ptr(const Ctx& c, long long qv) { return &c.fwd[qv]; }

// A lower bound on the min delay from v to t from the landmark delays lv
// (of v) and lt (of t): d(v, t) >= d(L, t) - d(L, v) on any graph, and with
// twin arcs (delays symmetric) also >= d(L, v) - d(L, t).  Both are
// consistent (h(u) <= delay(u, v) + h(v) along an arc), and exact on integer
// delays below 2^24.
__device__ __forceinline__ float landmark_bound(float4 lv, float4 lt, bool twin) {
  const float a[4] = {lt.x, lt.y, lt.z, lt.w}, b[4] = {lv.x, lv.y, lv.z, lv.w};
  float h = 0.f;
#pragma unroll
  for (int j = 0; j < 4; j++)
    if (a[j] >= 0.f && b[j] >= 0.f) {
      h = fmaxf(h, a[j] - b[j]);
      if (twin) h = fmaxf(h, b[j] - a[j]);
    }
  return h;
}

// Whether v lies in query q's bound region, i.e. was reached by the search
// of lbd (whose values are all <= the bound): one bit per vertex, set when
// lbd(v) first drops from +inf, so the later searches' region test reads a
// line that covers 1,024 vertices instead of v's 8-B lbd.
__device__ __forceinline__ bool in_region(const Ctx& c, int q, int v) {
  return (c.ell[(long long)q * c.ell_words + (v >> 5)] >> (v & 31)) & 1u;
}

template <int WHICH>
__device__ __forceinline__ float* bound_dist(const Ctx& c, long long qv) {
  return WHICH == 2 ? fwd_ptr(c, qv) : WHICH == 3 ? &c.lbw[qv] : &c.lbl[qv];
}

// WHICH = 4 keeps four Lagrangian distances per (query, vertex) in the label
// search's pset and cand_head words, which are free while the bound searches
// run (the workgroup restores them before it ends)
__device__ __forceinline__ uint32_t* grid_field(const Ctx& c, long long qv, int j) {
  return j < 2 ? (uint32_t*)&c.pset[qv] + j : (uint32_t*)&c.cand_head[qv] + (j - 2);
}

__device__ __forceinline__ uint32_t load_agent_u32p(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int WHICH>
__device__ void wg_search(const Ctx& c, int q, float delta, BoundLds& L, int& relaxed, int& rounds,
                          float lk = 0.f, float4 lk4 = make_float4(0.f, 0.f, 0.f, 0.f)) {
  // arcs a lane relaxes per step (their loads and atomics issued together);
  // two measured no faster and take registers that the other stream's label
  // search needs to share a CU
  constexpr int R = 1;
  const float lks[4] = {lk4.x, lk4.y, lk4.z, lk4.w};
  const int n = c.n;
  const long long qb = (long long)q * n;
  int32_t* que[2] = {c.wq + 2 * qb, c.wq + 2 * qb + n};  // HBM part of the two frontiers
  const float bnd = c.q_bound[q];
  const int tid = threadIdx.x;
  const int lane = lane_id();
  const int nw = blockDim.x / WAVE;
  const int wid = tid / WAVE;
  const int32_t* rp = (WHICH == 2) ? c.row_ptr : c.rrow_ptr;
  const int4* ep = (WHICH == 2) ? c.edges : c.redges;
  const int tq = c.q_dst[q];
  // the forward search queues only vertices v with fwd(v) + h(v) <= bound,
  // h the landmark bound on the delay from v to t: every vertex of the
  // region fwd + lbd <= bound passes, and (h being consistent) so does every
  // vertex of its shortest path from s, so fwd stays exact on the region
  const bool use_lmk = WHICH == 2 && c.lmk != nullptr;
  const float4 lmt = use_lmk ? c.lmk[tq] : make_float4(-1.f, -1.f, -1.f, -1.f);
  if (tid == 0) {
    const int root = (WHICH == 2) ? c.q_src[q] : c.q_dst[q];
    if (WHICH == 0) {
      c.lbd[qb + root] = pack2(0.f, 0.f);
      atomicOr(&c.ell[(long long)q * c.ell_words + (root >> 5)], 1u << (root & 31));
    }
    else if (WHICH == 4)
      for (int j = 0; j < 4; j++) *grid_field(c, qb + root, j) = 0u;
    else *bound_dist<WHICH>(c, qb + root) = 0.f;
    L.q[0][0] = make_int4(root, (int)ROW_UNKNOWN, (int)__float_as_uint(0.f), 0);
    L.n[0] = 1;
    L.n[1] = 0;
    L.kmin[0] = __float_as_uint(0.f);
    L.kmin[1] = __float_as_uint(FINF);
  }
  for (int i = tid; i < BQ_HASH; i += blockDim.x) L.hkey[i] = -1;
  __syncthreads();
  int cur = 0;
  for (int round = 0;; round++) {
    const int ncur = L.n[cur];
    if (ncur == 0) break;
    // every round settles the smallest frontier vertex for good, so rounds
    // are bounded by about n; this guard only stops a malformed run
    if (round > 4 * n + 1024) {
      if (tid == 0) atomicOr(&c.ct->overflow, OVF_SSSP);
      break;
    }
    rounds = round + 1;
    const int nxt = cur ^ 1;
    const float lim_key = __uint_as_float(L.kmin[cur]) + delta;
    float flim = bnd;  // WHICH 2: radius of the forward search this round
    if (WHICH == 2 && c.fwd_cut && bnd < 16777216.f) {
      const float dt = fminf(load_agent(fwd_ptr(c, qb + tq)), bnd);
      flim = (float)floor(0.5 * ((double)bnd + (double)dt));
    }
    const uint32_t stamp = L.stamp;
    const int32_t* qin = que[cur];
    int32_t* qout = que[nxt];
    // entries are dealt to the waves in turn (entry lane * nw + wid of each
    // block-sized slice), so a small frontier still spreads over every wave
    for (int base = 0; base < ncur; base += blockDim.x) {
      const int i = base + lane * nw + wid;
      // the wave's last entry index is past the LDS part (and exists)
      const bool hbm_wave = min(base + (WAVE - 1) * nw + wid, ncur - 1) >= BQ_LDS;
      int deg = 0, row = 0, u = 0;
      float du = 0.f, dl2 = 0.f;  // distance (delay for WHICH 0 and 2, len for 1); len tie key
      uint32_t dub = 0;           // WHICH 3: the integer distance (du is its float value)
      uint32_t dg[4] = {0u, 0u, 0u, 0u};  // WHICH 4: the four distances (du orders by the first)
      bool carry = false;
      int4 ent = make_int4(0, (int)ROW_UNKNOWN, 0, 0);
      if (i < ncur) {
        ent = i < BQ_LDS ? L.q[cur][i] : make_int4(qin[i - BQ_LDS], (int)ROW_UNKNOWN, 0, 1);
        u = ent.x;
        const long long qv = qb + u;
        // the vertex's current distance (atomics update it in L2)
        if (WHICH == 0) {
          const unsigned long long pk = __hip_atomic_load(&c.lbd[qv], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          du = hi_of(pk);
          dl2 = lo_of(pk);
        } else if (WHICH == 3) {
          dub = __float_as_uint(load_agent(bound_dist<WHICH>(c, qv)));
          du = (float)dub;
        } else if (WHICH == 4) {
#pragma unroll
          for (int j = 0; j < 4; j++) dg[j] = load_agent_u32p(grid_field(c, qv, j));
          du = (float)dg[0];
        } else {
          du = load_agent(bound_dist<WHICH>(c, qv));
        }
        // taken by its key when queued, so the arc reads need not wait for
        // the distance.  HBM entries keep no key and go by their distance,
        // within four windows: a frontier that overflows the LDS part is
        // worked through in fewer rounds, and the overflow never runs far
        // ahead of the window (taking every HBM entry made a search with a
        // 512-entry LDS part six times slower).  The test on them is in the
        // waves that hold any (wave-uniform), so the others never wait.
        if (hbm_wave) carry = ent.w ? du > __uint_as_float(L.kmin[cur]) + 4.f * delta : __int_as_float(ent.z) > lim_key;
        else carry = __int_as_float(ent.z) > lim_key;
        if (!carry) {
          if (row_known((uint32_t)ent.y)) {
            row = (int)((uint32_t)ent.y >> 8);
            deg = ent.y & 255;
          } else {
            row = rp[u];
            deg = rp[u + 1] - row;
          }
        }
      }
      // the wave's arcs, R per lane per step
      int incl = deg;
#pragma unroll
      for (int o = 1; o < WAVE; o <<= 1) {
        const int t = __shfl_up(incl, o, WAVE);
        if (lane >= o) incl += t;
      }
      const int excl = incl - deg;
      const int total = __shfl(incl, WAVE - 1, WAVE);
      for (int k = 0; k < total; k += R * WAVE) {
        int own[R], pos[R];
        bool act[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
          const int e = k + r * WAVE + lane;
          act[r] = e < total;
          const int ec = act[r] ? e : total - 1;
          int lo = 0;
#pragma unroll
          for (int step = 32; step > 0; step >>= 1) {
            const int v = __shfl(excl, lo + step, WAVE);
            if (v <= ec) lo += step;
          }
          own[r] = lo;
          pos[r] = __shfl(row, lo, WAVE) + (ec - __shfl(excl, lo, WAVE));
        }
        int4 ed[R];
#pragma unroll
        for (int r = 0; r < R; r++) ed[r] = act[r] ? ep[pos[r]] : make_int4(0, 0, 0, 0);
        // candidate distances, and the target's current value and region test
        float nd[R];
        bool go[R];
        unsigned long long pk[R];
        uint32_t ndg[R][4];
#pragma unroll
        for (int r = 0; r < R; r++) {
          const float odu = __shfl(du, own[r], WAVE);
          const float odl = __shfl(dl2, own[r], WAVE);
          const long long wv = qb + ed[r].x;
          go[r] = false;
          nd[r] = FINF;
          pk[r] = 0ull;
          if (WHICH == 2) {
            nd[r] = odu + __int_as_float(ed[r].z);
            go[r] = act[r] && nd[r] <= flim && nd[r] < *bound_dist<2>(c, wv);
          } else if (WHICH == 1) {
            nd[r] = odu + __int_as_float(ed[r].y);
            go[r] = act[r] && in_region(c, q, ed[r].x) && nd[r] < *bound_dist<1>(c, wv);
          } else if (WHICH == 3) {
            const uint32_t odub = (uint32_t)__shfl((int)dub, own[r], WAVE);
            const unsigned long long s64 = (unsigned long long)odub + lagr_weight(ed[r], lk);
            ndg[r][0] = s64 < (unsigned long long)LBW_NONE ? (uint32_t)s64 : LBW_NONE;
            nd[r] = (float)ndg[r][0];
            go[r] = act[r] && in_region(c, q, ed[r].x) && ndg[r][0] < *(const uint32_t*)bound_dist<3>(c, wv);
          } else if (WHICH == 4) {
            go[r] = act[r] && in_region(c, q, ed[r].x);
#pragma unroll
            for (int j = 0; j < 4; j++) {
              const uint32_t od = (uint32_t)__shfl((int)dg[j], own[r], WAVE);
              const unsigned long long s64 = (unsigned long long)od + lagr_weight(ed[r], lks[j]);
              ndg[r][j] = s64 < (unsigned long long)LBW_NONE ? (uint32_t)s64 : LBW_NONE;
            }
            nd[r] = (float)ndg[r][0];
          } else {
            nd[r] = odu + __int_as_float(ed[r].z);
            pk[r] = pack2(nd[r], odl + __int_as_float(ed[r].y));
            // with the forward search done, only vertices with fwd + lbd <=
            // bound: every vertex of a feasible s-t path lies there, and so
            // does every vertex of such a vertex's shortest path to t
            const float lim = c.ellipse ? bnd - *fwd_ptr(c, wv) : bnd;
            go[r] = act[r] && nd[r] <= lim && pk[r] < c.lbd[wv];
          }
        }
        bool lowered[R];
        uint32_t nkey[R];  // key of a lowered target (WHICH 4: its first field after the update)
#pragma unroll
        for (int r = 0; r < R; r++) {
          const long long wv = qb + ed[r].x;
          lowered[r] = false;
          nkey[r] = __float_as_uint(nd[r]);
          if (act[r]) relaxed++;
          if (!go[r]) {
          } else if (WHICH == 2 || WHICH == 1) {
            const unsigned old = atomicMin((unsigned*)bound_dist<WHICH>(c, wv), __float_as_uint(nd[r]));
            lowered[r] = __float_as_uint(nd[r]) < old;
          } else if (WHICH == 3) {
            const unsigned old = atomicMin((unsigned*)bound_dist<3>(c, wv), ndg[r][0]);
            lowered[r] = ndg[r][0] < old;
          } else if (WHICH == 4) {
            uint32_t f0 = LBW_NONE;
#pragma unroll
            for (int j = 0; j < 4; j++) {
              uint32_t* fp = grid_field(c, wv, j);
              const uint32_t cf = *fp;
              uint32_t nf = cf;
              if (ndg[r][j] < cf) {
                const unsigned old = atomicMin(fp, ndg[r][j]);
                lowered[r] = lowered[r] || ndg[r][j] < old;
                nf = min(old, ndg[r][j]);
              }
              if (j == 0) f0 = nf;
            }
            nkey[r] = __float_as_uint((float)f0);
          } else {
            const unsigned long long old = atomicMin(&c.lbd[wv], pk[r]);
            lowered[r] = pk[r] < old;
            if (lowered[r] && old == LBD_NONE)
              atomicOr(&c.ell[(long long)q * c.ell_words + (ed[r].x >> 5)], 1u << (ed[r].x & 31));
          }
        }
        if (use_lmk) {
#pragma unroll
          for (int r = 0; r < R; r++)
            if (lowered[r]) lowered[r] = nd[r] + landmark_bound(c.lmk[ed[r].x], lmt, c.fwd_cut != 0) <= bnd;
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
          if (lowered[r] && nkey[r] < L.kmin[nxt]) atomicMin(&L.kmin[nxt], nkey[r]);
          bq_push(c, L, qb, nxt, qout, stamp, lowered[r], ed[r].x, (uint32_t)ed[r].w, nkey[r]);
        }
      }
      // carried entries, queued again with their current key
      if (carry && __float_as_uint(du) < L.kmin[nxt]) atomicMin(&L.kmin[nxt], __float_as_uint(du));
      bq_push(c, L, qb, nxt, qout, stamp, carry, u, (uint32_t)ent.y, __float_as_uint(du));
    }
    __syncthreads();
    if (tid == 0) {
      L.n[cur] = 0;
      L.kmin[cur] = __float_as_uint(FINF);
      L.stamp++;
    }
    for (int i = tid; i < BQ_HASH; i += blockDim.x) L.hkey[i] = -1;
    cur = nxt;
    __syncthreads();
  }
}

// The multiplier of the Lagrangian bound, chosen by a few steps of the
// classical hull iteration once lbl and lbd are final.  Two points of the
// query's (delay, len) trade-off start it: a minimum-delay path (D0, UB), from
// lbd(s), and a shortest path inside the bound region (D1, L1), whose delay is
// summed by walking from s along tight arcs.  The slope between the feasible
// point P0 and the infeasible point P1 is the multiplier at which both cost the
// same; a search under it yields a multiplier-shortest path, walked the same
// way, which replaces P0 or P1 by the side of the bound it falls on.  The
// multiplier whose bound at s is largest is kept.  Any multiplier gives a valid
// bound, so the choice only affects how much is pruned.  No Lagrangian search
// (k = 0) when the shortest path is feasible or as short as the minimum-delay
// one, or when a walk does not reach t within WALK_MAX steps.
constexpr int WALK_MAX = 1 << 17;
constexpr float LK_MAX = 65536.f;

struct LagrState {
  double best_lb;     // largest bound at s so far
  float p0d, p0l;     // feasible point (delay, len)
  float p1d, p1l;     // infeasible point
  float k;            // multiplier to search next (LSCALE units)
  float best_k;       // multiplier of best_lb (0: none)
  float last_k;       // multiplier whose lbw is in memory
  int stop;
};

// Walk from s to t along tight arcs of lbl (WHICH 1), of lbw under
// multiplier k (WHICH 3) or of the four-multiplier search's field j (WHICH 4),
// by one wave; sums of len and delay, false if t is not reached.
template <int WHICH>
__device__ bool walk_tight(const Ctx& c, int q, float k, double& sum_len, double& sum_delay, int j = 0) {
  const int lane = lane_id();
  const long long qb = (long long)q * c.n;
  const int t = c.q_dst[q];
  int v = c.q_src[q];
  float lv = c.lbl[qb + v];
  uint32_t wv = WHICH == 4 ? *grid_field(c, qb + v, j) : __float_as_uint(c.lbw[qb + v]);
  sum_len = 0.0;
  sum_delay = 0.0;
  for (int step = 0; v != t; step++) {
    if (step >= WALK_MAX) return false;
    const int r0 = c.row_ptr[v], r1 = c.row_ptr[v + 1];
    int next = -1;
    float nl = 0.f, nd = 0.f, nlv = 0.f;
    uint32_t nwv = 0;
    for (int base = r0; base < r1 && next < 0; base += WAVE) {
      const int p = base + lane;
      int4 e = make_int4(0, 0, 0, 0);
      float2 lw = make_float2(FINF, FINF);
      bool tight = false;
      if (p < r1) {
        e = c.edges[p];
        lw = lbl_pair(c, qb + e.x);
        if (WHICH == 4) lw.y = __uint_as_float(*grid_field(c, qb + e.x, j));
        if (WHICH == 1) {
          tight = lw.x < FINF && __int_as_float(e.y) + lw.x == lv;
        } else {
          const uint32_t w = __float_as_uint(lw.y);
          tight = w < LBW_NONE && (unsigned long long)w + lagr_weight(e, k) == (unsigned long long)wv;
        }
      }
      const unsigned long long m = __ballot(tight);
      if (m) {
        const int l = __ffsll((long long)m) - 1;
        next = __shfl(e.x, l, WAVE);
        nl = __shfl(__int_as_float(e.y), l, WAVE);
        nd = __shfl(__int_as_float(e.z), l, WAVE);
        nlv = __shfl(lw.x, l, WAVE);
        nwv = (uint32_t)__shfl((int)__float_as_uint(lw.y), l, WAVE);
      }
    }
    if (next < 0) return false;
    sum_len += nl;
    sum_delay += nd;
    v = next;
    lv = nlv;
    wv = nwv;
  }
  return true;
}

__device__ __forceinline__ float hull_multiplier(const LagrState& ls) {
  const float k = rintf(LSCALE * (ls.p0l - ls.p1l) / (ls.p1d - ls.p0d));
  return fminf(fmaxf(k, 1.f), LK_MAX);
}

// first wave: the starting points and the first multiplier
__device__ void lagr_begin(const Ctx& c, int q, LagrState& ls) {
  const long long qb = (long long)q * c.n;
  const int s = c.q_src[q];
  const float bnd = c.q_bound[q];
  const float L1 = c.lbl[qb + s];
  const unsigned long long pk = c.lbd[qb + s];
  const float D0 = hi_of(pk), UB = lo_of(pk);
  double wl = 0.0, wd = 0.0;
  bool go = L1 < FINF && D0 <= bnd && UB > L1 && walk_tight<1>(c, q, 0.f, wl, wd);
  go = go && wd > (double)bnd && wd > (double)D0;
  if (lane_id() == 0) {
    ls.best_lb = -1e300;
    ls.best_k = 0.f;
    ls.last_k = 0.f;
    ls.p0d = D0;
    ls.p0l = UB;
    ls.p1d = (float)wd;
    ls.p1l = L1;
    ls.stop = go ? 0 : 1;
    ls.k = go ? hull_multiplier(ls) : 0.f;
  }
}

// first wave, after the search under ls.k: keep the better bound, move the
// hull point, next multiplier (or stop)
__device__ void lagr_step(const Ctx& c, int q, LagrState& ls) {
  const long long qb = (long long)q * c.n;
  const int s = c.q_src[q];
  const float bnd = c.q_bound[q];
  const float k = ls.k;
  const uint32_t ws = __float_as_uint(c.lbw[qb + s]);
  double wl = 0.0, wd = 0.0;
  const bool ok = ws < LBW_NONE && walk_tight<3>(c, q, k, wl, wd);
  if (lane_id() == 0) {
    ls.last_k = k;
    if (ws < LBW_NONE) {
      const double lb = ((double)ws - (double)k * (double)bnd) / (double)LSCALE;
      if (lb > ls.best_lb) {
        ls.best_lb = lb;
        ls.best_k = k;
      }
    }
    if (!ok) {
      ls.stop = 1;
      return;
    }
    if (wd > (double)bnd) {
      ls.p1d = (float)wd;
      ls.p1l = (float)wl;
    } else {
      ls.p0d = (float)wd;
      ls.p0l = (float)wl;
    }
    float nk = k;
    if (ls.p1d > ls.p0d && ls.p0l > ls.p1l) nk = hull_multiplier(ls);
    ls.stop = (nk == k || nk == ls.best_k) ? 1 : 0;
    ls.k = nk;
  }
}

// Workgroups are dispatched in blockIdx order, so the slots are taken
// largest delay bound first (a bound of r searches a ball of radius r): the
// longest searches start at once instead of in the last wave of workgroups.
// One workgroup sorts (bound, slot) pairs bitonically in LDS.
constexpr int ORDER_MAX = 8192;

__global__ void __launch_bounds__(1024) k_bound_order(Ctx c, int nq) {
  __shared__ float sb[ORDER_MAX];
  __shared__ int si[ORDER_MAX];
  int np = 1;
  while (np < nq) np <<= 1;
  for (int i = threadIdx.x; i < np; i += blockDim.x) {
    sb[i] = i < nq ? c.q_bound[i] : -FINF;
    si[i] = i;
  }
  __syncthreads();
  for (int k = 2; k <= np; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < np; i += blockDim.x) {
        const int l = i ^ j;
        if (l > i) {
          // descending by bound (ties by slot) in blocks where (i & k) == 0
          const bool before = sb[i] > sb[l] || (sb[i] == sb[l] && si[i] < si[l]);
          if (((i & k) == 0) != before) {
            const float tb = sb[i];
            sb[i] = sb[l];
            sb[l] = tb;
            const int ti = si[i];
            si[i] = si[l];
            si[l] = ti;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < nq; i += blockDim.x) c.q_order[i] = si[i];
}

// lagrange_bound = LAGR_GRID: the multiplier from one search under four
// multipliers around the starting slope k0 (GRID_FRAC x k0, all relaxed in the
// same rounds), the four multiplier-shortest paths walked by four waves at once,
// then one hull step from the points nearest the bound on either side; the
// multiplier with the largest bound at s is kept.  Two search passes instead of
// one per multiplier tried.  Returns the rounds of the searches.
constexpr int LAGR_GRID = 2;

__device__ int lagr_grid(const Ctx& c, int q, LagrState& ls, float delta_len, float delta_delay, BoundLds& L,
                         int& relaxed) {
  __shared__ double s_glb[4];
  __shared__ float s_gd[4], s_gl[4];
  __shared__ int s_gok[4];
  __shared__ float s_kh;
  const int tid = threadIdx.x, wid = tid / WAVE;
  const long long qb = (long long)q * c.n;
  const int s = c.q_src[q];
  const float bnd = c.q_bound[q];
  const float k0 = ls.k;
  const float4 lk4 = make_float4(fmaxf(1.f, rintf(0.25f * k0)), fmaxf(1.f, rintf(0.4f * k0)),
                                 fmaxf(1.f, rintf(0.55f * k0)), fmaxf(1.f, rintf(0.75f * k0)));
  const float lks[4] = {lk4.x, lk4.y, lk4.z, lk4.w};
  // the four fields start unreached
  const unsigned long long none2 = ((unsigned long long)LBW_NONE << 32) | LBW_NONE;
  for (int j = tid; j < c.n; j += blockDim.x) {
    c.pset[qb + j] = make_int2((int)LBW_NONE, (int)LBW_NONE);
    c.cand_head[qb + j] = none2;
  }
  __threadfence();
  __syncthreads();
  int rounds = 0, rr = 0;
  wg_search<4>(c, q, LSCALE * delta_len + lk4.x * delta_delay, L, relaxed, rr, 0.f, lk4);
  rounds += rr;
  __threadfence();
  __syncthreads();
  if (wid < 4) {
    double wl = 0.0, wd = 0.0;
    const float kw = wid == 0 ? lk4.x : wid == 1 ? lk4.y : wid == 2 ? lk4.z : lk4.w;
    const uint32_t ws = *grid_field(c, qb + s, wid);
    const bool ok = ws < LBW_NONE && walk_tight<4>(c, q, kw, wl, wd, wid);
    if (lane_id() == 0) {
      s_gok[wid] = ok ? 1 : 0;
      s_glb[wid] = ws < LBW_NONE ? ((double)ws - (double)kw * (double)bnd) / (double)LSCALE : -1e300;
      s_gd[wid] = (float)wd;
      s_gl[wid] = (float)wl;
    }
  }
  __syncthreads();
  if (tid == 0) {
    int jb = -1;
    for (int j = 0; j < 4; j++) {
      if (s_glb[j] > ls.best_lb) {
        ls.best_lb = s_glb[j];
        ls.best_k = lks[j];
        jb = j;
      }
      if (s_gok[j]) {
        if (s_gd[j] > bnd) {
          if (s_gd[j] < ls.p1d) {
            ls.p1d = s_gd[j];
            ls.p1l = s_gl[j];
          }
        } else if (s_gd[j] > ls.p0d) {
          ls.p0d = s_gd[j];
          ls.p0l = s_gl[j];
        }
      }
    }
    ls.last_k = jb;  // here: the grid field holding best_k (-1: none)
    float kh = 0.f;
    if (ls.p1d > ls.p0d && ls.p0l > ls.p1l) kh = hull_multiplier(ls);
    for (int j = 0; j < 4; j++)
      if (kh == lks[j]) kh = 0.f;
    s_kh = kh;
  }
  __syncthreads();
  const float kh = s_kh;
  float* lq = c.lbw + qb;
  bool from_grid = true;
  if (kh > 0.f) {
    // hull step into lbw (still unreached from the batch reset)
    wg_search<3>(c, q, LSCALE * delta_len + kh * delta_delay, L, relaxed, rr, kh);
    rounds += rr;
    __threadfence();
    __syncthreads();
    if (tid == 0) {
      const uint32_t ws = __float_as_uint(lq[s]);
      const double lb = ws < LBW_NONE ? ((double)ws - (double)kh * (double)bnd) / (double)LSCALE : -1e300;
      if (lb >= ls.best_lb) {
        ls.best_lb = lb;
        ls.best_k = kh;
        ls.last_k = -1.f;
      }
    }
    __syncthreads();
    from_grid = ls.last_k >= 0.f;
  }
  const int jb = (int)ls.last_k;
  // lbw from the chosen grid field (or cleared when the hull step is kept but
  // no bound was found), then the label search's words back to their reset
  // values
  for (int j = tid; j < c.n; j += blockDim.x) {
    if (from_grid) lq[j] = jb >= 0 ? __uint_as_float(*grid_field(c, qb + j, jb)) : __uint_as_float(LBW_NONE);
    c.pset[qb + j] = make_int2(0, 0);
    c.cand_head[qb + j] = ~0ull;
  }
  if (tid == 0 && from_grid && jb < 0) ls.best_k = 0.f;
  __syncthreads();
  return rounds;
}

__global__ void __launch_bounds__(BOUND_BLOCK) k_bounds(Ctx c, int nq, float delta_delay, float delta_len) {
  __shared__ BoundLds L;
  if ((int)blockIdx.x >= nq) return;
  if (threadIdx.x == 0) L.stamp = 1;
  const int q = c.q_order ? c.q_order[blockIdx.x] : (int)blockIdx.x;
  int relaxed = 0, r0 = 0, r1 = 0, r2 = 0;
  if (c.ellipse) {
    wg_search<2>(c, q, delta_delay, L, relaxed, r2);
    // later reads of fwd must not hit lines cached while it was written
    __threadfence();
    __syncthreads();
  }
  wg_search<0>(c, q, delta_delay, L, relaxed, r0);
  __syncthreads();
  __threadfence();
  __syncthreads();
  wg_search<1>(c, q, delta_len, L, relaxed, r1);
  int r3 = 0;
  if (c.lagrange) {
    __shared__ LagrState ls;
    __threadfence();
    __syncthreads();
    if (threadIdx.x < WAVE) lagr_begin(c, q, ls);
    __syncthreads();
    float* lq = c.lbw + (long long)q * c.n;
    if (c.lagrange == LAGR_GRID) {
      if (!ls.stop) r3 = lagr_grid(c, q, ls, delta_len, delta_delay, L, relaxed);
      if (threadIdx.x == 0) c.q_lk[q] = ls.stop ? 0.f : ls.best_k;
      wave_add(&c.ct->sssp_relaxed, relaxed);
      if (threadIdx.x == 0) atomicMax(&c.ct->sround, r0 + r1 + r2 + r3);
      return;
    }
    for (int it = 0; it < c.lagrange && !ls.stop; it++) {
      const float k = ls.k;
      if (it > 0) {
        // the last search's lbw is replaced
        for (int j = threadIdx.x; j < c.n; j += blockDim.x) lq[j] = __uint_as_float(LBW_NONE);
        __threadfence();
        __syncthreads();
      }
      int rr = 0;
      wg_search<3>(c, q, LSCALE * delta_len + k * delta_delay, L, relaxed, rr, k);
      r3 += rr;
      __threadfence();
      __syncthreads();
      if (threadIdx.x < WAVE) lagr_step(c, q, ls);
      __syncthreads();
    }
    const float bk = ls.best_k;
    if (bk != ls.last_k) {
      // lbw in memory belongs to another multiplier: cleared, and searched
      // again under the best one if there is one
      for (int j = threadIdx.x; j < c.n; j += blockDim.x) lq[j] = __uint_as_float(LBW_NONE);
      __threadfence();
      __syncthreads();
      if (bk > 0.f) {
        int rr = 0;
        wg_search<3>(c, q, LSCALE * delta_len + bk * delta_delay, L, relaxed, rr, bk);
        r3 += rr;
      }
    }
    if (threadIdx.x == 0) c.q_lk[q] = bk;
  } else if (threadIdx.x == 0) {
    c.q_lk[q] = 0.f;
  }
  wave_add(&c.ct->sssp_relaxed, relaxed);
  if (threadIdx.x == 0) atomicMax(&c.ct->sround, r0 + r1 + r2 + r3);
}

// ---------------------------------------------------------------------------
// Append a carried entry to bucket fb of query q's far tier: lock-free on the
// bucket's chunk list (as the candidate lists).  Entries go into the head
// chunk while it has room.  The entry that takes slot 3/4 of a head chunk
// also links an empty successor twice its size in front, so that the many
// waves appending to one bucket move on to it instead of each allocating a
// chunk of their own when the head fills up.  Entries that still find the
// head full (or the bucket empty) retry once on a newer head, then are
// grouped by bucket within the wave, and one lane per bucket allocates a
// chunk just large enough for them.  Called by the lanes with an entry to
// park (any subset of the wave).  The caller counts the entry in far_cnt[q].
__device__ __forceinline__ int far_fit_class(int n) {
  int k = 0;
  while ((FU << k) < n && k < FCLS_MAX) k++;
  return k;
}

// LOCAL: only the calling workgroup appends to and reads this query's far
// tier (the workgroup engine), so its own order suffices; otherwise waves on
// every CU do, and the records must reach memory before the link.
template <bool LOCAL>
__device__ __forceinline__ void far_link(const Ctx& c, int* headp, int q, int fb, int blk, int cls, int count,
                                         int expected) {
  const int head = (blk << 4) | cls;
  for (;;) {
    c.fhd[blk] = make_int2(count, expected);
    // records and header before the link
    if (LOCAL) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    else __threadfence();
    const int prev = atomicCAS(headp, expected, head);
    if (prev == expected) break;
    expected = prev;
  }
  // a bucket becomes non-empty only through a CAS from -1
  if (expected < 0) atomicOr(&c.far_occ[(long long)q * NBW + (fb >> 6)], 1ull << (fb & 63));
}

template <bool LOCAL = false>
__device__ __forceinline__ void far_append(const Ctx& c, Counters* ct, int q, int fb, const int4& crec4) {
  int* headp = &c.far_head[(long long)q * NBK + fb];
  int hv = __hip_atomic_load(headp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  bool done = false;
  for (int attempt = 0; attempt < 2 && hv >= 0 && !done; attempt++) {
    const int blk = hv >> 4, cls = hv & 15;
    const int cap = FU << cls;
    const int slot = atomicAdd(&c.fhd[blk].x, 1);
    if (slot < cap) {
      c.frec[(long long)blk * FU + slot] = crec4;
      done = true;
      if (slot == cap - cap / 4 && cls < FCLS_MAX) {
        // successor of this chunk, linked empty
        const int ncls = cls + 1;
        const int nb = atomicAdd(&ct->far_chunks, 1 << ncls);
        if (nb < 0 || nb > c.far_cap - (1 << ncls)) atomicOr(&ct->overflow, OVF_FAR);
        else far_link<LOCAL>(c, headp, q, fb, nb, ncls, 0, hv);
      }
    } else {
      const int hv2 = __hip_atomic_load(headp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (hv2 == hv) break;
      hv = hv2;
    }
  }
  const int lane = lane_id();
  unsigned long long pend = __ballot(!done);
  while (pend) {
    const int leader = __ffsll((long long)pend) - 1;
    const int lq = __shfl(q, leader, WAVE), lfb = __shfl(fb, leader, WAVE);
    const bool mine = !done && q == lq && fb == lfb;
    const unsigned long long mm = __ballot(mine);
    const int n_same = __popcll(mm);
    const int ncls = far_fit_class(n_same);
    int blk = 0;
    if (lane == leader) blk = atomicAdd(&ct->far_chunks, 1 << ncls);  // bump allocator, reset per batch
    blk = __shfl(blk, leader, WAVE);
    if (blk < 0 || blk > c.far_cap - (1 << ncls)) {
      if (lane == leader) atomicOr(&ct->overflow, OVF_FAR);
    } else {
      if (mine) c.frec[(long long)blk * FU + __popcll(mm & ((1ull << lane) - 1ull))] = crec4;
      if (lane == leader) far_link<LOCAL>(c, headp, q, fb, blk, ncls, n_same, hv);
    }
    if (mine) done = true;
    pend &= ~mm;
  }
}

// Append a group of candidate records for one (query, vertex) to its HBM
// list: `gm` holds the group's lanes (those with ok set), `lead` one of them,
// `hw` the head word the lead read.  Everything appenders of one kernel share
// lives in the 64-bit head word (head chunk << 32 | records claimed in it), so
// no fence is needed: the lead claims slots in the head chunk for the whole
// group with one 64-bit atomicAdd on the word, and records that do not fit go
// into new chunks that one CAS publishes as the new head (the chunk links and
// the replaced head's final count are plain stores, read by merge after the
// kernel boundary).  A list that was empty is recorded as touched.
// Wave-uniform call.
constexpr unsigned long long HEAD_EMPTY = ~0ull;
__device__ __forceinline__ int head_chunk(unsigned long long w) { return (int)(uint32_t)(w >> 32); }
__device__ __forceinline__ int head_count(unsigned long long w) { return (int)min((uint32_t)w, (uint32_t)CCHUNK); }

__device__ __forceinline__ void append_group(const Ctx& c, Counters* ct, int cur, bool ok, long long qv,
                                             unsigned long long hw, int lead, unsigned long long gm,
                                             const int4& rec) {
  const int lane = lane_id();
  const int gn = __popcll(gm);
  const int rank = __popcll(gm & ((1ull << lane) - 1ull));
  const int hidx = __shfl(head_chunk(hw), lead, WAVE);
  int s0 = CCHUNK;  // first slot the group claimed in the head chunk
  int hcur = -1;
  if (ok && lane == lead && hidx >= 0) {
    const unsigned long long old = atomicAdd(&c.cand_head[qv], (unsigned long long)gn);
    hcur = head_chunk(old);  // the head now (it never returns to empty within a round)
    s0 = (int)min((uint32_t)old, (uint32_t)CCHUNK);
  }
  s0 = __shfl(s0, lead, WAVE);
  hcur = __shfl(hcur, lead, WAVE);
  const int nfit = min(max(CCHUNK - s0, 0), gn);
  if (ok && rank < nfit) c.crec[(long long)hcur * CCHUNK + s0 + rank] = rec;
  const int rest = gn - nfit;  // records for new chunks
  bool first = false;
  int nc = -1;
  if (ok && lane == lead && rest > 0) {
    const int k = (rest + CCHUNK - 1) / CCHUNK;
    nc = atomicAdd(&ct->cand_ctr[cur], k);
    if (nc < 0 || nc > c.chunk_cap - k) {
      atomicOr(&ct->overflow, OVF_CAND);
      nc = -1;
    } else {
      // chunks nc .. nc+k-1, full but the last; the last is the new head
      // and links back to nc, which links to the list as it was.  A head
      // chunk's count lives in the head word: its header's count is written
      // once, by the wave that replaces it as head (one writer per field, as
      // plain stores from different XCDs reach memory in no fixed order)
      const int last_cnt = rest - (k - 1) * CCHUNK;
      for (int j = 1; j < k - 1; j++) c.chd[nc + j] = make_int2(CCHUNK, nc + j - 1);
      if (k > 1) c.chd[nc + k - 1].y = nc + k - 2;
      const unsigned long long neww = ((unsigned long long)(uint32_t)(nc + k - 1) << 32) | (uint32_t)last_cnt;
      unsigned long long expected =
          __hip_atomic_load(&c.cand_head[qv], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      for (;;) {
        if (k == 1) c.chd[nc].y = head_chunk(expected);
        else c.chd[nc] = make_int2(CCHUNK, head_chunk(expected));
        const unsigned long long prev = atomicCAS(&c.cand_head[qv], expected, neww);
        if (prev == expected) break;
        expected = prev;  // another wave published or claimed first: retry on its word
      }
      // the replaced head's final count, for merge
      if (head_chunk(expected) >= 0) c.chd[head_chunk(expected)].x = head_count(expected);
      first = head_chunk(expected) < 0;
    }
  }
  nc = __shfl(nc, lead, WAVE);
  if (ok && rank >= nfit && nc >= 0) {
    const int j = rank - nfit;
    c.crec[(long long)(nc + j / CCHUNK) * CCHUNK + j % CCHUNK] = rec;
  }
  const int tix = wave_alloc32(&ct->touched_n[cur], first);
  if (first) {
    if (tix < c.chunk_cap) c.touched[tix] = qv;
    else atomicOr(&ct->overflow, OVF_TOUCH);
  }
}

// First-level candidate table in LDS, the counterpart of the reference's
// per-block candidate hash (algorithm/bucket.cuh:29-41, 79-119): candidates of
// one workgroup for the same (query, vertex) meet in a slot holding a small
// Pareto set, so candidates dominated by another one of the workgroup never
// reach HBM.  A slot is claimed by a 64-bit CAS on its key (linear probing,
// HASH_PROBES probes), updated under a per-slot LDS lock; a candidate that
// finds no slot, a full slot or a lock held too long is appended to HBM
// directly (the overflow path).  The slots are flushed to the HBM lists at the
// end of the kernel.
constexpr int HASH_K = 4;          // candidates kept per slot
constexpr int HASH_PROBES = 8;     // linear probes before a candidate overflows
constexpr int HASH_SPIN = 64;      // lock attempts before a candidate overflows
constexpr int HASH_MAX_SLOTS = 1024;
constexpr unsigned long long HASH_EMPTY = ~0ull;
struct HashSlot {
  unsigned long long key;  // (query, vertex) index, HASH_EMPTY = free
  int lock;
  int cnt;
  int4 ent[HASH_K];        // {d bits, delay bits, parent label, 0}
};

__device__ __forceinline__ int hash_slot_of(unsigned long long key, int slots) {
  return (int)(((key * 0x9E3779B97F4A7C15ull) >> 40) & (unsigned long long)(slots - 1));
}

// The slot of `key` (claimed if free), or -1 after HASH_PROBES probes.
__device__ __forceinline__ int hash_find(HashSlot* T, int slots, unsigned long long key) {
  int h = hash_slot_of(key, slots);
  for (int p = 0; p < HASH_PROBES; p++) {
    unsigned long long k = T[h].key;
    if (k == HASH_EMPTY) k = atomicCAS(&T[h].key, HASH_EMPTY, key);
    if (k == HASH_EMPTY || k == key) return h;
    h = (h + 1) & (slots - 1);
  }
  return -1;
}

// Offer candidate x to slot h: 1 = kept or dominated (handled in LDS),
// 0 = the slot is full or busy (the caller appends x to HBM).  *dropped is set
// when x was dominated by a slot entry.
template <typename Label_t>
__device__ __forceinline__ int hash_offer(HashSlot& S, int vid, const int4& x, bool* dropped) {
  using LT = label_traits<Label_t>;
  int spin = 0;
  while (atomicCAS(&S.lock, 0, 1) != 0)
    if (++spin >= HASH_SPIN) return 0;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  const Label_t xl = LT::unpack(vid, make_float2(__int_as_float(x.x), __int_as_float(x.y)));
  int n = S.cnt, res = 1;
  bool dom = false;
  for (int j = 0; j < n && !dom; j++)
    dom = Label_t::dominate(LT::unpack(vid, make_float2(__int_as_float(S.ent[j].x), __int_as_float(S.ent[j].y))), xl);
  if (dom) {
    *dropped = true;
  } else {
    int w = 0;  // entries x dominates leave the slot
    for (int j = 0; j < n; j++) {
      const int4 e = S.ent[j];
      if (!Label_t::dominate(xl, LT::unpack(vid, make_float2(__int_as_float(e.x), __int_as_float(e.y)))))
        S.ent[w++] = e;
    }
    if (w < HASH_K) S.ent[w++] = x;
    else res = 0;
    S.cnt = w;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  atomicExch(&S.lock, 0);
  return res;
}

// Expand (PRUNE: VINE_PRUNE_*).  Every frontier entry is checked; the live
// ones with key <= the query's smallest key + window are expanded, the rest
// carried to the near or far tier.  The window plays the role of the
// reference's DOP count (expand the last DOP entries, carry the rest,
// algorithm/cspp.cuh:179-216) in key units.
template <typename Label_t, typename Constraint_t, typename Edge_t, int PRUNE>
__global__ void __launch_bounds__(MAX_BLOCK) k_expand(Ctx c, int lpw) {
  using LT = label_traits<Label_t>;
  Counters* ct = c.ct;
  const int r = ct->round_e;
  const int cur = r & 1, nxt = cur ^ 1;
  const int ncarry = min(ct->carry_n[cur], c.carry_cap);
  // input: near-tier carry records, then the far segments pull queued (64
  // lanes each, some empty), then the labels merge inserted last round
  const long long nd64 = (long long)min(ct->drain_n, c.far_cap) * FSEG;
  const long long nold = (long long)ncarry + nd64;
  const long long lstart = ct->lstart[cur];
  const long long lend = min((long long)ct->label_ctr, c.label_cap);
  const long long fn = nold + (lend - lstart);
  // count semantics: only the last dop_count entries may be expanded
  const long long first_exp = (c.dop_count >= 0 && fn > c.dop_count) ? fn - c.dop_count : 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ct->lstart[nxt] = lend;
    ct->round_m = r;
    c.ts[r & (TS_CAP - 1)] = wall_clock64();  // start of round r
  }
  const int lane = lane_id();
  // per-wave tables grouping a step's candidates by (query, vertex)
  __shared__ int s_lead[MAX_WAVES][GROUP_SLOTS];
  __shared__ unsigned long long s_mask[MAX_WAVES][GROUP_SLOTS];
  for (int j = threadIdx.x; j < MAX_WAVES * GROUP_SLOTS; j += blockDim.x) (&s_mask[0][0])[j] = 0ull;
  // first-level candidate table (dynamic LDS of hash_slots slots); the
  // learned policy sizes its active part from the expansions it predicts for
  // this round: about two slots per candidate of a workgroup
  extern __shared__ __align__(16) unsigned char dyn_smem[];
  HashSlot* T = (HashSlot*)dyn_smem;
  int hslots = c.hash_slots;
  if (hslots > 0 && c.hash_mlp) {
    const float pred = c.hash_pred[r & 1];
    if (pred >= 0.f) {
      const float per_wg = pred * c.avg_deg / (float)gridDim.x;
      int want = 16;
      while ((float)want < 2.f * per_wg && want < hslots) want <<= 1;
      hslots = want;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) c.hash_pred[(r + 1) & 1] = 0.f;
  }
  for (int j = threadIdx.x; j < hslots; j += blockDim.x) {
    T[j].key = HASH_EMPTY;
    T[j].lock = 0;
    T[j].cnt = 0;
  }
  __syncthreads();
  const int wid = threadIdx.x / WAVE;
  const long long nwaves = (long long)gridDim.x * (blockDim.x / WAVE);
  int n_exp = 0, n_rel = 0, n_cand = 0, n_far = 0, n_near = 0, n_scan = 0, n_hdrop = 0, n_hover = 0;
  const uint32_t* fmin_cur = c.q_fmin + (size_t)cur * c.B;
  uint32_t* fmin_nxt = c.q_fmin + (size_t)nxt * c.B;

  // lpw frontier entries per wave step (lanes >= lpw only help with the arcs):
  // fewer entries per wave spread a heavy round's arcs over more waves
  for (long long base = ((long long)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE) * lpw;
       base < fn; base += nwaves * lpw) {
    const long long i = lane < lpw ? base + lane : fn;
    int id = -1, deg = 0, row = 0, q = 0;
    Label_t lab;
    bool carry = false, expd = false;
    float key = 0.f, rd = 0.f;
    if (i < fn) {
      // carried entries bring their query, key and pruning delay along, so a
      // label that stays outside the window costs one 16-byte read per round
      bool fresh = i >= nold;
      if (!fresh) {
        int4 cr = make_int4(-1, 0, 0, 0);
        if (i < ncarry) {
          cr = c.carry[cur][i];
        } else {
          const long long j = i - ncarry;
          const int2 sg = c.dl[j / FSEG];  // {first unit, records}
          const int l = (int)(j % FSEG);
          if (l < sg.y) cr = c.frec[(long long)sg.x * FU + l];
        }
        id = cr.x;
        q = cr.y;
        key = __int_as_float(cr.z);
        rd = __int_as_float(cr.w);
      } else {
        id = (int)(lstart + (i - nold));
      }
      if (id >= 0) {
        // A tail round is a chain of dependent loads, so the loads an entry
        // needs are issued as soon as their addresses are known.  A fresh
        // label (made by the last merge) is known only by its id: its record
        // and meta, then everything by query and vertex.  A carried entry
        // brings its query, key and pruning delay, so its window and pruning
        // checks need only per-query state; its label and row are read only
        // when it is expanded (most carried entries are carried again, and
        // their label loads would be wasted).
        const bool dead = c.dead[id] != 0;
        int4 L = make_int4(0, 0, 0, 0);
        int2 meta = make_int2(q, __float_as_int(key));
        if (fresh) {
          L = c.labs[id];
          meta = c.lmeta[id];
        }
        // an id reserved by merge but never filled is dead and holds stale
        // contents: its query and vertex are used only clamped into range,
        // for loads whose values are then ignored
        const int qq = min(max(meta.x, 0), c.B - 1);
        const float2 b = c.q_best[qq];
        const float lim = __uint_as_float(fmin_cur[qq]) + c.q_delta[qq];
        int qdst = 0, row0 = 0, row1 = 0;
        float qbound = 0.f, lbdv = 0.f;
        if (fresh) {
          const int vv = min(max(L.x, 0), c.n - 1);
          qdst = c.q_dst[qq];
          qbound = c.q_bound[qq];
          if (PRUNE == 2) lbdv = lbd_of(c, (long long)qq * c.n + vv);
          row0 = c.row_ptr[vv];
          row1 = c.row_ptr[vv + 1];
        }
        bool live = !dead;
        if (live) {
          q = meta.x;
          key = __int_as_float(meta.y);
        }
        if (live && fresh) {
          lab = LT::unpack(L.x, make_float2(__int_as_float(L.y), __int_as_float(L.z)));
          rd = lab.delay;
          if (PRUNE != 0 && lab.vid == qdst) live = false;  // answers are not extended
          if (live && PRUNE == 2) {
            rd = lab.delay + lbdv;
            if (rd > qbound) live = false;
          }
        }
        // (key, rd) is (d, delay) without bounds and (d + lbl, delay + lbd) with them
        if (live && PRUNE != 0 && lex_ge(key, rd, b.x, b.y)) live = false;
        if (live) {
          c.q_last[q] = r;
          if (!(key > lim) && i >= first_exp) {
            if (!fresh) {
              L = c.labs[id];
              lab = LT::unpack(L.x, make_float2(__int_as_float(L.y), __int_as_float(L.z)));
              const int vv = min(max(L.x, 0), c.n - 1);
              row0 = c.row_ptr[vv];
              row1 = c.row_ptr[vv + 1];
            }
            row = row0;
            deg = row1 - row0;
            n_exp++;
            expd = true;
          } else {
            carry = true;
          }
        }
      }
    }
    if (c.stats_on) {
      wave_add_keyed(c.q_exp, q, expd);
    }
    // carried labels: near tier if within far_mult windows of the front,
    // else the query's far tier (not scanned again until pulled)
    bool far = false;
    int fb = 0;
    if (carry) {
      fb = far_bucket(c, q, key);
      far = fb >= c.far_next[q];
    }
    const bool near = carry && !far;
    int cslot = wave_alloc32(&ct->carry_n[nxt], near);
    const int4 crec4 = make_int4(id, q, __float_as_int(key), __float_as_int(rd));
    if (c.stats_on) wave_add_keyed(c.q_near, q, near);
    if (near) {
      if (cslot < c.carry_cap) c.carry[nxt][cslot] = crec4;
      else atomicOr(&ct->overflow, OVF_CARRY);
    }
    // one atomic per query of the wave for the front key and the far count
    wave_min_keyed(fmin_nxt, q, __float_as_uint(key), near);
    wave_add_keyed(c.far_cnt, q, far);
    if (far) far_append(c, ct, q, fb, crec4);
    n_far += far ? 1 : 0;
    n_near += near ? 1 : 0;
    n_scan += i < fn ? 1 : 0;
    float ld = lab.d, ldl = lab.delay;
    // the owner's query bound and best answer travel with the label
    const float obound = (i < fn && id >= 0) ? c.q_bound[q] : 0.f;
    const float olk = (PRUNE == 2 && i < fn && id >= 0) ? c.q_lk[q] : 0.f;
    const float2 obest = (PRUNE != 0 && i < fn && id >= 0) ? c.q_best[q] : make_float2(0.f, 0.f);
    wave_flat(deg, row, [&](int owner, int pos, bool act) {
      // shuffles by every lane (a lane reads its owner's registers)
      Label_t src = LT::unpack(0, make_float2(__shfl(ld, owner, WAVE), __shfl(ldl, owner, WAVE)));
      const int oq = __shfl(q, owner, WAVE);
      const int oid = __shfl(id, owner, WAVE);
      const float2 b = make_float2(__shfl(obest.x, owner, WAVE), __shfl(obest.y, owner, WAVE));
      const float bound = __shfl(obound, owner, WAVE);
      const float lk = PRUNE == 2 ? __shfl(olk, owner, WAVE) : 0.f;
      bool ok = false;
      Label_t son;
      long long qv = 0;
      unsigned long long hw = HEAD_EMPTY;
      if (act) {
        n_rel++;
        int4 e = c.edges[pos];
        Edge_t ed(__int_as_float(e.y), __int_as_float(e.z));
        src.expand(e.x, ed, &son);
        Constraint_t C(bound);
        ok = C.pass_constraints_check(son);
        qv = (long long)oq * c.n + son.vid;
        // everything the checks and the append read at (query, vertex), at once
        const int2 ps = c.pset[qv];
        hw = __hip_atomic_load(&c.cand_head[qv], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const float lbdv = PRUNE == 2 ? lbd_of(c, qv) : 0.f;
        const float2 lw = PRUNE == 2 ? lbl_pair(c, qv) : make_float2(0.f, FINF);
        if (ok && PRUNE == 1 && son.vid != c.q_dst[oq] && lex_ge(son.d, son.delay, b.x, b.y)) ok = false;
        if (ok && PRUNE == 2) {
          const float rd = son.delay + lbdv;
          if (rd > C.delay_L || lex_ge(lower_key(son.d, son.delay, lw, lk, bound), rd, b.x, b.y)) ok = false;
        }
        // read-only dominance check against the vertex's front (no writer
        // runs now); candidates of this round are checked against each
        // other in merge
        if (ok && front_dominates<Label_t>(c.pool_cd + (uint32_t)ps.x, pset_cnt(ps), son.vid, son)) ok = false;
      }
      // The wave's candidates for one (query, vertex) form a group (several
      // labels of one vertex reach the same neighbours in the same step).
      int lead = lane;
      unsigned long long gm = 1ull << lane;
      wave_group_by_key(s_lead[wid], s_mask[wid], ok, qv, lane, lead, gm);
      int4 rec = make_int4(0, 0, -1, 0);
      if (ok) {
        const float2 pk = LT::pack(son);
        rec = make_int4(__float_as_int(pk.x), __float_as_int(pk.y), oid, 0);
      }
      if (hslots > 0) {
        // first level: the group's lead finds the slot, then its members are
        // offered to it one at a time (members of one group share a slot;
        // other groups' slots are distinct)
        int slot = -1;
        if (ok && lane == lead) slot = hash_find(T, hslots, (unsigned long long)qv);
        slot = __shfl(slot, lead, WAVE);
        const int rank = __popcll(gm & ((1ull << lane) - 1ull));
        int gmax = ok ? __popcll(gm) : 0;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) gmax = max(gmax, __shfl_xor(gmax, o, WAVE));
        bool held = false, dropped = false;
        for (int k = 0; k < gmax; k++)
          if (ok && slot >= 0 && rank == k) held = hash_offer<Label_t>(T[slot], son.vid, rec, &dropped) != 0;
        n_hdrop += dropped ? 1 : 0;
        n_hover += (ok && !held) ? 1 : 0;
        // the rest go to HBM now, as a group
        ok = ok && !held;
        gm = __ballot(ok) & gm;
        lead = gm ? __ffsll((long long)gm) - 1 : lane;
      }
      n_cand += ok ? 1 : 0;
      append_group(c, ct, cur, ok, qv, hw, lead, gm, rec);
    });
  }
  if (hslots > 0) {
    // flush the table: 16 slots per wave step, lanes 4j .. 4j+3 holding slot
    // j's entries, one group per slot
    __syncthreads();
    const int nw = blockDim.x / WAVE;
    for (int base = wid * (WAVE / HASH_K); base < hslots; base += nw * (WAVE / HASH_K)) {
      const int sl = base + lane / HASH_K, e = lane % HASH_K;
      bool valid = false;
      long long qv = 0;
      int4 rec = make_int4(0, 0, -1, 0);
      unsigned long long hw = HEAD_EMPTY;
      if (sl < hslots) {
        const unsigned long long key = T[sl].key;
        if (key != HASH_EMPTY && e < T[sl].cnt) {
          valid = true;
          qv = (long long)key;
          rec = T[sl].ent[e];
          hw = __hip_atomic_load(&c.cand_head[qv], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
      const unsigned long long gm = __ballot(valid) & (((1ull << HASH_K) - 1ull) << (lane & ~(HASH_K - 1)));
      const int lead = gm ? __ffsll((long long)gm) - 1 : lane;
      n_cand += valid ? 1 : 0;
      append_group(c, ct, cur, valid, qv, hw, lead, gm, rec);
    }
  }
  unsigned long long* const ctrs[9] = {&ct->expanded,     &ct->relaxed,      &ct->cands,
                                       (unsigned long long*)&ct->far_live,   &ct->far_appended,
                                       &ct->near_carried, &ct->scanned,      &ct->hash_dropped,
                                       &ct->hash_overflow};
  const long long vals[9] = {n_exp, n_rel, n_cand, n_far, n_far, n_near, n_scan, n_hdrop, n_hover};
  block_add<9>(ctrs, vals);
}

// ---------------------------------------------------------------------------
// Merge: fold each touched (query, vertex)'s candidates into its Pareto front,
// kept sorted by d ascending (delay strictly descending).  A group of MG lanes
// owns one (query, vertex); for each candidate (broadcast from the lane that
// loaded its chunk record) the group locates its position (the last entry
// first, then MG entries per step backwards), decides "dominated" from the
// entry before it, counts the entries it dominates (a contiguous run after
// it) and replaces that run by the candidate, shifting the tail MG entries a
// step.  The front grows into storage of twice its size when full.
// Candidate loops are wave-uniform, predicates group-uniform, so label and
// pool allocations aggregate over the whole wave.
constexpr int MG = 16;  // lanes per (query, vertex) group


template <typename Label_t, int PRUNE>
__global__ void __launch_bounds__(MAX_BLOCK) k_merge(Ctx c) {
  using LT = label_traits<Label_t>;
  Counters* ct = c.ct;
  const int r = ct->round_m;
  const int cur = r & 1, nxt = cur ^ 1;
  const int tn = min(ct->touched_n[cur], c.chunk_cap);
  // a list never holds more chunks than this round allocated; the walk is cut
  // there so that state left by an overflowed round cannot loop
  const int max_steps = min(ct->cand_ctr[cur], c.chunk_cap);
  const long long gtid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gtid == 0) {
    ct->touched_n[nxt] = 0;
    ct->cand_ctr[nxt] = 0;
    ct->carry_n[cur] = 0;
    ct->drain_n = 0;  // expand of this round has scanned the segments
    ct->round_e = r + 1;  // expand of this round has finished reading it
  }
  for (long long q = gtid; q < c.B; q += (long long)gridDim.x * blockDim.x)
    c.q_fmin[(size_t)cur * c.B + q] = __float_as_uint(FINF);
  uint32_t* fmin_nxt = c.q_fmin + (size_t)nxt * c.B;
  const int lane = lane_id();
  const int gl = lane & (MG - 1);
  const int gbase = lane & ~(MG - 1);
  const unsigned long long gmask = ((1ull << MG) - 1ull) << gbase;
  constexpr int GPW = WAVE / MG;  // groups per wave
  const long long nwaves = (long long)gridDim.x * (blockDim.x / WAVE);
  int n_ins = 0, scanned = 0, steps_max = 0, n_touch = 0, n_chunks = 0;

  for (long long base = ((long long)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE) * GPW;
       base < tn; base += nwaves * GPW) {
    const long long ti = base + lane / MG;
    const bool have = ti < tn;
    long long qv = 0;
    int q = 0, v = 0, h = -1, cnt = 0, cap = 0, tq = -1;
    unsigned long long hw = HEAD_EMPTY;
    uint32_t off = 0;
    float2 lb = make_float2(0.f, FINF);
    float qlk = 0.f, qbnd = 0.f;
    if (have) {
      qv = c.touched[ti];
      q = (int)(qv / c.n);
      v = (int)(qv - (long long)q * c.n);
      hw = c.cand_head[qv];
      h = head_chunk(hw);
      int2 ps = c.pset[qv];
      off = (uint32_t)ps.x;
      cnt = pset_cnt(ps);
      cap = pset_cap(ps);
      tq = c.q_dst[q];
      if (PRUNE == 2) {
        lb = lbl_pair(c, qv);
        qlk = c.q_lk[q];
        qbnd = c.q_bound[q];
      }
    }
    if (have && gl == 0) c.cand_head[qv] = HEAD_EMPTY;
    // the front's last entry stays in registers across the candidates
    float2 lastv = (have && cnt > 0) ? c.pool_cd[off + cnt - 1] : make_float2(-1.f, FINF);
    // smallest key and count of the labels this group inserts, published once
    uint32_t gkey = __float_as_uint(FINF);
    int gins = 0;
    int steps = 0;
    // The list is walked one chunk ahead: the next chunk's header and records
    // are in flight while this chunk's label ids are allocated and its
    // candidates folded.  A chunk's records are read with its header, not
    // after it (lanes past the count read slots that are ignored).  Chunk
    // links always name a chunk of this round or -1.
    int2 hd = make_int2(0, -1);
    int4 hrec = make_int4(0, 0, -1, 0);
    if (h >= 0) {
      hd = c.chd[h];
      hd.x = head_count(hw);  // the head chunk's count lives in the head word
      if (gl < CCHUNK) hrec = c.crec[(long long)h * CCHUNK + gl];
    }
    while (__ballot(h >= 0)) {
      if (h >= 0 && ++steps > max_steps) {
        atomicOr(&ct->overflow, OVF_CAND);
        h = -1;
      }
      int n8 = 0, nexth = -1;
      int4 myrec = make_int4(0, 0, -1, 0);
      if (h >= 0) {
        myrec = hrec;
        n8 = min(hd.x, CCHUNK);
        nexth = hd.y;
        n_chunks += gl == 0 ? 1 : 0;
      }
      int kmax = n8;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) kmax = max(kmax, __shfl_xor(kmax, o, WAVE));
      // label ids for the whole chunk in one wave-wide allocation; the ids
      // of candidates that do not survive are marked dead below
      long long rbase = wave_alloc64_n(&ct->label_ctr, gl == 0 ? n8 : 0);
      hd = make_int2(0, -1);
      hrec = make_int4(0, 0, -1, 0);
      if (nexth >= 0) {
        hd = c.chd[nexth];
        if (gl < CCHUNK) hrec = c.crec[(long long)nexth * CCHUNK + gl];
      }
      rbase = __shfl(rbase, gbase, WAVE);
      int used = 0;
      for (int k = 0; k < kmax; k++) {
        const bool act = k < n8;  // group-uniform
        const float cd = __int_as_float(__shfl(myrec.x, gbase + k, WAVE));
        const float cdl = __int_as_float(__shfl(myrec.y, gbase + k, WAVE));
        const int par = __shfl(myrec.z, gbase + k, WAVE);
        const Label_t cl = LT::unpack(v, make_float2(cd, cdl));
        // The set is a Pareto front sorted by d ascending, so delay strictly
        // descends.  p = entries with d <= cd; the candidate is dominated iff
        // entry p-1 has delay <= cdl.  Children reach a vertex in nearly
        // ascending d (the window follows d + lbl, and lbl is fixed per
        // vertex), so the last entry usually settles it.
        int p = 0;
        float2 prev = make_float2(-1.f, FINF);
        if (act && cnt > 0) {
          const float2 last = lastv;
          if (last.x <= cd) {
            p = cnt;
            prev = last;
          } else {
            // group scan backwards, MG entries per step (sorted: the entries
            // with d <= cd are a prefix of each block)
            int hi = cnt - 1;
            for (;;) {
              const int lo = max(0, hi - MG);
              const int j = lo + gl;
              const bool in = j < hi;
              const float2 e = in ? c.pool_cd[off + j] : make_float2(FINF, FINF);
              const int np = __popcll(__ballot(in && e.x <= cd) & gmask);
              scanned++;
              if (np > 0 || lo == 0) {
                p = lo + np;
                break;
              }
              hi = lo;
            }
            if (p > 0) prev = c.pool_cd[off + p - 1];
          }
        }
        const bool acc = act && !(p > 0 && Label_t::dominate(LT::unpack(v, prev), cl));
        // entries the candidate dominates: from `start` (entry p-1 when it has
        // the same d and a larger delay) those with delay >= cdl, a prefix
        int start = p, kdom = 0;
        if (acc) {
          if (p > 0 && prev.x == cd) start = p - 1;
          for (int lo = start; lo < cnt; lo += MG) {
            const int j = lo + gl;
            const bool in = j < cnt;
            const float2 e = in ? c.pool_cd[off + j] : make_float2(FINF, -FINF);
            const bool dm = in && Label_t::dominate(cl, LT::unpack(v, e));
            if (dm) c.dead[c.pool_id[off + j]] = 1;
            const int nd = __popcll(__ballot(dm) & gmask);
            kdom += nd;
            scanned++;
            if (nd < MG) break;
          }
        }
        const int newcnt = cnt - kdom + 1;
        const bool grow = acc && newcnt > cap;
        int newcap = 0;
        if (grow) {
          newcap = INIT_CAP;
          while (newcap < 2 * newcnt) newcap <<= 1;
        }
        long long poff = -1;
        if (__ballot(grow && gl == 0)) {
          poff = wave_alloc64_n(&ct->pool_ctr, (grow && gl == 0) ? newcap : 0);
          poff = __shfl(poff, gbase, WAVE);
        }
        const long long lid = rbase + used;
        bool fits = acc;
        if (grow && (poff < 0 || (unsigned long long)(poff + newcap) > c.pool_cap || poff + newcap > 0xFFFFFFFFll)) {
          if (gl == 0) atomicOr(&ct->overflow, OVF_POOL);
          fits = false;
        }
        if (fits && lid >= c.label_cap) {
          if (gl == 0) atomicOr(&ct->overflow, OVF_LABEL);
          fits = false;
        }
        if (fits) {
          if (grow) {
            // new storage: [0, start), the candidate, [start + kdom, cnt)
            for (int j = gl; j < cnt; j += MG) {
              if (j < start || j >= start + kdom) {
                const int dj = j < start ? j : j - kdom + 1;
                c.pool_cd[poff + dj] = c.pool_cd[off + j];
                c.pool_id[poff + dj] = c.pool_id[off + j];
              }
            }
            off = (uint32_t)poff;
            cap = newcap;
          } else if (kdom == 0 && start < cnt) {
            // open slot `start`: shift [start, cnt) up by one, last block
            // first (each block is read whole before it is written)
            for (int hi = cnt; hi > start; hi -= MG) {
              const int lo = max(start, hi - MG);
              const int j = lo + gl;
              if (j < hi) {
                const float2 e = c.pool_cd[off + j];
                const int32_t id = c.pool_id[off + j];
                c.pool_cd[off + j + 1] = e;
                c.pool_id[off + j + 1] = id;
              }
            }
          } else if (kdom > 1) {
            // close the gap: shift [start + kdom, cnt) down by kdom - 1
            for (int lo = start + kdom; lo < cnt; lo += MG) {
              const int j = lo + gl;
              if (j < cnt) {
                const float2 e = c.pool_cd[off + j];
                const int32_t id = c.pool_id[off + j];
                c.pool_cd[off + j - kdom + 1] = e;
                c.pool_id[off + j - kdom + 1] = id;
              }
            }
          }
          if (gl == 0) {
            n_ins++;
            c.pool_cd[off + start] = make_float2(cd, cdl);
            c.pool_id[off + start] = (int)lid;
            const float key = PRUNE == 2 ? lower_key(cd, cdl, lb, qlk, qbnd) : LT::key(cl);
            c.labs[lid] = make_int4(v, __float_as_int(cd), __float_as_int(cdl), par);
            c.lmeta[lid] = make_int2(q, __float_as_int(key));
            c.dead[lid] = 0;
            if (!(cd < EXACT_LIMIT) || !(cdl < EXACT_LIMIT)) ct->range_flag = 1;
            if (v == tq) {
              float2 b = c.q_best[q];
              if (cd < b.x || (cd == b.x && cdl < b.y)) {
                c.q_best[q] = make_float2(cd, cdl);
                c.q_best_id[q] = (int)lid;
              }
            }
            if (!(PRUNE != 0 && v == tq)) gkey = min(gkey, __float_as_uint(key));
          }
          gins++;
          if (start + 1 == newcnt) lastv = make_float2(cd, cdl);
          cnt = newcnt;
          used++;
        }
      }
      // reserved ids no candidate took: dead, so the next expand skips them
      if (gl >= used && gl < n8 && rbase + gl < c.label_cap) c.dead[rbase + gl] = 1;
      h = nexth;
    }
    if (have && gl == 0) c.pset[qv] = pset_make(off, cnt, cap);
    // the groups of a wave often belong to one query: one atomic per query
    const bool lead = have && gl == 0;
    wave_min_keyed(fmin_nxt, q, gkey, lead && gkey != __float_as_uint(FINF));
    if (c.stats_on) {
      wave_add_keyed_n(c.q_ins, q, gins, lead && gins > 0);
      wave_add_keyed_n(c.q_near, q, gins, lead && gins > 0);
    }
    steps_max = max(steps_max, steps);
    n_touch += (have && gl == 0) ? 1 : 0;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) steps_max = max(steps_max, __shfl_xor(steps_max, o, WAVE));
  if (lane == 0 && steps_max > 0 && steps_max > load_agent_u32((const uint32_t*)&ct->merge_chain_max))
    atomicMax(&ct->merge_chain_max, steps_max);
  unsigned long long* const ctrs[4] = {&ct->touched_total, &ct->merge_scanned, &ct->inserted, &ct->cand_chunks};
  const long long vals[4] = {n_touch, scanned, n_ins, n_chunks};
  block_add<4>(ctrs, vals);
}

// ---------------------------------------------------------------------------
// Pull: between merge of round r and expand of round r+1.  For each query the
// buckets whose key range comes within reach of the front (front + reach) are
// handed whole to the next expand; if the near tier is empty, the next
// non-empty bucket is.  A far entry is therefore moved at most once, and
// buckets whose keys the upper bound prunes are dropped without being read.
// The pull itself only chases chunk headers (chunks double in size along a
// bucket, so even a dense bucket is a short list) and queues each chunk as
// drain segments of up to 64 records; the records are read by the next expand across
// the whole grid, where dead and pruned ones are dropped.  An occupancy
// bitmap (one bit per bucket) finds the non-empty buckets of a key range
// 4,096 buckets per load step.  One wave per query.
template <int PRUNE>
__device__ void pull_query(const Ctx& c, long long q) {
  Counters* ct = c.ct;
  const int rn = ct->round_e;  // the round that follows
  const int nc = rn & 1;
  uint32_t* fmin = c.q_fmin + (size_t)nc * c.B;
  const int lane = lane_id();
  // the query's state in one step of loads (a pull is a chain of dependent
  // reads: state, occupancy words, bucket heads, chunk headers)
  const int nb = c.far_next[q];
  const float front = __uint_as_float(load_agent_u32(&fmin[q]));
  const float qdelta = c.q_delta[q];
  const int fcnt = c.far_cnt[q];
  const float2 best = c.q_best[q];
  const float fbase = c.far_base[q], fbw = c.far_bw[q];
  if (nb >= NBK) return;
  const float reach = fmaxf(c.far_mult * qdelta, c.far_floor);
  if (fcnt <= 0) {
    // nothing parked: only keep the near/far boundary up with the front
    if (lane == 0 && front < FINF) c.far_next[q] = max(nb, far_bucket(c, (int)q, front + reach) + 1);
    return;
  }
  if (PRUNE != 0 && fbase + (float)nb * fbw > best.x) {
    // every parked key exceeds t's best answer: the whole far tier is
    // pruned without reading it (its chunks are reclaimed at the next batch)
    if (lane == 0) {
      const int left = fcnt;
      c.far_cnt[q] = 0;
      c.far_next[q] = NBK;
      atomicAdd((unsigned long long*)&ct->far_live, (unsigned long long)(-(long long)left));
    }
    return;
  }
  unsigned long long* occ = c.far_occ + q * NBW;
  int target;
  if (front < FINF) {
    target = far_bucket(c, (int)q, front + reach);
  } else {
    // near tier empty: the first non-empty bucket, if any (64 buckets a word)
    target = -1;
    for (int w0 = nb >> 6; w0 < NBW && target < 0; w0 += WAVE) {
      const int w = w0 + lane;
      unsigned long long bits = 0;
      if (w < NBW) {
        bits = occ[w];
        if (w == (nb >> 6)) bits &= ~0ull << (nb & 63);
      }
      const unsigned long long m = __ballot(bits != 0);
      if (m) {
        const int l = __ffsll((long long)m) - 1;
        const unsigned long long wb = __shfl(bits, l, WAVE);
        target = ((w0 + l) << 6) + __ffsll((long long)wb) - 1;
      }
    }
    if (target < 0) return;
  }
  if (target < nb) return;
  // walk the chunk headers of the occupied buckets of [nb, target] and queue
  // their blocks as drain segments: the next expand scans them like carried
  // entries, spread over the whole grid.  Only headers are chased here.
  long long queued = 0;
  const int wlast = target >> 6;
  for (int w0 = nb >> 6; w0 <= wlast; w0 += WAVE) {
    // 64 occupancy words at a time; their bits are cleared here (expand
    // appends only to buckets >= the new far_next)
    const int w = w0 + lane;
    unsigned long long bits = 0;
    if (w <= wlast) {
      const unsigned long long word = occ[w];
      bits = word;
      if (w == (nb >> 6)) bits &= ~0ull << (nb & 63);
      if (w == wlast && (target & 63) != 63) bits &= (1ull << ((target & 63) + 1)) - 1ull;
      if (bits) occ[w] = word & ~bits;
    }
    unsigned long long lanes = __ballot(bits != 0);
    while (lanes) {
      const int l = __ffsll((long long)lanes) - 1;
      lanes &= lanes - 1;
      unsigned long long wb = __shfl(bits, l, WAVE);
      // the word's 64 bucket heads in one load; drained heads reset at once
      int* hrow = c.far_head + q * NBK + ((long long)(w0 + l) << 6);
      const int myhead = hrow[lane];
      if ((wb >> lane) & 1ull) hrow[lane] = -1;
      while (wb) {
        const int bit = __ffsll((long long)wb) - 1;
        wb &= wb - 1;
        int hv = __shfl(myhead, bit, WAVE);
        int guard = 0;
        while (hv >= 0 && guard++ < c.far_cap) {
          const int blk = hv >> 4;
          const int2 hd = c.fhd[blk];
          const int cnt = min(hd.x, FU << (hv & 15));
          const int nseg = (cnt + FSEG - 1) / FSEG;
          int base = 0;
          if (lane == 0) base = atomicAdd(&ct->drain_n, nseg);
          base = __shfl(base, 0, WAVE);
          for (int k = lane; k < nseg; k += WAVE) {
            if (base + k < c.far_cap) c.dl[base + k] = make_int2(blk + k * (FSEG / FU), min(FSEG, cnt - k * FSEG));
            else atomicOr(&ct->overflow, OVF_FAR);
          }
          queued += cnt;
          hv = hd.y;
        }
      }
    }
  }
  if (lane == 0) {
    c.far_next[q] = target + 1;
    if (queued) {
      c.far_cnt[q] -= (int)queued;
      c.q_last[q] = rn;
      if (c.stats_on) c.q_near[q] += (int)queued;
      atomicAdd((unsigned long long*)&ct->far_live, (unsigned long long)(-queued));
      atomicAdd(&ct->far_pulled, (unsigned long long)queued);
    }
  }
}

// ---------------------------------------------------------------------------
// Scheduler step, after pull: per query, the 8 features of the state the next
// round starts from, optionally recorded as a trace row, and (VINE_DELTA_MLP)
// the next window width from the MLP:
//   x0 log2(1 + near-tier entries queued this round)
//   x1 log2(1 + key range left between the front and t's best bound)
//   x2 log2(1 + labels expanded this round)
//   x3 log2(1 + labels inserted this round)
//   x4 progress of the front from the smallest key to t's best bound, in [0, 1.5]
//   x5 log2(1 + round)
//   x6 log2(1 + labels the query has expanded so far)
//   x7 1 if an answer has been found
// Output 0 gives the next window as 2^y0, clamped to the range the network is
// trained on, [2^MLP_LOG2_MIN, 2^MLP_LOG2_MAX]; output 1
// is a forecast of log2(1 + next round's expansions).
__device__ void policy_query(const Ctx& c, int q, int nq) {
  Counters* ct = c.ct;
  const int lane = lane_id();
  const int rn = ct->round_e;
  const int nc = rn & 1;
  // per-query counters: read and reset by lane 0 (it also wrote q_near and
  // far_cnt in the pull just before), broadcast to the wave
  int exp_ = 0, ins = 0, near = 0, far = 0;
  long long tot = 0;
  if (lane == 0) {
    exp_ = c.q_exp[q];
    ins = c.q_ins[q];
    near = c.q_near[q];
    far = c.far_cnt[q];
    c.q_exp[q] = 0;
    c.q_ins[q] = 0;
    c.q_near[q] = 0;
    tot = c.q_exptot[q] + exp_;
    c.q_exptot[q] = tot;
  }
  exp_ = __shfl(exp_, 0, WAVE);
  ins = __shfl(ins, 0, WAVE);
  near = __shfl(near, 0, WAVE);
  far = __shfl(far, 0, WAVE);
  tot = __shfl(tot, 0, WAVE);
  const bool active = q < nq && (near + far + exp_) > 0;
  if (!active) return;
  const float fmin = __uint_as_float(c.q_fmin[(size_t)nc * c.B + q]);
  const float base = c.far_base[q];
  const float2 best = c.q_best[q];
  float prog = 1.5f;
  if (fmin < FINF && best.x < FINF && best.x > base) prog = fminf(fmaxf((fmin - base) / (best.x - base), 0.f), 1.5f);
  float x[mlp::NIN];
  x[0] = log2f(1.f + (float)near);
  x[1] = (fmin < FINF && best.x < FINF) ? log2f(1.f + fmaxf(best.x - fmin, 0.f)) : 0.f;
  x[2] = log2f(1.f + (float)exp_);
  x[3] = log2f(1.f + (float)ins);
  x[4] = prog;
  x[5] = log2f(1.f + (float)rn);
  x[6] = log2f(1.f + (float)tot);
  x[7] = c.q_best_id[q] >= 0 ? 1.f : 0.f;
  if (lane == 0 && c.trace && c.trace_every > 0 && rn % c.trace_every == 0) {
    long long row = atomicAdd((unsigned long long*)&ct->trace_n, 1ull);
    if (row < c.trace_cap) {
      float* out = c.trace + row * 10;
      out[0] = (float)q;
      out[1] = (float)rn;
#pragma unroll
      for (int j = 0; j < mlp::NIN; j++) out[2 + j] = x[j];
    }
  }
  if (c.policy == 3 && c.mlp) {
    float y[mlp::NOUT];
    mlp::apply_wave(c.mlp, x, y);
    float l = y[0];
    if (!(l == l)) l = log2f(fmaxf(c.far_floor, 1.f));  // NaN output: fall back to a small window
    l = fminf(fmaxf(l, MLP_LOG2_MIN), MLP_LOG2_MAX);
    if (lane == 0) {
      c.q_delta[q] = exp2f(l);
      c.q_fcast[q] = y[1];
      // the second output forecasts log2(1 + expansions): their sum over the
      // queries sizes the next round's first-level candidate table
      if (c.hash_mlp && y[1] == y[1]) atomicAdd(&c.hash_pred[rn & 1], fmaxf(exp2f(fminf(y[1], 30.f)) - 1.f, 0.f));
    }
  }
}

// After merge: per query (one wave each) the far-tier pull, then the policy
// step when statistics are on (learned window or trace).
template <int PRUNE>
__global__ void __launch_bounds__(MAX_BLOCK) k_pull_policy(Ctx c, int nq) {
  const long long q = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
  if (q >= c.B) return;  // whole waves exit together
  if (c.policy != 0) pull_query<PRUNE>(c, q);
  if (c.stats_on) policy_query(c, (int)q, nq);
}

// ---------------------------------------------------------------------------
// Answers and paths.
__global__ void k_answers(Ctx c, int nq, float* cost, float* delay, int32_t* feasible,
                          int32_t* plen) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  int id = c.q_best_id[q];
  float2 b = c.q_best[q];
  if (cost) cost[q] = id >= 0 ? b.x : FINF;
  if (delay) delay[q] = id >= 0 ? b.y : FINF;
  if (feasible) feasible[q] = id >= 0 ? 1 : 0;
  if (plen) {
    int h = 0;
    for (int x = id; x >= 0; x = c.labs[x].w) h++;
    plen[q] = h;
  }
}

// Paths are written in the caller's vertex ids (vinv: internal -> caller).
__global__ void k_paths(Ctx c, int nq, const long long* off, int32_t* path) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  long long o = off[q], e = off[q + 1];
  long long k = e - 1;
  for (int x = c.q_best_id[q]; x >= 0 && k >= o; x = c.labs[x].w) path[k--] = c.vinv[c.labs[x].x];
}

__global__ void k_stamp_start(Ctx c) { c.ct->t_start = wall_clock64(); }
__global__ void k_stamp_round(Ctx c, int r) { c.ts[r & (TS_CAP - 1)] = wall_clock64(); }

// Query endpoints from the caller's vertex ids to internal ones.
__global__ void k_map_ids(const int32_t* in, const int32_t* perm, int32_t* out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = perm[in[i]];
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

}  // namespace dev
}  // namespace vine
