// Workgroup engine of the label search (VINE_PRUNE_BOUNDS with a fixed or
// learned window): one workgroup runs one query slot's whole search, round
// after round, separated by workgroup barriers instead of grid-wide launches.
// A short query's workgroup retires early and its CU takes the next query; a
// long one never waits for the others.  Per round, inside the workgroup:
//   pull    -- wave 0 moves the far-tier buckets within reach of the front
//              (or, with an empty near tier, the first occupied one) into
//              this round's near list, whole chunks at a time, stopping at a
//              chunk that does not fit (the rest of its bucket stays far);
//   expand  -- the near list: dead / bound / best checks, in-window labels
//              reserve LDS arena room for their arcs (labels that do not fit
//              wait a round) and expand them (wave_flat over the arcs, the
//              same filters and sorted-front pre-check as k_expand); each
//              candidate goes into an LDS per-vertex hash table: open
//              addressing on the vertex id, the slot heading a linked list
//              of arena records;
//   merge   -- 16-lane groups over the touched vertices merge each
//              vertex's candidates into its sorted Pareto front a chunk at a
//              time (every pair compared in registers, survivors placed by
//              d, the changed tail written back once; a front of FC entries
//              or more takes the rest one at a time in the pool); live new
//              labels join the next near list;
//   policy  -- wave 0: the query's 8 features, an optional trace row and,
//              for VINE_DELTA_MLP, the next window from the network.
// All per-query state (near lists, fronts, label records of the query, far
// tier, best) is written and read by this workgroup alone, i.e. through one
// CU's L1.  Atomics execute in L2 and do not refresh a line another wave
// holds in L1, so every value updated by atomics during the search (far
// heads, chunk counts, the occupancy bitmap, far counts) is read with
// agent-scope loads and updated only by atomics.
#pragma once

#include "solver_kernels.hpp"

namespace vine {
namespace dev {

constexpr int WG_BLOCK = 512;
constexpr int HSLOTS = 2048;  // LDS hash slots (vertices touched per round <= ARENA < HSLOTS)
constexpr int ARENA = 1024;   // LDS candidate records per round; arcs of a label never split
constexpr int WG_ROUND_GUARD = 1 << 21;  // rounds of one query (real searches need ~1e4)
constexpr int WG_GROUPS = WG_BLOCK / MG;  // merge groups of a workgroup
constexpr int FC = 64;        // front + candidates of one vertex merged at once (registers of a merge group)
constexpr int FCC = 64;       // candidates of one vertex staged in LDS for that merge
// Label ids are handed out from a block reserved per query (one global atomic
// per block instead of one per inserted label).  A block is refilled at the
// end of a round once fewer than ARENA ids are left (a round inserts at most
// ARENA labels), with at least LID_MIN ids and about as many as the query has
// used so far (at most LID_MAX); the ids left in a replaced or final block
// stay unused.
constexpr int LID_MIN = 2 * ARENA;
constexpr int LID_MAX = 1 << 16;
// Pareto-pool entries likewise: a block per query, refilled at the end of a
// round once fewer than POOL_LOW entries are left (POOL_MIN, then about as
// many as the query has reserved so far, at most POOL_MAX); an allocation the
// block cannot take goes to the global counter.
constexpr int POOL_LOW = 4096;
constexpr long long POOL_MIN = 8192;
constexpr long long POOL_MAX = 1 << 20;

struct WgLds {
  int key[HSLOTS];    // vertex id, -1 = free
  int head[HSLOTS];   // first arena record of the vertex's candidate list, -1 = none
  int touched[ARENA]; // slots claimed this round
  int4 rec[ARENA];    // {d bits, delay bits, parent label, next record}
  int n_arena;
  int n_touched;
  int n_near[2];
  uint32_t fmin[2];   // smallest key of each near list (float bits)
  float delta;
  float2 best;
  int best_id;
  int n_exp, n_ins;
  int more;           // far tier still holds records (decided by thread 0)
  long long exp_tot;
  long long lid_base;  // label ids [lid_base + lid_n, lid_base + lid_lim) are free
  long long lid_total; // ids reserved so far by this query
  int lid_n, lid_lim;
  long long pool_base;  // pool entries [pool_base + pool_n, pool_base + pool_lim) are free
  long long pool_total; // entries reserved so far by this query
  int pool_n, pool_lim;
  float2 fcd[WG_GROUPS][FCC];  // merge: a vertex's candidates staged in LDS, (d, delay) ...
  int fid[WG_GROUPS][FCC];     // ... and parent labels
};

// Reserve the next block of label ids (thread 0, between phases).
__device__ __forceinline__ void wg_lid_refill(const Ctx& c, Counters* ct, WgLds& S) {
  const long long want = min(max((long long)LID_MIN, S.lid_total), (long long)LID_MAX);
  const long long base = (long long)atomicAdd(&ct->label_ctr, (unsigned long long)want);
  S.lid_n = 0;
  if (base + want > c.label_cap) {
    atomicOr(&ct->overflow, OVF_LABEL);
    S.lid_lim = 0;
    return;
  }
  S.lid_base = base;
  S.lid_lim = (int)want;
  S.lid_total += want;
}

// Reserve the next block of pool entries (thread 0, between phases).
__device__ __forceinline__ void wg_pool_refill(const Ctx& c, Counters* ct, WgLds& S) {
  const long long want = min(max(POOL_MIN, S.pool_total), POOL_MAX);
  const long long base = (long long)atomicAdd(&ct->pool_ctr, (unsigned long long)want);
  S.pool_n = 0;
  S.pool_lim = 0;
  if (base < 0 || (unsigned long long)(base + want) > c.pool_cap || base + want > 0xFFFFFFFFll) return;
  S.pool_base = base;
  S.pool_lim = (int)want;
  S.pool_total += want;
}

// n pool entries for one front (one lane): from the query's block, else from
// the global counter (the caller checks that range against the capacity).
__device__ __forceinline__ long long wg_pool_alloc(Counters* ct, WgLds& S, int n) {
  const int o = atomicAdd(&S.pool_n, n);
  if (o + n <= S.pool_lim) return S.pool_base + o;
  return (long long)atomicAdd(&ct->pool_ctr, (unsigned long long)n);
}

__device__ __forceinline__ int wg_hash(int v) { return (int)(((uint32_t)v * 2654435761u) >> 21); }

// 64-bit agent-scope load of a chunk header (its count is bumped by atomics).
__device__ __forceinline__ int2 load_hdr(const int2* p) {
  const unsigned long long u =
      __hip_atomic_load((const unsigned long long*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return make_int2((int)(uint32_t)u, (int)(uint32_t)(u >> 32));
}

// Push a carried or new entry: into the next near list, or past the near
// boundary (or when the list is full) into the far tier.
__device__ __forceinline__ void wg_carry(const Ctx& c, Counters* ct, int q, WgLds& S, int4* near_nxt, int nxt,
                                         bool pred, const int4& rec, float key) {
  const int nfn = c.far_next[q];
  int fb = 0;
  bool to_far = false;
  if (pred) {
    fb = far_bucket(c, q, key);
    to_far = fb >= nfn;
  }
  const int slot = wave_alloc32(&S.n_near[nxt], pred && !to_far);
  if (pred && !to_far) {
    if (slot < c.wn_cap) {
      near_nxt[slot] = rec;
      if (__float_as_uint(key) < S.fmin[nxt]) atomicMin(&S.fmin[nxt], __float_as_uint(key));
    } else {
      to_far = true;  // list full: the boundary bucket takes it (pulled next)
      fb = max(fb, nfn);
    }
  }
  // far_cnt counts the wave's far entries with one atomic
  const unsigned long long mf = __ballot(pred && to_far && nfn < NBK);
  if (mf && lane_id() == __ffsll((long long)mf) - 1) atomicAdd(&c.far_cnt[q], __popcll(mf));
  if (pred && to_far) {
    if (nfn < NBK) far_append<true>(c, ct, q, fb, rec);
    else atomicOr(&ct->overflow, OVF_CARRY);
  }
}

// Pull, run by wave 0 only (the other waves wait at the barrier after it).
template <int PRUNE>
__device__ void wg_pull(const Ctx& c, Counters* ct, int q, WgLds& S, int4* near_cur, int cur) {
  const int lane = lane_id();
  const int nb = c.far_next[q];
  if (nb >= NBK) return;
  const float front = __uint_as_float(S.fmin[cur]);
  const float reach = fmaxf(c.far_mult * S.delta, c.far_floor);
  const int left = __hip_atomic_load(&c.far_cnt[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (left <= 0) {
    if (lane == 0 && front < FINF) c.far_next[q] = max(nb, far_bucket(c, q, front + reach) + 1);
    return;
  }
  const float2 best = S.best;
  if (PRUNE != 0 && c.far_base[q] + (float)nb * c.far_bw[q] > best.x) {
    // every parked key exceeds t's best answer: dropped without reading
    if (lane == 0) {
      atomicSub(&c.far_cnt[q], left);
      c.far_next[q] = NBK;
    }
    return;
  }
  unsigned long long* occ = c.far_occ + (long long)q * NBW;
  int target = -1;
  if (front < FINF) {
    target = far_bucket(c, q, front + reach);
  } else {
    for (int w0 = nb >> 6; w0 < NBW && target < 0; w0 += WAVE) {
      const int w = w0 + lane;
      unsigned long long bits = 0;
      if (w < NBW) {
        // bits are set by atomics (in L2) during this kernel: read past L1
        bits = __hip_atomic_load(&occ[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (w == (nb >> 6)) bits &= ~0ull << (nb & 63);
      }
      const unsigned long long m = __ballot(bits != 0);
      if (m) {
        const int l = __ffsll((long long)m) - 1;
        const unsigned long long wb = __shfl(bits, l, WAVE);
        target = ((w0 + l) << 6) + __ffsll((long long)wb) - 1;
      }
    }
  }
  if (target < nb) return;
  int moved = 0;     // records taken out of the far tier (live or not)
  int next_b = target + 1;
  bool stop = false;
  for (int w = nb >> 6; w <= (target >> 6) && !stop; w++) {
    unsigned long long bits = __hip_atomic_load(&occ[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (w == (nb >> 6)) bits &= ~0ull << (nb & 63);
    if (w == (target >> 6) && (target & 63) != 63) bits &= (1ull << ((target & 63) + 1)) - 1ull;
    while (bits && !stop) {
      const int b = (w << 6) + __ffsll((long long)bits) - 1;
      bits &= bits - 1;
      int* headp = c.far_head + (long long)q * NBK + b;
      int hv = __hip_atomic_load(headp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      int guard = 0;
      while (hv >= 0 && guard++ < c.far_cap) {
        const int blk = hv >> 4;
        const int2 hd = load_hdr(&c.fhd[blk]);
        const int cnt = min(hd.x, FU << (hv & 15));
        if (cnt > (int)c.wn_cap - min(S.n_near[cur], (int)c.wn_cap)) {
          // this chunk does not fit: the rest of the bucket stays far
          stop = true;
          break;
        }
        for (int k = lane; k < cnt; k += WAVE) {
          const int4 rec = c.frec[(long long)blk * FU + k];
          const float key = __int_as_float(rec.z), rd = __int_as_float(rec.w);
          const bool live = !c.dead[rec.x] && !(PRUNE != 0 && lex_ge(key, rd, best.x, best.y));
          const int slot = wave_alloc32(&S.n_near[cur], live);
          if (live) {
            near_cur[slot] = rec;
            if (__float_as_uint(key) < S.fmin[cur]) atomicMin(&S.fmin[cur], __float_as_uint(key));
          }
        }
        moved += cnt;
        hv = hd.y;
      }
      if (lane == 0) *headp = hv;  // -1 once the bucket is empty
      if (hv >= 0) {
        stop = true;
        next_b = b;  // partly drained: the bucket stays the near/far boundary
      } else if (lane == 0) {
        atomicAnd(&occ[w], ~(1ull << (b & 63)));
      }
    }
  }
  if (lane == 0) {
    c.far_next[q] = max(nb, next_b);
    if (moved) atomicSub(&c.far_cnt[q], moved);
  }
}

// Fold all k candidates of vertex v into its Pareto front at once.  Entries
// x < cnt are the front in the pool at off (sorted by d, delay strictly
// falling), entries x >= cnt the candidates, staged in LDS at E/I[x - cnt]
// in list order (parent label in I); m = cnt + k <= FC, k <= FCC.  Lane gl holds
// entries gl + MG*r in registers.  The result
// is the sequential one (candidates folded in list order, a front entry or
// an earlier candidate winning a tie): an entry is dropped iff another one
// is componentwise <= it and differs from it or precedes it, since the
// dominance of two-criterion labels is a componentwise order.  Candidates
// dropped this way get no label.  Survivors are placed by d (distinct among
// them), the changed tail is written back to the pool (a new allocation if
// the front outgrew it), new labels get ids from the query's block and join
// the next near list.  off, cnt and cap return the new front's header.
constexpr int NSL = FC / MG;

template <typename Label_t>
__device__ __forceinline__ void wg_merge_batch(const Ctx& c, Counters* ct, WgLds& S, int q, int v, long long qv,
                                               int tq, float2 lb, float lbdv, float bnd, float lk, float2 best,
                                               int4* near_nxt, int nxt, const float2* E, int32_t* I, uint32_t& off,
                                               int& cnt, int& cap, int k, int gl, int gbase,
                                               unsigned long long gmask, int& my_ins, int& n_scan) {
  const int m = cnt + k;
  const int nsl = (m + MG - 1) / MG;
  // front entries straight from the pool (all loads issued together),
  // candidates from the stage
  float xd[NSL], xl[NSL];
  int32_t xid[NSL];  // front: label id; candidate: parent label
  bool alive[NSL];
#pragma unroll
  for (int r = 0; r < NSL; r++) {
    const int x = gl + MG * r;
    alive[r] = x < m;
    xd[r] = FINF;
    xl[r] = FINF;
    xid[r] = 0;
    if (x < cnt) {
      const float2 e = c.pool_cd[off + x];
      xd[r] = e.x;
      xl[r] = e.y;
      xid[r] = c.pool_id[off + x];
    } else if (x < m) {
      const float2 e = E[x - cnt];
      xd[r] = e.x;
      xl[r] = e.y;
      xid[r] = I[x - cnt];
    }
  }
  // every candidate against every entry, in both directions
  for (int j = cnt; j < m; j++) {
    const float2 cj = E[j - cnt];  // the same address for all lanes
    bool hit = false;              // some entry of this lane drops candidate j
#pragma unroll
    for (int r = 0; r < NSL; r++) {
      if (r >= nsl) break;
      const int x = gl + MG * r;
      if (x < m && x != j) {
        const bool same = xd[r] == cj.x && xl[r] == cj.y;
        // x <= cj: drops j unless x is a later candidate of equal value
        if (xd[r] <= cj.x && xl[r] <= cj.y && (!same || x < j)) hit = true;
        // cj <= x: drops x unless x is a front entry or an earlier candidate of equal value
        if (cj.x <= xd[r] && cj.y <= xl[r] && (!same || (x >= cnt && j < x))) alive[r] = false;
      }
    }
    const bool dropped = (__ballot(hit) & gmask) != 0ull;
    if (dropped) {
#pragma unroll
      for (int r = 0; r < NSL; r++)
        if (gl + MG * r == j) alive[r] = false;
    }
  }
  n_scan += k;
  // survivor masks per slot (bit l: entry l + MG*r), front and candidates
  uint32_t fm[NSL], cm[NSL];
  int nf = 0, nc = 0;
#pragma unroll
  for (int r = 0; r < NSL; r++) {
    fm[r] = 0u;
    cm[r] = 0u;
    if (r >= nsl) break;
    const int x = gl + MG * r;
    fm[r] = (uint32_t)((__ballot(alive[r] && x < cnt) & gmask) >> gbase);
    cm[r] = (uint32_t)((__ballot(alive[r] && x >= cnt) & gmask) >> gbase);
    nf += __popc(fm[r]);
    nc += __popc(cm[r]);
  }
  // candidates' survival into the stage (sign bit of the parent field: free,
  // parents are label ids >= 0), then the order statistics: for every held
  // entry the surviving candidates with a smaller d, for every surviving
  // candidate the surviving front entries with a smaller d
  int less[NSL], nfl[NSL];
#pragma unroll
  for (int r = 0; r < NSL; r++) {
    less[r] = 0;
    nfl[r] = 0;
  }
  if (nc > 0) {
#pragma unroll
    for (int r = 0; r < NSL; r++) {
      const int x = gl + MG * r;
      if (x >= cnt && x < m && alive[r]) I[x - cnt] = xid[r] | (int32_t)0x80000000;
    }
    wave_lds_order();
    for (int j = cnt; j < m; j++) {
      const float2 cj = E[j - cnt];
      const bool sj = I[j - cnt] < 0;
      int nfj = 0;
#pragma unroll
      for (int r = 0; r < NSL; r++) {
        if (r >= nsl) break;
        const int x = gl + MG * r;
        if (sj && cj.x < xd[r]) less[r]++;
        nfj += __popc((uint32_t)((__ballot(alive[r] && x < cnt && xd[r] < cj.x) & gmask) >> gbase));
      }
#pragma unroll
      for (int r = 0; r < NSL; r++)
        if (gl + MG * r == j) nfl[r] = nfj;
    }
  }
  const int newcnt = nf + nc;
  // positions, the first changed one, label ranks
  int pos[NSL], crank[NSL];
  int first = newcnt;
  {
    int fbefore = 0, cbefore = 0;
    const uint32_t below = (1u << gl) - 1u;
#pragma unroll
    for (int r = 0; r < NSL; r++) {
      const int x = gl + MG * r;
      const int fr = fbefore + __popc(fm[r] & below);
      crank[r] = cbefore + __popc(cm[r] & below);
      pos[r] = -1;
      if (alive[r]) {
        pos[r] = (x < cnt ? fr : nfl[r]) + less[r];
        if (x >= cnt) first = min(first, pos[r]);
      } else if (x < cnt) {
        first = min(first, x);  // a front entry dropped: everything from it on may move
      }
      fbefore += __popc(fm[r]);
      cbefore += __popc(cm[r]);
    }
  }
#pragma unroll
  for (int o = MG / 2; o > 0; o >>= 1) first = min(first, __shfl_xor(first, o, WAVE));
  if (first >= newcnt && newcnt == cnt) return;  // nothing changed
  // label ids for the surviving candidates
  int lbase = 0;
  if (gl == 0 && nc > 0) lbase = atomicAdd(&S.lid_n, nc);
  lbase = __shfl(lbase, gbase, WAVE);
  if (nc > 0 && lbase + nc > S.lid_lim) {
    if (gl == 0) atomicOr(&ct->overflow, OVF_LABEL);
    return;  // the solve fails
  }
  // a front that outgrew its allocation moves, written whole
  if (newcnt > cap) {
    int newcap = INIT_CAP;
    while (newcap < 2 * newcnt) newcap <<= 1;
    long long poff = 0;
    if (gl == 0) poff = wg_pool_alloc(ct, S, newcap);
    poff = __shfl(poff, gbase, WAVE);
    if (poff < 0 || (unsigned long long)(poff + newcap) > c.pool_cap || poff + newcap > 0xFFFFFFFFll) {
      if (gl == 0) atomicOr(&ct->overflow, OVF_POOL);
      return;
    }
    off = (uint32_t)poff;
    cap = newcap;
    first = 0;
  }
  float2 nb = make_float2(FINF, FINF);
  int nbid = -1;
#pragma unroll
  for (int r = 0; r < NSL; r++) {
    const int x = gl + MG * r;
    bool push = false;
    int4 nrec = make_int4(-1, 0, 0, 0);
    float nkey = 0.f;
    if (x < m) {
      const int32_t idp = xid[r];  // front: label id; candidate: parent
      if (x < cnt) {
        if (!alive[r]) c.dead[idp] = 1;
        else if (pos[r] >= first) {
          c.pool_cd[off + pos[r]] = make_float2(xd[r], xl[r]);
          c.pool_id[off + pos[r]] = idp;
        }
      } else if (alive[r]) {
        const long long lid = S.lid_base + lbase + crank[r];
        c.pool_cd[off + pos[r]] = make_float2(xd[r], xl[r]);
        c.pool_id[off + pos[r]] = (int)lid;
        nkey = lower_key(xd[r], xl[r], lb, lk, bnd);
        c.labs[lid] = make_int4(v, __float_as_int(xd[r]), __float_as_int(xl[r]), idp);
        c.lmeta[lid] = make_int2(q, __float_as_int(nkey));
        c.dead[lid] = 0;
        if (!(xd[r] < EXACT_LIMIT) || !(xl[r] < EXACT_LIMIT)) ct->range_flag = 1;
        if (v == tq) {
          if (xd[r] < nb.x) {
            nb = make_float2(xd[r], xl[r]);
            nbid = (int)lid;
          }
        } else {
          const float nrd = xl[r] + lbdv;
          if (!(nrd > bnd) && !lex_ge(nkey, nrd, best.x, best.y)) {
            push = true;
            nrec = make_int4((int)lid, q, __float_as_int(nkey), __float_as_int(nrd));
          }
        }
      }
    }
    if (r < nsl) wg_carry(c, ct, q, S, near_nxt, nxt, push, nrec, nkey);
  }
  if (v == tq) {
    // the new labels at t: the smallest d among them improves the best answer
#pragma unroll
    for (int o = MG / 2; o > 0; o >>= 1) {
      const float od = __shfl_xor(nb.x, o, WAVE), ol = __shfl_xor(nb.y, o, WAVE);
      const int oid = __shfl_xor(nbid, o, WAVE);
      if (od < nb.x) {
        nb = make_float2(od, ol);
        nbid = oid;
      }
    }
    if (gl == 0 && nbid >= 0) {
      const float2 b = S.best;
      if (nb.x < b.x || (nb.x == b.x && nb.y < b.y)) {
        S.best = nb;
        S.best_id = nbid;
      }
    }
  }
  if (gl == 0) {
    my_ins += nc;
    c.pset[qv] = pset_make(off, newcnt, cap);
  }
  cnt = newcnt;
}

template <typename Label_t, typename Constraint_t, typename Edge_t>
__global__ void __launch_bounds__(WG_BLOCK) k_search(Ctx c, int nq) {
  constexpr int PRUNE = 2;
  using LT = label_traits<Label_t>;
  __shared__ WgLds S;
  if ((int)blockIdx.x >= nq) return;
  const int q = c.q_order ? c.q_order[blockIdx.x] : (int)blockIdx.x;
  Counters* ct = c.ct;
  const int tid = threadIdx.x;
  const int lane = lane_id();
  const int wid = tid / WAVE;
  const long long qb = (long long)q * c.n;
  const int tq = c.q_dst[q];
  const float bnd = c.q_bound[q];
  const float lk = c.q_lk[q];
  int4* nearb[2] = {c.wnear + (2LL * q) * c.wn_cap, c.wnear + (2LL * q + 1) * c.wn_cap};
  for (int i = tid; i < HSLOTS; i += WG_BLOCK) {
    S.key[i] = -1;
    S.head[i] = -1;
  }
  if (tid == 0) {
    S.n_arena = 0;
    S.n_touched = 0;
    S.n_near[0] = 0;
    S.n_near[1] = 0;
    S.fmin[0] = __float_as_uint(FINF);
    S.fmin[1] = __float_as_uint(FINF);
    S.delta = c.q_delta[q];
    S.best = c.q_best[q];
    S.best_id = c.q_best_id[q];
    S.n_exp = 0;
    S.n_ins = 0;
    S.exp_tot = 0;
    S.lid_total = 0;
    wg_lid_refill(c, ct, S);
    S.pool_total = 0;
    wg_pool_refill(c, ct, S);
    // the source label (id q, written by k_init_batch) starts the near list
    const int s = c.q_src[q];
    const float key0 = __int_as_float(c.lmeta[q].y);
    const float rd0 = lbd_of(c, qb + s);
    if (s != tq && key0 < FINF && !(rd0 > bnd) && !lex_ge(key0, rd0, S.best.x, S.best.y)) {
      nearb[0][0] = make_int4(q, q, __float_as_int(key0), __float_as_int(rd0));
      S.n_near[0] = 1;
      S.fmin[0] = __float_as_uint(key0);
    }
  }
  __syncthreads();
  int n_rel = 0, n_cand = 0, n_scan = 0, n_ins_tot = 0, n_exp_tot = 0, scanned = 0;
  int cur = 0;
  int round = 0;
  for (;; round++) {
    const int nxt = cur ^ 1;
    if (wid == 0) wg_pull<PRUNE>(c, ct, q, S, nearb[cur], cur);
    __syncthreads();
    if (round >= WG_ROUND_GUARD) {
      if (tid == 0) atomicOr(&ct->overflow, OVF_CARRY);
      break;
    }
    // with an empty near list the pull takes the first occupied far bucket;
    // if every record it held was dead or pruned the list is still empty and
    // the next round takes the next bucket.  The query is done when both
    // tiers are empty (a dropped or disabled far tier has far_cnt = 0).
    const int ncur = min(S.n_near[cur], (int)c.wn_cap);
    if (ncur == 0) {
      if (tid == 0) S.more = __hip_atomic_load(&c.far_cnt[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > 0;
      // also: every thread has read n_near before the next pull writes it
      __syncthreads();
      if (!S.more) break;
      continue;
    }
    const float lim = __uint_as_float(S.fmin[cur]) + S.delta;
    const float2 best = S.best;  // snapshot: merge may improve it during the round
    int4* near_cur = nearb[cur];
    int4* near_nxt = nearb[nxt];

    // ---------------- expand
    int my_exp = 0;
    for (int base = wid * WAVE; base < ncur; base += WG_BLOCK) {
      const int i = base + lane;
      int4 rec = make_int4(-1, 0, 0, 0);
      if (i < ncur) rec = near_cur[i];
      const float key = __int_as_float(rec.z), rd = __int_as_float(rec.w);
      const bool live = rec.x >= 0 && !c.dead[rec.x] && !lex_ge(key, rd, best.x, best.y);
      const bool inw = live && !(key > lim);
      Label_t lab;
      int row = 0, deg = 0;
      if (inw) {
        const int4 L = c.labs[rec.x];
        lab = LT::unpack(L.x, make_float2(__int_as_float(L.y), __int_as_float(L.z)));
        row = c.row_ptr[lab.vid];
        deg = c.row_ptr[lab.vid + 1] - row;
      }
      // arena room for all arcs of a label, or the label waits a round
      int incl = inw ? deg : 0;
#pragma unroll
      for (int o = 1; o < WAVE; o <<= 1) {
        const int t = __shfl_up(incl, o, WAVE);
        if (lane >= o) incl += t;
      }
      const int total = __shfl(incl, WAVE - 1, WAVE);
      int abase = 0;
      if (total > 0 && lane == 0) abase = atomicAdd(&S.n_arena, total);
      abase = __shfl(abase, 0, WAVE);
      const int my_at = abase + incl - (inw ? deg : 0);
      const bool fit = inw && my_at + deg <= ARENA;
      my_exp += fit ? 1 : 0;
      scanned += (i < ncur) ? 1 : 0;
      wg_carry(c, ct, q, S, near_nxt, nxt, live && !fit, rec, key);
      const float ld = lab.d, ldl = lab.delay;
      const int pid = rec.x;
      wave_flat(fit ? deg : 0, row, [&](int owner, int pos, bool act) {
        const Label_t src = LT::unpack(0, make_float2(__shfl(ld, owner, WAVE), __shfl(ldl, owner, WAVE)));
        const int oid = __shfl(pid, owner, WAVE);
        const int ridx = __shfl(my_at, owner, WAVE) + (pos - __shfl(row, owner, WAVE));
        if (!act) return;
        n_rel++;
        const int4 e = c.edges[pos];
        const Edge_t ed(__int_as_float(e.y), __int_as_float(e.z));
        Label_t son;
        src.expand(e.x, ed, &son);
        const Constraint_t C(bnd);
        if (!C.pass_constraints_check(son)) return;
        const long long qv = qb + son.vid;
        const float srd = son.delay + lbd_of(c, qv);
        if (srd > C.delay_L) return;
        if (lex_ge(lower_key(son.d, son.delay, lbl_pair(c, qv), lk, bnd), srd, best.x, best.y)) return;
        const int2 ps = c.pset[qv];
        if (front_dominates<Label_t>(c.pool_cd + (uint32_t)ps.x, pset_cnt(ps), son.vid, son)) return;
        n_cand++;
        // LDS hash: find or claim the vertex's slot, then link the record
        int h = wg_hash(son.vid);
        for (int probe = 0; probe < HSLOTS; probe++) {
          const int k = S.key[h];
          if (k == son.vid) break;
          if (k == -1) {
            const int prev = atomicCAS(&S.key[h], -1, son.vid);
            if (prev == -1) {
              S.touched[atomicAdd(&S.n_touched, 1)] = h;
              break;
            }
            if (prev == son.vid) break;
          }
          h = (h + 1) & (HSLOTS - 1);
        }
        const float2 pk = LT::pack(son);
        S.rec[ridx] = make_int4(__float_as_int(pk.x), __float_as_int(pk.y), oid, atomicExch(&S.head[h], ridx));
      });
    }
    if (my_exp) atomicAdd(&S.n_exp, my_exp);
    __syncthreads();

    // ---------------- merge: MG-lane groups over the touched vertices.
    // Candidates are merged into a vertex's front a chunk at a time
    // (wg_merge_batch); once the front has FC entries, the rest are folded
    // one at a time in the pool.
    const int nt = S.n_touched;
    int my_ins = 0;
    {
      const int gl = lane & (MG - 1);
      const int gbase = lane & ~(MG - 1);
      const unsigned long long gmask = ((1ull << MG) - 1ull) << gbase;
      const int grp = tid / MG;
      float2* const fcd = S.fcd[grp];
      int32_t* const fid = S.fid[grp];
      for (int g = grp; g < nt; g += WG_GROUPS) {
        const int h = S.touched[g];
        const int v = S.key[h];
        const long long qv = qb + v;
        const int2 ps = c.pset[qv];
        uint32_t off = (uint32_t)ps.x;
        int cnt = pset_cnt(ps), cap = pset_cap(ps);
        const float2 lb = lbl_pair(c, qv);
        const float lbdv = lbd_of(c, qv);
        // the candidates in list order, folded in chunks that fit the
        // registers (front + chunk <= FC) and the stage (chunk <= FCC)
        int kk = S.head[h];
        while (kk >= 0 && cnt < FC) {
          const int room = min(FCC, FC - cnt);
          int k = 0;
          for (; kk >= 0 && k < room; k++) {
            const int4 r = S.rec[kk];
            kk = r.w;
            if ((k & (MG - 1)) == gl) {
              fcd[k] = make_float2(__int_as_float(r.x), __int_as_float(r.y));
              fid[k] = r.z;
            }
          }
          wave_lds_order();
          wg_merge_batch<Label_t>(c, ct, S, q, v, qv, tq, lb, lbdv, bnd, lk, best, near_nxt, nxt, fcd, fid, off, cnt,
                                  cap, k, gl, gbase, gmask, my_ins, n_scan);
          wave_lds_order();  // the stage is reused by the next chunk
        }
        if (kk < 0) continue;
        // a front of FC entries or more: the rest folded one at a time in the pool
        for (int idx = kk; idx >= 0;) {
          const int4 r = S.rec[idx];
          idx = r.w;
          const float cd = __int_as_float(r.x), cdl = __int_as_float(r.y);
          const int par = r.z;
          const Label_t cl = LT::unpack(v, make_float2(cd, cdl));
          int p = 0;
          float2 prev = make_float2(-1.f, FINF);
          if (cnt > 0) {
            const float2 last = c.pool_cd[off + cnt - 1];
            n_scan++;
            if (last.x <= cd) {
              p = cnt;
              prev = last;
            } else {
              int hi = cnt - 1;
              for (;;) {
                const int lo = max(0, hi - MG);
                const int j = lo + gl;
                const bool in = j < hi;
                const float2 e = in ? c.pool_cd[off + j] : make_float2(FINF, FINF);
                const int np = __popcll(__ballot(in && e.x <= cd) & gmask);
                n_scan++;
                if (np > 0 || lo == 0) {
                  p = lo + np;
                  break;
                }
                hi = lo;
              }
              if (p > 0) prev = c.pool_cd[off + p - 1];
            }
          }
          const bool acc = !(p > 0 && Label_t::dominate(LT::unpack(v, prev), cl));
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
              n_scan++;
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
          // groups run divergently here: the leader allocates alone
          long long poff = 0;
          if (grow && gl == 0) poff = wg_pool_alloc(ct, S, newcap);
          poff = __shfl(poff, gbase, WAVE);
          // the label id, from the query's reserved block
          int lslot = 0;
          if (acc && gl == 0) lslot = atomicAdd(&S.lid_n, 1);
          lslot = __shfl(lslot, gbase, WAVE);
          bool fits = acc;
          if (grow && (poff < 0 || (unsigned long long)(poff + newcap) > c.pool_cap || poff + newcap > 0xFFFFFFFFll)) {
            if (gl == 0) atomicOr(&ct->overflow, OVF_POOL);
            fits = false;
          }
          if (fits && lslot >= S.lid_lim) {
            if (gl == 0) atomicOr(&ct->overflow, OVF_LABEL);
            fits = false;
          }
          const long long lid = S.lid_base + lslot;
          bool push = false;
          int4 nrec = make_int4(-1, 0, 0, 0);
          float nkey = 0.f;
          if (fits) {
            if (grow) {
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
              for (int hi = cnt; hi > start; hi -= MG) {
                const int lo = max(start, hi - MG);
                const int j = lo + gl;
                if (j < hi) {
                  const float2 e = c.pool_cd[off + j];
                  const int32_t id = c.pool_id[off + j];
                  c.pool_cd[off + j + 1] = e;
                  c.pool_id[off + j + 1] = id;
                }
                __builtin_amdgcn_wave_barrier();
              }
            } else if (kdom > 1) {
              for (int lo = start + kdom; lo < cnt; lo += MG) {
                const int j = lo + gl;
                if (j < cnt) {
                  const float2 e = c.pool_cd[off + j];
                  const int32_t id = c.pool_id[off + j];
                  c.pool_cd[off + j - kdom + 1] = e;
                  c.pool_id[off + j - kdom + 1] = id;
                }
                __builtin_amdgcn_wave_barrier();
              }
            }
            nkey = lower_key(cd, cdl, lb, lk, bnd);
            if (gl == 0) {
              my_ins++;
              c.pool_cd[off + start] = make_float2(cd, cdl);
              c.pool_id[off + start] = (int)lid;
              c.labs[lid] = make_int4(v, __float_as_int(cd), __float_as_int(cdl), par);
              c.lmeta[lid] = make_int2(q, __float_as_int(nkey));
              c.dead[lid] = 0;
              if (!(cd < EXACT_LIMIT) || !(cdl < EXACT_LIMIT)) ct->range_flag = 1;
              if (v == tq) {
                // only this group handles vertex t this round
                const float2 b = S.best;
                if (cd < b.x || (cd == b.x && cdl < b.y)) {
                  S.best = make_float2(cd, cdl);
                  S.best_id = (int)lid;
                }
              } else {
                const float nrd = cdl + lbdv;
                if (!(nrd > bnd) && !lex_ge(nkey, nrd, best.x, best.y)) {
                  push = true;
                  nrec = make_int4((int)lid, q, __float_as_int(nkey), __float_as_int(nrd));
                }
              }
            }
            cnt = newcnt;
          }
          wg_carry(c, ct, q, S, near_nxt, nxt, push, nrec, nkey);
        }
        if (gl == 0) c.pset[qv] = pset_make(off, cnt, cap);
      }
    }
    if (my_ins) atomicAdd(&S.n_ins, my_ins);
    n_ins_tot += my_ins;
    __syncthreads();

    // ---------------- round end
    for (int i = tid; i < nt; i += WG_BLOCK) {
      const int h = S.touched[i];
      S.key[h] = -1;
      S.head[h] = -1;
    }
    n_exp_tot += my_exp;
    __syncthreads();
    if (wid == 0) {
      // policy step, by wave 0 (the network is evaluated wave-wide)
      const int exp_ = S.n_exp;
      const long long tot = S.exp_tot + exp_;
      const int near_q = min(S.n_near[nxt], (int)c.wn_cap);
      const float fm = __uint_as_float(S.fmin[nxt]);
      const float fbase = c.far_base[q];
      const float2 bb = S.best;
      float prog = 1.5f;
      if (fm < FINF && bb.x < FINF && bb.x > fbase) prog = fminf(fmaxf((fm - fbase) / (bb.x - fbase), 0.f), 1.5f);
      float x[mlp::NIN];
      x[0] = log2f(1.f + (float)near_q);
      x[1] = (fm < FINF && bb.x < FINF) ? log2f(1.f + fmaxf(bb.x - fm, 0.f)) : 0.f;
      x[2] = log2f(1.f + (float)exp_);
      x[3] = log2f(1.f + (float)S.n_ins);
      x[4] = prog;
      x[5] = log2f(1.f + (float)(round + 1));
      x[6] = log2f(1.f + (float)tot);
      x[7] = S.best_id >= 0 ? 1.f : 0.f;
      if (lane == 0 && c.trace && c.trace_every > 0 && (round + 1) % c.trace_every == 0) {
        const long long row = atomicAdd((unsigned long long*)&ct->trace_n, 1ull);
        if (row < c.trace_cap) {
          float* out = c.trace + row * 10;
          out[0] = (float)q;
          out[1] = (float)(round + 1);
#pragma unroll
          for (int j = 0; j < mlp::NIN; j++) out[2 + j] = x[j];
        }
      }
      float ndelta = S.delta;
      if (c.policy == 3 && c.mlp) {
        float y[mlp::NOUT];
        mlp::apply_wave(c.mlp, x, y);
        float l = y[0];
        if (!(l == l)) l = log2f(fmaxf(c.far_floor, 1.f));
        l = fminf(fmaxf(l, MLP_LOG2_MIN), MLP_LOG2_MAX);
        ndelta = exp2f(l);
      }
      if (lane == 0) {
        S.delta = ndelta;
        S.exp_tot = tot;
        S.n_exp = 0;
        S.n_ins = 0;
        S.n_arena = 0;
        S.n_touched = 0;
        S.n_near[cur] = 0;
        S.fmin[cur] = __float_as_uint(FINF);
        // the next round inserts at most ARENA labels
        if (S.lid_lim - S.lid_n < ARENA && S.lid_lim > 0) wg_lid_refill(c, ct, S);
        if (S.pool_lim - S.pool_n < POOL_LOW) wg_pool_refill(c, ct, S);
      }
    }
    __syncthreads();
    cur = nxt;
  }
  // per-query results and statistics
  if (tid == 0) {
    c.q_best[q] = S.best;
    c.q_best_id[q] = S.best_id;
    c.q_last[q] = round - 1;
    c.q_exptot[q] = S.exp_tot;
    c.q_delta[q] = S.delta;
    c.q_fin[q] = wall_clock64();
    atomicMax(&ct->round_e, round);
  }
  wave_add(&ct->expanded, n_exp_tot);
  wave_add(&ct->relaxed, n_rel);
  wave_add(&ct->cands, n_cand);
  wave_add(&ct->inserted, n_ins_tot);
  wave_add(&ct->merge_scanned, n_scan);
  wave_add(&ct->scanned, scanned);
}

}  // namespace dev
}  // namespace vine
