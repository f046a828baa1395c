/*
 * cspp_oracle.c -- CPU oracle for the exact constrained-shortest-path solver.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline).  The product path never calls it.
 *
 * Contents and what pins each function (tests/test_oracle.py holds the checks):
 *   oracle_cspp          restatement of RUN_CSPP_CPU (algorithm/cspp_cpu.hpp:44-125,
 *                        add_new_frontier :16-42).  Pinned by: lexicographic
 *                        Dijkstra closed forms (bound >= delay of the (len,delay)-
 *                        shortest path => cost = Dijkstra on len; bound = min delay
 *                        => cost = second key of Dijkstra on (delay,len);
 *                        bound < min delay => infeasible), the resource-expanded
 *                        DP below, and brute-force path enumeration on tiny graphs.
 *   oracle_cspp_astar    a second, independent exact algorithm for the same
 *                        answer: lexicographic bi-criteria A* (label-setting,
 *                        the fast-dominance scheme of bi-objective A*, BOA*:
 *                        Hernandez, Yeoh, Baier, Zhang, Koenig, Salzman,
 *                        SoCS 2020) with exact reverse-Dijkstra heuristics.  Not
 *                        a restatement of the reference: it expands each
 *                        non-dominated label once, so it finishes the long
 *                        queries the round-synchronous restatement cannot.
 *                        Pinned by: equality with oracle_cspp on every query
 *                        both finish, the DP and brute force below.
 *   oracle_dijkstra_lex  textbook Dijkstra on (w1, w2) pairs, lexicographic.
 *   oracle_resource_dp   f[b][v] = min len with delay <= b (integer delays >= 1).
 *   oracle_bruteforce    DFS over all simple paths (tiny graphs).
 *   oracle_ar_burg       restatement of AR::AutoRegression with the Burg
 *                        (max-entropy) method (include/AR.cpp:6-147); pinned
 *                        against the compiled reference AR.cpp (golden vectors).
 *   oracle_ar_lsq        restatement of the least-squares variant
 *                        (include/AR.cpp:154-259); same pin.
 *
 * Build rules applied on top of the literal reference (each a reference defect,
 * documented with its line):
 *   - expand is additive, as the device c_label::expand (include/label.hpp:213-222);
 *     the host expand_cpu multiplies (include/label.hpp:224-232), which would keep
 *     every label at (0,0).
 *   - the source label is placed in its own Pareto set before the first round,
 *     as ck_init does on the GPU (algorithm/cspp.cuh:63-67), so cost(s,s) = 0.
 *   - the result is extracted: the answer is the min-d label of t's Pareto set
 *     (the commented-out extraction at algorithm/cspp.cuh:305-314); a path is
 *     recovered from parent links; infeasible <=> t's set is empty
 *     (algorithm/cspp.cuh:337-340).
 * Semantics kept literally: round-synchronous frontier; every out-edge in
 * adjacency order; `continue` on an infeasible label (algorithm/cspp_cpu.hpp:80-82);
 * non-strict dominance (include/label.hpp:234-236); dominated entries erased in
 * order; next round deduplicated per vertex and concatenated by ascending vertex
 * id (algorithm/cspp_cpu.hpp:111-115); constraint delay <= bound
 * (include/constraint.hpp:99-104).
 */
#define _POSIX_C_SOURCE 199309L /* clock_gettime */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef struct {
  int32_t vid;
  float d;
  float delay;
  int32_t parent;
} olabel;

/* CPU time of the calling thread in seconds (queries solved on several
 * threads each get their own time limit) */
static double thread_cpu_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_THREAD_CPUTIME_ID, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

typedef struct {
  int32_t* a;
  int32_t n, cap;
} ivec;

static int ivec_push(ivec* v, int32_t x) {
  if (v->n == v->cap) {
    int32_t nc = v->cap ? v->cap * 2 : 4;
    int32_t* na = (int32_t*)realloc(v->a, sizeof(int32_t) * (size_t)nc);
    if (!na) return -1;
    v->a = na;
    v->cap = nc;
  }
  v->a[v->n++] = x;
  return 0;
}

static int dominate(const olabel* a, const olabel* b) { return a->d <= b->d && a->delay <= b->delay; }

static int cmp_i32(const void* x, const void* y) {
  int32_t a = *(const int32_t*)x, b = *(const int32_t*)y;
  return (a > b) - (a < b);
}

static int reverse_bounds(int32_t n, const int32_t* row_ptr, const int32_t* dst, const float* len,
                          const float* delay, int32_t t, double* lbl, double* lbd);
int oracle_dijkstra_lex(int32_t n, const int32_t* row_ptr, const int32_t* dst, const float* w1,
                        const float* w2, int32_t s, double* dist1, double* dist2);

/* counters[0] = labels expanded (X), [1] = arcs relaxed (E), [2] = labels
 * inserted into Pareto sets (I), [3] = rounds, [4] = max frontier size,
 * [5] = sum of frontier sizes. */
enum { OC_X = 0, OC_E, OC_I, OC_ROUNDS, OC_MAXF, OC_SUMF, OC_NCOUNT };

/* Returns 1 if feasible, 0 if infeasible, negative on error.
 * prune_target selects extra pruning that leaves the answer (cost, delay)
 * unchanged while shrinking the search; 0 is the reference's one-to-all search:
 *   1: drop a candidate at v != t when the best label found at t so far is
 *      lexicographically no worse (d > bd, or d == bd and delay >= bdl);
 *   2: lower-bound pruning with lbl/lbd = shortest remaining len/delay from v to
 *      t (Dijkstra on the transposed graph): drop when delay + lbd > bound, or
 *      when (d + lbl, delay + lbd) is lexicographically >= (bd, bdl).
 * out_pareto receives t's Pareto set sorted by d as (d, delay) pairs.
 * frontier_trace (optional, length trace_cap) receives the frontier size after
 * every round.  time_limit_s > 0 aborts with -4 once the calling thread has spent that much CPU time
 * (checked once per round). */
int oracle_cspp(int32_t n, const int32_t* row_ptr, const int32_t* dst, const float* len,
                const float* delay, int32_t s, int32_t t, float bound, int32_t prune_target,
                float* out_cost, float* out_delay, int32_t* out_path, int32_t path_cap,
                int32_t* out_path_len, float* out_pareto, int32_t pareto_cap,
                int32_t* out_pareto_len, int64_t* counters, int32_t* frontier_trace,
                int32_t trace_cap, double time_limit_s) {
  if (n <= 0 || s < 0 || s >= n || t < 0 || t >= n) return -1;
  int64_t cnt[OC_NCOUNT] = {0};
  ivec* dom = (ivec*)calloc((size_t)n, sizeof(ivec));
  ivec* nxt = (ivec*)calloc((size_t)n, sizeof(ivec));
  char* in_next = (char*)calloc((size_t)n, 1);
  olabel* log = NULL;
  int64_t log_n = 0, log_cap = 0;
  ivec frontier = {0}, touched = {0}, newfront = {0};
  double* lbl = NULL; /* prune_target == 2: lower bounds on the remaining len / delay to t */
  double* lbd = NULL;
  int rc = 0;
  if (!dom || !nxt || !in_next) {
    rc = -2;
    goto done;
  }
#define LOG_PUSH(L)                                                          \
  do {                                                                       \
    if (log_n == log_cap) {                                                  \
      int64_t nc = log_cap ? log_cap * 2 : 1024;                             \
      olabel* nl = (olabel*)realloc(log, sizeof(olabel) * (size_t)nc);       \
      if (!nl) {                                                             \
        rc = -2;                                                             \
        goto done;                                                           \
      }                                                                      \
      log = nl;                                                              \
      log_cap = nc;                                                          \
    }                                                                        \
    log[log_n++] = (L);                                                      \
  } while (0)

  {
    olabel src = {s, 0.0f, 0.0f, -1};
    LOG_PUSH(src);
    if (ivec_push(&dom[s], 0) || ivec_push(&frontier, 0)) {
      rc = -2;
      goto done;
    }
  }
  float bd = INFINITY, bdl = INFINITY; /* best label at t so far */
  if (prune_target == 2) {
    lbl = (double*)malloc(sizeof(double) * (size_t)n);
    lbd = (double*)malloc(sizeof(double) * (size_t)n);
    if (!lbl || !lbd || reverse_bounds(n, row_ptr, dst, len, delay, t, lbl, lbd) != 0) {
      rc = -2;
      goto done;
    }
  }
  if (s == t) {
    bd = 0.0f;
    bdl = 0.0f;
  }
  int32_t rounds = 0;
  const double t_begin = thread_cpu_s();
  while (frontier.n > 0) {
    rounds++;
    if (time_limit_s > 0 && thread_cpu_s() - t_begin > time_limit_s) {
      rc = -4; /* time limit reached: no answer */
      goto done;
    }
    for (int32_t ff = 0; ff < frontier.n; ff++) {
      int32_t fid = frontier.a[ff];
      olabel f = log[fid];
      cnt[OC_X]++;
      for (int32_t p = row_ptr[f.vid]; p < row_ptr[f.vid + 1]; p++) {
        cnt[OC_E]++;
        olabel nl;
        nl.vid = dst[p];
        nl.d = f.d + len[p];
        nl.delay = f.delay + delay[p];
        nl.parent = fid;
        if (nl.delay > bound) continue; /* !pass_constraints_check */
        if (prune_target == 1 && nl.vid != t && (nl.d > bd || (nl.d == bd && nl.delay >= bdl))) continue;
        if (prune_target == 2) {
          double rl = (double)nl.d + lbl[nl.vid], rd = (double)nl.delay + lbd[nl.vid];
          if (rd > (double)bound) continue;
          if (rl > (double)bd || (rl == (double)bd && rd >= (double)bdl)) continue;
        }
        ivec* L = &dom[nl.vid];
        int dominated = 0;
        for (int32_t i = 0; i < L->n;) {
          const olabel* e = &log[L->a[i]];
          if (dominate(e, &nl)) {
            dominated = 1;
            break;
          }
          if (dominate(&nl, e)) {
            memmove(&L->a[i], &L->a[i + 1], sizeof(int32_t) * (size_t)(L->n - i - 1));
            L->n--;
          } else {
            i++;
          }
        }
        if (dominated) continue;
        LOG_PUSH(nl);
        int32_t nid = (int32_t)(log_n - 1);
        if (ivec_push(L, nid)) {
          rc = -2;
          goto done;
        }
        cnt[OC_I]++;
        if (nl.vid == t && (nl.d < bd || (nl.d == bd && nl.delay < bdl))) {
          bd = nl.d;
          bdl = nl.delay;
        }
        /* add_new_frontier */
        ivec* X = &nxt[nl.vid];
        if (!in_next[nl.vid]) {
          in_next[nl.vid] = 1;
          if (ivec_push(&touched, nl.vid)) {
            rc = -2;
            goto done;
          }
        }
        int keep = 1;
        for (int32_t i = 0; i < X->n;) {
          const olabel* e = &log[X->a[i]];
          if (dominate(e, &nl)) {
            keep = 0;
            break;
          }
          if (dominate(&nl, e)) {
            memmove(&X->a[i], &X->a[i + 1], sizeof(int32_t) * (size_t)(X->n - i - 1));
            X->n--;
          } else {
            i++;
          }
        }
        if (keep && ivec_push(X, nid)) {
          rc = -2;
          goto done;
        }
      }
    }
    /* frontiers = concat(next_round by ascending vid) */
    qsort(touched.a, (size_t)touched.n, sizeof(int32_t), cmp_i32);
    newfront.n = 0;
    for (int32_t i = 0; i < touched.n; i++) {
      int32_t v = touched.a[i];
      for (int32_t j = 0; j < nxt[v].n; j++)
        if (ivec_push(&newfront, nxt[v].a[j])) {
          rc = -2;
          goto done;
        }
      nxt[v].n = 0;
      in_next[v] = 0;
    }
    touched.n = 0;
    ivec tmp = frontier;
    frontier = newfront;
    newfront = tmp;
    if (frontier_trace && rounds - 1 < trace_cap) frontier_trace[rounds - 1] = frontier.n;
    cnt[OC_SUMF] += frontier.n;
    if (frontier.n > cnt[OC_MAXF]) cnt[OC_MAXF] = frontier.n;
  }
  cnt[OC_ROUNDS] = rounds;

  /* result extraction: min-d label of t's Pareto set (unique under non-strict dominance) */
  {
    ivec* T = &dom[t];
    int32_t best = -1;
    for (int32_t i = 0; i < T->n; i++) {
      const olabel* e = &log[T->a[i]];
      if (best < 0 || e->d < log[best].d || (e->d == log[best].d && e->delay < log[best].delay))
        best = T->a[i];
    }
    if (out_pareto_len) {
      /* sort t's set by d (insertion sort; sets are small) */
      int32_t m = T->n;
      int32_t* ids = (int32_t*)malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1));
      if (!ids) {
        rc = -2;
        goto done;
      }
      memcpy(ids, T->a, sizeof(int32_t) * (size_t)m);
      for (int32_t i = 1; i < m; i++) {
        int32_t x = ids[i], j = i - 1;
        while (j >= 0 && (log[ids[j]].d > log[x].d ||
                          (log[ids[j]].d == log[x].d && log[ids[j]].delay > log[x].delay))) {
          ids[j + 1] = ids[j];
          j--;
        }
        ids[j + 1] = x;
      }
      *out_pareto_len = m;
      for (int32_t i = 0; i < m && i < pareto_cap; i++) {
        out_pareto[2 * i] = log[ids[i]].d;
        out_pareto[2 * i + 1] = log[ids[i]].delay;
      }
      free(ids);
    }
    if (best < 0) {
      if (out_cost) *out_cost = INFINITY;
      if (out_delay) *out_delay = INFINITY;
      if (out_path_len) *out_path_len = 0;
      rc = 0;
    } else {
      if (out_cost) *out_cost = log[best].d;
      if (out_delay) *out_delay = log[best].delay;
      int32_t hops = 0;
      for (int32_t c = best; c >= 0; c = log[c].parent) hops++;
      if (out_path_len) *out_path_len = hops;
      if (out_path && hops <= path_cap) {
        int32_t k = hops - 1;
        for (int32_t c = best; c >= 0; c = log[c].parent) out_path[k--] = log[c].vid;
      }
      rc = 1;
    }
  }
done:
  free(lbl);
  free(lbd);
  if (counters)
    for (int i = 0; i < OC_NCOUNT; i++) counters[i] = cnt[i];
  if (dom)
    for (int32_t v = 0; v < n; v++) free(dom[v].a);
  if (nxt)
    for (int32_t v = 0; v < n; v++) free(nxt[v].a);
  free(dom);
  free(nxt);
  free(in_next);
  free(log);
  free(frontier.a);
  free(touched.a);
  free(newfront.a);
  return rc;
#undef LOG_PUSH
}

/* ---------------------------------------------------------------------------
 * Textbook Dijkstra with lexicographic (w1, w2) keys; integer-valued weights
 * are summed in double (exact).  dist1/dist2 receive the keys of every vertex
 * (INFINITY if unreachable).  Binary heap with lazy deletion.
 */
typedef struct {
  double k1, k2;
  int32_t v;
} hitem;

static int hless(const hitem* a, const hitem* b) {
  return a->k1 < b->k1 || (a->k1 == b->k1 && (a->k2 < b->k2 || (a->k2 == b->k2 && a->v < b->v)));
}

int oracle_dijkstra_lex(int32_t n, const int32_t* row_ptr, const int32_t* dst, const float* w1,
                        const float* w2, int32_t s, double* dist1, double* dist2) {
  if (n <= 0 || s < 0 || s >= n) return -1;
  int64_t m = row_ptr[n];
  hitem* h = (hitem*)malloc(sizeof(hitem) * (size_t)(m + 2));
  char* done = (char*)calloc((size_t)n, 1);
  if (!h || !done) {
    free(h);
    free(done);
    return -2;
  }
  for (int32_t v = 0; v < n; v++) dist1[v] = dist2[v] = INFINITY;
  int64_t hn = 0;
  dist1[s] = dist2[s] = 0.0;
  h[hn++] = (hitem){0.0, 0.0, s};
  while (hn > 0) {
    hitem top = h[0];
    h[0] = h[--hn];
    for (int64_t i = 0;;) { /* sift down */
      int64_t l = 2 * i + 1, r = l + 1, b = i;
      if (l < hn && hless(&h[l], &h[b])) b = l;
      if (r < hn && hless(&h[r], &h[b])) b = r;
      if (b == i) break;
      hitem x = h[i];
      h[i] = h[b];
      h[b] = x;
      i = b;
    }
    int32_t u = top.v;
    if (done[u]) continue;
    done[u] = 1;
    for (int32_t p = row_ptr[u]; p < row_ptr[u + 1]; p++) {
      int32_t v = dst[p];
      double a = dist1[u] + (double)w1[p], b = dist2[u] + (double)w2[p];
      if (!done[v] && (a < dist1[v] || (a == dist1[v] && b < dist2[v]))) {
        dist1[v] = a;
        dist2[v] = b;
        int64_t i = hn++; /* sift up */
        h[i] = (hitem){a, b, v};
        while (i > 0) {
          int64_t par = (i - 1) / 2;
          if (!hless(&h[i], &h[par])) break;
          hitem x = h[i];
          h[i] = h[par];
          h[par] = x;
          i = par;
        }
      }
    }
  }
  free(h);
  free(done);
  return 0;
}

/* Lower bounds to t: Dijkstra on the transposed graph from t, once on len and
 * once on delay (a completion from v costs at least lbl[v] and lbd[v]). */
static int reverse_bounds(int32_t n, const int32_t* row_ptr, const int32_t* dst, const float* len,
                          const float* delay, int32_t t, double* lbl, double* lbd) {
  int64_t m = row_ptr[n];
  int32_t* rp = (int32_t*)calloc((size_t)n + 1, sizeof(int32_t));
  int32_t* rd = (int32_t*)malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1));
  float* rl = (float*)malloc(sizeof(float) * (size_t)(m > 0 ? m : 1));
  float* rdl = (float*)malloc(sizeof(float) * (size_t)(m > 0 ? m : 1));
  double* scratch = (double*)malloc(sizeof(double) * (size_t)n);
  int rc = -2;
  if (!rp || !rd || !rl || !rdl || !scratch) goto out;
  for (int64_t p = 0; p < m; p++) rp[dst[p] + 1]++;
  for (int32_t v = 0; v < n; v++) rp[v + 1] += rp[v];
  {
    int32_t* fill = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    if (!fill) goto out;
    memcpy(fill, rp, sizeof(int32_t) * (size_t)n);
    for (int32_t u = 0; u < n; u++)
      for (int32_t p = row_ptr[u]; p < row_ptr[u + 1]; p++) {
        int32_t q = fill[dst[p]]++;
        rd[q] = u;
        rl[q] = len[p];
        rdl[q] = delay[p];
      }
    free(fill);
  }
  if (oracle_dijkstra_lex(n, rp, rd, rl, rl, t, lbl, scratch) != 0) goto out;
  if (oracle_dijkstra_lex(n, rp, rd, rdl, rdl, t, lbd, scratch) != 0) goto out;
  rc = 0;
out:
  free(rp);
  free(rd);
  free(rl);
  free(rdl);
  free(scratch);
  return rc;
}

/* ---------------------------------------------------------------------------
 * Lexicographic bi-criteria A* (label-setting).  A label (v, d, delay) has the
 * key f = (d + lbl[v], delay + lbd[v]) with lbl / lbd the exact shortest len /
 * delay from v to t; both are consistent heuristics, so keys never decrease
 * lexicographically along an arc and labels leave the heap in non-decreasing
 * key order.  Labels of one vertex therefore leave in non-decreasing (d,
 * delay) order, and a label is dominated (non-strictly, as c_label::dominate)
 * by an earlier one at its vertex iff its delay is >= the smallest delay
 * expanded there (g2min[v]).  A label whose delay plus lbd exceeds the bound
 * cannot complete (the constraint delay <= bound).  The first label taken at t
 * is the lexicographic minimum (d, delay) over feasible walks: the answer of
 * oracle_cspp.  counters[0..2] = labels expanded, arcs relaxed, labels pushed.
 * Returns 1 feasible, 0 infeasible, negative on error (-4: time limit).
 */
typedef struct {
  double f1, f2;
  int32_t id;
} aitem;

static int aless(const aitem* a, const aitem* b) {
  return a->f1 < b->f1 || (a->f1 == b->f1 && (a->f2 < b->f2 || (a->f2 == b->f2 && a->id < b->id)));
}

int oracle_cspp_astar(int32_t n, const int32_t* row_ptr, const int32_t* dst, const float* len,
                      const float* delay, int32_t s, int32_t t, float bound, float* out_cost,
                      float* out_delay, int32_t* out_path, int32_t path_cap, int32_t* out_path_len,
                      int64_t* counters, double time_limit_s) {
  if (n <= 0 || s < 0 || s >= n || t < 0 || t >= n) return -1;
  int64_t cx = 0, ce = 0, ci = 0;
  double* lbl = (double*)malloc(sizeof(double) * (size_t)n);
  double* lbd = (double*)malloc(sizeof(double) * (size_t)n);
  float* g2min = (float*)malloc(sizeof(float) * (size_t)n);
  olabel* log = NULL;
  aitem* h = NULL;
  int64_t log_n = 0, log_cap = 0, hn = 0, hcap = 0;
  int rc = 0, best = -1;
  if (!lbl || !lbd || !g2min || reverse_bounds(n, row_ptr, dst, len, delay, t, lbl, lbd) != 0) {
    rc = -2;
    goto done;
  }
  for (int32_t v = 0; v < n; v++) g2min[v] = INFINITY;
#define A_PUSH(L, F1, F2)                                                                 \
  do {                                                                                    \
    if (log_n == log_cap) {                                                               \
      int64_t nc = log_cap ? log_cap * 2 : 4096;                                          \
      olabel* nl = (olabel*)realloc(log, sizeof(olabel) * (size_t)nc);                    \
      if (!nl) {                                                                          \
        rc = -2;                                                                          \
        goto done;                                                                        \
      }                                                                                   \
      log = nl;                                                                           \
      log_cap = nc;                                                                       \
    }                                                                                     \
    log[log_n] = (L);                                                                     \
    if (hn == hcap) {                                                                     \
      int64_t nc = hcap ? hcap * 2 : 4096;                                                \
      aitem* nh = (aitem*)realloc(h, sizeof(aitem) * (size_t)nc);                         \
      if (!nh) {                                                                          \
        rc = -2;                                                                          \
        goto done;                                                                        \
      }                                                                                   \
      h = nh;                                                                             \
      hcap = nc;                                                                          \
    }                                                                                     \
    int64_t i_ = hn++;                                                                    \
    h[i_] = (aitem){(F1), (F2), (int32_t)log_n};                                          \
    while (i_ > 0) {                                                                      \
      int64_t p_ = (i_ - 1) / 2;                                                          \
      if (!aless(&h[i_], &h[p_])) break;                                                  \
      aitem x_ = h[i_];                                                                   \
      h[i_] = h[p_];                                                                      \
      h[p_] = x_;                                                                         \
      i_ = p_;                                                                            \
    }                                                                                     \
    log_n++;                                                                              \
    ci++;                                                                                 \
  } while (0)

  if (lbd[s] <= (double)bound) {
    olabel src = {s, 0.0f, 0.0f, -1};
    A_PUSH(src, lbl[s], lbd[s]);
  }
  const double t_begin = thread_cpu_s();
  while (hn > 0) {
    aitem top = h[0];
    h[0] = h[--hn];
    for (int64_t i = 0;;) {
      int64_t l = 2 * i + 1, r = l + 1, b = i;
      if (l < hn && aless(&h[l], &h[b])) b = l;
      if (r < hn && aless(&h[r], &h[b])) b = r;
      if (b == i) break;
      aitem x = h[i];
      h[i] = h[b];
      h[b] = x;
      i = b;
    }
    olabel f = log[top.id];
    if (f.delay >= g2min[f.vid]) continue; /* dominated by a label expanded earlier at f.vid */
    g2min[f.vid] = f.delay;
    if (f.vid == t) {
      best = top.id;
      break;
    }
    cx++;
    if (time_limit_s > 0 && (cx & 0xffff) == 0 && thread_cpu_s() - t_begin > time_limit_s) {
      rc = -4;
      goto done;
    }
    for (int32_t p = row_ptr[f.vid]; p < row_ptr[f.vid + 1]; p++) {
      ce++;
      olabel nl;
      nl.vid = dst[p];
      nl.d = f.d + len[p];
      nl.delay = f.delay + delay[p];
      nl.parent = top.id;
      if (nl.delay > bound) continue;                           /* !pass_constraints_check */
      double f2 = (double)nl.delay + lbd[nl.vid];
      if (f2 > (double)bound) continue;                         /* no feasible completion */
      if (nl.delay >= g2min[nl.vid]) continue;                  /* dominated at its vertex */
      A_PUSH(nl, (double)nl.d + lbl[nl.vid], f2);
    }
  }
  if (best < 0) {
    if (out_cost) *out_cost = INFINITY;
    if (out_delay) *out_delay = INFINITY;
    if (out_path_len) *out_path_len = 0;
    rc = 0;
  } else {
    if (out_cost) *out_cost = log[best].d;
    if (out_delay) *out_delay = log[best].delay;
    int32_t hops = 0;
    for (int32_t c = best; c >= 0; c = log[c].parent) hops++;
    if (out_path_len) *out_path_len = hops;
    if (out_path && hops <= path_cap) {
      int32_t k = hops - 1;
      for (int32_t c = best; c >= 0; c = log[c].parent) out_path[k--] = log[c].vid;
    }
    rc = 1;
  }
done:
  if (counters) {
    counters[0] = cx;
    counters[1] = ce;
    counters[2] = ci;
  }
  free(lbl);
  free(lbd);
  free(g2min);
  free(log);
  free(h);
  return rc;
#undef A_PUSH
}

/* ---------------------------------------------------------------------------
 * Resource-expanded dynamic programme, independent of label algorithms:
 *   f[b][v] = min total len of an s->v walk with total delay <= b,
 * for integer delays >= 1 and integer b in [0, B].  f[b][s] = 0.
 * Returns (cost, lex-min delay) at t: delay = min b with f[b][t] == f[B][t].
 * Memory is two rolling arrays is not enough (need f[b - c] for c up to max
 * delay), so a full (B+1) x n table of doubles is used: keep B*n small.
 */
int oracle_resource_dp(int32_t n, const int32_t* row_ptr, const int32_t* dst, const float* len,
                       const float* delay, int32_t s, int32_t t, int32_t B, double* out_cost,
                       double* out_delay) {
  if (n <= 0 || s < 0 || s >= n || t < 0 || t >= n || B < 0) return -1;
  int64_t m = row_ptr[n];
  for (int64_t p = 0; p < m; p++)
    if (delay[p] < 1.0f || delay[p] != floorf(delay[p])) return -1;
  double* f = (double*)malloc(sizeof(double) * (size_t)(B + 1) * (size_t)n);
  if (!f) return -2;
  for (int32_t b = 0; b <= B; b++) {
    double* fb = f + (size_t)b * n;
    for (int32_t v = 0; v < n; v++) fb[v] = (b > 0) ? f[(size_t)(b - 1) * n + v] : INFINITY;
    fb[s] = 0.0;
    for (int32_t u = 0; u < n; u++)
      for (int32_t p = row_ptr[u]; p < row_ptr[u + 1]; p++) {
        int32_t c = (int32_t)delay[p];
        if (c > b) continue;
        double prev = f[(size_t)(b - c) * n + u];
        if (prev + (double)len[p] < fb[dst[p]]) fb[dst[p]] = prev + (double)len[p];
      }
  }
  double best = f[(size_t)B * n + t];
  *out_cost = best;
  *out_delay = INFINITY;
  if (best < INFINITY)
    for (int32_t b = 0; b <= B; b++)
      if (f[(size_t)b * n + t] == best) {
        *out_delay = b;
        break;
      }
  free(f);
  return best < INFINITY ? 1 : 0;
}

/* ---------------------------------------------------------------------------
 * Brute force: enumerate every simple s->t path (tiny graphs only) and return
 * the lexicographically smallest (len, delay) among those with delay <= bound.
 * Walks with cycles never beat their simple sub-path (weights >= 0).
 */
static void bf_dfs(int32_t u, int32_t t, const int32_t* row_ptr, const int32_t* dst,
                   const float* len, const float* delay, double bound, char* on, double d,
                   double dl, double* bc, double* bdl, int64_t* budget) {
  if (--(*budget) < 0) return;
  if (u == t) {
    if (d < *bc || (d == *bc && dl < *bdl)) {
      *bc = d;
      *bdl = dl;
    }
    return;
  }
  for (int32_t p = row_ptr[u]; p < row_ptr[u + 1]; p++) {
    int32_t v = dst[p];
    double ndl = dl + delay[p];
    if (on[v] || ndl > bound) continue;
    on[v] = 1;
    bf_dfs(v, t, row_ptr, dst, len, delay, bound, on, d + len[p], ndl, bc, bdl, budget);
    on[v] = 0;
  }
}

int oracle_bruteforce(int32_t n, const int32_t* row_ptr, const int32_t* dst, const float* len,
                      const float* delay, int32_t s, int32_t t, float bound, double* out_cost,
                      double* out_delay) {
  if (n <= 0 || n > 64 || s < 0 || s >= n || t < 0 || t >= n) return -1;
  char on[64] = {0};
  double bc = INFINITY, bdl = INFINITY;
  int64_t budget = 50000000;
  on[s] = 1;
  bf_dfs(s, t, row_ptr, dst, len, delay, (double)bound, on, 0.0, 0.0, &bc, &bdl, &budget);
  if (budget < 0) return -3;
  *out_cost = bc;
  *out_delay = bdl;
  return bc < INFINITY ? 1 : 0;
}

/* ---------------------------------------------------------------------------
 * AR model restatements (scheduling only; never affect costs).
 * oracle_ar_burg follows AR::AutoRegression(method=MAXENTROPY) and
 * AR::ARMaxEntropy (include/AR.cpp:6-147): subtract the mean, run Burg's
 * recursion, coefficients[i-1] = -ar[degree][i].
 */
int oracle_ar_burg(const double* x, int32_t length, int32_t degree, double* coef) {
  if (length < 2 || degree < 1 || degree >= length) return -1;
  double* w = (double*)malloc(sizeof(double) * (size_t)length);
  double* fwd = (double*)calloc((size_t)length + 1, sizeof(double));
  double* bwd = (double*)calloc((size_t)length + 1, sizeof(double));
  double* g = (double*)calloc((size_t)degree + 2, sizeof(double));
  double* hh = (double*)calloc((size_t)degree + 2, sizeof(double));
  double* last = (double*)calloc((size_t)degree + 2, sizeof(double));
  if (!w || !fwd || !bwd || !g || !hh || !last) {
    free(w);
    free(fwd);
    free(bwd);
    free(g);
    free(hh);
    free(last);
    return -2;
  }
  double mean = 0.0;
  for (int32_t i = 0; i < length; i++) mean += x[i];
  mean /= (double)length;
  for (int32_t i = 0; i < length; i++) w[i] = x[i] - mean;
  /* Burg recursion in the reference's 1-based indexing. */
  for (int32_t order = 1; order <= degree; order++) {
    int32_t n0 = order - 1;
    int32_t span = length - n0 - 1;
    double num = 0.0, den = 0.0;
    for (int32_t j = 1; j <= span; j++) {
      double a = w[j + n0] + fwd[j];
      double b = w[j - 1] + bwd[j];
      num -= 2.0 * a * b;
      den += a * a + b * b;
    }
    double k = num / den;
    g[order + 1] = k;
    if (n0 != 0) {
      for (int32_t j = 2; j <= order; j++) hh[j] = g[j] + k * g[n0 - j + 3];
      for (int32_t j = 2; j <= order; j++) g[j] = hh[j];
      span--;
    }
    for (int32_t j = 1; j <= span; j++) {
      bwd[j] += k * fwd[j] + k * w[j + order - 1];
      fwd[j] = fwd[j + 1] + k * bwd[j + 1] + k * w[j];
    }
    for (int32_t j = 2; j <= order + 1; j++) last[j - 1] = g[j];
  }
  for (int32_t i = 1; i <= degree; i++) coef[i - 1] = -last[i];
  free(w);
  free(fwd);
  free(bwd);
  free(g);
  free(hh);
  free(last);
  return 0;
}

/* Least-squares AR fit on the mean-subtracted series, Gaussian elimination with
 * partial pivoting (include/AR.cpp:154-259). */
int oracle_ar_lsq(const double* x, int32_t length, int32_t degree, double* coef) {
  if (length < 2 || degree < 1 || degree >= length || degree > 16) return -1;
  double w[4096];
  if (length > 4096) return -1;
  double mean = 0.0;
  for (int32_t i = 0; i < length; i++) mean += x[i];
  mean /= (double)length;
  for (int32_t i = 0; i < length; i++) w[i] = x[i] - mean;
  double mat[16][16] = {{0}};
  double* rows[16];
  for (int32_t i = 0; i < degree; i++) {
    coef[i] = 0.0;
    rows[i] = mat[i];
  }
  for (int32_t i = degree - 1; i < length - 1; i++) {
    for (int32_t j = 0; j < degree; j++) {
      coef[j] += w[i + 1] * w[i - j];
      for (int32_t k = j; k < degree; k++) mat[j][k] += w[i - j] * w[i - k];
    }
  }
  for (int32_t i = 0; i < degree; i++) {
    coef[i] /= (double)(length - degree);
    for (int32_t j = i; j < degree; j++) {
      mat[i][j] /= (double)(length - degree);
      mat[j][i] = mat[i][j];
    }
  }
  int32_t nn = degree;
  for (int32_t i = 0; i < nn - 1; i++) {
    double mx = fabs(rows[i][i]);
    int32_t mi = i;
    for (int32_t j = i + 1; j < nn; j++)
      if (fabs(rows[j][i]) > mx) {
        mx = fabs(rows[j][i]);
        mi = j;
      }
    if (mi != i) {
      double* r = rows[i];
      rows[i] = rows[mi];
      rows[mi] = r;
      double v = coef[i];
      coef[i] = coef[mi];
      coef[mi] = v;
    }
    double piv = rows[i][i];
    if (fabs(piv) == 0.0) return -3;
    for (int32_t j = i + 1; j < nn; j++) {
      double q = -rows[j][i] / piv;
      rows[j][i] = 0.0;
      for (int32_t k = i + 1; k < nn; k++) rows[j][k] += q * rows[i][k];
      coef[j] += q * coef[i];
    }
  }
  coef[nn - 1] /= rows[nn - 1][nn - 1];
  for (int32_t i = nn - 2; i >= 0; i--) {
    for (int32_t j = nn - 1; j > i; j--) coef[i] -= rows[i][j] * coef[j];
    coef[i] /= rows[i][i];
  }
  return 0;
}
