/*
 * families_oracle.c -- CPU oracle for the generic label families.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/ may load it, as the checker; the
 * product path never calls it.
 *
 *   oracle_family             restatement of RUN_CSPP_CPU (algorithm/cspp_cpu.hpp:44-125,
 *                             add_new_frontier :16-42) for one label family at a
 *                             time.  The family semantics follow the reference's
 *                             label and constraint classes:
 *                               1 / 4  CSP == 1 / 4: include/label.hpp:170-303,
 *                                      include/constraint.hpp:81-133
 *                               2      CSP == 2 (hop): label.hpp:46-105, constraint.hpp:21-49
 *                               3      CSP == 3 (hop + maxcost): label.hpp:106-169,
 *                                      constraint.hpp:50-80
 *                               5      CSP == 5 (social): label.hpp:304-381,
 *                                      constraint.hpp:134-161
 *                               6      default branch (QoS): label.hpp:382-461,
 *                                      constraint.hpp:162-186
 *                             Pinned by oracle_family_bruteforce on tiny random
 *                             graphs (every family, the whole Pareto set at t) and,
 *                             for QoS, by Dijkstra on the arcs whose delay meets
 *                             the bound (tests/test_oracle_families.py).
 *   oracle_family_bruteforce  every simple s->t path by depth-first search,
 *                             labels built arc by arc with the constraint checked
 *                             at each step (as the label search does), then the
 *                             non-dominated set.  Independent of the round
 *                             structure; valid where removing a cycle never makes a
 *                             label worse (all families; social with weights in
 *                             [0, 1]).
 *
 * Build rules on top of the literal reference (each a reference defect):
 * `continue` instead of `break` on an infeasible son (algorithm/bucket.cuh:76-78);
 * expand_cpu of social and QoS kept as in the device expand; the answer is the
 * best label of t's Pareto set in the order of family_traits
 * (include/vine/families.hpp): the reference extracts none.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum { F_CSP = 1, F_HOP = 2, F_HOPMAX = 3, F_2OSP = 4, F_SOCIAL = 5, F_QOS = 6 };

typedef struct {
  int32_t vid;
  int32_t parent;
  float f[4]; /* d, delay, hop, maxcost */
} flabel;

typedef struct {
  int32_t* a;
  int32_t n, cap;
} fvec;

static int fvec_push(fvec* v, int32_t x) {
  if (v->n == v->cap) {
    int32_t nc = v->cap ? 2 * v->cap : 4;
    int32_t* na = (int32_t*)realloc(v->a, sizeof(int32_t) * (size_t)nc);
    if (!na) return -1;
    v->a = na;
    v->cap = nc;
  }
  v->a[v->n++] = x;
  return 0;
}

static void f_source(int fam, int32_t s, flabel* l) {
  l->vid = s;
  l->parent = -1;
  l->f[0] = l->f[1] = l->f[2] = l->f[3] = 0.f;
  if (fam == F_SOCIAL) l->f[0] = l->f[1] = 1.f;
  if (fam == F_QOS) l->f[1] = 21474836.f;
}

static void f_expand(int fam, const flabel* l, int32_t v, float len, float delay, flabel* son) {
  *son = *l;
  son->vid = v;
  switch (fam) {
    case F_SOCIAL:
      son->f[0] = l->f[0] * len;
      son->f[1] = l->f[1] * delay;
      break;
    case F_QOS:
      son->f[0] = l->f[0] + len;
      son->f[1] = l->f[1] < delay ? l->f[1] : delay;
      break;
    default:
      son->f[0] = l->f[0] + len;
      son->f[1] = l->f[1] + delay;
      if (fam == F_HOP || fam == F_HOPMAX) son->f[2] = l->f[2] + 1.f;
      if (fam == F_HOPMAX) son->f[3] = l->f[3] > len ? l->f[3] : len;
  }
}

static int f_dominate(int fam, const flabel* a, const flabel* b) {
  switch (fam) {
    case F_HOP: return a->f[0] <= b->f[0] && a->f[1] <= b->f[1] && a->f[2] <= b->f[2];
    case F_HOPMAX:
      return a->f[0] <= b->f[0] && a->f[1] <= b->f[1] && a->f[2] <= b->f[2] && a->f[3] <= b->f[3];
    case F_SOCIAL: return a->f[0] >= b->f[0] && a->f[1] >= b->f[1];
    case F_QOS: return a->f[0] <= b->f[0] && a->f[1] >= b->f[1];
    default: return a->f[0] <= b->f[0] && a->f[1] <= b->f[1];
  }
}

static int f_pass(int fam, const flabel* l, const float* lim) {
  switch (fam) {
    case F_HOP: return !(l->f[1] > lim[0] || l->f[2] > (float)(int)lim[1]);
    case F_HOPMAX: return !(l->f[1] > lim[0] || l->f[2] > (float)(int)lim[1] || l->f[3] > lim[2]);
    case F_QOS: return !(l->f[1] < lim[0]);
    default: return !(l->f[1] > lim[0]);
  }
}

/* "a before b" in the family's answer order */
static int f_better(int fam, const flabel* a, const flabel* b) {
  int dir[4] = {1, 1, 1, 1}, k = 2;
  if (fam == F_HOP) k = 3;
  if (fam == F_HOPMAX) k = 4;
  if (fam == F_SOCIAL) dir[0] = dir[1] = -1;
  if (fam == F_QOS) dir[1] = -1;
  for (int i = 0; i < k; i++)
    if (a->f[i] != b->f[i]) return dir[i] > 0 ? a->f[i] < b->f[i] : a->f[i] > b->f[i];
  return 0;
}

static int cmp_vid(const void* x, const void* y) {
  int32_t a = *(const int32_t*)x, b = *(const int32_t*)y;
  return (a > b) - (a < b);
}

/* sorts ids by the answer order and writes up to cap labels */
static void f_emit(int fam, const flabel* log, int32_t* ids, int32_t m, float* out, int32_t cap) {
  for (int32_t i = 1; i < m; i++) {
    int32_t x = ids[i], j = i - 1;
    while (j >= 0 && f_better(fam, &log[x], &log[ids[j]])) {
      ids[j + 1] = ids[j];
      j--;
    }
    ids[j + 1] = x;
  }
  for (int32_t i = 0; i < m && i < cap; i++) memcpy(out + 4 * i, log[ids[i]].f, sizeof(float) * 4);
}

/* Returns 1 feasible, 0 infeasible, <0 error (-1 arguments, -2 memory, -3 label
 * limit).  out_best: the answer's 4 fields; out_front: t's Pareto set in the
 * answer order; counters: rounds, labels expanded, arcs relaxed, labels inserted. */
int oracle_family(int32_t fam, int32_t n, const int32_t* row_ptr, const int32_t* dst, const float* len,
                  const float* delay, int32_t s, int32_t t, const float* lim, float* out_best, int32_t* out_path,
                  int32_t path_cap, int32_t* out_path_len, float* out_front, int32_t front_cap,
                  int32_t* out_front_len, int64_t* counters, int64_t max_labels) {
  if (fam < F_CSP || fam > F_QOS || n <= 0 || s < 0 || s >= n || t < 0 || t >= n) return -1;
  fvec* dom = (fvec*)calloc((size_t)n, sizeof(fvec));
  fvec* nxt = (fvec*)calloc((size_t)n, sizeof(fvec));
  char* in_next = (char*)calloc((size_t)n, 1);
  fvec frontier = {0}, touched = {0}, newfront = {0};
  flabel* log = NULL;
  int64_t log_n = 0, log_cap = 0, cnt[4] = {0, 0, 0, 0};
  int rc = 0;
  if (!dom || !nxt || !in_next) {
    rc = -2;
    goto done;
  }
  log_cap = 1024;
  log = (flabel*)malloc(sizeof(flabel) * (size_t)log_cap);
  if (!log) {
    rc = -2;
    goto done;
  }
  f_source(fam, s, &log[0]);
  log_n = 1;
  if (fvec_push(&dom[s], 0) || fvec_push(&frontier, 0)) {
    rc = -2;
    goto done;
  }
  while (frontier.n > 0) {
    cnt[0]++;
    for (int32_t k = 0; k < frontier.n; k++) {
      const int32_t fid = frontier.a[k];
      cnt[1]++;
      for (int32_t p = row_ptr[log[fid].vid]; p < row_ptr[log[fid].vid + 1]; p++) {
        cnt[2]++;
        flabel son;
        f_expand(fam, &log[fid], dst[p], len[p], delay[p], &son);
        son.parent = fid;
        if (!f_pass(fam, &son, lim)) continue;
        fvec* L = &dom[son.vid];
        int dominated = 0;
        for (int32_t i = 0; i < L->n;) {
          if (f_dominate(fam, &log[L->a[i]], &son)) {
            dominated = 1;
            break;
          }
          if (f_dominate(fam, &son, &log[L->a[i]])) {
            memmove(&L->a[i], &L->a[i + 1], sizeof(int32_t) * (size_t)(L->n - i - 1));
            L->n--;
          } else {
            i++;
          }
        }
        if (dominated) continue;
        if (log_n == log_cap) {
          if (log_cap >= max_labels) {
            rc = -3;
            goto done;
          }
          flabel* nl = (flabel*)realloc(log, sizeof(flabel) * (size_t)(2 * log_cap));
          if (!nl) {
            rc = -2;
            goto done;
          }
          log = nl;
          log_cap *= 2;
        }
        const int32_t nid = (int32_t)log_n;
        log[log_n++] = son;
        cnt[3]++;
        if (fvec_push(L, nid)) {
          rc = -2;
          goto done;
        }
        /* add_new_frontier: per-vertex next-round list, dominated entries dropped */
        fvec* X = &nxt[son.vid];
        if (!in_next[son.vid]) {
          in_next[son.vid] = 1;
          if (fvec_push(&touched, son.vid)) {
            rc = -2;
            goto done;
          }
        }
        int keep = 1;
        for (int32_t i = 0; i < X->n;) {
          if (f_dominate(fam, &log[X->a[i]], &son)) {
            keep = 0;
            break;
          }
          if (f_dominate(fam, &son, &log[X->a[i]])) {
            memmove(&X->a[i], &X->a[i + 1], sizeof(int32_t) * (size_t)(X->n - i - 1));
            X->n--;
          } else {
            i++;
          }
        }
        if (keep && fvec_push(X, nid)) {
          rc = -2;
          goto done;
        }
      }
    }
    qsort(touched.a, (size_t)touched.n, sizeof(int32_t), cmp_vid);
    newfront.n = 0;
    for (int32_t i = 0; i < touched.n; i++) {
      const int32_t v = touched.a[i];
      for (int32_t j = 0; j < nxt[v].n; j++)
        if (fvec_push(&newfront, nxt[v].a[j])) {
          rc = -2;
          goto done;
        }
      nxt[v].n = 0;
      in_next[v] = 0;
    }
    touched.n = 0;
    fvec tmp = frontier;
    frontier = newfront;
    newfront = tmp;
  }
  {
    fvec* T = &dom[t];
    int32_t best = -1;
    for (int32_t i = 0; i < T->n; i++)
      if (best < 0 || f_better(fam, &log[T->a[i]], &log[best])) best = T->a[i];
    if (out_front_len) {
      int32_t* ids = (int32_t*)malloc(sizeof(int32_t) * (size_t)(T->n > 0 ? T->n : 1));
      if (!ids) {
        rc = -2;
        goto done;
      }
      memcpy(ids, T->a, sizeof(int32_t) * (size_t)T->n);
      f_emit(fam, log, ids, T->n, out_front, front_cap);
      *out_front_len = T->n;
      free(ids);
    }
    if (best < 0) {
      for (int k = 0; k < 4; k++) out_best[k] = INFINITY;
      if (out_path_len) *out_path_len = 0;
      rc = 0;
    } else {
      memcpy(out_best, log[best].f, sizeof(float) * 4);
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
  if (counters)
    for (int k = 0; k < 4; k++) counters[k] = cnt[k];
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
}

/* ---- brute force over simple paths ------------------------------------- */
typedef struct {
  flabel* a;
  int32_t n, cap;
} lvec;

static void bf_family(int fam, int32_t u, int32_t t, const int32_t* row_ptr, const int32_t* dst, const float* len,
                      const float* delay, const float* lim, char* on, const flabel* cur, lvec* found,
                      int64_t* budget) {
  if (--(*budget) < 0) return;
  if (u == t) {
    if (found->n < found->cap) found->a[found->n++] = *cur;
    else *budget = -1;
    return;
  }
  for (int32_t p = row_ptr[u]; p < row_ptr[u + 1]; p++) {
    const int32_t v = dst[p];
    if (on[v]) continue;
    flabel son;
    f_expand(fam, cur, v, len[p], delay[p], &son);
    if (!f_pass(fam, &son, lim)) continue;
    on[v] = 1;
    bf_family(fam, v, t, row_ptr, dst, len, delay, lim, on, &son, found, budget);
    on[v] = 0;
  }
}

/* t's non-dominated set over simple s->t paths, in the answer order (one label
 * per distinct value).  Returns the set size, -1 on bad arguments, -3 when the
 * enumeration is too large. */
int oracle_family_bruteforce(int32_t fam, int32_t n, const int32_t* row_ptr, const int32_t* dst, const float* len,
                             const float* delay, int32_t s, int32_t t, const float* lim, float* out_front,
                             int32_t cap) {
  if (fam < F_CSP || fam > F_QOS || n <= 0 || n > 64 || s < 0 || s >= n || t < 0 || t >= n) return -1;
  char on[64] = {0};
  lvec found = {(flabel*)malloc(sizeof(flabel) * 200000), 0, 200000};
  if (!found.a) return -2;
  int64_t budget = 20000000;
  flabel src;
  f_source(fam, s, &src);
  on[s] = 1;
  bf_family(fam, s, t, row_ptr, dst, len, delay, lim, on, &src, &found, &budget);
  if (budget < 0) {
    free(found.a);
    return -3;
  }
  /* keep labels no other label dominates; among equal values keep the first */
  int32_t* ids = (int32_t*)malloc(sizeof(int32_t) * (size_t)(found.n > 0 ? found.n : 1));
  int32_t m = 0;
  for (int32_t i = 0; i < found.n; i++) {
    int keep = 1;
    for (int32_t j = 0; j < found.n && keep; j++) {
      if (i == j || !f_dominate(fam, &found.a[j], &found.a[i])) continue;
      const int equal = f_dominate(fam, &found.a[i], &found.a[j]);
      if (!equal || j < i) keep = 0;
    }
    if (keep) ids[m++] = i;
  }
  f_emit(fam, found.a, ids, m, out_front, cap);
  free(ids);
  free(found.a);
  return m;
}
