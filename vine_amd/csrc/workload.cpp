// Seeded synthetic workloads for the CSP solver: graphs (grid, k-NN geometric,
// R-MAT), query sets and the Dijkstra / BFS helpers that derive query bounds.
// Host-only C++17.  Output CSR keeps adjacency insertion order, as the
// reference's Graph::adj_to_csr does (include/graph.hpp:136-167).
#include "vine_workload.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <numeric>
#include <queue>
#include <thread>
#include <vector>

#include "host_graph.hpp"

namespace {

// splitmix64: small, seedable, stream-splittable generator.
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  // uniform integer in [lo, hi]
  int64_t range(int64_t lo, int64_t hi) {
    uint64_t span = (uint64_t)(hi - lo) + 1;
    return lo + (int64_t)(((unsigned __int128)next() * span) >> 64);
  }
  double unit() { return (double)(next() >> 11) * 0x1.0p-53; }
};

uint64_t stream_seed(uint64_t seed, uint64_t stream) {
  Rng r(seed ^ (0xD1B54A32D192ED03ull * (stream + 1)));
  return r.next();
}

struct UEdge {
  int32_t u, v;
  float len, delay;
};

// Adjacency-insertion-order CSR from an ordered undirected edge list.
vine_host_graph* build_undirected(int32_t n, const std::vector<UEdge>& edges) {
  auto* g = new vine_host_graph();
  g->n = n;
  std::vector<int32_t> deg(n, 0);
  for (const auto& e : edges) {
    deg[e.u]++;
    deg[e.v]++;
  }
  g->row_ptr.assign(n + 1, 0);
  for (int32_t i = 0; i < n; i++) g->row_ptr[i + 1] = g->row_ptr[i] + deg[i];
  int64_t m = g->row_ptr[n];
  g->dst.resize(m);
  g->len.resize(m);
  g->delay.resize(m);
  std::vector<int32_t> fill(g->row_ptr.begin(), g->row_ptr.end() - 1);
  for (const auto& e : edges) {
    int32_t p = fill[e.u]++;
    g->dst[p] = e.v;
    g->len[p] = e.len;
    g->delay[p] = e.delay;
    p = fill[e.v]++;
    g->dst[p] = e.u;
    g->len[p] = e.len;
    g->delay[p] = e.delay;
  }
  return g;
}

void parallel_for(int64_t n, int nthreads, const std::function<void(int64_t)>& fn) {
  if (nthreads <= 1 || n < 2) {
    for (int64_t i = 0; i < n; i++) fn(i);
    return;
  }
  std::atomic<int64_t> next{0};
  std::vector<std::thread> pool;
  for (int t = 0; t < nthreads; t++) {
    pool.emplace_back([&]() {
      for (;;) {
        int64_t i = next.fetch_add(1);
        if (i >= n) break;
        fn(i);
      }
    });
  }
  for (auto& th : pool) th.join();
}

int default_threads() {
  unsigned hc = std::thread::hardware_concurrency();
  if (hc == 0) hc = 1;
  return (int)std::min(hc, 16u);
}

// Dijkstra on integer-valued delays with early exit at t.
int64_t dijkstra_delay(const vine_host_graph* g, int32_t s, int32_t t, std::vector<int64_t>& dist,
                       std::vector<int32_t>& touched) {
  using Item = std::pair<int64_t, int32_t>;
  std::priority_queue<Item, std::vector<Item>, std::greater<Item>> pq;
  for (int32_t v : touched) dist[v] = INT64_MAX;
  touched.clear();
  dist[s] = 0;
  touched.push_back(s);
  pq.push({0, s});
  int64_t ans = -1;
  while (!pq.empty()) {
    auto [d, u] = pq.top();
    pq.pop();
    if (d != dist[u]) continue;
    if (u == t) {
      ans = d;
      break;
    }
    for (int32_t p = g->row_ptr[u]; p < g->row_ptr[u + 1]; p++) {
      int32_t v = g->dst[p];
      int64_t nd = d + (int64_t)g->delay[p];
      if (nd < dist[v]) {
        if (dist[v] == INT64_MAX) touched.push_back(v);
        dist[v] = nd;
        pq.push({nd, v});
      }
    }
  }
  return ans;
}

float bound_of(int64_t mindelay, int64_t num, int64_t den) {
  if (mindelay < 0) return 0.0f;
  return (float)((mindelay * num) / den);
}

std::vector<int32_t> largest_component(const vine_host_graph* g) {
  int32_t n = g->n;
  std::vector<int32_t> comp(n, -1);
  int32_t best = -1, best_size = 0, c = 0;
  std::vector<int32_t> stack;
  for (int32_t s = 0; s < n; s++) {
    if (comp[s] >= 0) continue;
    int32_t size = 0;
    comp[s] = c;
    stack.push_back(s);
    while (!stack.empty()) {
      int32_t u = stack.back();
      stack.pop_back();
      size++;
      for (int32_t p = g->row_ptr[u]; p < g->row_ptr[u + 1]; p++) {
        int32_t v = g->dst[p];
        if (comp[v] < 0) {
          comp[v] = c;
          stack.push_back(v);
        }
      }
    }
    if (size > best_size) {
      best_size = size;
      best = c;
    }
    c++;
  }
  std::vector<int32_t> out;
  out.reserve(best_size);
  for (int32_t v = 0; v < n; v++)
    if (comp[v] == best) out.push_back(v);
  return out;
}

}  // namespace

extern "C" {

int vine_gen_grid(int32_t rows, int32_t cols, int32_t delete_permille, int32_t len_lo,
                  int32_t len_hi, int32_t dly_lo, int32_t dly_hi, uint64_t seed,
                  vine_host_graph** out) {
  if (!out || rows <= 0 || cols <= 0 || delete_permille < 0 || delete_permille > 1000 ||
      len_lo < 0 || len_hi < len_lo || dly_lo < 0 || dly_hi < dly_lo)
    return -1;
  std::vector<std::pair<int32_t, int32_t>> all;
  for (int32_t r = 0; r < rows; r++)
    for (int32_t c = 0; c < cols; c++) {
      int32_t id = r * cols + c;
      if (c + 1 < cols) all.push_back({id, id + 1});
      if (r + 1 < rows) all.push_back({id, id + cols});
    }
  int64_t e = (int64_t)all.size();
  int64_t ndel = e * delete_permille / 1000;
  std::vector<int64_t> perm(e);
  std::iota(perm.begin(), perm.end(), 0);
  Rng shuf(stream_seed(seed, 0));
  for (int64_t i = e - 1; i > 0; i--) std::swap(perm[i], perm[shuf.range(0, i)]);
  std::vector<char> deleted(e, 0);
  for (int64_t i = 0; i < ndel; i++) deleted[perm[i]] = 1;
  Rng w(stream_seed(seed, 1));
  std::vector<UEdge> edges;
  edges.reserve(e - ndel);
  for (int64_t i = 0; i < e; i++) {
    if (deleted[i]) continue;
    float l = (float)w.range(len_lo, len_hi);
    float d = (float)w.range(dly_lo, dly_hi);
    edges.push_back({all[i].first, all[i].second, l, d});
  }
  *out = build_undirected(rows * cols, edges);
  return 0;
}

int vine_gen_knn(int64_t n, int32_t k, double len_scale, uint64_t seed, vine_host_graph** out) {
  if (!out || n < 2 || n > (1ll << 30) || k < 1 || k >= n || !(len_scale > 0)) return -1;
  std::vector<double> xs(n), ys(n);
  Rng pr(stream_seed(seed, 0));
  for (int64_t i = 0; i < n; i++) {
    xs[i] = pr.unit();
    ys[i] = pr.unit();
  }
  // Uniform bucket grid with about two points per cell.
  int64_t side = std::max<int64_t>(1, (int64_t)std::sqrt((double)n / 2.0));
  double cell = 1.0 / (double)side;
  std::vector<int64_t> cstart(side * side + 1, 0);
  std::vector<int32_t> cpts(n);
  auto cell_of = [&](double v) {
    int64_t c = (int64_t)(v * side);
    return c >= side ? side - 1 : c;
  };
  for (int64_t i = 0; i < n; i++) cstart[cell_of(ys[i]) * side + cell_of(xs[i]) + 1]++;
  for (int64_t c = 0; c < side * side; c++) cstart[c + 1] += cstart[c];
  {
    std::vector<int64_t> fill(cstart.begin(), cstart.end() - 1);
    for (int64_t i = 0; i < n; i++) cpts[fill[cell_of(ys[i]) * side + cell_of(xs[i])]++] = (int32_t)i;
  }
  std::vector<int32_t> nbr(n * k);
  parallel_for(n, default_threads(), [&](int64_t i) {
    // best k by (dist2, index), kept sorted ascending
    std::vector<std::pair<double, int32_t>> best;
    best.reserve(k + 1);
    int64_t cx = cell_of(xs[i]), cy = cell_of(ys[i]);
    for (int64_t r = 0;; r++) {
      for (int64_t yy = cy - r; yy <= cy + r; yy++) {
        if (yy < 0 || yy >= side) continue;
        for (int64_t xx = cx - r; xx <= cx + r; xx++) {
          if (xx < 0 || xx >= side) continue;
          if (std::max(std::llabs(xx - cx), std::llabs(yy - cy)) != r) continue;
          int64_t c = yy * side + xx;
          for (int64_t p = cstart[c]; p < cstart[c + 1]; p++) {
            int32_t j = cpts[p];
            if (j == i) continue;
            double dx = xs[j] - xs[i], dy = ys[j] - ys[i];
            std::pair<double, int32_t> cand{dx * dx + dy * dy, j};
            if ((int)best.size() < k || cand < best.back()) {
              best.insert(std::upper_bound(best.begin(), best.end(), cand), cand);
              if ((int)best.size() > k) best.pop_back();
            }
          }
        }
      }
      double reach = (double)r * cell;
      if ((int)best.size() == k && best.back().first <= reach * reach) break;
      if (r > 2 * side + 2) break;
    }
    for (int32_t q = 0; q < k; q++) nbr[i * k + q] = best[q].second;
  });
  std::vector<std::pair<int32_t, int32_t>> pairs;
  pairs.reserve(n * k);
  for (int64_t i = 0; i < n; i++)
    for (int32_t q = 0; q < k; q++) {
      int32_t j = nbr[i * k + q];
      pairs.push_back({(int32_t)std::min<int64_t>(i, j), (int32_t)std::max<int64_t>(i, j)});
    }
  std::sort(pairs.begin(), pairs.end());
  pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
  Rng w(stream_seed(seed, 1));
  std::vector<UEdge> edges;
  edges.reserve(pairs.size());
  for (auto& pq : pairs) {
    double dx = xs[pq.first] - xs[pq.second], dy = ys[pq.first] - ys[pq.second];
    int64_t l = std::llround(std::sqrt(dx * dx + dy * dy) * len_scale);
    if (l < 1) l = 1;
    double f = 0.5 + 1.5 * w.unit();
    int64_t d = std::llround((double)l * f);
    if (d < 1) d = 1;
    edges.push_back({pq.first, pq.second, (float)l, (float)d});
  }
  *out = build_undirected((int32_t)n, edges);
  return 0;
}

int vine_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c, int32_t lo,
                  int32_t hi, uint64_t seed, vine_host_graph** out) {
  if (!out || scale < 1 || scale > 28 || edge_factor < 1 || a < 0 || b < 0 || c < 0 ||
      a + b + c > 1.0 || lo < 0 || hi < lo)
    return -1;
  int64_t n = 1ll << scale;
  int64_t ne = (int64_t)edge_factor * n;
  Rng r(stream_seed(seed, 0));
  std::vector<std::pair<int32_t, int32_t>> pairs;
  pairs.reserve(ne);
  for (int64_t e = 0; e < ne; e++) {
    int64_t u = 0, v = 0;
    for (int32_t lvl = 0; lvl < scale; lvl++) {
      double x = r.unit();
      int bu, bv;
      if (x < a) {
        bu = 0;
        bv = 0;
      } else if (x < a + b) {
        bu = 0;
        bv = 1;
      } else if (x < a + b + c) {
        bu = 1;
        bv = 0;
      } else {
        bu = 1;
        bv = 1;
      }
      u = (u << 1) | bu;
      v = (v << 1) | bv;
    }
    if (u == v) continue;
    pairs.push_back({(int32_t)std::min(u, v), (int32_t)std::max(u, v)});
  }
  std::sort(pairs.begin(), pairs.end());
  pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
  Rng w(stream_seed(seed, 1));
  std::vector<UEdge> edges;
  edges.reserve(pairs.size());
  for (auto& pq : pairs) {
    float l = (float)w.range(lo, hi);
    float d = (float)w.range(lo, hi);
    edges.push_back({pq.first, pq.second, l, d});
  }
  *out = build_undirected((int32_t)n, edges);
  return 0;
}

int vine_host_graph_from_csr(int32_t n, int64_t m, const int32_t* row_ptr, const int32_t* dst,
                             const float* len, const float* delay, vine_host_graph** out) {
  if (!out || n < 0 || m < 0 || !row_ptr || (m > 0 && (!dst || !len || !delay))) return -1;
  if (row_ptr[0] != 0 || row_ptr[n] != m) return -1;
  for (int32_t i = 0; i < n; i++)
    if (row_ptr[i + 1] < row_ptr[i]) return -1;
  for (int64_t p = 0; p < m; p++)
    if (dst[p] < 0 || dst[p] >= n) return -1;
  auto* g = new vine_host_graph();
  g->n = n;
  g->row_ptr.assign(row_ptr, row_ptr + n + 1);
  g->dst.assign(dst, dst + m);
  g->len.assign(len, len + m);
  g->delay.assign(delay, delay + m);
  *out = g;
  return 0;
}

void vine_host_graph_free(vine_host_graph* g) { delete g; }
int32_t vine_host_graph_n(const vine_host_graph* g) { return g ? g->n : -1; }
int64_t vine_host_graph_m(const vine_host_graph* g) { return g ? (int64_t)g->dst.size() : -1; }
const int32_t* vine_host_graph_row_ptr(const vine_host_graph* g) { return g->row_ptr.data(); }
const int32_t* vine_host_graph_dst(const vine_host_graph* g) { return g->dst.data(); }
const float* vine_host_graph_len(const vine_host_graph* g) { return g->len.data(); }
const float* vine_host_graph_delay(const vine_host_graph* g) { return g->delay.data(); }

int64_t vine_min_delay(const vine_host_graph* g, int32_t s, int32_t t) {
  if (!g || s < 0 || s >= g->n || t < 0 || t >= g->n) return -2;
  std::vector<int64_t> dist(g->n, INT64_MAX);
  std::vector<int32_t> touched;
  return dijkstra_delay(g, s, t, dist, touched);
}

int vine_gen_queries_random(const vine_host_graph* g, int32_t nq, int32_t lcc_only, int32_t num,
                            int32_t den, uint64_t seed, int32_t nthreads, int32_t* out_src,
                            int32_t* out_dst, float* out_bound, int64_t* out_mindelay) {
  if (!g || nq < 0 || g->n < 2 || num <= 0 || den <= 0 || !out_src || !out_dst || !out_bound)
    return -1;
  std::vector<int32_t> pool;
  if (lcc_only) {
    pool = largest_component(g);
    if (pool.size() < 2) return -1;
  }
  Rng r(stream_seed(seed, 7));
  for (int32_t q = 0; q < nq; q++) {
    int32_t s, t;
    do {
      if (lcc_only) {
        s = pool[r.range(0, (int64_t)pool.size() - 1)];
        t = pool[r.range(0, (int64_t)pool.size() - 1)];
      } else {
        s = (int32_t)r.range(0, g->n - 1);
        t = (int32_t)r.range(0, g->n - 1);
      }
    } while (s == t);
    out_src[q] = s;
    out_dst[q] = t;
  }
  if (nthreads <= 0) nthreads = default_threads();
  std::vector<int64_t> md(nq);
  std::vector<std::vector<int64_t>> dists(nthreads);
  std::vector<std::vector<int32_t>> touch(nthreads);
  std::atomic<int32_t> next{0};
  std::vector<std::thread> pool_t;
  for (int th = 0; th < nthreads; th++) {
    pool_t.emplace_back([&, th]() {
      dists[th].assign(g->n, INT64_MAX);
      for (;;) {
        int32_t q = next.fetch_add(1);
        if (q >= nq) break;
        md[q] = dijkstra_delay(g, out_src[q], out_dst[q], dists[th], touch[th]);
      }
    });
  }
  for (auto& th : pool_t) th.join();
  for (int32_t q = 0; q < nq; q++) {
    out_bound[q] = bound_of(md[q], num, den);
    if (out_mindelay) out_mindelay[q] = md[q];
  }
  return 0;
}

int vine_gen_queries_hops(const vine_host_graph* g, int32_t nq, int32_t hop_lo, int32_t hop_hi,
                          int32_t f_lo_permille, int32_t f_hi_permille, uint64_t seed,
                          int32_t nthreads, int32_t* out_src, int32_t* out_dst, float* out_bound,
                          int64_t* out_mindelay, int32_t* out_hops) {
  if (!g || nq < 0 || g->n < 2 || hop_lo < 1 || hop_hi < hop_lo || f_lo_permille < 1000 ||
      f_hi_permille < f_lo_permille || !out_src || !out_dst || !out_bound)
    return -1;
  if (nthreads <= 0) nthreads = default_threads();
  const double llo = std::log((double)hop_lo), lhi = std::log((double)hop_hi);
  std::vector<int64_t> md(nq);
  std::vector<int32_t> fperm(nq), hops(nq);
  std::atomic<int32_t> next{0};
  std::atomic<int> failed{0};
  std::vector<std::thread> pool_t;
  for (int th = 0; th < nthreads; th++) {
    pool_t.emplace_back([&]() {
      std::vector<int32_t> depth(g->n, -1);
      std::vector<int64_t> dist(g->n, INT64_MAX);
      std::vector<int32_t> dtouch, frontier, nextf, ring;
      for (;;) {
        int32_t q = next.fetch_add(1);
        if (q >= nq) break;
        Rng r(stream_seed(seed, 1000 + (uint64_t)q));
        int32_t s = -1, t = -1, h = 0;
        for (int attempt = 0; attempt < 100000 && t < 0; attempt++) {
          s = (int32_t)r.range(0, g->n - 1);
          h = (int32_t)std::floor(std::exp(llo + (lhi - llo) * r.unit()));
          if (h < hop_lo) h = hop_lo;
          if (h > hop_hi) h = hop_hi;
          // BFS from s up to depth h
          std::vector<int32_t> seen;
          frontier.assign(1, s);
          depth[s] = 0;
          seen.push_back(s);
          int32_t lvl = 0;
          while (!frontier.empty() && lvl < h) {
            nextf.clear();
            for (int32_t u : frontier)
              for (int32_t p = g->row_ptr[u]; p < g->row_ptr[u + 1]; p++) {
                int32_t v = g->dst[p];
                if (depth[v] < 0) {
                  depth[v] = lvl + 1;
                  seen.push_back(v);
                  nextf.push_back(v);
                }
              }
            frontier.swap(nextf);
            lvl++;
          }
          ring.clear();
          if (lvl == h)
            for (int32_t v : frontier) ring.push_back(v);
          for (int32_t v : seen) depth[v] = -1;
          if (!ring.empty()) {
            std::sort(ring.begin(), ring.end());
            t = ring[r.range(0, (int64_t)ring.size() - 1)];
          }
        }
        if (t < 0) {
          failed = 1;
          continue;
        }
        out_src[q] = s;
        out_dst[q] = t;
        hops[q] = h;
        fperm[q] = (int32_t)r.range(f_lo_permille, f_hi_permille);
        md[q] = dijkstra_delay(g, s, t, dist, dtouch);
      }
    });
  }
  for (auto& th : pool_t) th.join();
  if (failed) return -2;
  for (int32_t q = 0; q < nq; q++) {
    out_bound[q] = bound_of(md[q], fperm[q], 1000);
    if (out_mindelay) out_mindelay[q] = md[q];
    if (out_hops) out_hops[q] = hops[q];
  }
  return 0;
}

int vine_gen_preset_graph(int32_t id, uint64_t seed, vine_host_graph** out) {
  switch (id) {
    case 1:
      return vine_gen_grid(64, 64, 100, 1, 1000, 1, 100, seed, out);
    case 2:
      return vine_gen_knn(264346, 3, 1e5, seed, out);
    case 3:
      return vine_gen_knn(1524453, 3, 1e5, seed, out);
    case 4:
      return vine_gen_rmat(17, 8, 0.57, 0.19, 0.19, 1, 100, seed, out);
    default:
      return -1;
  }
}

}  // extern "C"
