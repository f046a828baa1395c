// Vertex order for HBM locality.
//
// The solver keeps per-(query, vertex) arrays (Pareto-front headers, bounds,
// search flags) and reads them at the heads of the arcs it relaxes.  With the
// caller's vertex ids those heads are scattered, so every 4-8 byte access
// moves a whole cache line.  vine_graph_create therefore renumbers vertices
// so that neighbours get nearby ids: per connected component (arcs taken as
// undirected), two far-apart vertices A and B are found by breadth-first
// search and every vertex is placed on a Hilbert curve over its hop
// distances (d(A, v), d(B, v)), which act as two coordinates on road-like
// graphs.  Components follow each other in the order of their smallest id.
// Any order is correct; this one only changes which lines are shared.
#include <algorithm>
#include <cstdint>
#include <vector>

#include "layout.hpp"
#include "vine.h"

namespace vine {
namespace {

// Position of (x, y) along the Hilbert curve filling a side x side square
// (side a power of two).
uint64_t hilbert_index(uint32_t side, uint32_t x, uint32_t y) {
  uint64_t d = 0;
  for (uint32_t s = side / 2; s > 0; s /= 2) {
    const uint32_t rx = (x & s) ? 1u : 0u;
    const uint32_t ry = (y & s) ? 1u : 0u;
    d += (uint64_t)s * s * ((3u * rx) ^ ry);
    if (ry == 0) {
      if (rx == 1) {
        x = side - 1 - x;
        y = side - 1 - y;
      }
      std::swap(x, y);
    }
  }
  return d;
}

struct Undirected {
  std::vector<int64_t> off;
  std::vector<int32_t> adj;
};

Undirected undirected(int32_t n, const int32_t* row, const int32_t* dst) {
  Undirected u;
  u.off.assign((size_t)n + 1, 0);
  for (int32_t v = 0; v < n; v++)
    for (int32_t p = row[v]; p < row[v + 1]; p++) {
      u.off[v + 1]++;
      u.off[dst[p] + 1]++;
    }
  for (int32_t v = 0; v < n; v++) u.off[v + 1] += u.off[v];
  u.adj.resize((size_t)u.off[n]);
  std::vector<int64_t> fill(u.off.begin(), u.off.end() - 1);
  for (int32_t v = 0; v < n; v++)
    for (int32_t p = row[v]; p < row[v + 1]; p++) {
      u.adj[fill[v]++] = dst[p];
      u.adj[fill[dst[p]]++] = v;
    }
  return u;
}

// Hop distances from `src` within its component; returns the farthest vertex.
// `dist` must hold -1 for the component's vertices on entry.
int32_t bfs(const Undirected& g, int32_t src, std::vector<int32_t>& dist, std::vector<int32_t>& queue) {
  queue.clear();
  queue.push_back(src);
  dist[src] = 0;
  int32_t far = src;
  for (size_t h = 0; h < queue.size(); h++) {
    const int32_t v = queue[h];
    if (dist[v] > dist[far]) far = v;
    for (int64_t p = g.off[v]; p < g.off[v + 1]; p++) {
      const int32_t w = g.adj[p];
      if (dist[w] < 0) {
        dist[w] = dist[v] + 1;
        queue.push_back(w);
      }
    }
  }
  return far;
}

}  // namespace

std::vector<int32_t> locality_order(int32_t n, const int32_t* row, const int32_t* dst) {
  const Undirected g = undirected(n, row, dst);
  std::vector<int32_t> order;
  order.reserve((size_t)n);
  std::vector<int32_t> d0((size_t)n, -1), da((size_t)n, -1), db((size_t)n, -1), queue;
  std::vector<std::pair<uint64_t, int32_t>> keyed;
  for (int32_t r = 0; r < n; r++) {
    if (d0[r] >= 0) continue;
    const int32_t a = bfs(g, r, d0, queue);  // queue = the component
    const std::vector<int32_t> comp = queue;
    const int32_t b = bfs(g, a, da, queue);
    bfs(g, b, db, queue);
    int32_t span = 1;
    for (int32_t v : comp) span = std::max(span, std::max(da[v], db[v]) + 1);
    uint32_t side = 1;
    while (side < (uint32_t)span) side <<= 1;
    keyed.clear();
    for (int32_t v : comp) keyed.push_back({hilbert_index(side, (uint32_t)da[v], (uint32_t)db[v]), v});
    std::sort(keyed.begin(), keyed.end());
    for (const auto& kv : keyed) order.push_back(kv.second);
  }
  return order;
}

namespace {

// Min delays from src (-1: unreachable), by Dijkstra with a binary heap;
// sums in double, so integer delays stay exact.
void dijkstra(int32_t n, const int32_t* row, const int32_t* dst, const float* delay, int32_t src,
              std::vector<double>& d) {
  d.assign((size_t)n, -1.0);
  std::vector<std::pair<double, int32_t>> heap;
  auto cmp = [](const std::pair<double, int32_t>& a, const std::pair<double, int32_t>& b) { return a.first > b.first; };
  d[src] = 0.0;
  heap.push_back({0.0, src});
  while (!heap.empty()) {
    std::pop_heap(heap.begin(), heap.end(), cmp);
    const auto [du, u] = heap.back();
    heap.pop_back();
    if (du > d[u]) continue;
    for (int32_t p = row[u]; p < row[u + 1]; p++) {
      const int32_t w = dst[p];
      const double nd = du + (double)delay[p];
      if (d[w] < 0.0 || nd < d[w]) {
        d[w] = nd;
        heap.push_back({nd, w});
        std::push_heap(heap.begin(), heap.end(), cmp);
      }
    }
  }
}

}  // namespace

void landmark_distances(int32_t n, const int32_t* row, const int32_t* dst, const float* delay, int K,
                        std::vector<float>& out) {
  out.assign((size_t)n * 4, -1.f);
  std::vector<double> d, nearest;
  dijkstra(n, row, dst, delay, 0, d);
  int32_t next = 0;
  for (int32_t v = 0; v < n; v++)
    if (d[v] > d[next]) next = v;
  nearest.assign((size_t)n, -1.0);
  for (int j = 0; j < K; j++) {
    dijkstra(n, row, dst, delay, next, d);
    for (int32_t v = 0; v < n; v++) {
      out[(size_t)v * 4 + j] = (float)d[v];
      if (d[v] >= 0.0 && (nearest[v] < 0.0 || d[v] < nearest[v])) nearest[v] = d[v];
    }
    for (int32_t v = 0; v < n; v++)
      if (nearest[v] > nearest[next]) next = v;
  }
}

}  // namespace vine

extern "C" int32_t vine_locality_order(int32_t n, const int32_t* row_ptr, const int32_t* dst, int32_t* out_order) {
  if (n <= 0 || !row_ptr || !out_order || (row_ptr[n] > 0 && !dst)) return VINE_E_ARG;
  const std::vector<int32_t> o = vine::locality_order(n, row_ptr, dst);
  std::copy(o.begin(), o.end(), out_order);
  return VINE_OK;
}

extern "C" int32_t vine_landmark_delays(int32_t n, const int32_t* row_ptr, const int32_t* dst, const float* delay,
                                        float* out) {
  if (n <= 0 || !row_ptr || !out || (row_ptr[n] > 0 && (!dst || !delay))) return VINE_E_ARG;
  std::vector<float> o;
  vine::landmark_distances(n, row_ptr, dst, delay, 4, o);
  std::copy(o.begin(), o.end(), out);
  return VINE_OK;
}
