// Graph and query files: the input formats of the reference's read_graph
// (include/io.hpp:22-91) and task::read_query / task::load_query
// (include/task.hpp:27-76), read into the host CSR of vine_workload.h.
//
// Differences from the reference parsers, all on inputs they mishandle:
// unreadable files, ids outside [0, N), malformed rows, negative weights and
// weights at or above 2^24 (no longer exact in float32) are reported through
// vine_last_error() instead of crashing or corrupting memory (io.hpp:28 reads
// a NULL FILE*, io.hpp:26 indexes edges with an uninitialised id); and the
// SNAP form also accepts the MatrixMarket files the reference ships
// (inputs/tech-*.mtx.txt), which its parser cannot read.
#include <cctype>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "errors.hpp"
#include "host_graph.hpp"
#include "vine.h"
#include "vine_workload.h"

namespace {

constexpr int64_t kExact = 1 << 24;  // integer weights below this are exact in float32

// The whole file in memory, read as a stream of integers and separators.
struct Text {
  std::string buf;
  size_t pos = 0;
  int64_t line = 1;

  bool load(const char* path) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    char chunk[1 << 16];
    size_t got;
    while ((got = std::fread(chunk, 1, sizeof(chunk), f)) > 0) buf.append(chunk, got);
    std::fclose(f);
    return true;
  }
  void skip_space() {
    while (pos < buf.size() && std::isspace((unsigned char)buf[pos])) {
      if (buf[pos] == '\n') line++;
      pos++;
    }
  }
  bool at_end() {
    skip_space();
    return pos >= buf.size();
  }
  // next non-space character, 0 at the end
  char peek() {
    skip_space();
    return pos < buf.size() ? buf[pos] : 0;
  }
  void skip_line() {
    while (pos < buf.size() && buf[pos] != '\n') pos++;
  }
  // optional sign and decimal digits after optional white space
  bool integer(int64_t* v) {
    skip_space();
    size_t p = pos;
    bool neg = false;
    if (p < buf.size() && (buf[p] == '-' || buf[p] == '+')) neg = buf[p++] == '-';
    if (p >= buf.size() || !std::isdigit((unsigned char)buf[p])) return false;
    int64_t x = 0;
    while (p < buf.size() && std::isdigit((unsigned char)buf[p])) {
      if (x > (int64_t)1 << 40) return false;
      x = x * 10 + (buf[p++] - '0');
    }
    pos = p;
    *v = neg ? -x : x;
    return true;
  }
  // one literal character after optional white space
  bool expect(char c) {
    if (peek() != c) return false;
    pos++;
    return true;
  }
};

struct Arc {
  int32_t u, v;
  float len, delay;
};

int fail(const char* path, const Text& t, const char* what) {
  vine::set_error("%s:%lld: %s", path, (long long)t.line, what);
  return VINE_E_ARG;
}

// CSR in adjacency insertion order (Graph::adj_to_csr, include/graph.hpp:136-167).
vine_host_graph* to_csr(int32_t n, const std::vector<Arc>& arcs) {
  auto* g = new vine_host_graph();
  g->n = n;
  g->row_ptr.assign((size_t)n + 1, 0);
  for (const Arc& a : arcs) g->row_ptr[a.u + 1]++;
  for (int32_t v = 0; v < n; v++) g->row_ptr[v + 1] += g->row_ptr[v];
  std::vector<int32_t> fill(g->row_ptr.begin(), g->row_ptr.end() - 1);
  g->dst.resize(arcs.size());
  g->len.resize(arcs.size());
  g->delay.resize(arcs.size());
  for (const Arc& a : arcs) {
    const int32_t p = fill[a.u]++;
    g->dst[p] = a.v;
    g->len[p] = a.len;
    g->delay[p] = a.delay;
  }
  return g;
}

bool weight_ok(int64_t w) { return w >= 0 && w < kExact; }

// COLA road network (io.hpp:61-91): the vertex count, then per line
// "u: v w c v w c ... -1", each triple an arc u -> v with len w and delay c,
// directed as listed.
int read_cola(const char* path, Text& t, int32_t* n_out, std::vector<Arc>& arcs) {
  int64_t n;
  if (!t.integer(&n) || n <= 0 || n > INT32_MAX) return fail(path, t, "expected the vertex count");
  while (!t.at_end()) {
    int64_t u;
    if (!t.integer(&u) || !t.expect(':')) return fail(path, t, "expected \"u:\"");
    if (u < 0 || u >= n) return fail(path, t, "vertex id out of range");
    for (;;) {
      int64_t v, w, c;
      if (!t.integer(&v)) return fail(path, t, "expected a neighbour id or -1");
      if (v == -1) break;
      if (v < 0 || v >= n) return fail(path, t, "vertex id out of range");
      if (!t.integer(&w) || !t.integer(&c)) return fail(path, t, "expected the arc's length and delay");
      if (!weight_ok(w) || !weight_ok(c)) return fail(path, t, "weights must lie in [0, 2^24)");
      arcs.push_back({(int32_t)u, (int32_t)v, (float)w, (float)c});
    }
  }
  *n_out = (int32_t)n;
  return 0;
}

// SNAP edge lists, each row one undirected edge stored as both arcs with the
// same weights (io.hpp:31-59).  Two layouts:
//   the reference's: "N?M?" (two counts, each followed by one separator
//   character), then M rows "u v w c";
//   MatrixMarket, as shipped in inputs/tech-*.mtx.txt: '%' comment lines,
//   "N N M", then M rows "id u v w c" with 0-based ids.
int read_snap(const char* path, Text& t, int32_t* n_out, std::vector<Arc>& arcs) {
  const bool mtx = t.peek() == '%';
  while (t.peek() == '%') t.skip_line();
  int64_t n, m;
  if (mtx) {
    int64_t cols;
    if (!t.integer(&n) || !t.integer(&cols) || !t.integer(&m)) return fail(path, t, "expected \"N N M\"");
    if (cols != n) return fail(path, t, "the matrix must be square");
  } else {
    if (!t.integer(&n)) return fail(path, t, "expected the vertex count");
    if (t.pos < t.buf.size() && !std::isspace((unsigned char)t.buf[t.pos])) t.pos++;
    if (!t.integer(&m)) return fail(path, t, "expected the edge count");
    if (t.pos < t.buf.size() && !std::isspace((unsigned char)t.buf[t.pos])) t.pos++;
  }
  if (n <= 0 || n > INT32_MAX) return fail(path, t, "vertex count out of range");
  if (m < 0 || 2 * m > INT32_MAX) return fail(path, t, "edge count out of range");
  arcs.reserve((size_t)(2 * m));
  for (int64_t e = 0; e < m; e++) {
    int64_t id, u, v, w, c;
    if (mtx && !t.integer(&id)) return fail(path, t, "expected \"id u v w c\"");
    if (!t.integer(&u) || !t.integer(&v) || !t.integer(&w) || !t.integer(&c))
      return fail(path, t, mtx ? "expected \"id u v w c\"" : "expected \"u v w c\"");
    if (u < 0 || u >= n || v < 0 || v >= n) return fail(path, t, "vertex id out of range");
    if (!weight_ok(w) || !weight_ok(c)) return fail(path, t, "weights must lie in [0, 2^24)");
    arcs.push_back({(int32_t)u, (int32_t)v, (float)w, (float)c});
    arcs.push_back({(int32_t)v, (int32_t)u, (float)w, (float)c});
  }
  if (!t.at_end()) return fail(path, t, "more rows than the header's edge count");
  *n_out = (int32_t)n;
  return 0;
}

}  // namespace

extern "C" {

int vine_read_graph(const char* path, int32_t format, vine_host_graph** out) {
  if (!path || !out || (format != VINE_FILE_SNAP && format != VINE_FILE_COLA)) {
    vine::set_error("vine_read_graph: invalid arguments");
    return VINE_E_ARG;
  }
  Text t;
  if (!t.load(path)) {
    vine::set_error("vine_read_graph: cannot open %s", path);
    return VINE_E_ARG;
  }
  std::vector<Arc> arcs;
  int32_t n = 0;
  const int rc = format == VINE_FILE_COLA ? read_cola(path, t, &n, arcs) : read_snap(path, t, &n, arcs);
  if (rc != 0) return rc;
  if ((int64_t)arcs.size() > INT32_MAX) {
    vine::set_error("vine_read_graph: %s has more arcs than 32-bit CSR offsets hold", path);
    return VINE_E_ARG;
  }
  *out = to_csr(n, arcs);
  return 0;
}

int64_t vine_read_queries(const char* path, int32_t format, int64_t cap, int32_t* out_src, int32_t* out_dst,
                          float* out_bound, int32_t* out_aux) {
  if (!path || cap < 0 || (cap > 0 && (!out_src || !out_dst || !out_bound)) ||
      (format != VINE_FILE_SNAP && format != VINE_FILE_COLA)) {
    vine::set_error("vine_read_queries: invalid arguments");
    return VINE_E_ARG;
  }
  Text t;
  if (!t.load(path)) {
    vine::set_error("vine_read_queries: cannot open %s", path);
    return VINE_E_ARG;
  }
  if (format == VINE_FILE_SNAP) t.skip_line();  // header line
  int64_t count = 0;
  while (!t.at_end()) {
    int64_t f[6];
    const int cols = format == VINE_FILE_COLA ? 5 : 6;
    for (int k = 0; k < cols; k++) {
      if (format == VINE_FILE_SNAP && k > 0 && !t.expect(','))
        return fail(path, t, "expected \"id,src,snk,delay,hop,cost\"");
      if (!t.integer(&f[k]))
        return fail(path, t, format == VINE_FILE_COLA ? "expected \"u v lower upper limit\""
                                                      : "expected \"id,src,snk,delay,hop,cost\"");
    }
    // COLA: u v lower upper limit, bound = limit; CSV: id src snk delay hop cost, bound = delay
    const int64_t s = format == VINE_FILE_COLA ? f[0] : f[1];
    const int64_t d = format == VINE_FILE_COLA ? f[1] : f[2];
    const int64_t b = format == VINE_FILE_COLA ? f[4] : f[3];
    const int64_t a0 = format == VINE_FILE_COLA ? f[2] : f[4];
    const int64_t a1 = format == VINE_FILE_COLA ? f[3] : f[5];
    if (s < 0 || s > INT32_MAX || d < 0 || d > INT32_MAX) return fail(path, t, "vertex id out of range");
    if (!weight_ok(b)) return fail(path, t, "the bound must lie in [0, 2^24)");
    if (a0 < INT32_MIN || a0 > INT32_MAX || a1 < INT32_MIN || a1 > INT32_MAX)
      return fail(path, t, "value out of range");
    if (count < cap) {
      out_src[count] = (int32_t)s;
      out_dst[count] = (int32_t)d;
      out_bound[count] = (float)b;
      if (out_aux) {
        out_aux[2 * count] = (int32_t)a0;
        out_aux[2 * count + 1] = (int32_t)a1;
      }
    }
    count++;
  }
  return count;
}

}  // extern "C"
