// vine/cspp.hpp -- C++ host facade over the C ABI (include/vine.h).
//
// Mirrors the reference's solver facade cspp<Label_t, Edge_t>
// (algorithm/cspp.hpp:11-69) and its graph builder Graph<Edge_t>
// (include/graph.hpp:90-195): build a graph from arcs, construct the solver
// with the number of tasks, the source labels, the sinks and one constraint
// per task, call solve().  Results land on the object (cost_l, delay_l,
// feasible_l, path_l, size_of_result_l); the reference allocates path_l but
// never fills it.  Differences: solve() returns the vine_status instead of
// always 0, and every task uses its own constraint (the reference passes
// C_l[0] for all tasks, algorithm/cspp.cuh:196,234).
//
// cspp<c_label, c_edge> runs the batched solver (vine_solve); the other label
// families of vine/families.hpp (hop-bounded, hop + max arc length, 2-OSP,
// social, QoS) run the generic search (vine_family_solve), with the answer's
// fields in label_l.  A new label type needs a family_traits specialisation
// and the kernel templates of vine_amd/csrc/families.hip compiled for it.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "vine.h"
#include "vine/families.hpp"
#include "vine/label.hpp"
#include "vine_workload.h"

namespace vine {

// Host graph in adjacency insertion order (the CSR order of
// Graph::adj_to_csr, include/graph.hpp:136-167).
template <typename Edge_t>
class Graph {
 public:
  explicit Graph(int n) : N(n), M(0), adj_(n) {}
  // directed arc u -> v
  void add_edge(int u, int v, const Edge_t& e) {
    if (u < 0 || u >= N || v < 0 || v >= N) throw std::out_of_range("vertex id");
    adj_[u].push_back({v, e});
    M++;
  }
  // both arcs with the same weights (the reference's undirected loader,
  // include/io.hpp:50-60)
  void add_undirected(int u, int v, const Edge_t& e) {
    add_edge(u, v, e);
    add_edge(v, u, e);
  }
  void adj_to_csr(std::vector<int32_t>& row, std::vector<int32_t>& dst, std::vector<float>& len,
                  std::vector<float>& delay) const {
    row.assign(N + 1, 0);
    dst.clear();
    len.clear();
    delay.clear();
    for (int u = 0; u < N; u++) {
      for (const auto& a : adj_[u]) {
        dst.push_back(a.first);
        len.push_back(a.second.len);
        delay.push_back(a.second.delay);
      }
      row[u + 1] = (int32_t)dst.size();
    }
  }
  int N;
  int64_t M;

 private:
  std::vector<std::vector<std::pair<int, Edge_t>>> adj_;
};

// Label_t: c_label (the batched solver of vine_solve) or any label family of
// vine/families.hpp (the generic search of vine_family_solve); the
// constraint type is the family's c_constraint<Label_t>.
template <typename Label_t, typename Edge_t>
class cspp {
  static_assert(std::is_same<Edge_t, c_edge>::value, "the C ABI carries c_edge arcs");
  using Traits = family_traits<Label_t>;
  using Constraint_t = typename Traits::constraint;

 public:
  cspp(Graph<Edge_t>* g, int ntask, std::vector<Label_t>& source_state, std::vector<int>& sink_l,
       std::vector<Constraint_t>& C_l)
      : g_(g), ntask(ntask), init_state(source_state), sink_l(sink_l), C_l(C_l) {
    vine_options_default(&options);
    if ((int)init_state.size() != ntask || (int)sink_l.size() != ntask || (int)C_l.size() != ntask)
      throw std::invalid_argument("ntask must match the sources, sinks and constraints");
  }

  // Solve every task on the GPU.  Returns VINE_OK or the failing vine_status
  // (message in last_error).
  int solve() {
    std::vector<int32_t> row, dst;
    std::vector<float> len, delay;
    g_->adj_to_csr(row, dst, len, delay);
    vine_graph* dg = nullptr;
    vine_status st = vine_graph_create(g_->N, (int64_t)dst.size(), row.data(), dst.data(), len.data(),
                                       delay.data(), &dg);
    if (st != VINE_OK) return fail(st);
    if (Traits::family != VINE_FAMILY_CSP) {
      const int fst = solve_family(dg);
      vine_graph_destroy(dg);
      return fst;
    }
    vine_solver* sv = nullptr;
    st = vine_solver_create(dg, &options, &sv);
    if (st != VINE_OK) {
      vine_graph_destroy(dg);
      return fail(st);
    }
    std::vector<int32_t> src(ntask), snk(sink_l.begin(), sink_l.end());
    std::vector<float> bound(ntask);
    for (int i = 0; i < ntask; i++) {
      src[i] = init_state[i].vid;
      bound[i] = C_l[i].delay_L;
    }
    cost_l.assign(ntask, 0.f);
    delay_l.assign(ntask, 0.f);
    std::vector<int32_t> feas(ntask, 0);
    std::vector<int64_t> off(ntask + 1, 0);
    std::vector<int32_t> path((size_t)ntask * 256 + 1);  // grown to the reported size if short
    st = vine_solve(sv, ntask, src.data(), snk.data(), bound.data(), cost_l.data(), delay_l.data(), feas.data(),
                    off.data(), path.data(), (int64_t)path.size(), &stats);
    if (st == VINE_E_OVERFLOW && off[ntask] > (int64_t)path.size()) {
      path.resize(off[ntask]);
      if (ntask <= options.max_batch)  // one batch: its paths are still on the device
        st = vine_solver_paths(sv, ntask, off.data(), path.data(), (int64_t)path.size());
      else
        st = vine_solve(sv, ntask, src.data(), snk.data(), bound.data(), cost_l.data(), delay_l.data(),
                        feas.data(), off.data(), path.data(), (int64_t)path.size(), &stats);
    }
    if (st == VINE_OK) {
      feasible_l.assign(feas.begin(), feas.end());
      path_l.assign(ntask, {});
      size_of_result_l.assign(ntask, 0);
      for (int i = 0; i < ntask; i++) {
        path_l[i].assign(path.begin() + off[i], path.begin() + off[i + 1]);
        size_of_result_l[i] = (int)path_l[i].size();
      }
    } else {
      last_error = vine_last_error();
    }
    vine_solver_destroy(sv);
    vine_graph_destroy(dg);
    return st;
  }

  vine_options options;
  vine_stats stats{};
  int ntask;
  std::vector<Label_t> init_state;
  std::vector<int> sink_l;
  std::vector<Constraint_t> C_l;
  std::vector<float> cost_l, delay_l;
  std::vector<int> feasible_l;
  std::vector<std::vector<int>> path_l;
  std::vector<int> size_of_result_l;
  // other families: the answer's (d, delay, hop, maxcost) per task
  std::vector<float> label_l;
  std::string last_error;

 private:
  int fail(vine_status st) {
    last_error = vine_last_error();
    return st;
  }

  // the generic label-correcting search for the other families
  int solve_family(vine_graph* dg) {
    vine_family_solver* fs = nullptr;
    const int batch = options.max_batch > 1 ? options.max_batch : (ntask < 64 ? (ntask > 0 ? ntask : 1) : 64);
    // only the answers are returned: target pruning unless the options ask
    // for the reference's one-to-all search (VINE_PRUNE_NONE)
    // (a fixed window, VINE_DELTA_FIXED with delta_fixed > 0, orders the rounds)
    const float window = options.delta_policy == VINE_DELTA_FIXED ? options.delta_fixed : 0.f;
    vine_status st = vine_family_solver_create(dg, Traits::family, batch, 0, options.prune != VINE_PRUNE_NONE ? 1 : 0,
                                               window, &fs);
    if (st != VINE_OK) return fail(st);
    std::vector<int32_t> src(ntask), snk(sink_l.begin(), sink_l.end());
    std::vector<float> lim(3 * (size_t)ntask);
    for (int i = 0; i < ntask; i++) {
      src[i] = init_state[i].vid;
      Traits::limits(C_l[i], &lim[3 * (size_t)i]);
    }
    label_l.assign(4 * (size_t)ntask, 0.f);
    std::vector<int32_t> feas(ntask, 0);
    std::vector<int64_t> off(ntask + 1, 0);
    std::vector<int32_t> path((size_t)ntask * 256 + 1);  // retried at the reported size if short
    st = vine_family_solve(fs, ntask, src.data(), snk.data(), lim.data(), label_l.data(), feas.data(), off.data(),
                           path.data(), (int64_t)path.size(), nullptr);
    if (st == VINE_E_OVERFLOW && off[ntask] > (int64_t)path.size()) {
      path.resize(off[ntask]);
      st = vine_family_solve(fs, ntask, src.data(), snk.data(), lim.data(), label_l.data(), feas.data(),
                             off.data(), path.data(), (int64_t)path.size(), nullptr);
    }
    if (st == VINE_OK) {
      cost_l.assign(ntask, 0.f);
      delay_l.assign(ntask, 0.f);
      feasible_l.assign(feas.begin(), feas.end());
      path_l.assign(ntask, {});
      size_of_result_l.assign(ntask, 0);
      for (int i = 0; i < ntask; i++) {
        cost_l[i] = label_l[4 * (size_t)i];
        delay_l[i] = label_l[4 * (size_t)i + 1];
        path_l[i].assign(path.begin() + off[i], path.begin() + off[i + 1]);
        size_of_result_l[i] = (int)path_l[i].size();
      }
    } else {
      last_error = vine_last_error();
    }
    vine_family_solver_destroy(fs);
    return st;
  }
  Graph<Edge_t>* g_;
};

// read_graph (include/io.hpp:22-91): replaces *g with the graph in file_name.
// graph_type 2 = COLA road network (arcs as listed), 1 = SNAP edge list
// (undirected; the reference's "N?M?" layout or the shipped MatrixMarket
// files).  Returns a vine_status (the reference returns 0 and crashes on a
// missing file); the message is in vine_last_error().
template <typename Edge_t>
int read_graph(const char* file_name, Graph<Edge_t>* g, int graph_type) {
  vine_host_graph* h = nullptr;
  const int st = vine_read_graph(file_name, graph_type, &h);
  if (st != 0) return st;
  const int32_t n = vine_host_graph_n(h);
  const int32_t* row = vine_host_graph_row_ptr(h);
  const int32_t* dst = vine_host_graph_dst(h);
  const float* len = vine_host_graph_len(h);
  const float* delay = vine_host_graph_delay(h);
  *g = Graph<Edge_t>(n);
  for (int32_t u = 0; u < n; u++)
    for (int32_t p = row[u]; p < row[u + 1]; p++) {
      Edge_t e;
      e.len = len[p];
      e.delay = delay[p];
      g->add_edge(u, dst[p], e);
    }
  vine_host_graph_free(h);
  return VINE_OK;
}

// task (include/task.hpp:9-79): the query list of a run.  read_query with
// graph_type 2 (COLA rows "u v lower upper limit") or 1 (a header line, then
// "id,src,snk,delay,hop,cost"); load_query is the COLA form.  delay_l_ holds
// the delay bounds; hop_l_ and cost_l_ are filled by the CSV form.
class task {
 public:
  task() = default;
  explicit task(int n) : ntask(n) {}

  int read_query(const char* file_name, int graph_type) {
    const int64_t n = vine_read_queries(file_name, graph_type, 0, nullptr, nullptr, nullptr, nullptr);
    if (n < 0) return (int)n;
    std::vector<int32_t> src((size_t)n), dst((size_t)n), aux(2 * (size_t)n);
    std::vector<float> bound((size_t)n);
    const int64_t got = vine_read_queries(file_name, graph_type, n, src.data(), dst.data(), bound.data(),
                                          aux.data());
    if (got != n) return got < 0 ? (int)got : VINE_E_ARG;
    for (int64_t i = 0; i < n; i++) {
      source_l_.push_back(src[i]);
      sink_l_.push_back(dst[i]);
      delay_l_.push_back((int)bound[i]);
      if (graph_type == VINE_FILE_SNAP) {
        hop_l_.push_back(aux[2 * i]);
        cost_l_.push_back(aux[2 * i + 1]);
      }
    }
    ntask = (int)source_l_.size();
    return VINE_OK;
  }
  int load_query(const char* file_name) { return read_query(file_name, VINE_FILE_COLA); }

  int ntask = 0;
  std::vector<int> source_l_;
  std::vector<int> sink_l_;
  std::vector<int> delay_l_;
  std::vector<int> hop_l_;
  std::vector<int> cost_l_;
};

}  // namespace vine
