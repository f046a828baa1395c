// Builds a 4x4 grid through the C++ facade (include/vine/cspp.hpp), solves a
// few tasks and prints one line per task: feasible cost delay path...
// With argument "args" it only checks the argument errors (no GPU needed);
// with "files ..." it reads a graph and a query file (see run_files); with
// "hop" or "qos" it solves the grid for another label family (run_family).
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "vine/cspp.hpp"

// files <graph> <graph_type> <queries> <graph_type> [solve]: reads both files
// through read_graph / task::read_query, prints "graph N M" and "tasks n",
// then with "solve" one line per task as above (path omitted).
static int run_files(char** argv, bool solve) {
  using namespace vine;
  Graph<c_edge> g(1);
  int st = read_graph(argv[2], &g, std::atoi(argv[3]));
  if (st != VINE_OK) {
    std::printf("error %d %s\n", st, vine_last_error());
    return 2;
  }
  task tasks;
  st = tasks.read_query(argv[4], std::atoi(argv[5]));
  if (st != VINE_OK) {
    std::printf("error %d %s\n", st, vine_last_error());
    return 2;
  }
  if (std::atoi(argv[5]) == 1)  // main.cu:19-26 folds SNAP task ids into range
    for (int i = 0; i < tasks.ntask; i++) {
      tasks.source_l_[i] %= g.N;
      tasks.sink_l_[i] %= g.N;
    }
  std::printf("graph %d %lld\ntasks %d\n", g.N, (long long)g.M, tasks.ntask);
  if (!solve) return 0;
  std::vector<c_label> src;
  std::vector<c_constraint<c_label>> C;
  for (int i = 0; i < tasks.ntask; i++) {
    src.push_back(c_label(tasks.source_l_[i]));
    C.push_back(c_constraint<c_label>((float)tasks.delay_l_[i]));
  }
  cspp<c_label, c_edge> solver(&g, tasks.ntask, src, tasks.sink_l_, C);
  st = solver.solve();
  if (st != VINE_OK) {
    std::printf("error %d %s\n", st, solver.last_error.c_str());
    return 2;
  }
  for (int i = 0; i < tasks.ntask; i++)
    std::printf("%d %.0f %.0f\n", solver.feasible_l[i], solver.cost_l[i], solver.delay_l[i]);
  return 0;
}

static vine::Graph<vine::c_edge> grid4x4() {
  using namespace vine;
  Graph<c_edge> g(16);
  for (int r = 0; r < 4; r++)
    for (int c = 0; c < 4; c++) {
      int id = r * 4 + c;
      if (c + 1 < 4) g.add_undirected(id, id + 1, c_edge(1.f + (float)((id * 7) % 5), 1.f + (float)((id * 3) % 4)));
      if (r + 1 < 4) g.add_undirected(id, id + 4, c_edge(2.f + (float)((id * 5) % 3), 1.f + (float)((id * 11) % 3)));
    }
  return g;
}

// hop | qos: the same grid through cspp<Label_t> for another label family;
// one line per task: feasible d delay hop path...
template <typename Label_t>
static int run_family(std::vector<typename vine::family_traits<Label_t>::constraint> C) {
  using namespace vine;
  Graph<c_edge> g = grid4x4();
  std::vector<Label_t> src = {Label_t(0), Label_t(0), Label_t(5), Label_t(3)};
  std::vector<int> snk = {15, 15, 10, 3};
  cspp<Label_t, c_edge> solver(&g, 4, src, snk, C);
  int st = solver.solve();
  if (st != VINE_OK) {
    std::printf("error %d %s\n", st, solver.last_error.c_str());
    return 2;
  }
  for (int i = 0; i < 4; i++) {
    std::printf("%d %.0f %.0f %.0f", solver.feasible_l[i], solver.label_l[4 * i], solver.label_l[4 * i + 1],
                solver.label_l[4 * i + 2]);
    for (int v : solver.path_l[i]) std::printf(" %d", v);
    std::printf("\n");
  }
  return 0;
}

int main(int argc, char** argv) {
  using namespace vine;
  if (argc >= 6 && std::strcmp(argv[1], "files") == 0)
    return run_files(argv, argc > 6 && std::strcmp(argv[6], "solve") == 0);
  if (argc > 1 && std::strcmp(argv[1], "hop") == 0) {
    using C = hop::c_constraint<hop::c_label>;
    return run_family<hop::c_label>({C(100.f, 6), C(100.f, 5), C(4.f), C(0.f)});
  }
  if (argc > 1 && std::strcmp(argv[1], "qos") == 0) {
    using C = qos::c_constraint<qos::c_label>;
    return run_family<qos::c_label>({C(1.f), C(2.f), C(3.f), C(1.f)});
  }
  Graph<c_edge> g = grid4x4();
  std::vector<c_label> src = {c_label(0), c_label(0), c_label(5), c_label(3)};
  std::vector<int> snk = {15, 15, 10, 3};
  std::vector<c_constraint<c_label>> C = {c_constraint<c_label>(100.f), c_constraint<c_label>(7.f),
                                          c_constraint<c_label>(2.f), c_constraint<c_label>(0.f)};
  if (argc > 1 && std::strcmp(argv[1], "args") == 0) {
    std::vector<int> bad = {15, 15, 10};
    try {
      cspp<c_label, c_edge> s(&g, 4, src, bad, C);
      std::printf("no exception\n");
      return 1;
    } catch (const std::invalid_argument&) {
      std::printf("args ok\n");
      return 0;
    }
  }
  cspp<c_label, c_edge> solver(&g, 4, src, snk, C);
  int st = solver.solve();
  if (st != VINE_OK) {
    std::printf("error %d %s\n", st, solver.last_error.c_str());
    return 2;
  }
  for (int i = 0; i < 4; i++) {
    std::printf("%d %.0f %.0f", solver.feasible_l[i], solver.cost_l[i], solver.delay_l[i]);
    for (int v : solver.path_l[i]) std::printf(" %d", v);
    std::printf("\n");
  }
  return 0;
}
