// Host-side CSR behind the vine_host_graph handle of include/vine_workload.h.
#pragma once

#include <cstdint>
#include <vector>

struct vine_host_graph {
  int32_t n = 0;
  std::vector<int32_t> row_ptr;
  std::vector<int32_t> dst;
  std::vector<float> len;
  std::vector<float> delay;
};
