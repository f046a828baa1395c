// Device-resident graph behind the vine_graph handle (include/vine.h),
// shared by the solver engines.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

struct vine_graph {
  int32_t n = 0;
  int64_t m = 0;
  float mean_len = 1.f;
  float mean_delay = 1.f;
  int32_t max_deg = 0;
  float max_len = 0.f;
  float max_delay = 0.f;
  bool integer_weights = true;  // every len and delay an integer below 2^24 (path sums exact)
  // every arc u->v has a twin v->u of no larger delay, so the min delay from
  // t to any v is at most the min delay from v to t
  bool twin_delay = false;
  int device = 0;
  // vertices are renumbered for locality (layout.cpp): perm[caller id] =
  // internal id, d_inv the inverse; every array below uses internal ids
  std::vector<int32_t> perm;
  int32_t* d_perm = nullptr;
  int32_t* d_inv = nullptr;
  int32_t* d_row = nullptr;
  int4* d_edges = nullptr;
  int32_t* d_rrow = nullptr;
  int4* d_redges = nullptr;
  // min delays from four landmarks to every vertex (-1: unreachable), for the
  // forward bound search's region test; null unless every delay is an
  // integer and every such distance is below 2^24 (exact as floats)
  float4* d_lmk = nullptr;
};
