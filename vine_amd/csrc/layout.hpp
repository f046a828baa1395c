// Vertex order for HBM locality (layout.cpp).
#pragma once

#include <cstdint>
#include <vector>

namespace vine {

// order[new id] = caller's id: a permutation of 0..n-1 that gives the
// vertices of each connected component nearby ids, along a Hilbert curve over
// hop distances from two far-apart vertices.
std::vector<int32_t> locality_order(int32_t n, const int32_t* row, const int32_t* dst);

// Delay distances from K landmarks (K <= 4), chosen farthest-first: the
// first is the vertex farthest from vertex 0, each next one the vertex whose
// nearest landmark is farthest.  out[v * 4 + j] = min delay from landmark j to
// v, -1 where v is not reachable (and for j >= K).  Dijkstra over the CSR.
void landmark_distances(int32_t n, const int32_t* row, const int32_t* dst, const float* delay, int K,
                        std::vector<float>& out);

}  // namespace vine
