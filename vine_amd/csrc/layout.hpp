// Vertex order for HBM locality (layout.cpp).
#pragma once

#include <cstdint>
#include <vector>

namespace vine {

// order[new id] = caller's id: a permutation of 0..n-1 that gives the
// vertices of each connected component nearby ids, along a Hilbert curve over
// hop distances from two far-apart vertices.
std::vector<int32_t> locality_order(int32_t n, const int32_t* row, const int32_t* dst);

}  // namespace vine
