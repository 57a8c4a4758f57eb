// Median-split bounding-box tree for point-mesh squared distance (host side,
// plain C++; built by wn_aabb_tree.cpp, uploaded and walked by wn_aabb.hip).
#pragma once

#include <stdint.h>

#include <vector>

namespace wn {

// The reference's tree arrays (python/igl/__init__.py:286-319): per node in
// preorder, boxes[6 n .. 6 n + 5] = (lo xyz, hi xyz) and nodes[4 n .. 4 n + 3]
// = (left, right, start, count); count > 0 marks a leaf whose faces are
// order[start .. start + count). stack_size = deepest level + 1 bounds the
// descent stack.
struct AabbTree {
  std::vector<double> boxes;
  std::vector<int64_t> nodes;
  std::vector<int64_t> order;
  int64_t stack_size = 0;
};

// Node count of the tree over nf faces with leaves of <= leaf faces.
int64_t aabb_node_count(int64_t nf, int64_t leaf);

// Builds the tree of faces F (nf x 3, fdtype 0 = int64, 1 = int32) over V
// (nv x 3 float64). Indices must already be validated.
void aabb_build(const double* V, const void* F, int fdtype, int64_t nf, int64_t leaf,
                AabbTree& out);

}  // namespace wn
