// Hierarchical ("fast") winding number: per-node far-field expansions over
// the median-split tree (host side, plain C++; wn_fwn_tree.cpp builds it and
// runs the CPU query, wn_fwn.hip the GPU query).
//
// No counterpart in the reference (README.md:29 lists the fast winding
// number as not covered); the method is Barill et al., "Fast Winding Numbers
// for Soups and Clouds" (SIGGRAPH 2018): the solid angle of a node's
// triangles seen from q,
//   Omega(q) = sum_t int_t n . (x - q) / |x - q|^3 dA,
// Taylor-expanded to second order about the node's area-weighted centroid c
// (r = c - q, y = x - c):
//   Omega ~ N . f(r) + sum_ij J_ij(r) C_ij + 1/2 sum_ijk H_ijk(r) D_ijk
//   f = r / |r|^3, J = d f / d x, H = d^2 f / d x^2,
//   N_i = sum_t A_t n_ti,  C_ij = sum_t A_t n_ti ybar_tj,
//   D_ijk = sum_t A_t n_ti (1/12) (sum_v y_vj y_vk + s_j s_k),  s = sum_v y_v
// (the last is the exact integral of y y^T over a triangle). A node is used
// through its expansion when |q - c| > beta * R (R = farthest corner of its
// faces from c); leaves closer than that are summed exactly with the
// per-pair formula of src/capi.mojo:218-235. w = sum / (4 pi).
#pragma once

#include <stdint.h>

#include <vector>

namespace wn {

// Per node (preorder, same numbering as AabbTree): expansion centre (x, y, z)
// and radius R, N (3), C (3 x 3 row-major), D (3 x 6: i, then (j, k) in
// 00 01 02 11 12 22 order), then children and leaf range.
struct FwnNode {
  double cx, cy, cz, radius;
  double n[3];
  double c[9];
  double d[18];
  int64_t left, right, start, count;  // count > 0: leaf of faces order[start..)
};

struct FwnTree {
  std::vector<FwnNode> nodes;
  std::vector<int64_t> order;    // leaf-order face ids
  std::vector<double> corners;   // 9 per leaf-order face (A, B, C)
  int64_t stack_size = 0;
};

// Builds the tree (leaves of <= leaf faces) and every node's expansion.
// Indices must already be validated.
void fwn_build(const double* V, const void* F, int fdtype, int64_t nf, int64_t leaf,
               FwnTree& out);

// w for queries q[q0:q1] (host), accuracy parameter beta (> 0).
void fwn_query_cpu(const FwnTree& t, const double* q, double* w, double beta, int64_t q0,
                   int64_t q1);

}  // namespace wn
