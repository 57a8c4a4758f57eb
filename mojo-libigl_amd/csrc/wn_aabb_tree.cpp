// Point-mesh squared distance, host side: the median-split tree builder and
// the CPU query path behind the reference's C ABI.
//
//   aabb_build                 python/igl/__init__.py:286-319 (the tree the
//                              reference builds in NumPy, here in C++; same
//                              node order, boxes and leaf order)
//   mli_aabb_squared_distance  src/capi.mojo:950-959 ABI over the descent of
//                              src/capi.mojo:508-587
//
// Built with -ffp-contract=off; closest_on_triangle / box_distance2 come from
// wn_closest.h, the same code the GPU kernel runs.

#include "wn_aabb.h"

#include <algorithm>

#include "wn_closest.h"

namespace wn {

namespace {

struct Builder {
  const std::vector<double>& lo;   // per face: corner minimum (3)
  const std::vector<double>& hi;   // per face: corner maximum (3)
  const std::vector<double>& mid;  // per face: 0.5 (lo + hi) (3)
  int64_t leaf;
  AabbTree& t;
  int64_t deepest = 0;

  int64_t node(int64_t* ids, int64_t n, int64_t depth) {
    deepest = std::max(deepest, depth);
    const int64_t me = (int64_t)t.nodes.size() / 4;
    t.nodes.insert(t.nodes.end(), {-1, -1, 0, 0});
    double box[6] = {lo[3 * ids[0]], lo[3 * ids[0] + 1], lo[3 * ids[0] + 2],
                     hi[3 * ids[0]], hi[3 * ids[0] + 1], hi[3 * ids[0] + 2]};
    for (int64_t k = 1; k < n; ++k)
      for (int d = 0; d < 3; ++d) {
        box[d] = std::min(box[d], lo[3 * ids[k] + d]);
        box[3 + d] = std::max(box[3 + d], hi[3 * ids[k] + d]);
      }
    t.boxes.insert(t.boxes.end(), box, box + 6);
    if (n <= leaf) {
      t.nodes[4 * me + 2] = (int64_t)t.order.size();
      t.nodes[4 * me + 3] = n;
      t.order.insert(t.order.end(), ids, ids + n);
      return me;
    }
    // split axis: largest spread of the face-box centres, first on ties
    int axis = 0;
    double best = -1.0;
    for (int d = 0; d < 3; ++d) {
      double a = mid[3 * ids[0] + d], b = a;
      for (int64_t k = 1; k < n; ++k) {
        a = std::min(a, mid[3 * ids[k] + d]);
        b = std::max(b, mid[3 * ids[k] + d]);
      }
      if (d == 0 || b - a > best) {
        best = b - a;
        axis = d;
      }
    }
    const double* key = mid.data() + axis;
    std::stable_sort(ids, ids + n, [key](int64_t x, int64_t y) { return key[3 * x] < key[3 * y]; });
    const int64_t half = n / 2;
    const int64_t l = node(ids, half, depth + 1);
    const int64_t r = node(ids + half, n - half, depth + 1);
    t.nodes[4 * me] = l;
    t.nodes[4 * me + 1] = r;
    return me;
  }
};

template <typename FT>
void face_bounds(const double* V, const FT* F, int64_t nf, std::vector<double>& lo,
                 std::vector<double>& hi, std::vector<double>& mid) {
  lo.resize(3 * nf);
  hi.resize(3 * nf);
  mid.resize(3 * nf);
  for (int64_t f = 0; f < nf; ++f)
    for (int d = 0; d < 3; ++d) {
      const double a = V[3 * (int64_t)F[3 * f] + d], b = V[3 * (int64_t)F[3 * f + 1] + d],
                   c = V[3 * (int64_t)F[3 * f + 2] + d];
      lo[3 * f + d] = std::min(std::min(a, b), c);
      hi[3 * f + d] = std::max(std::max(a, b), c);
      mid[3 * f + d] = 0.5 * (lo[3 * f + d] + hi[3 * f + d]);
    }
}

}  // namespace

int64_t aabb_node_count(int64_t nf, int64_t leaf) {
  if (nf <= 0) return 0;
  if (nf <= leaf) return 1;
  return 1 + aabb_node_count(nf / 2, leaf) + aabb_node_count(nf - nf / 2, leaf);
}

void aabb_build(const double* V, const void* F, int fdtype, int64_t nf, int64_t leaf,
                AabbTree& out) {
  out = AabbTree{};
  if (nf <= 0) return;
  std::vector<double> lo, hi, mid;
  if (fdtype == 0)
    face_bounds(V, (const int64_t*)F, nf, lo, hi, mid);
  else
    face_bounds(V, (const int32_t*)F, nf, lo, hi, mid);
  const int64_t count = aabb_node_count(nf, leaf);
  out.boxes.reserve(6 * count);
  out.nodes.reserve(4 * count);
  out.order.reserve(nf);
  std::vector<int64_t> ids(nf);
  for (int64_t f = 0; f < nf; ++f) ids[f] = f;
  Builder b{lo, hi, mid, leaf, out};
  b.node(ids.data(), nf, 1);
  out.stack_size = b.deepest + 1;
}

}  // namespace wn

extern "C" void mli_aabb_squared_distance(const double* p, const double* v, const int64_t* f,
                                          const double* boxes, const int64_t* nodes,
                                          const int64_t* order, double* sqrd, int64_t* indices,
                                          double* closest, int64_t* stack, int64_t np,
                                          int64_t nnodes, int64_t q0, int64_t q1) {
  (void)np;
  for (int64_t qi = q0; qi < q1; ++qi) {
    const double px = p[3 * qi], py = p[3 * qi + 1], pz = p[3 * qi + 2];
    double best = 1.0e300, cx = 0.0, cy = 0.0, cz = 0.0;
    int64_t best_f = -1, top = 0;
    stack[top++] = 0;
    while (top > 0) {
      const int64_t n = stack[--top];
      if (n < 0 || n >= nnodes) continue;
      const double* b = boxes + 6 * n;
      if (wn::box_distance2(px, py, pz, b[0], b[1], b[2], b[3], b[4], b[5]) > best) continue;
      const int64_t* nd = nodes + 4 * n;
      if (nd[3] > 0) {
        for (int64_t k = 0; k < nd[3]; ++k) {
          const int64_t face = order[nd[2] + k];
          const double* A = v + 3 * f[3 * face];
          const double* B = v + 3 * f[3 * face + 1];
          const double* C = v + 3 * f[3 * face + 2];
          const wn::ClosestHit h = wn::closest_on_triangle(px, py, pz, A[0], A[1], A[2], B[0],
                                                           B[1], B[2], C[0], C[1], C[2]);
          if (h.d2 < best) {
            best = h.d2;
            best_f = face;
            cx = h.x; cy = h.y; cz = h.z;
          }
        }
      } else {
        const double* bl = boxes + 6 * nd[0];
        const double* br = boxes + 6 * nd[1];
        const double dl = wn::box_distance2(px, py, pz, bl[0], bl[1], bl[2], bl[3], bl[4], bl[5]);
        const double dr = wn::box_distance2(px, py, pz, br[0], br[1], br[2], br[3], br[4], br[5]);
        // the nearer child goes on top of the stack
        if (dl < dr) {
          if (dr <= best) stack[top++] = nd[1];
          if (dl <= best) stack[top++] = nd[0];
        } else {
          if (dl <= best) stack[top++] = nd[0];
          if (dr <= best) stack[top++] = nd[1];
        }
      }
    }
    sqrd[qi] = best;
    indices[qi] = best_f;
    closest[3 * qi] = cx;
    closest[3 * qi + 1] = cy;
    closest[3 * qi + 2] = cz;
  }
}
