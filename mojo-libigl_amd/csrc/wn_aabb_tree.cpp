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
#include <atomic>
#include <thread>
#include <utility>

#include "wn_closest.h"

namespace wn {

namespace {

// Preorder numbering is fixed by the subtree sizes alone (aabb_node_count),
// and leaves take the face ids in the order the recursion leaves them, so
// every subtree knows its node index and its slice of `order` up front and
// large subtrees are built on their own threads with identical results.
struct Builder {
  const std::vector<double>& lo;   // per face: corner minimum (3)
  const std::vector<double>& hi;   // per face: corner maximum (3)
  const std::vector<double>& mid;  // per face: 0.5 (lo + hi) (3)
  int64_t leaf;
  AabbTree& t;
  int64_t* ids;                    // face ids, reordered in place; becomes `order`
  std::atomic<int64_t> deepest{0};

  void node(int64_t me, int64_t b, int64_t n, int64_t depth, int spawn_levels) {
    int64_t seen = deepest.load(std::memory_order_relaxed);
    while (depth > seen && !deepest.compare_exchange_weak(seen, depth)) {
    }
    const int64_t* id = ids + b;
    double box[6] = {lo[3 * id[0]], lo[3 * id[0] + 1], lo[3 * id[0] + 2],
                     hi[3 * id[0]], hi[3 * id[0] + 1], hi[3 * id[0] + 2]};
    for (int64_t k = 1; k < n; ++k)
      for (int d = 0; d < 3; ++d) {
        box[d] = std::min(box[d], lo[3 * id[k] + d]);
        box[3 + d] = std::max(box[3 + d], hi[3 * id[k] + d]);
      }
    std::copy(box, box + 6, t.boxes.begin() + 6 * me);
    if (n <= leaf) {
      t.nodes[4 * me] = -1;
      t.nodes[4 * me + 1] = -1;
      t.nodes[4 * me + 2] = b;
      t.nodes[4 * me + 3] = n;
      return;
    }
    // split axis: largest spread of the face-box centres, first on ties
    int axis = 0;
    double best = -1.0;
    for (int d = 0; d < 3; ++d) {
      double a = mid[3 * id[0] + d], c = a;
      for (int64_t k = 1; k < n; ++k) {
        a = std::min(a, mid[3 * id[k] + d]);
        c = std::max(c, mid[3 * id[k] + d]);
      }
      if (d == 0 || c - a > best) {
        best = c - a;
        axis = d;
      }
    }
    // stable sort by centre along the axis: sort (key, position) pairs, which
    // orders exactly as a stable sort of the ids does
    std::vector<std::pair<double, int64_t>> kp((size_t)n);
    for (int64_t k = 0; k < n; ++k) kp[(size_t)k] = {mid[3 * id[k] + axis], k};
    std::sort(kp.begin(), kp.end());
    std::vector<int64_t> tmp(id, id + n);
    for (int64_t k = 0; k < n; ++k) ids[b + k] = tmp[(size_t)kp[(size_t)k].second];
    const int64_t half = n / 2;
    const int64_t l = me + 1;
    const int64_t r = me + 1 + aabb_node_count(half, leaf);
    t.nodes[4 * me] = l;
    t.nodes[4 * me + 1] = r;
    t.nodes[4 * me + 2] = 0;
    t.nodes[4 * me + 3] = 0;
    if (spawn_levels > 0 && n > 65536) {
      std::thread right([&, r, b, half, n, depth, spawn_levels] {
        node(r, b + half, n - half, depth + 1, spawn_levels - 1);
      });
      node(l, b, half, depth + 1, spawn_levels - 1);
      right.join();
    } else {
      node(l, b, half, depth + 1, 0);
      node(r, b + half, n - half, depth + 1, 0);
    }
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
  // leaves(n) = 1 for n <= leaf, else leaves(n / 2) + leaves(n - n / 2); at
  // each depth the sizes take at most two adjacent values, so the recursion
  // is walked level by level over (size, multiplicity) pairs
  if (nf <= 0) return 0;
  std::vector<std::pair<int64_t, int64_t>> level{{nf, 1}};
  int64_t leaves = 0;
  while (!level.empty()) {
    std::vector<std::pair<int64_t, int64_t>> next;
    auto add = [&next](int64_t size, int64_t times) {
      for (auto& p : next)
        if (p.first == size) {
          p.second += times;
          return;
        }
      next.push_back({size, times});
    };
    for (const auto& p : level) {
      if (p.first <= leaf) {
        leaves += p.second;
      } else {
        add(p.first / 2, p.second);
        add(p.first - p.first / 2, p.second);
      }
    }
    level.swap(next);
  }
  return 2 * leaves - 1;
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
  out.boxes.assign(6 * count, 0.0);
  out.nodes.assign(4 * count, 0);
  out.order.resize(nf);
  for (int64_t f = 0; f < nf; ++f) out.order[f] = f;
  Builder b{lo, hi, mid, leaf, out, out.order.data()};
  b.node(0, 0, nf, 1, 3);  // up to 8 threads on the top three levels
  out.stack_size = b.deepest.load() + 1;
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
