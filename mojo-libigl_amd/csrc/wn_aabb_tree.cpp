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
#include <cstring>
#include <thread>
#include <utility>

#include "wn_closest.h"

namespace wn {

namespace {

// Face-box centre and face id, permuted in place by the node sorts: every
// per-node pass reads a contiguous slice.
struct FaceKey {
  double mid[3];
  int64_t id;
};

// Runs fn(begin, end) over [0, n) in `workers` contiguous chunks, the first
// on the calling thread.
template <typename Fn>
void parallel_chunks(int64_t n, int workers, Fn fn) {
  if (workers <= 1 || n < 2 * 4096) {
    fn((int64_t)0, n);
    return;
  }
  std::vector<std::thread> pool;
  for (int w = 1; w < workers; ++w)
    pool.emplace_back([&, w] { fn(n * w / workers, n * (w + 1) / workers); });
  fn((int64_t)0, n / workers);
  for (auto& th : pool) th.join();
}

// Order-preserving 64-bit key of a finite double (-0 and +0 map alike, as
// they compare equal).
inline uint64_t sort_key(double x) {
  uint64_t u;
  const double y = x + 0.0;  // -0 -> +0
  std::memcpy(&u, &y, sizeof u);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

using KeyPos = std::pair<uint64_t, int64_t>;

// Sorts kp[0 .. n) with `workers` threads: chunks sorted on their own, then
// merged pairwise. The pairs are distinct (positions are), so the result is
// the one sorted order whatever the split.
void sort_pairs(KeyPos* kp, int64_t n, int workers) {
  if (workers <= 1 || n < 65536) {
    std::sort(kp, kp + n);
    return;
  }
  int chunks = 1;
  while (chunks * 2 <= workers) chunks *= 2;
  std::vector<int64_t> cut((size_t)chunks + 1);
  for (int c = 0; c <= chunks; ++c) cut[(size_t)c] = n * c / chunks;
  {
    std::vector<std::thread> pool;
    for (int c = 1; c < chunks; ++c)
      pool.emplace_back([&, c] { std::sort(kp + cut[(size_t)c], kp + cut[(size_t)c + 1]); });
    std::sort(kp, kp + cut[1]);
    for (auto& th : pool) th.join();
  }
  std::vector<KeyPos> buf((size_t)n);
  KeyPos* src = kp;
  KeyPos* dst = buf.data();
  for (int width = 1; width < chunks; width *= 2) {
    std::vector<std::thread> pool;
    for (int c = 0; c < chunks; c += 2 * width) {
      const int64_t lo = cut[(size_t)c], mid = cut[(size_t)std::min(c + width, chunks)],
                    hi = cut[(size_t)std::min(c + 2 * width, chunks)];
      pool.emplace_back([=] { std::merge(src + lo, src + mid, src + mid, src + hi, dst + lo); });
    }
    for (auto& th : pool) th.join();
    std::swap(src, dst);
  }
  if (src != kp) std::copy(src, src + n, kp);
}

// Stable sort of r[0 .. n) by mid[axis] (ties keep their order), through
// `tmp`: the order a stable sort of the ids by centre gives, as the one
// sorted order of the distinct (key, position) pairs.
void sort_by_axis(FaceKey* r, FaceKey* tmp, int64_t n, int axis, int workers) {
  std::vector<KeyPos> kp((size_t)n);
  parallel_chunks(n, workers, [&](int64_t k0, int64_t k1) {
    for (int64_t k = k0; k < k1; ++k) kp[(size_t)k] = {sort_key(r[k].mid[axis]), k};
  });
  sort_pairs(kp.data(), n, workers);
  parallel_chunks(n, workers, [&](int64_t k0, int64_t k1) {
    for (int64_t k = k0; k < k1; ++k) tmp[k] = r[kp[(size_t)k].second];
  });
  parallel_chunks(n, workers, [&](int64_t k0, int64_t k1) { std::copy(tmp + k0, tmp + k1, r + k0); });
}

// Preorder numbering is fixed by the subtree sizes alone (aabb_node_count),
// and leaves take the face ids in the order the recursion leaves them, so
// every subtree knows its node index and its slice of `order` up front and
// large subtrees are built on their own threads with identical results: a
// node with a budget of w threads sorts with them and hands w / 2 to each
// child. Boxes are filled afterwards, bottom-up (box_pass): a node's box is
// the union of its children's, the same min / max as over its faces.
struct Builder {
  FaceKey* keys;     // reordered in place; its ids become `order`
  FaceKey* scratch;  // as long as keys
  int64_t leaf;
  AabbTree& t;
  std::atomic<int64_t> deepest{0};

  // parent_axis: the axis the slice is already stably sorted by (-1: none)
  void node(int64_t me, int64_t b, int64_t n, int64_t depth, int workers, int parent_axis) {
    int64_t seen = deepest.load(std::memory_order_relaxed);
    while (depth > seen && !deepest.compare_exchange_weak(seen, depth)) {
    }
    if (n <= leaf) {
      t.nodes[4 * me] = -1;
      t.nodes[4 * me + 1] = -1;
      t.nodes[4 * me + 2] = b;
      t.nodes[4 * me + 3] = n;
      return;
    }
    FaceKey* r = keys + b;
    // split axis: largest spread of the face-box centres, first on ties
    double mn[3] = {r[0].mid[0], r[0].mid[1], r[0].mid[2]};
    double mx[3] = {mn[0], mn[1], mn[2]};
    for (int64_t k = 1; k < n; ++k)
      for (int d = 0; d < 3; ++d) {
        mn[d] = std::min(mn[d], r[k].mid[d]);
        mx[d] = std::max(mx[d], r[k].mid[d]);
      }
    int axis = 0;
    double best = -1.0;
    for (int d = 0; d < 3; ++d)
      if (d == 0 || mx[d] - mn[d] > best) {
        best = mx[d] - mn[d];
        axis = d;
      }
    // a slice of the parent's sorted order is sorted by the parent's key with
    // ties in their order: a stable sort by the same key leaves it unchanged
    if (axis != parent_axis) sort_by_axis(r, scratch + b, n, axis, workers);
    const int64_t half = n / 2;
    const int64_t l = me + 1;
    const int64_t rr = me + 1 + aabb_node_count(half, leaf);
    t.nodes[4 * me] = l;
    t.nodes[4 * me + 1] = rr;
    t.nodes[4 * me + 2] = 0;
    t.nodes[4 * me + 3] = 0;
    if (workers > 1 && n > 65536) {
      const int wr = workers / 2;
      std::thread right([&, rr, b, half, n, depth, wr, axis] {
        node(rr, b + half, n - half, depth + 1, wr, axis);
      });
      node(l, b, half, depth + 1, workers - wr, axis);
      right.join();
    } else {
      node(l, b, half, depth + 1, 1, axis);
      node(rr, b + half, n - half, depth + 1, 1, axis);
    }
  }
};

// Leaf boxes from their faces' bounds (leaves in parallel), then every
// internal node (children follow their parent in preorder) as the union of
// its children.
void box_pass(const std::vector<double>& lo, const std::vector<double>& hi, AabbTree& t,
              int workers) {
  const int64_t count = (int64_t)t.nodes.size() / 4;
  parallel_chunks(count, workers, [&](int64_t n0, int64_t n1) {
    for (int64_t me = n0; me < n1; ++me) {
      const int64_t* d = t.nodes.data() + 4 * me;
      if (d[3] <= 0) continue;
      double* box = t.boxes.data() + 6 * me;
      const int64_t* id = t.order.data() + d[2];
      for (int c = 0; c < 3; ++c) {
        box[c] = lo[3 * id[0] + c];
        box[3 + c] = hi[3 * id[0] + c];
      }
      for (int64_t k = 1; k < d[3]; ++k)
        for (int c = 0; c < 3; ++c) {
          box[c] = std::min(box[c], lo[3 * id[k] + c]);
          box[3 + c] = std::max(box[3 + c], hi[3 * id[k] + c]);
        }
    }
  });
  for (int64_t me = count - 1; me >= 0; --me) {
    const int64_t* d = t.nodes.data() + 4 * me;
    if (d[3] > 0) continue;
    double* box = t.boxes.data() + 6 * me;
    const double* bl = t.boxes.data() + 6 * d[0];
    const double* br = t.boxes.data() + 6 * d[1];
    for (int c = 0; c < 3; ++c) {
      box[c] = std::min(bl[c], br[c]);
      box[3 + c] = std::max(bl[3 + c], br[3 + c]);
    }
  }
}

template <typename FT>
void face_bounds(const double* V, const FT* F, int64_t nf, std::vector<double>& lo,
                 std::vector<double>& hi, std::vector<double>& mid, int workers) {
  lo.resize(3 * nf);
  hi.resize(3 * nf);
  mid.resize(3 * nf);
  parallel_chunks(nf, workers, [&](int64_t f0, int64_t f1) {
    for (int64_t f = f0; f < f1; ++f)
      for (int d = 0; d < 3; ++d) {
        const double a = V[3 * (int64_t)F[3 * f] + d], b = V[3 * (int64_t)F[3 * f + 1] + d],
                     c = V[3 * (int64_t)F[3 * f + 2] + d];
        lo[3 * f + d] = std::min(std::min(a, b), c);
        hi[3 * f + d] = std::max(std::max(a, b), c);
        mid[3 * f + d] = 0.5 * (lo[3 * f + d] + hi[3 * f + d]);
      }
  });
}

}  // namespace

int64_t aabb_node_count(int64_t nf, int64_t leaf) {
  // leaves(n) = 1 for n <= leaf, else leaves(n / 2) + leaves(n - n / 2); at
  // each depth the sizes take at most two adjacent values, so the recursion
  // is walked level by level over (size, multiplicity) pairs (at most two
  // per level; four slots)
  if (nf <= 0) return 0;
  int64_t size[4] = {nf, 0, 0, 0}, times[4] = {1, 0, 0, 0};
  int slots = 1;
  int64_t leaves = 0;
  while (slots > 0) {
    int64_t nsize[4] = {0, 0, 0, 0}, ntimes[4] = {0, 0, 0, 0};
    int nslots = 0;
    auto add = [&](int64_t sz, int64_t t) {
      for (int i = 0; i < nslots; ++i)
        if (nsize[i] == sz) {
          ntimes[i] += t;
          return;
        }
      if (nslots == 4) return;  // unreachable: sizes at one depth differ by <= 1
      nsize[nslots] = sz;
      ntimes[nslots] = t;
      ++nslots;
    };
    for (int i = 0; i < slots; ++i) {
      if (size[i] <= leaf) {
        leaves += times[i];
      } else {
        add(size[i] / 2, times[i]);
        add(size[i] - size[i] / 2, times[i]);
      }
    }
    for (int i = 0; i < 4; ++i) {
      size[i] = nsize[i];
      times[i] = ntimes[i];
    }
    slots = nslots;
  }
  return 2 * leaves - 1;
}

void aabb_build(const double* V, const void* F, int fdtype, int64_t nf, int64_t leaf,
                AabbTree& out) {
  out = AabbTree{};
  if (nf <= 0) return;
  // host threads: up to 16
  const int workers = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<double> lo, hi, mid;
  if (fdtype == 0)
    face_bounds(V, (const int64_t*)F, nf, lo, hi, mid, workers);
  else
    face_bounds(V, (const int32_t*)F, nf, lo, hi, mid, workers);
  const int64_t count = aabb_node_count(nf, leaf);
  out.boxes.assign(6 * count, 0.0);
  out.nodes.assign(4 * count, 0);
  std::vector<FaceKey> keys((size_t)nf), scratch((size_t)nf);
  parallel_chunks(nf, workers, [&](int64_t f0, int64_t f1) {
    for (int64_t f = f0; f < f1; ++f)
      keys[(size_t)f] = {{mid[3 * f], mid[3 * f + 1], mid[3 * f + 2]}, f};
  });
  std::vector<double>().swap(mid);
  Builder b{keys.data(), scratch.data(), leaf, out};
  b.node(0, 0, nf, 1, workers, -1);
  out.stack_size = b.deepest.load() + 1;
  std::vector<FaceKey>().swap(scratch);
  out.order.resize(nf);
  parallel_chunks(nf, workers, [&](int64_t k0, int64_t k1) {
    for (int64_t k = k0; k < k1; ++k) out.order[(size_t)k] = keys[(size_t)k].id;
  });
  box_pass(lo, hi, out, workers);
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
