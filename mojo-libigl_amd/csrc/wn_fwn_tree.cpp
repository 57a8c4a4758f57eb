// Hierarchical winding number, host side: per-node expansions over the
// median-split tree, and the CPU query (formulas in wn_fwn.h).
//
// The tree is the distance tree of wn_aabb_tree.cpp (python/igl/__init__.py:
// 286-319): leaves are appended to `order` in preorder, so every node, leaf or
// not, owns a contiguous range of `order`, and its moments are plain sums over
// that range (O(n log n) in total).

#include "wn_fwn.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <thread>

#include "wn_aabb.h"
#include "wn_fwn_eval.h"

namespace wn {

namespace {

constexpr double kPi = 3.14159265358979323846264338327950288;

// [first, last) of `order` under node n, for every node of its subtree.
void node_ranges(const AabbTree& t, int64_t n, std::vector<int64_t>& lo, std::vector<int64_t>& hi) {
  const int64_t* d = t.nodes.data() + 4 * n;
  if (d[3] > 0) {
    lo[n] = d[2];
    hi[n] = d[2] + d[3];
    return;
  }
  node_ranges(t, d[0], lo, hi);
  node_ranges(t, d[1], lo, hi);
  lo[n] = lo[d[0]];
  hi[n] = hi[d[1]];
}

void node_expansion(const double* corners, int64_t first, int64_t last, FwnNode& e) {
  // centre: area-weighted mean of the face centroids (plain mean if the
  // faces have no area)
  double area = 0.0, sx = 0.0, sy = 0.0, sz = 0.0, mx = 0.0, my = 0.0, mz = 0.0;
  for (int64_t k = first; k < last; ++k) {
    const double* t = corners + 9 * k;
    const double ux = t[3] - t[0], uy = t[4] - t[1], uz = t[5] - t[2];
    const double vx = t[6] - t[0], vy = t[7] - t[1], vz = t[8] - t[2];
    const double nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
    const double a = 0.5 * std::sqrt(nx * nx + ny * ny + nz * nz);
    const double gx = (t[0] + t[3] + t[6]) / 3.0, gy = (t[1] + t[4] + t[7]) / 3.0,
                 gz = (t[2] + t[5] + t[8]) / 3.0;
    area += a;
    sx += a * gx; sy += a * gy; sz += a * gz;
    mx += gx; my += gy; mz += gz;
  }
  const double cnt = (double)(last - first);
  if (area > 0.0) {
    e.cx = sx / area; e.cy = sy / area; e.cz = sz / area;
  } else {
    e.cx = mx / cnt; e.cy = my / cnt; e.cz = mz / cnt;
  }
  double radius2 = 0.0;
  for (int i = 0; i < 3; ++i) e.n[i] = 0.0;
  for (int i = 0; i < 9; ++i) e.c[i] = 0.0;
  for (int i = 0; i < 18; ++i) e.d[i] = 0.0;
  for (int64_t k = first; k < last; ++k) {
    const double* t = corners + 9 * k;
    double y[3][3];
    for (int v = 0; v < 3; ++v) {
      y[v][0] = t[3 * v] - e.cx;
      y[v][1] = t[3 * v + 1] - e.cy;
      y[v][2] = t[3 * v + 2] - e.cz;
      radius2 = std::max(radius2, y[v][0] * y[v][0] + y[v][1] * y[v][1] + y[v][2] * y[v][2]);
    }
    // vector area A n = (B - A) x (C - A) / 2
    const double ux = t[3] - t[0], uy = t[4] - t[1], uz = t[5] - t[2];
    const double vx = t[6] - t[0], vy = t[7] - t[1], vz = t[8] - t[2];
    const double an[3] = {0.5 * (uy * vz - uz * vy), 0.5 * (uz * vx - ux * vz),
                          0.5 * (ux * vy - uy * vx)};
    const double s[3] = {y[0][0] + y[1][0] + y[2][0], y[0][1] + y[1][1] + y[2][1],
                         y[0][2] + y[1][2] + y[2][2]};
    // int_t y y^T dA = (A / 12) (sum_v y_v y_v^T + s s^T)
    double m[6];
    const int jj[6] = {0, 0, 0, 1, 1, 2}, kk[6] = {0, 1, 2, 1, 2, 2};
    for (int p = 0; p < 6; ++p) {
      const int j = jj[p], l = kk[p];
      m[p] = (y[0][j] * y[0][l] + y[1][j] * y[1][l] + y[2][j] * y[2][l] + s[j] * s[l]) / 12.0;
    }
    for (int i = 0; i < 3; ++i) {
      e.n[i] += an[i];
      for (int j = 0; j < 3; ++j) e.c[3 * i + j] += an[i] * s[j] / 3.0;
      for (int p = 0; p < 6; ++p) e.d[6 * i + p] += an[i] * m[p];
    }
  }
  e.radius = std::sqrt(radius2);
}

inline double exact_omega(const double* t, double qx, double qy, double qz) {
  const double ax = t[0] - qx, ay = t[1] - qy, az = t[2] - qz;
  const double bx = t[3] - qx, by = t[4] - qy, bz = t[5] - qz;
  const double cx = t[6] - qx, cy = t[7] - qy, cz = t[8] - qz;
  const double la = std::sqrt(ax * ax + ay * ay + az * az);
  const double lb = std::sqrt(bx * bx + by * by + bz * bz);
  const double lc = std::sqrt(cx * cx + cy * cy + cz * cz);
  const double det =
      ax * (by * cz - bz * cy) - ay * (bx * cz - bz * cx) + az * (bx * cy - by * cx);
  double den = la * lb * lc;
  den += (ax * bx + ay * by + az * bz) * lc;
  den += (bx * cx + by * cy + bz * cz) * la;
  den += (cx * ax + cy * ay + cz * az) * lb;
  return 2.0 * std::atan2(det, den);
}

}  // namespace

void fwn_build(const double* V, const void* F, int fdtype, int64_t nf, int64_t leaf,
               FwnTree& out) {
  out = FwnTree{};
  if (nf <= 0) return;
  AabbTree t;
  aabb_build(V, F, fdtype, nf, leaf, t);
  const int64_t nn = (int64_t)t.nodes.size() / 4;
  out.order = t.order;
  out.stack_size = t.stack_size;
  out.corners.resize(9 * (size_t)nf);
  // faces and nodes are independent: up to 16 threads
  const int64_t want = std::min<int64_t>(16, std::max<int64_t>(1, nf / 32768));
  const int workers = (int)std::min<int64_t>(
      want, std::max<unsigned>(1u, std::thread::hardware_concurrency()));
  auto gather = [&](int w) {
    for (int64_t k = nf * w / workers; k < nf * (w + 1) / workers; ++k) {
      const int64_t f = t.order[(size_t)k];
      for (int c = 0; c < 3; ++c) {
        const int64_t v = fdtype == 0 ? ((const int64_t*)F)[3 * f + c]
                                      : (int64_t)((const int32_t*)F)[3 * f + c];
        for (int d = 0; d < 3; ++d) out.corners[(size_t)(9 * k + 3 * c + d)] = V[3 * v + d];
      }
    }
  };
  {
    std::vector<std::thread> pool;
    for (int w = 1; w < workers; ++w) pool.emplace_back(gather, w);
    gather(0);
    for (auto& th : pool) th.join();
  }
  std::vector<int64_t> lo((size_t)nn), hi((size_t)nn);
  node_ranges(t, 0, lo, hi);
  out.nodes.resize((size_t)nn);
  // nodes, largest first (breadth-first order), handed out by an atomic
  // counter: one at a time for the top levels, where one node can hold a
  // large share of the work, in chunks of 256 below (each node's expansion
  // is computed alike whichever thread takes it)
  std::vector<int64_t> bfs;
  bfs.reserve((size_t)nn);
  bfs.push_back(0);
  for (size_t i = 0; i < bfs.size(); ++i) {
    const int64_t* d = t.nodes.data() + 4 * bfs[i];
    if (d[3] == 0) {
      bfs.push_back(d[0]);
      bfs.push_back(d[1]);
    }
  }
  constexpr int64_t kSingles = 4096, kChunk = 256;
  std::atomic<int64_t> next{0};
  auto run = [&](int) {
    for (;;) {
      const int64_t at = next.load(std::memory_order_relaxed);
      const int64_t take = at < kSingles ? 1 : kChunk;
      const int64_t b = next.fetch_add(take);
      if (b >= nn) break;
      for (int64_t i = b; i < std::min(nn, b + take); ++i) {
        const int64_t n = bfs[(size_t)i];
        FwnNode& e = out.nodes[(size_t)n];
        node_expansion(out.corners.data(), lo[n], hi[n], e);
        const int64_t* d = t.nodes.data() + 4 * n;
        e.left = d[0];
        e.right = d[1];
        e.start = d[2];
        e.count = d[3];
      }
    }
  };
  std::vector<std::thread> pool;
  for (int w = 1; w < workers; ++w) pool.emplace_back(run, w);
  run(0);
  for (auto& th : pool) th.join();
}

void fwn_query_cpu(const FwnTree& t, const double* q, double* w, double beta, int64_t q0,
                   int64_t q1) {
  const int64_t nn = (int64_t)t.nodes.size();
  std::vector<int64_t> stack((size_t)std::max<int64_t>(t.stack_size, 1));
  const double b2 = beta * beta;
  for (int64_t i = q0; i < q1; ++i) {
    const double qx = q[3 * i], qy = q[3 * i + 1], qz = q[3 * i + 2];
    double omega = 0.0;
    int64_t top = 0;
    if (nn > 0) stack[top++] = 0;
    while (top > 0) {
      const FwnNode& e = t.nodes[(size_t)stack[--top]];
      const double dx = qx - e.cx, dy = qy - e.cy, dz = qz - e.cz;
      if (dx * dx + dy * dy + dz * dz > b2 * e.radius * e.radius) {
        omega += fwn_far_omega(&e.cx, qx, qy, qz);
      } else if (e.count > 0) {
        for (int64_t k = e.start; k < e.start + e.count; ++k)
          omega += exact_omega(t.corners.data() + 9 * k, qx, qy, qz);
      } else {
        stack[top++] = e.right;
        stack[top++] = e.left;
      }
    }
    w[i] = omega / (4.0 * kPi);
  }
}

}  // namespace wn
