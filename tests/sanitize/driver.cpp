// Sanitizer driver for the host-side C++ of the library (the CPU winding
// sweep behind mli_winding, the stripifier, boundary edges, the bounding-box
// and fast-winding trees and their CPU queries). Built twice by the Makefile
// beside it, with g++ -fsanitize=address,undefined and with -fsanitize=thread,
// and run by tests/test_sanitizers.py. The threaded parts run the way the
// library's callers run them: mli_winding and the tree queries from several
// threads on disjoint query spans (the reference's fan-out, python/igl/
// _lib.py:72-95), the tree builders above their parallel thresholds. Every
// check is against a serial or brute-force result; exit status 0 = clean.

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <thread>
#include <vector>

#include "wn_aabb.h"
#include "wn_fwn.h"

extern "C" void mli_winding(const double* tri, const double* q, double* dst, int64_t nf,
                            int64_t nq, int64_t q0, int64_t q1);
extern "C" void mli_aabb_squared_distance(const double* p, const double* v, const int64_t* f,
                                          const double* boxes, const int64_t* nodes,
                                          const int64_t* order, double* sqrd, int64_t* indices,
                                          double* closest, int64_t* stack, int64_t np,
                                          int64_t nnodes, int64_t q0, int64_t q1);

// Host entry points of wn_strips.cpp and wn_edges.cpp (their declarations
// live beside the HIP launchers; a signature change fails to link here).
namespace wn {
int64_t strip_records(const double* V, const void* F, int64_t nf, int fdtype,
                      std::vector<double>& out);
int64_t strip_topology(const void* F, int64_t nf, int fdtype, std::vector<int32_t>& idx,
                       std::vector<int8_t>& flag);
int64_t soa_vertices(const double* tri, int64_t nf, std::vector<double>& V,
                     std::vector<int32_t>& F);
int64_t boundary_edges(const void* F, int64_t nf, int fdtype, std::vector<int32_t>& edges,
                       std::vector<double>& weight);
int64_t strip_boundary(const void* F, int64_t nf, int fdtype, int64_t nv, const int32_t* idx,
                       int64_t nrec, std::vector<int32_t>& pairs, std::vector<double>& weight);
}  // namespace wn

namespace {

int failures = 0;

void check(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "FAILED: %s\n", what);
    ++failures;
  }
}

struct Mesh {
  std::vector<double> V;
  std::vector<int64_t> F;
  int64_t nv() const { return (int64_t)V.size() / 3; }
  int64_t nf() const { return (int64_t)F.size() / 3; }
};

// Closed torus (R = 1, r = 0.35) on an nu x nv wrap-around grid, outward.
Mesh torus(int nu, int nv) {
  Mesh m;
  for (int i = 0; i < nu; ++i)
    for (int j = 0; j < nv; ++j) {
      const double u = 2 * M_PI * i / nu, v = 2 * M_PI * j / nv;
      m.V.insert(m.V.end(), {(1 + 0.35 * std::cos(v)) * std::cos(u),
                             (1 + 0.35 * std::cos(v)) * std::sin(u), 0.35 * std::sin(v)});
    }
  for (int i = 0; i < nu; ++i)
    for (int j = 0; j < nv; ++j) {
      const int64_t a = i * nv + j, b = ((i + 1) % nu) * nv + j;
      const int64_t c = ((i + 1) % nu) * nv + (j + 1) % nv, d = i * nv + (j + 1) % nv;
      m.F.insert(m.F.end(), {a, b, c, a, c, d});
    }
  return m;
}

// Open height field with holes, a repeated-vertex face and a duplicate face
// (the stripifier's and the edge counter's awkward cases).
Mesh open_grid(int n) {
  Mesh m;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i)
      m.V.insert(m.V.end(), {(double)i, (double)j, 0.1 * std::sin(i / 3.0) * std::cos(j / 5.0)});
  for (int j = 0; j + 1 < n; ++j)
    for (int i = 0; i + 1 < n; ++i) {
      if ((i * 7 + j * 3) % 17 == 0) continue;  // holes
      const int64_t a = j * n + i;
      m.F.insert(m.F.end(), {a, a + 1, a + n + 1, a, a + n + 1, a + n});
    }
  m.F.insert(m.F.end(), {0, 0, 1});                  // repeated vertex
  m.F.insert(m.F.end(), {m.F[3], m.F[4], m.F[5]});   // duplicate face
  return m;
}

std::vector<double> soa(const Mesh& m) {
  const int64_t nf = m.nf();
  std::vector<double> tri(9 * nf);
  for (int64_t f = 0; f < nf; ++f)
    for (int c = 0; c < 3; ++c)
      for (int k = 0; k < 3; ++k) tri[(3 * c + k) * nf + f] = m.V[3 * m.F[3 * f + c] + k];
  return tri;
}

std::vector<double> queries(int64_t n, double lo, double hi, unsigned seed) {
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> u(lo, hi);
  std::vector<double> q(3 * n);
  for (auto& x : q) x = u(rng);
  return q;
}

// fn(q0, q1) on `workers` threads over contiguous spans of [0, n).
template <typename Fn>
void fan_out(int64_t n, int workers, Fn fn) {
  std::vector<std::thread> pool;
  const int64_t step = (n + workers - 1) / workers;
  for (int64_t q0 = 0; q0 < n; q0 += step) pool.emplace_back(fn, q0, std::min(n, q0 + step));
  for (auto& t : pool) t.join();
}

void winding_sweep() {
  const Mesh m = torus(48, 40);
  const std::vector<double> tri = soa(m);
  const int64_t nf = m.nf(), nq = 301;
  std::vector<double> q = queries(nq, -1.6, 1.6, 7);
  q[0] = 1.0; q[1] = 0.0; q[2] = 0.0;  // inside the tube: w = 1
  std::vector<double> serial(nq), parallel(nq, -7.0);
  mli_winding(tri.data(), q.data(), serial.data(), nf, nq, 0, nq);
  fan_out(nq, 4, [&](int64_t a, int64_t b) {
    mli_winding(tri.data(), q.data(), parallel.data(), nf, nq, a, b);
  });
  bool same = true, integral = true;
  for (int64_t i = 0; i < nq; ++i) {
    same = same && serial[i] == parallel[i];
    integral = integral && std::fabs(serial[i] - std::round(serial[i])) < 1e-9;
  }
  check(same, "mli_winding: 4 threads == serial, bit for bit");
  check(integral, "mli_winding: closed torus gives integer w");
  check(std::fabs(serial[0] - 1.0) < 1e-12, "mli_winding: w = 1 inside the tube");
}

void strips_and_edges() {
  // the torus (358,400 faces) is above the stripifier's threaded threshold
  // (131,072 faces) and its 1,075,200 edge keys above the radix sort's (2^20)
  for (int which = 0; which < 2; ++which) {
    const Mesh m = which ? open_grid(60) : torus(640, 280);
    const int64_t nf = m.nf();
    std::vector<int32_t> F32(m.F.begin(), m.F.end());
    std::vector<double> rec64, rec32;
    const int64_t n64 = wn::strip_records(m.V.data(), m.F.data(), nf, 0, rec64);
    const int64_t n32 = wn::strip_records(m.V.data(), F32.data(), nf, 1, rec32);
    check(n64 > 0 && n64 == n32 && rec64 == rec32, "strip_records: int64 == int32 faces");
    check(n64 % 255 == 0, "strip_records: whole tiles");
    int64_t faces = 0;
    for (int64_t k = 0; k < n64; ++k) faces += rec64[4 * k + 3] != 0.0;
    check(faces == nf, "strip_records: every face closed once");
    std::vector<int32_t> idx;
    std::vector<int8_t> flag;
    const int64_t nt = wn::strip_topology(m.F.data(), nf, 0, idx, flag);
    check(nt == n64 && (int64_t)idx.size() == nt, "strip_topology: same record count");
    std::vector<int32_t> edges, pairs;
    std::vector<double> weight, pweight;
    const int64_t ne = wn::boundary_edges(m.F.data(), nf, 0, edges, weight);
    check(which ? ne > 0 : ne == 0, "boundary_edges: open grid has edges, torus none");
    const int64_t np = wn::strip_boundary(m.F.data(), nf, 0, m.nv(), idx.data(), nt, pairs,
                                          pweight);
    check(np == ne, "strip_boundary: one record pair per boundary edge");
    std::vector<double> V2;
    std::vector<int32_t> F2;
    const std::vector<double> tri = soa(m);
    const int64_t nv2 = wn::soa_vertices(tri.data(), nf, V2, F2);
    check(nv2 > 0 && nv2 <= m.nv(), "soa_vertices: distinct corners");
  }
}

void trees() {
  const Mesh m = torus(330, 110);  // 72,600 faces: above the parallel thresholds
  const int64_t nf = m.nf(), nq = 512;
  wn::AabbTree t;
  wn::aabb_build(m.V.data(), m.F.data(), 0, nf, 8, t);
  const int64_t nnodes = (int64_t)t.nodes.size() / 4;
  check(nnodes == wn::aabb_node_count(nf, 8), "aabb_build: node count");
  const std::vector<double> p = queries(nq, -1.6, 1.6, 11);
  std::vector<double> d2(nq), cl(3 * nq);
  std::vector<int64_t> id(nq);
  fan_out(nq, 4, [&](int64_t a, int64_t b) {
    std::vector<int64_t> stack(t.stack_size + 1);
    mli_aabb_squared_distance(p.data(), m.V.data(), m.F.data(), t.boxes.data(), t.nodes.data(),
                              t.order.data(), d2.data(), id.data(), cl.data(), stack.data(), nq,
                              nnodes, a, b);
  });
  // brute force over the vertices bounds the distance from above
  bool ok = true;
  for (int64_t i = 0; i < nq; i += 37) {
    double best = 1e300;
    for (int64_t v = 0; v < m.nv(); ++v) {
      const double dx = p[3 * i] - m.V[3 * v], dy = p[3 * i + 1] - m.V[3 * v + 1],
                   dz = p[3 * i + 2] - m.V[3 * v + 2];
      best = std::min(best, dx * dx + dy * dy + dz * dz);
    }
    ok = ok && d2[i] <= best && id[i] >= 0 && id[i] < nf;
  }
  check(ok, "mli_aabb_squared_distance: within the nearest vertex, valid face ids");

  wn::FwnTree ft;
  wn::fwn_build(m.V.data(), m.F.data(), 0, nf, 8, ft);
  std::vector<double> w(nq, -7.0), w1(nq);
  fan_out(nq, 4, [&](int64_t a, int64_t b) { wn::fwn_query_cpu(ft, p.data(), w.data(), 2.0, a, b); });
  wn::fwn_query_cpu(ft, p.data(), w1.data(), 2.0, 0, nq);
  bool same = true, near = true;
  for (int64_t i = 0; i < nq; ++i) {
    same = same && w[i] == w1[i];
    near = near && std::fabs(w[i] - std::round(w[i])) < 0.05;
  }
  check(same, "fwn_query_cpu: 4 threads == serial, bit for bit");
  check(near, "fwn_query_cpu: closed torus within 0.05 of an integer at beta 2");
}

}  // namespace

int main() {
  winding_sweep();
  strips_and_edges();
  trees();
  if (failures) return 1;
  std::printf("sanitizer driver: all checks passed\n");
  return 0;
}
