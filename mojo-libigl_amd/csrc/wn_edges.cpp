// Net directed boundary edges of a triangle mesh (host). No counterpart in
// the reference: this is the topology the edge adjoint sums over.
//
// Each face (i, j, k) contributes the directed edges i->j, j->k, k->i. For an
// undirected edge {u, v} (u < v) the net count is #(u->v) - #(v->u); edges
// shared by two consistently oriented faces cancel, so a closed, consistently
// oriented mesh has none. An edge with net count n != 0 is emitted once,
// directed u->v with weight n when n > 0 and v->u with weight -n otherwise;
// self-loops (a repeated vertex in a face) are dropped. Order: by (min id,
// max id).

#include <algorithm>
#include <cstdint>
#include <vector>

#include "wn_radix.h"

namespace wn {

int64_t boundary_edges(const void* F, int64_t nf, int fdtype, std::vector<int32_t>& edges,
                       std::vector<double>& weight) {
  edges.clear();
  weight.clear();
  // key = (min << 32 | max) << 1 | (1 when the edge runs max -> min)
  std::vector<uint64_t> keys;
  keys.reserve((size_t)(3 * nf));
  auto corner = [&](int64_t f, int c) -> uint32_t {
    return fdtype == 0 ? (uint32_t)((const int64_t*)F)[3 * f + c]
                       : (uint32_t)((const int32_t*)F)[3 * f + c];
  };
  for (int64_t f = 0; f < nf; ++f) {
    const uint32_t v[3] = {corner(f, 0), corner(f, 1), corner(f, 2)};
    for (int c = 0; c < 3; ++c) {
      const uint32_t s = v[c], t = v[(c + 1) % 3];
      if (s == t) continue;
      const uint64_t lo = s < t ? s : t, hi = s < t ? t : s;
      keys.push_back(((lo << 32 | hi) << 1) | (s > t ? 1u : 0u));
    }
  }
  radix_sort(keys, nullptr);
  size_t i = 0;
  while (i < keys.size()) {
    const uint64_t e = keys[i] >> 1;
    int64_t net = 0;
    for (; i < keys.size() && keys[i] >> 1 == e; ++i) net += (keys[i] & 1) ? -1 : 1;
    if (net == 0) continue;
    const int32_t lo = (int32_t)(e >> 32), hi = (int32_t)(e & 0xffffffffu);
    edges.push_back(net > 0 ? lo : hi);
    edges.push_back(net > 0 ? hi : lo);
    weight.push_back((double)(net > 0 ? net : -net));
  }
  return (int64_t)weight.size();
}

int64_t strip_boundary(const void* F, int64_t nf, int fdtype, int64_t nv, const int32_t* idx,
                       int64_t nrec, std::vector<int32_t>& pairs, std::vector<double>& weight) {
  std::vector<int32_t> edges;
  const int64_t n = boundary_edges(F, nf, fdtype, edges, weight);
  pairs.assign(edges.size(), -1);
  if (n == 0) return 0;
  // first record of each vertex (every face corner is some record's vertex)
  std::vector<int32_t> first((size_t)nv, -1);
  for (int64_t k = nrec - 1; k >= 0; --k) {
    const int32_t v = idx[k];
    if (v >= 0 && v < nv) first[(size_t)v] = (int32_t)k;
  }
  for (size_t i = 0; i < edges.size(); ++i) pairs[i] = first[(size_t)edges[i]];
  return n;
}

}  // namespace wn
