// Triangle-strip records for the strip forward kernel (host side).
//
// No counterpart in the reference (src/capi.mojo:205-294 evaluates every face
// from its three corners). A strip w0, w1, w2, w3, ... lists faces
// (w0,w1,w2), (w1,w2,w3), ...: each face after the first adds one vertex, so
// a kernel walking the strip evaluates per-vertex quantities (offset from the
// query, its length) once per vertex instead of three times per face.
//
// Records: one per strip vertex, (x, y, z, flag), flag = 0 for the first two
// vertices of a strip segment (no face ends there), +1 / -1 when the record
// closes face (w_{k-2}, w_{k-1}, w_k) whose own corner order is a cyclic
// rotation of that order (+1) or of its reverse (-1).
// Tiles: records are grouped in tiles of kStripTile = 255 (a multiple of 3,
// for the kernel's 3-way register rotation); every tile starts a new strip
// segment, so tiles are independent. A strip crossing a tile boundary is cut
// and continued in the next tile from its last two vertices; the tail of a
// tile that cannot hold another face is padded with flag-0 records.
//
// Stripification: greedy. Faces are adjacent across an edge shared by exactly
// two faces (edges of more than two faces are treated as borders); the edge
// matching is a radix sort of packed edge keys. From each
// unvisited face (in index order) the walk tries the three rotations of its
// corners and keeps the longest strip of unvisited faces, always crossing the
// edge (w_{k-1}, w_k). The key build and the matching of large meshes run
// on several threads (host_threads) with the serial result.

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <thread>
#include <unordered_map>
#include <vector>

#include "wn_radix.h"

namespace wn {

constexpr int64_t kStripTile = 255;

namespace {

template <typename FT>
int64_t corner(const FT* F, int64_t f, int c) {
  return (int64_t)F[3 * f + c];
}

// Host threads for a pass over nf faces: one below 131,072, else up to 16.
int host_threads(int64_t nf) {
  if (nf < 131072) return 1;
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  return (int)std::min<unsigned>(16u, hw);
}

// fn(t) for t in [0, T), on T - 1 new threads and the caller.
template <typename Fn>
void run_threads(int T, Fn&& fn) {
  std::vector<std::thread> pool;
  for (int t = 1; t < T; ++t) pool.emplace_back(fn, t);
  fn(0);
  for (auto& th : pool) th.join();
}

template <typename FT>
struct Stripifier {
  const FT* F;
  int64_t nf;
  std::vector<int64_t> nbr;  // 3 per face: face across local edge e, -1 if none
  std::vector<uint8_t> used;

  // Vertex ids are below 2^31 (checked by every caller), so an edge is the
  // key (lo << 32 | hi); the payload is (face << 2 | local edge). Entries are
  // generated in face order and the radix sort is stable, so each edge's
  // faces stay in ascending order, as a sort by (lo, hi, face) leaves them.
  // Large meshes use several threads (host_threads) for the key build and
  // the matching: keys are written in face order (per-thread face ranges at
  // counted offsets) and the sorted runs are split at run starts, so the
  // result does not depend on the thread count.
  Stripifier(const FT* F_, int64_t nf_) : F(F_), nf(nf_), nbr(3 * nf_, -1), used(nf_, 0) {
    const int T = host_threads(nf);
    std::vector<int64_t> fcut(T + 1);
    for (int t = 0; t <= T; ++t) fcut[t] = nf * t / T;
    std::vector<size_t> at(T + 1, 0);
    run_threads(T, [&](int t) {
      size_t c = 0;
      for (int64_t f = fcut[t]; f < fcut[t + 1]; ++f)
        for (int k = 0; k < 3; ++k) c += corner(F, f, k) != corner(F, f, (k + 1) % 3);
      at[t + 1] = c;
    });
    for (int t = 0; t < T; ++t) at[t + 1] += at[t];
    std::vector<uint64_t> key(at[T]), pay(at[T]);
    run_threads(T, [&](int t) {
      size_t i = at[t];
      for (int64_t f = fcut[t]; f < fcut[t + 1]; ++f)
        for (int k = 0; k < 3; ++k) {
          const int64_t a = corner(F, f, k), b = corner(F, f, (k + 1) % 3);
          if (a == b) continue;
          key[i] = (uint64_t)std::min(a, b) << 32 | (uint64_t)std::max(a, b);
          pay[i] = (uint64_t)f << 2 | (uint64_t)k;
          ++i;
        }
    });
    radix_sort(key, &pay);
    // runs of equal keys, split between threads at run starts; an edge
    // shared by exactly two faces links them (each (face, edge) slot is
    // written by the one thread holding its run)
    const size_t n = key.size();
    std::vector<size_t> kcut(T + 1);
    for (int t = 0; t <= T; ++t) {
      size_t c = n * (size_t)t / (size_t)T;
      while (c > 0 && c < n && key[c] == key[c - 1]) ++c;
      kcut[t] = c;
    }
    run_threads(T, [&](int t) {
      for (size_t i = kcut[t]; i < kcut[t + 1];) {
        size_t j = i;
        while (j < n && key[j] == key[i]) ++j;
        if (j - i == 2) {
          const int64_t f0 = (int64_t)(pay[i] >> 2), f1 = (int64_t)(pay[i + 1] >> 2);
          nbr[3 * f0 + (int64_t)(pay[i] & 3)] = f1;
          nbr[3 * f1 + (int64_t)(pay[i + 1] & 3)] = f0;
        }
        i = j;
      }
    });
  }

  // Local edge index of edge {a, b} in face f, -1 if absent.
  int local_edge(int64_t f, int64_t a, int64_t b) const {
    for (int k = 0; k < 3; ++k) {
      const int64_t u = corner(F, f, k), v = corner(F, f, (k + 1) % 3);
      if ((u == a && v == b) || (u == b && v == a)) return k;
    }
    return -1;
  }

  int64_t third(int64_t f, int64_t a, int64_t b) const {
    for (int k = 0; k < 3; ++k) {
      const int64_t c = corner(F, f, k);
      if (c != a && c != b) return c;
    }
    return corner(F, f, 0);
  }

  // +1 if (p, q, r) is a cyclic rotation of face f's corner order, else -1.
  int orientation(int64_t f, int64_t p, int64_t q) const {
    for (int k = 0; k < 3; ++k)
      if (corner(F, f, k) == p) return corner(F, f, (k + 1) % 3) == q ? 1 : -1;
    return 1;
  }
};

// Strip walks inside the face range [lo, hi) (faces outside it are
// borders).
template <typename FT>
struct Walker {
  Stripifier<FT>& s;
  int64_t lo, hi;
  std::vector<int64_t> touched;  // faces a trial walk marked, to unmark

  // Walk from face f with strip order (p, q, r); returns faces and vertices.
  // mark = false leaves the used flags as they were.
  void walk(int64_t f, int64_t p, int64_t q, int64_t r, bool mark, std::vector<int64_t>& faces,
            std::vector<int64_t>& verts) {
    faces.assign(1, f);
    verts.assign({p, q, r});
    touched.clear();
    if (!mark) touched.push_back(f);
    s.used[f] = 1;
    int64_t cur = f;
    for (;;) {
      const int64_t a = verts[verts.size() - 2], b = verts.back();
      const int le = s.local_edge(cur, a, b);
      if (le < 0) break;
      const int64_t g = s.nbr[3 * cur + le];
      if (g < lo || g >= hi || s.used[g]) break;
      const int64_t t = s.third(g, a, b);
      if (t == a || t == b) break;
      s.used[g] = 1;
      if (!mark) touched.push_back(g);
      faces.push_back(g);
      verts.push_back(t);
      cur = g;
    }
    if (!mark)
      for (int64_t u : touched) s.used[u] = 0;
  }
};

// Appends (vertex id, flag) records, tracking the position in the tile.
struct Emitter {
  std::vector<int32_t>* idx;
  std::vector<int8_t>* flag;
  int64_t in_tile = 0;

  void record(int64_t v, int f) {
    idx->push_back((int32_t)v);
    flag->push_back((int8_t)f);
    if (++in_tile == kStripTile) in_tile = 0;
  }
  void pad_tile(int64_t v) {
    while (in_tile != 0) record(v, 0);
  }
};

// Strips of the faces in [lo, hi), appended to idx / flag from a tile start
// and padded to a tile end.
template <typename FT>
void build_range(Stripifier<FT>& s, int64_t lo, int64_t hi, std::vector<int32_t>& idx,
                 std::vector<int8_t>& flag) {
  const FT* F = s.F;
  Walker<FT> wk{s, lo, hi, {}};
  Emitter em{&idx, &flag};
  std::vector<int64_t> faces, verts, best_faces, best_verts;
  int64_t last_v = corner(F, lo, 0);
  for (int64_t f = lo; f < hi; ++f) {
    if (s.used[f]) continue;
    // try the three rotations (no marking), keep the longest, mark its faces
    best_faces.clear();
    for (int rot = 0; rot < 3; ++rot) {
      wk.walk(f, corner(F, f, rot), corner(F, f, (rot + 1) % 3), corner(F, f, (rot + 2) % 3),
              false, faces, verts);
      if (faces.size() > best_faces.size()) {
        best_faces.swap(faces);
        best_verts.swap(verts);
      }
    }
    for (int64_t g : best_faces) s.used[g] = 1;
    // emit in segments that fit the current tile
    size_t i = 0;  // index of the first face of the segment
    while (i < best_faces.size()) {
      int64_t space = kStripTile - em.in_tile;
      if (space < 3) {
        em.pad_tile(last_v);
        space = kStripTile;
      }
      const size_t m = std::min<size_t>(best_faces.size() - i, (size_t)(space - 2));
      em.record(best_verts[i], 0);
      em.record(best_verts[i + 1], 0);
      for (size_t k = 0; k < m; ++k) {
        const int64_t face = best_faces[i + k];
        em.record(best_verts[i + k + 2],
                  s.orientation(face, best_verts[i + k], best_verts[i + k + 1]));
      }
      last_v = best_verts[i + m + 1];
      i += m;
    }
  }
  em.pad_tile(last_v);
}

template <typename FT>
int64_t build_topology(const FT* F, int64_t nf, std::vector<int32_t>& idx,
                       std::vector<int8_t>& flag) {
  Stripifier<FT> s(F, nf);
  build_range(s, 0, nf, idx, flag);
  return (int64_t)idx.size();
}

}  // namespace

// Strip topology of mesh F (nf x 3 of fdtype 0 = int64, 1 = int32; indices
// in [0, nv) with nv < 2^31): per record the vertex id and the flag (0, +1,
// -1). Returns the record count (a multiple of kStripTile; 0 for no faces).
int64_t strip_topology(const void* F, int64_t nf, int fdtype, std::vector<int32_t>& idx,
                       std::vector<int8_t>& flag) {
  idx.clear();
  flag.clear();
  if (nf <= 0) return 0;
  idx.reserve((size_t)(nf + nf / 4 + kStripTile));
  flag.reserve((size_t)(nf + nf / 4 + kStripTile));
  return fdtype == 0 ? build_topology((const int64_t*)F, nf, idx, flag)
                     : build_topology((const int32_t*)F, nf, idx, flag);
}

// Records (x, y, z, flag) of mesh (V: nv x 3 float64, F as above): the
// topology's vertex ids gathered from V. Returns the record count.
int64_t strip_records(const double* V, const void* F, int64_t nf, int fdtype,
                      std::vector<double>& out) {
  std::vector<int32_t> idx;
  std::vector<int8_t> flag;
  const int64_t n = strip_topology(F, nf, fdtype, idx, flag);
  out.resize((size_t)(4 * n));
  for (int64_t k = 0; k < n; ++k) {
    const int64_t v = idx[(size_t)k];
    out[(size_t)(4 * k)] = V[3 * v];
    out[(size_t)(4 * k + 1)] = V[3 * v + 1];
    out[(size_t)(4 * k + 2)] = V[3 * v + 2];
    out[(size_t)(4 * k + 3)] = (double)flag[(size_t)k];
  }
  return n;
}

// Vertices of a face SoA (9 x nf, rows ax..cz) by exact coordinate equality
// (-0 and +0 count as equal): V gets each distinct corner once, in order of
// first appearance (face by face, corners a, b, c), F the corner ids. The
// corners that the kernels round identically are exactly these.
int64_t soa_vertices(const double* tri, int64_t nf, std::vector<double>& V,
                     std::vector<int32_t>& F) {
  struct Key {
    uint64_t x, y, z;
    bool operator==(const Key& o) const { return x == o.x && y == o.y && z == o.z; }
  };
  struct Hash {
    size_t operator()(const Key& k) const {
      uint64_t h = k.x * 0x9e3779b97f4a7c15ull;
      h ^= (k.y + 0x632be59bd9b4e019ull + (h << 6) + (h >> 2));
      h *= 0xbf58476d1ce4e5b9ull;
      h ^= (k.z + 0x94d049bb133111ebull + (h << 6) + (h >> 2));
      return (size_t)(h ^ (h >> 31));
    }
  };
  auto bits = [](double v) {
    v += 0.0;  // -0 -> +0
    uint64_t b;
    std::memcpy(&b, &v, sizeof b);
    return b;
  };
  std::unordered_map<Key, int32_t, Hash> ids;
  ids.reserve((size_t)(nf * 3 / 2 + 16));
  V.clear();
  F.assign((size_t)(3 * nf), 0);
  for (int64_t f = 0; f < nf; ++f)
    for (int c = 0; c < 3; ++c) {
      const double x = tri[(3 * c) * nf + f], y = tri[(3 * c + 1) * nf + f],
                   z = tri[(3 * c + 2) * nf + f];
      const Key k{bits(x), bits(y), bits(z)};
      auto it = ids.find(k);
      int32_t id;
      if (it == ids.end()) {
        id = (int32_t)(V.size() / 3);
        ids.emplace(k, id);
        V.push_back(x + 0.0);
        V.push_back(y + 0.0);
        V.push_back(z + 0.0);
      } else {
        id = it->second;
      }
      F[(size_t)(3 * f + c)] = id;
    }
  return (int64_t)(V.size() / 3);
}

}  // namespace wn
