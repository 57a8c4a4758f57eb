// Point-mesh squared distance on the GPU: the reference's tree descent
// (src/capi.mojo:523-587), one query per lane, FP64.
//
// Every lane runs exactly the reference's depth-first walk -- same tree
// (wn_aabb_tree.cpp builds it as python/igl/__init__.py:286-319 does), same pop
// order (nearer child on top), same pruning ('>' on nodes, '<=' on pushes,
// strict '<' on faces), same closest-point arithmetic (wn_closest.h, FMA
// contraction off) -- so distance, face index and closest point equal the
// CPU path's bit for bit, ties included.
//
// Layout in HBM (built once per mesh by aabb_upload):
//   AabbNodeDev[nnodes]   64 B per node: box lo/hi, left, right, start, count
//   double[9 * nf]        face corners in leaf order (a leaf's faces are
//                         contiguous: one 72 B record each)
//   int64_t[nf]           original face id of each leaf-order face
// Blocks are one wave (64 lanes): walks diverge per lane, and small query
// counts (the reference's 10k-query benchmark) still spread over many CUs.
// The descent stack lives in LDS, [depth][lane] so lanes hit distinct banks.
// Optional Morton ordering of the queries keeps a wave's walks close.

#include "wn_aabb.h"
#include "wn_closest.h"
#include "wn_internal.h"

namespace wn {

struct alignas(64) AabbNodeDev {
  double lox, loy, loz, hix, hiy, hiz;
  int32_t left, right, start, count;
};

namespace {

constexpr int kAabbLanes = 64;

__global__ __launch_bounds__(kAabbLanes) void aabb_query_kernel(
    const AabbNodeDev* __restrict__ nodes, int64_t nnodes, const double* __restrict__ corners,
    const int64_t* __restrict__ face_id, const double* __restrict__ q,
    const int32_t* __restrict__ perm, double* __restrict__ sqrd, int64_t* __restrict__ index,
    double* __restrict__ closest, int64_t np) {
  __shared__ int32_t stack[kAabbStack][kAabbLanes];
  const int lane = threadIdx.x;
  const int64_t slot = (int64_t)blockIdx.x * kAabbLanes + lane;
  if (slot >= np) return;  // no barriers below
  const double px = q[3 * slot], py = q[3 * slot + 1], pz = q[3 * slot + 2];
  double best = 1.0e300, cx = 0.0, cy = 0.0, cz = 0.0;
  int64_t best_f = -1;
  int top = 0;
  stack[top++][lane] = 0;
  while (top > 0) {
    const int32_t n = stack[--top][lane];
    if (n < 0 || n >= nnodes) continue;
    const AabbNodeDev nd = nodes[n];
    if (box_distance2(px, py, pz, nd.lox, nd.loy, nd.loz, nd.hix, nd.hiy, nd.hiz) > best)
      continue;
    if (nd.count > 0) {
      for (int32_t k = nd.start; k < nd.start + nd.count; ++k) {
        const double* t = corners + 9 * (int64_t)k;
        const ClosestHit h =
            closest_on_triangle(px, py, pz, t[0], t[1], t[2], t[3], t[4], t[5], t[6], t[7], t[8]);
        if (h.d2 < best) {
          best = h.d2;
          best_f = k;  // leaf-order position, mapped to the face id below
          cx = h.x; cy = h.y; cz = h.z;
        }
      }
    } else {
      const AabbNodeDev l = nodes[nd.left], r = nodes[nd.right];
      const double dl = box_distance2(px, py, pz, l.lox, l.loy, l.loz, l.hix, l.hiy, l.hiz);
      const double dr = box_distance2(px, py, pz, r.lox, r.loy, r.loz, r.hix, r.hiy, r.hiz);
      // the nearer child goes on top of the stack
      if (dl < dr) {
        if (dr <= best) stack[top++][lane] = nd.right;
        if (dl <= best) stack[top++][lane] = nd.left;
      } else {
        if (dl <= best) stack[top++][lane] = nd.left;
        if (dr <= best) stack[top++][lane] = nd.right;
      }
    }
  }
  const int64_t out = perm ? (int64_t)perm[slot] : slot;
  sqrd[out] = best;
  index[out] = best_f >= 0 ? face_id[best_f] : -1;
  closest[3 * out] = cx;
  closest[3 * out + 1] = cy;
  closest[3 * out + 2] = cz;
}

// A mesh without faces: +inf, -1, origin.
__global__ __launch_bounds__(kBlockThreads) void aabb_empty_kernel(double* __restrict__ sqrd,
                                                                    int64_t* __restrict__ index,
                                                                    double* __restrict__ closest,
                                                                    int64_t np) {
  const int64_t i = (int64_t)blockIdx.x * kBlockThreads + threadIdx.x;
  if (i >= np) return;
  sqrd[i] = __builtin_inf();
  index[i] = -1;
  closest[3 * i] = closest[3 * i + 1] = closest[3 * i + 2] = 0.0;
}

}  // namespace

size_t aabb_device_bytes(int64_t nnodes, int64_t nf) {
  return align_up(sizeof(AabbNodeDev) * (size_t)nnodes) + align_up(sizeof(double) * 9 * (size_t)nf) +
         align_up(sizeof(int64_t) * (size_t)nf);
}

hipError_t aabb_upload(const AabbTree& t, const double* V, const void* F, int fdtype, void* dev,
                       hipStream_t stream) {
  const int64_t nnodes = (int64_t)t.nodes.size() / 4;
  const int64_t nf = (int64_t)t.order.size();
  std::vector<AabbNodeDev> nodes((size_t)nnodes);
  for (int64_t n = 0; n < nnodes; ++n) {
    const double* b = t.boxes.data() + 6 * n;
    const int64_t* d = t.nodes.data() + 4 * n;
    nodes[(size_t)n] = AabbNodeDev{b[0], b[1], b[2], b[3], b[4], b[5], (int32_t)d[0],
                                   (int32_t)d[1], (int32_t)d[2], (int32_t)d[3]};
  }
  std::vector<double> corners((size_t)(9 * nf));
  for (int64_t k = 0; k < nf; ++k) {
    const int64_t f = t.order[(size_t)k];
    for (int c = 0; c < 3; ++c) {
      const int64_t v = fdtype == 0 ? ((const int64_t*)F)[3 * f + c]
                                    : (int64_t)((const int32_t*)F)[3 * f + c];
      for (int d = 0; d < 3; ++d) corners[(size_t)(9 * k + 3 * c + d)] = V[3 * v + d];
    }
  }
  char* p = (char*)dev;
  hipError_t e = hipMemcpyAsync(p, nodes.data(), sizeof(AabbNodeDev) * nodes.size(),
                                hipMemcpyHostToDevice, stream);
  p += align_up(sizeof(AabbNodeDev) * (size_t)nnodes);
  if (e == hipSuccess)
    e = hipMemcpyAsync(p, corners.data(), sizeof(double) * corners.size(), hipMemcpyHostToDevice,
                       stream);
  p += align_up(sizeof(double) * 9 * (size_t)nf);
  if (e == hipSuccess)
    e = hipMemcpyAsync(p, t.order.data(), sizeof(int64_t) * (size_t)nf, hipMemcpyHostToDevice,
                       stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);  // host vectors go out of scope
  return e;
}

size_t aabb_workspace_bytes(int64_t np, int sort) {
  if (!sort) return 0;
  return align_up((size_t)np * sizeof(int32_t)) + align_up((size_t)np * 3 * sizeof(double)) +
         morton_workspace_bytes(np);
}

hipError_t aabb_query_launch(const void* dev, int64_t nnodes, int64_t nf, const double* q,
                             double* sqrd, int64_t* index, double* closest, int64_t np, int sort,
                             void* ws, size_t ws_bytes, hipStream_t stream) {
  if (np == 0) return hipSuccess;
  if (nnodes == 0) {
    aabb_empty_kernel<<<dim3((unsigned)((np + kBlockThreads - 1) / kBlockThreads)), kBlockThreads,
                        0, stream>>>(sqrd, index, closest, np);
    return hipGetLastError();
  }
  if (ws_bytes < aabb_workspace_bytes(np, sort)) return hipErrorInvalidValue;
  const char* p = (const char*)dev;
  const AabbNodeDev* nodes = (const AabbNodeDev*)p;
  p += align_up(sizeof(AabbNodeDev) * (size_t)nnodes);
  const double* corners = (const double*)p;
  p += align_up(sizeof(double) * 9 * (size_t)nf);
  const int64_t* face_id = (const int64_t*)p;
  int32_t* perm = nullptr;
  if (sort) {
    char* c = (char*)ws;
    perm = (int32_t*)c;
    c += align_up((size_t)np * sizeof(int32_t));
    double* qs = (double*)c;
    c += align_up((size_t)np * 3 * sizeof(double));
    hipError_t e = morton_order(q, np, perm, c, morton_workspace_bytes(np), stream);
    if (e == hipSuccess) e = gather_queries_launch(q, perm, np, qs, stream);
    if (e != hipSuccess) return e;
    q = qs;
  }
  aabb_query_kernel<<<dim3((unsigned)((np + kAabbLanes - 1) / kAabbLanes)), kAabbLanes, 0,
                      stream>>>(nodes, nnodes, corners, face_id, q, perm, sqrd, index, closest,
                                np);
  return hipGetLastError();
}

}  // namespace wn
