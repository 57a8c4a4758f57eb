// Hierarchical winding number on the GPU (formulas in wn_fwn.h). Each query
// walks the tree depth-first: a node whose expansion centre is farther than
// beta * radius contributes its second-order expansion (wn_fwn_eval.h, the
// code the CPU path runs), a nearer leaf contributes its faces exactly (FP64
// per-pair formula), a nearer inner node pushes its children. Same visiting
// order and arithmetic as the CPU path; the exact faces use the device's
// sqrt/atan2, so CPU and GPU agree to rounding.
//
// The 64 lanes of a wave walk together: one wave-uniform stack of (node,
// lanes still descending) entries, so every node is read once per wave, at a
// wave-uniform address (scalar loads), instead of once per lane. At a node,
// the lanes of its mask that find it far add its expansion, the others
// continue into its faces (leaf) or children (inner node, pushed with the
// mask of those lanes). A lane therefore adds exactly the contributions, in
// exactly the order, of its own depth-first walk; nodes no lane needs are
// never visited. The stack holds at most depth + 1 entries, as a per-lane
// walk's would (checked on the host against kAabbStack).
//
// Layout in HBM (one allocation per tree): FwnNode[nnodes] (304 B: centre,
// radius, 30 expansion coefficients, children and leaf range), then the face
// corners in leaf order (72 B per face). One-wave blocks; queries are walked
// in Morton order by default, so a wave's lanes share most of their walks
// (queries left in the given order use a per-lane walk instead).

#include "wn_device.h"
#include "wn_fwn.h"
#include "wn_fwn_eval.h"
#include "wn_internal.h"

namespace wn {

namespace {

constexpr int kFwnLanes = 64;

__device__ __forceinline__ uint64_t uniform_u64(uint64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

__global__ __launch_bounds__(kFwnLanes) void fwn_query_wave_kernel(
    const FwnNode* __restrict__ nodes, int64_t nnodes, const double* __restrict__ corners,
    const double* __restrict__ q, const int32_t* __restrict__ perm, double* __restrict__ w,
    int64_t nq, double beta2) {
  __shared__ int32_t s_node[kAabbStack];
  __shared__ uint64_t s_mask[kAabbStack];
  const int lane = threadIdx.x;
  const uint64_t bit = 1ull << lane;
  const int64_t i = (int64_t)blockIdx.x * kFwnLanes + lane;
  const bool valid = i < nq;
  const double qx = valid ? q[3 * i] : 0.0;
  const double qy = valid ? q[3 * i + 1] : 0.0;
  const double qz = valid ? q[3 * i + 2] : 0.0;
  double omega = 0.0;
  const uint64_t all = __builtin_amdgcn_ballot_w64(valid);
  int top = 0;  // wave-uniform
  if (nnodes > 0 && all != 0) {
    if (lane == 0) {
      s_node[0] = 0;
      s_mask[0] = all;
    }
    top = 1;
  }
  while (top > 0) {
    --top;
    // the entry was written by lane 0 of this wave: LDS operations of one
    // wave complete in order, so every lane reads the written value
    const int32_t n = __builtin_amdgcn_readfirstlane(s_node[top]);
    const uint64_t m = uniform_u64(s_mask[top]);
    const FwnNode* e = nodes + n;
    const bool mine = (m & bit) != 0;
    const double dx = qx - e->cx, dy = qy - e->cy, dz = qz - e->cz;
    const bool far = mine && dx * dx + dy * dy + dz * dz > beta2 * e->radius * e->radius;
    if (far) omega += fwn_far_omega(&e->cx, qx, qy, qz);
    const uint64_t near = __builtin_amdgcn_ballot_w64(mine && !far);
    if (near == 0) continue;
    if (e->count > 0) {
      const bool me = (near & bit) != 0;
      for (int64_t k = e->start; k < e->start + e->count; ++k) {
        const double* t = corners + 9 * k;
        if (me)
          omega += 2.0 * pair_f64(t[0] - qx, t[1] - qy, t[2] - qz, t[3] - qx, t[4] - qy,
                                  t[5] - qz, t[6] - qx, t[7] - qy, t[8] - qz);
      }
    } else {
      if (lane == 0) {
        s_node[top] = (int32_t)e->right;
        s_mask[top] = near;
        s_node[top + 1] = (int32_t)e->left;
        s_mask[top + 1] = near;
      }
      top += 2;
    }
  }
  if (valid) w[perm ? (int64_t)perm[i] : i] = omega * (0.25 / 3.14159265358979323846264338327950288);
}

// One lane per query, each with its own stack ([depth][lane] in LDS): for
// queries in the given order, whose lanes share little of their walks.
__global__ __launch_bounds__(kFwnLanes) void fwn_query_lane_kernel(
    const FwnNode* __restrict__ nodes, int64_t nnodes, const double* __restrict__ corners,
    const double* __restrict__ q, const int32_t* __restrict__ perm, double* __restrict__ w,
    int64_t nq, double beta2) {
  __shared__ int32_t stack[kAabbStack][kFwnLanes];
  const int lane = threadIdx.x;
  const int64_t i = (int64_t)blockIdx.x * kFwnLanes + lane;
  if (i >= nq) return;  // no barriers below
  const double qx = q[3 * i], qy = q[3 * i + 1], qz = q[3 * i + 2];
  double omega = 0.0;
  int top = 0;
  if (nnodes > 0) stack[top++][lane] = 0;
  while (top > 0) {
    const FwnNode* e = nodes + stack[--top][lane];
    const double dx = qx - e->cx, dy = qy - e->cy, dz = qz - e->cz;
    if (dx * dx + dy * dy + dz * dz > beta2 * e->radius * e->radius) {
      omega += fwn_far_omega(&e->cx, qx, qy, qz);
    } else if (e->count > 0) {
      for (int64_t k = e->start; k < e->start + e->count; ++k) {
        const double* t = corners + 9 * k;
        omega += 2.0 * pair_f64(t[0] - qx, t[1] - qy, t[2] - qz, t[3] - qx, t[4] - qy, t[5] - qz,
                                t[6] - qx, t[7] - qy, t[8] - qz);
      }
    } else {
      stack[top++][lane] = (int32_t)e->right;
      stack[top++][lane] = (int32_t)e->left;
    }
  }
  w[perm ? (int64_t)perm[i] : i] = omega * (0.25 / 3.14159265358979323846264338327950288);
}

}  // namespace

size_t fwn_device_bytes(int64_t nnodes, int64_t nf) {
  return align_up(sizeof(FwnNode) * (size_t)nnodes) + align_up(sizeof(double) * 9 * (size_t)nf);
}

hipError_t fwn_upload(const FwnTree& t, void* dev, hipStream_t stream) {
  char* p = (char*)dev;
  hipError_t e = hipMemcpyAsync(p, t.nodes.data(), sizeof(FwnNode) * t.nodes.size(),
                                hipMemcpyHostToDevice, stream);
  p += align_up(sizeof(FwnNode) * t.nodes.size());
  if (e == hipSuccess && !t.corners.empty())
    e = hipMemcpyAsync(p, t.corners.data(), sizeof(double) * t.corners.size(),
                       hipMemcpyHostToDevice, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);  // host vectors may go away after
  return e;
}

size_t fwn_workspace_bytes(int64_t nq, int sort) {
  if (!sort) return 0;
  return align_up((size_t)nq * sizeof(int32_t)) + align_up((size_t)nq * 3 * sizeof(double)) +
         morton_workspace_bytes(nq);
}

hipError_t fwn_query_launch(const void* dev, int64_t nnodes, const double* q, double* w,
                            int64_t nq, double beta, int sort, void* ws, size_t ws_bytes,
                            hipStream_t stream) {
  if (nq == 0) return hipSuccess;
  if (ws_bytes < fwn_workspace_bytes(nq, sort)) return hipErrorInvalidValue;
  const FwnNode* nodes = (const FwnNode*)dev;
  const double* corners = (const double*)((const char*)dev + align_up(sizeof(FwnNode) * (size_t)nnodes));
  int32_t* perm = nullptr;
  if (sort) {
    // queries in Morton order: neighbouring lanes walk the same nodes
    char* c = (char*)ws;
    perm = (int32_t*)c;
    c += align_up((size_t)nq * sizeof(int32_t));
    double* qs = (double*)c;
    c += align_up((size_t)nq * 3 * sizeof(double));
    hipError_t e = morton_order(q, nq, perm, c, morton_workspace_bytes(nq), stream);
    if (e == hipSuccess) e = gather_queries_launch(q, perm, nq, qs, stream);
    if (e != hipSuccess) return e;
    q = qs;
  }
  // Morton order: a wave's lanes walk mostly the same nodes, so they walk
  // them together; in the given order a wave's union of walks is several
  // times one walk (1.31M faces x 1M queries at beta 2: 39.5 ms together
  // against 6.8 ms per lane), so each lane walks alone
  const dim3 grid((unsigned)((nq + kFwnLanes - 1) / kFwnLanes));
  if (sort)
    fwn_query_wave_kernel<<<grid, kFwnLanes, 0, stream>>>(nodes, nnodes, corners, q, perm, w, nq,
                                                          beta * beta);
  else
    fwn_query_lane_kernel<<<grid, kFwnLanes, 0, stream>>>(nodes, nnodes, corners, q, perm, w, nq,
                                                          beta * beta);
  return hipGetLastError();
}

}  // namespace wn
