// Hierarchical winding number on the GPU (formulas in wn_fwn.h): one query
// per lane walks the tree depth-first -- a node whose expansion centre is
// farther than beta * radius contributes its second-order expansion
// (wn_fwn_eval.h, the code the CPU path runs), a nearer leaf contributes its
// faces exactly (FP64 per-pair formula), a nearer inner node pushes its
// children. Same visiting order and arithmetic as the CPU path; the exact
// faces use the device's sqrt/atan2, so CPU and GPU agree to rounding.
//
// Layout in HBM (one allocation per tree): FwnNode[nnodes] (304 B: centre,
// radius, 30 expansion coefficients, children and leaf range), then the face
// corners in leaf order (72 B per face). One-wave blocks (walks diverge per
// lane); descent stack in LDS, [depth][lane]. Queries are walked in Morton
// order by default (1.4-1.5x faster on a 1.31M-face mesh x 1M queries).

#include "wn_device.h"
#include "wn_fwn.h"
#include "wn_fwn_eval.h"
#include "wn_internal.h"

namespace wn {

namespace {

constexpr int kFwnLanes = 64;

__global__ __launch_bounds__(kFwnLanes) void fwn_query_kernel(
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
  fwn_query_kernel<<<dim3((unsigned)((nq + kFwnLanes - 1) / kFwnLanes)), kFwnLanes, 0, stream>>>(
      nodes, nnodes, corners, q, perm, w, nq, beta * beta);
  return hipGetLastError();
}

}  // namespace wn
