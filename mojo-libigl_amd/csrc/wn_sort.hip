// Morton (Z-order) ordering of query points, so that a block of consecutive
// slots covers a compact region of space. Used by the FP32-compensated
// forward pass: its per-block local origin then sits close to every query in
// the block, which shrinks the FP32 rounding of the block's mesh copy near
// those queries, and ill-conditioned (FP64) pairs cluster in few blocks.
//
// bbox (order-independent atomic min/max on order-preserving integer keys, so
// deterministic) -> 30-bit codes -> rocPRIM/hipCUB radix sort of (code, index).

#include <hipcub/hipcub.hpp>

#include "wn_internal.h"

namespace wn {

namespace {

__device__ __forceinline__ unsigned long long order_key(double x) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__device__ __forceinline__ double from_order_key(unsigned long long k) {
  const unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double((long long)b);
}

__global__ __launch_bounds__(kBlockThreads) void bbox_init_kernel(unsigned long long* box) {
  if (threadIdx.x < 3) box[threadIdx.x] = ~0ull;   // min slots
  else if (threadIdx.x < 6) box[threadIdx.x] = 0ull;  // max slots
}

__global__ __launch_bounds__(kBlockThreads) void bbox_kernel(const double* __restrict__ q,
                                                             int64_t nq,
                                                             unsigned long long* box) {
  __shared__ unsigned long long s[6][kBlockThreads];
  const int tid = threadIdx.x;
  unsigned long long mn[3] = {~0ull, ~0ull, ~0ull}, mx[3] = {0ull, 0ull, 0ull};
  for (int64_t i = (int64_t)blockIdx.x * kBlockThreads + tid; i < nq;
       i += (int64_t)gridDim.x * kBlockThreads) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const unsigned long long k = order_key(q[3 * i + d]);
      mn[d] = k < mn[d] ? k : mn[d];
      mx[d] = k > mx[d] ? k : mx[d];
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    s[d][tid] = mn[d];
    s[3 + d][tid] = mx[d];
  }
  __syncthreads();
  for (int w = kBlockThreads / 2; w > 0; w >>= 1) {
    if (tid < w) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        s[d][tid] = s[d][tid + w] < s[d][tid] ? s[d][tid + w] : s[d][tid];
        s[3 + d][tid] = s[3 + d][tid + w] > s[3 + d][tid] ? s[3 + d][tid + w] : s[3 + d][tid];
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      atomicMin(&box[d], s[d][0]);
      atomicMax(&box[3 + d], s[3 + d][0]);
    }
  }
}

// Blocks of the bbox pass: each ends in 6 global atomics on the same 6
// words, which serialise, so the grid stays small (<= 256 blocks, >= 16
// points per thread); the reads are a few microseconds of HBM either way.
inline unsigned bbox_blocks(int64_t nq) {
  int64_t blocks = (nq + 16 * kBlockThreads - 1) / (16 * kBlockThreads);
  if (blocks > 256) blocks = 256;
  return (unsigned)(blocks < 1 ? 1 : blocks);
}

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

__global__ __launch_bounds__(kBlockThreads) void morton_kernel(
    const double* __restrict__ q, int64_t nq, const unsigned long long* __restrict__ box,
    uint32_t* __restrict__ keys, int32_t* __restrict__ vals) {
  const int64_t i = (int64_t)blockIdx.x * kBlockThreads + threadIdx.x;
  if (i >= nq) return;
  uint32_t c[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double lo = from_order_key(box[d]);
    const double hi = from_order_key(box[3 + d]);
    const double ext = hi - lo;
    double t = ext > 0.0 ? (q[3 * i + d] - lo) / ext : 0.0;
    t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
    c[d] = (uint32_t)(t * 1023.0);
  }
  keys[i] = spread10(c[0]) | (spread10(c[1]) << 1) | (spread10(c[2]) << 2);
  vals[i] = (int32_t)i;
}

size_t sort_temp_bytes(int64_t nq) {
  size_t temp = 0;
  hipcub::DeviceRadixSort::SortPairs(nullptr, temp, (const uint32_t*)nullptr,
                                     (uint32_t*)nullptr, (const int32_t*)nullptr,
                                     (int32_t*)nullptr, (int)nq, 0, 30);
  return temp;
}

}  // namespace

__global__ __launch_bounds__(kBlockThreads) void gather_queries_kernel(
    const double* __restrict__ q, const int32_t* __restrict__ perm, int64_t nq,
    double* __restrict__ qs) {
  const int64_t slot = (int64_t)blockIdx.x * kBlockThreads + threadIdx.x;
  if (slot >= nq) return;
  const int64_t i = perm[slot];
  qs[3 * slot] = q[3 * i];
  qs[3 * slot + 1] = q[3 * i + 1];
  qs[3 * slot + 2] = q[3 * i + 2];
}

hipError_t gather_queries_launch(const double* q, const int32_t* perm, int64_t nq, double* qs,
                                 hipStream_t stream) {
  if (nq <= 0) return hipSuccess;
  gather_queries_kernel<<<(unsigned)((nq + kBlockThreads - 1) / kBlockThreads), kBlockThreads, 0,
                          stream>>>(q, perm, nq, qs);
  return hipGetLastError();
}

size_t morton_workspace_bytes(int64_t nq) {
  size_t b = align_up(6 * sizeof(unsigned long long));
  b += 3 * align_up((size_t)nq * sizeof(uint32_t));
  b += align_up(sort_temp_bytes(nq));
  return b;
}

hipError_t morton_order(const double* q, int64_t nq, int32_t* perm, void* ws,
                        size_t ws_bytes, hipStream_t stream) {
  if (nq <= 0) return hipSuccess;
  if (ws_bytes < morton_workspace_bytes(nq)) return hipErrorInvalidValue;
  char* cur = (char*)ws;
  unsigned long long* box = (unsigned long long*)cur;
  cur += align_up(6 * sizeof(unsigned long long));
  uint32_t* keys_in = (uint32_t*)cur;
  cur += align_up((size_t)nq * sizeof(uint32_t));
  uint32_t* keys_out = (uint32_t*)cur;
  cur += align_up((size_t)nq * sizeof(uint32_t));
  int32_t* vals_in = (int32_t*)cur;
  cur += align_up((size_t)nq * sizeof(uint32_t));
  size_t temp = sort_temp_bytes(nq);

  bbox_init_kernel<<<1, kBlockThreads, 0, stream>>>(box);
  bbox_kernel<<<bbox_blocks(nq), kBlockThreads, 0, stream>>>(q, nq, box);
  morton_kernel<<<(unsigned)((nq + kBlockThreads - 1) / kBlockThreads), kBlockThreads, 0,
                  stream>>>(q, nq, box, keys_in, vals_in);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return hipcub::DeviceRadixSort::SortPairs(cur, temp, keys_in, keys_out, vals_in, perm,
                                            (int)nq, 0, 30, stream);
}

// ---------------------------------------------------------------------------
// Segmented variant: key = (segment << 30) | code, codes on the bbox of all
// queries; the segment bits keep every segment in place.
// ---------------------------------------------------------------------------

namespace {

__global__ __launch_bounds__(kBlockThreads) void morton_seg_kernel(
    const double* __restrict__ q, int64_t nq, const unsigned long long* __restrict__ box,
    const int64_t* __restrict__ seg_off, int nseg, unsigned long long* __restrict__ keys,
    int32_t* __restrict__ vals) {
  const int64_t i = (int64_t)blockIdx.x * kBlockThreads + threadIdx.x;
  if (i >= nq) return;
  int lo = 0, hi = nseg;  // seg_off[lo] <= i < seg_off[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (seg_off[mid] <= i) lo = mid; else hi = mid;
  }
  uint32_t c[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double a = from_order_key(box[d]);
    const double b = from_order_key(box[3 + d]);
    const double ext = b - a;
    double t = ext > 0.0 ? (q[3 * i + d] - a) / ext : 0.0;
    t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
    c[d] = (uint32_t)(t * 1023.0);
  }
  const uint32_t code = spread10(c[0]) | (spread10(c[1]) << 1) | (spread10(c[2]) << 2);
  keys[i] = ((unsigned long long)lo << 30) | code;
  vals[i] = (int32_t)i;
}

size_t seg_sort_temp_bytes(int64_t nq) {
  size_t temp = 0;
  hipcub::DeviceRadixSort::SortPairs(nullptr, temp, (const unsigned long long*)nullptr,
                                     (unsigned long long*)nullptr, (const int32_t*)nullptr,
                                     (int32_t*)nullptr, (int)nq, 0, 64);
  return temp;
}

int bits_for(int n) {
  int b = 0;
  while ((1ll << b) < (long long)n) ++b;
  return b;
}

}  // namespace

size_t morton_segmented_workspace_bytes(int64_t nq) {
  size_t b = align_up(6 * sizeof(unsigned long long));
  b += 2 * align_up((size_t)nq * sizeof(unsigned long long));
  b += align_up((size_t)nq * sizeof(int32_t));
  b += align_up(seg_sort_temp_bytes(nq));
  return b;
}

hipError_t morton_order_segmented(const double* q, int64_t nq, const int64_t* seg_off, int nseg,
                                  int32_t* perm, void* ws, size_t ws_bytes, hipStream_t stream) {
  if (nq <= 0) return hipSuccess;
  if (ws_bytes < morton_segmented_workspace_bytes(nq)) return hipErrorInvalidValue;
  char* cur = (char*)ws;
  unsigned long long* box = (unsigned long long*)cur;
  cur += align_up(6 * sizeof(unsigned long long));
  unsigned long long* keys_in = (unsigned long long*)cur;
  cur += align_up((size_t)nq * sizeof(unsigned long long));
  unsigned long long* keys_out = (unsigned long long*)cur;
  cur += align_up((size_t)nq * sizeof(unsigned long long));
  int32_t* vals_in = (int32_t*)cur;
  cur += align_up((size_t)nq * sizeof(int32_t));
  size_t temp = seg_sort_temp_bytes(nq);

  bbox_init_kernel<<<1, kBlockThreads, 0, stream>>>(box);
  bbox_kernel<<<bbox_blocks(nq), kBlockThreads, 0, stream>>>(q, nq, box);
  morton_seg_kernel<<<(unsigned)((nq + kBlockThreads - 1) / kBlockThreads), kBlockThreads, 0,
                      stream>>>(q, nq, box, seg_off, nseg, keys_in, vals_in);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int end_bit = 30 + bits_for(nseg);
  return hipcub::DeviceRadixSort::SortPairs(cur, temp, keys_in, keys_out, vals_in, perm,
                                            (int)nq, 0, end_bit > 64 ? 64 : end_bit, stream);
}

}  // namespace wn
