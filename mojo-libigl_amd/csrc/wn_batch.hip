// Batched forward over many meshes of uneven sizes in one persistent launch.
//
// No counterpart in the reference (one mesh per call, src/capi.mojo:366-398).
//
// Layout: all meshes' faces in one 9 x F_total SoA (mesh m owns faces
// [face_off[m], face_off[m+1])), all queries in one Q_total x 3 array (mesh m
// owns [query_off[m], query_off[m+1])), one Q_total output.
//
// Plan (host, once per batch shape): work items (mesh, query tile of
// 256*QPL queries, face chunk of `chunk_faces` faces), sorted by decreasing
// pair count (longest-processing-time first) so the tail of the persistent
// loop holds only small items. Persistent kernel: CUs x occupancy blocks,
// each pulls the next item index from an atomic counter until the list is
// exhausted (every block reaches the exit test each iteration). Items of a
// single-chunk mesh write w directly; multi-chunk meshes write unscaled
// partials to fixed slots, summed in chunk order by a reduce kernel, so the
// result does not depend on which block took which item.
//
// Strip plans hold every mesh as triangle-strip records (wn_strips.cpp) in
// one buffer, each mesh a whole number of 255-record tiles; items then cover
// chunks of tiles and run the strip kernel's sweep (wn_strip.hip).
//
// The fill-counter / persistent / reduce sequence can be captured once into
// a hipGraph and replayed (fill kernel, no memset node). The optional Morton
// ordering of queries (segmented by mesh) runs eagerly before the graph.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "wn_internal.h"
#include "wn_sweep.h"

namespace wn {

namespace {

__global__ void counter_reset_kernel(int32_t* counter) {
  if (threadIdx.x == 0) counter[0] = 0;
}

// FP32-compensated face plans (FP64 face plans run batch_f64f_kernel of
// wn_forward.hip over FP64 face records).
template <int QPL>
__global__ __launch_bounds__(kBlockThreads) void batch_kernel(
    const double* __restrict__ tri, int64_t ftot, const double* __restrict__ q,
    const int32_t* __restrict__ perm, double* __restrict__ dst, double* __restrict__ partial,
    const BatchMesh* __restrict__ meshes, const BatchItem* __restrict__ items, int64_t n_items,
    int32_t* __restrict__ counter, int64_t chunk_faces, float kappa) {
  __shared__ SweepF32Smem sm;
  __shared__ int32_t s_item;
  const int tid = threadIdx.x;
  constexpr int QB = kBlockThreads * QPL;
  for (;;) {
    __syncthreads();  // every thread has read the previous s_item
    if (tid == 0) s_item = atomicAdd(counter, 1);
    __syncthreads();
    const int32_t it = s_item;
    if (it >= n_items) break;  // block-uniform exit
    const BatchItem item = items[it];
    const BatchMesh m = meshes[item.mesh];
    const int64_t slot0 = m.query_off + (int64_t)item.qtile * QB;
    const int64_t slot_end = m.query_off + m.nq;
    const int64_t f0 = m.face_off + (int64_t)item.chunk * chunk_faces;
    int64_t f1 = f0 + chunk_faces;
    if (f1 > m.face_off + m.nf) f1 = m.face_off + m.nf;

    int64_t qidx[QPL];
    bool valid[QPL];
    double Ox[QPL], Oy[QPL], Oz[QPL], acc[QPL];
#pragma unroll
    for (int k = 0; k < QPL; ++k) {
      const int64_t slot = slot0 + k * kBlockThreads + tid;
      valid[k] = slot < slot_end;
      qidx[k] = valid[k] ? (perm ? (int64_t)perm[slot] : slot) : 0;
      // q is in launch order (gathered through perm beforehand when sorted)
      Ox[k] = valid[k] ? q[3 * slot] : 0.0;
      Oy[k] = valid[k] ? q[3 * slot + 1] : 0.0;
      Oz[k] = valid[k] ? q[3 * slot + 2] : 0.0;
      acc[k] = 0.0;
    }
    double gx, gy, gz;
    int er;
    block_origin<QPL>(Ox, Oy, Oz, valid, sm, gx, gy, gz, er);
    const double so = pow2_f64(2 - er);
    float ox[QPL], oy[QPL], oz[QPL];
#pragma unroll
    for (int k = 0; k < QPL; ++k) {
      ox[k] = valid[k] ? (float)((Ox[k] - gx) * so) : 0.0f;
      oy[k] = valid[k] ? (float)((Oy[k] - gy) * so) : 0.0f;
      oz[k] = valid[k] ? (float)((Oz[k] - gz) * so) : 0.0f;
    }
    sweep_f32c<QPL>(tri, ftot, f0, f1, gx, gy, gz, er, ox, oy, oz, kappa, acc, sm);
#pragma unroll
    for (int k = 0; k < QPL; ++k) {
      if (valid[k]) {
        const int64_t slot = slot0 + k * kBlockThreads + tid;
        if (m.chunks == 1)
          dst[qidx[k]] = acc[k] * kInv2Pi;
        else  // chunk-major, slot order: a block's writes are contiguous
          partial[m.pbase + (int64_t)item.chunk * m.nq + (slot - m.query_off)] = acc[k];
      }
    }
  }
}

// One thread per query slot: sums its mesh's chunk partials in chunk order and
// writes w through the query permutation.
__global__ __launch_bounds__(kBlockThreads) void batch_reduce_kernel(
    const double* __restrict__ partial, const int32_t* __restrict__ perm,
    double* __restrict__ dst, const BatchMesh* __restrict__ meshes,
    const int64_t* __restrict__ query_off, int n_meshes, int64_t qtot) {
  const int64_t slot = (int64_t)blockIdx.x * kBlockThreads + threadIdx.x;
  if (slot >= qtot) return;
  int lo = 0, hi = n_meshes;  // query_off[lo] <= slot < query_off[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (query_off[mid] <= slot) lo = mid; else hi = mid;
  }
  const BatchMesh m = meshes[lo];
  if (m.chunks <= 1) return;
  const double* p = partial + m.pbase + (slot - m.query_off);
  double s = 0.0;
  for (int k = 0; k < m.chunks; ++k) s += p[(int64_t)k * m.nq];
  dst[perm ? (int64_t)perm[slot] : slot] = s * kInv2Pi;
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

struct BatchPlan {
  bool strips = false;  // strip plan: face units are 255-record tiles
  int precision = kFP32C;
  int qpl = 1;
  int sort = 0;
  float kappa = 1.0f;
  int64_t chunk_faces = 2048;
  int n_meshes = 0;
  int64_t ftot = 0, qtot = 0, n_items = 0, n_partials = 0;
  int grid = 0;
  std::vector<int64_t> query_off;
  // device
  void* ws = nullptr;
  size_t ws_bytes = 0;
  int32_t* d_counter = nullptr;
  BatchMesh* d_meshes = nullptr;
  BatchItem* d_items = nullptr;
  int64_t* d_query_off = nullptr;
  double* d_partial = nullptr;
  int32_t* d_perm = nullptr;
  double* d_qs = nullptr;  // queries in launch order (sorted plans)
  void* d_sort_ws = nullptr;
  size_t sort_bytes = 0;
  StripRec64* d_rec64 = nullptr;  // FP64 strip plans: sweep records, rebuilt per run
  FaceRec64* d_frec64 = nullptr;  // FP64 face plans: face records, rebuilt per run
  // Boundary correction of FP32-compensated plans (wn_ribbon.hip): one block
  // per (mesh, query tile), each mesh's edge range. Face plans find their
  // boundary edges from the SoA on every run (the pass before the graph);
  // strip plans take them from wn_batch_set_boundary.
  bool boundary = false;
  int64_t n_rblocks = 0;
  void* ws2 = nullptr;  // rblocks, edge ranges, face offsets, the SoA pass
  int2* d_rblocks = nullptr;
  int32_t* d_e_begin = nullptr;
  int32_t* d_e_end = nullptr;
  int64_t* d_face_off = nullptr;
  void* d_sb_ws = nullptr;
  size_t sb_bytes = 0;
  SoaBoundary sb{};
  const int32_t* b_ids = nullptr;  // strip plans: caller's edges (record pairs)
  const double* b_w = nullptr;
  int64_t b_n = 0;
  // graph cache
  hipStream_t cap_stream = nullptr;
  hipGraphExec_t exec = nullptr;
  const void* g_tri = nullptr;
  const void* g_q = nullptr;
  void* g_dst = nullptr;
};

static const void* batch_kernel_ptr(int precision, int qpl) {
  if (precision == kFP64) return batch_face64_kernel(qpl);
  return qpl == 1 ? (const void*)batch_kernel<1> : (const void*)batch_kernel<2>;
}

const char* batch_check(const int64_t* face_off, const int64_t* query_off, int64_t n_meshes,
                        int precision, int qpl, int64_t chunk_faces) {
  if (n_meshes < 1 || n_meshes > (1 << 24)) return "n_meshes must be in [1, 2^24]";
  if (!face_off || !query_off) return "null offsets";
  if (precision != kFP64 && precision != kFP32C) return "unknown precision";
  if (qpl != 1 && qpl != 2) return "batch queries_per_lane must be 1 or 2";
  if (chunk_faces < kFaceTile || chunk_faces % kFaceTile) return "chunk_faces must be a positive multiple of 256";
  if (face_off[0] != 0 || query_off[0] != 0) return "offsets must start at 0";
  for (int64_t m = 0; m < n_meshes; ++m) {
    if (face_off[m + 1] < face_off[m] || query_off[m + 1] < query_off[m]) return "offsets must be non-decreasing";
    if (face_off[m + 1] - face_off[m] > 0x7fffffff || query_off[m + 1] - query_off[m] > 0x7fffffff)
      return "a mesh has more than 2^31-1 faces or queries";
  }
  if (query_off[n_meshes] > 0x7fffffff) return "more than 2^31-1 queries in the batch";
  return nullptr;
}

namespace {

// Items sorted by decreasing cost (longest processing time first), the
// persistent grid (resident blocks of the plan's kernel, at most one per
// item), then one device allocation holding counter, meshes, items,
// offsets, partial slots and the sort buffers.
hipError_t plan_finish(BatchPlan* p, const std::vector<BatchMesh>& meshes,
                       std::vector<std::pair<int64_t, BatchItem>>& items, int64_t n_partials,
                       const int64_t* query_off, int resident) {
  std::stable_sort(items.begin(), items.end(),
                   [](const auto& a, const auto& b) { return a.first > b.first; });
  p->n_items = (int64_t)items.size();
  p->n_partials = n_partials;
  std::vector<BatchItem> flat(items.size());
  for (size_t i = 0; i < items.size(); ++i) flat[i] = items[i].second;
  p->grid = (int)std::max<int64_t>(1, std::min<int64_t>(std::max(1, resident),
                                                        std::max<int64_t>(1, p->n_items)));
  const int64_t n_meshes = p->n_meshes;
  const size_t b_counter = align_up(sizeof(int32_t));
  const size_t b_meshes = align_up(sizeof(BatchMesh) * (size_t)n_meshes);
  const size_t b_items = align_up(sizeof(BatchItem) * std::max<size_t>(1, flat.size()));
  const size_t b_qoff = align_up(sizeof(int64_t) * (size_t)(n_meshes + 1));
  const size_t b_partial = align_up(sizeof(double) * (size_t)std::max<int64_t>(1, n_partials));
  const size_t b_perm = p->sort ? align_up(sizeof(int32_t) * (size_t)std::max<int64_t>(1, p->qtot)) : 0;
  const size_t b_qs = p->sort ? align_up(3 * sizeof(double) * (size_t)std::max<int64_t>(1, p->qtot)) : 0;
  p->sort_bytes = p->sort ? morton_segmented_workspace_bytes(p->qtot) : 0;
  p->ws_bytes = b_counter + b_meshes + b_items + b_qoff + b_partial + b_perm + b_qs + p->sort_bytes;
  hipError_t e = hipMalloc(&p->ws, p->ws_bytes);
  if (e != hipSuccess) {
    p->ws = nullptr;
    return e;
  }
  char* c = (char*)p->ws;
  p->d_counter = (int32_t*)c; c += b_counter;
  p->d_meshes = (BatchMesh*)c; c += b_meshes;
  p->d_items = (BatchItem*)c; c += b_items;
  p->d_query_off = (int64_t*)c; c += b_qoff;
  p->d_partial = (double*)c; c += b_partial;
  if (p->sort) {
    p->d_perm = (int32_t*)c; c += b_perm;
    p->d_qs = (double*)c; c += b_qs;
    p->d_sort_ws = c;
  }
  e = hipMemcpy(p->d_meshes, meshes.data(), sizeof(BatchMesh) * meshes.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess && !flat.empty())
    e = hipMemcpy(p->d_items, flat.data(), sizeof(BatchItem) * flat.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(p->d_query_off, query_off, sizeof(int64_t) * (size_t)(n_meshes + 1),
                  hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking);
  return e;
}

// Mesh table and work items for units of `unit` faces (1 for face plans,
// 255 records for strip plans): chunks of `chunk` units, query tiles of QB.
int64_t plan_items(const int64_t* unit_off, const int64_t* query_off, int64_t n_meshes,
                   int64_t chunk, int64_t QB, int64_t faces_per_unit, int64_t overhead,
                   std::vector<BatchMesh>& meshes,
                   std::vector<std::pair<int64_t, BatchItem>>& items) {
  meshes.assign((size_t)n_meshes, BatchMesh{});
  int64_t pbase = 0;
  for (int64_t m = 0; m < n_meshes; ++m) {
    BatchMesh bm{};
    bm.face_off = unit_off[m];
    bm.query_off = query_off[m];
    bm.nf = (int32_t)(unit_off[m + 1] - unit_off[m]);
    bm.nq = (int32_t)(query_off[m + 1] - query_off[m]);
    bm.chunks = (int32_t)std::max<int64_t>(1, (bm.nf + chunk - 1) / chunk);
    bm.pbase = pbase;
    if (bm.chunks > 1) pbase += (int64_t)bm.nq * bm.chunks;
    meshes[(size_t)m] = bm;
    const int64_t tiles = (bm.nq + QB - 1) / QB;
    for (int64_t j = 0; j < tiles; ++j) {
      const int64_t qn = std::min<int64_t>(QB, bm.nq - j * QB);
      for (int32_t k = 0; k < bm.chunks; ++k) {
        const int64_t fn = std::min<int64_t>(chunk, bm.nf - (int64_t)k * chunk) * faces_per_unit;
        // cost in pairs, plus a fixed per-item overhead
        const int64_t cost = qn * std::max<int64_t>(fn, 0) + overhead;
        items.push_back({cost, BatchItem{(int32_t)m, (int32_t)j, k, 0}});
      }
    }
  }
  return pbase;
}

// The boundary correction's block table ((mesh, query tile) for query
// tiles of QB queries) plus per-mesh edge ranges, and, for a face plan, its
// face offsets and the SoA boundary pass's workspace: one allocation.
hipError_t plan_boundary(BatchPlan* p, const int64_t* face_off, int64_t QB) {
  std::vector<int2> blocks;
  for (int m = 0; m < p->n_meshes; ++m) {
    const int64_t nq = p->query_off[(size_t)m + 1] - p->query_off[(size_t)m];
    for (int64_t j = 0; j * QB < nq; ++j) blocks.push_back(make_int2(m, (int)j));
  }
  p->n_rblocks = (int64_t)blocks.size();
  const size_t b_blocks = align_up(sizeof(int2) * std::max<size_t>(1, blocks.size()));
  const size_t b_range = align_up(sizeof(int32_t) * (size_t)p->n_meshes);
  const size_t b_foff = face_off ? align_up(sizeof(int64_t) * (size_t)(p->n_meshes + 1)) : 0;
  p->sb_bytes = face_off ? soa_boundary_workspace_bytes(p->ftot, p->n_meshes) : 0;
  hipError_t e = hipMalloc(&p->ws2, b_blocks + 2 * b_range + b_foff + p->sb_bytes);
  if (e != hipSuccess) {
    p->ws2 = nullptr;
    return e;
  }
  char* c = (char*)p->ws2;
  p->d_rblocks = (int2*)c; c += b_blocks;
  p->d_e_begin = (int32_t*)c; c += b_range;
  p->d_e_end = (int32_t*)c; c += b_range;
  if (face_off) {
    p->d_face_off = (int64_t*)c; c += b_foff;
    p->d_sb_ws = c;
  }
  if (!blocks.empty())
    e = hipMemcpy(p->d_rblocks, blocks.data(), sizeof(int2) * blocks.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess && face_off)
    e = hipMemcpy(p->d_face_off, face_off, sizeof(int64_t) * (size_t)(p->n_meshes + 1),
                  hipMemcpyHostToDevice);
  return e;
}

}  // namespace

hipError_t batch_create(const int64_t* face_off, const int64_t* query_off, int64_t n_meshes,
                        int precision, int qpl, int64_t chunk_faces, int sort, float kappa,
                        BatchPlan** out) {
  BatchPlan* p = new BatchPlan();
  p->precision = precision;
  p->qpl = qpl;
  p->sort = sort;
  p->kappa = kappa;
  p->chunk_faces = chunk_faces;
  p->n_meshes = (int)n_meshes;
  p->ftot = face_off[n_meshes];
  p->qtot = query_off[n_meshes];
  p->query_off.assign(query_off, query_off + n_meshes + 1);

  const int64_t QB = (int64_t)kBlockThreads * qpl;
  std::vector<BatchMesh> meshes;
  std::vector<std::pair<int64_t, BatchItem>> items;
  const int64_t pbase = plan_items(face_off, query_off, n_meshes, chunk_faces, QB, 1,
                                   QB * kFaceTile, meshes, items);

  // persistent grid: CUs x resident blocks of this kernel instance
  int dev = 0, cus = 0, per_cu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel_ptr(precision, qpl),
                                                     kBlockThreads, 0);
  if (e == hipSuccess)
    e = plan_finish(p, meshes, items, pbase, query_off, std::max(1, cus) * std::max(1, per_cu));
  if (e == hipSuccess && precision == kFP32C) {
    e = plan_boundary(p, face_off, QB);
    p->boundary = e == hipSuccess;
  }
  if (e == hipSuccess && precision == kFP64 && p->ftot > 0)
    e = hipMalloc(&p->d_frec64, (size_t)p->ftot * sizeof(FaceRec64));
  if (e != hipSuccess) {
    if (p->ws) (void)hipFree(p->ws);
    if (p->ws2) (void)hipFree(p->ws2);
    if (p->d_frec64) (void)hipFree(p->d_frec64);
    delete p;
    return e;
  }
  *out = p;
  return hipSuccess;
}

const char* batch_strip_check(const int64_t* rec_off, const int64_t* query_off, int64_t n_meshes,
                              int64_t chunk_tiles) {
  if (n_meshes < 1 || n_meshes > (1 << 24)) return "n_meshes must be in [1, 2^24]";
  if (!rec_off || !query_off) return "null offsets";
  if (chunk_tiles < 1) return "chunk_tiles must be positive";
  if (rec_off[0] != 0 || query_off[0] != 0) return "offsets must start at 0";
  for (int64_t m = 0; m < n_meshes; ++m) {
    if (rec_off[m + 1] < rec_off[m] || query_off[m + 1] < query_off[m]) return "offsets must be non-decreasing";
    if ((rec_off[m + 1] - rec_off[m]) % 255) return "each mesh's record count must be a multiple of 255";
    if ((rec_off[m + 1] - rec_off[m]) / 255 > 0x7fffffff || query_off[m + 1] - query_off[m] > 0x7fffffff)
      return "a mesh has more than 2^31-1 tiles or queries";
  }
  if (query_off[n_meshes] > 0x7fffffff) return "more than 2^31-1 queries in the batch";
  return nullptr;
}

hipError_t batch_create_strips(const int64_t* rec_off, const int64_t* query_off, int64_t n_meshes,
                               int64_t chunk_tiles, int sort, float kappa, int precision,
                               BatchPlan** out) {
  BatchPlan* p = new BatchPlan();
  p->strips = true;
  p->precision = precision;
  p->qpl = 1;
  p->sort = sort;
  p->kappa = kappa;
  p->chunk_faces = chunk_tiles;  // in tiles
  p->n_meshes = (int)n_meshes;
  p->ftot = rec_off[n_meshes] / 255;  // tiles
  p->qtot = query_off[n_meshes];
  p->query_off.assign(query_off, query_off + n_meshes + 1);

  std::vector<int64_t> tile_off((size_t)n_meshes + 1);
  for (int64_t m = 0; m <= n_meshes; ++m) tile_off[(size_t)m] = rec_off[m] / 255;
  std::vector<BatchMesh> meshes;
  std::vector<std::pair<int64_t, BatchItem>> items;
  const int64_t pbase = plan_items(tile_off.data(), query_off, n_meshes, chunk_tiles,
                                   kBlockThreads, 255, kBlockThreads * 255, meshes, items);
  const int resident =
      precision == kFP64 ? batch_strip64_resident_blocks() : batch_strip_resident_blocks();
  hipError_t e = resident > 0 ? hipSuccess : hipErrorNoDevice;
  if (e == hipSuccess) e = plan_finish(p, meshes, items, pbase, query_off, resident);
  if (e == hipSuccess && precision == kFP64 && p->ftot > 0)
    e = hipMalloc(&p->d_rec64, (size_t)p->ftot * 255 * sizeof(StripRec64));
  if (e != hipSuccess) {
    if (p->ws) (void)hipFree(p->ws);
    if (p->d_rec64) (void)hipFree(p->d_rec64);
    delete p;
    return e;
  }
  *out = p;
  return hipSuccess;
}

static hipError_t face_batch_launch(const BatchPlan* p, const double* tri, const double* q,
                                    const int32_t* perm, double* dst, hipStream_t s) {
  if (p->precision == kFP64) {
    // (no faces at all: no records, and every item writes w = 0)
    hipError_t e = face_rec64_launch(tri, p->ftot, p->d_frec64, s);
    if (e == hipSuccess)
      e = batch_face64_enqueue(p->d_frec64, q, perm, dst, p->d_partial, p->d_meshes, p->d_items,
                               p->n_items, p->d_counter, p->chunk_faces, p->qpl, p->grid, s);
    return e;
  }
  const dim3 grid((unsigned)p->grid);
  if (p->qpl == 1)
    batch_kernel<1><<<grid, kBlockThreads, 0, s>>>(tri, p->ftot, q, perm, dst, p->d_partial,
                                                   p->d_meshes, p->d_items, p->n_items,
                                                   p->d_counter, p->chunk_faces, p->kappa);
  else
    batch_kernel<2><<<grid, kBlockThreads, 0, s>>>(tri, p->ftot, q, perm, dst, p->d_partial,
                                                   p->d_meshes, p->d_items, p->n_items,
                                                   p->d_counter, p->chunk_faces, p->kappa);
  return hipGetLastError();
}

// tri: the 9 x F_total face SoA, or the strip records of a strip plan.
static hipError_t batch_enqueue(BatchPlan* p, const double* tri, const double* q, double* dst,
                                hipStream_t s) {
  counter_reset_kernel<<<1, 64, 0, s>>>(p->d_counter);
  const int32_t* perm = p->sort ? p->d_perm : nullptr;
  if (p->sort) q = p->d_qs;
  if (p->strips && p->precision == kFP64) {
    hipError_t e = strip64_prep_launch((const double4*)tri, p->ftot * 255, p->d_rec64, s);
    if (e == hipSuccess)
      e = batch_strip64_enqueue(p->d_rec64, (const double4*)tri, q, perm, dst, p->d_partial,
                                p->d_meshes, p->d_items, p->n_items, p->d_counter,
                                p->chunk_faces, p->grid, s);
    if (e != hipSuccess) return e;
  } else if (p->strips) {
    const hipError_t e = batch_strip_enqueue((const double4*)tri, q, perm, dst, p->d_partial,
                                             p->d_meshes, p->d_items, p->n_items, p->d_counter,
                                             p->chunk_faces, p->kappa, p->grid, s);
    if (e != hipSuccess) return e;
  } else {
    const hipError_t e = face_batch_launch(p, tri, q, perm, dst, s);
    if (e != hipSuccess) return e;
  }
  if (p->n_partials > 0 && p->qtot > 0)
    batch_reduce_kernel<<<dim3((unsigned)((p->qtot + kBlockThreads - 1) / kBlockThreads)),
                          kBlockThreads, 0, s>>>(p->d_partial, perm, dst, p->d_meshes,
                                                 p->d_query_off, p->n_meshes, p->qtot);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !p->boundary || p->precision != kFP32C) return e;
  // boundary correction of the open meshes, for the same query blocks
  RibbonSrc src;
  src.pts = tri;
  if (p->strips) {
    src.ids = p->b_ids;
    src.w = p->b_w;
    src.n = p->b_n;
    return ribbon_batch_launch(src, 0, 1, p->d_e_begin, p->d_e_end, p->d_rblocks, p->n_rblocks,
                               p->d_meshes, q, perm, dst, s);
  }
  src.stride = p->ftot;
  src.ids = p->sb.ids;
  src.w = p->sb.w;
  src.count = p->sb.count;
  return ribbon_batch_launch(src, 1, p->qpl, p->sb.e_begin, p->sb.e_end, p->d_rblocks,
                             p->n_rblocks, p->d_meshes, q, perm, dst, s);
}

hipError_t batch_run(BatchPlan* p, const double* tri, const double* q, double* dst,
                     bool use_graph, hipStream_t stream) {
  if (p->qtot == 0) return hipSuccess;
  hipError_t e;
  if (p->boundary && !p->strips) {
    // boundary edges of this run's faces (the buffers keep their addresses)
    e = soa_boundary_launch(tri, p->ftot, p->d_face_off, p->n_meshes, p->d_sb_ws, p->sb_bytes,
                            &p->sb, stream);
    if (e != hipSuccess) return e;
  }
  if (p->sort) {
    e = morton_order_segmented(q, p->qtot, p->d_query_off, p->n_meshes, p->d_perm,
                               p->d_sort_ws, p->sort_bytes, stream);
    if (e == hipSuccess) e = gather_queries_launch(q, p->d_perm, p->qtot, p->d_qs, stream);
    if (e != hipSuccess) return e;
  }
  if (!use_graph) return batch_enqueue(p, tri, q, dst, stream);
  if (!p->exec || p->g_tri != tri || p->g_q != q || p->g_dst != dst) {
    if (p->exec) {
      (void)hipGraphExecDestroy(p->exec);
      p->exec = nullptr;
    }
    e = hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    hipError_t le = batch_enqueue(p, tri, q, dst, p->cap_stream);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(p->cap_stream, &graph);
    if (le != hipSuccess) e = le;
    if (e == hipSuccess) e = hipGraphInstantiate(&p->exec, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
      p->exec = nullptr;
      return e;
    }
    p->g_tri = tri;
    p->g_q = q;
    p->g_dst = dst;
  }
  return hipGraphLaunch(p->exec, stream);
}

void batch_destroy(BatchPlan* p) {
  if (!p) return;
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->cap_stream) {
    (void)hipStreamSynchronize(p->cap_stream);
    (void)hipStreamDestroy(p->cap_stream);
  }
  if (p->ws || p->d_rec64 || p->d_frec64 || p->ws2) (void)hipDeviceSynchronize();
  if (p->ws) (void)hipFree(p->ws);
  if (p->d_rec64) (void)hipFree(p->d_rec64);
  if (p->d_frec64) (void)hipFree(p->d_frec64);
  if (p->ws2) (void)hipFree(p->ws2);
  delete p;
}

const char* batch_boundary_check(const BatchPlan* p, int64_t ne, const int64_t* edge_off) {
  if (!p->strips) return "boundary edges are set on strip plans (face plans find their own)";
  if (ne < 0 || ne > 0x7fffffff) return "edge count out of range";
  if (!edge_off) return "null edge offsets";
  if (edge_off[0] != 0 || edge_off[p->n_meshes] != ne) return "edge offsets must run from 0 to the edge count";
  for (int m = 0; m < p->n_meshes; ++m)
    if (edge_off[m + 1] < edge_off[m]) return "edge offsets must be non-decreasing";
  return nullptr;
}

hipError_t batch_set_boundary(BatchPlan* p, const int32_t* ids, const double* w, int64_t ne,
                              const int64_t* edge_off) {
  if (p->exec) {  // a captured graph holds the previous edges
    (void)hipGraphExecDestroy(p->exec);
    p->exec = nullptr;
  }
  p->boundary = false;
  if (!p->ws2) {
    const hipError_t e = plan_boundary(p, nullptr, kBlockThreads);
    if (e != hipSuccess) return e;
  }
  std::vector<int32_t> b((size_t)p->n_meshes), en((size_t)p->n_meshes);
  for (int m = 0; m < p->n_meshes; ++m) {
    b[(size_t)m] = (int32_t)edge_off[m];
    en[(size_t)m] = (int32_t)edge_off[m + 1];
  }
  hipError_t e = hipMemcpy(p->d_e_begin, b.data(), sizeof(int32_t) * b.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(p->d_e_end, en.data(), sizeof(int32_t) * en.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) return e;
  p->b_ids = ids;
  p->b_w = w;
  p->b_n = ne;
  p->boundary = ne > 0;
  return hipSuccess;
}

void batch_info(const BatchPlan* p, int64_t* out) {
  out[0] = p->n_items;
  out[1] = p->n_partials;
  out[2] = (int64_t)p->ws_bytes;
  out[3] = p->grid;
  out[4] = p->ftot;
  out[5] = p->qtot;
}

}  // namespace wn
