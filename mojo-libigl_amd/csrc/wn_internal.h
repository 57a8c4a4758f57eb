// Host-side interfaces between the C-ABI layer and the kernel launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "wn_aabb.h"
#include "wn_fwn.h"

namespace wn {

// kFP64: FP64 with the far-pair fast path (refined square roots, an
// arctangent polynomial for far pairs, the reference formula for the rest);
// kFP32C: FP32 pairs, FP64 accumulation and re-evaluation; kFP64Plain: the
// reference formula in FP64 for every pair (face kernels only).
enum Precision : int { kFP64 = 0, kFP32C = 1, kFP64Plain = 2 };

constexpr int kBlockThreads = 256;  // 4 waves
constexpr int kFaceTile = 256;      // faces staged in LDS per tile (one per thread)
constexpr int kStripTileRecords = 255;  // strip records per tile (strips cut at tile ends)

// FP64 sweep record of strip record k (wn_strip64.hip), built on the device
// from the records: its vertex, minus half the squared lengths to the
// previous two vertices of its tile, and the normal of the face it closes,
// oriented by its flag (zero when it closes none).
struct alignas(64) StripRec64 {
  double x, y, z, e1, e2, nx, ny, nz;
};

// FP64 face record of the FP64 face kernel (wn_forward.hip), built on the
// device from the face SoA: its corners, its normal (B - A) x (C - A),
// and minus half its squared edge lengths |B-A|^2, |C-B|^2, |A-C|^2.
struct alignas(128) FaceRec64 {
  double ax, ay, az, bx, by, bz, cx, cy, cz, nx, ny, nz, eab, ebc, eca, pad;
};

struct FwdConfig {
  int precision = kFP64;  // a Precision
  int qpl = 1;            // queries per lane (1, 2 or 4)
  int splits = 1;         // face splits (split-K over faces), >= 1
  int sort = 0;           // kFP32C: process queries in Morton order
  float kappa = 1.0f;     // kFP32C: conditioning threshold for the FP64 re-evaluation
};

// Device timing of the dominant kernel of each forward call, off unless
// enabled (wn_kernel_timer): HIP events recorded on the launch stream right
// before and after the kernel; nothing is recorded on a capturing stream.
void timer_begin(hipStream_t s);
void timer_end(hipStream_t s);

// Rounds to 256 bytes so sub-buffers stay aligned.
inline size_t align_up(size_t n, size_t a = 256) { return (n + a - 1) / a * a; }

// p[0:n] = v; dst[perm[slot]] = sum over splits of partial[s * nq + slot] / (2 pi)
// (perm may be null: identity).
hipError_t fill_launch(double* p, int64_t n, double v, hipStream_t stream);
hipError_t reduce_splits_launch(const double* partial, const int32_t* perm, double* dst,
                                int64_t nq, int splits, hipStream_t stream);

// Strip forward (wn_strips.cpp builds records on the host, wn_strip.hip runs
// them). Records: 4 doubles each (x, y, z, flag), a multiple of 255 per mesh.
int64_t strip_records(const double* V, const void* F, int64_t nf, int fdtype,
                      std::vector<double>& out);
// The same records as vertex ids and flags (host), and the device gather
// that turns them into records for device vertices V (nv x 3; vdtype 0 =
// float64, 1 = float32): rec[k] = (V[idx[k]], flag[k]).
int64_t strip_topology(const void* F, int64_t nf, int fdtype, std::vector<int32_t>& idx,
                       std::vector<int8_t>& flag);
// Distinct corners of a face SoA (tri, 9 x nf) by exact coordinate equality
// as V (nv x 3) and the faces' corner ids as F (nf x 3 int32); returns nv.
int64_t soa_vertices(const double* tri, int64_t nf, std::vector<double>& V,
                     std::vector<int32_t>& F);
hipError_t strip_gather_launch(const void* V, int vdtype, int64_t nv, const int32_t* idx,
                               const int8_t* flag, int64_t nrec, double4* rec,
                               hipStream_t stream);
struct RibbonSrc;
int strip_resident_blocks();
// nbe: boundary edges of the mesh (the correction's partials when it splits)
size_t strip_workspace_bytes(int64_t nq, int splits, int sort, int64_t nbe = 0);
// FP64 strip forward (wn_strip64.hip): sweep records from the records, and
// the launch (same workspace layout and size as strip_launch).
hipError_t strip64_prep_launch(const double4* rec, int64_t nrec, StripRec64* out,
                               hipStream_t stream);
hipError_t strip64_launch(const StripRec64* rec64, const double4* rec, int64_t nrec,
                          const double* q, double* dst, int64_t nq, int splits, int sort,
                          void* ws, size_t ws_bytes, hipStream_t stream);
// boundary (optional): the mesh's boundary edges as record pairs (pts is
// taken to be rec); the correction runs after the sweep when it has edges.
hipError_t strip_launch(const double4* rec, int64_t nrec, const double* q, double* dst,
                        int64_t nq, int splits, int sort, float kappa, void* ws, size_t ws_bytes,
                        hipStream_t stream, const RibbonSrc* boundary = nullptr);

// Forward pass (wn_forward.hip).
// Blocks of the forward kernel (precision, qpl) resident at once on the
// current device (CUs x blocks per CU), 0 if no device answers.
int fwd_resident_blocks(int precision, int qpl);
size_t fwd_workspace_bytes(int64_t nf, int64_t nq, const FwdConfig& cfg);
const char* fwd_check(int64_t nf, int64_t nq, const FwdConfig& cfg);  // nullptr if ok
hipError_t fwd_launch(const double* tri, const double* q, double* dst, int64_t nf,
                      int64_t nq, const FwdConfig& cfg, void* ws, size_t ws_bytes,
                      hipStream_t stream);

// Morton ordering of query points (wn_sort.hip).
size_t morton_workspace_bytes(int64_t nq);
hipError_t morton_order(const double* q, int64_t nq, int32_t* perm, void* ws,
                        size_t ws_bytes, hipStream_t stream);
// qs[slot] = q[perm[slot]] (nq x 3 float64): queries in launch order, so that
// kernels read them contiguously.
hipError_t gather_queries_launch(const double* q, const int32_t* perm, int64_t nq, double* qs,
                                 hipStream_t stream);

// Device-side gather of V[F] into the 9 x nf FP64 face SoA (wn_forward.hip).
// vdtype: 0 = float64, 1 = float32; fdtype: 0 = int64, 1 = int32.
hipError_t pack_soa_launch(const void* V, const void* F, int64_t nv, int64_t nf,
                           int vdtype, int fdtype, double* tri, int32_t* bad_index,
                           hipStream_t stream);

}  // namespace wn

namespace wn {

// Adjoint (wn_backward.hip). grad_q (nq x 3) and grad_tri (9 x nf) are
// overwritten.
const char* bwd_check(int64_t nf, int64_t nq, int qpl, int splits);
hipError_t bwd_launch(const double* tri, const double* q, const double* gw, double* gq,
                      double* gtri, int64_t nf, int64_t nq, int qpl, int splits,
                      hipStream_t stream);
// grad_V[F[f, c]] += grad_tri corner c of face f; grad_V (nv x 3) overwritten.
hipError_t scatter_corners_launch(const double* gtri, const void* F, int fdtype, int64_t nf,
                                  int64_t nv, double* gv, hipStream_t stream);
// Net directed boundary edges of F (wn_edges.cpp): edges (2 x ne, directed
// vertex ids) and their positive weights (net counts); returns ne.
int64_t boundary_edges(const void* F, int64_t nf, int fdtype, std::vector<int32_t>& edges,
                       std::vector<double>& weight);
// The same edges as strip record indices (interleaved pairs: the first
// record holding each endpoint; idx = vertex id per record), for the
// boundary correction of the strip paths.
int64_t strip_boundary(const void* F, int64_t nf, int fdtype, int64_t nv, const int32_t* idx,
                       int64_t nrec, std::vector<int32_t>& pairs, std::vector<double>& weight);
// Edge adjoint (wn_backward.hip): grad_q (nq x 3) and grad_V (nv x 3, vdtype
// 0 = float64, 1 = float32 V) overwritten from the ne weighted edges.
const char* bwd_edge_check(int64_t ne, int64_t nq, int qpl, int splits);
hipError_t bwd_edge_launch(const void* V, int vdtype, int64_t nv, const int32_t* edges,
                           const double* weight, int64_t ne, const double* q, const double* gw,
                           double* gq, double* gv, int64_t nq, int qpl, int splits,
                           hipStream_t stream);

}  // namespace wn

namespace wn {

// Boundary correction of the FP32-compensated mode (wn_ribbon.hip): adds,
// per block of queries, the angle of the thin quads between each boundary
// edge and its FP32-rounded copy (rounded relative to the block's origin as
// the forward kernels round it) to dst, w += sum / (2 pi) through perm.
//
// Edge source: interleaved id pairs and weights. Records (soa = 0): both ids
// index strip records (4 doubles each) in pts. Face SoA (soa = 1): (face,
// corner) of the 9-row SoA pts with row stride `stride`; the edge runs corner
// c -> c + 1 mod 3. The edge count is *count (device) when count is set,
// else n.
struct RibbonSrc {
  const double* pts = nullptr;
  int64_t stride = 0;
  const int32_t* ids = nullptr;
  const double* w = nullptr;
  const int32_t* count = nullptr;
  int64_t n = 0;
};

// Edge splits of a single-mesh correction (deterministic in its arguments)
// and the partials it needs.
int ribbon_splits(int64_t nq, int qpl, int64_t ne_bound);
size_t ribbon_workspace_bytes(int64_t nq, int splits);
// Single mesh: query blocks of 256 x qpl queries of q (launch order), the
// blocks of the forward kernel that ran before it on the same stream.
hipError_t ribbon_launch(const RibbonSrc& src, int soa, int qpl, const double* q,
                         const int32_t* perm, double* dst, int64_t nq, int splits, void* ws,
                         size_t ws_bytes, hipStream_t s);
// Batch: one block per (mesh, query tile) in blocks; mesh m's edges are
// [e_begin[m], e_end[m]).
struct BatchMesh;
hipError_t ribbon_batch_launch(const RibbonSrc& src, int soa, int qpl, const int32_t* e_begin,
                               const int32_t* e_end, const int2* blocks, int64_t nblocks,
                               const BatchMesh* meshes, const double* q, const int32_t* perm,
                               double* dst, hipStream_t s);

// Boundary edges of a face SoA from its corner coordinates (device):
// interleaved (face, corner) pairs, weights, the count, and per segment
// (faces [face_off[s], face_off[s+1]); face_off null = one segment) the
// range [e_begin[s], e_end[s]) of its edges. Capacity 3 nf edges.
struct SoaBoundary {
  int32_t* ids = nullptr;
  double* w = nullptr;
  int32_t* count = nullptr;
  int32_t* e_begin = nullptr;
  int32_t* e_end = nullptr;
  int64_t cap = 0;
};
size_t soa_boundary_workspace_bytes(int64_t nf, int nseg);
hipError_t soa_boundary_launch(const double* tri, int64_t nf, const int64_t* face_off, int nseg,
                               void* ws, size_t ws_bytes, SoaBoundary* out, hipStream_t s);

}  // namespace wn

namespace wn {

// Batched persistent forward (wn_batch.hip).
//
// Per mesh: its first face (first strip tile for strip plans), first query,
// first partial slot (chunks > 1), and counts. Face units are faces for face
// plans and 255-record tiles for strip plans.
struct BatchMesh {
  int64_t face_off;
  int64_t query_off;
  int64_t pbase;
  int32_t nf;
  int32_t nq;
  int32_t chunks;
  int32_t pad;
};

// One work item: a query tile of a mesh against one face (tile) chunk.
struct BatchItem {
  int32_t mesh;
  int32_t qtile;
  int32_t chunk;
  int32_t pad;
};

// FP64 face records of a 9 x nf SoA (wn_forward.hip), and the persistent
// FP64 face kernel of batch face plans over them (qpl 1 or 2).
hipError_t face_rec64_launch(const double* tri, int64_t nf, FaceRec64* out, hipStream_t stream);
const void* batch_face64_kernel(int qpl);
hipError_t batch_face64_enqueue(const FaceRec64* rec, const double* q, const int32_t* perm,
                                double* dst, double* partial, const BatchMesh* meshes,
                                const BatchItem* items, int64_t n_items, int32_t* counter,
                                int64_t chunk_faces, int qpl, int grid, hipStream_t stream);

// Persistent strip kernel over a work list (wn_strip.hip): one query per
// lane, blocks of kBlockThreads queries, chunks of chunk_tiles tiles.
int batch_strip_resident_blocks();
// The FP64 strip counterpart (wn_strip64.hip) over the plan's sweep records.
int batch_strip64_resident_blocks();
hipError_t batch_strip64_enqueue(const StripRec64* rec64, const double4* rec, const double* q,
                                 const int32_t* perm, double* dst, double* partial,
                                 const BatchMesh* meshes, const BatchItem* items, int64_t n_items,
                                 int32_t* counter, int64_t chunk_tiles, int grid,
                                 hipStream_t stream);
hipError_t batch_strip_enqueue(const double4* rec, const double* q, const int32_t* perm,
                               double* dst, double* partial, const BatchMesh* meshes,
                               const BatchItem* items, int64_t n_items, int32_t* counter,
                               int64_t chunk_tiles, float kappa, int grid, hipStream_t stream);

struct BatchPlan;
const char* batch_check(const int64_t* face_off, const int64_t* query_off, int64_t n_meshes,
                        int precision, int qpl, int64_t chunk_faces);
hipError_t batch_create(const int64_t* face_off, const int64_t* query_off, int64_t n_meshes,
                        int precision, int qpl, int64_t chunk_faces, int sort, float kappa,
                        BatchPlan** out);
hipError_t batch_run(BatchPlan* p, const double* tri, const double* q, double* dst,
                     bool use_graph, hipStream_t stream);
// Strip plans: rec_off in records (each mesh a multiple of 255), chunks of
// chunk_tiles tiles; FP32-compensated or FP64, one query per lane.
const char* batch_strip_check(const int64_t* rec_off, const int64_t* query_off, int64_t n_meshes,
                              int64_t chunk_tiles);
hipError_t batch_create_strips(const int64_t* rec_off, const int64_t* query_off, int64_t n_meshes,
                               int64_t chunk_tiles, int sort, float kappa, int precision,
                               BatchPlan** out);
void batch_destroy(BatchPlan* p);
// Strip plans: the meshes' boundary edges for the FP32-compensated boundary
// correction, as interleaved record pairs (ids, 2 x ne) and weights (w) in
// device memory the caller keeps; mesh m owns edges [edge_off[m],
// edge_off[m+1]) (host). nullptr from the check if the arguments fit.
const char* batch_boundary_check(const BatchPlan* p, int64_t ne, const int64_t* edge_off);
hipError_t batch_set_boundary(BatchPlan* p, const int32_t* ids, const double* w, int64_t ne,
                              const int64_t* edge_off);
// out6 = {items, partial slots, workspace bytes, persistent grid, faces, queries}
void batch_info(const BatchPlan* p, int64_t* out6);

// Morton ordering inside segments [seg_off[s], seg_off[s+1]) (wn_sort.hip):
// each segment stays in place, its queries are reordered along a Z-curve.
size_t morton_segmented_workspace_bytes(int64_t nq);
hipError_t morton_order_segmented(const double* q, int64_t nq, const int64_t* seg_off, int nseg,
                                  int32_t* perm, void* ws, size_t ws_bytes, hipStream_t stream);

}  // namespace wn

namespace wn {

// Point-mesh squared distance on the GPU (wn_aabb.hip). The device copy of a
// tree is one allocation of aabb_device_bytes() filled by aabb_upload(); a
// descent needs at most stack_size <= kAabbStack entries per query.
constexpr int kAabbStack = 40;
size_t aabb_device_bytes(int64_t nnodes, int64_t nf);
hipError_t aabb_upload(const AabbTree& t, const double* V, const void* F, int fdtype, void* dev,
                       hipStream_t stream);
size_t aabb_workspace_bytes(int64_t np, int sort);
// q, sqrd, index (int64), closest: device buffers of np queries.
hipError_t aabb_query_launch(const void* dev, int64_t nnodes, int64_t nf, const double* q,
                             double* sqrd, int64_t* index, double* closest, int64_t np, int sort,
                             void* ws, size_t ws_bytes, hipStream_t stream);

}  // namespace wn

namespace wn {

// Hierarchical winding number on the GPU (wn_fwn.hip): the device copy of a
// tree is one allocation of fwn_device_bytes() filled by fwn_upload().
size_t fwn_device_bytes(int64_t nnodes, int64_t nf);
hipError_t fwn_upload(const FwnTree& t, void* dev, hipStream_t stream);
// q (nq x 3) and w (nq): device buffers; sort: Morton order (workspace).
size_t fwn_workspace_bytes(int64_t nq, int sort);
hipError_t fwn_query_launch(const void* dev, int64_t nnodes, const double* q, double* w,
                            int64_t nq, double beta, int sort, void* ws, size_t ws_bytes,
                            hipStream_t stream);

}  // namespace wn

