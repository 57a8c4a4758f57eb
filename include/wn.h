/*
 * wn.h -- C ABI of the MI355X (gfx950) generalized-winding-number engine,
 * library mojo-libigl_amd/lib/libmojo_libigl_amd.so.
 *
 * Plain pointers and sizes only. Host pointers are marked (host), device
 * (HBM) pointers (device). Every int64_t-returning wn_* call returns 0 (or a
 * non-negative value where stated) on success and a negative status on
 * failure, with a message in wn_last_error():
 *   -1 bad argument, -2 no HIP device, -3 HIP runtime error,
 *   -4 workspace too small, -5 failure injected through WN_FAULT.
 *
 * Face layout ("tri", the reference's): 9 x nf float64 SoA, row k in
 * tri[k*nf : (k+1)*nf], rows ax,ay,az,bx,by,bz,cx,cy,cz
 * (reference python/igl/__init__.py:248-250). Queries: nq x 3 float64,
 * row-major. Output: nq float64 winding numbers.
 *
 * precision: 0 = FP64: a wave whose pairs are all far (|det| < den / 8,
 *                den > |a||b||c|) takes atan(det / den) from a polynomial,
 *                with square roots and the quotient refined to FP64
 *                rounding; other pairs take the reference formula (within
 *                ~1e-15 of precision 2, 2x its speed);
 *            1 = FP32-compensated (FP32 pairs, FP64 accumulation, FP64
 *                re-evaluation of ill-conditioned pairs);
 *            2 = FP64 plain: the reference formula in FP64 for every pair,
 *                with the library's sqrt and atan2 (face-SoA entry points
 *                only: wn_winding_gpu_ex, wn_forward, wn_resolve_config,
 *                wn_forward_workspace_bytes).
 * Tuning arguments (0 or -1 = library default): qpl = queries per lane
 * (1, 2, 4), splits = face splits (split-K, 1..65535), sort = Morton-order
 * the queries (FP32-compensated only; -1 auto), kappa = conditioning
 * threshold of the FP64 re-evaluation (< 0: default 1.0).
 */
#ifndef MOJO_LIBIGL_AMD_WN_H
#define MOJO_LIBIGL_AMD_WN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Reference-compatible entry points --------------------------------- */

/* Replaces mli_winding, src/capi.mojo:938-942 (bound in the reference at
 * python/igl/_lib.py:22 as 7 x c_int64). CPU, FP64, writes dst[q0:q1] only.
 * (host) tri, q, dst. Reentrant; spans are fanned out by the caller. */
void mli_winding(const double *tri, const double *q, double *dst, int64_t nf,
                 int64_t nq, int64_t q0, int64_t q1);

/* Replaces mli_winding_gpu, src/capi.mojo:945-947 (python/igl/_lib.py:23,
 * 5 x c_int64 -> c_int64). GPU, FP64. (host) tri, q, dst. Below 262,144
 * queries the FP64 face kernel (precision 0); from there the FP64 strip
 * kernel (the faces' shared corners found by coordinate equality, the strip
 * mesh kept for a repeated call as wn_winding_gpu_mesh keeps it). Both agree
 * with the plain formula (precision 2) to ~1e-15. Returns 1 on success, 0 on any failure (reason
 * in wn_last_error()), as the reference's try_winding_gpu does
 * (src/capi.mojo:366-398). Thread-safe: the host-buffer entries run on a
 * pool of four contexts (stream, device buffer, kept strip mesh), so calls
 * from several threads overlap. */
int64_t mli_winding_gpu(const double *tri, const double *q, double *dst,
                        int64_t nf, int64_t nq);

/* Replaces mli_aabb_squared_distance, src/capi.mojo:950-959 (bound at
 * python/igl/_lib.py:24 as 14 x c_int64). CPU, FP64: for queries
 * p[q0:q1] (np x 3), the descent of src/capi.mojo:523-587 over the tree
 * arrays boxes (nnodes x 6), nodes (nnodes x 4), order (wn_aabb_build or the
 * reference's python/igl/__init__.py:286-319) writes sqrd, indices (int64)
 * and closest (np x 3). (host) all buffers; stack holds >= stack_size
 * entries and is owned by one caller at a time. */
void mli_aabb_squared_distance(const double *p, const double *v, const int64_t *f,
                               const double *boxes, const int64_t *nodes,
                               const int64_t *order, double *sqrd, int64_t *indices,
                               double *closest, int64_t *stack, int64_t np,
                               int64_t nnodes, int64_t q0, int64_t q1);

/* ---- Status ------------------------------------------------------------- */

/* No reference counterpart (the reference probes nvidia-smi from Python,
 * python/igl/__init__.py:213-229). Number of visible HIP devices (0 if none). */
int64_t wn_device_count(void);

/* Free HBM bytes on the current device, or a negative status. Replaces the
 * memory.free probe of python/igl/__init__.py:215-228. */
int64_t wn_device_free_bytes(void);

/* Message of the calling thread's last failure. Replaces the reference's
 * swallow-all except (src/capi.mojo:397-398), which keeps no reason. */
const char *wn_last_error(void);

/* ABI version (2: adds wn_mesh_forward_noalloc). */
int64_t wn_version(void);

/* No reference counterpart: measurement. op 1 starts timing the dominant
 * kernel of every forward call (face, strip and FP64 strip kernels; HIP
 * events recorded on the launch stream right before and after the kernel;
 * nothing on a stream under graph capture), op 0 stops, op 2 writes out2 =
 * {launches, total milliseconds} of the launches timed since the last op
 * (waiting for their events) and starts a new count. */
int64_t wn_kernel_timer(int64_t op, double *out2);

/* ---- Extended forward --------------------------------------------------- */

/* Extends mli_winding_gpu (src/capi.mojo:945-947) with precision and tuning.
 * (host) tri, q, dst; synchronous. */
int64_t wn_winding_gpu_ex(const double *tri, const double *q, double *dst,
                          int64_t nf, int64_t nq, int64_t precision, int64_t qpl,
                          int64_t splits, int64_t sort, double kappa);

/* Replaces the host SoA pack (python/igl/__init__.py:248-250) followed by
 * mli_winding_gpu (src/capi.mojo:945-947): (host) V (nv x 3 float64) and F
 * (nf x 3; fdtype 0 = int64, 1 = int32) are uploaded as they are and gathered
 * into the face SoA on the device, then the forward of wn_winding_gpu_ex runs
 * with its default tuning. Same results as packing on the host, bit for bit.
 * A corner outside [0, nv) fails with -1 before anything is uploaded (the
 * reference raises IndexError in Python, python/igl/__init__.py:119-124).
 * (host) q, dst; synchronous; 0 or a negative status. */
int64_t wn_winding_gpu_vf(const double *V, const void *F, int64_t nv, int64_t nf,
                          int64_t fdtype, const double *q, double *dst, int64_t nq,
                          int64_t precision);

/* Resolved launch configuration for the arguments: out4 = {qpl, splits, sort,
 * precision}. */
int64_t wn_resolve_config(int64_t nf, int64_t nq, int64_t precision, int64_t qpl,
                          int64_t splits, int64_t sort, double kappa, int64_t *out4);

/* Bytes of device workspace wn_forward needs (>= 0), or a negative status. */
int64_t wn_forward_workspace_bytes(int64_t nf, int64_t nq, int64_t precision,
                                   int64_t qpl, int64_t splits, int64_t sort);

/* Device-resident forward pass, the path of the reference's GPU kernels
 * winding_angles_gpu + winding_reduce_gpu (src/capi.mojo:297-363) without the
 * nf x nq angle buffer. (device) tri, q, dst, workspace. Asynchronous on
 * `stream` (a hipStream_t; NULL = default stream). Allocates nothing. */
int64_t wn_forward(const double *tri, const double *q, double *dst, int64_t nf,
                   int64_t nq, int64_t precision, int64_t qpl, int64_t splits,
                   int64_t sort, double kappa, void *workspace,
                   int64_t workspace_bytes, void *stream);

/* Device-side ingest replacing the host SoA pack of
 * python/igl/__init__.py:248-250: tri = V[F] as 9 x nf float64. (device) V
 * (nv x 3; vdtype 0 = float64, 1 = float32), F (nf x 3; fdtype 0 = int64,
 * 1 = int32), tri, bad_index (optional int32 the caller initialises to
 * INT32_MAX; receives the smallest face index with a corner outside [0, nv)).
 * Asynchronous on `stream`. */
int64_t wn_pack_soa(const void *V, const void *F, int64_t nv, int64_t nf,
                    int64_t vdtype, int64_t fdtype, double *tri,
                    int32_t *bad_index, void *stream);

/* ---- Device-resident mesh (triangle strips) ------------------------------ */

/* No reference counterpart (the reference re-packs and re-uploads the mesh on
 * every call, python/igl/__init__.py:248-250, src/capi.mojo:371-376). Turns
 * (host) V (nv x 3 float64) and F (nf x 3; fdtype 0 = int64, 1 = int32) into
 * triangle-strip records, 4 doubles each (x, y, z, flag), in tiles of 255 that
 * each start a strip; flag 0 = no face ends at the record, +1 / -1 = the
 * record closes face (w[k-2], w[k-1], w[k]) with that orientation relative to
 * the face's own corner order. Writes them to (host) out when capacity (in
 * records) suffices; returns the record count, or a negative status. */
int64_t wn_mesh_records(const double *V, const void *F, int64_t nv, int64_t nf,
                        int64_t fdtype, double *out, int64_t capacity);

/* Builds the strip records of (host) V, F and uploads them with the strip
 * topology (vertex id and flag per record; the only allocation of the mesh
 * path). nv < 2^31. Returns a handle (> 0) or a negative status (-1 also for
 * a face index outside [0, nv)). */
int64_t wn_mesh_create(const double *V, const void *F, int64_t nv, int64_t nf,
                       int64_t fdtype);

/* No reference counterpart. A mesh from its faces alone (host F, nv
 * vertices): the strip topology is built and uploaded, the records are set
 * by wn_mesh_set_vertices. Returns a handle (> 0) or a negative status. */
int64_t wn_mesh_create_topology(const void *F, int64_t nv, int64_t nf, int64_t fdtype);

/* No reference counterpart. Re-gathers a mesh's records from (device) V
 * (nv x 3, vdtype 0 = float64, 1 = float32; nv as at creation) on `stream`:
 * new vertex positions without re-stripifying or host copies. Asynchronous;
 * allocates nothing. */
int64_t wn_mesh_set_vertices(int64_t mesh, const void *V, int64_t nv, int64_t vdtype,
                             void *stream);

/* out3 = {faces, records, tiles}. */
int64_t wn_mesh_info(int64_t mesh, int64_t *out3);

/* Resolved {splits, sort} for nq queries (splits 0 / sort -1 = default). */
int64_t wn_mesh_resolve(int64_t mesh, int64_t nq, int64_t splits, int64_t sort,
                        int64_t *out2);

/* Device workspace bytes of wn_mesh_forward. */
int64_t wn_mesh_workspace_bytes(int64_t mesh, int64_t nq, int64_t splits,
                                int64_t sort);

/* FP32-compensated forward over the mesh's strips for (device) q (nq x 3
 * float64) into (device) dst (nq). Same arithmetic and accuracy contract as
 * wn_forward precision 1; one vertex offset, length and square root per face
 * instead of three. Asynchronous on `stream`; allocates nothing. */
int64_t wn_mesh_forward(int64_t mesh, const double *q, double *dst, int64_t nq,
                        int64_t splits, int64_t sort, double kappa,
                        void *workspace, int64_t workspace_bytes, void *stream);

/* No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). wn_mesh_forward without the workspace argument: its scratch
 * (split partials, Morton permutation, queries in launch order, sort
 * temporaries) lives in a static 64 MiB device array that each GPU holds
 * until the process ends. Calls nothing that allocates. splits 0 picks the
 * default split count, lowered until the scratch fits; a configuration that
 * needs more than 64 MiB (explicit splits, or more than ~1.9M queries with
 * Morton order) fails with -4 and a message. Same results as wn_mesh_forward
 * with the same splits and sort, bit for bit. Every call on one GPU shares
 * that GPU's scratch: calls that overlap on one GPU (other streams, other
 * threads) race and must be ordered by the caller. Asynchronous on
 * `stream`. */
int64_t wn_mesh_forward_noalloc(int64_t mesh, const double *q, double *dst, int64_t nq,
                                int64_t splits, int64_t sort, double kappa, void *stream);

/* The same query in FP64 (no reference counterpart beyond
 * src/capi.mojo:218-235,286 itself): FP64 over the mesh's strip records, one
 * square root per face, far pairs from a polynomial accurate to FP64
 * rounding, every other pair with the plain FP64 formula; agrees with the
 * plain FP64 kernel to ~1e-15. Workspace as wn_mesh_workspace_bytes. */
int64_t wn_mesh_forward64(int64_t mesh, const double *q, double *dst, int64_t nq,
                          int64_t splits, int64_t sort, void *workspace,
                          int64_t workspace_bytes, void *stream);

/* Frees a mesh (synchronises the device first). */
int64_t wn_mesh_destroy(int64_t mesh);

/* Host-buffer strip path (no reference counterpart; the reference's GPU entry
 * is mli_winding_gpu above): host V (nv x 3 float64) and F (nf x 3, fdtype 0
 * = int64, 1 = int32) are stripified and uploaded, (host) q (nq x 3) runs
 * through the FP64 (precision 0) or FP32-compensated (1) strip kernel into
 * (host) dst (nq). Synchronous; returns 0 or a negative status. The strip
 * mesh of the last call stays on the device with host copies of its V and F:
 * a call with byte-identical F reuses its topology (no stripifier, no
 * upload), and with byte-identical V too its records; results are those of a
 * fresh mesh, bit for bit. wn_release_host_buffers frees it. */
int64_t wn_winding_gpu_mesh(const double *V, const void *F, int64_t nv, int64_t nf,
                            int64_t fdtype, const double *q, double *dst, int64_t nq,
                            int64_t precision);

/* No reference counterpart (the reference frees everything per call,
 * src/capi.mojo:366-398). Frees the device memory the host-buffer entries
 * keep between calls (mli_winding_gpu, wn_winding_gpu_ex, wn_winding_gpu_vf:
 * the grow-only buffer; wn_winding_gpu_mesh: also the kept strip mesh).
 * The next call allocates again. Returns 0. */
int64_t wn_release_host_buffers(void);

/* ---- Batched forward ------------------------------------------------------ */

/* No reference counterpart (the reference evaluates one mesh per call,
 * src/capi.mojo:366-398). Plans a batch of n_meshes meshes: all faces in one
 * 9 x F_total tri SoA, mesh m owning faces [face_off[m], face_off[m+1]) and
 * queries [query_off[m], query_off[m+1]) of one Q_total x 3 array. (host)
 * face_off, query_off (n_meshes + 1 entries each, starting at 0). Builds the
 * longest-first work list of (mesh, query tile, face chunk) items, uploads it
 * and allocates the plan's device workspace (the only allocation of the
 * batch path). precision 0/1; qpl 1 or 2 (0 = 1); chunk_faces a multiple of
 * 256 (0 = 2048); sort -1 = Morton order for FP32-compensated. Returns a
 * handle (> 0) or a negative status. */
int64_t wn_batch_create(const int64_t *face_off, const int64_t *query_off,
                        int64_t n_meshes, int64_t precision, int64_t qpl,
                        int64_t chunk_faces, int64_t sort, double kappa);

/* No reference counterpart. A batch plan over meshes in triangle-strip
 * record form (wn_mesh_records per mesh, concatenated): mesh m owns records
 * [rec_off[m], rec_off[m+1]) (a multiple of 255 each) and queries
 * [query_off[m], query_off[m+1]). FP32-compensated, one query per lane;
 * chunk_tiles = 255-record tiles per work item (0 = 8); sort -1 = Morton
 * order. Run with wn_batch_run, passing the records (4 doubles each) as
 * `tri`. Returns a handle (> 0) or a negative status. */
int64_t wn_batch_create_strips(const int64_t *rec_off, const int64_t *query_off,
                               int64_t n_meshes, int64_t chunk_tiles, int64_t sort,
                               double kappa);

/* The same strip plan in FP64 (the FP64 strip kernel of wn_mesh_forward64):
 * allocates 64 bytes of sweep records per record, rebuilt from the records
 * by every wn_batch_run. No kappa. */
int64_t wn_batch_create_strips64(const int64_t *rec_off, const int64_t *query_off,
                                 int64_t n_meshes, int64_t chunk_tiles, int64_t sort);

/* Runs a plan: (device) tri (9 x F_total; the records of a strip plan),
 * q (Q_total x 3), dst (Q_total).
 * use_graph != 0 replays a hipGraph of the persistent launch, captured on
 * first use for these buffer addresses. Asynchronous on `stream`. Calls on
 * one handle must not overlap. */
int64_t wn_batch_run(int64_t handle, const double *tri, const double *q,
                     double *dst, int64_t use_graph, void *stream);

/* No reference counterpart. Boundary edges of a strip plan's meshes for the
 * FP32-compensated boundary correction (an open mesh's FP32-rounded boundary
 * differs from the true one by thin quads; the correction adds their angle,
 * see wn_mesh_forward): (device) pairs (2 x ne int32, interleaved record
 * indices into the plan's records, as wn_strip_boundary gives them per mesh,
 * offset by the mesh's first record) and weight (ne float64), both kept by
 * the caller while the plan runs; (host) edge_off (n_meshes + 1, from 0 to
 * ne): mesh m owns edges [edge_off[m], edge_off[m+1]). Without it a plan runs
 * no correction (open meshes then carry the FP32 boundary rounding, ~1e-6).
 * FP64 strip plans accept and ignore it; face plans find their boundary
 * edges themselves. Drops a captured graph. */
int64_t wn_batch_set_boundary(int64_t handle, const int32_t *pairs, const double *weight,
                              int64_t ne, const int64_t *edge_off);

/* out6 = {work items, partial slots, workspace bytes, persistent grid,
 * F_total (tiles for a strip plan), Q_total}. */
int64_t wn_batch_info(int64_t handle, int64_t *out6);

/* Frees a plan (synchronises the device first). */
int64_t wn_batch_destroy(int64_t handle);

/* No reference counterpart. The strip topology of (host) F (nv vertices,
 * nv < 2^31): per record the vertex id and the flag (0, +1, -1) of
 * wn_mesh_records' format. Fills (host) idx and flag when capacity suffices;
 * returns the record count or a negative status. */
int64_t wn_strip_topology(const void *F, int64_t nv, int64_t nf, int64_t fdtype,
                          int32_t *idx, int8_t *flag, int64_t capacity);

/* No reference counterpart. Net directed boundary edges of (host) F (as
 * wn_boundary_edges) expressed as strip record indices: per edge the first
 * record of (host) idx (the strip topology of F from wn_strip_topology, nrec
 * records) holding each endpoint, as interleaved pairs (2 x count int32), and
 * the weights (count float64); fills them when capacity suffices. Returns the
 * count or a negative status. */
int64_t wn_strip_boundary(const void *F, int64_t nv, int64_t nf, int64_t fdtype,
                          const int32_t *idx, int64_t nrec, int32_t *pairs, double *weight,
                          int64_t capacity);

/* No reference counterpart. Device gather of strip records: rec[k] =
 * (V[idx[k]], flag[k]) for (device) V (nv x 3, vdtype 0 = float64,
 * 1 = float32), idx, flag and rec (nrec x 4 float64). Ids outside [0, nv)
 * leave their record untouched. Asynchronous on `stream`; allocates
 * nothing. */
int64_t wn_strip_gather(const void *V, int64_t nv, int64_t vdtype, const int32_t *idx,
                        const int8_t *flag, int64_t nrec, double *rec, void *stream);

/* ---- Point-mesh squared distance -------------------------------------- */

/* Replaces the reference's NumPy tree build, python/igl/__init__.py:286-319
 * (AABB.init): the median-split tree over V (nv x 3 f64, host) and F (nf x 3,
 * fdtype 0 = int64, 1 = int32), leaves of <= leaf_size faces (0 = 8). Returns
 * the node count; when boxes (count x 6), nodes (count x 4), order (nf) are
 * all given and capacity >= count, fills them and *stack_size. Negative
 * status on bad arguments (IndexError-class: a face index outside V). */
int64_t wn_aabb_build(const double *V, const void *F, int64_t nv, int64_t nf,
                      int64_t fdtype, int64_t leaf_size, double *boxes, int64_t *nodes,
                      int64_t *order, int64_t capacity, int64_t *stack_size);

/* No reference counterpart (the reference has no GPU distance query). Builds
 * the same tree and keeps it on the current device (nodes, face corners in
 * leaf order, face ids). Returns a handle (> 0) or a negative status. */
int64_t wn_aabb_create(const double *V, const void *F, int64_t nv, int64_t nf,
                       int64_t fdtype, int64_t leaf_size);

/* out4 = {faces, nodes, stack size, device bytes}. */
int64_t wn_aabb_info(int64_t handle, int64_t *out4);

/* Workspace bytes for wn_aabb_query. sort 1 = walk the points in Morton
 * order (a Z-curve sort first); 0 or -1 (default) = in the given order. */
int64_t wn_aabb_workspace_bytes(int64_t handle, int64_t np, int64_t sort);

/* GPU counterpart of mli_aabb_squared_distance over a wn_aabb_create tree:
 * (device) p (np x 3 f64), sqrd (np), indices (np int64), closest (np x 3);
 * every output equals the CPU path's bit for bit (same descent, same
 * arithmetic, FMA contraction off). A tree without faces gives +inf, -1, 0.
 * Asynchronous on `stream`; allocates nothing. */
int64_t wn_aabb_query(int64_t handle, const double *p, double *sqrd, int64_t *indices,
                      double *closest, int64_t np, int64_t sort, void *workspace,
                      int64_t workspace_bytes, void *stream);

/* Frees the device tree (synchronises the device first). */
int64_t wn_aabb_destroy(int64_t handle);

/* ---- Hierarchical ("fast") winding number ------------------------------- */

/* No reference counterpart (README.md:29 lists it as not covered). Barill et
 * al. 2018: per node of the median-split tree (the wn_aabb_build tree), the
 * second-order expansion of its faces' solid angle about their area-weighted
 * centroid; a query uses a node's expansion when farther than beta x the
 * node's radius, else recurses, and sums leaf faces exactly
 * (src/capi.mojo:218-235). V (nv x 3 f64), F (nf x 3, fdtype 0 = int64,
 * 1 = int32) on the host; leaf_size 0 = 8; upload != 0 also copies the tree
 * to the current device. Returns a handle (> 0) or a negative status. */
int64_t wn_fwn_create(const double *V, const void *F, int64_t nv, int64_t nf,
                      int64_t fdtype, int64_t leaf_size, int64_t upload);

/* out4 = {faces, nodes, stack size, device bytes (0 without a device copy)}. */
int64_t wn_fwn_info(int64_t handle, int64_t *out4);

/* Per node: expansions (nodes x 34 doubles: centre xyz, radius, N (3),
 * C (3 x 3), D (3 x 6 symmetric blocks 00 01 02 11 12 22)) and links
 * (nodes x 4: left, right, start, count). Returns the node count; fills the
 * outputs when both are given and capacity >= it. (host) */
int64_t wn_fwn_export(int64_t handle, double *expansions, int64_t *links,
                      int64_t capacity);

/* CPU query: w[q0:q1] for q (nq x 3), accuracy parameter beta > 0 (2 is the
 * published default; larger is more accurate and slower). (host) */
int64_t wn_fwn_query_cpu(int64_t handle, const double *q, double *w, int64_t nq,
                         double beta, int64_t q0, int64_t q1);

/* Workspace bytes for wn_fwn_query (sort -1 = Morton order when nq >= 4096,
 * 1 = always, 0 = never). */
int64_t wn_fwn_workspace_bytes(int64_t handle, int64_t nq, int64_t sort);

/* GPU query over the device copy: (device) q (nq x 3), w (nq), workspace.
 * Same walk and expansion arithmetic as the CPU query (exact faces agree to
 * rounding); with sort, queries are walked in Morton order and w is written
 * back in the given order. Asynchronous on `stream`; allocates nothing. */
int64_t wn_fwn_query(int64_t handle, const double *q, double *w, int64_t nq,
                     double beta, int64_t sort, void *workspace,
                     int64_t workspace_bytes, void *stream);

/* Frees the tree (synchronises the device first when it has a device copy). */
int64_t wn_fwn_destroy(int64_t handle);

/* ---- Adjoint ------------------------------------------------------------ */

/* No reference counterpart (the reference has no derivative of
 * src/capi.mojo:218-235,286). Given grad_w = dL/dw (nq), writes
 * grad_q = dL/dO (nq x 3) and grad_tri = dL/d(face corners) (9 x nf, the tri
 * layout). FP64 (precision must be 0). (device) all buffers; asynchronous on
 * `stream`; allocates nothing. grad_tri is accumulated with FP64 atomics, so
 * its rounding depends on scheduling; grad_q is exact-order when splits == 1.
 * qpl (queries per lane) 1, 2 or 4, 0 = default (4); splits 0 = default. */
int64_t wn_backward(const double *tri, const double *q, const double *grad_w,
                    double *grad_q, double *grad_tri, int64_t nf, int64_t nq,
                    int64_t precision, int64_t qpl, int64_t splits, void *stream);

/* No reference counterpart. Net directed boundary edges of (host) F over nv
 * vertices (nv < 2^31): an edge {u, v} used n times as u->v and m times as
 * v->u with n != m is one edge, directed u->v with weight n - m when n > m
 * and v->u with weight m - n otherwise; a closed, consistently oriented mesh
 * has none. Fills (host) edges (2 x count int32, source then target) and
 * weight (count float64) when capacity suffices; returns the count or a
 * negative status. */
int64_t wn_boundary_edges(const void *F, int64_t nv, int64_t nf, int64_t fdtype,
                          int32_t *edges, double *weight, int64_t capacity);

/* No reference counterpart. dL/dO and dL/dV of the winding number from the
 * boundary edges alone: per face the gradient of atan2(det, den) is a sum
 * over its three directed edges, whose terms cancel on edges shared by
 * oppositely directed faces, so this equals wn_backward followed by
 * wn_scatter_corners (to rounding). (device) V (nv x 3; vdtype 0 = float64,
 * 1 = float32), edges/weight from wn_boundary_edges, q, grad_w (nq);
 * grad_q (nq x 3) and grad_V (nv x 3) overwritten, in FP64. Edge ids outside
 * [0, nv) contribute nothing. Asynchronous on `stream`; allocates nothing.
 * qpl 1, 2 or 4 (0 = 4); splits 0 = default. grad_V is accumulated with FP64
 * atomics; grad_q is exact-order when one split runs. */
int64_t wn_backward_edges(const void *V, int64_t nv, int64_t vdtype, const int32_t *edges,
                          const double *weight, int64_t ne, const double *q,
                          const double *grad_w, double *grad_q, double *grad_V, int64_t nq,
                          int64_t qpl, int64_t splits, void *stream);

/* grad_V (nv x 3, overwritten) = per-vertex sum of the corner gradients of
 * grad_tri through F (nf x 3; fdtype 0 = int64, 1 = int32). (device);
 * asynchronous on `stream`. */
int64_t wn_scatter_corners(const double *grad_tri, const void *F, int64_t fdtype,
                           int64_t nf, int64_t nv, double *grad_V, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* MOJO_LIBIGL_AMD_WN_H */
