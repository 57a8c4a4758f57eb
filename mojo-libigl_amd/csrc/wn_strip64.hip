// FP64 forward over triangle-strip records: the winding number of
// src/capi.mojo:218-235,286 in FP64 at the strip kernel's per-face cost.
//
// Per face the lane computes one vertex offset, one squared length and one
// square root (the other two vertices are kept from the previous records),
// the dot products from the law of cosines with the halved squared edge
// lengths of the sweep record, and det = N . c with the face normal of the
// sweep record (exact: N . (a - c) = 0). Records go in groups of three
// whose angles are taken together as the argument of the product of their
// faces' (den + i det) (sum3_f64; an 11-term Remez polynomial of the
// product's tan, relative error 1.6e-17 on |x| <= 0.4) when every lane
// meets its conditions; otherwise a wave whose 64 pairs are all far
// (|det| < den / 8 and den > la lb lc, the FP32 strip kernel's test) takes
// each angle from a 7-term polynomial of det / den (1.8e-18 on |x| <= 1/8,
// below FP64 rounding; gen_atan_coeffs.py), and any other wave evaluates
// every lane's pair with the plain FP64 formula (pair_f64, the plain
// kernel's arithmetic). Square roots and the quotient
// are refined from the hardware estimates to FP64 rounding (Goldschmidt for
// the root, one Newton step and a remainder correction for the quotient).
//
// The sweep record is read at a wave-uniform address, so it arrives in
// scalar registers through scalar loads: no LDS, no tile staging, no
// barriers -- every wave streams the records on its own. Queries are used as
// given (no block-local origin: FP64 offsets V - O are the plain kernel's
// own). Splits, Morton order and the split reduction are strip_launch's.

#include "wn_device.h"
#include "wn_f64.h"
#include "wn_internal.h"

namespace wn {

namespace {

// Builds the sweep records of one tile-aligned record array. Neighbours are
// taken inside the record's tile (the first two records of a tile close no
// face: strips are cut at tile ends), as the FP32 staging does.
__global__ __launch_bounds__(kBlockThreads) void strip64_prep_kernel(
    const double4* __restrict__ rec, int64_t nrec, StripRec64* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * kBlockThreads + threadIdx.x;
  if (k >= nrec) return;
  const int64_t j = k % kStripTileRecords;
  const int64_t base = k - j;
  const double4 c = rec[k];
  const double4 p = rec[base + (j >= 1 ? j - 1 : j)];
  const double4 pp = rec[base + (j >= 2 ? j - 2 : j)];
  const double d1x = c.x - p.x, d1y = c.y - p.y, d1z = c.z - p.z;
  const double d2x = c.x - pp.x, d2y = c.y - pp.y, d2z = c.z - pp.z;
  const double ex = p.x - pp.x, ey = p.y - pp.y, ez = p.z - pp.z;
  const double f = c.w;  // orientation flag (0 = no face)
  StripRec64 r;
  r.x = c.x;
  r.y = c.y;
  r.z = c.z;
  r.e1 = -0.5 * (d1x * d1x + d1y * d1y + d1z * d1z);
  r.e2 = -0.5 * (d2x * d2x + d2y * d2y + d2z * d2z);
  r.nx = f * (ey * d2z - ez * d2y);
  r.ny = f * (ez * d2x - ex * d2z);
  r.nz = f * (ex * d2y - ey * d2x);
  out[k] = r;
}

struct Slot64 {
  double ax, ay, az, l2, l;
};

// The plain FP64 pair, oriented; out of line so that its registers are not
// reserved for the whole sweep.
__device__ __noinline__ double strip64_pair(double ax, double ay, double az, double bx,
                                            double by, double bz, double cx, double cy,
                                            double cz, bool flip) {
  const double v = pair_f64(ax, ay, az, bx, by, bz, cx, cy, cz);
  return flip ? -v : v;
}

// Geometry of record j: R takes its vertex, and for face (P, Q, R) det, den
// and dp = den - la lb lc (meaningful where the record closes a
// face; a record that closes none, flag 0, has N = 0 and is never summed).
// s_carry is the dot product of the previous record's two newest vertices
// (Q, R of the last record = P, Q here). Dot products come from the law of
// cosines, b.c = (|b|^2 + |c|^2) / 2 - |b - c|^2 / 2, with -|b - c|^2 / 2
// stored in the record: one fma, and den = pq lc + dp needs no doubling.
struct Geom64 {
  double det, den, dp;
};

__device__ __forceinline__ Geom64 strip64_geom(const StripRec64& r, const Slot64& P,
                                               const Slot64& Q, Slot64& R, double& s_carry,
                                               double ox, double oy, double oz) {
  R.ax = r.x - ox;
  R.ay = r.y - oy;
  R.az = r.z - oz;
  R.l2 = __builtin_fma(R.az, R.az, __builtin_fma(R.ay, R.ay, R.ax * R.ax));
  R.l = sqrt_gs_f64(R.l2);
  const double s_qr = half_sum_plus(Q.l2, R.l2, r.e1);  // b.c
  const double s_pr = half_sum_plus(P.l2, R.l2, r.e2);  // c.a
  Geom64 g;
  g.det = __builtin_fma(r.nz, R.az, __builtin_fma(r.ny, R.ay, r.nx * R.ax));
  const double pq = P.l * Q.l;
  g.dp = __builtin_fma(R.l, s_carry, __builtin_fma(P.l, s_qr, Q.l * s_pr));
  g.den = __builtin_fma(pq, R.l, g.dp);
  s_carry = s_qr;
  return g;
}

// Face j (closed by a record with flag_bits != 0) on its own: far on every
// lane (|det / den| < 1/8, den > la lb lc), the 7-term polynomial into acc;
// otherwise the plain formula into acc_slow (two sums, so that each is only
// ever updated in place), from the corners' offsets rebuilt from records
// j-2, j-1, j of the tile (the values the register slots held).
__device__ __forceinline__ void strip64_face(const StripRec64* __restrict__ r, int j,
                                             long long flag_bits, const Geom64& g, double& acc,
                                             double& acc_slow, double ox, double oy, double oz) {
  const double det8 = 8.0 * g.det;  // exact
  if (__builtin_expect(wave_all_far_x8_f64(det8, g.den, g.dp), 1)) {
    acc += atan_small_x8_f64(div_f64(det8, g.den));
  } else {
    asm volatile("" ::);  // keep a real wave-uniform branch
    const StripRec64 &a = r[j - 2], &b = r[j - 1], &c = r[j];
    // flag -1.0 has the sign bit set: the face runs against the strip order
    acc_slow += strip64_pair(a.x - ox, a.y - oy, a.z - oz, b.x - ox, b.y - oy, b.z - oz,
                             c.x - ox, c.y - oy, c.z - oz, flag_bits < 0);
  }
}

// Geometry of the face closed by record j (j >= 2 in its tile) rebuilt from
// records j-2, j-1, j with strip64_geom's operations: the values the sweep
// formed for it.
__device__ __forceinline__ Geom64 strip64_face_geom(const StripRec64* __restrict__ r, int j,
                                                    double ox, double oy, double oz) {
  Slot64 P, Q, R;
  double s_carry = 0.0;
  P.ax = r[j - 2].x - ox;
  P.ay = r[j - 2].y - oy;
  P.az = r[j - 2].z - oz;
  P.l2 = __builtin_fma(P.az, P.az, __builtin_fma(P.ay, P.ay, P.ax * P.ax));
  P.l = sqrt_gs_f64(P.l2);
  Q.ax = r[j - 1].x - ox;
  Q.ay = r[j - 1].y - oy;
  Q.az = r[j - 1].z - oz;
  Q.l2 = __builtin_fma(Q.az, Q.az, __builtin_fma(Q.ay, Q.ay, Q.ax * Q.ax));
  Q.l = sqrt_gs_f64(Q.l2);
  s_carry = half_sum_plus(P.l2, Q.l2, r[j - 1].e1);  // a.b, as record j-1 formed it
  return strip64_geom(r[j], P, Q, R, s_carry, ox, oy, oz);
}

// Records j, j+1, j+2 whose product t and smallest dp the sweep formed: the
// group's angle when all three close a face and every lane meets the group
// test, else face by face (strip64_face, geometry rebuilt from the records).
// Out of line (the fallback of a failed group of six): its registers are not
// reserved for the sweep. Returns (fast-path sum, plain-formula sum).
__device__ __noinline__ double2 strip64_triple_of(const StripRec64* __restrict__ r, int j,
                                                  long long f1, long long f2, long long f3,
                                                  double t_re, double t_im, double dp_min,
                                                  double ox, double oy, double oz) {
  double acc = 0.0, acc_slow = 0.0;
  if (f1 != 0 && f2 != 0 && f3 != 0) {
    const double x = div_f64(t_im, t_re);  // used only where t_re is positive normal
    if (wave_all_group_f64(dp_min, t_re, t_re, x)) return make_double2(atan_sum_f64(x), 0.0);
  }
  const long long f[3] = {f1, f2, f3};
  for (int i = 0; i < 3; ++i)
    if (f[i] != 0)
      strip64_face(r, j + i, f[i], strip64_face_geom(r, j + i, ox, oy, oz), acc, acc_slow, ox,
                   oy, oz);
  return make_double2(acc, acc_slow);
}

// Unscaled FP64 angle sum of one query over strip tiles [t_begin, t_end).
// Records go in groups of six (two turns of the three register slots; a
// tile's 255 = 42 x 6 + 3). When all six close a face (flags: scalar) and
// every lane meets the group test, the group's angle is the argument of the
// product of its faces' (den + i det): one quotient and one 11-term
// polynomial for the six; otherwise as two groups of three
// (strip64_triple_of), and a triple that fails face by face. The last three
// records of a tile go as a group of three.
__device__ __forceinline__ double strip64_sweep(const StripRec64* __restrict__ rec64,
                                                const double4* __restrict__ rec, int64_t t_begin,
                                                int64_t t_end, double ox, double oy, double oz) {
  double acc = 0.0, acc_slow = 0.0, s_carry = 0.0;
  Slot64 A{0.0, 0.0, 0.0, 0.0, 0.0}, B = A, C = A;
  for (int64_t t = t_begin; t < t_end; ++t) {
    const StripRec64* r = rec64 + t * kStripTileRecords;
    const long long* flags = reinterpret_cast<const long long*>(rec + t * kStripTileRecords) + 3;
    int j = 0;
    for (; j + 6 <= kStripTileRecords; j += 6) {
      const long long f1 = flags[4 * j], f2 = flags[4 * (j + 1)], f3 = flags[4 * (j + 2)];
      const long long f4 = flags[4 * (j + 3)], f5 = flags[4 * (j + 4)], f6 = flags[4 * (j + 5)];
      const Geom64 g1 = strip64_geom(r[j], A, B, C, s_carry, ox, oy, oz);
      const Geom64 g2 = strip64_geom(r[j + 1], B, C, A, s_carry, ox, oy, oz);
      const Geom64 g3 = strip64_geom(r[j + 2], C, A, B, s_carry, ox, oy, oz);
      const Prod3F64 t1 = prod3_f64(g1.det, g1.den, g2.det, g2.den, g3.det, g3.den);
      const double m1 = fmin(fmin(g1.dp, g2.dp), g3.dp);
      const Geom64 g4 = strip64_geom(r[j + 3], A, B, C, s_carry, ox, oy, oz);
      const Geom64 g5 = strip64_geom(r[j + 4], B, C, A, s_carry, ox, oy, oz);
      const Geom64 g6 = strip64_geom(r[j + 5], C, A, B, s_carry, ox, oy, oz);
      const Prod3F64 t2 = prod3_f64(g4.det, g4.den, g5.det, g5.den, g6.det, g6.den);
      const double m2 = fmin(fmin(g4.dp, g5.dp), g6.dp);
      if (__builtin_expect(f1 != 0 && f2 != 0 && f3 != 0 && f4 != 0 && f5 != 0 && f6 != 0, 1)) {
        const double re = __builtin_fma(-t1.im, t2.im, t1.re * t2.re);
        const double im = __builtin_fma(t1.re, t2.im, t1.im * t2.re);
        const double x = div_f64(im, re);  // used only where re is positive normal
        if (__builtin_expect(wave_all_group_f64(fmin(m1, m2), fmin(t1.re, t2.re), re, x), 1)) {
          acc += atan_sum_f64(x);
          continue;
        }
      }
      asm volatile("" ::);  // keep a real wave-uniform branch
      const double2 a = strip64_triple_of(r, j, f1, f2, f3, t1.re, t1.im, m1, ox, oy, oz);
      const double2 b = strip64_triple_of(r, j + 3, f4, f5, f6, t2.re, t2.im, m2, ox, oy, oz);
      acc += a.x;
      acc += b.x;
      acc_slow += a.y;
      acc_slow += b.y;
    }
    // the tile's last three records
    const long long f1 = flags[4 * j], f2 = flags[4 * (j + 1)], f3 = flags[4 * (j + 2)];
    const Geom64 g1 = strip64_geom(r[j], A, B, C, s_carry, ox, oy, oz);
    const Geom64 g2 = strip64_geom(r[j + 1], B, C, A, s_carry, ox, oy, oz);
    const Geom64 g3 = strip64_geom(r[j + 2], C, A, B, s_carry, ox, oy, oz);
    const Prod3F64 tl = prod3_f64(g1.det, g1.den, g2.det, g2.den, g3.det, g3.den);
    const double2 a = strip64_triple_of(r, j, f1, f2, f3, tl.re, tl.im,
                                        fmin(fmin(g1.dp, g2.dp), g3.dp), ox, oy, oz);
    acc += a.x;
    acc_slow += a.y;
  }
  return acc + acc_slow;
}

// 5 waves per SIMD: 96 VGPRs, no spills on the fast path (the slow paths
// are calls, around which registers are saved; at 6 waves and 80 VGPRs the
// groups of six spilled in the loop, 6.5 % slower)
__global__ __launch_bounds__(kBlockThreads) __attribute__((amdgpu_waves_per_eu(5))) void fwd_strip64_kernel(
    const StripRec64* __restrict__ rec64, const double4* __restrict__ rec, int64_t ntiles,
    const double* __restrict__ q, const int32_t* __restrict__ perm, double* __restrict__ out,
    int64_t nq, int64_t tiles_per_split, int splits) {
  const int64_t slot = (int64_t)blockIdx.x * kBlockThreads + threadIdx.x;
  const bool valid = slot < nq;
  // q is in launch order (gathered through perm beforehand when sorted)
  const double ox = valid ? q[3 * slot] : 0.0;
  const double oy = valid ? q[3 * slot + 1] : 0.0;
  const double oz = valid ? q[3 * slot + 2] : 0.0;
  const int64_t t_begin = (int64_t)blockIdx.y * tiles_per_split;
  const int64_t t_end = t_begin + tiles_per_split < ntiles ? t_begin + tiles_per_split : ntiles;
  const double acc = strip64_sweep(rec64, rec, t_begin, t_end, ox, oy, oz);
  // the output slot rebuilt after the sweep (kept live across it, it was
  // spilled to scratch at every block start)
  const int64_t s = (int64_t)blockIdx.x * kBlockThreads + thread_index(wave_base());
  if (s < nq) {
    if (splits == 1)
      out[perm ? (int64_t)perm[s] : s] = acc * kInv2Pi;
    else
      out[(int64_t)blockIdx.y * nq + s] = acc;
  }
}

// Persistent FP64 strip kernel over a batch plan's work list (mesh, query
// tile, tile chunk), with the FP32 strip batch kernel's item protocol
// (atomic counter, block-uniform exit) and output rule (single-chunk meshes
// write w, others unscaled chunk-major partials summed in chunk order).
__global__ __launch_bounds__(kBlockThreads) __attribute__((amdgpu_waves_per_eu(5))) void batch_strip64_kernel(
    const StripRec64* __restrict__ rec64, const double4* __restrict__ rec,
    const double* __restrict__ q, const int32_t* __restrict__ perm, double* __restrict__ dst,
    double* __restrict__ partial, const BatchMesh* __restrict__ meshes,
    const BatchItem* __restrict__ items, int64_t n_items, int32_t* __restrict__ counter,
    int64_t chunk_tiles) {
  __shared__ int32_t s_item;
  const int tid = threadIdx.x;
  for (;;) {
    __syncthreads();  // every thread has read the previous s_item
    if (tid == 0) s_item = atomicAdd(counter, 1);
    __syncthreads();
    const int32_t it = __builtin_amdgcn_readfirstlane(s_item);
    if (it >= n_items) break;  // block-uniform exit
    const BatchItem item = items[it];
    const BatchMesh m = meshes[item.mesh];
    const int32_t local = item.qtile * kBlockThreads + tid;
    const bool valid = local < m.nq;
    const int64_t slot = m.query_off + local;
    const double ox = valid ? q[3 * slot] : 0.0;
    const double oy = valid ? q[3 * slot + 1] : 0.0;
    const double oz = valid ? q[3 * slot + 2] : 0.0;
    const int64_t t0 = m.face_off + (int64_t)item.chunk * chunk_tiles;
    int64_t t1 = t0 + chunk_tiles;
    if (t1 > m.face_off + m.nf) t1 = m.face_off + m.nf;
    const double acc = strip64_sweep(rec64, rec, t0, t1, ox, oy, oz);
    if (valid) {
      if (m.chunks == 1)
        dst[perm ? (int64_t)perm[slot] : slot] = acc * kInv2Pi;
      else
        partial[m.pbase + (int64_t)item.chunk * m.nq + local] = acc;
    }
  }
}

}  // namespace

hipError_t strip64_prep_launch(const double4* rec, int64_t nrec, StripRec64* out,
                               hipStream_t stream) {
  if (nrec == 0) return hipSuccess;
  strip64_prep_kernel<<<(unsigned)((nrec + kBlockThreads - 1) / kBlockThreads), kBlockThreads,
                        0, stream>>>(rec, nrec, out);
  return hipGetLastError();
}

int batch_strip64_resident_blocks() {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)batch_strip64_kernel,
                                                   kBlockThreads, 0) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return cus * per_cu;
}

hipError_t batch_strip64_enqueue(const StripRec64* rec64, const double4* rec, const double* q,
                                 const int32_t* perm, double* dst, double* partial,
                                 const BatchMesh* meshes, const BatchItem* items, int64_t n_items,
                                 int32_t* counter, int64_t chunk_tiles, int grid,
                                 hipStream_t stream) {
  batch_strip64_kernel<<<dim3((unsigned)grid), kBlockThreads, 0, stream>>>(
      rec64, rec, q, perm, dst, partial, meshes, items, n_items, counter, chunk_tiles);
  return hipGetLastError();
}

hipError_t strip64_launch(const StripRec64* rec64, const double4* rec, int64_t nrec,
                          const double* q, double* dst, int64_t nq, int splits, int sort,
                          void* ws, size_t ws_bytes, hipStream_t stream) {
  if (nq == 0) return hipSuccess;
  if (ws_bytes < strip_workspace_bytes(nq, splits, sort)) return hipErrorInvalidValue;
  const int64_t ntiles = nrec / kStripTileRecords;
  char* cursor = (char*)ws;
  double* partial = nullptr;
  if (splits > 1) {
    partial = (double*)cursor;
    cursor += align_up((size_t)nq * (size_t)splits * sizeof(double));
  }
  int32_t* perm = nullptr;
  if (sort) {
    perm = (int32_t*)cursor;
    cursor += align_up((size_t)nq * sizeof(int32_t));
    double* qs = (double*)cursor;
    cursor += align_up((size_t)nq * 3 * sizeof(double));
    hipError_t e = morton_order(q, nq, perm, cursor, morton_workspace_bytes(nq), stream);
    if (e == hipSuccess) e = gather_queries_launch(q, perm, nq, qs, stream);
    if (e != hipSuccess) return e;
    q = qs;
  }
  if (ntiles == 0) return fill_launch(dst, nq, 0.0, stream);
  const int64_t tps = (ntiles + splits - 1) / splits;
  const int used_splits = (int)((ntiles + tps - 1) / tps);
  const dim3 grid((unsigned)((nq + kBlockThreads - 1) / kBlockThreads), (unsigned)used_splits);
  timer_begin(stream);
  fwd_strip64_kernel<<<grid, kBlockThreads, 0, stream>>>(
      rec64, rec, ntiles, q, perm, used_splits > 1 ? partial : dst, nq, tps, used_splits);
  hipError_t e = hipGetLastError();
  timer_end(stream);
  if (e != hipSuccess || used_splits == 1) return e;
  return reduce_splits_launch(partial, perm, dst, nq, used_splits, stream);
}

}  // namespace wn
