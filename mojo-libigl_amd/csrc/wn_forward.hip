// Forward winding-number kernels for gfx950.
//
// All three kernels run one query per lane (QPL per lane with the template
// argument) over a face range. The plain FP64 and the FP32-compensated
// kernels broadcast faces from LDS: a block of 256 threads owns 256*QPL
// queries, streams the face list through LDS tiles of 256 faces (one face
// staged per thread, coalesced loads from the 9 x nf SoA), and every lane
// sweeps each tile with all its queries. The FP64 kernel with the far-pair
// fast path reads prepared 128-byte face records through scalar loads
// instead. Accumulation stays in registers in fixed face order, so results
// are bitwise reproducible.
//
// Split-K: grid.y splits the face range. Each (block, split) writes unscaled
// partials to partial[s * nq + slot] (slot = the query's position in launch
// order, so a block's writes are contiguous); the reduce pass sums the splits
// of a slot in fixed order -- one thread per slot for up to 16 splits, one
// wave per slot with a DPP/ds_swizzle tree beyond -- applies 1/(2 pi) and
// writes dst through the query permutation. With splits == 1 the kernel
// writes dst directly.
//
// The reference instead stores an nf x nq FP64 angle buffer and reduces it
// serially per query (src/capi.mojo:297-363, 366-398); nothing of nf*nq size
// is materialised here.

#include "wn_device.h"
#include "wn_f64.h"
#include "wn_internal.h"
#include "wn_sweep.h"

namespace wn {

// --------------------------------------------------------------------------
// FP64 plain: the reference formula in FP64 for every pair.
// --------------------------------------------------------------------------
template <int QPL>
__global__ __launch_bounds__(kBlockThreads) void fwd_f64_kernel(
    const double* __restrict__ tri, const double* __restrict__ q,
    double* __restrict__ out, int64_t nf, int64_t nq, int64_t faces_per_split,
    int splits) {
  __shared__ SweepF64Smem sm;
  const int tid = threadIdx.x;
  const int64_t qbase = (int64_t)blockIdx.x * (kBlockThreads * QPL);

  double ox[QPL], oy[QPL], oz[QPL], acc[QPL];
#pragma unroll
  for (int k = 0; k < QPL; ++k) {
    const int64_t qi = qbase + k * kBlockThreads + tid;
    const bool ok = qi < nq;
    ox[k] = ok ? q[3 * qi] : 0.0;
    oy[k] = ok ? q[3 * qi + 1] : 0.0;
    oz[k] = ok ? q[3 * qi + 2] : 0.0;
    acc[k] = 0.0;
  }
  const int64_t f0 = (int64_t)blockIdx.y * faces_per_split;
  const int64_t f1 = f0 + faces_per_split < nf ? f0 + faces_per_split : nf;
  sweep_f64<QPL>(tri, nf, f0, f1, ox, oy, oz, acc, sm);

#pragma unroll
  for (int k = 0; k < QPL; ++k) {
    const int64_t qi = qbase + k * kBlockThreads + tid;
    if (qi < nq) {
      if (splits == 1)
        out[qi] = acc[k] * kInv2Pi;
      else
        out[(int64_t)blockIdx.y * nq + qi] = acc[k];
    }
  }
}

// --------------------------------------------------------------------------
// FP64 with the far-pair fast path: the FP64 strip kernel's per-pair
// arithmetic on face records. Per pair: three offsets, three squared lengths
// and three refined square roots, the law-of-cosines dot products with the
// record's halved squared edge lengths, and det = N . a with the record's
// normal (exact: N . (b - a) = N . (c - a) = 0). A wave whose pairs are all
// far (|det| < den / 8 and den > la lb lc) takes atan(det / den) from the
// 7-term polynomial; any other wave evaluates every lane's pair with the
// plain formula (pair_f64 on the same offsets, the plain kernel's own
// arithmetic). Records are read at wave-uniform addresses, so they arrive in
// scalar registers through scalar loads: no LDS, no tile staging, no
// barriers.
// --------------------------------------------------------------------------
__global__ __launch_bounds__(kBlockThreads) void face_rec64_kernel(
    const double* __restrict__ tri, int64_t nf, FaceRec64* __restrict__ out) {
  const int64_t f = (int64_t)blockIdx.x * kBlockThreads + threadIdx.x;
  if (f >= nf) return;
  FaceRec64 r;
  r.ax = tri[f];
  r.ay = tri[nf + f];
  r.az = tri[2 * nf + f];
  r.bx = tri[3 * nf + f];
  r.by = tri[4 * nf + f];
  r.bz = tri[5 * nf + f];
  r.cx = tri[6 * nf + f];
  r.cy = tri[7 * nf + f];
  r.cz = tri[8 * nf + f];
  const double e1x = r.bx - r.ax, e1y = r.by - r.ay, e1z = r.bz - r.az;  // B - A
  const double e2x = r.cx - r.ax, e2y = r.cy - r.ay, e2z = r.cz - r.az;  // C - A
  const double e3x = r.cx - r.bx, e3y = r.cy - r.by, e3z = r.cz - r.bz;  // C - B
  r.nx = e1y * e2z - e1z * e2y;
  r.ny = e1z * e2x - e1x * e2z;
  r.nz = e1x * e2y - e1y * e2x;
  r.eab = -0.5 * (e1x * e1x + e1y * e1y + e1z * e1z);
  r.ebc = -0.5 * (e3x * e3x + e3y * e3y + e3z * e3z);
  r.eca = -0.5 * (e2x * e2x + e2y * e2y + e2z * e2z);
  r.pad = 0.0;
  out[f] = r;
}

// The plain FP64 pair of face r and query o, out of line (offsets formed
// inside: the sweep keeps no registers for them).
__device__ __noinline__ double face64_pair(const FaceRec64* r, double ox, double oy, double oz) {
  return pair_f64(r->ax - ox, r->ay - oy, r->az - oz, r->bx - ox, r->by - oy, r->bz - oz,
                  r->cx - ox, r->cy - oy, r->cz - oz);
}

struct FaceGeom64 {
  double det, den, dp;
};

// det, den and dp = den - la lb lc of face r for query o.
__device__ __forceinline__ FaceGeom64 face64_geom(const FaceRec64& r, double ox, double oy,
                                                  double oz) {
  const double ax = r.ax - ox, ay = r.ay - oy, az = r.az - oz;
  const double bx = r.bx - ox, by = r.by - oy, bz = r.bz - oz;
  const double cx = r.cx - ox, cy = r.cy - oy, cz = r.cz - oz;
  const double a2 = __builtin_fma(az, az, __builtin_fma(ay, ay, ax * ax));
  const double b2 = __builtin_fma(bz, bz, __builtin_fma(by, by, bx * bx));
  const double c2 = __builtin_fma(cz, cz, __builtin_fma(cy, cy, cx * cx));
  const double la = sqrt_gs_f64(a2), lb = sqrt_gs_f64(b2), lc = sqrt_gs_f64(c2);
  const double s_ab = half_sum_plus(a2, b2, r.eab);  // a.b
  const double s_bc = half_sum_plus(b2, c2, r.ebc);  // b.c
  const double s_ca = half_sum_plus(c2, a2, r.eca);  // c.a
  FaceGeom64 g;
  g.det = __builtin_fma(r.nz, az, __builtin_fma(r.ny, ay, r.nx * ax));
  g.dp = __builtin_fma(lc, s_ab, __builtin_fma(la, s_bc, lb * s_ca));
  g.den = __builtin_fma(la * lb, lc, g.dp);
  return g;
}

// One face on its own for one query: a far pair on every lane into acc, a
// pair of the plain formula (offsets formed again from the record: the same
// values) into acc_slow (two sums, each only ever updated in place).
__device__ __forceinline__ void face64_single(const FaceRec64* r, const FaceGeom64& g, double ox,
                                              double oy, double oz, double& acc,
                                              double& acc_slow) {
  const double det8 = 8.0 * g.det;  // exact
  if (__builtin_expect(wave_all_far_x8_f64(det8, g.den, g.dp), 1)) {
    acc += atan_small_x8_f64(div_f64(det8, g.den));
  } else {
    asm volatile("" ::);  // keep a real wave-uniform branch
    acc_slow += face64_pair(r, ox, oy, oz);
  }
}

// One face for the lane's QPL queries.
template <int QPL>
__device__ __forceinline__ void face64_step(const FaceRec64& r, const double (&ox)[QPL],
                                            const double (&oy)[QPL], const double (&oz)[QPL],
                                            double (&acc)[QPL], double (&acc_slow)[QPL]) {
#pragma unroll
  for (int k = 0; k < QPL; ++k)
    face64_single(&r, face64_geom(r, ox[k], oy[k], oz[k]), ox[k], oy[k], oz[k], acc[k],
                  acc_slow[k]);
}

// Three faces for the lane's QPL queries: when every lane meets
// sum3_f64's conditions their angles are taken together as the argument
// of the product of their (den + i det), one quotient and one polynomial for
// the three; otherwise face by face.
template <int QPL>
__device__ __forceinline__ void face64_triple(const FaceRec64* __restrict__ r,
                                              const double (&ox)[QPL], const double (&oy)[QPL],
                                              const double (&oz)[QPL], double (&acc)[QPL],
                                              double (&acc_slow)[QPL]) {
#pragma unroll
  for (int k = 0; k < QPL; ++k) {
    const FaceGeom64 g1 = face64_geom(r[0], ox[k], oy[k], oz[k]);
    const FaceGeom64 g2 = face64_geom(r[1], ox[k], oy[k], oz[k]);
    const FaceGeom64 g3 = face64_geom(r[2], ox[k], oy[k], oz[k]);
    const Sum3F64 s = sum3_f64(g1.det, g1.den, g2.det, g2.den, g3.det, g3.den);
    if (__builtin_expect(wave_all_sum3_f64(g1.dp, g2.dp, g3.dp, s), 1)) {
      acc[k] += atan_sum_f64(s.x);
    } else {
      asm volatile("" ::);  // keep a real wave-uniform branch
      face64_single(r, g1, ox[k], oy[k], oz[k], acc[k], acc_slow[k]);
      face64_single(r + 1, g2, ox[k], oy[k], oz[k], acc[k], acc_slow[k]);
      face64_single(r + 2, g3, ox[k], oy[k], oz[k], acc[k], acc_slow[k]);
    }
  }
}

// Three faces whose product t and smallest dp the caller formed: the group's
// angle when every lane meets the group test, else face by face, each
// face's geometry rebuilt from its record (the values the caller formed).
// Out of line (the fallback of a failed group of six, rare): its registers
// are not reserved for the sweep. Returns (fast-path sum, plain-formula sum).
__device__ __noinline__ double2 face64_triple_of(const FaceRec64* __restrict__ r, double t_re,
                                                 double t_im, double dp_min, double ox,
                                                 double oy, double oz) {
  double acc = 0.0, acc_slow = 0.0;
  const double x = div_f64(t_im, t_re);  // used only where t_re is positive normal
  if (__builtin_expect(wave_all_group_f64(dp_min, t_re, t_re, x), 1)) {
    acc = atan_sum_f64(x);
  } else {
    asm volatile("" ::);  // keep a real wave-uniform branch
    for (int i = 0; i < 3; ++i)
      face64_single(r + i, face64_geom(r[i], ox, oy, oz), ox, oy, oz, acc, acc_slow);
  }
  return make_double2(acc, acc_slow);
}

// Six faces for the lane's QPL queries: when every lane meets the group test
// their angles are taken together as the argument of the product of their
// (den + i det), one quotient and one polynomial for the six; otherwise as
// two triples (face64_triple_of). Only the triples' products and smallest
// dp stay live.
template <int QPL>
__device__ __forceinline__ void face64_hexad(const FaceRec64* __restrict__ r,
                                             const double (&ox)[QPL], const double (&oy)[QPL],
                                             const double (&oz)[QPL], double (&acc)[QPL],
                                             double (&acc_slow)[QPL]) {
#pragma unroll
  for (int k = 0; k < QPL; ++k) {
    const FaceGeom64 g1 = face64_geom(r[0], ox[k], oy[k], oz[k]);
    const FaceGeom64 g2 = face64_geom(r[1], ox[k], oy[k], oz[k]);
    const FaceGeom64 g3 = face64_geom(r[2], ox[k], oy[k], oz[k]);
    const Prod3F64 t1 = prod3_f64(g1.det, g1.den, g2.det, g2.den, g3.det, g3.den);
    const double m1 = fmin(fmin(g1.dp, g2.dp), g3.dp);
    const FaceGeom64 g4 = face64_geom(r[3], ox[k], oy[k], oz[k]);
    const FaceGeom64 g5 = face64_geom(r[4], ox[k], oy[k], oz[k]);
    const FaceGeom64 g6 = face64_geom(r[5], ox[k], oy[k], oz[k]);
    const Prod3F64 t2 = prod3_f64(g4.det, g4.den, g5.det, g5.den, g6.det, g6.den);
    const double m2 = fmin(fmin(g4.dp, g5.dp), g6.dp);
    const double re = __builtin_fma(-t1.im, t2.im, t1.re * t2.re);
    const double im = __builtin_fma(t1.re, t2.im, t1.im * t2.re);
    const double x = div_f64(im, re);  // used only where re is positive normal
    if (__builtin_expect(wave_all_group_f64(fmin(m1, m2), fmin(t1.re, t2.re), re, x), 1)) {
      acc[k] += atan_sum_f64(x);
    } else {
      asm volatile("" ::);  // keep a real wave-uniform branch
      const double2 a = face64_triple_of(r, t1.re, t1.im, m1, ox[k], oy[k], oz[k]);
      const double2 b = face64_triple_of(r + 3, t2.re, t2.im, m2, ox[k], oy[k], oz[k]);
      acc[k] += a.x;
      acc[k] += b.x;
      acc_slow[k] += a.y;
      acc_slow[k] += b.y;
    }
  }
}

template <int QPL>
__global__ __launch_bounds__(kBlockThreads) __attribute__((amdgpu_waves_per_eu(QPL == 1 ? 6 : 4))) void fwd_f64f_kernel(
    const FaceRec64* __restrict__ rec, const double* __restrict__ q, double* __restrict__ out,
    int64_t nf, int64_t nq, int64_t faces_per_split, int splits) {
  const int tid = threadIdx.x;
  const int64_t qbase = (int64_t)blockIdx.x * (kBlockThreads * QPL);
  double ox[QPL], oy[QPL], oz[QPL], acc[QPL], acc_slow[QPL];
#pragma unroll
  for (int k = 0; k < QPL; ++k) {
    const int64_t qi = qbase + k * kBlockThreads + tid;
    const bool ok = qi < nq;
    ox[k] = ok ? q[3 * qi] : 0.0;
    oy[k] = ok ? q[3 * qi + 1] : 0.0;
    oz[k] = ok ? q[3 * qi + 2] : 0.0;
    acc[k] = 0.0;
    acc_slow[k] = 0.0;
  }
  const int64_t f0 = (int64_t)blockIdx.y * faces_per_split;
  const int64_t f1 = f0 + faces_per_split < nf ? f0 + faces_per_split : nf;
  // 32-bit inner counter: the loop test stays a scalar compare
  for (int64_t c0 = f0; c0 < f1; c0 += (int64_t)1 << 30) {
    const FaceRec64* r = rec + c0;
    const int cnt = (int)(f1 - c0 < ((int64_t)1 << 30) ? f1 - c0 : ((int64_t)1 << 30));
    int j = 0;
    for (; j + 6 <= cnt; j += 6) face64_hexad<QPL>(r + j, ox, oy, oz, acc, acc_slow);
    for (; j + 3 <= cnt; j += 3) face64_triple<QPL>(r + j, ox, oy, oz, acc, acc_slow);
    for (; j < cnt; ++j) face64_step<QPL>(r[j], ox, oy, oz, acc, acc_slow);
  }
#pragma unroll
  for (int k = 0; k < QPL; ++k) {
    const int64_t qi = qbase + k * kBlockThreads + tid;
    if (qi < nq) {
      const double a = acc[k] + acc_slow[k];
      if (splits == 1)
        out[qi] = a * kInv2Pi;
      else
        out[(int64_t)blockIdx.y * nq + qi] = a;
    }
  }
}

// Persistent FP64 face kernel over a batch plan's work list (mesh, query
// tile of 256*QPL queries, face chunk), with the batch kernels' item
// protocol (atomic counter, block-uniform exit) and output rule (single-chunk
// meshes write w, others unscaled chunk-major partials summed in chunk
// order). rec holds the records of the batch's 9 x F_total SoA.
template <int QPL>
__global__ __launch_bounds__(kBlockThreads) __attribute__((amdgpu_waves_per_eu(QPL == 1 ? 6 : 4))) void batch_f64f_kernel(
    const FaceRec64* __restrict__ rec, const double* __restrict__ q,
    const int32_t* __restrict__ perm, double* __restrict__ dst, double* __restrict__ partial,
    const BatchMesh* __restrict__ meshes, const BatchItem* __restrict__ items, int64_t n_items,
    int32_t* __restrict__ counter, int64_t chunk_faces) {
  __shared__ int32_t s_item;
  const int tid = threadIdx.x;
  constexpr int QB = kBlockThreads * QPL;
  for (;;) {
    __syncthreads();  // every thread has read the previous s_item
    if (tid == 0) s_item = atomicAdd(counter, 1);
    __syncthreads();
    const int32_t it = __builtin_amdgcn_readfirstlane(s_item);
    if (it >= n_items) break;  // block-uniform exit
    const BatchItem item = items[it];
    const BatchMesh m = meshes[item.mesh];
    const int64_t slot0 = m.query_off + (int64_t)item.qtile * QB;
    const int64_t slot_end = m.query_off + m.nq;
    const int64_t f0 = m.face_off + (int64_t)item.chunk * chunk_faces;
    int64_t f1 = f0 + chunk_faces;
    if (f1 > m.face_off + m.nf) f1 = m.face_off + m.nf;
    bool valid[QPL];
    int64_t qidx[QPL];
    double ox[QPL], oy[QPL], oz[QPL], acc[QPL], acc_slow[QPL];
#pragma unroll
    for (int k = 0; k < QPL; ++k) {
      const int64_t slot = slot0 + k * kBlockThreads + tid;
      valid[k] = slot < slot_end;
      qidx[k] = valid[k] ? (perm ? (int64_t)perm[slot] : slot) : 0;
      // q is in launch order (gathered through perm beforehand when sorted)
      ox[k] = valid[k] ? q[3 * slot] : 0.0;
      oy[k] = valid[k] ? q[3 * slot + 1] : 0.0;
      oz[k] = valid[k] ? q[3 * slot + 2] : 0.0;
      acc[k] = 0.0;
      acc_slow[k] = 0.0;
    }
    const FaceRec64* r = rec + f0;
    const int cnt = (int)(f1 - f0);  // a mesh has at most 2^31-1 faces
    int j = 0;
    for (; j + 6 <= cnt; j += 6) face64_hexad<QPL>(r + j, ox, oy, oz, acc, acc_slow);
    for (; j + 3 <= cnt; j += 3) face64_triple<QPL>(r + j, ox, oy, oz, acc, acc_slow);
    for (; j < cnt; ++j) face64_step<QPL>(r[j], ox, oy, oz, acc, acc_slow);
#pragma unroll
    for (int k = 0; k < QPL; ++k) {
      if (valid[k]) {
        const double a = acc[k] + acc_slow[k];
        const int64_t slot = slot0 + k * kBlockThreads + tid;
        if (m.chunks == 1)
          dst[qidx[k]] = a * kInv2Pi;
        else  // chunk-major, slot order: a block's writes are contiguous
          partial[m.pbase + (int64_t)item.chunk * m.nq + (slot - m.query_off)] = a;
      }
    }
  }
}

// --------------------------------------------------------------------------
// FP32-compensated.
//  * local origin: the block's query bbox centre g (FP64). Query offsets
//    o = O - g and face corners A - g are rounded to FP32 once, so the whole
//    block works on one consistently rounded copy of the mesh whose rounding
//    error is eps * |A - g|: small exactly where it matters (faces near the
//    block's queries). With Morton-ordered queries (perm) blocks are compact.
//  * per pair: FP32 arithmetic (pair_rec32), angle accumulated into an FP64
//    running sum (an FP32 value converts to FP64 exactly; the FP64 sum of the
//    per-face terms keeps ~1e-16 relative accumulation error, which is the
//    compensation a Kahan/TwoSum FP32 sum would provide, for 2 VALU ops).
//  * ill-conditioned pairs (risky) are re-evaluated in FP64 from the same
//    FP32 record (pair_rec32_f64).
// --------------------------------------------------------------------------
template <int QPL>
__global__ __launch_bounds__(kBlockThreads) __attribute__((amdgpu_waves_per_eu(QPL == 1 ? 8 : 4))) void fwd_f32c_kernel(
    const double* __restrict__ tri, const double* __restrict__ q,
    const int32_t* __restrict__ perm, double* __restrict__ out, int64_t nf,
    int64_t nq, int64_t faces_per_split, int splits, float kappa) {
  __shared__ SweepF32Smem sm;
  const int tid = threadIdx.x;
  const int64_t qbase = (int64_t)blockIdx.x * (kBlockThreads * QPL);

  bool valid[QPL];
  double Ox[QPL], Oy[QPL], Oz[QPL];
#pragma unroll
  for (int k = 0; k < QPL; ++k) {
    const int64_t slot = qbase + k * kBlockThreads + tid;
    valid[k] = slot < nq;
    // q is in launch order (gathered through perm beforehand when sorted)
    Ox[k] = valid[k] ? q[3 * slot] : 0.0;
    Oy[k] = valid[k] ? q[3 * slot + 1] : 0.0;
    Oz[k] = valid[k] ? q[3 * slot + 2] : 0.0;
  }
  double gx, gy, gz;
  int er;
  block_origin<QPL>(Ox, Oy, Oz, valid, sm, gx, gy, gz, er);
  const double so = pow2_f64(2 - er);

  float ox[QPL], oy[QPL], oz[QPL];
  double acc[QPL];
#pragma unroll
  for (int k = 0; k < QPL; ++k) {
    ox[k] = valid[k] ? (float)((Ox[k] - gx) * so) : 0.0f;
    oy[k] = valid[k] ? (float)((Oy[k] - gy) * so) : 0.0f;
    oz[k] = valid[k] ? (float)((Oz[k] - gz) * so) : 0.0f;
    acc[k] = 0.0;
  }
  const int64_t f0 = (int64_t)blockIdx.y * faces_per_split;
  const int64_t f1 = f0 + faces_per_split < nf ? f0 + faces_per_split : nf;
  sweep_f32c<QPL>(tri, nf, f0, f1, gx, gy, gz, er, ox, oy, oz, kappa, acc, sm);

  // output slots rebuilt after the sweep (kept live across it, the 64-bit
  // output index was spilled to scratch at every block start)
  const int32_t t = thread_index(wave_base());
#pragma unroll
  for (int k = 0; k < QPL; ++k) {
    const int64_t slot = (int64_t)blockIdx.x * (kBlockThreads * QPL) + k * kBlockThreads + t;
    if (slot < nq) {
      if (splits == 1)
        out[perm ? (int64_t)perm[slot] : slot] = acc[k] * kInv2Pi;
      else
        out[(int64_t)blockIdx.y * nq + slot] = acc[k];
    }
  }
}

__global__ __launch_bounds__(kBlockThreads) void fill_f64_kernel(double* __restrict__ p,
                                                                 int64_t n, double v) {
  const int64_t i = (int64_t)blockIdx.x * kBlockThreads + threadIdx.x;
  if (i < n) p[i] = v;
}

// Reduction of split partials partial[s * nq + slot] -> dst[perm[slot]].
// Few splits: one thread per slot, splits summed in order.
__global__ __launch_bounds__(kBlockThreads) void reduce_splits_kernel(
    const double* __restrict__ partial, const int32_t* __restrict__ perm,
    double* __restrict__ dst, int64_t nq, int splits) {
  const int64_t slot = (int64_t)blockIdx.x * kBlockThreads + threadIdx.x;
  if (slot >= nq) return;
  double s = 0.0;
  for (int j = 0; j < splits; ++j) s += partial[(int64_t)j * nq + slot];
  dst[perm ? (int64_t)perm[slot] : slot] = s * kInv2Pi;
}

// Many splits: one wave per slot, lanes stride over the splits, then a fixed
// DPP/ds_swizzle tree (wave_sum_f64); lane 0 writes.
__global__ __launch_bounds__(kBlockThreads) void reduce_splits_wave_kernel(
    const double* __restrict__ partial, const int32_t* __restrict__ perm,
    double* __restrict__ dst, int64_t nq, int splits) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t slot = (int64_t)blockIdx.x * (kBlockThreads / kWave) + threadIdx.x / kWave;
  if (slot >= nq) return;  // wave-uniform
  double s = 0.0;
  for (int j = lane; j < splits; j += kWave) s += partial[(int64_t)j * nq + slot];
  s = wave_sum_f64(s);
  if (lane == 0) dst[perm ? (int64_t)perm[slot] : slot] = s * kInv2Pi;
}

// --------------------------------------------------------------------------
// Device ingest: tri[k*nf + f] = V[F[f, k/3], k%3]. Out-of-range indices are
// reported through bad_index (first offending face, via atomicMin) and the
// corner is zeroed.
// --------------------------------------------------------------------------
template <typename VT, typename FT>
__global__ __launch_bounds__(kBlockThreads) void pack_soa_kernel(
    const VT* __restrict__ V, const FT* __restrict__ F, int64_t nv, int64_t nf,
    double* __restrict__ tri, int32_t* __restrict__ bad_index) {
  const int64_t f = (int64_t)blockIdx.x * kBlockThreads + threadIdx.x;
  if (f >= nf) return;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int64_t vi = (int64_t)F[3 * f + c];
    const bool ok = vi >= 0 && vi < nv;
    if (!ok && bad_index) atomicMin(bad_index, (int32_t)(f < 0x7fffffff ? f : 0x7fffffff));
#pragma unroll
    for (int d = 0; d < 3; ++d)
      tri[(3 * c + d) * nf + f] = ok ? (double)V[3 * vi + d] : 0.0;
  }
}

hipError_t pack_soa_launch(const void* V, const void* F, int64_t nv, int64_t nf,
                           int vdtype, int fdtype, double* tri, int32_t* bad_index,
                           hipStream_t stream) {
  if (nf <= 0) return hipSuccess;
  const dim3 grid((unsigned)((nf + kBlockThreads - 1) / kBlockThreads));
  if (vdtype == 0 && fdtype == 0)
    pack_soa_kernel<double, int64_t><<<grid, kBlockThreads, 0, stream>>>(
        (const double*)V, (const int64_t*)F, nv, nf, tri, bad_index);
  else if (vdtype == 0 && fdtype == 1)
    pack_soa_kernel<double, int32_t><<<grid, kBlockThreads, 0, stream>>>(
        (const double*)V, (const int32_t*)F, nv, nf, tri, bad_index);
  else if (vdtype == 1 && fdtype == 0)
    pack_soa_kernel<float, int64_t><<<grid, kBlockThreads, 0, stream>>>(
        (const float*)V, (const int64_t*)F, nv, nf, tri, bad_index);
  else
    pack_soa_kernel<float, int32_t><<<grid, kBlockThreads, 0, stream>>>(
        (const float*)V, (const int32_t*)F, nv, nf, tri, bad_index);
  return hipGetLastError();
}

// --------------------------------------------------------------------------
// Host launchers.
// --------------------------------------------------------------------------
hipError_t fill_launch(double* p, int64_t n, double v, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  fill_f64_kernel<<<dim3((unsigned)((n + kBlockThreads - 1) / kBlockThreads)), kBlockThreads, 0,
                    stream>>>(p, n, v);
  return hipGetLastError();
}

hipError_t face_rec64_launch(const double* tri, int64_t nf, FaceRec64* out, hipStream_t stream) {
  if (nf <= 0) return hipSuccess;
  face_rec64_kernel<<<dim3((unsigned)((nf + kBlockThreads - 1) / kBlockThreads)), kBlockThreads, 0,
                      stream>>>(tri, nf, out);
  return hipGetLastError();
}

const void* batch_face64_kernel(int qpl) {
  return qpl == 1 ? (const void*)batch_f64f_kernel<1> : (const void*)batch_f64f_kernel<2>;
}

hipError_t batch_face64_enqueue(const FaceRec64* rec, const double* q, const int32_t* perm,
                                double* dst, double* partial, const BatchMesh* meshes,
                                const BatchItem* items, int64_t n_items, int32_t* counter,
                                int64_t chunk_faces, int qpl, int grid, hipStream_t stream) {
  if (qpl == 1)
    batch_f64f_kernel<1><<<dim3((unsigned)grid), kBlockThreads, 0, stream>>>(
        rec, q, perm, dst, partial, meshes, items, n_items, counter, chunk_faces);
  else
    batch_f64f_kernel<2><<<dim3((unsigned)grid), kBlockThreads, 0, stream>>>(
        rec, q, perm, dst, partial, meshes, items, n_items, counter, chunk_faces);
  return hipGetLastError();
}

hipError_t reduce_splits_launch(const double* partial, const int32_t* perm, double* dst,
                                int64_t nq, int splits, hipStream_t stream) {
  if (nq <= 0) return hipSuccess;
  if (splits <= 16) {
    reduce_splits_kernel<<<dim3((unsigned)((nq + kBlockThreads - 1) / kBlockThreads)),
                           kBlockThreads, 0, stream>>>(partial, perm, dst, nq, splits);
  } else {
    const int per_block = kBlockThreads / kWave;
    reduce_splits_wave_kernel<<<dim3((unsigned)((nq + per_block - 1) / per_block)),
                                kBlockThreads, 0, stream>>>(partial, perm, dst, nq, splits);
  }
  return hipGetLastError();
}

int fwd_resident_blocks(int precision, int qpl) {
  const void* fn = nullptr;
  if (precision == kFP64Plain)
    fn = qpl == 1 ? (const void*)fwd_f64_kernel<1> : (const void*)fwd_f64_kernel<2>;
  else if (precision == kFP64)
    fn = qpl == 1 ? (const void*)fwd_f64f_kernel<1> : (const void*)fwd_f64f_kernel<2>;
  else
    fn = qpl == 1 ? (const void*)fwd_f32c_kernel<1>
                  : (qpl == 2 ? (const void*)fwd_f32c_kernel<2> : (const void*)fwd_f32c_kernel<4>);
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kBlockThreads, 0) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return cus * per_cu;
}

static int64_t faces_per_split_of(int64_t nf, int splits) {
  int64_t per = (nf + splits - 1) / splits;
  per = (per + kFaceTile - 1) / kFaceTile * kFaceTile;
  return per > 0 ? per : kFaceTile;
}

const char* fwd_check(int64_t nf, int64_t nq, const FwdConfig& cfg) {
  if (nf < 0 || nq < 0) return "negative size";
  if (cfg.precision != kFP64 && cfg.precision != kFP32C && cfg.precision != kFP64Plain)
    return "unknown precision";
  if (cfg.qpl != 1 && cfg.qpl != 2 && cfg.qpl != 4) return "queries_per_lane must be 1, 2 or 4";
  if (cfg.precision != kFP32C && cfg.qpl == 4) return "fp64 supports queries_per_lane 1 or 2";
  if (cfg.splits < 1 || cfg.splits > 65535) return "splits must be in [1, 65535]";
  if (nq > (int64_t)0x7fffffff) return "more than 2^31-1 queries per call";
  const int64_t qblock = (int64_t)kBlockThreads * cfg.qpl;
  if ((nq + qblock - 1) / qblock > (int64_t)0x7fffffff) return "too many query blocks";
  if (!(cfg.kappa >= 0.0f)) return "kappa must be >= 0";
  return nullptr;
}

size_t fwd_workspace_bytes(int64_t nf, int64_t nq, const FwdConfig& cfg) {
  size_t bytes = 0;
  if (cfg.splits > 1) bytes += align_up((size_t)nq * (size_t)cfg.splits * sizeof(double));
  if (cfg.precision == kFP64 && nf > 0) bytes += align_up((size_t)nf * sizeof(FaceRec64));
  if (cfg.precision == kFP32C && cfg.sort) {
    bytes += align_up((size_t)nq * sizeof(int32_t));
    bytes += align_up((size_t)nq * 3 * sizeof(double));
    bytes += morton_workspace_bytes(nq);
  }
  if (cfg.precision == kFP32C && nf > 0) {
    // boundary edges from the corner coordinates, then the correction
    bytes += soa_boundary_workspace_bytes(nf, 0);
    bytes += ribbon_workspace_bytes(nq, ribbon_splits(nq, cfg.qpl, 3 * nf));
  }
  return bytes;
}

hipError_t fwd_launch(const double* tri, const double* q, double* dst, int64_t nf,
                      int64_t nq, const FwdConfig& cfg, void* ws, size_t ws_bytes,
                      hipStream_t stream) {
  if (nq == 0) return hipSuccess;
  if (ws_bytes < fwd_workspace_bytes(nf, nq, cfg)) return hipErrorInvalidValue;
  char* cursor = (char*)ws;
  double* partial = nullptr;
  if (cfg.splits > 1) {
    partial = (double*)cursor;
    cursor += align_up((size_t)nq * (size_t)cfg.splits * sizeof(double));
  }
  int32_t* perm = nullptr;
  if (cfg.precision == kFP32C && cfg.sort) {
    perm = (int32_t*)cursor;
    cursor += align_up((size_t)nq * sizeof(int32_t));
    double* qs = (double*)cursor;
    cursor += align_up((size_t)nq * 3 * sizeof(double));
    hipError_t e = morton_order(q, nq, perm, cursor, morton_workspace_bytes(nq), stream);
    if (e == hipSuccess) e = gather_queries_launch(q, perm, nq, qs, stream);
    if (e != hipSuccess) return e;
    cursor += morton_workspace_bytes(nq);
    q = qs;
  }
  double* out = cfg.splits > 1 ? partial : dst;
  const int64_t fps = faces_per_split_of(nf, cfg.splits);
  const int64_t qblock = (int64_t)kBlockThreads * cfg.qpl;
  const dim3 grid((unsigned)((nq + qblock - 1) / qblock), (unsigned)cfg.splits);
  if (nf == 0) {
    // Empty mesh: w = 0 everywhere (a fill kernel, not a memset node, so the
    // launch sequence stays safe to capture in a graph).
    fill_f64_kernel<<<dim3((unsigned)((nq + kBlockThreads - 1) / kBlockThreads)),
                      kBlockThreads, 0, stream>>>(dst, nq, 0.0);
    return hipGetLastError();
  }
  FaceRec64* rec = nullptr;
  if (cfg.precision == kFP64) {
    rec = (FaceRec64*)cursor;
    cursor += align_up((size_t)nf * sizeof(FaceRec64));
    const hipError_t e = face_rec64_launch(tri, nf, rec, stream);
    if (e != hipSuccess) return e;
  }
  timer_begin(stream);
  if (cfg.precision == kFP64Plain) {
    if (cfg.qpl == 1)
      fwd_f64_kernel<1><<<grid, kBlockThreads, 0, stream>>>(tri, q, out, nf, nq, fps, cfg.splits);
    else
      fwd_f64_kernel<2><<<grid, kBlockThreads, 0, stream>>>(tri, q, out, nf, nq, fps, cfg.splits);
  } else if (cfg.precision == kFP64) {
    if (cfg.qpl == 1)
      fwd_f64f_kernel<1><<<grid, kBlockThreads, 0, stream>>>(rec, q, out, nf, nq, fps, cfg.splits);
    else
      fwd_f64f_kernel<2><<<grid, kBlockThreads, 0, stream>>>(rec, q, out, nf, nq, fps, cfg.splits);
  } else {
    if (cfg.qpl == 1)
      fwd_f32c_kernel<1><<<grid, kBlockThreads, 0, stream>>>(tri, q, perm, out, nf, nq, fps,
                                                             cfg.splits, cfg.kappa);
    else if (cfg.qpl == 2)
      fwd_f32c_kernel<2><<<grid, kBlockThreads, 0, stream>>>(tri, q, perm, out, nf, nq, fps,
                                                             cfg.splits, cfg.kappa);
    else
      fwd_f32c_kernel<4><<<grid, kBlockThreads, 0, stream>>>(tri, q, perm, out, nf, nq, fps,
                                                             cfg.splits, cfg.kappa);
  }
  hipError_t e = hipGetLastError();
  timer_end(stream);
  if (e != hipSuccess) return e;
  if (cfg.splits > 1) e = reduce_splits_launch(partial, perm, dst, nq, cfg.splits, stream);
  if (e != hipSuccess || cfg.precision != kFP32C) return e;
  // boundary correction of an open mesh: its boundary edges from the corner
  // coordinates (none for a closed mesh), then the same query blocks
  const size_t sb_bytes = soa_boundary_workspace_bytes(nf, 0);
  SoaBoundary sb;
  e = soa_boundary_launch(tri, nf, nullptr, 0, cursor, sb_bytes, &sb, stream);
  if (e != hipSuccess) return e;
  cursor += sb_bytes;
  RibbonSrc src;
  src.pts = tri;
  src.stride = nf;
  src.ids = sb.ids;
  src.w = sb.w;
  src.count = sb.count;
  const int rs = ribbon_splits(nq, cfg.qpl, 3 * nf);
  return ribbon_launch(src, 1, cfg.qpl, q, perm, dst, nq, rs, cursor,
                       ribbon_workspace_bytes(nq, rs), stream);
}

}  // namespace wn
