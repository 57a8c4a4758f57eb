// Block-level building blocks shared by the single-mesh forward kernels and
// the batched persistent kernel: the local-origin reduction and the sweep of
// a face range through LDS tiles for one block's queries.
//
// Contract of a sweep: all 256 threads of the block call it with the same
// face range [f0, f1) of a 9-row SoA whose rows are `stride` apart; every
// thread accumulates its QPL queries' angle sums (unscaled) into acc[] in
// face order. The function begins and ends with __syncthreads().
#pragma once

#include "wn_device.h"
#include "wn_internal.h"

namespace wn {

struct SweepF64Smem {
  double2 face[kFaceTile][5];
};

struct SweepF32Smem {
  FaceRec32 face[kFaceTile];
  double red[6][kBlockThreads / kWave];
  uint32_t tmax[kBlockThreads / kWave];  // per wave: abs_hi_word max |corner - g| of the next tile
};

template <int QPL>
__device__ __forceinline__ void sweep_f64(const double* __restrict__ tri, int64_t stride,
                                          int64_t f0, int64_t f1, const double (&ox)[QPL],
                                          const double (&oy)[QPL], const double (&oz)[QPL],
                                          double (&acc)[QPL], SweepF64Smem& sm) {
  const int tid = threadIdx.x;
  for (int64_t t0 = f0; t0 < f1; t0 += kFaceTile) {
    const int cnt = (int)(f1 - t0 < kFaceTile ? f1 - t0 : kFaceTile);
    __syncthreads();
    if (tid < cnt) {
      const int64_t f = t0 + tid;
      sm.face[tid][0] = make_double2(tri[f], tri[stride + f]);
      sm.face[tid][1] = make_double2(tri[2 * stride + f], tri[3 * stride + f]);
      sm.face[tid][2] = make_double2(tri[4 * stride + f], tri[5 * stride + f]);
      sm.face[tid][3] = make_double2(tri[6 * stride + f], tri[7 * stride + f]);
      sm.face[tid][4] = make_double2(tri[8 * stride + f], 0.0);
    }
    __syncthreads();
    for (int j = 0; j < cnt; ++j) {
      const double2 p0 = sm.face[j][0], p1 = sm.face[j][1], p2 = sm.face[j][2],
                    p3 = sm.face[j][3], p4 = sm.face[j][4];
#pragma unroll
      for (int k = 0; k < QPL; ++k) {
        acc[k] += pair_f64(p0.x - ox[k], p0.y - oy[k], p1.x - oz[k], p1.y - ox[k],
                           p2.x - oy[k], p2.y - oz[k], p3.x - ox[k], p3.y - oy[k],
                           p4.x - oz[k]);
      }
    }
  }
  __syncthreads();
}

// Centre of the bbox of the block's valid queries (valid[k] false = ignored)
// and er = frame_exponent of the bbox's largest side (|query - g| < 2^er).
// At least one thread of the block must hold a valid query.
template <int QPL>
__device__ __forceinline__ void block_origin(const double (&Ox)[QPL], const double (&Oy)[QPL],
                                             const double (&Oz)[QPL], const bool (&valid)[QPL],
                                             SweepF32Smem& sm, double& gx, double& gy,
                                             double& gz, int& er) {
  const int tid = threadIdx.x;
  double lo0 = __builtin_inf(), lo1 = __builtin_inf(), lo2 = __builtin_inf();
  double hi0 = -__builtin_inf(), hi1 = -__builtin_inf(), hi2 = -__builtin_inf();
#pragma unroll
  for (int k = 0; k < QPL; ++k) {
    if (valid[k]) {
      lo0 = fmin(lo0, Ox[k]); hi0 = fmax(hi0, Ox[k]);
      lo1 = fmin(lo1, Oy[k]); hi1 = fmax(hi1, Oy[k]);
      lo2 = fmin(lo2, Oz[k]); hi2 = fmax(hi2, Oz[k]);
    }
  }
  lo0 = wave_min_f64(lo0); lo1 = wave_min_f64(lo1); lo2 = wave_min_f64(lo2);
  hi0 = wave_max_f64(hi0); hi1 = wave_max_f64(hi1); hi2 = wave_max_f64(hi2);
  const int wave = tid / kWave;
  __syncthreads();
  if ((tid & (kWave - 1)) == 0) {
    sm.red[0][wave] = lo0; sm.red[1][wave] = lo1; sm.red[2][wave] = lo2;
    sm.red[3][wave] = hi0; sm.red[4][wave] = hi1; sm.red[5][wave] = hi2;
  }
  __syncthreads();
#pragma unroll
  for (int w = 0; w < kBlockThreads / kWave; ++w) {
    lo0 = fmin(lo0, sm.red[0][w]); lo1 = fmin(lo1, sm.red[1][w]); lo2 = fmin(lo2, sm.red[2][w]);
    hi0 = fmax(hi0, sm.red[3][w]); hi1 = fmax(hi1, sm.red[4][w]); hi2 = fmax(hi2, sm.red[5][w]);
  }
  gx = 0.5 * (lo0 + hi0);
  gy = 0.5 * (lo1 + hi1);
  gz = 0.5 * (lo2 + hi2);
  er = frame_exponent(fmax(fmax(hi0 - lo0, hi1 - lo1), hi2 - lo2));
}

// One face of a tile for QPL queries: its FP32 angle is added to part[k],
// the FP32 sum of at most three consecutive faces' angles that the caller
// adds to the FP64 sum. Far pairs on every lane of the wave (|det / den| < 1/8 and
// den > la lb lc): the small-angle arctangent; otherwise the octant-reduced
// atan2, and ill-conditioned pairs (max(|det|, |den|) < kappa la lb lc: a
// query near an edge line, or close to the face inside its outline) are
// re-evaluated in FP64 from the same record and rounded to FP32 once.
template <int QPL>
__device__ __forceinline__ void face_step_f32c(const FaceRec32& rec, const float (&ox)[QPL],
                                               const float (&oy)[QPL], const float (&oz)[QPL],
                                               float kappa, float (&part)[QPL]) {
#pragma clang fp contract(off)  // part += angle rounds the angle first
  const float4 r0 = rec.r0, r1 = rec.r1, r2 = rec.r2, r3 = rec.r3;
#pragma unroll
  for (int k = 0; k < QPL; ++k) {
    const PairGeom32 g = pair_geom32(r0, r1, r2, r3, ox[k], oy[k], oz[k]);
    const float det8 = 8.0f * g.det;  // exact
    float ang = atan_small_x8_f32(det8 * __builtin_amdgcn_rcpf(g.den));
    if (__builtin_expect(!wave_all_far_x8(det8, g.den, g.dp), 0)) {
      // An empty volatile asm gives this path a side effect, so the
      // compiler keeps a real (wave-uniform) branch instead of computing
      // both paths and selecting.
      asm volatile("" ::);
      float mx;
      ang = atan2_f32(g.det, g.den, mx);
      if (mx < kappa * g.pql) ang = (float)pair_rec32_f64(r0, r1, r2, ox[k], oy[k], oz[k]);
    }
    part[k] += ang;
  }
}

// Three faces of a tile for QPL queries: the FP32 sum of their angles into
// sum[k], as the argument of the product of their (den + i det) when every
// lane meets sum3_f32's conditions (one reciprocal and one polynomial for
// the three); otherwise face by face as face_step_f32c.
template <int QPL>
__device__ __forceinline__ void face_triple_f32c(const FaceRec32& ra, const FaceRec32& rb,
                                                 const FaceRec32& rc, const float (&ox)[QPL],
                                                 const float (&oy)[QPL], const float (&oz)[QPL],
                                                 float kappa, float (&sum)[QPL]) {
#pragma unroll
  for (int k = 0; k < QPL; ++k) {
    const PairGeom32 g1 = pair_geom32(ra.r0, ra.r1, ra.r2, ra.r3, ox[k], oy[k], oz[k]);
    const PairGeom32 g2 = pair_geom32(rb.r0, rb.r1, rb.r2, rb.r3, ox[k], oy[k], oz[k]);
    const PairGeom32 g3 = pair_geom32(rc.r0, rc.r1, rc.r2, rc.r3, ox[k], oy[k], oz[k]);
    const AngleSum s = sum3_f32(g1.det, g1.den, g2.det, g2.den, g3.det, g3.den);
    float part;
    if (__builtin_expect(wave_all_sum3(g1.dp, g2.dp, g3.dp, s), 1)) {
      part = atan_sum_f32(s.x);
    } else {
      asm volatile("" ::);  // keep a real wave-uniform branch
      float one[1] = {-0.0f};
      const float qx[1] = {ox[k]}, qy[1] = {oy[k]}, qz[1] = {oz[k]};
      face_step_f32c<1>(ra, qx, qy, qz, kappa, one);
      face_step_f32c<1>(rb, qx, qy, qz, kappa, one);
      face_step_f32c<1>(rc, qx, qy, qz, kappa, one);
      part = one[0];
    }
    sum[k] = part;
  }
}

// Three faces whose product t (prod3_f32) and smallest dp the caller formed:
// the FP32 sum of their angles by sum3_f32's test, else face by face as
// face_step_f32c (from the records again).
__device__ __forceinline__ float face_triple_of(const FaceRec32& ra, const FaceRec32& rb,
                                                const FaceRec32& rc, const Prod3& t,
                                                float dp_min, float ox, float oy, float oz,
                                                float kappa) {
  const AngleSum s = angle_sum_of(t.re, t.im);
  if (__builtin_expect(wave_all_group(dp_min, s), 1)) return atan_sum_f32(s.x);
  asm volatile("" ::);  // keep a real wave-uniform branch
  float one[1] = {-0.0f};
  const float qx[1] = {ox}, qy[1] = {oy}, qz[1] = {oz};
  face_step_f32c<1>(ra, qx, qy, qz, kappa, one);
  face_step_f32c<1>(rb, qx, qy, qz, kappa, one);
  face_step_f32c<1>(rc, qx, qy, qz, kappa, one);
  return one[0];
}

// Six faces for QPL queries: the FP32 sum of their angles into sum[k], as the
// argument of the product of their (den + i det) when every lane meets
// wave_all_sum6 (one reciprocal and one polynomial for the six), otherwise
// as two groups of three (face_triple_of). Only the two triples' products
// and smallest dp stay live; a group that fails rebuilds its faces' geometry
// from the records.
template <int QPL>
__device__ __forceinline__ void face_hexad_f32c(const FaceRec32* rec, const float (&ox)[QPL],
                                                const float (&oy)[QPL], const float (&oz)[QPL],
                                                float kappa, float (&sum)[QPL]) {
#pragma clang fp contract(off)  // the two triples' sums round before they add
#pragma unroll
  for (int k = 0; k < QPL; ++k) {
    const PairGeom32 g1 = pair_geom32(rec[0].r0, rec[0].r1, rec[0].r2, rec[0].r3, ox[k], oy[k], oz[k]);
    const PairGeom32 g2 = pair_geom32(rec[1].r0, rec[1].r1, rec[1].r2, rec[1].r3, ox[k], oy[k], oz[k]);
    const PairGeom32 g3 = pair_geom32(rec[2].r0, rec[2].r1, rec[2].r2, rec[2].r3, ox[k], oy[k], oz[k]);
    const Prod3 t1 = prod3_f32(g1.det, g1.den, g2.det, g2.den, g3.det, g3.den);
    const float m1 = __builtin_fminf(__builtin_fminf(g1.dp, g2.dp), g3.dp);
    const PairGeom32 g4 = pair_geom32(rec[3].r0, rec[3].r1, rec[3].r2, rec[3].r3, ox[k], oy[k], oz[k]);
    const PairGeom32 g5 = pair_geom32(rec[4].r0, rec[4].r1, rec[4].r2, rec[4].r3, ox[k], oy[k], oz[k]);
    const PairGeom32 g6 = pair_geom32(rec[5].r0, rec[5].r1, rec[5].r2, rec[5].r3, ox[k], oy[k], oz[k]);
    const Prod3 t2 = prod3_f32(g4.det, g4.den, g5.det, g5.den, g6.det, g6.den);
    const float m2 = __builtin_fminf(__builtin_fminf(g4.dp, g5.dp), g6.dp);
    const AngleSum s = sum6_f32(t1, t2);
    float part;
    if (__builtin_expect(wave_all_sum6(m1, m2, t1, t2, s), 1)) {
      part = atan_sum_f32(s.x);
    } else {
      asm volatile("" ::);  // keep a real wave-uniform branch
      part = face_triple_of(rec[0], rec[1], rec[2], t1, m1, ox[k], oy[k], oz[k], kappa) +
             face_triple_of(rec[3], rec[4], rec[5], t2, m2, ox[k], oy[k], oz[k], kappa);
    }
    sum[k] = part;
  }
}

// Six faces' product of their (den + i det) and its lane mask (hex_of), the
// geometry computed from the records for one query offset.
__device__ __forceinline__ Hex32 face_hex_product(const FaceRec32* rec, float ox, float oy,
                                                  float oz) {
  const PairGeom32 g1 = pair_geom32(rec[0].r0, rec[0].r1, rec[0].r2, rec[0].r3, ox, oy, oz);
  const PairGeom32 g2 = pair_geom32(rec[1].r0, rec[1].r1, rec[1].r2, rec[1].r3, ox, oy, oz);
  const PairGeom32 g3 = pair_geom32(rec[2].r0, rec[2].r1, rec[2].r2, rec[2].r3, ox, oy, oz);
  const Prod3 t1 = prod3_f32(g1.det, g1.den, g2.det, g2.den, g3.det, g3.den);
  const float m1 = __builtin_fminf(__builtin_fminf(g1.dp, g2.dp), g3.dp);
  const PairGeom32 g4 = pair_geom32(rec[3].r0, rec[3].r1, rec[3].r2, rec[3].r3, ox, oy, oz);
  const PairGeom32 g5 = pair_geom32(rec[4].r0, rec[4].r1, rec[4].r2, rec[4].r3, ox, oy, oz);
  const PairGeom32 g6 = pair_geom32(rec[5].r0, rec[5].r1, rec[5].r2, rec[5].r3, ox, oy, oz);
  const Prod3 t2 = prod3_f32(g4.det, g4.den, g5.det, g5.den, g6.det, g6.den);
  const float m2 = __builtin_fminf(__builtin_fminf(g4.dp, g5.dp), g6.dp);
  return hex_of(t1, m1, t2, m2);
}

// Three faces rebuilt from the records: their product and smallest dp, then
// face_triple_of's test and fallback.
__device__ __forceinline__ float face_triple_rebuilt(const FaceRec32* rec, float ox, float oy,
                                                     float oz, float kappa) {
  const PairGeom32 g1 = pair_geom32(rec[0].r0, rec[0].r1, rec[0].r2, rec[0].r3, ox, oy, oz);
  const PairGeom32 g2 = pair_geom32(rec[1].r0, rec[1].r1, rec[1].r2, rec[1].r3, ox, oy, oz);
  const PairGeom32 g3 = pair_geom32(rec[2].r0, rec[2].r1, rec[2].r2, rec[2].r3, ox, oy, oz);
  const Prod3 t = prod3_f32(g1.det, g1.den, g2.det, g2.den, g3.det, g3.den);
  const float m = __builtin_fminf(__builtin_fminf(g1.dp, g2.dp), g3.dp);
  return face_triple_of(rec[0], rec[1], rec[2], t, m, ox, oy, oz, kappa);
}

// Six faces whose product h was formed: their FP32 angle sum when every lane
// meets the six-group test, else as two triples rebuilt from the records.
__device__ __forceinline__ float face_hexad_of(const FaceRec32* rec, const Hex32& h, float ox,
                                               float oy, float oz, float kappa) {
#pragma clang fp contract(off)  // the two triples' sums round before they add
  float x;
  if (__builtin_expect(wave_all_hex(h, x), 1)) return atan_sum_f32(x);
  asm volatile("" ::);  // keep a real wave-uniform branch
  return face_triple_rebuilt(rec, ox, oy, oz, kappa) +
         face_triple_rebuilt(rec + 3, ox, oy, oz, kappa);
}

// Twelve faces for QPL queries: the FP32 sum of their angles into sum[k], as
// the argument of the product of their (den + i det) when every lane meets
// both six-groups' conditions, 1/re positive normal and |x| < 0.4 (one
// reciprocal and one polynomial for the twelve), otherwise as two groups of
// six (face_hexad_of). Only each six's (re, im) and lane mask stay live
// across the second six.
template <int QPL>
__device__ __forceinline__ void face_dodecad_f32c(const FaceRec32* rec, const float (&ox)[QPL],
                                                  const float (&oy)[QPL], const float (&oz)[QPL],
                                                  float kappa, float (&sum)[QPL]) {
#pragma clang fp contract(off)  // the two sixes' sums round before they add
#pragma unroll
  for (int k = 0; k < QPL; ++k) {
    const Hex32 h1 = face_hex_product(rec, ox[k], oy[k], oz[k]);
    const Hex32 h2 = face_hex_product(rec + 6, ox[k], oy[k], oz[k]);
    float x, part;
    if (__builtin_expect(wave_all_dodecad(h1, h2, x), 1)) {
      part = atan_sum_f32(x);
    } else {
      asm volatile("" ::);  // keep a real wave-uniform branch
      part = face_hexad_of(rec, h1, ox[k], oy[k], oz[k], kappa) +
             face_hexad_of(rec + 6, h2, ox[k], oy[k], oz[k], kappa);
    }
    sum[k] = part;
  }
}

// Faces [f0, f1) for QPL queries per lane; o (ox0 ..) the queries' block-local
// FP32 offsets in the frame 2^(2 - er) (block_origin). Each tile is staged in
// its own power-of-two frame sc = 2^(-E), 2^E bounding |corner - g| over the
// tile and |query - g| over the block (so that |corner - query| < 2 sqrt(3)
// ~ 3.5 in it, and a product of twelve faces' (den + i det), each
// den <= 4 la lb lc < 2^7.4, stays within the FP32 range), and the offsets
// are rescaled into it by the exact factor 2^(er - E - 2) <= 1/4 (see
// frame_exponent).
template <int QPL>
__device__ __forceinline__ void sweep_f32c(const double* __restrict__ tri, int64_t stride,
                                           int64_t f0, int64_t f1, double gx, double gy,
                                           double gz, int er, const float (&ox0)[QPL],
                                           const float (&oy0)[QPL], const float (&oz0)[QPL],
                                           float kappa, double (&acc)[QPL], SweepF32Smem& sm) {
  const int tid = threadIdx.x;
  for (int64_t t0 = f0; t0 < f1; t0 += kFaceTile) {
    const int cnt = (int)(f1 - t0 < kFaceTile ? f1 - t0 : kFaceTile);
    // corners relative to g in FP64 (exact rounding of the difference, as
    // before any scaling); threads without a face hold zeros
    double d[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (tid < cnt) {
      const int64_t f = t0 + tid;
#pragma unroll
      for (int k = 0; k < 9; ++k) d[k] = tri[k * stride + f] - (k % 3 == 0 ? gx : k % 3 == 1 ? gy : gz);
    }
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < 9; ++k) m = max(m, abs_hi_word(d[k]));
    m = wave_max_uniform_u32(m);
    if ((tid & (kWave - 1)) == 0) sm.tmax[tid / kWave] = m;
    __syncthreads();  // the previous tile is swept; sm.tmax holds this tile's extents
    uint32_t mt = sm.tmax[0];
#pragma unroll
    for (int w = 1; w < kBlockThreads / kWave; ++w) mt = max(mt, sm.tmax[w]);
    mt = (uint32_t)__builtin_amdgcn_readfirstlane((int)mt);
    const int ef = frame_exponent_hi(mt) + 1;
    const int E = ef > er ? ef : er;
    const double sc = pow2_f64(-E);
    if (tid < cnt)
      sm.face[tid] = make_face_rec32(d[0] * sc, d[1] * sc, d[2] * sc, d[3] * sc, d[4] * sc,
                                     d[5] * sc, d[6] * sc, d[7] * sc, d[8] * sc, 0.0, 0.0, 0.0);
    __syncthreads();
    const float fs = pow2_f32(er - E - 2);  // exact rescaling: a power of two <= 1/4
    float ox[QPL], oy[QPL], oz[QPL];
#pragma unroll
    for (int k = 0; k < QPL; ++k) {
      ox[k] = ox0[k] * fs;
      oy[k] = oy0[k] * fs;
      oz[k] = oz0[k] * fs;
    }
    // the angles of twelve consecutive faces are summed in FP32 (the
    // argument of their product, or two groups of six, each the argument of
    // its product or two groups of three: a triple's product or three FP32
    // angles added; at most a few ulps of their magnitudes, the bound of a
    // compensated sum) and joined to the FP64 sum once
    int j = 0;
    for (; j + 12 <= cnt; j += 12) {
      float s12[QPL];
      face_dodecad_f32c<QPL>(&sm.face[j], ox, oy, oz, kappa, s12);
#pragma unroll
      for (int k = 0; k < QPL; ++k) acc[k] += (double)s12[k];
    }
    for (; j + 6 <= cnt; j += 6) {
      float s6[QPL];
      face_hexad_f32c<QPL>(&sm.face[j], ox, oy, oz, kappa, s6);
#pragma unroll
      for (int k = 0; k < QPL; ++k) acc[k] += (double)s6[k];
    }
    for (; j + 3 <= cnt; j += 3) {
      float s1[QPL];
      face_triple_f32c<QPL>(sm.face[j], sm.face[j + 1], sm.face[j + 2], ox, oy, oz, kappa, s1);
#pragma unroll
      for (int k = 0; k < QPL; ++k) acc[k] += (double)s1[k];
    }
    if (j < cnt) {
      float part[QPL];
#pragma unroll
      for (int k = 0; k < QPL; ++k) part[k] = -0.0f;
      for (; j < cnt; ++j) face_step_f32c<QPL>(sm.face[j], ox, oy, oz, kappa, part);
#pragma unroll
      for (int k = 0; k < QPL; ++k) acc[k] += (double)part[k];
    }
  }
  __syncthreads();
}

}  // namespace wn
