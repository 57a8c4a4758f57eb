// FP32-compensated forward over triangle-strip records (records built by
// wn_strips.cpp; format documented there).
//
// Same arithmetic as fwd_f32c_kernel (block-local origin, FP32 pairs, FP64
// accumulation, far-pair fast path, FP64 re-evaluation of ill-conditioned
// pairs, the fast path on |det / den| < 1/8 with a 3-term arctangent),
// except that the angles of twelve consecutive records are summed in FP32
// before each FP64 add (error a few ulps of the twelve angles' magnitudes,
// the bound of a compensated sum, with a twelfth of the FP64 instructions),
// and
// the re-evaluated angle is rounded to FP32 once before that sum. Faces
// arrive as a strip: record k brings vertex w_k and closes
// face (w_{k-2}, w_{k-1}, w_k). Per face the lane computes one offset, one
// squared length and one square root (the other two vertices are kept in
// registers from the previous records) and two law-of-cosines dot products
// (the third is carried), instead of three of each.
//
// LDS tile (255 records, a multiple of 3 so the three register slots rotate
// statically through a 3-way unrolled loop):
//   R0[j] = (w.x, w.y, w.z, -|w_j - w_{j-1}|^2 / 2)  (block-local FP32 coords)
//   R1[j] = (-|w_j - w_{j-2}|^2 / 2, N.x, N.y, N.z)  (N = face normal)
// in the tile's power-of-two frame (frame_exponent in wn_device.h: here the
// block-local coordinates times 2^(-E), |corner - query| < 2 sqrt(3) ~ 3.5
// in it, so that the products of twelve faces below stay in the FP32 range
// at any input scale);
// so that a dot product is one add and one fma with 0.5 (law of cosines;
// the halving is exact) and den is formed undoubled.
// The face orientation (record flag: 0 = no face ends at j, +1 / -1 = face
// orientation vs strip order) is folded into N; a record with flag 0 is
// staged with N = 0 and R1.x = +2^20, which makes its angle 0 without a
// branch. The flag itself is read (from the records, at a wave-uniform
// address) only on the rare slow path.
// Fast path: for each group of twelve records, the sum of their angles is the
// argument of the product of the faces' (den + i det) (two sixes, each two
// triple products): one reciprocal and one 5-term polynomial per twelve faces
// instead of twelve of each. A group that fails its test takes each six by its
// own (strip_hexad_of), a six that fails is rebuilt from the LDS tile as two
// groups of three (sum3_f32), and a group of three that fails goes face by
// face; the rebuilt faces come from the LDS tile, so the fast path keeps no
// per-face values live.
// Squared lengths and N are computed in FP64 from the rounded FP32 corners.

#include <algorithm>
#include <map>
#include <mutex>

#include "wn_device.h"
#include "wn_internal.h"
#include "wn_sweep.h"

namespace wn {

struct StripSmem {
  float4 r0[kStripTileRecords];
  float4 r1[kStripTileRecords];
  double red[6][kBlockThreads / kWave];
  uint32_t tmax[kBlockThreads / kWave];  // per wave: abs_hi_word max |corner - g| of the next tile
};

namespace {

struct VSlot {
  float ax, ay, az, l2, l;
};


// One tile's staging, in two steps around a block barrier. strip_tile_extent:
// every thread of the block (all lanes: the wave reduction needs them) reads
// its record j, and its wave's max over the tile of |corner - g| (infinity
// norm) goes to sm.tmax. After the barrier, strip_tile_store reads records
// j, j-1, j-2 (the record loads of the first step, again: nothing is held
// across the barrier), scales the tile into the frame sc = 2^(-E)
// (frame_exponent: 2^E bounds |corner - g| over the tile and, through er,
// |query - g| over the block) and writes the LDS records; it returns the
// factor 2^(er - E) that takes the lanes' offsets (staged at 2^(-er))
// into the same frame.
__device__ __forceinline__ void strip_tile_extent(const double4* __restrict__ rec, int64_t base,
                                                  int cnt, double gx, double gy, double gz,
                                                  int32_t wbase, StripSmem& sm) {
  const int j = thread_index(wbase);
  uint32_t m = 0;
  if (j < cnt) {
    const double4 c = rec[base + j];
    const uint32_t mx = abs_hi_word(c.x - gx), my = abs_hi_word(c.y - gy),
                   mz = abs_hi_word(c.z - gz);
    m = max(max(mx, my), mz);
  }
  m = wave_max_uniform_u32(m);
  if ((j & (kWave - 1)) == 0) sm.tmax[j / kWave] = m;
}

__device__ __forceinline__ float strip_tile_store(const double4* __restrict__ rec, int64_t base,
                                                  int cnt, double gx, double gy, double gz,
                                                  int er, int32_t wbase, StripSmem& sm) {
  uint32_t mt = sm.tmax[0];
#pragma unroll
  for (int w = 1; w < kBlockThreads / kWave; ++w) mt = max(mt, sm.tmax[w]);
  mt = (uint32_t)__builtin_amdgcn_readfirstlane((int)mt);
  const int ef = frame_exponent_hi(mt) + 1;
  const int E = ef > er ? ef : er;
  const double sc = pow2_f64(-E);
  const int j = thread_index(wbase);
  if (j < cnt) {
    const double4 c = rec[base + j];
    const double4 p = rec[base + (j >= 1 ? j - 1 : j)];
    const double4 pp = rec[base + (j >= 2 ? j - 2 : j)];
    const float cx = (float)((c.x - gx) * sc), cy = (float)((c.y - gy) * sc),
                cz = (float)((c.z - gz) * sc);
    const float px = (float)((p.x - gx) * sc), py = (float)((p.y - gy) * sc),
                pz = (float)((p.z - gz) * sc);
    const float qx = (float)((pp.x - gx) * sc), qy = (float)((pp.y - gy) * sc),
                qz = (float)((pp.z - gz) * sc);
    // FP64 on the rounded corners: w_k - w_{k-1}, w_k - w_{k-2}, w_{k-1} - w_{k-2}
    const double d1x = (double)cx - px, d1y = (double)cy - py, d1z = (double)cz - pz;
    const double d2x = (double)cx - qx, d2y = (double)cy - qy, d2z = (double)cz - qz;
    const double ex = (double)px - qx, ey = (double)py - qy, ez = (double)pz - qz;
    // N = (w_{k-1} - w_{k-2}) x (w_k - w_{k-2}), oriented by the flag: the
    // lane's dot product with it is det
    // A record that closes no face (flag 0) gets N = 0 and +2^20 in place of
    // -|w_k - w_{k-2}|^2 / 2: det is then 0 exactly and c.a ~ 2^20 makes dp
    // and den positive and finite (see strip_record), so both fast paths give
    // it angle 0 (0 / den, and a real factor den + 0i of a group product).
    const double f = c.w;
    sm.r0[j] = make_float4(cx, cy, cz, (float)(-0.5 * (d1x * d1x + d1y * d1y + d1z * d1z)));
    sm.r1[j] = make_float4(c.w == 0.0 ? 1048576.0f
                                      : (float)(-0.5 * (d2x * d2x + d2y * d2y + d2z * d2z)),
                           (float)(f * (ey * d2z - ez * d2y)), (float)(f * (ez * d2x - ex * d2z)),
                           (float)(f * (ex * d2y - ey * d2x)));
  }
  return pow2_f32(er - E);
}

// FP64 re-evaluation of face (j-2, j-1, j) from the LDS corners, oriented.
__device__ __noinline__ double strip_pair_f64(const StripSmem& sm, int j, bool flip, float ox,
                                              float oy, float oz) {
  const float4 a = sm.r0[j - 2], b = sm.r0[j - 1], c = sm.r0[j];
  const double qx = ox, qy = oy, qz = oz;
  const double v = pair_f64((double)a.x - qx, (double)a.y - qy, (double)a.z - qz,
                            (double)b.x - qx, (double)b.y - qy, (double)b.z - qz,
                            (double)c.x - qx, (double)c.y - qy, (double)c.z - qz);
  return flip ? -v : v;
}

// Offset, squared length and length of the lane's query to record vertex
// r0 (block-local FP32 coordinates).
__device__ __forceinline__ void strip_vertex(const float4& r0, float ox, float oy, float oz,
                                             VSlot& R) {
  R.ax = r0.x - ox;
  R.ay = r0.y - oy;
  R.az = r0.z - oz;
  R.l2 = __builtin_fmaf(R.az, R.az, __builtin_fmaf(R.ay, R.ay, R.ax * R.ax));
  R.l = __builtin_amdgcn_sqrtf(R.l2);
}

// Face (a, b, c) = (P, Q, R) closed by a record with tile data r0, r1; a.b
// = s_carry (carried from the previous record), b.c = s_qr (carried to the
// next), c.a from r1.x (law of cosines); den = la lb lc + dp with
// dp = lc a.b + la b.c + lb c.a.
struct StripGeom {
  float det, den, dp, pq, s_qr;
};

__device__ __forceinline__ StripGeom strip_geom(const float4& r0, const float4& r1, const VSlot& P,
                                                const VSlot& Q, const VSlot& R, float s_carry) {
  StripGeom g;
  g.s_qr = __builtin_fmaf(Q.l2 + R.l2, 0.5f, r0.w);
  const float s_pr = __builtin_fmaf(P.l2 + R.l2, 0.5f, r1.x);
  g.det = __builtin_fmaf(r1.w, R.az, __builtin_fmaf(r1.z, R.ay, r1.y * R.ax));
  g.pq = P.l * Q.l;
  g.dp = __builtin_fmaf(R.l, s_carry, __builtin_fmaf(P.l, g.s_qr, Q.l * s_pr));
  g.den = __builtin_fmaf(g.pq, R.l, g.dp);
  return g;
}

// Record j of the tile closes face (P, Q, R) (P and Q from the two previous
// records, R computed here): its FP32 geometry. Every record takes the same
// instructions: one that closes no face (flag 0: the first two of a strip,
// tile padding) was staged (strip_tile_store) with N = 0 and c.a = +2^20,
// so det = 0 exactly and, with Q's length above ~1e-4 of the frame's unit
// (|corner - query| < 3.5 in it), dp and den are positive: its angle is 0
// on the fast paths (a real factor of at most 2^21.8 in a group's product:
// two of them and ten far faces' 4 la lb lc <= 2^7.4 stay below 2^126).
// boff: byte offset of record j in r0 and r1, held in a VGPR (see strip_sweep).
__device__ __forceinline__ StripGeom strip_record(const StripSmem& sm, int boff, const VSlot& P,
                                                  const VSlot& Q, VSlot& R, float& s_carry,
                                                  float ox, float oy, float oz) {
  const float4 r0 = *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(sm.r0) + boff);
  strip_vertex(r0, ox, oy, oz, R);
  const float4 r1 = *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(sm.r1) + boff);
  const StripGeom g = strip_geom(r0, r1, P, Q, R, s_carry);
  s_carry = g.s_qr;
  return g;
}

// The FP32 angle of face j on its own (geometry g, rl = the length of its
// newest vertex): the small-angle arctangent when every lane is far
// (|det / den| < 1/8 and den > la lb lc), else atan2 in FP32 and, for an
// ill-conditioned pair, the FP64 angle rounded once. The record flag (raw
// bits of the record's w; read from global memory at a wave-uniform
// address) is looked at only here, off the fast path.
__device__ __forceinline__ float strip_angle(const StripSmem& sm,
                                             const long long* __restrict__ flags, int j,
                                             const StripGeom& g, float rl, float ox, float oy,
                                             float oz, float kappa) {
  const float det8 = 8.0f * g.det;  // exact
  float ang = atan_small_x8_f32(det8 * __builtin_amdgcn_rcpf(g.den));
  if (__builtin_expect(!wave_all_far_x8(det8, g.den, g.dp), 0)) {
    asm volatile("" ::);  // keep a real wave-uniform branch
    const long long flag_bits = flags[4 * j];
    if (flag_bits == 0) {
      ang = 0.0f;
    } else {
      float mx;
      ang = atan2_f32(g.det, g.den, mx);
      // ill-conditioned (mx < kappa x la lb lc): the FP64 angle, rounded
      // once, replaces the FP32 one. Flag -1.0 has the sign bit set: the
      // face runs against the strip order.
      if (mx < kappa * (g.pq * rl))
        ang = (float)strip_pair_f64(sm, j, flag_bits < 0, ox, oy, oz);
    }
  }
  return ang;
}

// Records j, j+1, j+2 (slots A, B, C rotating as P, Q, R): the FP32 sum of
// their three angles. When every lane meets sum3_f32's conditions the sum is
// the argument of the product of the faces' (den + i det), one reciprocal
// and one polynomial for the three faces; otherwise each face takes
// strip_angle and the three FP32 angles are added (-0 + x == x: the first
// add folds away). Used for the last three records of a tile.
__device__ __forceinline__ float strip_triple(const StripSmem& sm,
                                              const long long* __restrict__ flags, int j,
                                              int boff, VSlot& A, VSlot& B, VSlot& C,
                                              float& s_carry, float ox, float oy, float oz,
                                              float kappa) {
#pragma clang fp contract(off)  // part += angle rounds the angle first
  const StripGeom g1 = strip_record(sm, boff, A, B, C, s_carry, ox, oy, oz);
  const StripGeom g2 = strip_record(sm, boff + 16, B, C, A, s_carry, ox, oy, oz);
  const StripGeom g3 = strip_record(sm, boff + 32, C, A, B, s_carry, ox, oy, oz);
  const AngleSum s = sum3_f32(g1.det, g1.den, g2.det, g2.den, g3.det, g3.den);
  if (__builtin_expect(wave_all_sum3(g1.dp, g2.dp, g3.dp, s), 1)) return atan_sum_f32(s.x);
  asm volatile("" ::);  // keep a real wave-uniform branch
  float part = -0.0f;
  part += strip_angle(sm, flags, j, g1, C.l, ox, oy, oz, kappa);
  part += strip_angle(sm, flags, j + 1, g2, A.l, ox, oy, oz, kappa);
  part += strip_angle(sm, flags, j + 2, g3, B.l, ox, oy, oz, kappa);
  return part;
}

// The angle of record j's face on its own, its geometry rebuilt from the
// LDS tile (corners j-2, j-1, j). The operations are strip_record's, in
// its order (a.b as the previous record formed it), so the values are those
// the sweep computed. A record with flag 0 closes no face: angle 0. Records
// 0 and 1 of a tile have flag 0 (every tile starts a strip).
__device__ __forceinline__ float strip_face_lds(const StripSmem& sm,
                                                const long long* __restrict__ flags, int j,
                                                float ox, float oy, float oz, float kappa) {
  if (flags[4 * j] == 0) return 0.0f;  // wave-uniform
  const float4 p = sm.r0[j - 2], q = sm.r0[j - 1], c = sm.r0[j];
  VSlot P, Q, R;
  strip_vertex(p, ox, oy, oz, P);
  strip_vertex(q, ox, oy, oz, Q);
  strip_vertex(c, ox, oy, oz, R);
  const float s_pq = __builtin_fmaf(P.l2 + Q.l2, 0.5f, q.w);
  const StripGeom g = strip_geom(c, sm.r1[j], P, Q, R, s_pq);
  return strip_angle(sm, flags, j, g, R.l, ox, oy, oz, kappa);
}

// Records j, j+1, j+2 whose product t (prod3_f32) and smallest dp the sweep
// already formed: the group's FP32 angle sum by sum3_f32's test, else face
// by face from the tile.
__device__ __forceinline__ float strip_triple_lds(const StripSmem& sm,
                                                  const long long* __restrict__ flags, int j,
                                                  const Prod3& t, float dp_min, float ox,
                                                  float oy, float oz, float kappa) {
#pragma clang fp contract(off)  // part += angle rounds the angle first
  const AngleSum s = angle_sum_of(t.re, t.im);
  if (__builtin_expect(wave_all_group(dp_min, s), 1)) return atan_sum_f32(s.x);
  asm volatile("" ::);  // keep a real wave-uniform branch
  float part = -0.0f;
  part += strip_face_lds(sm, flags, j, ox, oy, oz, kappa);
  part += strip_face_lds(sm, flags, j + 1, ox, oy, oz, kappa);
  part += strip_face_lds(sm, flags, j + 2, ox, oy, oz, kappa);
  return part;
}

// The geometry of record j's face rebuilt from the LDS tile (corners j-2,
// j-1, j; j >= 2) with strip_record's operations in its order (a.b as the
// previous record formed it): the values the sweep computed. rl: the length
// of the newest corner.
__device__ __forceinline__ StripGeom strip_geom_lds(const StripSmem& sm, int j, float ox,
                                                    float oy, float oz, float& rl) {
  const float4 p = sm.r0[j - 2], q = sm.r0[j - 1], c = sm.r0[j];
  VSlot P, Q, R;
  strip_vertex(p, ox, oy, oz, P);
  strip_vertex(q, ox, oy, oz, Q);
  strip_vertex(c, ox, oy, oz, R);
  const float s_pq = __builtin_fmaf(P.l2 + Q.l2, 0.5f, q.w);
  rl = R.l;
  return strip_geom(c, sm.r1[j], P, Q, R, s_pq);
}

// Records j, j+1, j+2 rebuilt from the tile: the product of their faces'
// (den + i det) and the smallest dp, a record that closes no face (flag 0;
// read at wave-uniform addresses) taking the factor 1 and dp 1 (its angle is
// 0). Used when a group of six that the sweep formed fails its test.
__device__ __forceinline__ void strip_triple_rebuild(const StripSmem& sm,
                                                     const long long* __restrict__ flags,
                                                     int j, float ox, float oy, float oz,
                                                     Prod3& t, float& dp_min) {
  float det[3], den[3], dp[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    det[i] = 0.0f;
    den[i] = 1.0f;
    dp[i] = 1.0f;
    if (flags[4 * (j + i)] != 0) {
      float rl;
      const StripGeom g = strip_geom_lds(sm, j + i, ox, oy, oz, rl);
      det[i] = g.det;
      den[i] = g.den;
      dp[i] = g.dp;
    }
  }
  t = prod3_f32(det[0], den[0], det[1], den[1], det[2], den[2]);
  dp_min = __builtin_fminf(__builtin_fminf(dp[0], dp[1]), dp[2]);
}

// Six records' product of their faces' (den + i det) and its lane mask
// (hex_of).
__device__ __forceinline__ Hex32 strip_hex_product(const StripSmem& sm, int boff, VSlot& A,
                                                   VSlot& B, VSlot& C, float& s_carry, float ox,
                                                   float oy, float oz) {
  const StripGeom g1 = strip_record(sm, boff, A, B, C, s_carry, ox, oy, oz);
  const StripGeom g2 = strip_record(sm, boff + 16, B, C, A, s_carry, ox, oy, oz);
  const StripGeom g3 = strip_record(sm, boff + 32, C, A, B, s_carry, ox, oy, oz);
  const Prod3 t1 = prod3_f32(g1.det, g1.den, g2.det, g2.den, g3.det, g3.den);
  const float m1 = __builtin_fminf(__builtin_fminf(g1.dp, g2.dp), g3.dp);
  const StripGeom g4 = strip_record(sm, boff + 48, A, B, C, s_carry, ox, oy, oz);
  const StripGeom g5 = strip_record(sm, boff + 64, B, C, A, s_carry, ox, oy, oz);
  const StripGeom g6 = strip_record(sm, boff + 80, C, A, B, s_carry, ox, oy, oz);
  const Prod3 t2 = prod3_f32(g4.det, g4.den, g5.det, g5.den, g6.det, g6.den);
  const float m2 = __builtin_fminf(__builtin_fminf(g4.dp, g5.dp), g6.dp);
  return hex_of(t1, m1, t2, m2);
}

// Six records j .. j+5 whose product h the sweep formed: their FP32 angle sum
// when every lane meets the six-group test (h.ok, 1/re positive normal and
// |x| < 0.4), else as two triples rebuilt from the tile (strip_triple_lds).
__device__ __forceinline__ float strip_hexad_of(const StripSmem& sm,
                                                const long long* __restrict__ flags, int j,
                                                const Hex32& h, float ox, float oy, float oz,
                                                float kappa) {
#pragma clang fp contract(off)  // the two triples' sums round before they add
  float x;
  if (__builtin_expect(wave_all_hex(h, x), 1)) return atan_sum_f32(x);
  asm volatile("" ::);  // keep a real wave-uniform branch
  Prod3 t1, t2;
  float m1, m2;
  strip_triple_rebuild(sm, flags, j, ox, oy, oz, t1, m1);
  strip_triple_rebuild(sm, flags, j + 3, ox, oy, oz, t2, m2);
  return strip_triple_lds(sm, flags, j, t1, m1, ox, oy, oz, kappa) +
         strip_triple_lds(sm, flags, j + 3, t2, m2, ox, oy, oz, kappa);
}

// Records j .. j+11 (four turns of the slot rotation): the FP32 sum of their
// twelve angles, as the argument of the product of their (den + i det) when
// every lane meets both six-groups' conditions, 1/re positive normal and
// |x| < 0.4 (one reciprocal and one polynomial for the twelve), else as two
// groups of six (strip_hexad_of). Only each six's (re, im) and its lane mask
// stay live across the second six.
__device__ __forceinline__ float strip_dodecad(const StripSmem& sm,
                                               const long long* __restrict__ flags, int j,
                                               int boff, VSlot& A, VSlot& B, VSlot& C,
                                               float& s_carry, float ox, float oy, float oz,
                                               float kappa) {
#pragma clang fp contract(off)  // the two sixes' sums round before they add
  const Hex32 h1 = strip_hex_product(sm, boff, A, B, C, s_carry, ox, oy, oz);
  const Hex32 h2 = strip_hex_product(sm, boff + 96, A, B, C, s_carry, ox, oy, oz);
  float x;
  if (__builtin_expect(wave_all_dodecad(h1, h2, x), 1)) return atan_sum_f32(x);
  asm volatile("" ::);  // keep a real wave-uniform branch
  return strip_hexad_of(sm, flags, j, h1, ox, oy, oz, kappa) +
         strip_hexad_of(sm, flags, j + 6, h2, ox, oy, oz, kappa);
}

// Block-local origin: centre of the bbox of the block's valid queries (the
// same reduction as the face-record kernels), and er = frame_exponent of the
// bbox's largest side (|query - g| < 2^er). Begins and ends with a barrier.
__device__ __forceinline__ void strip_block_origin(bool valid, double Ox, double Oy, double Oz,
                                                   StripSmem& sm, double& gx, double& gy,
                                                   double& gz, int& er) {
  const int tid = threadIdx.x;
  double lo0 = valid ? Ox : __builtin_inf(), hi0 = valid ? Ox : -__builtin_inf();
  double lo1 = valid ? Oy : __builtin_inf(), hi1 = valid ? Oy : -__builtin_inf();
  double lo2 = valid ? Oz : __builtin_inf(), hi2 = valid ? Oz : -__builtin_inf();
  lo0 = wave_min_uniform_f64(lo0); lo1 = wave_min_uniform_f64(lo1); lo2 = wave_min_uniform_f64(lo2);
  hi0 = wave_max_uniform_f64(hi0); hi1 = wave_max_uniform_f64(hi1); hi2 = wave_max_uniform_f64(hi2);
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
  // every lane combined the same values in the same order: the origin is
  // block-uniform, so it lives in scalar registers for the rest of the sweep
  gx = uniform_f64(0.5 * (lo0 + hi0));
  gy = uniform_f64(0.5 * (lo1 + hi1));
  gz = uniform_f64(0.5 * (lo2 + hi2));
  er = frame_exponent(uniform_f64(fmax(fmax(hi0 - lo0, hi1 - lo1), hi2 - lo2)));
  __syncthreads();
}

// Unscaled angle sum of one query over strip tiles [t_begin, t_end); o is
// the query's block-local FP32 offset in the frame 2^(-er)
// (strip_block_origin), rescaled into each tile's frame (strip_tile_store).
// All threads of the block call it together; wbase = wave_base(). The
// staging rebuilds the thread index from wbase on every tile (thread_index),
// so no per-lane index or address stays live across the sweep (with the
// sweep's 64 registers, such values were spilled to scratch per block).
__device__ __forceinline__ double strip_sweep(const double4* __restrict__ rec, int64_t t_begin,
                                              int64_t t_end, double gx, double gy, double gz,
                                              int er, float ox0, float oy0, float oz0,
                                              float kappa, int32_t wbase, StripSmem& sm) {
  double acc = 0.0;
  VSlot A{0.f, 0.f, 0.f, 0.f, 0.f}, B = A, C = A;
  float s_carry = 0.0f;
  for (int64_t t = t_begin; t < t_end; ++t) {
    strip_tile_extent(rec, t * kStripTileRecords, kStripTileRecords, gx, gy, gz, wbase, sm);
    __syncthreads();  // the previous tile is swept; sm.tmax holds this tile's extents
    const float f = strip_tile_store(rec, t * kStripTileRecords, kStripTileRecords, gx, gy, gz,
                                     er, wbase, sm);
    __syncthreads();
    // exact: a power of two no larger than 1
    const float ox = ox0 * f, oy = oy0 * f, oz = oz0 * f;
    const long long* flags = reinterpret_cast<const long long*>(rec + t * kStripTileRecords) + 3;
    // The record offset is wave-uniform, but an LDS read takes its address
    // from a VGPR: a VGPR byte offset advanced once per twelve records and read
    // with immediate offsets. (The empty asm after the increment keeps it one
    // VGPR updated in place: placed before the use, it cost a copy per
    // iteration.)
    int boff = 0;
    asm volatile("" : "+v"(boff));
    // groups of twelve records, each group's FP32 angle sum joined to the
    // FP64 sum once; 255 = 21 x 12 + 3, the last three records as a group of
    // three
    int j = 0;
    for (; j + 12 <= kStripTileRecords; j += 12) {
      acc += (double)strip_dodecad(sm, flags, j, boff, A, B, C, s_carry, ox, oy, oz, kappa);
      boff += 12 * (int)sizeof(float4);
      asm volatile("" : "+v"(boff));
    }
    acc += (double)strip_triple(sm, flags, j, boff, A, B, C, s_carry, ox, oy, oz, kappa);
  }
  __syncthreads();
  return acc;
}

// One block per query block, or per (query block, face split) in the tail.
// Blocks [0, whole) sweep query block b over every tile and write w through
// perm. The remaining query blocks (the tail: one round of the device's
// resident blocks) are split over `splits` tile ranges, split-major: block
// whole + l takes split l / ntail_qb of tail query block l % ntail_qb, so
// that the blocks running together sweep the same tiles (the order of a
// (query block, split) grid; measured 4 % faster than dispatching one query
// block's splits together). Tail blocks write unscaled partials
// partial[split * ntail + (slot - 256 whole)], ntail = nq - 256 whole, summed
// in split order afterwards. The whole blocks fill the first round and the
// short split blocks the last one: the end is as even as with every query
// block split, for half the partials.
__global__ __launch_bounds__(kBlockThreads) __attribute__((amdgpu_waves_per_eu(8))) void fwd_strip_kernel(
    const double4* __restrict__ rec, int64_t ntiles, const double* __restrict__ q,
    const int32_t* __restrict__ perm, double* __restrict__ dst, double* __restrict__ partial,
    int64_t nq, int32_t whole, int64_t tiles_per_split, int splits, float kappa) {
  __shared__ StripSmem sm;
  const int32_t nqb = (int32_t)((nq + kBlockThreads - 1) / kBlockThreads);
  int32_t qb = (int32_t)blockIdx.x, split = -1;
  if (qb >= whole) {
    const int32_t l = qb - whole, tail = nqb - whole;
    qb = whole + l % tail;
    split = l / tail;
  }
  // 32-bit slot (nq < 2^31), and the output index read only after the sweep:
  // nothing 64-bit stays live across the sweep (it would be spilled)
  const int32_t base = wave_base();
  const int32_t slot = qb * kBlockThreads + (int32_t)threadIdx.x;
  const bool valid = slot < nq;
  // q is in launch order (gathered through perm beforehand when sorted)
  const double Ox = valid ? q[3 * (int64_t)slot] : 0.0;
  const double Oy = valid ? q[3 * (int64_t)slot + 1] : 0.0;
  const double Oz = valid ? q[3 * (int64_t)slot + 2] : 0.0;

  double gx, gy, gz;
  int er;
  strip_block_origin(valid, Ox, Oy, Oz, sm, gx, gy, gz, er);
  const double so = pow2_f64(-er);
  const float ox = valid ? (float)((Ox - gx) * so) : 0.0f;
  const float oy = valid ? (float)((Oy - gy) * so) : 0.0f;
  const float oz = valid ? (float)((Oz - gz) * so) : 0.0f;

  int64_t t_begin = 0, t_end = ntiles;
  if (split >= 0) {
    t_begin = (int64_t)split * tiles_per_split;
    t_end = t_begin + tiles_per_split < ntiles ? t_begin + tiles_per_split : ntiles;
  }
  const double acc = strip_sweep(rec, t_begin, t_end, gx, gy, gz, er, ox, oy, oz, kappa, base, sm);
  const int32_t oslot = qb * kBlockThreads + thread_index(base);
  if (oslot < nq) {
    if (split < 0)
      dst[perm ? (int64_t)perm[oslot] : (int64_t)oslot] = acc * kInv2Pi;
    else
      partial[(int64_t)split * (nq - (int64_t)whole * kBlockThreads) +
              (oslot - whole * kBlockThreads)] = acc;
  }
}

// Persistent strip kernel over a work list (mesh, query tile, tile chunk):
// the strip counterpart of the face-record batch kernel, with the same
// item protocol (atomic counter, block-uniform exit) and the same output
// rule (single-chunk meshes write w, others write unscaled chunk-major
// partials summed in chunk order afterwards).
__global__ __launch_bounds__(kBlockThreads) __attribute__((amdgpu_waves_per_eu(8))) void batch_strip_kernel(
    const double4* __restrict__ rec, const double* __restrict__ q,
    const int32_t* __restrict__ perm, double* __restrict__ dst, double* __restrict__ partial,
    const BatchMesh* __restrict__ meshes, const BatchItem* __restrict__ items, int64_t n_items,
    int32_t* __restrict__ counter, int64_t chunk_tiles, float kappa) {
  __shared__ StripSmem sm;
  __shared__ int32_t s_item;
  const int tid = threadIdx.x;
  const int32_t base = wave_base();
  for (;;) {
    __syncthreads();  // every thread has read the previous s_item
    if (tid == 0) s_item = atomicAdd(counter, 1);
    __syncthreads();
    // block-uniform: scalar registers, and the item and mesh records are
    // scalar loads (nothing per lane to keep live across the sweep)
    const int32_t it = __builtin_amdgcn_readfirstlane(s_item);
    if (it >= n_items) break;  // block-uniform exit
    const BatchItem item = items[it];
    const BatchMesh m = meshes[item.mesh];
    const int64_t slot = m.query_off + (int64_t)item.qtile * kBlockThreads + tid;
    const bool valid = slot < m.query_off + m.nq;
    // q is in launch order (gathered through perm beforehand when sorted)
    const double Ox = valid ? q[3 * slot] : 0.0;
    const double Oy = valid ? q[3 * slot + 1] : 0.0;
    const double Oz = valid ? q[3 * slot + 2] : 0.0;
    double gx, gy, gz;
    int er;
    strip_block_origin(valid, Ox, Oy, Oz, sm, gx, gy, gz, er);
    const double so = pow2_f64(-er);
    const float ox = valid ? (float)((Ox - gx) * so) : 0.0f;
    const float oy = valid ? (float)((Oy - gy) * so) : 0.0f;
    const float oz = valid ? (float)((Oz - gz) * so) : 0.0f;
    const int64_t t0 = m.face_off + (int64_t)item.chunk * chunk_tiles;
    int64_t t1 = t0 + chunk_tiles;
    if (t1 > m.face_off + m.nf) t1 = m.face_off + m.nf;
    const double acc = strip_sweep(rec, t0, t1, gx, gy, gz, er, ox, oy, oz, kappa, base, sm);
    const int32_t local = item.qtile * kBlockThreads + thread_index(base);  // slot - query_off
    if (local < m.nq) {
      const int64_t oslot = m.query_off + local;
      if (m.chunks == 1)
        dst[perm ? (int64_t)perm[oslot] : oslot] = acc * kInv2Pi;
      else  // chunk-major, slot order: a block's writes are contiguous
        partial[m.pbase + (int64_t)item.chunk * m.nq + local] = acc;
    }
  }
}

}  // namespace

int strip_resident_blocks() {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)fwd_strip_kernel,
                                                   kBlockThreads, 0) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return cus * per_cu;
}

// strip_resident_blocks() of the current device, asked once per device.
int strip_resident_blocks_cached() {
  static std::mutex mu;
  static std::map<int, int> by_device;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  std::lock_guard<std::mutex> lock(mu);
  auto it = by_device.find(dev);
  if (it != by_device.end()) return it->second;
  const int r = strip_resident_blocks();
  by_device[dev] = r;
  return r;
}

size_t strip_workspace_bytes(int64_t nq, int splits, int sort, int64_t nbe) {
  size_t bytes = 0;
  if (splits > 1) bytes += align_up((size_t)nq * (size_t)splits * sizeof(double));
  if (sort)
    bytes += align_up((size_t)nq * sizeof(int32_t)) + align_up((size_t)nq * 3 * sizeof(double)) +
             morton_workspace_bytes(nq);
  if (nbe > 0) bytes += ribbon_workspace_bytes(nq, ribbon_splits(nq, 1, nbe));
  return bytes;
}

hipError_t strip_launch(const double4* rec, int64_t nrec, const double* q, double* dst,
                        int64_t nq, int splits, int sort, float kappa, void* ws, size_t ws_bytes,
                        hipStream_t stream, const RibbonSrc* boundary) {
  if (nq == 0) return hipSuccess;
  const int64_t nbe = boundary ? boundary->n : 0;
  if (ws_bytes < strip_workspace_bytes(nq, splits, sort, nbe)) return hipErrorInvalidValue;
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
    cursor += morton_workspace_bytes(nq);
    q = qs;
  }
  if (ntiles == 0) return fill_launch(dst, nq, 0.0, stream);
  const int64_t tps = (ntiles + splits - 1) / splits;
  const int used_splits = (int)((ntiles + tps - 1) / tps);
  // whole query blocks up to one round of resident blocks before the end;
  // the last round's query blocks are split (none when there is one split)
  const int64_t nqb = (nq + kBlockThreads - 1) / kBlockThreads;
  const int64_t resident = std::max<int64_t>(strip_resident_blocks_cached(), 1);
  const int64_t whole = used_splits == 1 ? nqb : std::max<int64_t>(0, nqb - resident);
  const int64_t tail_qb = nqb - whole;
  const int64_t nblocks = whole + (int64_t)used_splits * tail_qb;
  timer_begin(stream);
  fwd_strip_kernel<<<dim3((unsigned)nblocks), kBlockThreads, 0, stream>>>(
      rec, ntiles, q, perm, dst, partial, nq, (int32_t)whole, tps, used_splits, kappa);
  hipError_t e = hipGetLastError();
  timer_end(stream);
  if (e == hipSuccess && tail_qb > 0) {
    const int64_t off = whole * kBlockThreads;
    e = reduce_splits_launch(partial, perm ? perm + off : nullptr, perm ? dst : dst + off,
                             nq - off, used_splits, stream);
  }
  if (e != hipSuccess || nbe <= 0) return e;
  // boundary correction of an open mesh, for the same query blocks
  RibbonSrc src = *boundary;
  src.pts = (const double*)rec;
  const int rs = ribbon_splits(nq, 1, nbe);
  return ribbon_launch(src, 0, 1, q, perm, dst, nq, rs, cursor, ribbon_workspace_bytes(nq, rs),
                       stream);
}

namespace {

template <typename VT>
__global__ __launch_bounds__(kBlockThreads) void strip_gather_kernel(
    const VT* __restrict__ V, int64_t nv, const int32_t* __restrict__ idx,
    const int8_t* __restrict__ flag, int64_t nrec, double4* __restrict__ rec) {
  const int64_t k = (int64_t)blockIdx.x * kBlockThreads + threadIdx.x;
  if (k >= nrec) return;
  const int64_t v = idx[k];
  if (v < 0 || v >= nv) return;  // ids come from the host stripifier: in range
  rec[k] = make_double4((double)V[3 * v], (double)V[3 * v + 1], (double)V[3 * v + 2],
                        (double)flag[k]);
}

}  // namespace

hipError_t strip_gather_launch(const void* V, int vdtype, int64_t nv, const int32_t* idx,
                               const int8_t* flag, int64_t nrec, double4* rec,
                               hipStream_t stream) {
  if (nrec == 0) return hipSuccess;
  const dim3 grid((unsigned)((nrec + kBlockThreads - 1) / kBlockThreads));
  if (vdtype == 0)
    strip_gather_kernel<double><<<grid, kBlockThreads, 0, stream>>>((const double*)V, nv, idx,
                                                                     flag, nrec, rec);
  else
    strip_gather_kernel<float><<<grid, kBlockThreads, 0, stream>>>((const float*)V, nv, idx,
                                                                    flag, nrec, rec);
  return hipGetLastError();
}

int batch_strip_resident_blocks() {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)batch_strip_kernel,
                                                   kBlockThreads, 0) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return cus * per_cu;
}

hipError_t batch_strip_enqueue(const double4* rec, const double* q, const int32_t* perm,
                               double* dst, double* partial, const BatchMesh* meshes,
                               const BatchItem* items, int64_t n_items, int32_t* counter,
                               int64_t chunk_tiles, float kappa, int grid, hipStream_t stream) {
  batch_strip_kernel<<<dim3((unsigned)grid), kBlockThreads, 0, stream>>>(
      rec, q, perm, dst, partial, meshes, items, n_items, counter, chunk_tiles, kappa);
  return hipGetLastError();
}

}  // namespace wn
