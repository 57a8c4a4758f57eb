// Device-side building blocks of the winding-number kernels (gfx950, wave64).
//
// Per query-face pair the reference evaluates (src/capi.mojo:218-235, 309-325)
//   a = A - O, b = B - O, c = C - O, la = |a|, lb = |b|, lc = |c|
//   det = a . (b x c)
//   den = la lb lc + (a.b) lc + (b.c) la + (c.a) lb
//   angle = atan2(det, den)            (half the signed solid angle)
// and w(O) = sum_f angle_f / (2 pi)  (src/capi.mojo:286).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wn {

constexpr int kWave = 64;
constexpr double kInv2Pi = 0.159154943091895335768883763372514362;

// ---------------------------------------------------------------------------
// FP32 atan2 with our own minimax polynomial.
//
// atan(r) ~= r * P(r^2) on r in [0, 1], P of degree 8 in r^2, coefficients
// from a Remez exchange minimising the RELATIVE error (max 1.52e-8, below
// half an FP32 ulp) computed at 40 digits in mpmath by gen_atan_coeffs.py in
// this directory. Octant reduction: r = min(|y|,|x|) / max(|y|,|x|), then
// pi/2 - t if |y| > |x|, pi - t if x < 0, sign of y. The division is
// v_rcp_f32 (1 ulp), the relative accuracy of the result is ~2 ulp.
// `mx` returns max(|y|, |x|, 1e-30) for the conditioning test of the caller.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float atan2_f32(float y, float x, float& mx) {
  const float ay = __builtin_fabsf(y);
  const float ax = __builtin_fabsf(x);
  mx = __builtin_fmaxf(__builtin_fmaxf(ay, ax), 1e-30f);
  const float mn = __builtin_fminf(ay, ax);
  const float r = mn * __builtin_amdgcn_rcpf(mx);
  const float t = r * r;
  float p = 0.0028498926605998343f;
  p = __builtin_fmaf(p, t, -0.016068640521845983f);
  p = __builtin_fmaf(p, t, 0.04269153699916155f);
  p = __builtin_fmaf(p, t, -0.075042959409187448f);
  p = __builtin_fmaf(p, t, 0.10640934631555586f);
  p = __builtin_fmaf(p, t, -0.14203644605898132f);
  p = __builtin_fmaf(p, t, 0.19992619406714787f);
  p = __builtin_fmaf(p, t, -0.33333073345588504f);
  p = __builtin_fmaf(p, t, 0.99999998476576588f);
  float a = r * p;
  a = (ay > ax) ? (1.57079632679489662f - a) : a;
  a = (x < 0.0f) ? (3.14159265358979324f - a) : a;
  return __builtin_copysignf(a, y);
}

// First thread of the calling wave, in a scalar register (all lanes active).
__device__ __forceinline__ int32_t wave_base() {
  return __builtin_amdgcn_readfirstlane((int32_t)threadIdx.x) & ~(kWave - 1);
}

// threadIdx.x rebuilt from the wave base and the lane's position in the
// wave, computed where it is called: an output index formed from it after a
// sweep needs no vector register kept live across the sweep. (The mbcnt pair
// sits in a volatile asm so the compiler cannot hoist it above the sweep.)
__device__ __forceinline__ int32_t thread_index(int32_t base) {
  int32_t lane;
  asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(lane));
  return base + lane;
}

// ---------------------------------------------------------------------------
// Power-of-two frames of the FP32-compensated kernels.
//
// Every FP32 quantity of a pair scales with a power of its coordinates (the
// lengths l, den ~ l^3, a product of six faces' (den + i det) ~ l^18), and
// the group products leave the FP32 range unless l stays within a few
// powers of two of 1. Each LDS tile is therefore staged in a frame scaled by
// sc = 2^(-E), where 2^E bounds both |corner - g| over the tile and
// |query - g| over the block (g the block origin, infinity norm): every
// pair's |corner - query| is then below 2 sqrt(3) ~ 3.5, and a product of
// twelve far faces' (den + i det) below 2^89. Multiplying by a
// power of two is exact and FP32 rounding commutes with it (for results in
// the normal range), so the rounded corners and query offsets are the
// unscaled ones times sc, every angle is unchanged, and a power-of-two
// rescaling of the whole input gives bit-identical results.
// ---------------------------------------------------------------------------

// floor(log2(m)) of m >= 0 from its exponent field, clamped to [-1000, 1000]
// (0 and subnormals give -1000), so that 2^(2 - e) and 2^(2 + 1000) stay
// normal FP64 numbers.
__device__ __forceinline__ int frame_exponent(double m) {
  const int e = (int)((uint64_t)__double_as_longlong(m) >> 52) - 1023;
  return e < -1000 ? -1000 : (e > 1000 ? 1000 : e);
}

// High 32 bits of |d| (sign cleared): monotonic in |d|, and its exponent
// field is d's. A max over these words is a max of |d| to ~20 mantissa bits.
__device__ __forceinline__ uint32_t abs_hi_word(double d) {
  return (uint32_t)((uint64_t)__double_as_longlong(d) >> 32) & 0x7fffffffu;
}

// frame_exponent of the double whose (sign-cleared) high word is w.
__device__ __forceinline__ int frame_exponent_hi(uint32_t w) {
  const int e = (int)(w >> 20) - 1023;
  return e < -1000 ? -1000 : (e > 1000 ? 1000 : e);
}

// 2^k as FP64, k in [-1022, 1023].
__device__ __forceinline__ double pow2_f64(int k) {
  return __longlong_as_double((long long)(k + 1023) << 52);
}

// 2^k as FP32 for k in [-126, 0]; 0 below (a query offset that small next to
// the tile's extent is below FP32 resolution of the pair).
__device__ __forceinline__ float pow2_f32(int k) {
  return k < -126 ? 0.0f : __int_as_float((k + 127) << 23);
}

// ---------------------------------------------------------------------------
// FP64 pair in the reference's exact arithmetic form (cross-product det,
// dot-product den), with OCML's correctly-rounded sqrt and atan2.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double pair_f64(double ax, double ay, double az,
                                           double bx, double by, double bz,
                                           double cx, double cy, double cz) {
  const double la = sqrt(ax * ax + ay * ay + az * az);
  const double lb = sqrt(bx * bx + by * by + bz * bz);
  const double lc = sqrt(cx * cx + cy * cy + cz * cz);
  const double det = ax * (by * cz - bz * cy) - ay * (bx * cz - bz * cx) +
                     az * (bx * cy - by * cx);
  double den = la * lb * lc;
  den += (ax * bx + ay * by + az * bz) * lc;
  den += (bx * cx + by * cy + bz * cz) * la;
  den += (cx * ax + cy * ay + cz * az) * lb;
  return atan2(det, den);
}

// ---------------------------------------------------------------------------
// FP32-compensated face record, 64 bytes, staged in LDS per block.
// Coordinates are relative to the block's local origin (its query bbox
// centre), rounded once to FP32; N = (B'-A') x (C'-A') and the halved
// squared edge lengths are computed in FP64 from those same rounded corners,
// so every quantity the pair uses describes one consistent (rounded)
// triangle. The lane's dot product with N is det, and a law-of-cosines dot
// product is one add and one fma with 0.5 (the halving is exact).
//   r0 = (A.x, A.y, A.z, B.x)   r1 = (B.y, B.z, C.x, C.y)
//   r2 = (C.z, N.x, N.y, N.z)   r3 = (-|B-A|^2/2, -|C-B|^2/2, -|A-C|^2/2, 0)
// ---------------------------------------------------------------------------
struct FaceRec32 {
  float4 r0, r1, r2, r3;
};

// Build the record from FP64 corners and the FP64 local origin (gx,gy,gz).
__device__ __forceinline__ FaceRec32 make_face_rec32(double Ax, double Ay, double Az,
                                                     double Bx, double By, double Bz,
                                                     double Cx, double Cy, double Cz,
                                                     double gx, double gy, double gz) {
  const float ax = (float)(Ax - gx), ay = (float)(Ay - gy), az = (float)(Az - gz);
  const float bx = (float)(Bx - gx), by = (float)(By - gy), bz = (float)(Bz - gz);
  const float cx = (float)(Cx - gx), cy = (float)(Cy - gy), cz = (float)(Cz - gz);
  const double e1x = (double)bx - (double)ax, e1y = (double)by - (double)ay,
               e1z = (double)bz - (double)az;
  const double e2x = (double)cx - (double)ax, e2y = (double)cy - (double)ay,
               e2z = (double)cz - (double)az;
  const double e3x = (double)cx - (double)bx, e3y = (double)cy - (double)by,
               e3z = (double)cz - (double)bz;
  const double nx = e1y * e2z - e1z * e2y;
  const double ny = e1z * e2x - e1x * e2z;
  const double nz = e1x * e2y - e1y * e2x;
  FaceRec32 r;
  r.r0 = make_float4(ax, ay, az, bx);
  r.r1 = make_float4(by, bz, cx, cy);
  r.r2 = make_float4(cz, (float)nx, (float)ny, (float)nz);
  r.r3 = make_float4((float)(-0.5 * (e1x * e1x + e1y * e1y + e1z * e1z)),
                     (float)(-0.5 * (e3x * e3x + e3y * e3y + e3z * e3z)),
                     (float)(-0.5 * (e2x * e2x + e2y * e2y + e2z * e2z)), 0.0f);
  return r;
}

// FP64 re-evaluation of one pair from the same FP32 record and FP32 query
// offset. Differences of FP32 values are exact in FP64, so this is the
// reference formula on the rounded triangle at ~1e-16 relative accuracy.
__device__ __noinline__ double pair_rec32_f64(float4 r0, float4 r1, float4 r2,
                                              float ox, float oy, float oz) {
  const double qx = ox, qy = oy, qz = oz;
  return pair_f64((double)r0.x - qx, (double)r0.y - qy, (double)r0.z - qz,
                  (double)r0.w - qx, (double)r1.x - qy, (double)r1.y - qz,
                  (double)r1.z - qx, (double)r1.w - qy, (double)r2.x - qz);
}

// atan(r) for |r| <= 1/8: r * Q(r^2), Q of degree 2 in r^2, relative-error
// minimax on [0, 1/8] (max 1.67e-8), fitted like atan2_f32's polynomial by
// gen_atan_coeffs.py.
__device__ __forceinline__ float atan_small3_f32(float r) {
  const float t = r * r;
  float p = 0.19669884933950198f;
  p = __builtin_fmaf(p, t, -0.33331408377741376f);
  p = __builtin_fmaf(p, t, 0.99999998333354555f);
  return r * p;
}

// atan(x / 8) for |x| <= 1: atan_small3_f32 with its coefficients scaled by
// powers of two so that it takes x = 8 r (the same values, bit for bit).
__device__ __forceinline__ float atan_small_x8_f32(float x) {
  const float t = x * x;
  float p = 0.19669884933950198f * (1.0f / 32768.0f);
  p = __builtin_fmaf(p, t, -0.33331408377741376f * (1.0f / 512.0f));
  p = __builtin_fmaf(p, t, 0.99999998333354555f * 0.125f);
  return x * p;
}

// atan(x) for |x| <= 0.4: x * Q(x^2), Q of degree 4 in x^2, relative-error
// minimax on (0, 0.4] (max 1.30e-8, gen_atan_coeffs.py, n = 5). The angle of
// a product of far faces' (den + i det) (sum3_f32, sum6_f32).
__device__ __forceinline__ float atan_sum_f32(float x) {
  const float t = x * x;
  float p = 0.08156783959325319f;
  p = __builtin_fmaf(p, t, -0.13901368582940957f);
  p = __builtin_fmaf(p, t, 0.19978842639359343f);
  p = __builtin_fmaf(p, t, -0.33332919866326131f);
  p = __builtin_fmaf(p, t, 0.99999998698746807f);
  return x * p;
}

// The angles of a group of faces, summed as the argument of the product of
// their w_k = den_k + i det_k. For three faces (prod3_f32):
//   w1 w2 = re12 + i im12:  re12 = d1 d2 - e1 e2,     im12 = d1 e2 + e1 d2
//   w1 w2 w3 = re + i im:   re = re12 d3 - im12 e3,   im = re12 e3 + im12 d3
// and x = im / re = tan(angle_1 + angle_2 + angle_3) (sum3_f32); for six
// faces the product of two such groups (sum6_f32). A face with
// den > la lb lc has |det| <= la lb lc < den (|a . (b x c)| <= |a||b||c|),
// so its angle lies within +-pi/4: three angles sum to within +-3pi/4, six
// to within +-3pi/2, and in both cases a product argument within +-0.39 is
// the sum itself (the other candidates lie 2pi away, beyond 5.8). The sum
// is then atan(x) when, on top of that, 1/re is a positive normal number
// and |x| < 0.4 (the polynomial's range; wave_all_group tests both). 1/re
// positive normal bounds re to [2^-128, 2^126] (and rejects re zero,
// negative, infinite or NaN): above 2^126 the reciprocal would leave the
// normal range, which the transcendental unit returns as 0, and x would
// read 0 whatever the angles. The absolute error of the argument is that of
// adding the FP32 angles: a few ulps of the sum of their magnitudes (each
// product term carries one relative rounding; re just below 2^-126, whose
// reciprocal still passes, keeps 21 of its 24 bits).
struct Prod3 {
  float re, im;
};

__device__ __forceinline__ Prod3 prod3_f32(float e1, float d1, float e2, float d2, float e3,
                                           float d3) {
  const float re12 = __builtin_fmaf(-e1, e2, d1 * d2);
  const float im12 = __builtin_fmaf(d1, e2, e1 * d2);
  Prod3 t;
  t.re = __builtin_fmaf(-im12, e3, re12 * d3);
  t.im = __builtin_fmaf(re12, e3, im12 * d3);
  return t;
}

struct AngleSum {
  float rc, x;  // 1 / re and x = im / re
};

__device__ __forceinline__ AngleSum angle_sum_of(float re, float im) {
  AngleSum s;
  s.rc = __builtin_amdgcn_rcpf(re);
  s.x = im * s.rc;
  return s;
}

__device__ __forceinline__ AngleSum sum3_f32(float e1, float d1, float e2, float d2, float e3,
                                             float d3) {
  const Prod3 t = prod3_f32(e1, d1, e2, d2, e3, d3);
  return angle_sum_of(t.re, t.im);
}

// Six faces as two groups of three: (t1.re + i t1.im)(t2.re + i t2.im).
__device__ __forceinline__ AngleSum sum6_f32(const Prod3& t1, const Prod3& t2) {
  return angle_sum_of(__builtin_fmaf(-t1.im, t2.im, t1.re * t2.re),
                      __builtin_fmaf(t1.re, t2.im, t1.im * t2.re));
}

// 1/re positive normal (see above): one v_cmp_class_f32 writing the lane
// mask (the class builtin goes through a VGPR select and a second compare
// before its ballot). Its operand is often a v_rcp_f32 result: a VALU
// instruction reading a transcendental's result needs one wait state after
// it, which the compiler inserts for the instructions it emits but not for
// inline asm operands, so the asm starts with its own s_nop (without it, a
// schedule that put the compare right after the reciprocal read a stale
// register, and the group test, and so the rounding path, varied from run
// to run).
__device__ __forceinline__ uint64_t lanes_positive_normal(float v) {
  uint64_t m;
  asm volatile("s_nop 0\n\tv_cmp_class_f32_e64 %0, %1, %2" : "=s"(m) : "v"(v), "s"(1 << 8));
  return m;
}

// Every active lane meets the group's conditions: dp_min, the smallest
// dp_k = den_k - la lb lc of the group, positive; 1/re positive normal;
// |x| < 0.4. A NaN dp_k, which a min skips, comes with a NaN den_k and so a
// NaN re and reciprocal, which the class test rejects.
__device__ __forceinline__ uint64_t group_lanes(float dp_min, const AngleSum& s) {
  const uint64_t m1 = __builtin_amdgcn_ballot_w64(dp_min > 0.0f);
  const uint64_t m2 = lanes_positive_normal(s.rc);
  const uint64_t m3 = __builtin_amdgcn_ballot_w64(__builtin_fabsf(s.x) < 0.4f);
  return m1 & m2 & m3;
}

__device__ __forceinline__ bool wave_all_group(float dp_min, const AngleSum& s) {
  return group_lanes(dp_min, s) == __builtin_amdgcn_read_exec();
}

// Three faces: the three dp tests are one min3 and the group's compare.
__device__ __forceinline__ bool wave_all_sum3(float dp1, float dp2, float dp3,
                                              const AngleSum& s) {
  return wave_all_group(__builtin_fminf(__builtin_fminf(dp1, dp2), dp3), s);
}

// Six faces (sum6_f32 of t1, t2; m1, m2 the groups' smallest dp): also
// t1.re and t2.re positive normal, so that neither group's product lost
// precision to underflow before the last multiplication.
__device__ __forceinline__ bool wave_all_sum6(float m1, float m2, const Prod3& t1,
                                              const Prod3& t2, const AngleSum& s) {
  const uint64_t m = lanes_positive_normal(__builtin_fminf(t1.re, t2.re));
  return (group_lanes(__builtin_fminf(m1, m2), s) & m) == __builtin_amdgcn_read_exec();
}

// Six faces as two triples t1, t2 (m1, m2 their smallest dp): their
// product (re, im) and the lanes meeting the six-group conditions other than
// the quotient's (every dp > 0, both triples' re and the product's re positive
// normal: no precision lost to underflow), as a lane mask in scalar
// registers. Twelve faces are the product of two sixes: each six's
// angle sum then lies within +-pi/2 (its re > 0 and, with every angle within
// +-pi/4, the sum within +-3pi/2), so the twelve sum within +-pi and a
// twelve-product argument within +-0.39 is that sum (wave_all_dodecad).
struct Hex32 {
  float re, im;
  uint64_t ok;
};

__device__ __forceinline__ Hex32 hex_of(const Prod3& t1, float m1, const Prod3& t2, float m2) {
  Hex32 h;
  h.re = __builtin_fmaf(-t1.im, t2.im, t1.re * t2.re);
  h.im = __builtin_fmaf(t1.re, t2.im, t1.im * t2.re);
  h.ok = __builtin_amdgcn_ballot_w64(__builtin_fminf(m1, m2) > 0.0f) &
         lanes_positive_normal(__builtin_fminf(t1.re, t2.re)) & lanes_positive_normal(h.re);
  return h;
}

// The angle sum of six faces whose Hex32 is h when every lane meets the
// six-group test (h.ok, 1/re positive normal, |x| < 0.4): true and *x set.
__device__ __forceinline__ bool wave_all_hex(const Hex32& h, float& x) {
  const AngleSum s = angle_sum_of(h.re, h.im);
  x = s.x;
  const uint64_t m = h.ok & lanes_positive_normal(s.rc) &
                     __builtin_amdgcn_ballot_w64(__builtin_fabsf(s.x) < 0.4f);
  return m == __builtin_amdgcn_read_exec();
}

// Twelve faces as two sixes h1, h2: true and *x set when every lane meets
// both sixes' conditions, 1/re of the product positive normal and |x| < 0.4.
__device__ __forceinline__ bool wave_all_dodecad(const Hex32& h1, const Hex32& h2, float& x) {
  const AngleSum s = angle_sum_of(__builtin_fmaf(-h1.im, h2.im, h1.re * h2.re),
                                  __builtin_fmaf(h1.re, h2.im, h1.im * h2.re));
  x = s.x;
  const uint64_t m = h1.ok & h2.ok & lanes_positive_normal(s.rc) &
                     __builtin_amdgcn_ballot_w64(__builtin_fabsf(s.x) < 0.4f);
  return m == __builtin_amdgcn_read_exec();
}

// Geometry of one pair in FP32 from a record. Algebra (exact identities,
// cheaper and better conditioned than the direct form):
//   det = N . a              with N = (B-A) x (C-A)  (det(a,b,c) = det(a,b-a,c-a))
//   a.b = (la^2 + lb^2)/2 - |B-A|^2/2   (law of cosines), likewise b.c, c.a
//   den = la lb lc + dp,  dp = lc a.b + la b.c + lb c.a
// pql = la lb lc, the scale of the conditioning test. A distant face has
// den -> 4 la lb lc.
struct PairGeom32 {
  float det, den, dp, pql;
};

__device__ __forceinline__ PairGeom32 pair_geom32(const float4 r0, const float4 r1,
                                                  const float4 r2, const float4 r3,
                                                  float ox, float oy, float oz) {
  const float ax = r0.x - ox, ay = r0.y - oy, az = r0.z - oz;
  const float bx = r0.w - ox, by = r1.x - oy, bz = r1.y - oz;
  const float cx = r1.z - ox, cy = r1.w - oy, cz = r2.x - oz;
  const float la2 = __builtin_fmaf(az, az, __builtin_fmaf(ay, ay, ax * ax));
  const float lb2 = __builtin_fmaf(bz, bz, __builtin_fmaf(by, by, bx * bx));
  const float lc2 = __builtin_fmaf(cz, cz, __builtin_fmaf(cy, cy, cx * cx));
  const float la = __builtin_amdgcn_sqrtf(la2);
  const float lb = __builtin_amdgcn_sqrtf(lb2);
  const float lc = __builtin_amdgcn_sqrtf(lc2);
  PairGeom32 g;
  g.det = __builtin_fmaf(r2.w, az, __builtin_fmaf(r2.z, ay, r2.y * ax));
  const float ab = __builtin_fmaf(la2 + lb2, 0.5f, r3.x);
  const float bc = __builtin_fmaf(lb2 + lc2, 0.5f, r3.y);
  const float ca = __builtin_fmaf(lc2 + la2, 0.5f, r3.z);
  const float pq = la * lb;
  g.dp = __builtin_fmaf(lc, ab, __builtin_fmaf(la, bc, lb * ca));
  g.den = __builtin_fmaf(pq, lc, g.dp);
  g.pql = pq * lc;
  return g;
}

// True on every lane when `pred` holds on every active lane of the wave
// (the ballot mask compared with exec: scalar work only).
__device__ __forceinline__ bool wave_all(bool pred) {
  return __builtin_amdgcn_ballot_w64(pred) == __builtin_amdgcn_read_exec();
}

// Far test of the strip and face-record kernels, for det8 = 8 det, den and
// dp = den - la lb lc: every active lane has |det8| < den (|det / den| < 1/8, the range of
// atan_small_x8_f32) and dp > 0 (den > la lb lc, away from every edge line).
// Each compare is balloted on its own so both masks stay in SGPRs.
__device__ __forceinline__ bool wave_all_far_x8(float det8, float den, float dp) {
  const uint64_t m1 = __builtin_amdgcn_ballot_w64(__builtin_fabsf(det8) < den);
  const uint64_t m2 = __builtin_amdgcn_ballot_w64(dp > 0.0f);
  return (m1 & m2) == __builtin_amdgcn_read_exec();
}

// ---------------------------------------------------------------------------
// Wavefront reductions.
// ---------------------------------------------------------------------------
// Sum over the 64 lanes of a wave; the total is valid in lane 0 (and, with a
// lane-dependent rounding order, in the others). Steps: DPP quad_perm xor 1,
// xor 2, row_ror 4, row_ror 8 inside each 16-lane row, then ds_swizzle xor 16
// inside each half, then a cross-half permute. Fixed order => deterministic.
template <int CTRL>
__device__ __forceinline__ double dpp_move_f64(double v) {
  const uint64_t bits = __double_as_longlong(v);
  const int lo = (int)(uint32_t)bits;
  const int hi = (int)(uint32_t)(bits >> 32);
  const int lo2 = __builtin_amdgcn_update_dpp(0, lo, CTRL, 0xF, 0xF, true);
  const int hi2 = __builtin_amdgcn_update_dpp(0, hi, CTRL, 0xF, 0xF, true);
  return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi2 << 32) | (uint32_t)lo2));
}

__device__ __forceinline__ double swizzle_xor16_f64(double v) {
  const uint64_t bits = __double_as_longlong(v);
  // ds_swizzle bit mode: and_mask = 0x1F, or_mask = 0, xor_mask = 0x10.
  const int lo = __builtin_amdgcn_ds_swizzle((int)(uint32_t)bits, 0x401F);
  const int hi = __builtin_amdgcn_ds_swizzle((int)(uint32_t)(bits >> 32), 0x401F);
  return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

// A value every lane holds identically, moved to scalar registers (lane 0's
// copy of each 32-bit half).
__device__ __forceinline__ double uniform_f64(double v) {
  const uint64_t bits = __double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)bits);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(bits >> 32));
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

__device__ __forceinline__ double wave_sum_f64(double v) {
  v += dpp_move_f64<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_move_f64<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_move_f64<0x124>(v);  // row_ror:4
  v += dpp_move_f64<0x128>(v);  // row_ror:8
  v += swizzle_xor16_f64(v);
  v += __shfl_xor(v, 32, kWave);
  return v;
}

// Lane `lane`'s copy of v, in scalar registers.
__device__ __forceinline__ double lane_f64(double v, int lane) {
  const uint64_t bits = __double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)bits, lane);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(bits >> 32), lane);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Wave minimum / maximum with every lane active, returned in scalar
// registers: DPP within each 16-lane row (quad_perm, row_ror), ds_swizzle
// across the two rows of each half-wave, then lanes 0 and 32 combined. Needs
// no per-lane address registers (unlike a shuffle-based reduction).
__device__ __forceinline__ double wave_min_uniform_f64(double v) {
  v = fmin(v, dpp_move_f64<0xB1>(v));   // quad_perm [1,0,3,2]
  v = fmin(v, dpp_move_f64<0x4E>(v));   // quad_perm [2,3,0,1]
  v = fmin(v, dpp_move_f64<0x124>(v));  // row_ror:4
  v = fmin(v, dpp_move_f64<0x128>(v));  // row_ror:8
  v = fmin(v, swizzle_xor16_f64(v));
  return fmin(lane_f64(v, 0), lane_f64(v, 32));
}

__device__ __forceinline__ double wave_max_uniform_f64(double v) {
  v = fmax(v, dpp_move_f64<0xB1>(v));
  v = fmax(v, dpp_move_f64<0x4E>(v));
  v = fmax(v, dpp_move_f64<0x124>(v));
  v = fmax(v, dpp_move_f64<0x128>(v));
  v = fmax(v, swizzle_xor16_f64(v));
  return fmax(lane_f64(v, 0), lane_f64(v, 32));
}

// Wave maximum of an unsigned 32-bit value, every lane active, in a scalar
// register (the steps of wave_max_uniform_f64 on one 32-bit word).
__device__ __forceinline__ uint32_t wave_max_uniform_u32(uint32_t v) {
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, true));
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, true));
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x124, 0xF, 0xF, true));
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xF, 0xF, true));
  v = max(v, (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x401F));
  const uint32_t a = (uint32_t)__builtin_amdgcn_readlane((int)v, 0);
  const uint32_t b = (uint32_t)__builtin_amdgcn_readlane((int)v, 32);
  return a > b ? a : b;
}

__device__ __forceinline__ double wave_min_f64(double v) {
  for (int m = 1; m < kWave; m <<= 1) v = fmin(v, __shfl_xor(v, m, kWave));
  return v;
}

__device__ __forceinline__ double wave_max_f64(double v) {
  for (int m = 1; m < kWave; m <<= 1) v = fmax(v, __shfl_xor(v, m, kWave));
  return v;
}

}  // namespace wn
