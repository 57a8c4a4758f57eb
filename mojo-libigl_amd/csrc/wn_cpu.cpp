// CPU path of igl.winding_number(..., device="cpu"): the reference's
// mli_winding ABI (src/capi.mojo:938-942) over the reference's per-query face
// sweep (src/capi.mojo:205-294).
//
// Summation order is the reference's: kLanes x kAcc partial sums over blocks
// of kLanes*kAcc faces (src/capi.mojo:215-244), kLanes-wide blocks into the
// first partial (:245-264), the ((p0 + p1) + (p2 + p3)) lane combine and a
// halving lane reduction (:265), the scalar tail (:266-285), then / 2 pi
// (:286). kLanes = 4 is the float64 SIMD width of the published AVX2 host,
// and the four lanes are AVX2 lanes here too: the per-pair formula runs on
// four faces at once in the reference's operation order (built with
// -ffp-contract=off: every operation rounds as written), with an own
// four-lane atan2 in place of libm's (as the reference uses its own SIMD
// atan2). Spans [q0, q1) are independent; the Python shim fans them out
// over threads.

#include <immintrin.h>

#include <cmath>
#include <cstdint>

namespace {

constexpr int kLanes = 4;
constexpr int kAcc = 4;
constexpr double kTwoPi = 6.28318530717958647692528676655900577;

struct Query {
  double x, y, z;
};

// atan2 on four lanes (AVX2). |y| and |x| give r = min / max in [0, 1];
// above tan(pi/8), atan(r) = pi/4 + atan((min - max) / (min + max)), so one
// division gives t with |t| <= tan(pi/8) either way. atan(t) = t + t^3
// P(t^2), P the 12-term relative-error minimax fit of atan on
// (0, tan(pi/8)] (max 1.2e-18, gen_atan_coeffs.py, n = 12) less its unit
// leading term; then pi/2 - a where |y| > |x|, pi - a where x has its sign
// bit set (-0 included), and the sign of y, with pi/4, pi/2 and pi as
// two-double constants. Within 2 ulps of a correctly rounded atan2 on 16M
// random argument pairs over 60 decades (libm's result differs in 20 % of
// them by 1-2 ulps); signed zeros follow IEEE atan2 (atan2(+-0, -0) =
// +-pi, atan2(+-0, +0) = +-0), NaN in gives NaN out; infinite arguments are
// outside the winding number's use (its det and den are then NaN anyway).
inline __m256d atan2_pd(__m256d y, __m256d x) {
  const __m256d sign = _mm256_set1_pd(-0.0);
  const __m256d ay = _mm256_andnot_pd(sign, y), ax = _mm256_andnot_pd(sign, x);
  const __m256d swap = _mm256_cmp_pd(ay, ax, _CMP_GT_OQ);
  const __m256d mx = _mm256_max_pd(ax, ay), mn = _mm256_min_pd(ax, ay);
  const __m256d big =
      _mm256_cmp_pd(mn, _mm256_mul_pd(mx, _mm256_set1_pd(0.41421356237309503)), _CMP_GT_OQ);
  const __m256d num = _mm256_blendv_pd(mn, _mm256_sub_pd(mn, mx), big);
  __m256d den = _mm256_blendv_pd(mx, _mm256_add_pd(mn, mx), big);
  den = _mm256_blendv_pd(den, _mm256_set1_pd(1.0), _mm256_cmp_pd(mx, _mm256_setzero_pd(), _CMP_EQ_OQ));
  const __m256d t = _mm256_div_pd(num, den);
  const __m256d t2 = _mm256_mul_pd(t, t);
  __m256d p = _mm256_set1_pd(-0.01783567611769065118095);
  p = _mm256_fmadd_pd(p, t2, _mm256_set1_pd(0.03799387238825216926295));
  p = _mm256_fmadd_pd(p, t2, _mm256_set1_pd(-0.05036270381980785347696));
  p = _mm256_fmadd_pd(p, t2, _mm256_set1_pd(0.05847147724306206998007));
  p = _mm256_fmadd_pd(p, t2, _mm256_set1_pd(-0.06662990414266017001494));
  p = _mm256_fmadd_pd(p, t2, _mm256_set1_pd(0.07692048884228758962105));
  p = _mm256_fmadd_pd(p, t2, _mm256_set1_pd(-0.09090897019181828460015));
  p = _mm256_fmadd_pd(p, t2, _mm256_set1_pd(0.1111111075266798763801));
  p = _mm256_fmadd_pd(p, t2, _mm256_set1_pd(-0.1428571427941699912346));
  p = _mm256_fmadd_pd(p, t2, _mm256_set1_pd(0.1999999999994273295678));
  p = _mm256_fmadd_pd(p, t2, _mm256_set1_pd(-0.333333333333331291005));
  __m256d a = _mm256_fmadd_pd(_mm256_mul_pd(t, t2), p, t);
  const __m256d pio4_hi = _mm256_set1_pd(0.78539816339744827900);
  const __m256d pio4_lo = _mm256_set1_pd(3.0616169978683830e-17);
  a = _mm256_blendv_pd(a, _mm256_add_pd(pio4_hi, _mm256_add_pd(a, pio4_lo)), big);
  const __m256d pio2_hi = _mm256_set1_pd(1.57079632679489655800);
  const __m256d pio2_lo = _mm256_set1_pd(6.1232339957367660e-17);
  a = _mm256_blendv_pd(a, _mm256_add_pd(pio2_hi, _mm256_sub_pd(pio2_lo, a)), swap);
  const __m256d pi_hi = _mm256_set1_pd(3.14159265358979311600);
  const __m256d pi_lo = _mm256_set1_pd(1.2246467991473532e-16);
  a = _mm256_blendv_pd(a, _mm256_add_pd(pi_hi, _mm256_sub_pd(pi_lo, a)), x);  // sign bit of x
  a = _mm256_or_pd(a, _mm256_and_pd(sign, y));                                 // a >= 0 here
  const __m256d s = _mm256_add_pd(x, y);
  return _mm256_blendv_pd(a, s, _mm256_cmp_pd(s, s, _CMP_UNORD_Q));
}

// The reference's per-pair formula (src/capi.mojo:218-235) for the four
// faces f .. f+3 (or face f on every lane), in the reference's operation
// order, every operation rounded as written (no contraction).
inline __m256d face_angles(const double* tri, int64_t nf, int64_t f, const Query& o, bool one) {
  auto row = [&](int k) {
    return one ? _mm256_set1_pd(tri[k * nf + f]) : _mm256_loadu_pd(tri + k * nf + f);
  };
  const __m256d ox = _mm256_set1_pd(o.x), oy = _mm256_set1_pd(o.y), oz = _mm256_set1_pd(o.z);
  const __m256d ax = _mm256_sub_pd(row(0), ox), ay = _mm256_sub_pd(row(1), oy),
                az = _mm256_sub_pd(row(2), oz);
  const __m256d bx = _mm256_sub_pd(row(3), ox), by = _mm256_sub_pd(row(4), oy),
                bz = _mm256_sub_pd(row(5), oz);
  const __m256d cx = _mm256_sub_pd(row(6), ox), cy = _mm256_sub_pd(row(7), oy),
                cz = _mm256_sub_pd(row(8), oz);
  auto dot = [](__m256d px, __m256d py, __m256d pz, __m256d qx, __m256d qy, __m256d qz) {
    return _mm256_add_pd(_mm256_add_pd(_mm256_mul_pd(px, qx), _mm256_mul_pd(py, qy)),
                         _mm256_mul_pd(pz, qz));
  };
  const __m256d la = _mm256_sqrt_pd(dot(ax, ay, az, ax, ay, az));
  const __m256d lb = _mm256_sqrt_pd(dot(bx, by, bz, bx, by, bz));
  const __m256d lc = _mm256_sqrt_pd(dot(cx, cy, cz, cx, cy, cz));
  // det = ax (by cz - bz cy) - ay (bx cz - bz cx) + az (bx cy - by cx)
  const __m256d det = _mm256_add_pd(
      _mm256_sub_pd(_mm256_mul_pd(ax, _mm256_sub_pd(_mm256_mul_pd(by, cz), _mm256_mul_pd(bz, cy))),
                    _mm256_mul_pd(ay, _mm256_sub_pd(_mm256_mul_pd(bx, cz), _mm256_mul_pd(bz, cx)))),
      _mm256_mul_pd(az, _mm256_sub_pd(_mm256_mul_pd(bx, cy), _mm256_mul_pd(by, cx))));
  __m256d den = _mm256_mul_pd(_mm256_mul_pd(la, lb), lc);
  den = _mm256_add_pd(den, _mm256_mul_pd(dot(ax, ay, az, bx, by, bz), lc));
  den = _mm256_add_pd(den, _mm256_mul_pd(dot(bx, by, bz, cx, cy, cz), la));
  den = _mm256_add_pd(den, _mm256_mul_pd(dot(cx, cy, cz, ax, ay, az), lb));
  return atan2_pd(det, den);
}

double winding_of(const double* tri, int64_t nf, const Query& o) {
  __m256d part[kAcc];
  for (int a = 0; a < kAcc; ++a) part[a] = _mm256_setzero_pd();
  int64_t f = 0;
  for (; f + kAcc * kLanes <= nf; f += kAcc * kLanes)
    for (int a = 0; a < kAcc; ++a)
      part[a] = _mm256_add_pd(part[a], face_angles(tri, nf, f + a * kLanes, o, false));
  for (; f + kLanes <= nf; f += kLanes)
    part[0] = _mm256_add_pd(part[0], face_angles(tri, nf, f, o, false));
  double lane[kLanes];
  _mm256_storeu_pd(lane, _mm256_add_pd(_mm256_add_pd(part[0], part[1]),
                                       _mm256_add_pd(part[2], part[3])));
  double total = (lane[0] + lane[2]) + (lane[1] + lane[3]);
  for (; f < nf; ++f) total += _mm256_cvtsd_f64(face_angles(tri, nf, f, o, true));
  return total / kTwoPi;
}

}  // namespace

extern "C" void mli_winding(const double* tri, const double* q, double* dst, int64_t nf,
                            int64_t nq, int64_t q0, int64_t q1) {
  (void)nq;
  for (int64_t i = q0; i < q1; ++i) dst[i] = winding_of(tri, nf, Query{q[3 * i], q[3 * i + 1], q[3 * i + 2]});
}
