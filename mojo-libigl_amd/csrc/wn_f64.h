// FP64 building blocks of the far-pair fast path shared by the FP64 face
// and strip kernels: square root and quotient refined from the hardware
// estimates to FP64 rounding, the arctangent polynomial on |det / den| < 1/8,
// and the wave-uniform far test.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wn {

// sqrt(x) for x > 0 finite: one Goldschmidt step from v_rsq_f64 (relative
// error d of the estimate -> 1.5 d^2), then the root corrected by its exact
// remainder x - g^2 times the unrefined half reciprocal (-> 1.5 d^3 before
// the final rounding). x == 0 gives NaN (0 * inf), which fails the far test
// and sends the wave to pair_f64.
__device__ __forceinline__ double sqrt_gs_f64(double x) {
  const double r = __builtin_amdgcn_rsq(x);
  const double g = x * r, h = 0.5 * r;
  const double e = __builtin_fma(-g, h, 0.5);
  const double g1 = __builtin_fma(g, e, g);
  return __builtin_fma(__builtin_fma(-g1, g1, x), h, g1);
}

// n / d for d > 0 finite: v_rcp_f64, one Newton step, then the quotient
// corrected by its remainder.
__device__ __forceinline__ double div_f64(double n, double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = __builtin_fma(r, __builtin_fma(-d, r, 1.0), r);
  const double q = n * r;
  return __builtin_fma(__builtin_fma(-d, q, n), r, q);
}

// atan(x / 8) for |x| <= 1: x * P(x^2) with the n = 7 relative-error
// minimax coefficients of atan on (0, 1/8], each divided by 8^(2k+1)
// (powers of two: exact). The Horner steps are one asm block of VOP3
// v_fma_f64 with the coefficients in scalar registers: compiled from C, the
// coefficients stay in VGPRs and each step copies one into a v_fmac_f64
// destination (extra FP64 moves per pair), and one asm block per step costs
// a wait state after each.
__device__ __forceinline__ double atan_small_x8_f64(double x) {
  const double t = x * x;
  double p;
  asm("v_fma_f64 %0, %2, %1, %3\n\t"
      "v_fma_f64 %0, %0, %1, %4\n\t"
      "v_fma_f64 %0, %0, %1, %5\n\t"
      "v_fma_f64 %0, %0, %1, %6\n\t"
      "v_fma_f64 %0, %0, %1, %7\n\t"
      "v_fma_f64 %0, %0, %1, %8"
      : "=&v"(p)
      : "v"(t), "v"(0.073382651424641993073 / 549755813888.0),  // x^13: 8^13
        "s"(-0.090833776841241786793 / 8589934592.0),             // x^11: 8^11
        "s"(0.11111031276417376124 / 134217728.0),                // x^9: 8^9
        "s"(-0.14285713850448952581 / 2097152.0),                 // x^7: 8^7
        "s"(0.19999999998868866968 / 32768.0),                    // x^5: 8^5
        "s"(-0.33333333333332230442 / 512.0),                     // x^3: 8^3
        "s"(0.99999999999999999824 / 8.0));                       // x^1: 8
  return x * p;
}

// h + e for h = (u + v) / 2, e in a scalar register: one VOP3 v_fma_f64
// (compiled from C, the addend is copied to VGPRs for a v_fmac_f64).
__device__ __forceinline__ double half_sum_plus(double u, double v, double e) {
  double r;
  asm("v_fma_f64 %0, %1, 0.5, %2" : "=v"(r) : "v"(u + v), "s"(e));
  return r;
}

// |det8| < den is |det / den| < 1/8 (det8 = 8 det); dp > 0 is
// den > la lb lc.
__device__ __forceinline__ bool wave_all_far_x8_f64(double det8, double den, double dp) {
  const uint64_t m1 = __builtin_amdgcn_ballot_w64(__builtin_fabs(det8) < den);
  const uint64_t m2 = __builtin_amdgcn_ballot_w64(dp > 0.0);
  return (m1 & m2) == __builtin_amdgcn_read_exec();
}

// atan(x) for |x| <= 0.4: x * P(x^2) with the n = 11 relative-error minimax
// coefficients of atan on (0, 0.4] (max 1.6e-17, gen_atan_coeffs.py). The
// angle of the product of three far faces' (den + i det) (sum3_f64).
__device__ __forceinline__ double atan_sum_f64(double x) {
  // one asm block of VOP3 v_fma_f64 with the coefficients in scalar
  // registers, as atan_small_x8_f64 (from C they are kept in VGPRs)
  const double t = x * x;
  double p;
  asm("v_fma_f64 %0, %2, %1, %3\n\t"
      "v_fma_f64 %0, %0, %1, %4\n\t"
      "v_fma_f64 %0, %0, %1, %5\n\t"
      "v_fma_f64 %0, %0, %1, %6\n\t"
      "v_fma_f64 %0, %0, %1, %7\n\t"
      "v_fma_f64 %0, %0, %1, %8\n\t"
      "v_fma_f64 %0, %0, %1, %9\n\t"
      "v_fma_f64 %0, %0, %1, %10\n\t"
      "v_fma_f64 %0, %0, %1, %11\n\t"
      "v_fma_f64 %0, %0, %1, %12"
      : "=&v"(p)
      : "v"(t), "v"(0.02229460232451987210308),  // x^21
        "s"(-0.04439518713060166475412),          // x^19
        "s"(0.0571895757608583000099),            // x^17
        "s"(-0.06645891376788731639366),          // x^15
        "s"(0.07690583722896371814793),           // x^13
        "s"(-0.09090816394011522252152),          // x^11
        "s"(0.1111110798684710014473),            // x^9
        "s"(-0.1428571422406846551985),           // x^7
        "s"(0.1999999999937472279326),            // x^5
        "s"(-0.333333333333308557907),            // x^3
        "s"(0.999999999999999983756));            // x^1
  return x * p;
}

// The FP64 counterpart of sum3_f32 (wn_device.h): from det and den of three
// faces, re = Re(w1 w2 w3) and x = Im / Re = tan(angle_1 + angle_2 +
// angle_3) for w_k = den_k + i det_k; valid when every face has dp_k =
// den_k - la lb lc > 0, re is a positive normal number and |x| < 0.4
// (wave_all_sum3_f64).
struct Sum3F64 {
  double re, x;
};

__device__ __forceinline__ Sum3F64 sum3_f64(double e1, double d1, double e2, double d2,
                                           double e3, double d3) {
  const double re12 = __builtin_fma(-e1, e2, d1 * d2);
  const double im12 = __builtin_fma(d1, e2, e1 * d2);
  Sum3F64 s;
  s.re = __builtin_fma(-im12, e3, re12 * d3);
  // (the quotient is used only where re is a positive normal number)
  s.x = div_f64(__builtin_fma(re12, e3, im12 * d3), s.re);
  return s;
}

// The product of three faces' (den + i det) (sum3_f64 without the
// quotient), for groups of six: the product of two such triples.
struct Prod3F64 {
  double re, im;
};

__device__ __forceinline__ Prod3F64 prod3_f64(double e1, double d1, double e2, double d2,
                                              double e3, double d3) {
  const double re12 = __builtin_fma(-e1, e2, d1 * d2);
  const double im12 = __builtin_fma(d1, e2, e1 * d2);
  Prod3F64 t;
  t.re = __builtin_fma(-im12, e3, re12 * d3);
  t.im = __builtin_fma(re12, e3, im12 * d3);
  return t;
}

// Lanes where v is a positive normal FP64 number: one v_cmp_class_f64. Its
// operands here are FMA / min results, never a transcendental's (which an
// inline asm operand would have to wait one state for, as in
// lanes_positive_normal).
__device__ __forceinline__ uint64_t lanes_positive_normal_f64(double v) {
  uint64_t m;
  asm volatile("v_cmp_class_f64_e64 %0, %1, %2" : "=s"(m) : "v"(v), "s"(1 << 8));
  return m;
}

// A group of faces (three, or six as two triples) on every active lane: the
// smallest dp_k positive, re (and, for six, each triple's re: no
// precision lost to underflow before the last product) a positive normal
// number, |x| < 0.4 for x = Im / Re. The group's angle sum is then atan(x)
// (the argument reasoning of sum3_f32 in wn_device.h: each angle within
// +-pi/4, six within +-3pi/2, and an argument within +-0.39 is the sum). A
// NaN dp, which a min skips, comes with a NaN den and re, which the class
// test rejects.
__device__ __forceinline__ bool wave_all_group_f64(double dp_min, double re_min, double re,
                                                   double x) {
  const uint64_t m1 = __builtin_amdgcn_ballot_w64(dp_min > 0.0);
  const uint64_t m2 = lanes_positive_normal_f64(re);
  const uint64_t m3 = lanes_positive_normal_f64(re_min);
  const uint64_t m4 = __builtin_amdgcn_ballot_w64(__builtin_fabs(x) < 0.4);
  return (m1 & m2 & m3 & m4) == __builtin_amdgcn_read_exec();
}

__device__ __forceinline__ bool wave_all_sum3_f64(double dp1, double dp2, double dp3,
                                                  const Sum3F64& s) {
  const uint64_t m1 = __builtin_amdgcn_ballot_w64(dp1 > 0.0);
  const uint64_t m2 = __builtin_amdgcn_ballot_w64(dp2 > 0.0);
  const uint64_t m3 = __builtin_amdgcn_ballot_w64(dp3 > 0.0);
  uint64_t m4;  // re positive normal: one v_cmp_class_f64 writing the lane mask
  asm volatile("v_cmp_class_f64_e64 %0, %1, %2" : "=s"(m4) : "v"(s.re), "s"(1 << 8));
  const uint64_t m5 = __builtin_amdgcn_ballot_w64(__builtin_fabs(s.x) < 0.4);
  return (m1 & m2 & m3 & m4 & m5) == __builtin_amdgcn_read_exec();
}

}  // namespace wn
