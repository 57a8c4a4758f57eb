// Far-field evaluation of one node's second-order expansion (see wn_fwn.h),
// shared by the CPU path (g++) and the GPU kernel (hipcc): the solid angle
// of the node's faces seen from q, unscaled (w = sum / (4 pi)).
#pragma once

#include <math.h>

#if defined(__HIPCC__)
#define WN_FWN_HD __host__ __device__
#else
#define WN_FWN_HD
#endif

namespace wn {

// e: 34 doubles = centre (3), radius, N (3), C (9, row-major), D (18:
// for i = 0..2 the symmetric (j, k) block 00 01 02 11 12 22).
WN_FWN_HD inline double fwn_far_omega(const double* e, double qx, double qy, double qz) {
  const double rx = e[0] - qx, ry = e[1] - qy, rz = e[2] - qz;
  const double r2 = rx * rx + ry * ry + rz * rz;
  const double inv_r = 1.0 / sqrt(r2);
  const double inv3 = inv_r / r2;
  const double inv5 = inv3 / r2;
  const double inv7 = inv5 / r2;
  const double* N = e + 4;
  const double* C = e + 7;
  const double* D = e + 16;
  // order 0: N . r / |r|^3
  const double t0 = (N[0] * rx + N[1] * ry + N[2] * rz) * inv3;
  // order 1: tr(C) / |r|^3 - 3 r^T C r / |r|^5
  const double rCr = rx * (C[0] * rx + C[1] * ry + C[2] * rz) +
                     ry * (C[3] * rx + C[4] * ry + C[5] * rz) +
                     rz * (C[6] * rx + C[7] * ry + C[8] * rz);
  const double t1 = (C[0] + C[4] + C[8]) * inv3 - 3.0 * rCr * inv5;
  // order 2: 1/2 [-3 (2 A + B) / |r|^5 + 15 S / |r|^7] with, per i and
  // M_i = D[i] (symmetric 3 x 3): A = sum_i (row i of M_i) . r,
  // B = sum_i r_i tr(M_i), S = sum_i r_i r^T M_i r
  const double r[3] = {rx, ry, rz};
  double A = 0.0, B = 0.0, S = 0.0;
  for (int i = 0; i < 3; ++i) {
    const double* m = D + 6 * i;  // m00 m01 m02 m11 m12 m22
    const double row0 = m[0] * rx + m[1] * ry + m[2] * rz;
    const double row1 = m[1] * rx + m[3] * ry + m[4] * rz;
    const double row2 = m[2] * rx + m[4] * ry + m[5] * rz;
    A += i == 0 ? row0 : (i == 1 ? row1 : row2);
    B += r[i] * (m[0] + m[3] + m[5]);
    S += r[i] * (rx * row0 + ry * row1 + rz * row2);
  }
  const double t2 = 0.5 * (-3.0 * (2.0 * A + B) * inv5 + 15.0 * S * inv7);
  return t0 + t1 + t2;
}

}  // namespace wn
