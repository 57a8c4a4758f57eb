// CPU path of igl.winding_number(..., device="cpu"): the reference's
// mli_winding ABI (src/capi.mojo:938-942) over the reference's per-query face
// sweep (src/capi.mojo:205-294).
//
// Summation order is the reference's: kLanes x kAcc partial sums over blocks
// of kLanes*kAcc faces (src/capi.mojo:215-244), kLanes-wide blocks into the
// first partial (:245-264), the ((p0 + p1) + (p2 + p3)) lane combine and a
// halving lane reduction (:265), the scalar tail (:266-285), then / 2 pi
// (:286). kLanes = 4 is the float64 SIMD width of the published AVX2 host.
// Built with -ffp-contract=off so every operation rounds as written. Spans
// [q0, q1) are independent; the Python shim fans them out over threads.

#include <cmath>
#include <cstdint>

namespace {

constexpr int kLanes = 4;
constexpr int kAcc = 4;
constexpr double kTwoPi = 6.28318530717958647692528676655900577;

struct Query {
  double x, y, z;
};

inline double face_angle(const double* tri, int64_t nf, int64_t f, const Query& o) {
  const double ax = tri[f] - o.x, ay = tri[nf + f] - o.y, az = tri[2 * nf + f] - o.z;
  const double bx = tri[3 * nf + f] - o.x, by = tri[4 * nf + f] - o.y,
               bz = tri[5 * nf + f] - o.z;
  const double cx = tri[6 * nf + f] - o.x, cy = tri[7 * nf + f] - o.y,
               cz = tri[8 * nf + f] - o.z;
  const double la = std::sqrt(ax * ax + ay * ay + az * az);
  const double lb = std::sqrt(bx * bx + by * by + bz * bz);
  const double lc = std::sqrt(cx * cx + cy * cy + cz * cz);
  const double det =
      ax * (by * cz - bz * cy) - ay * (bx * cz - bz * cx) + az * (bx * cy - by * cx);
  double den = la * lb * lc;
  den += (ax * bx + ay * by + az * bz) * lc;
  den += (bx * cx + by * cy + bz * cz) * la;
  den += (cx * ax + cy * ay + cz * az) * lb;
  return std::atan2(det, den);
}

double winding_of(const double* tri, int64_t nf, const Query& o) {
  double part[kAcc][kLanes] = {};
  int64_t f = 0;
  for (; f + kAcc * kLanes <= nf; f += kAcc * kLanes)
    for (int a = 0; a < kAcc; ++a)
      for (int l = 0; l < kLanes; ++l) part[a][l] += face_angle(tri, nf, f + a * kLanes + l, o);
  for (; f + kLanes <= nf; f += kLanes)
    for (int l = 0; l < kLanes; ++l) part[0][l] += face_angle(tri, nf, f + l, o);
  double lane[kLanes];
  for (int l = 0; l < kLanes; ++l) lane[l] = (part[0][l] + part[1][l]) + (part[2][l] + part[3][l]);
  double total = (lane[0] + lane[2]) + (lane[1] + lane[3]);
  for (; f < nf; ++f) total += face_angle(tri, nf, f, o);
  return total / kTwoPi;
}

}  // namespace

extern "C" void mli_winding(const double* tri, const double* q, double* dst, int64_t nf,
                            int64_t nq, int64_t q0, int64_t q1) {
  (void)nq;
  for (int64_t i = q0; i < q1; ++i) dst[i] = winding_of(tri, nf, Query{q[3 * i], q[3 * i + 1], q[3 * i + 2]});
}
