// Closest point on a triangle and point-box distance, shared by the CPU path
// (g++) and the GPU kernel (hipcc) so both round every operation alike.
//
// Follows src/capi.mojo:401-486 (closest_on_triangle) and :489-493
// (box_distance2): the Voronoi-region walk over the triangle's vertex, edge
// and face regions (Ericson, Real-Time Collision Detection, 5.1.5), with the
// reference's fallback for degenerate triangles (|va + vb + vc| <= 1e-30:
// clamp onto the longest edge, ties in the order ab, bc, ca). FMA
// contraction is off in this header on both compilers (the CPU objects are
// built with -ffp-contract=off), so CPU and GPU results are bitwise equal.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define WN_HD __host__ __device__
#else
#define WN_HD
#endif

namespace wn {

struct ClosestHit {
  double d2;       // squared distance
  double x, y, z;  // closest point
};

WN_HD inline double cl_max(double a, double b) { return a > b ? a : b; }
WN_HD inline double cl_min(double a, double b) { return a < b ? a : b; }

WN_HD inline ClosestHit closest_on_triangle(double px, double py, double pz, double ax,
                                            double ay, double az, double bx, double by,
                                            double bz, double cx, double cy, double cz) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  // edges from A and the query relative to A
  const double abx = bx - ax, aby = by - ay, abz = bz - az;
  const double acx = cx - ax, acy = cy - ay, acz = cz - az;
  const double apx = px - ax, apy = py - ay, apz = pz - az;
  const double d1 = abx * apx + aby * apy + abz * apz;
  const double d2 = acx * apx + acy * apy + acz * apz;
  double qx = ax, qy = ay, qz = az;  // vertex region A
  do {
    if (d1 <= 0.0 && d2 <= 0.0) break;
    const double bpx = px - bx, bpy = py - by, bpz = pz - bz;
    const double d3 = abx * bpx + aby * bpy + abz * bpz;
    const double d4 = acx * bpx + acy * bpy + acz * bpz;
    if (d3 >= 0.0 && d4 <= d3) {  // vertex region B
      qx = bx; qy = by; qz = bz;
      break;
    }
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {  // edge AB
      const double t = d1 / (d1 - d3);
      qx = ax + t * abx; qy = ay + t * aby; qz = az + t * abz;
      break;
    }
    const double cpx = px - cx, cpy = py - cy, cpz = pz - cz;
    const double d5 = abx * cpx + aby * cpy + abz * cpz;
    const double d6 = acx * cpx + acy * cpy + acz * cpz;
    if (d6 >= 0.0 && d5 <= d6) {  // vertex region C
      qx = cx; qy = cy; qz = cz;
      break;
    }
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {  // edge AC
      const double t = d2 / (d2 - d6);
      qx = ax + t * acx; qy = ay + t * acy; qz = az + t * acz;
      break;
    }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {  // edge BC
      const double t = (d4 - d3) / ((d4 - d3) + (d5 - d6));
      qx = bx + t * (cx - bx); qy = by + t * (cy - by); qz = bz + t * (cz - bz);
      break;
    }
    const double area = va + vb + vc;
    if ((area < 0.0 ? -area : area) <= 1.0e-30) {
      // degenerate: nearest point of the longest edge (segment clamp)
      const double bcx = cx - bx, bcy = cy - by, bcz = cz - bz;
      const double lab = abx * abx + aby * aby + abz * abz;
      const double lbc = bcx * bcx + bcy * bcy + bcz * bcz;
      const double lca = acx * acx + acy * acy + acz * acz;
      if (lab >= lbc && lab >= lca && lab > 0.0) {
        const double t = cl_max(0.0, cl_min(1.0, d1 / lab));
        qx = ax + t * abx; qy = ay + t * aby; qz = az + t * abz;
      } else if (lbc >= lca && lbc > 0.0) {
        const double t = cl_max(0.0, cl_min(1.0, (bpx * bcx + bpy * bcy + bpz * bcz) / lbc));
        qx = bx + t * bcx; qy = by + t * bcy; qz = bz + t * bcz;
      } else if (lca > 0.0) {
        const double t = cl_max(0.0, cl_min(1.0, d2 / lca));
        qx = ax + t * acx; qy = ay + t * acy; qz = az + t * acz;
      }
      break;
    }
    // face interior: barycentric (1 - v - w, v, w)
    const double inv = 1.0 / area;
    const double v = vb * inv, w = vc * inv;
    qx = ax + abx * v + acx * w;
    qy = ay + aby * v + acy * w;
    qz = az + abz * v + acz * w;
  } while (false);
  const double dx = px - qx, dy = py - qy, dz = pz - qz;
  ClosestHit h;
  h.d2 = dx * dx + dy * dy + dz * dz;
  h.x = qx; h.y = qy; h.z = qz;
  return h;
}

// Squared distance from a point to a box (lo, hi); 0 inside.
WN_HD inline double box_distance2(double px, double py, double pz, double lox, double loy,
                                  double loz, double hix, double hiy, double hiz) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  const double dx = cl_max(lox - px, cl_max(0.0, px - hix));
  const double dy = cl_max(loy - py, cl_max(0.0, py - hiy));
  const double dz = cl_max(loz - pz, cl_max(0.0, pz - hiz));
  return dx * dx + dy * dy + dz * dz;
}

}  // namespace wn
