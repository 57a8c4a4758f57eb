/*
 * aabb_oracle.c -- TEST INFRASTRUCTURE ONLY. Correctness oracle for the
 * point-mesh squared distance; never linked into, loaded by or called from the
 * product library. Only tests/ and the cpu_baseline leg of bench.py use it.
 *
 * Plain C, FP64, source-order arithmetic (-ffp-contract=off), restating:
 *
 *   closest_on_triangle   src/capi.mojo:401-486  Voronoi-region closest point
 *                         (Ericson, Real-Time Collision Detection, 5.1.5) with
 *                         the reference's fallback for degenerate triangles
 *                         (|va + vb + vc| <= 1e-30: clamp onto the longest
 *                         edge, ties in the order ab, bc, ca)
 *   box_distance2         src/capi.mojo:489-493
 *   aabb_query            src/capi.mojo:523-587  depth-first descent, nearer
 *                         child popped first, strict '<' on the best distance
 *   aabb_squared_distance src/capi.mojo:508-520  query span [q0, q1)
 *
 * and, independent of any tree, wn_oracle_brute_distance: the minimum of
 * closest_on_triangle over all faces in index order (strict '<', so the first
 * face reaching the minimum wins). The tree descent must reach the same
 * minimum distance; the face index can differ only between faces at exactly
 * that distance.
 *
 * Tree arrays are the reference's: boxes (nnodes x 6: lo xyz, hi xyz),
 * nodes (nnodes x 4: left, right, start, count; count > 0 marks a leaf),
 * order (leaf face ids), built by python/igl/__init__.py:286-319 (restated in
 * oracle/__init__.py).
 */
#include <math.h>
#include <stdint.h>

typedef struct {
  double d2, x, y, z;
} Hit;

static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }

static Hit closest_on_triangle(double px, double py, double pz, double ax, double ay, double az,
                               double bx, double by, double bz, double cx, double cy,
                               double cz) {
  const double e1x = bx - ax, e1y = by - ay, e1z = bz - az; /* ab */
  const double e2x = cx - ax, e2y = cy - ay, e2z = cz - az; /* ac */
  const double wax = px - ax, way = py - ay, waz = pz - az; /* ap */
  const double d1 = e1x * wax + e1y * way + e1z * waz;
  const double d2 = e2x * wax + e2y * way + e2z * waz;
  double qx = ax, qy = ay, qz = az;
  if (!(d1 <= 0.0 && d2 <= 0.0)) {
    const double wbx = px - bx, wby = py - by, wbz = pz - bz; /* bp */
    const double d3 = e1x * wbx + e1y * wby + e1z * wbz;
    const double d4 = e2x * wbx + e2y * wby + e2z * wbz;
    if (d3 >= 0.0 && d4 <= d3) {
      qx = bx; qy = by; qz = bz;
    } else {
      const double vc = d1 * d4 - d3 * d2;
      if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
        const double t = d1 / (d1 - d3);
        qx = ax + t * e1x; qy = ay + t * e1y; qz = az + t * e1z;
      } else {
        const double wcx = px - cx, wcy = py - cy, wcz = pz - cz; /* cp */
        const double d5 = e1x * wcx + e1y * wcy + e1z * wcz;
        const double d6 = e2x * wcx + e2y * wcy + e2z * wcz;
        if (d6 >= 0.0 && d5 <= d6) {
          qx = cx; qy = cy; qz = cz;
        } else {
          const double vb = d5 * d2 - d1 * d6;
          if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
            const double t = d2 / (d2 - d6);
            qx = ax + t * e2x; qy = ay + t * e2y; qz = az + t * e2z;
          } else {
            const double va = d3 * d6 - d5 * d4;
            if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
              const double t = (d4 - d3) / ((d4 - d3) + (d5 - d6));
              qx = bx + t * (cx - bx); qy = by + t * (cy - by); qz = bz + t * (cz - bz);
            } else {
              const double sum = va + vb + vc;
              if (fabs(sum) <= 1.0e-30) {
                const double fx = cx - bx, fy = cy - by, fz = cz - bz; /* bc */
                const double lab = e1x * e1x + e1y * e1y + e1z * e1z;
                const double lbc = fx * fx + fy * fy + fz * fz;
                const double lca = e2x * e2x + e2y * e2y + e2z * e2z;
                if (lab >= lbc && lab >= lca && lab > 0.0) {
                  const double t = dmax(0.0, dmin(1.0, d1 / lab));
                  qx = ax + t * e1x; qy = ay + t * e1y; qz = az + t * e1z;
                } else if (lbc >= lca && lbc > 0.0) {
                  const double t = dmax(0.0, dmin(1.0, (wbx * fx + wby * fy + wbz * fz) / lbc));
                  qx = bx + t * fx; qy = by + t * fy; qz = bz + t * fz;
                } else if (lca > 0.0) {
                  const double t = dmax(0.0, dmin(1.0, d2 / lca));
                  qx = ax + t * e2x; qy = ay + t * e2y; qz = az + t * e2z;
                }
              } else {
                const double inv = 1.0 / sum;
                const double v = vb * inv, w = vc * inv;
                qx = ax + e1x * v + e2x * w;
                qy = ay + e1y * v + e2y * w;
                qz = az + e1z * v + e2z * w;
              }
            }
          }
        }
      }
    }
  }
  const double dx = px - qx, dy = py - qy, dz = pz - qz;
  Hit h = {dx * dx + dy * dy + dz * dz, qx, qy, qz};
  return h;
}

static double box_distance2(double px, double py, double pz, const double *b) {
  const double dx = dmax(b[0] - px, dmax(0.0, px - b[3]));
  const double dy = dmax(b[1] - py, dmax(0.0, py - b[4]));
  const double dz = dmax(b[2] - pz, dmax(0.0, pz - b[5]));
  return dx * dx + dy * dy + dz * dz;
}

static void corner(const double *V, const int64_t *F, int64_t f, int k, double *x, double *y,
                   double *z) {
  const int64_t v = F[3 * f + k];
  *x = V[3 * v]; *y = V[3 * v + 1]; *z = V[3 * v + 2];
}

static Hit face_hit(const double *V, const int64_t *F, int64_t f, double px, double py,
                    double pz) {
  double ax, ay, az, bx, by, bz, cx, cy, cz;
  corner(V, F, f, 0, &ax, &ay, &az);
  corner(V, F, f, 1, &bx, &by, &bz);
  corner(V, F, f, 2, &cx, &cy, &cz);
  return closest_on_triangle(px, py, pz, ax, ay, az, bx, by, bz, cx, cy, cz);
}

/* One query of the reference descent; stack holds >= depth + 1 entries. */
static void query_one(const double *P, const double *V, const int64_t *F, const double *boxes,
                      const int64_t *nodes, const int64_t *order, double *sqrd,
                      int64_t *indices, double *closest, int64_t *stack, int64_t nnodes,
                      int64_t qi) {
  const double px = P[3 * qi], py = P[3 * qi + 1], pz = P[3 * qi + 2];
  double best = 1.0e300, bx = 0.0, by = 0.0, bz = 0.0;
  int64_t best_i = -1, top = 0;
  stack[top++] = 0;
  while (top > 0) {
    const int64_t node = stack[--top];
    if (node < 0 || node >= nnodes || box_distance2(px, py, pz, boxes + 6 * node) > best)
      continue;
    const int64_t *n = nodes + 4 * node;
    if (n[3] > 0) {
      for (int64_t k = 0; k < n[3]; ++k) {
        const int64_t f = order[n[2] + k];
        const Hit h = face_hit(V, F, f, px, py, pz);
        if (h.d2 < best) {
          best = h.d2; best_i = f; bx = h.x; by = h.y; bz = h.z;
        }
      }
    } else {
      const double dl = box_distance2(px, py, pz, boxes + 6 * n[0]);
      const double dr = box_distance2(px, py, pz, boxes + 6 * n[1]);
      /* push the farther child first so the nearer one is popped next */
      const int64_t far_c = dl < dr ? n[1] : n[0], near_c = dl < dr ? n[0] : n[1];
      const double far_d = dl < dr ? dr : dl, near_d = dl < dr ? dl : dr;
      if (far_d <= best) stack[top++] = far_c;
      if (near_d <= best) stack[top++] = near_c;
    }
  }
  sqrd[qi] = best;
  indices[qi] = best_i;
  closest[3 * qi] = bx; closest[3 * qi + 1] = by; closest[3 * qi + 2] = bz;
}

void wn_oracle_aabb_distance(const double *P, const double *V, const int64_t *F,
                             const double *boxes, const int64_t *nodes, const int64_t *order,
                             double *sqrd, int64_t *indices, double *closest, int64_t *stack,
                             int64_t nnodes, int64_t q0, int64_t q1) {
  for (int64_t qi = q0; qi < q1; ++qi)
    query_one(P, V, F, boxes, nodes, order, sqrd, indices, closest, stack, nnodes, qi);
}

void wn_oracle_brute_distance(const double *P, const double *V, const int64_t *F, int64_t nf,
                              double *sqrd, int64_t *indices, double *closest, int64_t np) {
  for (int64_t qi = 0; qi < np; ++qi) {
    const double px = P[3 * qi], py = P[3 * qi + 1], pz = P[3 * qi + 2];
    double best = 1.0e300, bx = 0.0, by = 0.0, bz = 0.0;
    int64_t best_i = -1;
    for (int64_t f = 0; f < nf; ++f) {
      const Hit h = face_hit(V, F, f, px, py, pz);
      if (h.d2 < best) {
        best = h.d2; best_i = f; bx = h.x; by = h.y; bz = h.z;
      }
    }
    sqrd[qi] = best;
    indices[qi] = best_i;
    closest[3 * qi] = bx; closest[3 * qi + 1] = by; closest[3 * qi + 2] = bz;
  }
}

/* closest_on_triangle for one point and one explicit triangle (9 doubles):
 * out4 = {d2, x, y, z}. */
void wn_oracle_closest_on_triangle(const double *p, const double *t9, double *out4) {
  const Hit h = closest_on_triangle(p[0], p[1], p[2], t9[0], t9[1], t9[2], t9[3], t9[4], t9[5],
                                    t9[6], t9[7], t9[8]);
  out4[0] = h.d2; out4[1] = h.x; out4[2] = h.y; out4[3] = h.z;
}
