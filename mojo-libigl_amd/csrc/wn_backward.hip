// Adjoint of the winding number (no counterpart in the reference, which has
// no derivative; the forward it differentiates is src/capi.mojo:218-235,286).
//
// For one pair, theta = atan2(det, den) with a = A - O, b = B - O, c = C - O:
//   d theta = (den d(det) - det d(den)) / (det^2 + den^2)
//   d det / da = b x c,  d det / db = c x a,  d det / dc = a x b
//   d den / da = a^ (lb lc + b.c) + b lc + c lb      (a^ = a / la)
//   d den / db = b^ (lc la + c.a) + c la + a lc
//   d den / dc = c^ (la lb + a.b) + a lb + b la
// and d theta / dO = -(d/da + d/db + d/dc). With w = sum theta / (2 pi) and
// upstream gradient g_q = dL/dw_q:
//   dL/dO_q = g_q / (2 pi) sum_f d theta_qf / dO
//   dL/dA_f = sum_q g_q / (2 pi) d theta_qf / da      (likewise B, C)
// Pairs with det^2 + den^2 == 0 or a zero corner distance (query on an edge
// line or at a corner, where theta is not differentiable) contribute 0.
//
// FP64 throughout; corner distances and the 1/(det^2 + den^2) weight come
// from the hardware reciprocal (square root) estimates refined by Newton
// steps. Face tiles are restaged from the 9 x nf SoA per block (no per-pair
// quantity is stored). Each lane holds 1, 2 or 4 queries, which amortise the
// per-face reduction of the face gradient over the wave. dL/dO stays in
// registers (atomics only when the face range is split); the face gradient
// is reduce-scattered over 16-lane rows (DPP), summed over the rows and
// waves of the block (LDS ds_add_f64), then added to grad_tri with one FP64
// atomic per face-corner component per block -- so grad_tri and grad_V are
// sums in an unspecified order (not bitwise reproducible), grad_q with
// splits == 1 is.
//
// bwd_edge_kernel computes dL/dO and dL/dV from the mesh's net boundary
// edges instead (the derivation is above the kernel); same tiling, queries
// per lane and reduction, six components per edge.

#include "wn_device.h"
#include "wn_internal.h"

namespace wn {

namespace {

struct PairGrad {
  double ax, ay, az, bx, by, bz, cx, cy, cz;  // d theta / d(a, b, c)
};

// 1/sqrt(x) and 1/x in FP64 from the hardware estimates plus kNewton Newton
// steps each (quadratic convergence: one step takes the estimate's ~2^-23
// relative error to ~2^-46); x > 0 finite.
constexpr int kNewton = 1;

__device__ __forceinline__ double rsqrt_nr(double x) {
  double r = __builtin_amdgcn_rsq(x);
#pragma unroll
  for (int i = 0; i < kNewton; ++i) {
    const double e = __builtin_fma(-x * r, r, 1.0);  // 1 - x r^2
    r = __builtin_fma(0.5 * r, e, r);
  }
  return r;
}

__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
#pragma unroll
  for (int i = 0; i < kNewton; ++i) {
    const double e = __builtin_fma(-x, r, 1.0);  // 1 - x r
    r = __builtin_fma(r, e, r);
  }
  return r;
}

// d theta / d(a, b, c): the formulas above, with reciprocal square roots for
// the corner distances and one reciprocal for det^2 + den^2 (Newton-refined
// to ~1e-14 relative) in place of square roots and divisions.
__device__ __forceinline__ PairGrad pair_grad_f64(double ax, double ay, double az, double bx,
                                                  double by, double bz, double cx, double cy,
                                                  double cz) {
  PairGrad g;
  const double la2 = ax * ax + ay * ay + az * az;
  const double lb2 = bx * bx + by * by + bz * bz;
  const double lc2 = cx * cx + cy * cy + cz * cz;
  const double bcx = by * cz - bz * cy, bcy = bz * cx - bx * cz, bcz = bx * cy - by * cx;
  const double cax = cy * az - cz * ay, cay = cz * ax - cx * az, caz = cx * ay - cy * ax;
  const double abx = ay * bz - az * by, aby = az * bx - ax * bz, abz = ax * by - ay * bx;
  const double det = ax * bcx + ay * bcy + az * bcz;
  const double ab = ax * bx + ay * by + az * bz;
  const double bc = bx * cx + by * cy + bz * cz;
  const double ca = cx * ax + cy * ay + cz * az;
  const double ra = rsqrt_nr(la2), rb = rsqrt_nr(lb2), rc = rsqrt_nr(lc2);
  const double la = la2 * ra, lb = lb2 * rb, lc = lc2 * rc;
  const double den = la * lb * lc + ab * lc + bc * la + ca * lb;
  const double r2 = det * det + den * den;
  // A zero corner distance makes la (or lb, lc) NaN and so r2; r2 == 0 puts
  // the query on an edge. Both contribute 0, as does an overflowed r2.
  if (!(r2 > 0.0 && r2 < __builtin_inf())) {
    g.ax = g.ay = g.az = g.bx = g.by = g.bz = g.cx = g.cy = g.cz = 0.0;
    return g;
  }
  const double inv = rcp_nr(r2);
  const double p = den * inv;   // weight of d det
  const double s = -det * inv;  // weight of d den
  const double ka = s * (lb * lc + bc) * ra;
  const double kb = s * (lc * la + ca) * rb;
  const double kc = s * (la * lb + ab) * rc;
  const double sla = s * la, slb = s * lb, slc = s * lc;
  g.ax = p * bcx + ka * ax + slc * bx + slb * cx;
  g.ay = p * bcy + ka * ay + slc * by + slb * cy;
  g.az = p * bcz + ka * az + slc * bz + slb * cz;
  g.bx = p * cax + kb * bx + sla * cx + slc * ax;
  g.by = p * cay + kb * by + sla * cy + slc * ay;
  g.bz = p * caz + kb * bz + sla * cz + slc * az;
  g.cx = p * abx + kc * cx + slb * ax + sla * bx;
  g.cy = p * aby + kc * cy + slb * ay + sla * by;
  g.cz = p * abz + kc * cz + slb * az + sla * bz;
  return g;
}

// Reduce-scatter of nine per-lane values over each 16-lane row. Four DPP
// exchanges with partners lane^15 (row_mirror), lane^7 (row_half_mirror),
// lane^2 and lane^1 (quad_perm) -- independent masks, so every result sums
// all 16 lanes once; at each exchange a lane keeps about half of its
// components (by the bit its partner differs in) and adds the partner's
// copy. Lane l of a row returns the row sum of component row_slot(l) (9 =
// none).
__device__ __forceinline__ int row_slot(int lane) {
  // Components kept per lane bit 3, 2, 1, 0: {0..4 | 5..8}, then the first
  // three | last two, the first two | the third, the first | the second.
  // Nibble l: {0, 1, 2, -, 3, 4, -, -, 5, 6, 7, -, 8, -, -, -}.
  constexpr uint64_t kSlots = 0x9998976599439210ull;
  return (int)((kSlots >> (4 * (lane & 15))) & 0xF);
}

__device__ __forceinline__ double row_reduce_scatter9(const double (&v)[9], int lane) {
  const bool b3 = lane & 8, b2 = lane & 4, b1 = lane & 2, b0 = lane & 1;
  double u[5];
#pragma unroll
  for (int p = 0; p < 4; ++p)
    u[p] = (b3 ? v[5 + p] : v[p]) + dpp_move_f64<0x140>(b3 ? v[p] : v[5 + p]);
  u[4] = v[4] + dpp_move_f64<0x140>(v[4]);
  double w[3];
#pragma unroll
  for (int p = 0; p < 2; ++p)
    w[p] = (b2 ? u[3 + p] : u[p]) + dpp_move_f64<0x141>(b2 ? u[p] : u[3 + p]);
  w[2] = u[2] + dpp_move_f64<0x141>(u[2]);
  const double x0 = (b1 ? w[2] : w[0]) + dpp_move_f64<0x4E>(b1 ? w[0] : w[2]);
  const double x1 = w[1] + dpp_move_f64<0x4E>(w[1]);
  return (b0 ? x1 : x0) + dpp_move_f64<0xB1>(b0 ? x0 : x1);
}

template <int QPL>
__global__ __launch_bounds__(kBlockThreads) void bwd_f64_kernel(
    const double* __restrict__ tri, const double* __restrict__ q,
    const double* __restrict__ gw, double* __restrict__ gq, double* __restrict__ gtri,
    int64_t nf, int64_t nq, int64_t faces_per_split, int splits) {
  __shared__ double2 sface[kFaceTile][5];
  __shared__ double sacc[kFaceTile][9];
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int slot = row_slot(lane);
  const int64_t qbase = (int64_t)blockIdx.x * (kBlockThreads * QPL);

  double ox[QPL], oy[QPL], oz[QPL], g[QPL], dx[QPL], dy[QPL], dz[QPL];
#pragma unroll
  for (int k = 0; k < QPL; ++k) {
    const int64_t qi = qbase + k * kBlockThreads + tid;
    const bool ok = qi < nq;
    ox[k] = ok ? q[3 * qi] : 0.0;
    oy[k] = ok ? q[3 * qi + 1] : 0.0;
    oz[k] = ok ? q[3 * qi + 2] : 0.0;
    g[k] = ok ? gw[qi] * kInv2Pi : 0.0;
    dx[k] = dy[k] = dz[k] = 0.0;
  }

  const int64_t f0 = (int64_t)blockIdx.y * faces_per_split;
  const int64_t f1 = f0 + faces_per_split < nf ? f0 + faces_per_split : nf;
  for (int64_t t0 = f0; t0 < f1; t0 += kFaceTile) {
    const int cnt = (int)(f1 - t0 < kFaceTile ? f1 - t0 : kFaceTile);
    __syncthreads();
    if (tid < cnt) {
      const int64_t f = t0 + tid;
      sface[tid][0] = make_double2(tri[f], tri[nf + f]);
      sface[tid][1] = make_double2(tri[2 * nf + f], tri[3 * nf + f]);
      sface[tid][2] = make_double2(tri[4 * nf + f], tri[5 * nf + f]);
      sface[tid][3] = make_double2(tri[6 * nf + f], tri[7 * nf + f]);
      sface[tid][4] = make_double2(tri[8 * nf + f], 0.0);
    }
    for (int i = tid; i < kFaceTile * 9; i += kBlockThreads) (&sacc[0][0])[i] = 0.0;
    __syncthreads();
    for (int j = 0; j < cnt; ++j) {
      const double2 p0 = sface[j][0], p1 = sface[j][1], p2 = sface[j][2], p3 = sface[j][3],
                    p4 = sface[j][4];
      double fa[9];
#pragma unroll
      for (int i = 0; i < 9; ++i) fa[i] = 0.0;
#pragma unroll
      for (int k = 0; k < QPL; ++k) {
        const PairGrad pg = pair_grad_f64(p0.x - ox[k], p0.y - oy[k], p1.x - oz[k],
                                          p1.y - ox[k], p2.x - oy[k], p2.y - oz[k],
                                          p3.x - ox[k], p3.y - oy[k], p4.x - oz[k]);
        const double gk = g[k];
        dx[k] -= gk * (pg.ax + pg.bx + pg.cx);
        dy[k] -= gk * (pg.ay + pg.by + pg.cy);
        dz[k] -= gk * (pg.az + pg.bz + pg.cz);
        fa[0] += gk * pg.ax; fa[1] += gk * pg.ay; fa[2] += gk * pg.az;
        fa[3] += gk * pg.bx; fa[4] += gk * pg.by; fa[5] += gk * pg.bz;
        fa[6] += gk * pg.cx; fa[7] += gk * pg.cy; fa[8] += gk * pg.cz;
      }
      // Nine lanes per row add their component's row sum (four rows: 4-way
      // same-address LDS atomics).
      const double rs = row_reduce_scatter9(fa, lane);
      if (slot < 9) atomicAdd(&sacc[j][slot], rs);
    }
    __syncthreads();
    if (tid < cnt) {
      const int64_t f = t0 + tid;
#pragma unroll
      for (int i = 0; i < 9; ++i) unsafeAtomicAdd(&gtri[i * nf + f], sacc[tid][i]);
    }
  }

#pragma unroll
  for (int k = 0; k < QPL; ++k) {
    const int64_t qi = qbase + k * kBlockThreads + tid;
    if (qi < nq) {
      if (splits == 1) {
        gq[3 * qi] = dx[k];
        gq[3 * qi + 1] = dy[k];
        gq[3 * qi + 2] = dz[k];
      } else {
        unsafeAtomicAdd(&gq[3 * qi], dx[k]);
        unsafeAtomicAdd(&gq[3 * qi + 1], dy[k]);
        unsafeAtomicAdd(&gq[3 * qi + 2], dz[k]);
      }
    }
  }
}

// Edge form of the same adjoint. theta_f is half the solid angle of face f,
// a function of its boundary loop only, and the variation of a loop's solid
// angle under a motion of one straight edge P1 -> P2 integrates in closed
// form: with a = P1 - O, b = P2 - O and t = la lb + a.b,
//   d theta / dP1 = -(a x b) / (2 la t),  d theta / dP2 = -(a x b) / (2 lb t),
// and d theta / dO = -(d/dP1 + d/dP2) summed over the face's three edges.
// Summed over a mesh, the terms of an edge shared by two faces with opposite
// directions cancel identically, so dL/dO and dL/dV are sums over the net
// directed boundary edges alone (weight = net count) -- zero for a closed,
// consistently oriented mesh. t is formed as |a x b|^2 / (la lb - a.b) when
// a.b < 0 (no cancellation near the edge); t == 0 (query on the edge) and
// zero distances contribute 0.

// Reduce-scatter of six per-lane values over each 16-lane row, with the
// exchanges of row_reduce_scatter9: components {0,1,2 | 3,4,5} by lane bit 3,
// then {first two | third} by bit 2, then {first | second} by bit 1 (lanes
// with bit 2 set keep their one component), then the lane^1 sum. Lane l
// returns the row sum of component edge_slot(l) (6 = none: the duplicates).
__device__ __forceinline__ int edge_slot(int lane) {
  // Nibble l: {0, -, 1, -, 2, -, -, -, 3, -, 4, -, 5, -, -, -}.
  constexpr uint64_t kSlots = 0x6665646366626160ull;
  return (int)((kSlots >> (4 * (lane & 15))) & 0xF);
}

__device__ __forceinline__ double row_reduce_scatter6(const double (&v)[6], int lane) {
  const bool b3 = lane & 8, b2 = lane & 4, b1 = lane & 2;
  double u[3];
#pragma unroll
  for (int p = 0; p < 3; ++p)
    u[p] = (b3 ? v[3 + p] : v[p]) + dpp_move_f64<0x140>(b3 ? v[p] : v[3 + p]);
  const double x0 = (b2 ? u[2] : u[0]) + dpp_move_f64<0x141>(b2 ? u[0] : u[2]);
  const double x1 = u[1] + dpp_move_f64<0x141>(u[1]);
  const double keep = b2 ? x0 : (b1 ? x1 : x0);
  const double give = b2 ? x0 : (b1 ? x0 : x1);
  const double y = keep + dpp_move_f64<0x4E>(give);
  return y + dpp_move_f64<0xB1>(y);
}

template <int QPL, typename VT>
__global__ __launch_bounds__(kBlockThreads) void bwd_edge_kernel(
    const VT* __restrict__ V, const int32_t* __restrict__ edges,
    const double* __restrict__ weight, const double* __restrict__ q,
    const double* __restrict__ gw, double* __restrict__ gq, double* __restrict__ gv, int64_t nv,
    int64_t ne, int64_t nq, int64_t edges_per_split, int splits) {
  __shared__ double2 sedge[kFaceTile][4];
  __shared__ double sacc[kFaceTile][6];
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int slot = edge_slot(lane);
  const int64_t qbase = (int64_t)blockIdx.x * (kBlockThreads * QPL);

  double ox[QPL], oy[QPL], oz[QPL], g[QPL], dx[QPL], dy[QPL], dz[QPL];
#pragma unroll
  for (int k = 0; k < QPL; ++k) {
    const int64_t qi = qbase + k * kBlockThreads + tid;
    const bool ok = qi < nq;
    ox[k] = ok ? q[3 * qi] : 0.0;
    oy[k] = ok ? q[3 * qi + 1] : 0.0;
    oz[k] = ok ? q[3 * qi + 2] : 0.0;
    g[k] = ok ? gw[qi] * kInv2Pi : 0.0;
    dx[k] = dy[k] = dz[k] = 0.0;
  }

  const int64_t e0 = (int64_t)blockIdx.y * edges_per_split;
  const int64_t e1 = e0 + edges_per_split < ne ? e0 + edges_per_split : ne;
  for (int64_t t0 = e0; t0 < e1; t0 += kFaceTile) {
    const int cnt = (int)(e1 - t0 < kFaceTile ? e1 - t0 : kFaceTile);
    __syncthreads();
    if (tid < cnt) {
      const int64_t e = t0 + tid;
      const int64_t i = edges[2 * e], j = edges[2 * e + 1];
      // an id outside V makes a zero-weight edge at the origin (contributes 0)
      const bool in = i >= 0 && i < nv && j >= 0 && j < nv;
      const int64_t ii = in ? i : 0, jj = in ? j : 0;
      const double s = in ? 1.0 : 0.0;
      sedge[tid][0] = make_double2(s * (double)V[3 * ii], s * (double)V[3 * ii + 1]);
      sedge[tid][1] = make_double2(s * (double)V[3 * ii + 2], s * (double)V[3 * jj]);
      sedge[tid][2] = make_double2(s * (double)V[3 * jj + 1], s * (double)V[3 * jj + 2]);
      sedge[tid][3] = make_double2(in ? -0.5 * weight[e] : 0.0, 0.0);
    }
    for (int i = tid; i < kFaceTile * 6; i += kBlockThreads) (&sacc[0][0])[i] = 0.0;
    __syncthreads();
    for (int j = 0; j < cnt; ++j) {
      const double2 p0 = sedge[j][0], p1 = sedge[j][1], p2 = sedge[j][2];
      const double hw = sedge[j][3].x;
      double fa[6];
#pragma unroll
      for (int i = 0; i < 6; ++i) fa[i] = 0.0;
#pragma unroll
      for (int k = 0; k < QPL; ++k) {
        const double ax = p0.x - ox[k], ay = p0.y - oy[k], az = p1.x - oz[k];
        const double bx = p1.y - ox[k], by = p2.x - oy[k], bz = p2.y - oz[k];
        const double ra = rsqrt_nr(ax * ax + ay * ay + az * az);
        const double rb = rsqrt_nr(bx * bx + by * by + bz * bz);
        const double la = (ax * ax + ay * ay + az * az) * ra;
        const double lb = (bx * bx + by * by + bz * bz) * rb;
        const double ab = ax * bx + ay * by + az * bz;
        const double xx = ay * bz - az * by, xy = az * bx - ax * bz, xz = ax * by - ay * bx;
        const double lalb = la * lb;
        const bool pos = ab >= 0.0;
        const double x2 = xx * xx + xy * xy + xz * xz;
        const double dd = pos ? lalb + ab : x2;  // t, or |a x b|^2 = t (la lb - a.b)
        const double num = pos ? 1.0 : lalb - ab;
        // a zero distance makes ra (and so dd) NaN; dd == 0 is a query on the edge
        const bool ok = dd > 0.0 && dd < __builtin_inf();
        const double kk = hw * g[k] * num * rcp_nr(ok ? dd : 1.0);
        const double k1 = ok ? kk * ra : 0.0, k2 = ok ? kk * rb : 0.0;
        fa[0] += k1 * xx; fa[1] += k1 * xy; fa[2] += k1 * xz;
        fa[3] += k2 * xx; fa[4] += k2 * xy; fa[5] += k2 * xz;
        const double k12 = k1 + k2;
        dx[k] -= k12 * xx;
        dy[k] -= k12 * xy;
        dz[k] -= k12 * xz;
      }
      const double rs = row_reduce_scatter6(fa, lane);
      if (slot < 6) atomicAdd(&sacc[j][slot], rs);
    }
    __syncthreads();
    if (tid < cnt) {
      const int64_t e = t0 + tid;
      const int64_t i = edges[2 * e], j = edges[2 * e + 1];
      if (i >= 0 && i < nv && j >= 0 && j < nv) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          unsafeAtomicAdd(&gv[3 * i + d], sacc[tid][d]);
          unsafeAtomicAdd(&gv[3 * j + d], sacc[tid][3 + d]);
        }
      }
    }
  }

#pragma unroll
  for (int k = 0; k < QPL; ++k) {
    const int64_t qi = qbase + k * kBlockThreads + tid;
    if (qi < nq) {
      if (splits == 1) {
        gq[3 * qi] = dx[k];
        gq[3 * qi + 1] = dy[k];
        gq[3 * qi + 2] = dz[k];
      } else {
        unsafeAtomicAdd(&gq[3 * qi], dx[k]);
        unsafeAtomicAdd(&gq[3 * qi + 1], dy[k]);
        unsafeAtomicAdd(&gq[3 * qi + 2], dz[k]);
      }
    }
  }
}

__global__ __launch_bounds__(kBlockThreads) void zero_kernel(double* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlockThreads + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

template <typename FT>
__global__ __launch_bounds__(kBlockThreads) void scatter_corners_kernel(
    const double* __restrict__ gtri, const FT* __restrict__ F, int64_t nf, int64_t nv,
    double* __restrict__ gv) {
  const int64_t f = (int64_t)blockIdx.x * kBlockThreads + threadIdx.x;
  if (f >= nf) return;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int64_t v = (int64_t)F[3 * f + c];
    if (v < 0 || v >= nv) continue;
#pragma unroll
    for (int d = 0; d < 3; ++d) unsafeAtomicAdd(&gv[3 * v + d], gtri[(3 * c + d) * nf + f]);
  }
}

inline dim3 grid1(int64_t n) { return dim3((unsigned)((n + kBlockThreads - 1) / kBlockThreads)); }

}  // namespace

const char* bwd_check(int64_t nf, int64_t nq, int qpl, int splits) {
  if (nf < 0 || nq < 0) return "negative size";
  if (qpl != 1 && qpl != 2 && qpl != 4) return "backward queries_per_lane must be 1, 2 or 4";
  if (splits < 1 || splits > 65535) return "splits must be in [1, 65535]";
  if (nq > (int64_t)0x7fffffff) return "more than 2^31-1 queries per call";
  return nullptr;
}

hipError_t bwd_launch(const double* tri, const double* q, const double* gw, double* gq,
                      double* gtri, int64_t nf, int64_t nq, int qpl, int splits,
                      hipStream_t stream) {
  if (nq > 0) zero_kernel<<<grid1(3 * nq), kBlockThreads, 0, stream>>>(gq, 3 * nq);
  if (nf > 0) zero_kernel<<<grid1(9 * nf), kBlockThreads, 0, stream>>>(gtri, 9 * nf);
  if (nq == 0 || nf == 0) return hipGetLastError();
  int64_t per = (nf + splits - 1) / splits;
  per = (per + kFaceTile - 1) / kFaceTile * kFaceTile;
  const int64_t qblock = (int64_t)kBlockThreads * qpl;
  const dim3 grid((unsigned)((nq + qblock - 1) / qblock), (unsigned)splits);
  if (qpl == 1)
    bwd_f64_kernel<1><<<grid, kBlockThreads, 0, stream>>>(tri, q, gw, gq, gtri, nf, nq, per, splits);
  else if (qpl == 2)
    bwd_f64_kernel<2><<<grid, kBlockThreads, 0, stream>>>(tri, q, gw, gq, gtri, nf, nq, per, splits);
  else
    bwd_f64_kernel<4><<<grid, kBlockThreads, 0, stream>>>(tri, q, gw, gq, gtri, nf, nq, per, splits);
  return hipGetLastError();
}

const char* bwd_edge_check(int64_t ne, int64_t nq, int qpl, int splits) {
  if (ne < 0 || nq < 0) return "negative size";
  if (qpl != 1 && qpl != 2 && qpl != 4) return "backward queries_per_lane must be 1, 2 or 4";
  if (splits < 1 || splits > 65535) return "splits must be in [1, 65535]";
  if (nq > (int64_t)0x7fffffff) return "more than 2^31-1 queries per call";
  return nullptr;
}

template <typename VT>
static void edge_dispatch(dim3 grid, const VT* V, int64_t nv, const int32_t* edges,
                          const double* weight, int64_t ne, const double* q, const double* gw, double* gq, double* gv,
                          int64_t nq, int qpl, int64_t per, int splits, hipStream_t stream) {
  if (qpl == 1)
    bwd_edge_kernel<1, VT><<<grid, kBlockThreads, 0, stream>>>(V, edges, weight, q, gw, gq, gv,
                                                               nv, ne, nq, per, splits);
  else if (qpl == 2)
    bwd_edge_kernel<2, VT><<<grid, kBlockThreads, 0, stream>>>(V, edges, weight, q, gw, gq, gv,
                                                               nv, ne, nq, per, splits);
  else
    bwd_edge_kernel<4, VT><<<grid, kBlockThreads, 0, stream>>>(V, edges, weight, q, gw, gq, gv,
                                                               nv, ne, nq, per, splits);
}

hipError_t bwd_edge_launch(const void* V, int vdtype, int64_t nv, const int32_t* edges,
                           const double* weight, int64_t ne, const double* q, const double* gw,
                           double* gq, double* gv, int64_t nq, int qpl, int splits,
                           hipStream_t stream) {
  if (nq > 0) zero_kernel<<<grid1(3 * nq), kBlockThreads, 0, stream>>>(gq, 3 * nq);
  if (nv > 0) zero_kernel<<<grid1(3 * nv), kBlockThreads, 0, stream>>>(gv, 3 * nv);
  // with no vertices every edge id is outside V: nothing to add
  if (nq == 0 || ne == 0 || nv == 0) return hipGetLastError();
  int64_t per = (ne + splits - 1) / splits;
  per = (per + kFaceTile - 1) / kFaceTile * kFaceTile;
  const int64_t qblock = (int64_t)kBlockThreads * qpl;
  const dim3 grid((unsigned)((nq + qblock - 1) / qblock), (unsigned)((ne + per - 1) / per));
  if (vdtype == 0)
    edge_dispatch(grid, (const double*)V, nv, edges, weight, ne, q, gw, gq, gv, nq, qpl, per,
                  (int)grid.y, stream);
  else
    edge_dispatch(grid, (const float*)V, nv, edges, weight, ne, q, gw, gq, gv, nq, qpl, per,
                  (int)grid.y, stream);
  return hipGetLastError();
}

hipError_t scatter_corners_launch(const double* gtri, const void* F, int fdtype, int64_t nf,
                                  int64_t nv, double* gv, hipStream_t stream) {
  if (nv > 0) zero_kernel<<<grid1(3 * nv), kBlockThreads, 0, stream>>>(gv, 3 * nv);
  if (nf > 0) {
    if (fdtype == 0)
      scatter_corners_kernel<int64_t><<<grid1(nf), kBlockThreads, 0, stream>>>(
          gtri, (const int64_t*)F, nf, nv, gv);
    else
      scatter_corners_kernel<int32_t><<<grid1(nf), kBlockThreads, 0, stream>>>(
          gtri, (const int32_t*)F, nf, nv, gv);
  }
  return hipGetLastError();
}

}  // namespace wn
