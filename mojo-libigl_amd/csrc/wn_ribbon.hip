// Boundary correction of the FP32-compensated mode ("ribbon" pass).
//
// The FP32-compensated kernels evaluate, per block of queries, the mesh with
// its corners rounded to FP32 relative to the block's origin g (r = fl32(P -
// g)) seen from the rounded query offset o = fl32(O - g). Shared corners round
// identically, so the rounded mesh S' keeps the topology of S, and for a
// closed mesh w(S') = w(S) off the surface. An open mesh differs: S - S'
// leaves, for every boundary edge P1 -> P2 (net directed use count n), a
// thin quad (P1, P2, P2', P1') between the true and the rounded boundary,
// where P' = P + D, D = e_P - e_O is the corner's rounding residual relative
// to the query's (e_P = r - (P - g), e_O = o - (O - g)). The chains
// S - S' - sum_e n_e quad_e are closed, so off the surface
//     theta(S) = theta(S') + sum_e n_e theta(quad_e),
// theta = half the solid angle = the angle the kernels sum. This pass adds the
// second term, for the same blocks (same queries, same origin g, same
// rounding), to the kernels' result. A closed, consistently oriented mesh has
// no boundary edges and no pass.
//
// Per (query, edge) with u_i = r_i - o (the rounded offsets) and l_i = |u_i|:
//   fast path (first order in D, FP32):
//     theta_e = (u1 x u2) . (D1 / l1 + D2 / l2) / (2 t),  t = l1 l2 + u1 . u2,
//   valid when 32 t >= l1 l2 (away from the edge's line) and
//   2^21 |D| < min(l1, l2) (far from both corners in units of the rounding):
//   the neglected terms are below ~1e-10 rad there and the FP32 arithmetic
//   is relative to theta itself (|theta_e| < 5e-6 on this path).
//   slow path (any lane of the wave outside that range, FP64): the quad as
//   the two thin triangles (P1, P2, P2') and (P1, P2', P1') in closed forms
//   without cancellation: for a triangle (x, y, y + D),
//     det = D . (x x y),  den = t (ly + l(y+D)) + lx (y . D) + ly (x . D),
//   t = lx ly + x . y taken as |x x y|^2 / (lx ly - x . y) when x . y < 0,
//   and the angle atan2(det, den).
// Boundary edges come from the mesh topology (strip meshes and strip batch
// plans: record indices of both corners) or, for meshes given as a face SoA,
// from a device pass over the faces' corner coordinates: corners that are
// equal round equally, so coordinate equality is the topology that matters.

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "wn_device.h"
#include "wn_internal.h"

namespace wn {

namespace {

constexpr int kRibbonTile = 128;  // boundary edges staged in LDS per tile

// LDS record of one boundary edge for one block origin.
struct RibbonEdge {
  float4 p;  // r1.xyz, hw (weight / 2)
  float4 q;  // r2.xyz, eps (bound of |e1|, |e2| times 1.75 x 2^21)
  float4 e;  // e1.xyz, e2.x
  float2 f;  // e2.y, e2.z
  float w;   // weight
  float pad;
};

struct RibbonSmem {
  RibbonEdge edge[kRibbonTile];
  double red[6][kBlockThreads / kWave];
};

constexpr float kRibbonFar = 2097152.0f * 1.75f;  // 2^21 x (sqrt 3 bound of a max-norm)

__device__ __forceinline__ void ribbon_stage(double p1x, double p1y, double p1z, double p2x,
                                             double p2y, double p2z, double w, double gx,
                                             double gy, double gz, RibbonEdge& E) {
  // the kernels' own rounding: fl32(fl64(P - g))
  const double d1x = p1x - gx, d1y = p1y - gy, d1z = p1z - gz;
  const double d2x = p2x - gx, d2y = p2y - gy, d2z = p2z - gz;
  const float r1x = (float)d1x, r1y = (float)d1y, r1z = (float)d1z;
  const float r2x = (float)d2x, r2y = (float)d2y, r2z = (float)d2z;
  const float e1x = (float)((double)r1x - d1x), e1y = (float)((double)r1y - d1y),
              e1z = (float)((double)r1z - d1z);
  const float e2x = (float)((double)r2x - d2x), e2y = (float)((double)r2y - d2y),
              e2z = (float)((double)r2z - d2z);
  const float m = fmaxf(fmaxf(fmaxf(fabsf(e1x), fabsf(e1y)), fmaxf(fabsf(e1z), fabsf(e2x))),
                        fmaxf(fabsf(e2y), fabsf(e2z)));
  E.p = make_float4(r1x, r1y, r1z, (float)(0.5 * w));
  E.q = make_float4(r2x, r2y, r2z, m * kRibbonFar);
  E.e = make_float4(e1x, e1y, e1z, e2x);
  E.f = make_float2(e2y, e2z);
  E.w = (float)w;
  E.pad = 0.0f;
}

// t = lx ly + x . y without cancellation.
__device__ __forceinline__ double stable_t(double lx, double ly, double xy, double cx, double cy,
                                           double cz) {
  if (xy >= 0.0) return lx * ly + xy;
  return (cx * cx + cy * cy + cz * cz) / (lx * ly - xy);
}

// Angle of the thin triangle (x, y, y + D) (FP64 closed forms).
__device__ __forceinline__ double thin_angle(double xx, double xy_, double xz, double yx,
                                             double yy, double yz, double Dx, double Dy,
                                             double Dz) {
  const double lx = sqrt(xx * xx + xy_ * xy_ + xz * xz);
  const double ly = sqrt(yx * yx + yy * yy + yz * yz);
  const double zx = yx + Dx, zy = yy + Dy, zz = yz + Dz;
  const double lz = sqrt(zx * zx + zy * zy + zz * zz);
  const double cx = xy_ * yz - xz * yy, cy = xz * yx - xx * yz, cz = xx * yy - xy_ * yx;
  const double dot = xx * yx + xy_ * yy + xz * yz;
  const double t = stable_t(lx, ly, dot, cx, cy, cz);
  const double det = Dx * cx + Dy * cy + Dz * cz;
  const double den = t * (ly + lz) + lx * (yx * Dx + yy * Dy + yz * Dz) +
                     ly * (xx * Dx + xy_ * Dy + xz * Dz);
  return atan2(det, den);
}

// FP64 angle of edge E's quad for one lane (rounded offsets u = r - o,
// residuals D = e - eO), times the edge weight.
__device__ __noinline__ double ribbon_slow(const RibbonEdge& E, float ox, float oy, float oz,
                                           float eox, float eoy, float eoz) {
  const double u1x = (double)E.p.x - ox, u1y = (double)E.p.y - oy, u1z = (double)E.p.z - oz;
  const double u2x = (double)E.q.x - ox, u2y = (double)E.q.y - oy, u2z = (double)E.q.z - oz;
  const double D1x = (double)E.e.x - eox, D1y = (double)E.e.y - eoy, D1z = (double)E.e.z - eoz;
  const double D2x = (double)E.e.w - eox, D2y = (double)E.f.x - eoy, D2z = (double)E.f.y - eoz;
  // exact offsets of the true corners: a = P1 - O, b = P2 - O
  const double ax = u1x - D1x, ay = u1y - D1y, az = u1z - D1z;
  const double bx = u2x - D2x, by = u2y - D2y, bz = u2z - D2z;
  // (P1, P2, P2') = (a, b, b + D2): the form (x, y, y + D). (P1, P2', P1') =
  // (a, u2, a + D1) reversed and rotated is (u2, a, a + D1), the same form
  // with the opposite sign.
  const double t1 = thin_angle(ax, ay, az, bx, by, bz, D2x, D2y, D2z);
  const double t2 = -thin_angle(u2x, u2y, u2z, ax, ay, az, D1x, D1y, D1z);
  return (double)E.w * (t1 + t2);
}

// Fast-path angle of edge E for one lane; ok = inside the first-order range.
__device__ __forceinline__ float ribbon_fast(const RibbonEdge& E, float ox, float oy, float oz,
                                             float eox, float eoy, float eoz, float eos,
                                             bool& ok) {
  const float u1x = E.p.x - ox, u1y = E.p.y - oy, u1z = E.p.z - oz;
  const float u2x = E.q.x - ox, u2y = E.q.y - oy, u2z = E.q.z - oz;
  const float l1s = __builtin_fmaf(u1z, u1z, __builtin_fmaf(u1y, u1y, u1x * u1x));
  const float l2s = __builtin_fmaf(u2z, u2z, __builtin_fmaf(u2y, u2y, u2x * u2x));
  const float i1 = __builtin_amdgcn_rsqf(l1s), i2 = __builtin_amdgcn_rsqf(l2s);
  const float l1 = l1s * i1, l2 = l2s * i2;
  const float d = __builtin_fmaf(u1z, u2z, __builtin_fmaf(u1y, u2y, u1x * u2x));
  const float t = __builtin_fmaf(l1, l2, d);
  const float nx = __builtin_fmaf(u1y, u2z, -u1z * u2y);
  const float ny = __builtin_fmaf(u1z, u2x, -u1x * u2z);
  const float nz = __builtin_fmaf(u1x, u2y, -u1y * u2x);
  const float sx = __builtin_fmaf(E.e.x - eox, i1, (E.e.w - eox) * i2);
  const float sy = __builtin_fmaf(E.e.y - eoy, i1, (E.f.x - eoy) * i2);
  const float sz = __builtin_fmaf(E.e.z - eoz, i1, (E.f.y - eoz) * i2);
  const float num = __builtin_fmaf(nz, sz, __builtin_fmaf(ny, sy, nx * sx));
  ok = (32.0f * t >= l1 * l2) && (E.q.w + eos < fminf(l1, l2));
  return num * E.p.w * __builtin_amdgcn_rcpf(t);
}

// Sum over edges [e0, e1) of the block's staged tiles, for QPL queries per
// lane. Begins and ends with a barrier. SRC 0: edges as record index pairs
// into rec (double4 per record); SRC 1: (face, corner) of the 9-row face SoA
// tri (row stride `stride`), the edge running corner c -> corner c+1 mod 3.
template <int SRC, int QPL>
__device__ __forceinline__ void ribbon_sweep(const double* __restrict__ pts, int64_t stride,
                                             const int32_t* __restrict__ ids,
                                             const double* __restrict__ wts, int64_t e0,
                                             int64_t e1, double gx, double gy, double gz,
                                             const float (&ox)[QPL], const float (&oy)[QPL],
                                             const float (&oz)[QPL], const float (&eox)[QPL],
                                             const float (&eoy)[QPL], const float (&eoz)[QPL],
                                             const float (&eos)[QPL], double (&acc)[QPL],
                                             RibbonSmem& sm) {
  const int tid = threadIdx.x;
  for (int64_t t0 = e0; t0 < e1; t0 += kRibbonTile) {
    const int cnt = (int)(e1 - t0 < kRibbonTile ? e1 - t0 : kRibbonTile);
    __syncthreads();
    if (tid < cnt) {
      const int64_t i = t0 + tid;
      const int64_t a = ids[2 * i], b = ids[2 * i + 1];
      double p1x, p1y, p1z, p2x, p2y, p2z;
      if (SRC == 0) {
        const double* r1 = pts + 4 * a;
        const double* r2 = pts + 4 * b;
        p1x = r1[0]; p1y = r1[1]; p1z = r1[2];
        p2x = r2[0]; p2y = r2[1]; p2z = r2[2];
      } else {
        const int64_t c1 = b, c2 = b == 2 ? 0 : b + 1;
        p1x = pts[(3 * c1) * stride + a];
        p1y = pts[(3 * c1 + 1) * stride + a];
        p1z = pts[(3 * c1 + 2) * stride + a];
        p2x = pts[(3 * c2) * stride + a];
        p2y = pts[(3 * c2 + 1) * stride + a];
        p2z = pts[(3 * c2 + 2) * stride + a];
      }
      ribbon_stage(p1x, p1y, p1z, p2x, p2y, p2z, wts[i], gx, gy, gz, sm.edge[tid]);
    }
    __syncthreads();
    float part[QPL];
#pragma unroll
    for (int k = 0; k < QPL; ++k) part[k] = 0.0f;
    for (int j = 0; j < cnt; ++j) {
      const RibbonEdge& E = sm.edge[j];
#pragma unroll
      for (int k = 0; k < QPL; ++k) {
        bool ok;
        const float th = ribbon_fast(E, ox[k], oy[k], oz[k], eox[k], eoy[k], eoz[k], eos[k], ok);
        if (__builtin_expect(wave_all(ok), 1)) {
          part[k] += th;
        } else {
          asm volatile("" ::);  // keep a real wave-uniform branch
          acc[k] += ribbon_slow(E, ox[k], oy[k], oz[k], eox[k], eoy[k], eoz[k]);
        }
      }
    }
#pragma unroll
    for (int k = 0; k < QPL; ++k) acc[k] += (double)part[k];
  }
  __syncthreads();
}

// Block origin exactly as the forward kernels form it: the centre of the bbox
// of the block's valid queries (min / max are exact, so the order of the
// reduction does not matter). Begins and ends with a barrier.
template <int QPL>
__device__ __forceinline__ void ribbon_origin(const double (&Ox)[QPL], const double (&Oy)[QPL],
                                              const double (&Oz)[QPL], const bool (&valid)[QPL],
                                              RibbonSmem& sm, double& gx, double& gy,
                                              double& gz) {
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
  __syncthreads();
}

// One block: queries [slot0, slot_end) in blocks of 256 x QPL (slot0 + k *
// 256 + tid), edges [e0, e1). Writes the correction (angle units) to res[k].
template <int SRC, int QPL>
__device__ __forceinline__ void ribbon_block(const RibbonSrc& src, const double* __restrict__ q,
                                             int64_t slot0, int64_t slot_end, int64_t e0,
                                             int64_t e1, double (&res)[QPL], bool (&valid)[QPL],
                                             RibbonSmem& sm) {
  const int tid = threadIdx.x;
  double Ox[QPL], Oy[QPL], Oz[QPL];
#pragma unroll
  for (int k = 0; k < QPL; ++k) {
    const int64_t slot = slot0 + k * kBlockThreads + tid;
    valid[k] = slot < slot_end;
    Ox[k] = valid[k] ? q[3 * slot] : 0.0;
    Oy[k] = valid[k] ? q[3 * slot + 1] : 0.0;
    Oz[k] = valid[k] ? q[3 * slot + 2] : 0.0;
    res[k] = 0.0;
  }
  double gx, gy, gz;
  ribbon_origin<QPL>(Ox, Oy, Oz, valid, sm, gx, gy, gz);
  float ox[QPL], oy[QPL], oz[QPL], eox[QPL], eoy[QPL], eoz[QPL], eos[QPL];
#pragma unroll
  for (int k = 0; k < QPL; ++k) {
    const double dx = Ox[k] - gx, dy = Oy[k] - gy, dz = Oz[k] - gz;
    ox[k] = valid[k] ? (float)dx : 0.0f;
    oy[k] = valid[k] ? (float)dy : 0.0f;
    oz[k] = valid[k] ? (float)dz : 0.0f;
    eox[k] = valid[k] ? (float)((double)ox[k] - dx) : 0.0f;
    eoy[k] = valid[k] ? (float)((double)oy[k] - dy) : 0.0f;
    eoz[k] = valid[k] ? (float)((double)oz[k] - dz) : 0.0f;
    eos[k] = fmaxf(fmaxf(fabsf(eox[k]), fabsf(eoy[k])), fabsf(eoz[k])) * kRibbonFar;
  }
  ribbon_sweep<SRC, QPL>(src.pts, src.stride, src.ids, src.w, e0, e1, gx, gy, gz,
                         ox, oy, oz, eox, eoy, eoz, eos, res, sm);
}

// Single mesh: grid (query blocks, edge splits); edges [0, *count) shared
// evenly by the splits. One split adds w += sum / (2 pi) through perm; more
// write partials part[split * nq + slot] for ribbon_reduce_kernel.
template <int SRC, int QPL>
__global__ __launch_bounds__(kBlockThreads) __attribute__((amdgpu_waves_per_eu(8))) void ribbon_kernel(
    RibbonSrc src, const double* __restrict__ q, const int32_t* __restrict__ perm,
    double* __restrict__ dst, double* __restrict__ part, int64_t nq) {
  __shared__ RibbonSmem sm;
  const int64_t ne = src.count ? (int64_t)*src.count : src.n;
  const int splits = (int)gridDim.y;
  const int64_t per = (ne + splits - 1) / splits;
  const int64_t e0 = (int64_t)blockIdx.y * per;
  const int64_t e1 = e0 + per < ne ? e0 + per : ne;
  const int64_t slot0 = (int64_t)blockIdx.x * (kBlockThreads * QPL);
  if (splits == 1 && e0 >= e1) return;  // block-uniform: nothing to add
  double res[QPL];
  bool valid[QPL];
  if (e0 < e1) {
    ribbon_block<SRC, QPL>(src, q, slot0, nq, e0, e1, res, valid, sm);
  } else {
#pragma unroll
    for (int k = 0; k < QPL; ++k) {
      res[k] = 0.0;
      valid[k] = slot0 + k * kBlockThreads + (int64_t)threadIdx.x < nq;
    }
  }
#pragma unroll
  for (int k = 0; k < QPL; ++k) {
    const int64_t slot = slot0 + k * kBlockThreads + threadIdx.x;
    if (!valid[k]) continue;
    if (splits == 1)
      dst[perm ? (int64_t)perm[slot] : slot] += res[k] * kInv2Pi;
    else
      part[(int64_t)blockIdx.y * nq + slot] = res[k];
  }
}

__global__ __launch_bounds__(kBlockThreads) void ribbon_reduce_kernel(
    const double* __restrict__ part, const int32_t* __restrict__ perm, double* __restrict__ dst,
    int64_t nq, int splits) {
  const int64_t slot = (int64_t)blockIdx.x * kBlockThreads + threadIdx.x;
  if (slot >= nq) return;
  double s = 0.0;
  for (int j = 0; j < splits; ++j) s += part[(int64_t)j * nq + slot];
  dst[perm ? (int64_t)perm[slot] : slot] += s * kInv2Pi;
}

// Batch: one block per (mesh, query tile) of the plan's table; the mesh's
// edges [e_begin[m], e_end[m]).
template <int SRC, int QPL>
__global__ __launch_bounds__(kBlockThreads) __attribute__((amdgpu_waves_per_eu(8))) void ribbon_batch_kernel(
    RibbonSrc src, const int32_t* __restrict__ e_begin, const int32_t* __restrict__ e_end,
    const int2* __restrict__ blocks, const BatchMesh* __restrict__ meshes,
    const double* __restrict__ q, const int32_t* __restrict__ perm, double* __restrict__ dst) {
  __shared__ RibbonSmem sm;
  const int2 b = blocks[blockIdx.x];
  const BatchMesh m = meshes[b.x];
  const int64_t e0 = e_begin[b.x], e1 = e_end[b.x];
  if (e0 >= e1) return;  // block-uniform
  const int64_t slot0 = m.query_off + (int64_t)b.y * (kBlockThreads * QPL);
  double res[QPL];
  bool valid[QPL];
  ribbon_block<SRC, QPL>(src, q, slot0, m.query_off + m.nq, e0, e1, res, valid, sm);
#pragma unroll
  for (int k = 0; k < QPL; ++k) {
    const int64_t slot = slot0 + k * kBlockThreads + threadIdx.x;
    if (valid[k]) dst[perm ? (int64_t)perm[slot] : slot] += res[k] * kInv2Pi;
  }
}

// ---------------------------------------------------------------------------
// Boundary edges of a face SoA from corner coordinates.
//
// Directed face edge e = 3 f + c runs corner c -> corner c + 1 mod 3 of face
// f. Its undirected key hashes (segment, lower endpoint, upper endpoint)
// (lexicographic order of (x, y, z), -0 taken as +0); sign = +1 if it runs
// lower -> upper, -1 the other way, 0 for a degenerate edge. After a radix
// sort by key, each run of equal keys is split into groups of exactly equal
// coordinates (a hash collision only costs a longer run); a group whose
// signs do not cancel is one boundary edge, emitted as its first member
// (sorted order, so deterministic) with weight net x sign(member). The
// segment occupies the key's top 24 bits, so a batch's edges come out
// grouped by mesh.
// ---------------------------------------------------------------------------

struct EdgeEnds {
  double p[3], r[3];
};

__device__ __forceinline__ EdgeEnds edge_ends(const double* __restrict__ tri, int64_t stride,
                                              int64_t e) {
  const int64_t f = e / 3;
  const int c = (int)(e - 3 * f);
  const int c2 = c == 2 ? 0 : c + 1;
  EdgeEnds x;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    x.p[d] = tri[(3 * c + d) * stride + f] + 0.0;  // -0 -> +0
    x.r[d] = tri[(3 * c2 + d) * stride + f] + 0.0;
  }
  return x;
}

// -1, 0, +1: lexicographic comparison of a and b
__device__ __forceinline__ int lex_cmp(const double* a, const double* b) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (a[d] < b[d]) return -1;
    if (a[d] > b[d]) return 1;
  }
  return 0;
}

__device__ __forceinline__ uint64_t mix64(uint64_t h, uint64_t v) {
  h ^= v;
  h *= 0xbf58476d1ce4e5b9ull;
  h ^= h >> 31;
  h *= 0x94d049bb133111ebull;
  h ^= h >> 29;
  return h;
}

__device__ __forceinline__ int seg_of_face(const int64_t* __restrict__ face_off, int nseg,
                                           int64_t f) {
  if (!face_off) return 0;
  int lo = 0, hi = nseg;  // face_off[lo] <= f < face_off[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (face_off[mid] <= f) lo = mid; else hi = mid;
  }
  return lo;
}

// Undirected (lower, upper) ends and the sign of edge e.
__device__ __forceinline__ int edge_canon(const double* __restrict__ tri, int64_t stride,
                                          int64_t e, double* lo, double* hi) {
  const EdgeEnds x = edge_ends(tri, stride, e);
  const int c = lex_cmp(x.p, x.r);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    lo[d] = c <= 0 ? x.p[d] : x.r[d];
    hi[d] = c <= 0 ? x.r[d] : x.p[d];
  }
  return c < 0 ? 1 : (c > 0 ? -1 : 0);
}

__global__ __launch_bounds__(kBlockThreads) void edge_key_kernel(
    const double* __restrict__ tri, int64_t stride, int64_t ne, const int64_t* __restrict__ face_off,
    int nseg, unsigned long long* __restrict__ keys, int32_t* __restrict__ vals) {
  const int64_t e = (int64_t)blockIdx.x * kBlockThreads + threadIdx.x;
  if (e >= ne) return;
  double lo[3], hi[3];
  (void)edge_canon(tri, stride, e, lo, hi);
  uint64_t h = 0x243f6a8885a308d3ull;
#pragma unroll
  for (int d = 0; d < 3; ++d) h = mix64(h, (uint64_t)__double_as_longlong(lo[d]));
#pragma unroll
  for (int d = 0; d < 3; ++d) h = mix64(h, (uint64_t)__double_as_longlong(hi[d]));
  const uint64_t seg = (uint64_t)seg_of_face(face_off, nseg, e / 3);
  keys[e] = (seg << 40) | (h >> 24);
  vals[e] = (int32_t)e;
}

// For the run of equal keys starting at sorted position i (0 elsewhere): the
// number of boundary edges it holds; with `emit`, also writes them from
// out_off[i] on.
template <bool EMIT>
__global__ __launch_bounds__(kBlockThreads) void edge_net_kernel(
    const double* __restrict__ tri, int64_t stride, int64_t ne,
    const unsigned long long* __restrict__ keys, const int32_t* __restrict__ vals,
    int32_t* __restrict__ counts, const int32_t* __restrict__ out_off, int32_t* __restrict__ ids,
    double* __restrict__ w) {
  const int64_t i = (int64_t)blockIdx.x * kBlockThreads + threadIdx.x;
  if (i >= ne) return;
  if (i > 0 && keys[i] == keys[i - 1]) {
    if (!EMIT) counts[i] = 0;
    return;
  }
  int64_t end = i + 1;
  while (end < ne && keys[end] == keys[i]) ++end;
  int32_t n = 0;
  for (int64_t a = i; a < end; ++a) {
    double alo[3], ahi[3];
    const int sa = edge_canon(tri, stride, vals[a], alo, ahi);
    bool first = true;
    for (int64_t b = i; b < a && first; ++b) {
      double blo[3], bhi[3];
      (void)edge_canon(tri, stride, vals[b], blo, bhi);
      first = !(lex_cmp(alo, blo) == 0 && lex_cmp(ahi, bhi) == 0);
    }
    if (!first) continue;
    int net = sa;
    for (int64_t b = a + 1; b < end; ++b) {
      double blo[3], bhi[3];
      const int sb = edge_canon(tri, stride, vals[b], blo, bhi);
      if (lex_cmp(alo, blo) == 0 && lex_cmp(ahi, bhi) == 0) net += sb;
    }
    if (net == 0) continue;
    if (EMIT) {
      const int64_t o = out_off[i] + n;
      const int64_t e = vals[a];
      ids[2 * o] = (int32_t)(e / 3);
      ids[2 * o + 1] = (int32_t)(e - 3 * (e / 3));
      w[o] = (double)(net * sa);
    }
    ++n;
  }
  if (!EMIT) counts[i] = n;
}

__global__ void edge_total_kernel(const int32_t* __restrict__ counts,
                                  const int32_t* __restrict__ offs, int64_t ne,
                                  int32_t* __restrict__ total) {
  if (threadIdx.x == 0 && blockIdx.x == 0) total[0] = ne > 0 ? offs[ne - 1] + counts[ne - 1] : 0;
}

__global__ __launch_bounds__(kBlockThreads) void fill_i32_kernel(int32_t* __restrict__ p,
                                                                 int64_t n, int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * kBlockThreads + threadIdx.x;
  if (i < n) p[i] = v;
}

// Per-segment [begin, end) of the emitted edges (grouped by segment).
__global__ __launch_bounds__(kBlockThreads) void edge_bounds_kernel(
    const int32_t* __restrict__ ids, const int32_t* __restrict__ total,
    const int64_t* __restrict__ face_off, int nseg, int32_t* __restrict__ e_begin,
    int32_t* __restrict__ e_end, int64_t cap) {
  const int64_t i = (int64_t)blockIdx.x * kBlockThreads + threadIdx.x;
  const int64_t n = total[0];
  if (i >= n || i >= cap) return;
  const int s = seg_of_face(face_off, nseg, ids[2 * i]);
  if (i == 0 || seg_of_face(face_off, nseg, ids[2 * i - 2]) != s) e_begin[s] = (int32_t)i;
  if (i == n - 1 || seg_of_face(face_off, nseg, ids[2 * i + 2]) != s) e_end[s] = (int32_t)(i + 1);
}

// Key bits the sort must order: the 40-bit hash plus the segment's bits.
int key_bits(int nseg) {
  int b = 40;
  while (nseg > 1 && (1 << (b - 40)) < nseg) ++b;
  return b;
}

size_t sort_temp64(int64_t n, int bits) {
  size_t t = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, t, (const unsigned long long*)nullptr,
                                     (unsigned long long*)nullptr, (const int32_t*)nullptr,
                                     (int32_t*)nullptr, (int)n, 0, bits);
  return t;
}

size_t scan_temp(int64_t n) {
  size_t t = 0;
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, t, (const int32_t*)nullptr, (int32_t*)nullptr,
                                         (int)n);
  return t;
}

unsigned grid_of(int64_t n) { return (unsigned)((n + kBlockThreads - 1) / kBlockThreads); }

}  // namespace

size_t soa_boundary_workspace_bytes(int64_t nf, int nseg) {
  const int64_t ne = 3 * nf;
  if (ne == 0) return align_up(sizeof(int32_t)) + 2 * align_up(sizeof(int32_t) * (size_t)nseg);
  size_t b = 0;
  b += 2 * align_up(sizeof(unsigned long long) * (size_t)ne);  // keys in / out
  b += 2 * align_up(sizeof(int32_t) * (size_t)ne);             // vals in / out
  b += 2 * align_up(sizeof(int32_t) * (size_t)ne);             // counts, offsets
  b += align_up(std::max(sort_temp64(ne, key_bits(nseg)), scan_temp(ne)));
  b += 2 * align_up(sizeof(int32_t) * (size_t)ne);  // ids (2 x ne)
  b += align_up(sizeof(double) * (size_t)ne);       // weights
  b += align_up(sizeof(int32_t));                   // count
  b += 2 * align_up(sizeof(int32_t) * (size_t)std::max(nseg, 1));  // per-segment ranges
  return b;
}

hipError_t soa_boundary_launch(const double* tri, int64_t nf, const int64_t* face_off, int nseg,
                               void* ws, size_t ws_bytes, SoaBoundary* out, hipStream_t s) {
  if (ws_bytes < soa_boundary_workspace_bytes(nf, nseg)) return hipErrorInvalidValue;
  const int64_t ne = 3 * nf;
  char* c = (char*)ws;
  SoaBoundary r{};
  r.cap = ne;
  if (ne == 0) {
    r.count = (int32_t*)c; c += align_up(sizeof(int32_t));
    r.e_begin = (int32_t*)c; c += align_up(sizeof(int32_t) * (size_t)nseg);
    r.e_end = (int32_t*)c;
    fill_i32_kernel<<<1, kBlockThreads, 0, s>>>(r.count, 1, 0);
    if (nseg > 0) {
      fill_i32_kernel<<<grid_of(nseg), kBlockThreads, 0, s>>>(r.e_begin, nseg, 0);
      fill_i32_kernel<<<grid_of(nseg), kBlockThreads, 0, s>>>(r.e_end, nseg, 0);
    }
    *out = r;
    return hipGetLastError();
  }
  auto take = [&c](size_t bytes) { char* p = c; c += align_up(bytes); return (void*)p; };
  unsigned long long* k_in = (unsigned long long*)take(sizeof(unsigned long long) * ne);
  unsigned long long* k_out = (unsigned long long*)take(sizeof(unsigned long long) * ne);
  int32_t* v_in = (int32_t*)take(sizeof(int32_t) * ne);
  int32_t* v_out = (int32_t*)take(sizeof(int32_t) * ne);
  int32_t* counts = (int32_t*)take(sizeof(int32_t) * ne);
  int32_t* offs = (int32_t*)take(sizeof(int32_t) * ne);
  const int bits = key_bits(face_off ? nseg : 1);
  const size_t temp_b = std::max(sort_temp64(ne, bits), scan_temp(ne));
  void* temp = take(temp_b);
  r.ids = (int32_t*)take(2 * sizeof(int32_t) * ne);
  r.w = (double*)take(sizeof(double) * ne);
  r.count = (int32_t*)take(sizeof(int32_t));
  r.e_begin = (int32_t*)take(sizeof(int32_t) * (size_t)std::max(nseg, 1));
  r.e_end = (int32_t*)take(sizeof(int32_t) * (size_t)std::max(nseg, 1));

  edge_key_kernel<<<grid_of(ne), kBlockThreads, 0, s>>>(tri, nf, ne, face_off, nseg, k_in, v_in);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  size_t tb = temp_b;
  e = hipcub::DeviceRadixSort::SortPairs(temp, tb, k_in, k_out, v_in, v_out, (int)ne, 0, bits, s);
  if (e != hipSuccess) return e;
  edge_net_kernel<false><<<grid_of(ne), kBlockThreads, 0, s>>>(tri, nf, ne, k_out, v_out, counts,
                                                                nullptr, nullptr, nullptr);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  tb = temp_b;
  e = hipcub::DeviceScan::ExclusiveSum(temp, tb, counts, offs, (int)ne, s);
  if (e != hipSuccess) return e;
  edge_total_kernel<<<1, 64, 0, s>>>(counts, offs, ne, r.count);
  edge_net_kernel<true><<<grid_of(ne), kBlockThreads, 0, s>>>(tri, nf, ne, k_out, v_out, nullptr,
                                                               offs, r.ids, r.w);
  if (nseg > 0 && face_off) {
    fill_i32_kernel<<<grid_of(nseg), kBlockThreads, 0, s>>>(r.e_begin, nseg, 0);
    fill_i32_kernel<<<grid_of(nseg), kBlockThreads, 0, s>>>(r.e_end, nseg, 0);
    edge_bounds_kernel<<<grid_of(ne), kBlockThreads, 0, s>>>(r.ids, r.count, face_off, nseg,
                                                             r.e_begin, r.e_end, ne);
  }
  *out = r;
  return hipGetLastError();
}

int ribbon_splits(int64_t nq, int qpl, int64_t ne_bound) {
  const int64_t qb = (nq + (int64_t)kBlockThreads * qpl - 1) / ((int64_t)kBlockThreads * qpl);
  int64_t s = (512 + qb - 1) / std::max<int64_t>(qb, 1);
  const int64_t tiles = (ne_bound + kRibbonTile - 1) / kRibbonTile;
  s = std::min<int64_t>(s, std::max<int64_t>(tiles, 1));
  return (int)std::max<int64_t>(1, std::min<int64_t>(s, 64));
}

size_t ribbon_workspace_bytes(int64_t nq, int splits) {
  return splits > 1 ? align_up(sizeof(double) * (size_t)nq * (size_t)splits) : 0;
}

hipError_t ribbon_launch(const RibbonSrc& src, int soa, int qpl, const double* q,
                         const int32_t* perm, double* dst, int64_t nq, int splits, void* ws,
                         size_t ws_bytes, hipStream_t s) {
  if (nq <= 0) return hipSuccess;
  if (!src.count && src.n <= 0) return hipSuccess;
  if (ws_bytes < ribbon_workspace_bytes(nq, splits)) return hipErrorInvalidValue;
  double* part = splits > 1 ? (double*)ws : nullptr;
  const int64_t qb = (nq + (int64_t)kBlockThreads * qpl - 1) / ((int64_t)kBlockThreads * qpl);
  const dim3 grid((unsigned)qb, (unsigned)splits);
#define WN_RIBBON(SRC, QPL) \
  ribbon_kernel<SRC, QPL><<<grid, kBlockThreads, 0, s>>>(src, q, perm, dst, part, nq)
  if (soa) {
    if (qpl == 1) WN_RIBBON(1, 1); else if (qpl == 2) WN_RIBBON(1, 2); else WN_RIBBON(1, 4);
  } else {
    WN_RIBBON(0, 1);
  }
#undef WN_RIBBON
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || splits == 1) return e;
  ribbon_reduce_kernel<<<grid_of(nq), kBlockThreads, 0, s>>>(part, perm, dst, nq, splits);
  return hipGetLastError();
}

hipError_t ribbon_batch_launch(const RibbonSrc& src, int soa, int qpl, const int32_t* e_begin,
                               const int32_t* e_end, const int2* blocks, int64_t nblocks,
                               const BatchMesh* meshes, const double* q, const int32_t* perm,
                               double* dst, hipStream_t s) {
  if (nblocks <= 0) return hipSuccess;
#define WN_RIBBON_B(SRC, QPL)                                                              \
  ribbon_batch_kernel<SRC, QPL><<<dim3((unsigned)nblocks), kBlockThreads, 0, s>>>(         \
      src, e_begin, e_end, blocks, meshes, q, perm, dst)
  if (soa) {
    if (qpl == 1) WN_RIBBON_B(1, 1); else WN_RIBBON_B(1, 2);
  } else {
    WN_RIBBON_B(0, 1);
  }
#undef WN_RIBBON_B
  return hipGetLastError();
}

}  // namespace wn
