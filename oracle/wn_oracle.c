/*
 * wn_oracle.c -- TEST INFRASTRUCTURE ONLY. This is the correctness oracle and
 * the CPU baseline for the generalized winding number; it is never linked into,
 * loaded by, or called from the product library. Only tests/, the smoke check
 * in __graft_entry__.py and the cpu_baseline leg of bench.py may use it.
 *
 * It restates, in plain C and FP64 throughout, the reference CPU path:
 *
 *   winding_one   src/capi.mojo:205-286   per-query sweep over all faces
 *   winding       src/capi.mojo:289-294   query span [q0, q1)
 *   mli_winding   src/capi.mojo:938-942   C-ABI arguments (tri, q, dst, nf, nq, q0, q1)
 *
 * Summation order follows the reference exactly: W = 4 SIMD lanes (the AVX2
 * float64 width of the published Broadwell host; simd_width_of at
 * src/capi.mojo:206) times L = 4 partial accumulators over blocks of L*W faces
 * (src/capi.mojo:215-244), then W-wide blocks into the first accumulator
 * (:245-264), then ((t0 + t1) + (t2 + t3)).reduce_add() (:265), then a scalar
 * tail (:266-285), then division by 2*pi (:286). reduce_add of a 4-lane vector
 * is taken as the halving tree (v0 + v2) + (v1 + v3).
 *
 * atan2 and sqrt are the C library's (the reference calls Mojo std.math,
 * src/capi.mojo:9, whose source is not vendored; agreement with it is to
 * rounding, which the 1e-10 FP64 contract does not resolve).
 *
 * wn_oracle_ref_gpu_atan2 restates the reference GPU kernel's own arctangent
 * (src/capi.mojo:326-352: range reduction at tan(3pi/8), tan(pi/8), then a
 * 20-term Maclaurin series, then quadrant fix-ups), including its signed-zero
 * behaviour (det = -0.0, den < 0 gives +pi, where C's atan2 gives -pi).
 *
 * Build: oracle/Makefile (-O2 -ffp-contract=off: source-order IEEE arithmetic).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>

#define WN_W 4
#define WN_L 4

static const double WN_PI = 3.14159265358979323846264338327950288;

/* One query-face pair: src/capi.mojo:218-235 (identical to the GPU formula at
 * :309-325). tri is the 9 x nf SoA of face corners, rows ax,ay,az,bx,by,bz,cx,
 * cy,cz (python/igl/__init__.py:248-250). */
static inline double pair_angle(const double *tri, int64_t nf, int64_t f,
                                double ox, double oy, double oz) {
    double ax = tri[f] - ox;
    double ay = tri[nf + f] - oy;
    double az = tri[2 * nf + f] - oz;
    double bx = tri[3 * nf + f] - ox;
    double by = tri[4 * nf + f] - oy;
    double bz = tri[5 * nf + f] - oz;
    double cx = tri[6 * nf + f] - ox;
    double cy = tri[7 * nf + f] - oy;
    double cz = tri[8 * nf + f] - oz;
    double la = sqrt(ax * ax + ay * ay + az * az);
    double lb = sqrt(bx * bx + by * by + bz * bz);
    double lc = sqrt(cx * cx + cy * cy + cz * cz);
    double det = ax * (by * cz - bz * cy) - ay * (bx * cz - bz * cx) +
                 az * (bx * cy - by * cx);
    double den = la * lb * lc;
    den += (ax * bx + ay * by + az * bz) * lc;
    den += (bx * cx + by * cy + bz * cz) * la;
    den += (cx * ax + cy * ay + cz * az) * lb;
    return atan2(det, den);
}

double wn_oracle_pair_angle(const double *tri, int64_t nf, int64_t f,
                            double ox, double oy, double oz) {
    return pair_angle(tri, nf, f, ox, oy, oz);
}

/* src/capi.mojo:205-286 */
static double winding_one(const double *tri, double ox, double oy, double oz,
                          int64_t nf) {
    double t[WN_L][WN_W];
    for (int l = 0; l < WN_L; ++l)
        for (int k = 0; k < WN_W; ++k) t[l][k] = 0.0;
    int64_t fi = 0;
    while (fi + WN_L * WN_W <= nf) {
        for (int l = 0; l < WN_L; ++l) {
            int64_t at = fi + (int64_t)l * WN_W;
            for (int k = 0; k < WN_W; ++k)
                t[l][k] += pair_angle(tri, nf, at + k, ox, oy, oz);
        }
        fi += WN_L * WN_W;
    }
    while (fi + WN_W <= nf) {
        for (int k = 0; k < WN_W; ++k)
            t[0][k] += pair_angle(tri, nf, fi + k, ox, oy, oz);
        fi += WN_W;
    }
    double s[WN_W];
    for (int k = 0; k < WN_W; ++k) s[k] = (t[0][k] + t[1][k]) + (t[2][k] + t[3][k]);
    double total = (s[0] + s[2]) + (s[1] + s[3]);
    while (fi < nf) {
        total += pair_angle(tri, nf, fi, ox, oy, oz);
        fi += 1;
    }
    return total / (2.0 * WN_PI);
}

/* src/capi.mojo:289-294 and the mli_winding ABI at :938-942: writes only
 * dst[q0:q1]. */
void wn_oracle_winding(const double *tri, const double *q, double *dst,
                       int64_t nf, int64_t nq, int64_t q0, int64_t q1) {
    (void)nq;
    for (int64_t qi = q0; qi < q1; ++qi)
        dst[qi] = winding_one(tri, q[qi * 3], q[qi * 3 + 1], q[qi * 3 + 2], nf);
}

/* Thread fan-out for the CPU baseline: contiguous spans of ceil(nq/threads)
 * queries, the partition of python/igl/_lib.py:64-95. */
typedef struct {
    const double *tri, *q;
    double *dst;
    int64_t nf, nq, q0, q1;
} span_args;

static void *span_main(void *p) {
    span_args *a = (span_args *)p;
    wn_oracle_winding(a->tri, a->q, a->dst, a->nf, a->nq, a->q0, a->q1);
    return NULL;
}

int64_t wn_oracle_winding_threads(const double *tri, const double *q, double *dst,
                                  int64_t nf, int64_t nq, int64_t threads) {
    if (threads < 1) threads = 1;
    if (threads > nq) threads = nq > 0 ? nq : 1;
    if (threads == 1) {
        wn_oracle_winding(tri, q, dst, nf, nq, 0, nq);
        return 1;
    }
    int64_t step = (nq + threads - 1) / threads;
    pthread_t *tid = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
    span_args *args = (span_args *)malloc(sizeof(span_args) * (size_t)threads);
    int64_t used = 0;
    for (int64_t lo = 0; lo < nq; lo += step, ++used) {
        args[used].tri = tri;
        args[used].q = q;
        args[used].dst = dst;
        args[used].nf = nf;
        args[used].nq = nq;
        args[used].q0 = lo;
        args[used].q1 = lo + step < nq ? lo + step : nq;
        pthread_create(&tid[used], NULL, span_main, &args[used]);
    }
    for (int64_t i = 0; i < used; ++i) pthread_join(tid[i], NULL);
    free(tid);
    free(args);
    return used;
}

/* src/capi.mojo:326-352: the reference GPU kernel's own arctangent of
 * det/den with its quadrant logic. */
double wn_oracle_ref_gpu_atan2(double det, double den) {
    double ratio = den != 0.0 ? det / den : 0.0;
    double magnitude = fabs(ratio);
    double offset = 0.0;
    if (magnitude > 2.414213562373095) {
        magnitude = 1.0 / magnitude;
        offset = WN_PI / 2.0;
    } else if (magnitude > 0.414213562373095) {
        magnitude = (magnitude - 1.0) / (magnitude + 1.0);
        offset = WN_PI / 4.0;
    }
    double z2 = magnitude * magnitude;
    double polynomial = -1.0 / 39.0;
    for (int term = 18; term >= 0; --term) {
        double coefficient = 1.0 / (double)(2 * term + 1);
        if (term % 2 == 1) coefficient = -coefficient;
        polynomial = coefficient + z2 * polynomial;
    }
    double angle = magnitude * polynomial;
    if (offset == WN_PI / 2.0)
        angle = offset - angle;
    else
        angle += offset;
    if (ratio < 0.0) angle = -angle;
    if (den < 0.0)
        angle += det >= 0.0 ? WN_PI : -WN_PI;
    else if (den == 0.0)
        angle = det > 0.0 ? WN_PI / 2.0 : (det < 0.0 ? -WN_PI / 2.0 : 0.0);
    return angle;
}
