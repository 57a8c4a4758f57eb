"""Derive the arctangent polynomials of the kernels (wn_device.h, wn_strip64.hip).

Method: Remez exchange at 40 significant digits (mpmath) for
    atan(x) ~= x * P(x^2),   x in (0, R],
minimising the maximum RELATIVE error |x P(x^2) / atan(x) - 1|, so that small
angles (the far faces of a winding-number sum) keep relative accuracy.

    python gen_atan_coeffs.py
      n = 9 terms on (0, 1]    -> atan2_f32          (max rel. error 1.52e-8)
      n = 3 terms on (0, 1/8]  -> atan_small3_f32, atan_small_x8_f32  (max rel. error 1.67e-8)
      n = 7 terms on (0, 1/8]  -> atan_small_x8_f64  (max rel. error 1.8e-18)
      n = 5 terms on (0, 0.4]  -> atan_sum_x8_f32    (max rel. error 1.30e-8)
      n = 11 terms on (0, 0.4] -> atan_sum_x8_f64    (max rel. error 1.6e-17)
"""

import mpmath as mp

mp.mp.dps = 40


def remez(n: int, R, iters: int = 40, grid: int = 6000):
    m = n + 1
    xs = [R * mp.mpf(0.5) * (1 - mp.cos(mp.pi * (i + 0.5) / m)) for i in range(m)]
    err = None
    c = None
    for _ in range(iters):
        A = mp.matrix(m, m)
        b = mp.matrix(m, 1)
        for i, x in enumerate(xs):
            t = mp.atan(x)
            for k in range(n):
                A[i, k] = x ** (2 * k + 1) / t
            A[i, n] = (-1) ** i
            b[i] = 1
        sol = mp.lu_solve(A, b)
        c = [sol[k] for k in range(n)]
        pts = [R * mp.mpf(i + 1) / grid for i in range(grid)]
        vals = [sum(c[k] * x ** (2 * k + 1) for k in range(n)) / mp.atan(x) - 1 for x in pts]
        ext = [0] + [i for i in range(1, grid - 1)
                     if (vals[i] - vals[i - 1]) * (vals[i + 1] - vals[i]) <= 0] + [grid - 1]
        groups = []
        for i in ext:  # keep alternating-sign extrema of largest magnitude
            if groups and (vals[i] > 0) == (vals[groups[-1]] > 0):
                if abs(vals[i]) > abs(vals[groups[-1]]):
                    groups[-1] = i
            else:
                groups.append(i)
        while len(groups) > m:
            if abs(vals[groups[0]]) < abs(vals[groups[-1]]):
                groups.pop(0)
            else:
                groups.pop()
        if len(groups) == m:
            xs = [pts[i] for i in groups]
        err = max(abs(v) for v in vals)
    return c, err


if __name__ == "__main__":
    for n, R in ((9, mp.mpf(1)), (3, mp.mpf(1) / 8), (7, mp.mpf(1) / 8), (5, mp.mpf("0.4")),
                 (11, mp.mpf("0.4"))):
        c, err = remez(n, R)
        print(f"n={n} R={mp.nstr(R, 3)} max relative error {mp.nstr(err, 4)}")
        print("  coefficients (x^1, x^3, ...):", ", ".join(mp.nstr(v, 17) for v in c))
