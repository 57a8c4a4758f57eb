"""Regenerate tests/golden/winding_golden.npz and aabb_golden.npz.

Inputs are rebuilt from the reference's own test constructors and seeds
(reference tests/test_parity.py:12-34 grid/tetrahedron, :91-92 and :100
rng(31) queries, :420-422 rng(77) queries) and from the config-1 workload;
expected values are the per-pair formula of src/capi.mojo:218-235 evaluated
at 50 significant digits with mpmath, rounded once to float64. No reference
code is imported or run.

    python tests/golden/make_golden.py
"""

from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "mojo-libigl_amd", "python"))

import oracle  # noqa: E402
import wn_workloads as W  # noqa: E402


def mp_float(values):
    return np.array([float(v) for v in values], dtype=np.float64)


def main() -> None:
    out = {}
    V, F = W.tetrahedron()
    Q = np.array([[0.1, 0.1, 0.1], [2.0, 2.0, 2.0], [-1.0, 0.2, 0.2]])
    out.update(tet_V=V, tet_F=F, tet_Q=Q, tet_w=mp_float(oracle.winding_mpmath(V, F, Q, 50)))

    V, F = W.grid(5)
    rng = np.random.default_rng(31)
    Qt = rng.normal(size=(7, 3))
    out.update(tail_V=V, tail_F=F[:5], tail_Q=Qt,
               tail_w=mp_float(oracle.winding_mpmath(V, F[:5], Qt, 50)))
    V10, F10 = W.grid(10)
    Qp = rng.normal(size=(1619, 3)) + [4.5, 4.5, 0.0]
    pick = np.arange(0, 1619, 8)  # 203 of the 1619 queries
    out.update(par_V=V10, par_F=F10, par_Q=Qp, par_pick=pick,
               par_w=mp_float(oracle.winding_mpmath(V10, F10, Qp[pick], 50)))

    V1, F1, Q1 = W.config1()
    pick1 = np.arange(0, 1024, 32)  # 32 queries x 5,120 faces
    out.update(c1_pick=pick1, c1_w=mp_float(oracle.winding_mpmath(V1, F1, Q1[pick1], 50)))

    np.savez_compressed(os.path.join(HERE, "winding_golden.npz"), **out)

    # point-mesh squared distance: the reference test's grid(8) and 100
    # uniform points (tests/test_parity.py:12-26, 131-134), results of the oracle's
    # tree descent (checked against brute force before writing)
    Vg, Fg = W.grid(8)
    P = np.random.default_rng(4).uniform([-1.0, -1.0, -1.0], [8.0, 8.0, 2.0], size=(100, 3))
    sqrd, index, closest = oracle.aabb_distance(P, Vg, Fg)
    assert np.array_equal(sqrd, oracle.brute_distance(P, Vg, Fg)[0])
    np.savez_compressed(os.path.join(HERE, "aabb_golden.npz"), V=Vg, F=Fg, P=P, sqrd=sqrd,
                        index=index, closest=closest)
    print("wrote", os.path.join(HERE, "winding_golden.npz"))


if __name__ == "__main__":
    main()
