"""Every GPU path against the committed golden vectors (mpmath at 40+
digits, tests/golden/make_golden.py): the reference tests' own inputs --
the tetrahedron (tests/test_parity.py:29-34), the SIMD-tail soup and the
parallel-threshold grid -- and a subsample of the icosphere-4 workload.
FP64 paths within 1e-12, FP32-compensated paths within 1e-6 (closed meshes
and open soups alike)."""

import os

import numpy as np
import pytest

import igl
import wn_workloads as W

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

GOLDEN = np.load(os.path.join(os.path.dirname(__file__), "golden", "winding_golden.npz"))


def _cases():
    out = {}
    for key in ("tet", "tail", "par"):
        V, F, Q = GOLDEN[f"{key}_V"], GOLDEN[f"{key}_F"], GOLDEN[f"{key}_Q"]
        if key == "par":
            Q = Q[GOLDEN["par_pick"]]
        out[key] = (V, F, Q, GOLDEN[f"{key}_w"], key == "tet")
    V, F, Q = W.config1()
    out["c1"] = (V, F, Q[GOLDEN["c1_pick"]], GOLDEN["c1_w"], True)
    return out


CASES = _cases()


def _paths(V, F, Q):
    from igl.batch import winding_number_batch

    yield "fp64 face", 0, igl.winding_number(V, F, Q, device="gpu")
    yield "fp32c face", 1, igl.winding_number(V, F, Q, device="gpu", precision="fp32c")
    m = igl.Mesh(V, F)
    try:
        yield "fp32c strip", 1, m.winding_number(Q)
        yield "fp64 strip", 0, m.winding_number(Q, precision="fp64")
    finally:
        m.close()
    Fi = np.ascontiguousarray(F, dtype=np.int64)
    out = np.empty(len(Q))
    for mode in (0, 1):
        from igl._lib import addr, lib

        Qc = np.ascontiguousarray(Q, dtype=np.float64)
        Vc = np.ascontiguousarray(V, dtype=np.float64)
        assert lib().wn_winding_gpu_mesh(addr(Vc), addr(Fi), len(Vc), len(Fi), 0, addr(Qc),
                                         addr(out), len(Qc), mode) == 0
        yield f"host strip route mode {mode}", mode, out.copy()
    for prec, mode in (("fp32c", 1), ("fp64", 0)):
        yield f"strip batch {prec}", mode, winding_number_batch([V, V], [F, F], [Q, Q[::-1]],
                                                                 precision=prec, strips=True)[0]


@pytest.mark.parametrize("key", list(CASES))
def test_gpu_paths_match_golden(require_gpu, key):
    V, F, Q, ref, closed = CASES[key]
    for name, mode, w in _paths(V, F, Q):
        err = float(np.abs(np.asarray(w) - ref).max())
        bar = 1e-12 if mode == 0 else 1e-6
        print(f"{key} {name}: max|dw| {err:.2e} (bar {bar:.0e})", flush=True)
        assert err <= bar, f"{key} {name}"
