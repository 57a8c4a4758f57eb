"""Boundary edges and the edge form of the adjoint (CPU).

* wn_boundary_edges (C-ABI, host) against the counting oracle, and both
  against hand counts: closed meshes have none, an open grid's boundary is
  its rim, a flipped face in a closed mesh leaves its three edges with
  weight 2, repeated vertices drop out;
* oracle.edge_adjoint (the sum over boundary edges) against the autograd
  gradient of oracle.winding_torch (the per-face formula), FP64 on the CPU:
  open and closed meshes, relative error <= 1e-10.
"""

import numpy as np
import pytest

import oracle
import wn_workloads as W

torch = pytest.importorskip("torch")


def _capi_edges(F, nv):
    from igl._lib import addr, lib

    F = np.ascontiguousarray(F, dtype=np.int64)
    h = lib()
    n = h.wn_boundary_edges(addr(F), nv, len(F), 0, 0, 0, 0)
    assert n >= 0
    e = np.zeros((n, 2), np.int32)
    m = np.zeros(n)
    if n:
        assert h.wn_boundary_edges(addr(F), nv, len(F), 0, e.ctypes.data, m.ctypes.data, n) == n
    return {(int(u), int(v)): float(w) for (u, v), w in zip(e, m)}


def _meshes():
    cyl_V, cyl_F = W.open_cylinder(24, 8, 2, np.random.default_rng(9))
    tV, tF = W.tetrahedron()
    flipped = tF.copy()
    flipped[0] = flipped[0][::-1]
    return {
        "torus": W.torus(nu=30, nv=20),
        "ico2": W.icosphere(2),
        "grid": W.grid(6),
        "cylinder": (cyl_V, cyl_F),
        "tet_flipped": (tV, flipped),
        "degenerate": (np.eye(3), np.array([[0, 1, 2], [0, 0, 1], [2, 2, 2]])),
    }


@pytest.mark.parametrize("name", list(_meshes()))
def test_capi_edges_match_counting_oracle(name):
    V, F = _meshes()[name]
    assert _capi_edges(F, len(V)) == {k: float(v) for k, v in oracle.boundary_edges(F).items()}


def test_edge_hand_counts():
    for name in ("torus", "ico2"):
        V, F = _meshes()[name]
        assert oracle.boundary_edges(F) == {}
    # grid(n): n x n vertices, rim of 4 (n - 1) edges, each used once
    V, F = W.grid(6)
    E = oracle.boundary_edges(F)
    assert len(E) == 4 * 5 and set(E.values()) == {1}
    # a flipped face of a closed tetrahedron: its three edges, weight 2,
    # directed as the flipped face runs them
    V, F = _meshes()["tet_flipped"]
    E = oracle.boundary_edges(F)
    f = F[0].tolist()
    assert E == {(f[0], f[1]): 2, (f[1], f[2]): 2, (f[2], f[0]): 2}
    # repeated vertices: only the first face's edges remain
    assert oracle.boundary_edges(np.array([[0, 1, 2], [0, 0, 1], [2, 2, 2]])) == {
        (0, 1): 1, (1, 2): 1, (2, 0): 1}


def test_capi_rejects_bad_index():
    from igl._lib import addr, lib

    F = np.array([[0, 1, 5]], dtype=np.int64)
    assert lib().wn_boundary_edges(addr(F), 3, 1, 0, 0, 0, 0) < 0


@pytest.mark.parametrize("name", list(_meshes()))
def test_edge_adjoint_matches_face_autograd(name):
    V, F = _meshes()[name]
    V = np.asarray(V, np.float64)
    rng = np.random.default_rng(5)
    lo, hi = V.min(0), V.max(0)
    pad = 0.3 * (hi - lo) + 0.1
    Q = rng.uniform(lo - pad, hi + pad, size=(64, 3))
    g = rng.standard_normal(len(Q))
    gV, gO = oracle.edge_adjoint(V, F, Q, g)
    Vt = torch.tensor(V, requires_grad=True)
    Qt = torch.tensor(Q, requires_grad=True)
    (oracle.winding_torch(Vt, torch.tensor(np.asarray(F)), Qt) * torch.tensor(g)).sum().backward()
    rV, rO = Vt.grad.numpy(), Qt.grad.numpy()
    # scale: the per-face terms' size (closed meshes cancel to ~0)
    scale = max(np.abs(rV).max(), np.abs(rO).max(), 1e-3)
    assert np.abs(gV - rV).max() <= 1e-10 * scale
    assert np.abs(gO - rO).max() <= 1e-10 * scale
    if name in ("torus", "ico2"):
        assert not gV.any() and not gO.any()
