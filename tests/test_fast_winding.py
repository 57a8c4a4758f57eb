"""Hierarchical ("fast") winding number on the CPU: the oracle's expansion
pinned, then the library against the oracle and against the exact winding
number.

Pins of the second-order expansion (oracle.fwn_expansion / fwn_far_omega):
  * closed-surface identities (divergence theorem) on the root expansion of a
    closed mesh: N = sum A n = 0; tr C = int n . y dA = 3 vol;
    sum_i D_iik = int n_i y_i y_k dA = 4 vol (g - c)_k and
    sum_j D_kjj = int n_k |y|^2 dA = 2 vol (g - c)_k, where g is the solid's
    centroid (computed independently by tetrahedra) and c the expansion
    centre;
  * the truncation error of a single triangle's expansion against its exact
    solid angle falls as distance^-5 (second order: a slip in the first- or
    second-order term leaves a d^-3 or d^-4 error);
  * with beta beyond the mesh extent every face is summed exactly, so the
    result equals the exact winding number (oracle.winding).
Library: expansions exported by wn_fwn_export equal the oracle's; queries
equal oracle.fast_winding; the error against the exact winding number falls
with beta (measured max error on an icosphere: 5.8e-3 at beta 2, 1.7e-4 at
4, 8.6e-6 at 8).
"""

import numpy as np
import pytest

import igl
import oracle
import wn_workloads as W


def _volume_and_centroid(V, F):
    """Signed volume and centroid of a closed outward mesh, from tetrahedra
    (origin, A, B, C)."""
    A, B, C = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    vol = np.einsum("ij,ij->i", A, np.cross(B, C)) / 6.0
    cent = (A + B + C) / 4.0
    return vol.sum(), (vol[:, None] * cent).sum(0) / vol.sum()


@pytest.mark.parametrize("mesh", ["icosphere", "torus"])
def test_expansion_closed_surface_identities(mesh):
    if mesh == "icosphere":
        V, F = W.icosphere(3)
        V = V * [1.0, 1.3, 0.7] + [0.2, -0.1, 0.3]  # ellipsoid, off origin
    else:
        V, F = W.torus(nu=40, nv=30)
    c, _, N, C, D = oracle.fwn_expansion(V[F])
    vol, g = _volume_and_centroid(V, F)
    scale = np.abs(V).max()
    assert np.abs(N).max() <= 1e-12 * scale ** 2
    assert np.trace(C) == pytest.approx(3.0 * vol, rel=1e-12)
    np.testing.assert_allclose(np.einsum("iik->k", D), 4.0 * vol * (g - c), atol=1e-12 * scale ** 4)
    np.testing.assert_allclose(np.einsum("kjj->k", D), 2.0 * vol * (g - c), atol=1e-12 * scale ** 4)


def test_expansion_error_falls_as_d_to_the_minus_five():
    T = np.array([[[0.0, 0.0, 0.0], [1.0, 0.1, 0.2], [0.3, 0.9, -0.1]]])
    e = oracle.fwn_expansion(T)
    direction = np.array([0.3, -0.5, 0.8]) / np.linalg.norm([0.3, -0.5, 0.8])
    errs = []
    for d in (8.0, 16.0, 32.0):
        q = e[0] + d * direction
        errs.append(abs(oracle.fwn_far_omega(e, q) - oracle.exact_omega(T, q)))
    for a, b in zip(errs, errs[1:]):
        assert 24.0 < a / b < 40.0  # 2^5 = 32


def test_library_expansions_equal_oracle():
    V, F = W.icosphere(2)
    fw = igl.FastWindingNumber(V, F)
    e, links = fw.expansions()
    boxes, nodes, order, _ = oracle.build_aabb(V, F)
    np.testing.assert_array_equal(links, nodes)
    corners = V[F[order]]

    def span(n):
        l, r, s, k = nodes[n]
        return (s, s + k) if k > 0 else (span(l)[0], span(r)[1])

    idx = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    for n in range(len(nodes)):
        c, rad, N, C, D = oracle.fwn_expansion(corners[slice(*span(n))])
        np.testing.assert_allclose(e[n, :3], c, rtol=0, atol=1e-14)
        assert e[n, 3] == pytest.approx(rad, rel=1e-13)
        np.testing.assert_allclose(e[n, 4:7], N, rtol=0, atol=1e-14)
        np.testing.assert_allclose(e[n, 7:16].reshape(3, 3), C, rtol=0, atol=1e-14)
        Dp = e[n, 16:].reshape(3, 6)
        for i in range(3):
            for p, (j, k) in enumerate(idx):
                assert Dp[i, p] == pytest.approx(D[i, j, k], abs=1e-14)
    fw.close()


@pytest.mark.parametrize("beta", [1.0, 2.0, 5.0])
def test_library_equals_oracle_descent(beta):
    V, F = W.icosphere(2)
    Q = np.random.default_rng(3).uniform(-1.6, 1.6, size=(150, 3))
    np.testing.assert_allclose(igl.fast_winding_number(V, F, Q, beta=beta),
                               oracle.fast_winding(Q, V, F, beta), rtol=0, atol=1e-13)


@pytest.mark.parametrize("mesh", ["icosphere", "torus", "open_cylinder"])
def test_error_against_exact_falls_with_beta(mesh):
    if mesh == "icosphere":
        V, F = W.icosphere(4)
    elif mesh == "torus":
        V, F = W.torus(nu=60, nv=40)
    else:
        V, F = W.open_cylinder(48, 20, 2, np.random.default_rng(1))
    lo, hi = V.min(0), V.max(0)
    Q = np.random.default_rng(4).uniform(lo - 0.3 * (hi - lo), hi + 0.3 * (hi - lo), size=(1500, 3))
    exact = oracle.winding(V, F, Q)
    errs = {b: np.abs(igl.fast_winding_number(V, F, Q, beta=b) - exact).max()
            for b in (2.0, 4.0, 8.0, 1e9)}
    print(mesh, {k: f"{v:.1e}" for k, v in errs.items()})
    assert errs[2.0] <= 2e-2
    assert errs[4.0] <= errs[2.0] / 8 and errs[8.0] <= errs[4.0] / 8
    assert errs[1e9] <= 1e-12  # every face exact


def test_far_queries_and_contracts():
    V, F = W.icosphere(3)
    d = np.random.default_rng(5).normal(size=(50, 3))
    Q = 10.0 * d / np.linalg.norm(d, axis=1)[:, None]
    assert np.abs(igl.fast_winding_number(V, F, Q)).max() <= 1e-12
    fw = igl.FastWindingNumber(V, F)
    with pytest.raises(ValueError):
        fw.query(Q, beta=0.0)
    with pytest.raises(ValueError):
        fw.query(Q[:, :2])
    with pytest.raises(ValueError):
        fw.query([[np.nan, 0.0, 0.0]])
    with pytest.raises(ValueError):
        fw.query(Q, device="tpu")
    assert fw.query(np.zeros((0, 3))).shape == (0,)
    info = fw.info()
    assert info["faces"] == len(F) and info["nodes"] > 1 and info["device_bytes"] == 0
    fw.close()
    with pytest.raises(IndexError):
        igl.FastWindingNumber(V, np.array([[0, 1, len(V)]]))
    empty = igl.fast_winding_number(V, np.zeros((0, 3), dtype=np.int64), Q)
    np.testing.assert_array_equal(empty, np.zeros(len(Q)))


def test_serial_and_parallel_spans_identical(monkeypatch):
    import igl._lib as L

    V, F = W.icosphere(3)
    Q = np.random.default_rng(6).uniform(-1.5, 1.5, size=(3000, 3))
    fw = igl.FastWindingNumber(V, F)
    monkeypatch.setitem(L.PARALLEL_MIN_WORK, "winding", 1 << 62)
    serial = fw.query(Q)
    monkeypatch.setitem(L.PARALLEL_MIN_WORK, "winding", 0)
    parallel = fw.query(Q)
    assert np.array_equal(serial, parallel)
    fw.close()


def test_convenience_call_keeps_its_tree():
    """fast_winding_number keeps the last tree (per thread) for byte-identical
    V and F: repeated calls reuse it, an in-place change of V or another F
    builds a new one, and every result equals a freshly built tree's."""
    import igl.fast_winding as FW

    V, F = W.icosphere(3)
    V = np.ascontiguousarray(V, dtype=np.float64)
    Q = np.random.default_rng(3).uniform(-1.5, 1.5, size=(500, 3))

    def fresh(Vm, Fm):
        t = igl.FastWindingNumber(Vm, Fm)
        try:
            return t.query(Q)
        finally:
            t.close()

    FW.release()
    a = igl.fast_winding_number(V, F, Q)
    tree = FW._kept.tree
    b = igl.fast_winding_number(V, F, Q)
    assert FW._kept.tree is tree
    assert np.array_equal(a, b) and np.array_equal(a, fresh(V, F))
    V *= 1.1  # in place: the kept tree holds its own copy and must not be reused
    c = igl.fast_winding_number(V, F, Q)
    assert FW._kept.tree is not tree
    assert np.array_equal(c, fresh(V, F))
    d = igl.fast_winding_number(V, F[::-1].copy(), Q)
    assert np.array_equal(d, fresh(V, F[::-1].copy()))
    FW.release()
    assert FW._kept.tree is None
