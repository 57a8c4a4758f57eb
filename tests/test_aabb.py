"""Point-mesh squared distance on the CPU: the oracle pinned, then the
library's tree builder and CPU path against it.

Oracle pins (none of them retypes the oracle's own formula):
  * closed forms: points over a triangle's interior, in each vertex and edge
    region, on the triangle, and the reference's degenerate-triangle test
    (tests/test_parity.py:180-187) worked by hand;
  * an independent formulation: distance to a triangle as the minimum of the
    in-plane projection (when its barycentrics are all >= 0) and the three
    segment distances, vectorised in NumPy, on random triangles and points;
  * the tree descent against brute force over all faces (same minimum; the
    face index differs only between faces at exactly that distance).
Library checks: wn_aabb_build equals the NumPy restatement of the reference
builder array for array; the CPU path (mli_aabb_squared_distance through
igl.AABB) equals the oracle bit for bit, on the reference tests' inputs
(tests/test_parity.py:131-177, 425-445) and contracts (:270-290, 340-350).
"""

import os

import numpy as np
import pytest

import igl
import oracle
import wn_workloads as W


# The reference tests' own grid (tests/test_parity.py:12-26): z = 0.15 sin(i) cos(j).
grid = W.grid


# --- oracle pins ------------------------------------------------------------

TRI = np.array([[0.0, 0.0, 0.0], [4.0, 0.0, 0.0], [0.0, 3.0, 0.0]])


@pytest.mark.parametrize("p,d2,q", [
    ((1.0, 1.0, 2.0), 4.0, (1.0, 1.0, 0.0)),        # over the interior
    ((1.0, 1.0, 0.0), 0.0, (1.0, 1.0, 0.0)),        # on the face
    ((-1.0, -2.0, 0.0), 5.0, (0.0, 0.0, 0.0)),      # vertex region A
    ((6.0, -1.0, 1.0), 6.0, (4.0, 0.0, 0.0)),       # vertex region B
    ((-1.0, 5.0, 0.0), 5.0, (0.0, 3.0, 0.0)),       # vertex region C
    ((2.0, -3.0, 4.0), 25.0, (2.0, 0.0, 0.0)),      # edge AB
    ((-2.0, 1.5, 0.0), 4.0, (0.0, 1.5, 0.0)),       # edge AC
    # edge BC: the line 3x + 4y = 12; (4, 3) is 12/5 from it along (3, 4)/5
    ((4.0, 3.0, 0.0), 144.0 / 25.0, (4.0 - 36.0 / 25.0, 3.0 - 48.0 / 25.0, 0.0)),
])
def test_oracle_closest_on_triangle_closed_forms(p, d2, q):
    got, point = oracle.closest_on_triangle(p, TRI)
    assert got == pytest.approx(d2, rel=1e-15, abs=1e-15)
    np.testing.assert_allclose(point, q, rtol=0, atol=1e-15)


def test_oracle_degenerate_triangles_by_hand():
    """tests/test_parity.py:180-187: face 0 is the segment (0,0,0)-(2,0,0)
    (two equal corners), face 1 the point (0,1,0)."""
    V = np.array([[0.0, 0, 0], [2.0, 0, 0], [0.0, 1, 0]])
    F = np.array([[0, 1, 1], [2, 2, 2]], dtype=np.int64)
    P = np.array([[0.5, 1.0, 0.0], [0.0, 2.0, 0.0], [1.5, -2.0, 0.0]])
    sqrd, idx, cl = oracle.brute_distance(P, V, F)
    np.testing.assert_array_equal(sqrd, [0.25, 1.0, 4.0])
    np.testing.assert_array_equal(idx, [1, 1, 0])
    np.testing.assert_array_equal(cl, [[0.0, 1.0, 0.0], [0.0, 1.0, 0.0], [1.5, 0.0, 0.0]])
    # each degenerate face alone
    d0, q0 = oracle.closest_on_triangle([0.5, 1.0, 0.0], V[F[0]])
    assert d0 == 1.0 and tuple(q0) == (0.5, 0.0, 0.0)
    d1, q1 = oracle.closest_on_triangle([0.0, 2.0, 0.0], V[F[1]])
    assert d1 == 1.0 and tuple(q1) == (0.0, 1.0, 0.0)


def _segment_d2(p, a, b):
    ab = b - a
    den = (ab * ab).sum(-1)
    t = np.where(den > 0, ((p - a) * ab).sum(-1) / np.where(den > 0, den, 1.0), 0.0)
    t = np.clip(t, 0.0, 1.0)[..., None]
    d = p - (a + t * ab)
    return (d * d).sum(-1)


def _independent_d2(p, a, b, c):
    """Distance^2 to a triangle as min(projection if inside, three segments)."""
    n = np.cross(b - a, c - a)
    nn = (n * n).sum(-1)
    s = ((p - a) * n).sum(-1) / nn
    proj = p - s[..., None] * n
    # barycentrics of the projection
    def area(u, v, w):
        return (np.cross(v - u, w - u) * n).sum(-1) / nn
    la, lb, lc = area(proj, b, c), area(a, proj, c), area(a, b, proj)
    inside = (la >= 0) & (lb >= 0) & (lc >= 0)
    plane = s * s * nn
    edges = np.minimum(np.minimum(_segment_d2(p, a, b), _segment_d2(p, b, c)), _segment_d2(p, c, a))
    return np.where(inside, np.minimum(plane, edges), edges)


def test_oracle_matches_independent_formulation():
    rng = np.random.default_rng(11)
    worst = 0.0
    for _ in range(2000):
        t = rng.normal(size=(3, 3))
        p = rng.normal(size=3) * 2.0
        got, q = oracle.closest_on_triangle(p, t)
        ref = float(_independent_d2(p, t[0], t[1], t[2]))
        worst = max(worst, abs(got - ref) / max(ref, 1e-300))
        assert got == pytest.approx(ref, rel=1e-9, abs=1e-12)
        # the returned point realises the distance and lies on the triangle
        assert ((p - q) ** 2).sum() == pytest.approx(got, rel=1e-12, abs=1e-14)
        assert _independent_d2(q, t[0], t[1], t[2]) <= 1e-20 + 1e-12 * (t ** 2).sum()
    print(f"closest_on_triangle vs independent formulation: worst rel {worst:.2e}")


@pytest.mark.parametrize("n,seed", [(8, 4), (20, 5)])
def test_oracle_descent_equals_brute_force(n, seed):
    V, F = grid(n)
    P = np.random.default_rng(seed).uniform([-1, -1, -1], [n, n, 2], size=(300, 3))
    a = oracle.aabb_distance(P, V, F)
    b = oracle.brute_distance(P, V, F)
    np.testing.assert_array_equal(a[0], b[0])
    for i in np.nonzero(a[1] != b[1])[0]:  # exact ties only
        d, _ = oracle.closest_on_triangle(P[i], V[F[a[1][i]]])
        assert d == b[0][i]


# --- library: tree builder and CPU path --------------------------------------

def _soup(seed, nf=500):
    rng = np.random.default_rng(seed)
    V = rng.normal(size=(300, 3))
    F = rng.integers(0, 300, size=(nf, 3)).astype(np.int64)
    F[::17, 2] = F[::17, 1]  # some degenerate faces
    return V, F


@pytest.mark.parametrize("case", ["grid8", "grid60", "soup", "soup_int32", "one_face",
                                  "large_ties"])
def test_tree_builder_equals_reference_builder(case):
    """Same boxes, nodes, order and stack depth as the reference's builder;
    "large_ties" (140k faces on integer coordinates: many equal centres) runs
    the multi-threaded sorts of the top levels."""
    if case == "large_ties":
        rng = np.random.default_rng(9)
        V = np.round(rng.uniform(-20, 20, (4000, 3)))
        F = rng.integers(0, 4000, (140000, 3)).astype(np.int64)
    elif case == "grid8":
        V, F = grid(8)
    elif case == "grid60":
        V, F = grid(60)
    elif case.startswith("soup"):
        V, F = _soup(3)
    else:
        V, F = grid(2)
        F = F[:1]
    from igl._lib import lib

    Fi = np.ascontiguousarray(F.astype(np.int32) if case == "soup_int32" else F)
    fdtype = 1 if case == "soup_int32" else 0
    n = lib().wn_aabb_build(V.ctypes.data, Fi.ctypes.data, len(V), len(F), fdtype, 8, 0, 0, 0, 0, 0)
    boxes, nodes = np.empty((n, 6)), np.empty((n, 4), dtype=np.int64)
    order, stack = np.empty(len(F), dtype=np.int64), np.zeros(1, dtype=np.int64)
    assert lib().wn_aabb_build(V.ctypes.data, Fi.ctypes.data, len(V), len(F), fdtype, 8,
                               boxes.ctypes.data, nodes.ctypes.data, order.ctypes.data, n,
                               stack.ctypes.data) == n
    ours = (boxes, nodes, order, int(stack[0]))
    ref = oracle.build_aabb(V, F)
    for a, b in zip(ours[:3], ref[:3]):
        np.testing.assert_array_equal(a, b)
    assert ours[3] == ref[3]


@pytest.mark.parametrize("n,seed,kind", [(8, 4, "uniform"), (9, 12, "normal"), (10, 19, "uniform")])
def test_cpu_path_equals_oracle_on_reference_inputs(n, seed, kind):
    """tests/test_parity.py:131-177: grid(8) + 100 uniform points (seed 4),
    grid(9) + 80 normal points (seed 12), grid(10) + 257 uniform (seed 19)."""
    V, F = grid(n)
    rng = np.random.default_rng(seed)
    if kind == "normal":
        P = rng.normal(size=(80, 3)) * [3.0, 3.0, 1.0] + [4.0, 4.0, 0.0]
    else:
        P = rng.uniform([-1.0, -1.0, -1.0], [n, n, 2.0], size=(100 if n == 8 else 257, 3))
    ours = igl.point_mesh_squared_distance(P, V, F)
    ref = oracle.aabb_distance(P, V, F)
    for a, b in zip(ours, ref):
        np.testing.assert_array_equal(a, b)
    assert ours[1].dtype == np.int64 and ours[2].shape == (len(P), 3)
    # the chosen face realises the distance
    for p, d, f in zip(P, ours[0], ours[1]):
        got, _ = oracle.closest_on_triangle(p, V[F[f]])
        assert got == d


def test_cpu_path_degenerate_and_soup():
    V = np.array([[0.0, 0, 0], [2.0, 0, 0], [0.0, 1, 0]])
    F = np.array([[0, 1, 1], [2, 2, 2]], dtype=np.int64)
    P = np.array([[0.5, 1.0, 0.0], [0.0, 2.0, 0.0]])
    ours = igl.point_mesh_squared_distance(P, V, F)
    np.testing.assert_array_equal(ours[0], [0.25, 1.0])
    V, F = _soup(8)
    P = np.random.default_rng(9).normal(size=(400, 3)) * 1.5
    ours = igl.point_mesh_squared_distance(P, V, F)
    ref = oracle.aabb_distance(P, V, F)
    for a, b in zip(ours, ref):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(ours[0], oracle.brute_distance(P, V, F)[0])


def test_serial_and_parallel_spans_identical(monkeypatch):
    """tests/test_parity.py:431-445: grid(60), 2048 points (seed 78)."""
    import igl._lib as L

    V, F = grid(60)
    P = np.random.default_rng(78).uniform([-5.0, -5.0, -5.0], [65.0, 65.0, 5.0], size=(2048, 3))
    tree = igl.AABB()
    tree.init(V, F)
    monkeypatch.setitem(L.PARALLEL_MIN_WORK, "aabb", 1 << 62)
    serial = tree.squared_distance(V, F, P)
    monkeypatch.setitem(L.PARALLEL_MIN_WORK, "aabb", 0)
    for workers in (2, 4):
        monkeypatch.setattr(L, "MAX_WORKERS", workers)
        parallel = tree.squared_distance(V, F, P)
        for a, b in zip(parallel, serial):
            assert np.array_equal(a, b)
    assert set(serial[0]) != {np.inf}


def test_contracts():
    """tests/test_parity.py:270-290, 340-350."""
    V, F = grid(3)
    with pytest.raises(RuntimeError):
        igl.AABB().squared_distance(V, F, V[:1])
    with pytest.raises(ValueError):
        igl.point_mesh_squared_distance([[np.inf, 0.0, 0.0]], V, F)
    with pytest.raises(ValueError):
        igl.point_mesh_squared_distance([[0.0, 0.0]], V, F)
    with pytest.raises(IndexError):
        igl.point_mesh_squared_distance(V[:1], V, np.array([[0, 1, len(V)]]))
    with pytest.raises(TypeError):
        igl.point_mesh_squared_distance(V[:1], V, F.astype(np.float64))
    tree = igl.AABB()
    tree.init(V, F)
    changed = V.copy()
    changed[0, 0] += 1.0
    with pytest.raises(ValueError, match="AABB.init"):
        tree.squared_distance(changed, F, V[:1])
    with pytest.raises(ValueError):
        tree.squared_distance(V, F, V[:1], device="tpu")
    with pytest.raises(NotImplementedError):
        tree.find(V, F, V[:1])
    # no faces: +inf, -1, origin; no points: empty results
    s, i, c = igl.point_mesh_squared_distance(V[:2], V, np.zeros((0, 3), dtype=np.int64))
    assert np.isinf(s).all() and (i == -1).all() and (c == 0).all()
    s, i, c = tree.squared_distance(V, F, np.zeros((0, 3)))
    assert s.shape == (0,) and i.shape == (0,) and c.shape == (0, 3)


def test_golden_fixture():
    """tests/golden/aabb_golden.npz: grid(8) + 100 points (seed 4) with the
    oracle's results, written by tests/golden/make_golden.py."""
    import os

    path = os.path.join(os.path.dirname(__file__), "golden", "aabb_golden.npz")
    g = np.load(path)
    ours = igl.point_mesh_squared_distance(g["P"], g["V"], g["F"])
    np.testing.assert_array_equal(ours[0], g["sqrd"])
    np.testing.assert_array_equal(ours[1], g["index"])
    np.testing.assert_array_equal(ours[2], g["closest"])
    # the fixture itself agrees with the independent formulation
    V, F = g["V"], g["F"]
    ind = np.array([_independent_d2(p, *V[F[f]]) for p, f in zip(g["P"], g["index"])])
    np.testing.assert_allclose(g["sqrd"], ind, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("winding", ["exact", "fast"])
def test_signed_distance_closed_mesh(winding):
    """Sign from the winding number (inside negative), magnitude from the
    distance query; against the oracle's brute-force distance and exact
    winding number on a closed icosphere."""
    V, F = W.icosphere(3)
    rng = np.random.default_rng(21)
    d = rng.normal(size=(600, 3))
    P = d / np.linalg.norm(d, axis=1)[:, None] * rng.uniform(0.2, 1.8, size=(600, 1))
    S, I, C = igl.signed_distance(P, V, F, winding=winding, beta=8.0)
    sq, _, _ = oracle.brute_distance(P, V, F)
    w = oracle.winding(V, F, P)
    np.testing.assert_array_equal(np.abs(S), np.sqrt(sq))
    np.testing.assert_array_equal(S < 0, w >= 0.5)
    # radius 1 icosphere: points well inside or outside are signed by radius
    r = np.linalg.norm(P, axis=1)
    assert (S[r < 0.8] < 0).all() and (S[r > 1.2] > 0).all()
    with pytest.raises(ValueError):
        igl.signed_distance(P, V, F, winding="other")


def test_convenience_call_keeps_its_tree():
    """point_mesh_squared_distance keeps the last tree (per thread) for
    byte-identical V and Ele: repeated calls reuse it, an in-place change of
    V or another Ele builds a new one, and every result equals a freshly
    built tree's bit for bit."""
    import igl.aabb as AB

    V, F = grid(20)
    V = np.ascontiguousarray(V, dtype=np.float64)
    P = np.random.default_rng(8).uniform(V.min(0) - 1, V.max(0) + 1, size=(700, 3))

    def fresh(Vm, Fm):
        t = igl.AABB()
        t.init(Vm, Fm)
        return t.squared_distance(Vm, Fm, P)

    def same(a, b):
        return all(np.array_equal(x, y) for x, y in zip(a, b))

    AB.release()
    a = igl.point_mesh_squared_distance(P, V, F)
    tree = AB._kept.tree
    b = igl.point_mesh_squared_distance(P, V, F)
    assert AB._kept.tree is tree and same(a, b) and same(a, fresh(V, F))
    V[:, 2] += 0.25  # in place: the kept tree holds its own copy
    c = igl.point_mesh_squared_distance(P, V, F)
    assert AB._kept.tree is not tree and same(c, fresh(V, F))
    F2 = F[::-1].copy()
    assert same(igl.point_mesh_squared_distance(P, V, F2), fresh(V, F2))
    AB.release()
    assert AB._kept.tree is None


def _tree_golden():
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                             "aabb_tree_golden.npz"))
    names = sorted({k.split("_")[0] for k in g.files})
    return g, names


@pytest.mark.parametrize("leaf", [8, 3])
def test_tree_layout_matches_reference_builder(leaf):
    """The reference's own tree (its NumPy _build_aabb,
    python/igl/__init__.py:286-319, run once by
    tests/golden/make_aabb_tree_golden.py on the meshes of its AABB tests):
    boxes, preorder node links, face order (stable ties) and stack size equal
    bit for bit to the oracle's restatement and to the library builder
    (wn_aabb_build), which mli_aabb_squared_distance's callers receive."""
    from igl.aabb import _build_tree

    g, names = _tree_golden()
    assert {"grid8", "grid9", "grid10", "grid60", "flat12", "soup"} <= set(names)
    for name in names:
        V, F = g[f"{name}_V"], g[f"{name}_F"]
        key = f"{name}_leaf{leaf}"
        expect = (g[f"{key}_boxes"], g[f"{key}_nodes"], g[f"{key}_order"], int(g[f"{key}_stack"]))
        for got in (oracle.build_aabb(V, F, leaf), _build_tree(V, F, leaf)):
            assert np.array_equal(got[0], expect[0]), name
            assert np.array_equal(got[1], expect[1]), name
            assert np.array_equal(got[2], expect[2]), name
            assert got[3] == expect[3], name
