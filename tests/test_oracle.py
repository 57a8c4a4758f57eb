"""Pins the FP64 oracle (oracle/wn_oracle.c, restating src/capi.mojo:205-294)
before anything is checked against it.

Pins, chosen so that a dropped term, a wrong sign or index or a transposed
operand in the restatement fails at least one of them:
  * closed forms: the reference's tetrahedron (tests/test_parity.py:74-85),
    the octant triangle (w = 1/8), a rectangle over its centre
    (Omega = 4 atan(ab / (2h sqrt(4h^2 + a^2 + b^2)))), a polyhedral
    hemisphere seen from its centre (w = 1/2), closed meshes (integers);
  * 50-digit mpmath values on the reference tests' own inputs
    (tests/golden/winding_golden.npz, made by tests/golden/make_golden.py);
  * invariants: rigid motion and scale, cyclic corner order, face order,
    orientation reversal (w -> -w), additivity over face subsets;
  * an independent NumPy restatement and the reference's own GPU arctangent.
"""

import math
import os

import numpy as np
import pytest
from hypothesis import given, settings
from hypothesis import strategies as st

import oracle
import wn_workloads as W

GOLDEN = np.load(os.path.join(os.path.dirname(__file__), "golden", "winding_golden.npz"))


def test_tetrahedron_closed_form():
    V, F = W.tetrahedron()
    Q = np.array([[0.1, 0.1, 0.1], [2.0, 2.0, 2.0], [-1.0, 0.2, 0.2]])
    w = oracle.winding(V, F, Q)
    assert w[0] == pytest.approx(1.0, abs=1e-14)
    assert abs(w[1]) <= 1e-14 and abs(w[2]) <= 1e-14


def test_octant_triangle_is_one_eighth():
    V = np.eye(3)
    F = np.array([[0, 1, 2]])
    assert oracle.winding(V, F, [[0.0, 0.0, 0.0]])[0] == pytest.approx(0.125, abs=1e-15)
    assert oracle.winding(V, F[:, ::-1], [[0.0, 0.0, 0.0]])[0] == pytest.approx(-0.125, abs=1e-15)


@pytest.mark.parametrize("a,b,h", [(1.0, 1.0, 0.5), (2.0, 0.3, 0.1), (0.2, 5.0, 3.0), (1.0, 1.0, 1e-6)])
def test_rectangle_over_centre(a, b, h):
    V = np.array([[-a / 2, -b / 2, 0.0], [a / 2, -b / 2, 0.0], [a / 2, b / 2, 0.0], [-a / 2, b / 2, 0.0]])
    F = np.array([[0, 1, 2], [0, 2, 3]])
    omega = 4.0 * math.atan(a * b / (2.0 * h * math.sqrt(4 * h * h + a * a + b * b)))
    w = oracle.winding(V, F, [[0.0, 0.0, -h]])[0]
    # The query sits above the shared diagonal, at distance h from an edge
    # line, where the formula's FP64 conditioning is ~eps * size / h.
    assert w == pytest.approx(omega / (4 * math.pi), abs=1e-14 + 1e-15 * (a + b) / h)


def test_hemisphere_from_centre_is_one_half():
    # UV hemisphere: rings at polar angle pi/2 (equator) ... pole; vertices on
    # the unit sphere, so each flat face subtends its geodesic triangle.
    n_az, n_ring = 24, 8
    V = [[0.0, 0.0, 1.0]]
    for r in range(1, n_ring + 1):
        th = 0.5 * math.pi * r / n_ring
        for k in range(n_az):
            ph = 2 * math.pi * k / n_az
            V.append([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
    V = np.array(V)
    F = []
    for k in range(n_az):
        F.append([0, 1 + k, 1 + (k + 1) % n_az])
    for r in range(1, n_ring):
        base0, base1 = 1 + (r - 1) * n_az, 1 + r * n_az
        for k in range(n_az):
            a, b = base0 + k, base0 + (k + 1) % n_az
            c, d = base1 + k, base1 + (k + 1) % n_az
            F.append([a, c, d])
            F.append([a, d, b])
    F = np.array(F)
    w = oracle.winding(V, F, [[0.0, 0.0, 0.0]])[0]
    assert abs(w) == pytest.approx(0.5, abs=1e-13)


@pytest.mark.parametrize("mesh", ["ico3", "torus"])
def test_closed_meshes_give_integers(mesh):
    V, F = W.icosphere(3) if mesh == "ico3" else W.torus(nu=40, nv=30)
    Q = np.random.default_rng(3).uniform(-1.6, 1.6, size=(200, 3))
    w = oracle.winding(V, F, Q)
    assert np.abs(w - np.round(w)).max() <= 1e-12
    assert set(np.round(w).astype(int)) <= {0, 1}


@pytest.mark.parametrize("key", ["tet", "tail", "par"])
def test_golden_mpmath_reference_inputs(key):
    V, F, Q = GOLDEN[f"{key}_V"], GOLDEN[f"{key}_F"], GOLDEN[f"{key}_Q"]
    if key == "par":
        Q = Q[GOLDEN["par_pick"]]
    w = oracle.winding(V, F, Q)
    assert np.abs(w - GOLDEN[f"{key}_w"]).max() <= 1e-13


def test_golden_mpmath_config1():
    V, F, Q = W.config1()
    pick = GOLDEN["c1_pick"]
    w = oracle.winding(V, F, Q[pick])
    assert np.abs(w - GOLDEN["c1_w"]).max() <= 1e-12


def test_numpy_restatement_agrees():
    V, F, Q = W.config1()
    w = oracle.winding(V, F, Q[:64])
    assert np.abs(w - oracle.winding_numpy(V, F, Q[:64])).max() <= 1e-13


def _rotation(angles):
    a, b, c = angles
    Rx = np.array([[1, 0, 0], [0, math.cos(a), -math.sin(a)], [0, math.sin(a), math.cos(a)]])
    Ry = np.array([[math.cos(b), 0, math.sin(b)], [0, 1, 0], [-math.sin(b), 0, math.cos(b)]])
    Rz = np.array([[math.cos(c), -math.sin(c), 0], [math.sin(c), math.cos(c), 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


coords = st.floats(-2.0, 2.0, allow_nan=False, allow_infinity=False)


@settings(max_examples=40, deadline=None)
@given(st.lists(st.tuples(coords, coords, coords), min_size=3, max_size=3),
       st.tuples(st.floats(-3, 3), st.floats(-3, 3), st.floats(-3, 3)),
       st.floats(0.25, 4.0), st.tuples(coords, coords, coords))
def test_invariance_rigid_motion_and_scale(pts, angles, scale, shift):
    V, F = W.grid(4)
    Q = np.array(pts) + [1.5, 1.5, 0.0]
    w = oracle.winding(V, F, Q)
    R = _rotation(angles)
    t = np.array(shift)
    w2 = oracle.winding((V @ R.T) * scale + t, F, (Q @ R.T) * scale + t)
    # exclude queries that sit on the surface (w is discontinuous there)
    assert np.abs(w - w2).max() <= 1e-9 or np.min(np.abs(Q[:, 2])) < 1e-3


def test_corner_and_face_order_orientation_additivity():
    V, F, Q = W.config1()
    Q = Q[:64]
    tri = oracle.pack_soa(V, F)
    w = oracle.winding_soa(tri, Q)
    rng = np.random.default_rng(7)
    # cyclic corner permutation: same value
    assert np.abs(oracle.winding(V, F[:, [1, 2, 0]], Q) - w).max() <= 1e-12
    # reversed orientation (corner swap): negated
    assert np.abs(oracle.winding(V, F[:, [0, 2, 1]], Q) + w).max() <= 1e-12
    # face order: same value up to rounding of the sum
    assert np.abs(oracle.winding(V, F[rng.permutation(len(F))], Q) - w).max() <= 1e-12
    # additivity over a split of the faces
    cut = len(F) // 3
    parts = oracle.winding(V, F[:cut], Q) + oracle.winding(V, F[cut:], Q)
    assert np.abs(parts - w).max() <= 1e-12


def test_thread_fan_out_is_bitwise_serial():
    V, F = W.grid(20)
    Q = np.random.default_rng(77).uniform([0.0, 0.0, -2.0], [19.0, 19.0, 2.0], size=(512, 3))
    tri = oracle.pack_soa(V, F)
    serial = oracle.winding_threads(tri, Q, 1)
    for t in (2, 3, 8):
        assert np.array_equal(oracle.winding_threads(tri, Q, t), serial)


def test_reference_gpu_arctangent():
    """src/capi.mojo:326-352 agrees with atan2 away from the signed-zero case
    and gives +pi for det = -0.0, den < 0, where atan2 gives -pi."""
    rng = np.random.default_rng(11)
    for det, den in rng.normal(size=(2000, 2)) * 10.0 ** rng.uniform(-3, 3, size=(2000, 2)):
        assert oracle.ref_gpu_atan2(det, den) == pytest.approx(math.atan2(det, den), abs=2e-15)
    assert oracle.ref_gpu_atan2(-0.0, -1.0) == pytest.approx(math.pi)
    assert math.atan2(-0.0, -1.0) == pytest.approx(-math.pi)
    assert oracle.ref_gpu_atan2(0.0, 0.0) == 0.0
    assert oracle.ref_gpu_atan2(1.0, 0.0) == pytest.approx(math.pi / 2)


def test_mutation_is_caught():
    """A planted slip in a copy of the per-pair formula (den's (b.c) la term
    dropped) moves the tetrahedron closed form: the pins above can see it."""
    V, F = W.tetrahedron()
    Q = np.array([[0.1, 0.1, 0.1]])

    def mutated(V, F, Q):
        T = V[F]
        a, b, c = (T[None, :, k, :] - Q[:, None, :] for k in range(3))
        la, lb, lc = (np.linalg.norm(x, axis=-1) for x in (a, b, c))
        det = (a * np.cross(b, c)).sum(-1)
        den = la * lb * lc + (a * b).sum(-1) * lc + (c * a).sum(-1) * lb
        return np.arctan2(det, den).sum(-1) / (2 * np.pi)

    assert abs(mutated(V, F, Q)[0] - 1.0) > 1e-3
    assert abs(oracle.winding_numpy(V, F, Q)[0] - 1.0) <= 1e-14
