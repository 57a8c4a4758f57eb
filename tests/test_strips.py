"""Triangle-strip records (wn_mesh_records, the host stripifier behind the
strip kernel), checked on the CPU: every face exactly once with its own
orientation, tile structure, and an FP64 evaluation of the records equal to
the oracle."""

import numpy as np
import pytest

import oracle
import wn_workloads as W
from igl.mesh import strip_records

TILE = 255


def _meshes():
    rng = np.random.default_rng(3)
    soup_V = rng.normal(size=(60, 3))
    soup_F = rng.integers(0, 60, size=(40, 3))
    soup_F = soup_F[(soup_F[:, 0] != soup_F[:, 1]) & (soup_F[:, 1] != soup_F[:, 2]) & (soup_F[:, 0] != soup_F[:, 2])]
    # three faces on one edge (non-manifold) plus a normal neighbour
    nm_V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [1, 1, 1.0]])
    nm_F = np.array([[0, 1, 2], [1, 0, 3], [0, 1, 4], [1, 5, 2]])
    cyl_V, cyl_F = W.open_cylinder(40, 12, 2, np.random.default_rng(9))
    return {
        "torus": W.torus(nu=60, nv=40),
        "ico3": W.icosphere(3),
        "grid": W.grid(15),
        "cylinder": (cyl_V, cyl_F),
        "soup": (soup_V, soup_F),
        "nonmanifold": (nm_V, nm_F),
        "one_face": (np.eye(3), np.array([[0, 1, 2]])),
    }


def _canon(tri):
    # cyclic rotation starting at the smallest index (keeps orientation)
    k = int(np.argmin(tri))
    return tuple(int(tri[(k + i) % 3]) for i in range(3))


@pytest.mark.parametrize("name", list(_meshes()))
def test_records_cover_every_face_once(name):
    V, F = _meshes()[name]
    R = strip_records(V, F)
    assert len(R) % TILE == 0 and len(R) > 0
    tiles = R.reshape(-1, TILE, 4)
    assert (tiles[:, :2, 3] == 0).all()  # every tile starts a strip
    lookup = {tuple(v): i for i, v in enumerate(np.asarray(V, dtype=np.float64))}
    got = []
    for t in tiles:
        for k in range(2, TILE):
            if t[k, 3] == 0:
                continue
            ids = [lookup[tuple(t[k - 2 + i, :3])] for i in range(3)]
            if t[k, 3] < 0:
                ids = ids[::-1]
            got.append(_canon(ids))
    want = sorted(_canon(f) for f in np.asarray(F))
    assert sorted(got) == want


@pytest.mark.parametrize("name", ["torus", "ico3", "cylinder", "soup"])
def test_records_evaluate_to_the_oracle(name):
    V, F = _meshes()[name]
    R = strip_records(V, F)
    Q = np.random.default_rng(5).uniform(V.min(0) - 0.3, V.max(0) + 0.3, size=(40, 3))
    emit = np.nonzero(R[:, 3])[0]
    a = R[emit - 2, None, :3] - Q[None]
    b = R[emit - 1, None, :3] - Q[None]
    c = R[emit, None, :3] - Q[None]
    la, lb, lc = (np.linalg.norm(x, axis=-1) for x in (a, b, c))
    det = (a * np.cross(b, c)).sum(-1)
    den = la * lb * lc + (a * b).sum(-1) * lc + (b * c).sum(-1) * la + (c * a).sum(-1) * lb
    w = (R[emit, 3][:, None] * np.arctan2(det, den)).sum(0) / (2 * np.pi)
    assert np.abs(w - oracle.winding(V, F, Q)).max() <= 1e-12


def test_strips_are_long_on_regular_meshes():
    V, F = W.torus()
    R = strip_records(V, F)
    print(f"torus: {len(R)} records for {len(F)} faces ({len(R) / len(F):.3f} per face)")
    assert len(R) <= 1.05 * len(F)


def test_bad_index_is_rejected():
    V, F = W.tetrahedron()
    with pytest.raises(IndexError):
        strip_records(V, F + 10)


@pytest.mark.parametrize("name", list(_meshes()))
def test_topology_gathers_to_the_records(name):
    """wn_strip_topology's (vertex id, flag) per record, gathered on the host,
    is exactly strip_records' output: the device gather only needs these."""
    from igl._lib import addr, lib

    V, F = _meshes()[name]
    F = np.ascontiguousarray(F, dtype=np.int64)
    h = lib()
    n = h.wn_strip_topology(addr(F), len(V), len(F), 0, 0, 0, 0)
    assert n >= 0 and n % TILE == 0
    idx = np.empty(n, np.int32)
    flag = np.empty(n, np.int8)
    assert h.wn_strip_topology(addr(F), len(V), len(F), 0, idx.ctypes.data, flag.ctypes.data, n) == n
    R = strip_records(V, F)
    assert len(R) == n
    np.testing.assert_array_equal(R[:, :3], np.asarray(V, np.float64)[idx])
    np.testing.assert_array_equal(R[:, 3], flag.astype(np.float64))
    assert h.wn_strip_topology(addr(F), int(F.max()), len(F), 0, 0, 0, 0) < 0
