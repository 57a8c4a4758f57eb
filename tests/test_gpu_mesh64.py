"""FP64 strip kernel (igl.Mesh.winding_number(precision="fp64") ->
wn_mesh_forward64) against the oracle at the FP64 bar, max |dw| <= 1e-10:
closed and open meshes, degenerate faces, queries 1e-7..1e-2 from the
surface, every split / Morton-order variant, device vertex updates, and the
plain FP64 kernel at full size. (Queries exactly on the surface are left
out: there w jumps and the +-pi branch of a face's angle depends on the
signed-zero rounding of its vertex order, in the reference as anywhere.)"""

import numpy as np
import pytest

import oracle
import wn_workloads as W

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _plain_fp64(V, F, Q):
    """The plain FP64 face kernel (wn_winding_gpu_ex, precision 2: the
    reference formula for every pair; mli_winding_gpu and igl.winding_number
    route calls this large to the strip kernels themselves)."""
    import igl
    from igl import pack_faces
    from igl._lib import addr, lib

    tri = pack_faces(np.asarray(V, dtype=np.float64), F)
    Qc = np.ascontiguousarray(Q, dtype=np.float64)
    out = np.empty(len(Qc))
    assert lib().wn_winding_gpu_ex(addr(tri), addr(Qc), addr(out), tri.shape[1], len(Qc), 2, 0, 0,
                                   0, -1.0) == 0, igl._lib.last_error()
    return out


BAR = 1e-10


def _w64(V, F, Q, **kw):
    import igl

    m = igl.Mesh(V, F)
    try:
        return m.winding_number(Q, precision="fp64", **kw)
    finally:
        m.close()


@pytest.mark.parametrize("splits,sort", [(1, 0), (3, 1), (2, 0), (0, -1)])
def test_fp64_config1(require_gpu, splits, sort):
    V, F, Q = W.config1()
    w = _w64(V, F, Q, splits=splits, sort=sort)
    err = np.abs(w - oracle.winding(V, F, Q)).max()
    print(f"fp64 strip c1 splits={splits} sort={sort}: max|dw| {err:.3e}", flush=True)
    assert err <= BAR


def test_fp64_config2_and_determinism(require_gpu):
    import igl

    V, F, Q = W.config2(1024)
    m = igl.Mesh(V, F)
    try:
        a = m.winding_number(Q, precision="fp64")
        b = m.winding_number(Q, precision="fp64")
    finally:
        m.close()
    assert np.array_equal(a, b)
    err = np.abs(a - oracle.winding(V, F, Q)).max()
    print(f"fp64 strip c2@1024: max|dw| {err:.3e}", flush=True)
    assert err <= BAR


@pytest.mark.parametrize("nf_take,nq", [(1, 3), (7, 257), (300, 513), (5120, 700)])
def test_fp64_open_patches(require_gpu, nf_take, nq):
    V, F, Q = W.config1()
    Fs, Qs = F[:nf_take], Q[:nq]
    err = np.abs(_w64(V, Fs, Qs, splits=1, sort=0) - oracle.winding(V, Fs, Qs)).max()
    assert err <= BAR


def test_fp64_degenerate_faces(require_gpu):
    """Faces with a repeated vertex and collinear faces contribute 0, as in
    the plain formula."""
    V, F = W.grid(9)
    V = np.vstack([V, [[0.5, 0.5, 0.0], [1.0, 1.0, 0.0], [1.5, 1.5, 0.0]]])
    n = len(V)
    extra = np.array([[0, 0, 5], [3, 7, 3], [n - 3, n - 2, n - 1], [n - 1, n - 2, n - 3]])
    F = np.vstack([F, extra])
    Q = np.random.default_rng(4).uniform(V.min(0) - 1, V.max(0) + 1, size=(999, 3))
    err = np.abs(_w64(V, F, Q) - oracle.winding(V, F, Q)).max()
    assert err <= BAR


def test_fp64_near_surface(require_gpu):
    """Points at distance 1e-7..1e-2 on both sides of random faces of a
    closed sphere: FP64 bar against the oracle, and the side (0 or 1)."""
    V, F = W.icosphere(3)
    rng = np.random.default_rng(12)
    f = rng.integers(0, len(F), 4096)
    bc = rng.dirichlet([2.0, 2.0, 2.0], size=4096)
    T = V[F[f]]
    p = np.einsum("qk,qkd->qd", bc, T)
    n = np.cross(T[:, 1] - T[:, 0], T[:, 2] - T[:, 0])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    d = 10.0 ** rng.uniform(-7, -2, 4096) * np.where(rng.random(4096) < 0.5, -1.0, 1.0)
    Q = p + d[:, None] * n
    w = _w64(V, F, Q)
    ref = oracle.winding(V, F, Q)
    err = np.abs(w - ref).max()
    side = np.where(d < 0, 1.0, 0.0)  # outward normals: inside is -n
    print(f"fp64 strip near-surface: max|dw| {err:.3e}, max|w - side| "
          f"{np.abs(w - side).max():.3e}", flush=True)
    assert err <= BAR
    assert np.abs(w - side).max() <= 1e-8


def test_fp64_device_vertices(require_gpu):
    """Mesh.from_faces + set_vertices (float64 and float32 V): the sweep
    records are rebuilt with the records, so FP64 results equal those of a
    mesh built from the same vertices on the host, bit for bit."""
    import igl
    from igl.mesh import Mesh

    V, F, Q = W.config2(4096)
    Qd = torch.tensor(Q, device="cuda")
    topo = Mesh.from_faces(F, len(V))
    try:
        for Vn in (V, V * 1.3 + [0.1, -0.2, 0.05], V.astype(np.float32)):
            topo.set_vertices(torch.tensor(Vn, device="cuda"))
            fresh = igl.Mesh(np.asarray(Vn, dtype=np.float64), F)
            try:
                a = topo.winding_number(Qd, precision="fp64").cpu().numpy()
                b = fresh.winding_number(Qd, precision="fp64").cpu().numpy()
            finally:
                fresh.close()
            np.testing.assert_array_equal(a, b)
            assert np.abs(a - oracle.winding(np.asarray(Vn, np.float64), F, Q)).max() <= BAR
    finally:
        topo.close()


def test_fp64_full_size_vs_plain_kernel(require_gpu):
    """28,320-face torus x 1M queries: FP64 strip kernel vs the plain FP64
    kernel."""
    import igl

    V, F, Q = W.config2(1_048_576)
    m = igl.Mesh(V, F)
    try:
        w = m.winding_number(torch.tensor(Q, device="cuda"), precision="fp64").cpu().numpy()
    finally:
        m.close()
    w64 = _plain_fp64(V, F, Q)
    d = np.abs(w - w64)
    print(f"fp64 strip c2@1M vs plain fp64: max {d.max():.3e} mean {d.mean():.3e}", flush=True)
    assert d.max() <= 1e-12


def test_fp64_rejects_bad_precision(require_gpu):
    import igl

    V, F = W.tetrahedron()
    m = igl.Mesh(V, F)
    try:
        with pytest.raises(ValueError):
            m.winding_number(np.zeros((1, 3)), precision="fp16")
    finally:
        m.close()
