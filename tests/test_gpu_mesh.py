"""Strip forward kernel (igl.Mesh -> wn_mesh_forward) against the oracle.
Bars as for the FP32-compensated face kernel: 1e-6 on closed meshes and on
open patches (the boundary correction undoes the FP32 rounding of the
block-local boundary)."""

import numpy as np
import pytest

import oracle
import wn_workloads as W

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _plain_fp64(V, F, Q):
    """The plain FP64 face kernel through the reference's host entry
    (igl.winding_number routes calls this large to the strip kernels)."""
    from igl import pack_faces
    from igl._lib import addr, lib

    tri = pack_faces(np.asarray(V, dtype=np.float64), F)
    Qc = np.ascontiguousarray(Q, dtype=np.float64)
    out = np.empty(len(Qc))
    assert lib().mli_winding_gpu(addr(tri), addr(Qc), addr(out), tri.shape[1], len(Qc)) == 1
    return out


def _mesh(V, F):
    import igl

    return igl.Mesh(V, F)


@pytest.mark.parametrize("splits,sort", [(1, 0), (3, 1), (0, -1)])
def test_mesh_config1(require_gpu, splits, sort):
    V, F, Q = W.config1()
    m = _mesh(V, F)
    try:
        w = m.winding_number(Q, splits=splits, sort=sort)
    finally:
        m.close()
    err = np.abs(w - oracle.winding(V, F, Q)).max()
    print(f"mesh c1 splits={splits} sort={sort}: max|dw| {err:.3e}", flush=True)
    assert err <= 1e-6


def test_mesh_config2_and_determinism(require_gpu):
    V, F, Q = W.config2(1024)
    m = _mesh(V, F)
    try:
        a = m.winding_number(Q)
        b = m.winding_number(Q)
        allfp64 = m.winding_number(Q, kappa=1e30)
        info = m.info()
    finally:
        m.close()
    ref = oracle.winding(V, F, Q)
    assert np.array_equal(a, b)
    assert np.abs(a - ref).max() <= 1e-6
    assert np.abs(allfp64 - ref).max() <= 1e-6
    assert info["faces"] == len(F) and info["records"] % 255 == 0


@pytest.mark.parametrize("nf_take,nq", [(1, 3), (7, 257), (300, 513)])
def test_mesh_open_patches(require_gpu, nf_take, nq):
    V, F, Q = W.config1()
    Fs, Qs = F[:nf_take], Q[:nq]
    m = _mesh(V, Fs)
    try:
        w = m.winding_number(Qs, splits=1, sort=0)
    finally:
        m.close()
    assert np.abs(w - oracle.winding(V, Fs, Qs)).max() <= 1e-6


def test_mesh_matches_face_kernel_full_size(require_gpu):
    """28,320-face torus x 1M queries: strip kernel vs FP64 plain."""
    import igl

    V, F, Q = W.config2(1_048_576)
    m = _mesh(V, F)
    try:
        Qd = torch.tensor(Q, device="cuda")
        w = m.winding_number(Qd).cpu().numpy()
    finally:
        m.close()
    w64 = _plain_fp64(V, F, Q)
    d = np.abs(w - w64)
    print(f"mesh c2@1M vs fp64: max {d.max():.3e} mean {d.mean():.3e}", flush=True)
    assert d.max() <= 1e-6


@pytest.mark.parametrize("nq,splits,sort", [
    (1, 0, -1), (300, 3, -1), (524_287, 0, -1), (524_289, 0, -1), (533_765, 7, -1),
    (700_001, 64, -1), (1_048_593, 0, -1), (1_048_593, 1, -1), (600_000, 0, 0)])
def test_strip_layout_boundaries(require_gpu, nq, splits, sort):
    """The strip kernel's block layout: whole query blocks for all but the
    last round of resident blocks (2,048 on MI355X: 524,288 queries), that
    round's query blocks split over tile ranges. Query counts around that
    boundary, a partial last block, more splits than tiles (64 on this mesh's
    21 tiles), one split, and unsorted queries (partials written without the
    Morton permutation): FP32c within 1e-6 of the FP64 strip kernel (which
    splits every query block), and 1e-6 of the oracle on a subsample."""
    V, F, _ = W.config1()
    rng = np.random.default_rng(nq)
    Q = rng.uniform(-1.5, 1.5, size=(nq, 3))
    m = _mesh(V, F)
    try:
        print(f"strip layout: nq {nq}, splits {splits}, sort {sort}", flush=True)
        w32 = m.winding_number(Q, splits=splits, sort=sort)
        w64 = m.winding_number(Q, precision="fp64")
    finally:
        m.close()
    assert np.isfinite(w32).all()
    assert np.abs(w32 - w64).max() <= 1e-6
    idx = np.unique(np.concatenate([np.arange(min(nq, 64)), rng.integers(0, nq, 64), [nq - 1]]))
    ref = oracle.winding(V, F, Q[idx])
    assert np.abs(w32[idx] - ref).max() <= 1e-6


@pytest.mark.parametrize("precision", ["fp32c", "fp64"])
def test_strip_runs_are_bitwise_repeatable(require_gpu, precision):
    """Run-to-run bit equality of the strip kernels on an open cap with
    queries at 1e-6 .. 1e-1 from its faces, which send many groups of faces
    through every fallback of the group tests (two sixes, triples, single
    faces, the FP64 re-evaluation) and through the boundary correction: the
    path a wave takes must depend on the data only (a hazard that let a test
    read a stale register once made ~70 % of such queries differ between
    runs by a few 1e-8)."""
    V, F = W.config5()[:2]
    V = V.astype(np.float64)
    cz = V[F].mean(axis=1)[:, 2]
    Fs = F[np.sort(np.argsort(-cz)[:2000])]
    rng = np.random.default_rng(77)
    T = V[Fs[rng.integers(0, len(Fs), 100_000)]]
    bary = rng.dirichlet(np.ones(3), size=len(T))
    p = np.einsum("ij,ijk->ik", bary, T)
    n = np.cross(T[:, 1] - T[:, 0], T[:, 2] - T[:, 0])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    d = 10.0 ** rng.uniform(-6, -1, size=len(T)) * rng.choice([-1.0, 1.0], size=len(T))
    Q = np.vstack([p + d[:, None] * n, rng.uniform(V.min(0), V.max(0), size=(162_144, 3))])
    m = _mesh(V, Fs)
    try:
        runs = [m.winding_number(Q, precision=precision) for _ in range(3)]
    finally:
        m.close()
    for w in runs[1:]:
        assert np.array_equal(w, runs[0]), int((w != runs[0]).sum())


def test_mesh_rejects_bad_faces(require_gpu):
    import igl

    V, F = W.tetrahedron()
    with pytest.raises(IndexError):
        igl.Mesh(V, F + 4)


def test_topology_mesh_with_device_vertices(require_gpu):
    """Mesh.from_faces + set_vertices gathers the same records on the device
    as Mesh(V, F) builds on the host: identical w; moving the vertices and
    re-gathering equals a fresh mesh of the moved vertices."""
    from igl.mesh import Mesh

    V, F, Q = W.config2(4096)
    Qd = torch.tensor(Q, device="cuda")
    ref = _mesh(V, F)
    topo = Mesh.from_faces(F, len(V))
    try:
        topo.set_vertices(torch.tensor(V, device="cuda"))
        np.testing.assert_array_equal(topo.winding_number(Qd).cpu().numpy(),
                                      ref.winding_number(Qd).cpu().numpy())
        V2 = V * 1.3 + [0.1, -0.2, 0.05]
        topo.set_vertices(torch.tensor(V2, device="cuda"))
        fresh = _mesh(V2, F)
        np.testing.assert_array_equal(topo.winding_number(Qd).cpu().numpy(),
                                      fresh.winding_number(Qd).cpu().numpy())
        fresh.close()
        V3 = V.astype(np.float32)
        topo.set_vertices(torch.tensor(V3, device="cuda"))
        f32 = _mesh(V3.astype(np.float64), F)
        np.testing.assert_array_equal(topo.winding_number(Qd).cpu().numpy(),
                                      f32.winding_number(Qd).cpu().numpy())
        f32.close()
        with pytest.raises(ValueError):
            topo.set_vertices(torch.zeros((len(V) + 1, 3), device="cuda"))
        with pytest.raises(TypeError):
            topo.set_vertices(V)
    finally:
        topo.close()
        ref.close()
    with pytest.raises(IndexError):
        Mesh.from_faces(F, len(V) - 1)


def test_winding_field_forward_and_gradients(require_gpu):
    """WindingField (strip forward, FP64 adjoint) against winding_number:
    w within the FP32-compensated bar of FP64, gradients equal to the FP64
    path's (same adjoint kernel on the same faces) to rounding."""
    import igl.autograd as A

    V, F = W.icosphere(3)
    Q = np.random.default_rng(3).uniform(-1.4, 1.4, size=(3000, 3))
    t = np.random.default_rng(4).uniform(size=3000)
    field = A.WindingField(F, len(V))
    out = []
    for use_field in (True, False):
        Vd = torch.tensor(V, device="cuda", requires_grad=True)
        Od = torch.tensor(Q, device="cuda", requires_grad=True)
        w = field(Vd, Od) if use_field else A.winding_number(Vd, torch.tensor(F, device="cuda"), Od)
        ((w - torch.tensor(t, device="cuda")) ** 2).sum().backward()
        out.append((w.detach().cpu().numpy(), Vd.grad.cpu().numpy(), Od.grad.cpu().numpy()))
    field.close()
    (w1, gv1, go1), (w0, gv0, go0) = out
    assert np.abs(w1 - w0).max() <= 1e-6
    # dL/dw differs by 2 (w1 - w0) <= 2e-6 per query, so gradients agree to that scale
    assert np.abs(gv1 - gv0).max() <= 1e-4 * max(1.0, np.abs(gv0).max())
    assert np.abs(go1 - go0).max() <= 1e-4 * max(1.0, np.abs(go0).max())


def test_winding_field_fp64_forward(require_gpu):
    """WindingField(precision="fp64"): the FP64 strip forward, within the FP64
    bar of the oracle, with the same (edge) adjoint as the default field."""
    import igl.autograd as A

    V, F = W.icosphere(3)
    F = F[: len(F) - 40]  # an open surface: nonzero gradients
    Q = np.random.default_rng(5).uniform(-1.4, 1.4, size=(2000, 3))
    out = []
    for precision in ("fp64", "fp32c"):
        field = A.WindingField(F, len(V), precision=precision)
        Vd = torch.tensor(V, device="cuda", requires_grad=True)
        Od = torch.tensor(Q, device="cuda", requires_grad=True)
        w = field(Vd, Od)
        w.sum().backward()
        out.append((w.detach().cpu().numpy(), Vd.grad.cpu().numpy(), Od.grad.cpu().numpy()))
        field.close()
    (w64, gv64, go64), (w32, gv32, go32) = out
    assert np.abs(w64 - oracle.winding(V, F, Q)).max() <= 1e-10
    assert np.abs(w32 - w64).max() <= 1e-6
    # dL/dw = 1 in both: the same adjoint call (grad_V summed with atomics,
    # so equal to rounding)
    assert np.abs(gv64 - gv32).max() <= 1e-12 * np.abs(gv64).max()
    assert np.abs(go64 - go32).max() <= 1e-12 * np.abs(go64).max()
    with pytest.raises(ValueError):
        A.WindingField(F, len(V), precision="fp16")
