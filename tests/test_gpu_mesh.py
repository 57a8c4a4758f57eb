"""Strip forward kernel (igl.Mesh -> wn_mesh_forward) against the oracle.
Bars as for the FP32-compensated face kernel: 1e-6 on closed meshes, 1e-5 on
open patches (FP32 rounding of block-local coordinates near a boundary)."""

import numpy as np
import pytest

import oracle
import wn_workloads as W

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


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
    assert np.abs(w - oracle.winding(V, Fs, Qs)).max() <= 1e-5


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
    w64 = igl.winding_number(V, F, Q, device="gpu")
    d = np.abs(w - w64)
    print(f"mesh c2@1M vs fp64: max {d.max():.3e} mean {d.mean():.3e}", flush=True)
    assert d.max() <= 1e-6


def test_mesh_rejects_bad_faces(require_gpu):
    import igl

    V, F = W.tetrahedron()
    with pytest.raises(IndexError):
        igl.Mesh(V, F + 4)
