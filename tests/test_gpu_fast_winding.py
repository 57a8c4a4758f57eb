"""Hierarchical winding number on the GPU (wn_fwn_query) against the CPU
path (wn_fwn_query_cpu, itself equal to the oracle restatement, see
test_fast_winding.py): same tree, same walk, same expansion arithmetic; the
exact leaf faces use the device's sqrt/atan2, so the bar is 1e-12 absolute.
Also: the error against the exact winding number at full size (the 1.31M-face
displaced sphere), CUDA tensors in and out, and a mesh without faces."""

import numpy as np
import pytest

import igl
import oracle
import wn_workloads as W

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.mark.parametrize("mesh", ["icosphere", "torus", "open_cylinder", "soup"])
@pytest.mark.parametrize("beta", [2.0, 8.0, 1e9])
def test_gpu_equals_cpu(require_gpu, mesh, beta):
    rng = np.random.default_rng(7)
    if mesh == "icosphere":
        V, F = W.icosphere(4)
    elif mesh == "torus":
        V, F = W.torus(nu=60, nv=40)
    elif mesh == "open_cylinder":
        V, F = W.open_cylinder(48, 20, 2, np.random.default_rng(1))
    else:
        V = rng.normal(size=(400, 3))
        F = rng.integers(0, 400, size=(3000, 3)).astype(np.int64)
    lo, hi = V.min(0), V.max(0)
    Q = rng.uniform(lo - 0.3 * (hi - lo), hi + 0.3 * (hi - lo), size=(5000, 3))
    fw = igl.FastWindingNumber(V, F)
    cpu = fw.query(Q, beta=beta)
    gpu = fw.query(Q, beta=beta, device="gpu")  # 5000 queries: Morton order
    assert np.abs(gpu - cpu).max() <= 1e-12
    unsorted = fw.query(Q, beta=beta, device="gpu", sort=0)
    np.testing.assert_array_equal(unsorted, gpu)  # same walk per query
    if beta > 1e8:
        assert np.abs(gpu - oracle.winding(V, F, Q)).max() <= 1e-12
    assert fw.info()["device_bytes"] > 0
    fw.close()


def test_device_tensors_and_empty(require_gpu):
    V, F = W.icosphere(3)
    fw = igl.FastWindingNumber(V, F)
    Q = torch.tensor(np.random.default_rng(2).uniform(-1.5, 1.5, size=(3000, 3)), device="cuda")
    w = fw.query(Q, device="gpu")
    torch.cuda.synchronize()
    assert w.is_cuda and w.shape == (3000,)
    np.testing.assert_allclose(w.cpu().numpy(), fw.query(Q.cpu().numpy()), rtol=0, atol=1e-12)
    fw.close()
    Fe = np.zeros((0, 3), dtype=np.int64)
    out = igl.fast_winding_number(V, Fe, V[:10], device="gpu")
    np.testing.assert_array_equal(out, np.zeros(10))


def test_full_size_displaced_sphere(require_gpu):
    """1,310,720 faces; 65,536 of the stress workload's queries (half uniform,
    half at 1e-7..1e-2 from the surface): GPU against CPU on a subsample, and
    against the exact FP64 GPU winding number."""
    V, F = W.config3_mesh()
    Q, dist = W.config3_queries(V, F, 65536)
    fw = igl.FastWindingNumber(V, F)
    exact = igl.winding_number(V, F, Q, device="gpu")
    for beta, bound in ((2.0, 2e-2), (8.0, 1e-4)):
        gpu = fw.query(Q, beta=beta, device="gpu")
        err = np.abs(gpu - exact)
        print(f"beta {beta}: max err {err.max():.2e}, mean {err.mean():.2e}", flush=True)
        assert err.max() <= bound
        sub = slice(0, 512)
        assert np.abs(gpu[sub] - fw.query(Q[sub], beta=beta)).max() <= 1e-12
    fw.close()
