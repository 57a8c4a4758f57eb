"""Precision stress at full mesh size: the displaced level-8 icosphere
(1,310,720 faces) with 262,144 of its workload's queries (half uniform in the
box, half at |d| in [1e-7, 1e-2] along a face normal, so the exact answer is
0 outside and 1 inside). Bars: FP32-compensated strip kernel vs FP64 plain
<= 1e-6 on the uniform half and <= 1e-4 near the surface; every near-surface
query on its correct side in both precisions; reversing every face's corner
order negates w (size-independent property) to 2e-6."""

import numpy as np
import pytest

import igl
import wn_workloads as W

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def c3():
    V, F = W.config3_mesh()
    Q, d = W.config3_queries(V, F, 262_144)
    return V, F, Q, d


def _strip(V, F, Q):
    from igl.mesh import Mesh

    m = Mesh(V, F)
    try:
        return m.winding_number(torch.tensor(Q, device="cuda")).cpu().numpy()
    finally:
        m.close()


def test_stress_precision_and_sides(require_gpu, c3):
    V, F, Q, d = c3
    w64 = igl.winding_number(V, F, Q, device="gpu")
    w32 = _strip(V, F, Q)
    near = ~np.isnan(d)
    err = np.abs(w32 - w64)
    print(f"uniform max {err[~near].max():.2e}, near-surface max {err[near].max():.2e}", flush=True)
    assert err[~near].max() <= 1e-6
    assert err[near].max() <= 1e-4
    expect = (d[near] < 0).astype(np.float64)  # inside along the outward normal
    assert (np.abs(w64[near] - expect) < 0.5).all()
    assert (np.abs(w32[near] - expect) < 0.5).all()


def test_stress_orientation_reversal(require_gpu, c3):
    V, F, Q, _ = c3
    sub = slice(0, 65_536)
    w = _strip(V, F, Q[sub])
    w_rev = _strip(V, np.ascontiguousarray(F[:, ::-1]), Q[sub])
    assert np.abs(w + w_rev).max() <= 2e-6
