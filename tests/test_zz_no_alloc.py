"""wn_mesh_forward_noalloc, the entry point that allocates nothing (its
scratch is a static device array), against wn_mesh_forward, the entry point it
is a variant of, through the C ABI.

Bars: with the same splits and Morton order the two agree bit for bit; with
the default configuration (whose split count the static entry lowers until its
scratch fits) within 1e-7: the split count decides which groups of three
records start a split, and such a group sums its three FP32 angles face by
face where elsewhere it takes them as one product (FP32-level rounding,
~5e-9 per group); both within 1e-6 of the oracle on a subsample (a closed
torus). Each case runs two calls in a row, the second with fewer
queries, so the second reads a scratch the first left behind. A configuration
beyond the 64 MiB scratch fails with -4 before anything is launched, and a
call leaves the device's free memory unchanged.
"""

import numpy as np
import pytest

import igl
import oracle
import wn_workloads as W

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

SCRATCH_BYTES = 64 << 20


@pytest.fixture(scope="module")
def torus_mesh():
    from igl.mesh import Mesh

    V, F = W.torus()
    mesh = Mesh(V, F)
    yield V, F, mesh
    mesh.close()


def _workspace_entry(h, mesh, Qd, splits, sort):
    n = Qd.shape[0]
    nbytes = h.wn_mesh_workspace_bytes(mesh.handle, n, splits, sort)
    assert nbytes >= 0, igl._lib.last_error()
    ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device="cuda")
    out = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
    st = h.wn_mesh_forward(mesh.handle, Qd.data_ptr(), out.data_ptr(), n, splits, sort, -1.0,
                           ws.data_ptr(), nbytes, torch.cuda.current_stream().cuda_stream)
    assert st == 0, igl._lib.last_error()
    return out


def _noalloc_entry(h, mesh, Qd, splits, sort):
    n = Qd.shape[0]
    out = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
    st = h.wn_mesh_forward_noalloc(mesh.handle, Qd.data_ptr(), out.data_ptr(), n, splits, sort,
                                   -1.0, torch.cuda.current_stream().cuda_stream)
    assert st == 0, igl._lib.last_error()
    return out


@pytest.mark.parametrize("splits,sort", [(5, 1), (1, 0), (12, 0)])
def test_noalloc_equals_workspace_entry(require_gpu, torus_mesh, splits, sort):
    V, F, mesh = torus_mesh
    h = igl.lib()
    Q = W.config2(1 << 20)[2]
    for n in (200_000, 5_000):  # second call smaller: stale scratch from the first
        print(f"noalloc: {n} queries, splits {splits}, sort {sort}", flush=True)
        Qd = torch.tensor(Q[:n], device="cuda")
        ref = _workspace_entry(h, mesh, Qd, splits, sort)
        got = _noalloc_entry(h, mesh, Qd, splits, sort)
        torch.cuda.synchronize()
        assert torch.equal(got, ref)


def test_noalloc_default_config_and_oracle(require_gpu, torus_mesh):
    V, F, mesh = torus_mesh
    h = igl.lib()
    Q = W.config2(1 << 20)[2]
    sub = np.arange(0, 1 << 20, 4099)
    expect = oracle.winding(V, F, Q[sub])
    for n in (1 << 20, 70_000):
        print(f"noalloc default config: {n} queries", flush=True)
        Qd = torch.tensor(Q[:n], device="cuda")
        ref = _workspace_entry(h, mesh, Qd, 0, -1).cpu().numpy()
        got = _noalloc_entry(h, mesh, Qd, 0, -1).cpu().numpy()
        assert np.abs(got - ref).max() <= 1e-7
        s = sub[sub < n]
        assert np.abs(got[s] - expect[: len(s)]).max() <= 1e-6


def test_noalloc_rejects_configs_beyond_the_scratch(require_gpu, torus_mesh):
    V, F, mesh = torus_mesh
    h = igl.lib()
    n = 200_000
    Qd = torch.tensor(W.config2(1 << 20)[2][:n], device="cuda")
    out = torch.full((n,), -7.0, dtype=torch.float64, device="cuda")
    splits = 64  # 64 x 200,000 x 8 bytes of partials > 64 MiB
    assert n * splits * 8 > SCRATCH_BYTES
    st = h.wn_mesh_forward_noalloc(mesh.handle, Qd.data_ptr(), out.data_ptr(), n, splits, 1,
                                   -1.0, torch.cuda.current_stream().cuda_stream)
    assert st == -4
    assert "static scratch" in igl._lib.last_error()
    torch.cuda.synchronize()
    assert bool((out == -7.0).all())  # nothing ran
    assert h.wn_mesh_forward_noalloc(0, Qd.data_ptr(), out.data_ptr(), n, 0, -1, -1.0, 0) == -1


def test_noalloc_call_leaves_free_memory_unchanged(require_gpu, torus_mesh):
    V, F, mesh = torus_mesh
    h = igl.lib()
    Qd = torch.tensor(W.config2(1 << 20)[2][:100_000], device="cuda")
    out = torch.empty(100_000, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    assert h.wn_mesh_forward_noalloc(mesh.handle, Qd.data_ptr(), out.data_ptr(), 100_000, 0, -1,
                                     -1.0, stream) == 0
    torch.cuda.synchronize()
    before = h.wn_device_free_bytes()
    for _ in range(3):
        assert h.wn_mesh_forward_noalloc(mesh.handle, Qd.data_ptr(), out.data_ptr(), 100_000, 0,
                                         -1, -1.0, stream) == 0
    torch.cuda.synchronize()
    assert h.wn_device_free_bytes() == before
