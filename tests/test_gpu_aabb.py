"""Point-mesh squared distance on the GPU (wn_aabb_*) against the CPU path
(mli_aabb_squared_distance, itself equal to the oracle bit for bit, see
test_aabb.py). Bar: bitwise equality of distance, face index and closest
point -- the GPU runs the same descent with the same arithmetic.

Cases: the reference tests' inputs (tests/test_parity.py:131-187, 425-445),
a triangle soup with degenerate faces and shared corners (exact ties), a
mesh without faces, no points, Morton-sorted and unsorted launches, device
tensors in and out, and the reference benchmark's mesh (bench/bench.py:
grid(350), 243,602 faces) with its 10,000 points and with 1,048,576 points.
"""

import numpy as np
import pytest

import igl
import oracle
import wn_workloads as W

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _both(V, F, P, sort=-1):
    tree = igl.AABB()
    tree.init(V, F)
    cpu = tree.squared_distance(V, F, P)
    if sort == -1:
        gpu = tree.squared_distance(V, F, P, device="gpu")
    else:
        out = tree.query_device(torch.tensor(P, device="cuda"), sort=sort)
        gpu = tuple(t.cpu().numpy() for t in out)
    return cpu, gpu, tree


def _assert_same(cpu, gpu):
    for a, b in zip(cpu, gpu):
        assert a.dtype == b.dtype and a.shape == b.shape
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("n,seed,kind", [(8, 4, "uniform"), (9, 12, "normal"), (10, 19, "uniform")])
@pytest.mark.parametrize("sort", [0, 1])
def test_reference_inputs(require_gpu, n, seed, kind, sort):
    V, F = W.grid(n)
    rng = np.random.default_rng(seed)
    if kind == "normal":
        P = rng.normal(size=(80, 3)) * [3.0, 3.0, 1.0] + [4.0, 4.0, 0.0]
    else:
        P = rng.uniform([-1.0, -1.0, -1.0], [n, n, 2.0], size=(100 if n == 8 else 257, 3))
    cpu, gpu, _ = _both(V, F, P, sort)
    _assert_same(cpu, gpu)
    np.testing.assert_array_equal(gpu[0], oracle.brute_distance(P, V, F)[0])


def test_degenerate_and_ties(require_gpu):
    V = np.array([[0.0, 0, 0], [2.0, 0, 0], [0.0, 1, 0]])
    F = np.array([[0, 1, 1], [2, 2, 2]], dtype=np.int64)
    P = np.array([[0.5, 1.0, 0.0], [0.0, 2.0, 0.0]])
    cpu, gpu, _ = _both(V, F, P)
    _assert_same(cpu, gpu)
    np.testing.assert_array_equal(gpu[0], [0.25, 1.0])
    # soup: random faces over few vertices, degenerate faces, and points at
    # the vertices themselves (many faces at distance exactly 0)
    rng = np.random.default_rng(5)
    V = rng.normal(size=(200, 3))
    F = rng.integers(0, 200, size=(3000, 3)).astype(np.int64)
    F[::13, 2] = F[::13, 0]
    P = np.concatenate([rng.normal(size=(5000, 3)) * 1.5, V])
    for sort in (0, 1):
        cpu, gpu, _ = _both(V, F, P, sort)
        _assert_same(cpu, gpu)
    used = np.zeros(len(V), dtype=bool)
    used[F.ravel()] = True
    assert (gpu[0][-len(V):][used] == 0.0).all()


def test_empty_cases(require_gpu):
    V, F = W.grid(4)
    cpu, gpu, tree = _both(V, F, np.zeros((0, 3)))
    _assert_same(cpu, gpu)
    Fe = np.zeros((0, 3), dtype=np.int64)
    cpu, gpu, _ = _both(V, Fe, V[:5])
    _assert_same(cpu, gpu)
    assert np.isinf(gpu[0]).all() and (gpu[1] == -1).all()
    one = F[:1]
    cpu, gpu, tree = _both(V, one, V)
    _assert_same(cpu, gpu)
    assert tree.gpu_info()["nodes"] == 1


def test_device_tensors_and_reuse(require_gpu):
    V, F = W.grid(30)
    tree = igl.AABB()
    tree.init(V, F)
    rng = np.random.default_rng(3)
    for n, sort in ((4096, 1), (1000, -1)):  # Morton order, then the default; tree reused
        P = rng.uniform([-2, -2, -2], [32, 32, 2], size=(n, 3))
        Pd = torch.tensor(P, device="cuda")
        s, i, c = tree.query_device(Pd, sort=sort)
        torch.cuda.synchronize()
        assert s.is_cuda and i.dtype == torch.int64 and c.shape == (n, 3)
        cpu = tree.squared_distance(V, F, P)
        _assert_same(cpu, (s.cpu().numpy(), i.cpu().numpy(), c.cpu().numpy()))
    info = tree.gpu_info()
    assert info["faces"] == len(F) and info["stack_size"] <= 40


@pytest.mark.parametrize("npts", [10_000, 1_048_576])
def test_reference_benchmark_mesh(require_gpu, npts):
    """bench/bench.py:131-140: grid(350) (243,602 faces), points uniform in
    [-10, 360]^2 x [-5, 5]."""
    V, F = W.bench_grid(350)
    P = np.random.default_rng(1).uniform([-10, -10, -5], [360, 360, 5], size=(npts, 3))
    cpu, gpu, _ = _both(V, F, P)
    _assert_same(cpu, gpu)
    sub = np.random.default_rng(2).choice(npts, 64, replace=False)
    np.testing.assert_array_equal(gpu[0][sub], oracle.brute_distance(P[sub], V, F)[0])
    print(f"grid(350) x {npts}: max sqrd {gpu[0].max():.3f}", flush=True)


@pytest.mark.parametrize("winding", ["exact", "fast"])
def test_signed_distance_gpu_equals_cpu(require_gpu, winding):
    V, F = W.icosphere(4)
    rng = np.random.default_rng(22)
    d = rng.normal(size=(20000, 3))
    P = d / np.linalg.norm(d, axis=1)[:, None] * rng.uniform(0.2, 1.8, size=(20000, 1))
    cpu = igl.signed_distance(P, V, F, winding=winding, beta=8.0)
    gpu = igl.signed_distance(P, V, F, device="gpu", winding=winding, beta=8.0)
    for a, b in zip(cpu, gpu):
        np.testing.assert_array_equal(a, b)
