"""The C-ABI boundary and the Python drop-in, on the CPU.

Mirrors the reference's winding tests (tests/test_parity.py:74-128, 270-291,
294-340, 402-427) with the upstream-libigl comparisons replaced by the oracle.
No GPU compute is called here.
"""

import os
import re

import numpy as np
import pytest

import igl
import oracle
import wn_workloads as W
from igl import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "wn.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for m in re.finditer(r"^\s*(?:const\s+)?[\w\s\*]+?\b(\w+)\s*\(([^)]*)\)\s*;", text, flags=re.M):
        name, args = m.group(1), m.group(2).strip()
        out[name] = 0 if args in ("", "void") else len(args.split(","))
    return out


def test_library_exports_every_declared_symbol():
    import ctypes

    handle = ctypes.CDLL(_lib.LIB)
    declared = header_functions()
    assert {"mli_winding", "mli_winding_gpu", "wn_forward"} <= set(declared)
    for name in declared:
        assert hasattr(handle, name), name


def test_signatures_match_header_arity():
    """Reference tests/test_parity.py:294-340: every declared signature takes
    exactly the arguments callers pass."""
    declared = header_functions()
    assert set(_lib._SIGNATURES) == set(declared)
    for name, (args, _) in _lib._SIGNATURES.items():
        assert len(args) == declared[name], name
    # the two reference entry points keep the reference's arity
    assert len(_lib._SIGNATURES["mli_winding"][0]) == 3 + 4
    assert len(_lib._SIGNATURES["mli_winding_gpu"][0]) == 3 + 2


def test_cpu_path_matches_the_oracle_and_the_golden_vectors():
    """The CPU path runs the oracle's arithmetic in the oracle's summation
    order on four AVX2 lanes with its own four-lane atan2 (within 2 ulps of
    libm's, which the oracle calls): within 1e-14 of the oracle on C1 and on
    the near-vertex and near-edge queries where the pair formula is least
    well conditioned, and within the oracle's own bars of the 50-digit mpmath
    vectors on the reference tests' inputs (tests/golden)."""
    import os

    V, F, Q = W.config1()
    d = np.abs(igl.winding_number(V, F, Q) - oracle.winding(V, F, Q))
    assert d.max() <= 1e-14
    # queries 1e-9 from vertices and edge midpoints of a curved open grid
    Vg, Fg = W.grid(9)
    rng = np.random.default_rng(5)
    r = rng.normal(size=(len(Vg), 3))
    Qg = np.vstack([Vg + 1e-9 * r / np.linalg.norm(r, axis=1, keepdims=True),
                    0.5 * (Vg[Fg[:, 0]] + Vg[Fg[:, 1]]) + 1e-5 * rng.normal(size=(len(Fg), 3))])
    assert np.abs(igl.winding_number(Vg, Fg, Qg) - oracle.winding(Vg, Fg, Qg)).max() <= 1e-12
    G = np.load(os.path.join(os.path.dirname(__file__), "golden", "winding_golden.npz"))
    for key in ("tet", "tail", "par"):
        Vk, Fk, Qk = G[f"{key}_V"], G[f"{key}_F"], G[f"{key}_Q"]
        if key == "par":
            Qk = Qk[G["par_pick"]]
        assert np.abs(igl.winding_number(Vk, Fk, Qk) - G[f"{key}_w"]).max() <= 1e-13, key
    pick = G["c1_pick"]
    assert np.abs(igl.winding_number(V, F, Q[pick]) - G["c1_w"]).max() <= 1e-12


def test_tetrahedron_published_invariant():
    """Reference tests/test_parity.py:74-85."""
    V, F = W.tetrahedron()
    Q = np.array([[0.1, 0.1, 0.1], [2.0, 2.0, 2.0], [-1.0, 0.2, 0.2]])
    w = igl.winding_number(V, F, Q)
    assert w[0] == pytest.approx(1.0, abs=1e-14)
    assert np.allclose(w, oracle.winding(V, F, Q), atol=1e-14)
    assert igl.winding_number(V, F, Q[0]) == pytest.approx(1.0, abs=1e-14)
    assert isinstance(igl.winding_number(V, F, Q[0]), float)
    assert igl.winding_number_for_point(V, F, Q[0]) == pytest.approx(1.0, abs=1e-14)


def test_simd_tail_and_parallel_threshold():
    """Reference tests/test_parity.py:88-105 on the committed mpmath values."""
    g = np.load(os.path.join(ROOT, "tests", "golden", "winding_golden.npz"))
    w = igl.winding_number(g["tail_V"], g["tail_F"], g["tail_Q"])
    assert np.abs(w - g["tail_w"]).max() <= 1e-14
    w = igl.winding_number(g["par_V"], g["par_F"], g["par_Q"])
    assert len(w) == 1619
    assert np.abs(w[g["par_pick"]] - g["par_w"]).max() <= 1e-12


def _force(monkeypatch, workers):
    monkeypatch.setattr(igl, "worker_count", lambda items, work, kernel: max(1, min(items, workers)))


def test_serial_and_parallel_output_is_identical(monkeypatch):
    """Reference tests/test_parity.py:418-427 (grid(20) = 722 faces)."""
    V, F = W.grid(20)
    rng = np.random.default_rng(77)
    Q = rng.uniform([0.0, 0.0, -2.0], [19.0, 19.0, 2.0], size=(512, 3))
    _force(monkeypatch, 1)
    serial = igl.winding_number(V, F, Q)
    for workers in (1, 2, 4, 8):
        _force(monkeypatch, workers)
        assert np.array_equal(igl.winding_number(V, F, Q), serial)


def test_worker_policy():
    assert _lib.worker_count(1, 10**9, "winding") == 1
    assert _lib.worker_count(1000, 1000, "winding") == 1
    assert 1 <= _lib.worker_count(1000, 10**7, "winding") <= _lib.MAX_WORKERS
    assert _lib.spans(10, 4) == [(0, 3), (3, 6), (6, 9), (9, 10)]


def test_input_contracts():
    """Reference tests/test_parity.py:270-291 (winding rows) and the
    validation helpers of python/igl/__init__.py:72-124."""
    V, F = W.grid(3)
    with pytest.raises(ValueError):
        igl.winding_number(V, F, [[np.nan, 0.0, 0.0]])
    with pytest.raises(ValueError):
        igl.winding_number(V, F, [[np.inf, 0.0, 0.0]])
    with pytest.raises(IndexError):
        igl.winding_number(V, np.array([[0, 1, len(V)]]), [[0.0, 0.0, 0.0]])
    with pytest.raises(IndexError):
        igl.winding_number(V, np.array([[0, -1, 2]]), [[0.0, 0.0, 0.0]])
    with pytest.raises(TypeError):
        igl.winding_number(V, F.astype(np.float64), [[0.0, 0.0, 0.0]])
    with pytest.raises(TypeError):
        igl.winding_number(V.astype(complex), F, [[0.0, 0.0, 0.0]])
    with pytest.raises(ValueError):
        igl.winding_number(V[:, :2], F, [[0.0, 0.0, 0.0]])
    with pytest.raises(ValueError):
        igl.winding_number(V, F, [0.0, 0.0])
    with pytest.raises(ValueError):
        igl.winding_number(V, F, np.zeros((2, 2)))
    with pytest.raises(ValueError):
        igl.winding_number(V, F, [[0.0, 0.0, 0.0]], device="tpu")
    with pytest.raises(ValueError):
        igl.winding_number(V, F, [[0.0, 0.0, 0.0]], precision="fp16")
    with pytest.raises(ValueError):
        igl.winding_number(V, F, [[0.0, 0.0, 0.0]], precision="fp32c")  # GPU-only mode
    with pytest.raises(OverflowError):
        igl.winding_number(V, np.array([[0, 1, 2**63]], dtype=np.uint64), [[0.0, 0.0, 0.0]])
    with pytest.raises(ValueError):
        igl.winding_number(np.array([[2**60, 0, 0], [0, 1, 0], [0, 0, 1]]), [[0, 1, 2]], [[0, 0, 0]])
    # accepted: integer coordinates, int32 faces, empty queries, empty faces
    assert igl.winding_number(V.astype(np.int64), F.astype(np.int32), [[0.5, 0.5, 1.0]]).shape == (1,)
    assert igl.winding_number(V, F, np.zeros((0, 3))).shape == (0,)
    assert np.array_equal(igl.winding_number(V, np.zeros((0, 3), dtype=np.int64), [[0.0, 0, 0]]), [0.0])


def test_winding_gpu_request_and_cpu_fallback(monkeypatch):
    """Reference tests/test_parity.py:108-114 (no-device half; the GPU half
    is in tests/test_gpu_forward.py): with no usable device a GPU request
    returns the CPU path's bits (python/igl/__init__.py:251,264)."""
    V, F = W.tetrahedron()
    Q = np.array([[0.1, 0.1, 0.1], [2.0, 2.0, 2.0]])
    monkeypatch.delenv(igl.STRICT_ENV, raising=False)
    expected = igl.winding_number(V, F, Q)
    monkeypatch.setattr(igl, "_gpu_memory_available", lambda: False)
    assert np.array_equal(igl.winding_number(V, F, Q, device="gpu"), expected)
    assert np.array_equal(igl.winding_number(V, F, Q, device="gpu", precision="fp32c"), expected)
    assert igl.winding_number(V, F, Q[0], device="gpu") == float(expected[0])


def test_gpu_request_without_device_strict(monkeypatch):
    """WN_GPU_STRICT=1 turns the no-device fallback into RuntimeError (every
    GPU entry of the package); unset or 0 keeps the reference's fallback."""
    V, F = W.tetrahedron()
    P = np.array([[0.1, 0.1, 0.1], [2.0, 2.0, 2.0]])
    monkeypatch.setattr(igl, "_gpu_memory_available", lambda: False)
    monkeypatch.setenv(igl.STRICT_ENV, "1")
    with pytest.raises(RuntimeError, match="GPU"):
        igl.winding_number(V, F, P, device="gpu")
    tree = igl.AABB()
    tree.init(V, F)
    with pytest.raises(RuntimeError, match="GPU"):
        tree.squared_distance(V, F, P, device="gpu")
    with pytest.raises(RuntimeError, match="GPU"):
        igl.fast_winding_number(V, F, P, device="gpu")
    monkeypatch.setenv(igl.STRICT_ENV, "0")
    assert np.array_equal(igl.winding_number(V, F, P, device="gpu"), igl.winding_number(V, F, P))


def test_distance_and_fast_winding_cpu_fallback(monkeypatch):
    """AABB.squared_distance and fast_winding_number follow the same rule:
    no usable device -> the CPU path's results, bit for bit."""
    V, F = W.grid(8)
    rng = np.random.default_rng(4)
    P = rng.uniform(-1.0, 8.0, size=(50, 3))
    monkeypatch.delenv(igl.STRICT_ENV, raising=False)
    tree = igl.AABB()
    tree.init(V, F)
    cpu = tree.squared_distance(V, F, P)
    fast = igl.fast_winding_number(V, F, P)
    monkeypatch.setattr(igl, "_gpu_memory_available", lambda: False)
    gpu = tree.squared_distance(V, F, P, device="gpu")
    for a, b in zip(cpu, gpu):
        assert np.array_equal(a, b)
    assert np.array_equal(igl.point_mesh_squared_distance(P, V, F, device="gpu")[0], cpu[0])
    assert np.array_equal(igl.fast_winding_number(V, F, P, device="gpu"), fast)


def test_gpu_kernel_failure_is_not_swallowed(monkeypatch):
    """Reference tests/test_parity.py:117-128."""
    V, F = W.tetrahedron()

    class FailedGpu:
        @staticmethod
        def mli_winding_gpu(*args):
            return 0

        @staticmethod
        def wn_winding_gpu_ex(*args):
            return -3

        @staticmethod
        def wn_winding_gpu_vf(*args):
            return -3

    monkeypatch.setattr(igl, "_gpu_memory_available", lambda: True)
    monkeypatch.setattr(igl, "lib", lambda: FailedGpu())
    with pytest.raises(RuntimeError, match="GPU"):
        igl.winding_number(V, F, [[0.1, 0.1, 0.1]], device="gpu")
    with pytest.raises(RuntimeError, match="GPU"):
        igl.winding_number(V, F, [[0.1, 0.1, 0.1]], device="gpu", precision="fp32c")


def test_no_device_here_reports_through_the_abi(monkeypatch):
    h = igl.lib()
    if h.wn_device_count() > 0:
        pytest.skip("a HIP device is visible; this checks the no-device path")
    V, F = W.tetrahedron()
    tri = igl.pack_faces(V, F)
    Q = np.zeros((1, 3))
    out = np.zeros(1)
    assert h.mli_winding_gpu(tri.ctypes.data, Q.ctypes.data, out.ctypes.data, 4, 1) == 0
    assert "no HIP device" in _lib.last_error()
    assert h.wn_device_free_bytes() < 0
    # the real probe (not a monkeypatched one) sends the call to the CPU path
    monkeypatch.delenv(igl.STRICT_ENV, raising=False)
    assert igl.winding_number(V, F, Q, device="gpu") == igl.winding_number(V, F, Q)
    monkeypatch.setenv(igl.STRICT_ENV, "1")
    with pytest.raises(RuntimeError, match="GPU"):
        igl.winding_number(V, F, Q, device="gpu")


def test_config_validation_through_the_abi():
    h = igl.lib()
    cfg = np.zeros(4, dtype=np.int64)
    assert h.wn_resolve_config(1000, 10**6, 1, 0, 0, -1, -1.0, cfg.ctypes.data) == 0
    assert cfg[0] in (1, 2, 4) and cfg[1] >= 1 and cfg[2] == 1 and cfg[3] == 1
    assert h.wn_resolve_config(28320, 1024, 0, 0, 0, -1, -1.0, cfg.ctypes.data) == 0
    assert cfg[0] == 1 and cfg[1] > 1  # split-K for few queries
    assert h.wn_resolve_config(10, 10, 7, 0, 0, 0, 1.0, cfg.ctypes.data) == -1
    assert h.wn_resolve_config(10, 10, 3, 0, 0, 0, 1.0, cfg.ctypes.data) == -1
    # precision 2 (FP64 plain, the metric's comparison kernel): no sort, qpl 1 or 2
    assert h.wn_resolve_config(28320, 1024, 2, 0, 0, -1, -1.0, cfg.ctypes.data) == 0
    assert cfg[0] == 1 and cfg[1] > 1 and cfg[2] == 0 and cfg[3] == 2
    assert h.wn_resolve_config(10, 10, 2, 4, 0, 0, 1.0, cfg.ctypes.data) == -1
    # the FP64 face kernel's 128-byte records: workspace grows with nf, plain needs none
    assert h.wn_forward_workspace_bytes(1000, 256, 2, 1, 1, 0) == 0
    assert h.wn_forward_workspace_bytes(1000, 256, 0, 1, 1, 0) >= 128 * 1000
    assert h.wn_resolve_config(10, 10, 0, 3, 0, 0, 1.0, cfg.ctypes.data) == -1
    assert h.wn_resolve_config(10, 10, 0, 1, 70000, 0, 1.0, cfg.ctypes.data) == -1
    assert h.wn_resolve_config(10, 2**31, 0, 1, 1, 0, 1.0, cfg.ctypes.data) == -1
    assert h.wn_forward(0, 0, 0, 10, 10, 9, 1, 1, 0, 1.0, 0, 0, 0) == -1
    assert "precision" in _lib.last_error()


def test_new_entries_validate_arguments_before_the_device():
    """Argument errors of the strip, edge and batch entry points come back as
    -1 with a message, before any device work (no GPU needed)."""
    from igl._lib import addr

    h = igl.lib()
    V, F = W.tetrahedron()
    V = np.ascontiguousarray(V, dtype=np.float64)
    F = np.ascontiguousarray(F, dtype=np.int64)
    Q = np.zeros((2, 3))
    out = np.zeros(2)
    # host strip route: precision, then sizes
    assert h.wn_winding_gpu_mesh(addr(V), addr(F), 4, 4, 0, addr(Q), addr(out), 2, 5) == -1
    assert "precision" in _lib.last_error()
    assert h.wn_winding_gpu_mesh(addr(V), addr(F), 4, 4, 0, addr(Q), addr(out), -1, 0) == -1
    assert h.wn_winding_gpu_mesh(addr(V), addr(F), 4, 4, 0, addr(Q), addr(out), 0, 0) == 0
    # edge adjoint: dtype, lanes, splits
    assert h.wn_backward_edges(0, 4, 3, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0) == -1
    assert "vdtype" in _lib.last_error()
    assert h.wn_backward_edges(0, 4, 0, 0, 0, 0, 0, 0, 0, 0, 1, 3, 1, 0) == -1
    assert "queries_per_lane" in _lib.last_error()
    assert h.wn_backward_edges(0, 4, 0, 0, 0, 0, 0, 0, 0, 0, 1, 4, 70000, 0) == -1
    # device record gather: dtype
    assert h.wn_strip_gather(0, 4, 2, 0, 0, 10, 0, 0) == -1
    assert "vdtype" in _lib.last_error()
    # meshes: unknown handles
    assert h.wn_mesh_forward64(12345, addr(Q), addr(out), 2, 0, 0, 0, 0, 0) == -1
    assert "unknown mesh" in _lib.last_error()
    # FP64 strip plans: offsets must be whole tiles
    ro = np.array([0, 100], dtype=np.int64)
    qo = np.array([0, 10], dtype=np.int64)
    assert h.wn_batch_create_strips64(addr(ro), addr(qo), 1, 0, 0) == -1
    # topology and boundary edges: index range
    bad = F + 10
    assert h.wn_strip_topology(addr(bad), 4, 4, 0, 0, 0, 0) == -1
    assert h.wn_boundary_edges(addr(bad), 4, 4, 0, 0, 0, 0) == -1
    assert "outside V" in _lib.last_error()
    assert h.wn_boundary_edges(addr(F), 2**31, 4, 0, 0, 0, 0) == -1
    # host V/F forward: index range (int64 and int32 faces), face dtype, null mesh
    assert h.wn_winding_gpu_vf(addr(V), addr(bad), 4, 4, 0, addr(Q), addr(out), 2, 0) == -1
    assert "face 0 has a corner outside V" in _lib.last_error()
    bad32 = np.ascontiguousarray(F, dtype=np.int32)
    bad32[3, 2] = -1
    assert h.wn_winding_gpu_vf(addr(V), addr(bad32), 4, 4, 1, addr(Q), addr(out), 2, 0) == -1
    assert "face 3 has a corner outside V" in _lib.last_error()
    assert h.wn_winding_gpu_vf(addr(V), addr(F), 4, 4, 2, addr(Q), addr(out), 2, 0) == -1
    assert "fdtype" in _lib.last_error()
    assert h.wn_winding_gpu_vf(0, 0, 4, 4, 0, addr(Q), addr(out), 2, 0) == -1
    assert h.wn_winding_gpu_vf(addr(V), addr(F), 4, 4, 0, addr(Q), addr(out), 0, 0) == 0
