"""GPU parity of the forward kernels against the FP64 oracle (restatement of
the reference CPU path, src/capi.mojo:205-294).

Precision modes: 0 = FP64 (far pairs by the arctangent polynomial, the rest
by the reference formula), 1 = FP32-compensated, 2 = FP64 plain (the
reference formula for every pair). Bars: both FP64 modes within 1e-10 of the
oracle (the contract of the reference's FP64 mode) and FP64 within 1e-13 of
FP64 plain; FP32-compensated within 1e-6 of the oracle on every query of the
sizes run here. All calls go through the C ABI (include/wn.h).
"""

import numpy as np
import pytest

import igl
import oracle
import wn_workloads as W

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
FP32C_TOL = 1e-6


def host_call(tri, Q, precision, qpl=0, splits=0, sort=-1, kappa=-1.0):
    tri = np.ascontiguousarray(tri, dtype=np.float64)
    Q = np.ascontiguousarray(Q, dtype=np.float64).reshape(-1, 3)
    out = np.full(len(Q), np.nan)
    h = igl.lib()
    status = h.wn_winding_gpu_ex(tri.ctypes.data, Q.ctypes.data, out.ctypes.data,
                                 tri.shape[1], len(Q), precision, qpl, splits, sort, kappa)
    assert status == 0, igl._lib.last_error()
    return out


@pytest.fixture(scope="module")
def c1():
    V, F, Q = W.config1()
    tri = oracle.pack_soa(V, F)
    return V, F, Q, tri, oracle.winding_soa(tri, Q)


@pytest.fixture(scope="module")
def c2():
    V, F, Q = W.config2(1024)
    tri = oracle.pack_soa(V, F)
    return V, F, Q, tri, oracle.winding_soa(tri, Q)


def test_tetrahedron_gpu_matches_cpu(require_gpu):
    """Reference tests/test_parity.py:108-114: GPU == CPU within 1e-14."""
    V, F = W.tetrahedron()
    Q = np.array([[0.1, 0.1, 0.1], [2.0, 2.0, 2.0], [-1.0, 0.2, 0.2]])
    cpu = igl.winding_number(V, F, Q)
    gpu = igl.winding_number(V, F, Q, device="gpu")
    assert np.allclose(gpu, cpu, atol=1e-14, rtol=0)
    assert gpu[0] == pytest.approx(1.0, abs=1e-14)
    assert igl.winding_number(V, F, Q[0], device="gpu") == pytest.approx(1.0, abs=1e-14)


FAST_VS_PLAIN_TOL = 1e-13


@pytest.mark.parametrize("precision", [0, 2])
@pytest.mark.parametrize("qpl,splits", [(1, 1), (2, 1), (1, 7), (2, 20), (0, 0)])
def test_fp64_config1(require_gpu, c1, precision, qpl, splits):
    V, F, Q, tri, ref = c1
    out = host_call(tri, Q, precision, qpl, splits)
    print(f"fp64 mode {precision} c1 qpl={qpl} splits={splits} "
          f"max|dw|={np.abs(out - ref).max():.3e}", flush=True)
    assert np.abs(out - ref).max() <= FP64_TOL


def test_fp64_config2_1024(require_gpu, c2):
    V, F, Q, tri, ref = c2
    for precision in (0, 2):
        out = host_call(tri, Q, precision)
        assert np.abs(out - ref).max() <= FP64_TOL
        # closed torus: integer winding numbers
        assert np.abs(out - np.round(out)).max() <= FP64_TOL


def special_case_mesh():
    """An 8 x 8 curved grid (open) plus degenerate faces (a repeated corner,
    three collinear corners, a zero-area sliver 1e-12 wide), with queries
    1e-10 from the grid's vertices and 1e-5 from edge midpoints (random
    directions),
    on edge lines beyond the edge, 1e-9 and 1e-5 above face centroids, and
    uniform in a box around it. Points ON the surface are left out: there
    the reference formula's det and den are both zero for the faces through
    the point, and atan2 of the signed zeros that rounding leaves gives 0 or
    +-pi (w off by 1/2 or 1) in any implementation. Next to an edge, at
    distance d, det and den of its two faces both cancel to O(d): the
    formula's own error is ~1e-16 x (edge length / d) whatever the order of
    its operations (6e-8 at d = 1e-9, 1e-10 at 1e-6), hence 1e-5."""
    V, Fg = W.grid(8)
    V = np.vstack([V, [[0.3, 0.3, 0.0], [0.5, 0.5, 0.0], [0.7, 0.7, 0.0],
                       [0.2, 0.8, 0.1], [0.2 + 1e-12, 0.8, 0.1], [0.6, 0.1, 0.3]]])
    n = len(V) - 6
    F = np.vstack([Fg, [[0, 0, 5], [n, n + 1, n + 2], [n + 3, n + 4, n + 5]]])
    rng = np.random.default_rng(44)
    T = V[Fg]  # edge and face queries from the grid's own faces
    cen = T.mean(axis=1)
    nrm = np.cross(T[:, 1] - T[:, 0], T[:, 2] - T[:, 0])
    ln = np.linalg.norm(nrm, axis=1, keepdims=True)
    nrm = np.where(ln > 0, nrm / np.where(ln > 0, ln, 1.0), 0.0)
    def directions(k):
        d = rng.normal(size=(k, 3))
        return d / np.linalg.norm(d, axis=1, keepdims=True)

    Q = np.vstack([V[:n] + 1e-10 * directions(n),
                   0.5 * (T[:, 0] + T[:, 1]) + 1e-5 * directions(len(Fg)),
                   T[:, 0] + 0.5 * (T[:, 0] - T[:, 1]),
                   cen + 1e-9 * nrm, cen - 1e-5 * nrm,
                   rng.uniform(V.min(0) - 0.5, V.max(0) + 0.5, size=(2000, 3))])
    near = np.zeros(len(Q), bool)
    near[:n] = True  # 1e-10 from the grid's vertices
    return V, F, np.ascontiguousarray(Q), near


def test_fp64_fast_path_against_plain(require_gpu, c1, c2):
    """Mode 0 (far pairs by the polynomial, refined square roots and
    quotient) agrees with mode 2 (the reference formula for every pair) to
    1e-13 and both meet the oracle bar: on the reference-size meshes, on
    degenerate faces and queries next to vertices and edges, on the same mesh
    moved 1e6 from the origin (queries off the surface by more than the
    coordinates' rounding), and with queries that are NaN or infinite (NaN
    either way)."""
    cases = {"c1": (c1[3], c1[2]), "c2": (c2[3], c2[2])}
    V, F, Q, near = special_case_mesh()
    cases["special"] = (oracle.pack_soa(V, F), Q)
    # at 1e6 the coordinates' rounding (1.2e-10) would put the queries next to
    # vertices onto the surface: those are left out there
    shift = np.array([1e6, -3e5, 2e6])
    cases["special at 1e6"] = (oracle.pack_soa(V + shift, F), np.ascontiguousarray(Q[~near] + shift))
    for name, (tri, Qc) in cases.items():
        fast = host_call(tri, Qc, 0)
        plain = host_call(tri, Qc, 2)
        ref = oracle.winding_soa(tri, Qc)
        d = np.abs(fast - plain)
        print(f"{name}: fast vs plain {d.max():.3e}, fast vs oracle "
              f"{np.abs(fast - ref).max():.3e}, plain vs oracle {np.abs(plain - ref).max():.3e}",
              flush=True)
        assert d.max() <= FAST_VS_PLAIN_TOL, name
        assert np.abs(fast - ref).max() <= FP64_TOL, name
        assert np.abs(plain - ref).max() <= FP64_TOL, name
    tri, Qc = cases["c1"]
    Qn = Qc[:300].copy()
    Qn[7] = np.nan
    Qn[8, 1] = np.inf
    for precision in (0, 2):
        out = host_call(tri, Qn, precision)
        assert np.isnan(out[7]) and np.isnan(out[8]), precision
        keep = np.ones(len(Qn), bool)
        keep[[7, 8]] = False
        assert np.abs(out[keep] - c1[4][:300][keep]).max() <= FP64_TOL


@pytest.mark.parametrize("qpl,splits,sort", [(1, 1, 0), (2, 1, 0), (4, 1, 0), (2, 9, 1), (0, 0, -1)])
def test_fp32c_config1(require_gpu, c1, qpl, splits, sort):
    V, F, Q, tri, ref = c1
    out = host_call(tri, Q, 1, qpl, splits, sort)
    err = np.abs(out - ref).max()
    print(f"fp32c c1 qpl={qpl} splits={splits} sort={sort} max|dw|={err:.3e}", flush=True)
    assert err <= FP32C_TOL


def test_fp32c_config2_1024(require_gpu, c2):
    V, F, Q, tri, ref = c2
    for sort in (0, 1):
        out = host_call(tri, Q, 1, 0, 0, sort)
        err = np.abs(out - ref).max()
        print(f"fp32c c2@1024 sort={sort} max|dw|={err:.3e} mean={np.abs(out - ref).mean():.3e}", flush=True)
        assert err <= FP32C_TOL


def test_fp32c_all_pairs_fp64_fallback(require_gpu, c1):
    """kappa = 1e30 sends every pair through the FP64 re-evaluation; the
    result is then the FP64 formula on the block-rounded mesh."""
    V, F, Q, tri, ref = c1
    out = host_call(tri, Q, 1, 2, 1, 0, 1e30)
    assert np.abs(out - ref).max() <= FP32C_TOL


def test_fp32c_pure_fp32_is_worse_near_surface(require_gpu):
    """kappa = 0 disables the FP64 re-evaluation. On queries 1e-6 above the
    faces of a grid the pure-FP32 result must be no better than the default."""
    V, F = W.grid(12)
    rng = np.random.default_rng(5)
    f = rng.integers(0, len(F), 512)
    bary = rng.dirichlet(np.ones(3), 512)
    T = V[F[f]]
    p = np.einsum("nk,nkd->nd", bary, T)
    n = np.cross(T[:, 1] - T[:, 0], T[:, 2] - T[:, 0])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    Q = p + 1e-6 * n
    tri = oracle.pack_soa(V, F)
    ref = oracle.winding_soa(tri, Q)
    err_default = np.abs(host_call(tri, Q, 1, 1, 1, 0) - ref).max()
    err_fp32 = np.abs(host_call(tri, Q, 1, 1, 1, 0, 0.0) - ref).max()
    print(f"near-surface: default {err_default:.3e}, pure fp32 {err_fp32:.3e}", flush=True)
    assert err_default <= 1e-4
    assert err_fp32 >= err_default


@pytest.mark.parametrize("qpl", [1, 2])
def test_face_records_near_surface(require_gpu, qpl):
    """The face-record kernels on an open cap (the 2,000 highest faces of a
    displaced geodesic sphere) with 100,000 queries at 1e-6 .. 1e-1 from its
    faces and 62,144 uniform in its box: many groups of twelve faces fail
    their test and go through the sixes, triples, single faces and the FP64
    re-evaluation, and the open rim takes the boundary correction. Against
    the plain FP64 kernel: the FP32-compensated kernel within 1e-6 where a
    query is at least 1e-3 from the surface and within 1e-4 (the near-surface
    bar) closer, the FP64 face kernel within 1e-10; both bit for bit the same
    over three runs."""
    V, F = W.config5()[:2]
    V = V.astype(np.float64)
    cz = V[F].mean(axis=1)[:, 2]
    Fs = F[np.sort(np.argsort(-cz)[:2000])]
    rng = np.random.default_rng(78)
    T = V[Fs[rng.integers(0, len(Fs), 100_000)]]
    bary = rng.dirichlet(np.ones(3), size=len(T))
    p = np.einsum("ij,ijk->ik", bary, T)
    n = np.cross(T[:, 1] - T[:, 0], T[:, 2] - T[:, 0])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    d = 10.0 ** rng.uniform(-6, -1, size=len(T)) * rng.choice([-1.0, 1.0], size=len(T))
    box = V[np.unique(Fs)]
    Q = np.vstack([p + d[:, None] * n, rng.uniform(box.min(0), box.max(0), size=(62_144, 3))])
    tri = oracle.pack_soa(V, Fs)
    ref = host_call(tri, Q, 2)
    dist = np.sqrt(igl.point_mesh_squared_distance(Q, V, Fs)[0])
    far = dist >= 1e-3
    assert far.sum() > 50_000 and (~far).sum() > 50_000
    w32 = [host_call(tri, Q, 1, qpl) for _ in range(3)]
    w64 = [host_call(tri, Q, 0, qpl) for _ in range(3)]
    err = np.abs(w32[0] - ref)
    print(f"qpl {qpl}: fp32c max |dw| {err[far].max():.3e} (>= 1e-3 from the surface), "
          f"{err[~far].max():.3e} (closer); fp64 face {np.abs(w64[0] - ref).max():.3e}", flush=True)
    assert err[far].max() <= FP32C_TOL
    assert err[~far].max() <= 1e-4
    assert np.abs(w64[0] - ref).max() <= FP64_TOL
    for runs in (w32, w64):
        for w in runs[1:]:
            assert np.array_equal(w, runs[0]), int((w != runs[0]).sum())


@pytest.mark.parametrize("nf_take,nq", [(1, 1), (5, 7), (257, 255), (300, 513), (5120, 1)])
def test_ragged_sizes(require_gpu, c1, nf_take, nq):
    """Odd face and query counts (tails of the LDS tile and of the query
    block). Prefixes of the face list are OPEN patches: there w is not
    locally constant and the FP32 rounding of the block-local coordinates
    (~6e-8 * block extent, here the whole [-1.5, 1.5]^3 box) moves the
    boundary, which the boundary correction undoes; the FP32-compensated bar
    is 1e-6 as for closed meshes."""
    V, F, Q, tri, ref = c1
    Fs = F[:nf_take]
    Qs = Q[:nq]
    t = oracle.pack_soa(V, Fs)
    r = oracle.winding_soa(t, Qs)
    tol32 = FP32C_TOL
    for precision, tol in ((0, FP64_TOL), (1, tol32), (2, FP64_TOL)):
        for splits in (1, 3):
            out = host_call(t, Qs, precision, 1, splits, 0)
            assert np.abs(out - r).max() <= tol, (precision, splits)


def test_empty_mesh_and_queries(require_gpu):
    tri = np.zeros((9, 0))
    Q = np.random.default_rng(0).normal(size=(10, 3))
    for precision in (0, 1, 2):
        assert np.array_equal(host_call(tri, Q, precision), np.zeros(10))
    V, F = W.tetrahedron()
    assert igl.winding_number(V, F, np.zeros((0, 3)), device="gpu").shape == (0,)


def test_deterministic_run_to_run(require_gpu, c2):
    V, F, Q, tri, ref = c2
    for precision in (0, 1, 2):
        a = host_call(tri, Q, precision, 0, 16)
        b = host_call(tri, Q, precision, 0, 16)
        assert np.array_equal(a, b)


def test_device_api_matches_host_api(require_gpu, c1):
    """wn_pack_soa + wn_forward on device tensors (the path torch callers and
    the bench use) equals the host-buffer path."""
    import torch

    V, F, Q, tri, ref = c1
    h = igl.lib()
    dev = torch.device("cuda")
    Vd = torch.tensor(V, device=dev)
    Fd = torch.tensor(F.astype(np.int32), device=dev)
    Qd = torch.tensor(Q, device=dev)
    trid = torch.empty((9, len(F)), dtype=torch.float64, device=dev)
    bad = torch.full((1,), 2**31 - 1, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    assert h.wn_pack_soa(Vd.data_ptr(), Fd.data_ptr(), len(V), len(F), 0, 1, trid.data_ptr(),
                         bad.data_ptr(), stream) == 0
    torch.cuda.synchronize()
    assert int(bad.item()) == 2**31 - 1
    assert np.array_equal(trid.cpu().numpy(), tri)
    for precision in (0, 1, 2):
        nbytes = h.wn_forward_workspace_bytes(len(F), len(Q), precision, 0, 0, -1)
        assert nbytes >= 0
        ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=dev)
        out = torch.empty(len(Q), dtype=torch.float64, device=dev)
        st = h.wn_forward(trid.data_ptr(), Qd.data_ptr(), out.data_ptr(), len(F), len(Q),
                          precision, 0, 0, -1, -1.0, ws.data_ptr(), nbytes, stream)
        assert st == 0, igl._lib.last_error()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out.cpu().numpy(), host_call(tri, Q, precision))


def test_pack_soa_reports_bad_index(require_gpu):
    import torch

    h = igl.lib()
    dev = torch.device("cuda")
    V = torch.zeros((4, 3), dtype=torch.float32, device=dev)
    F = torch.tensor([[0, 1, 2], [0, 1, 9], [0, 7, 1]], dtype=torch.int64, device=dev)
    tri = torch.empty((9, 3), dtype=torch.float64, device=dev)
    bad = torch.full((1,), 2**31 - 1, dtype=torch.int32, device=dev)
    assert h.wn_pack_soa(V.data_ptr(), F.data_ptr(), 4, 3, 1, 0, tri.data_ptr(), bad.data_ptr(),
                         torch.cuda.current_stream().cuda_stream) == 0
    torch.cuda.synchronize()
    assert int(bad.item()) == 1


def test_injected_fault_raises(require_gpu, monkeypatch):
    """Reference tests/test_parity.py:117-128: a failed GPU kernel raises
    RuntimeError; here the failure is injected in the host code (WN_FAULT)."""
    V, F = W.tetrahedron()
    monkeypatch.setenv("WN_FAULT", "launch")
    with pytest.raises(RuntimeError, match="GPU"):
        igl.winding_number(V, F, [[0.1, 0.1, 0.1]], device="gpu")
    monkeypatch.delenv("WN_FAULT")
    assert igl.winding_number(V, F, [[0.1, 0.1, 0.1]], device="gpu")[0] == pytest.approx(1.0)


def test_config2_full_size_properties(require_gpu):
    """28,320 faces x 1,048,576 queries: FP32-compensated and FP64 against
    FP64 plain on the GPU (size-independent properties: agreement <= 1e-6 and
    1e-13 and, on the closed torus, integer values), plus the oracle on a
    strided subsample."""
    V, F, Q = W.config2(1_048_576)
    tri = oracle.pack_soa(V, F)
    plain = host_call(tri, Q, 2)
    w64 = host_call(tri, Q, 0)
    w32 = host_call(tri, Q, 1)
    d = np.abs(w32 - plain)
    print(f"c2@1M fp32c vs fp64 plain: max {d.max():.3e} mean {d.mean():.3e}; fp64 vs fp64 "
          f"plain {np.abs(w64 - plain).max():.3e}", flush=True)
    assert d.max() <= FP32C_TOL
    assert np.abs(w64 - plain).max() <= FAST_VS_PLAIN_TOL
    assert np.abs(w64 - np.round(w64)).max() <= 1e-9
    sub = np.arange(0, len(Q), 4099)
    ref = oracle.winding_soa(tri, Q[sub])
    assert np.abs(w64[sub] - ref).max() <= FP64_TOL


def test_host_entry_is_safe_under_concurrent_callers(require_gpu):
    """mli_winding_gpu from 8 Python threads at once (the reference's callers
    fan out over threads, python/igl/_lib.py:72-95): every call succeeds and
    returns the single-threaded result, bit for bit."""
    import threading

    from igl import pack_faces
    from igl._lib import addr, lib

    V, F, Q = W.config1()
    tri = pack_faces(V.astype(np.float64), F)
    Qs = [np.ascontiguousarray(Q[i::8]) for i in range(8)]
    ref = []
    for q in Qs:
        out = np.empty(len(q))
        assert lib().mli_winding_gpu(addr(tri), addr(q), addr(out), tri.shape[1], len(q)) == 1
        ref.append(out)
    got = [None] * 8
    errors = []

    def run(i):
        try:
            for _ in range(5):
                out = np.empty(len(Qs[i]))
                ok = lib().mli_winding_gpu(addr(tri), addr(Qs[i]), addr(out), tri.shape[1], len(Qs[i]))
                if ok != 1 or not np.array_equal(out, ref[i]):
                    errors.append(i)
            got[i] = True
        except Exception as e:  # pragma: no cover - reported below
            errors.append(repr(e))

    threads = [threading.Thread(target=run, args=(i,)) for i in range(8)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors and all(got)


def test_strip_route_under_concurrent_callers(require_gpu, monkeypatch):
    """wn_winding_gpu_mesh (the strip route of igl.winding_number) from 6
    threads at once, each on its own mesh and alternating FP64 and FP32c:
    the host entries run on a pool of contexts (own stream, buffers and kept
    strip mesh each), and every call returns the single-threaded result, bit
    for bit, whichever context it lands on."""
    import threading

    monkeypatch.setattr(igl, "STRIP_MIN_QUERIES", {0: 1000, 1: 1000})
    V0, F0, Q = W.config2(4096)
    meshes = [(V0 * (1.0 + 0.1 * i) + [0.05 * i, 0.0, -0.02 * i], F0) for i in range(6)]
    expect = {}
    for i, (V, F) in enumerate(meshes):
        for precision in ("fp64", "fp32c"):
            assert igl.lib().wn_release_host_buffers() == 0
            expect[i, precision] = igl.winding_number(V, F, Q, device="gpu", precision=precision)
    errors = []

    def run(i):
        V, F = meshes[i]
        try:
            for r in range(6):
                precision = ("fp64", "fp32c")[(i + r) % 2]
                w = igl.winding_number(V, F, Q, device="gpu", precision=precision)
                if not np.array_equal(w, expect[i, precision]):
                    errors.append((i, r, precision))
        except Exception as e:  # pragma: no cover - reported below
            errors.append(repr(e))

    threads = [threading.Thread(target=run, args=(i,)) for i in range(6)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors


def test_large_calls_use_strip_kernels(require_gpu, monkeypatch):
    """igl.winding_number sends calls with many queries through the strip
    kernels (wn_winding_gpu_mesh): FP64 agrees with the plain FP64 kernel to
    1e-12 and with the oracle to the FP64 bar; FP32c stays within its bar;
    the input contract is unchanged."""
    from igl import pack_faces
    from igl._lib import addr, lib

    V, F, Q = W.config2(4096)
    monkeypatch.setattr(igl, "STRIP_MIN_QUERIES", {0: 1000, 1: 1000})
    w64 = igl.winding_number(V, F, Q, device="gpu")
    w32 = igl.winding_number(V, F, Q, device="gpu", precision="fp32c")
    tri = pack_faces(V, F)
    plain = np.empty(len(Q))
    assert lib().mli_winding_gpu(addr(tri), addr(Q), addr(plain), tri.shape[1], len(Q)) == 1
    print(f"strip route: fp64 vs plain {np.abs(w64 - plain).max():.3e}, "
          f"fp32c vs plain {np.abs(w32 - plain).max():.3e}", flush=True)
    assert np.abs(w64 - plain).max() <= 1e-12
    assert np.abs(w32 - plain).max() <= FP32C_TOL
    sub = np.arange(0, len(Q), 17)
    assert np.abs(w64[sub] - oracle.winding(V, F, Q[sub])).max() <= FP64_TOL
    with pytest.raises(IndexError):
        igl.winding_number(V, F + len(V), Q, device="gpu")
    F32 = F.astype(np.int32)
    assert np.array_equal(igl.winding_number(V, F32, Q, device="gpu"), w64)


def test_reference_entry_large_calls_use_fp64_strips(require_gpu):
    """mli_winding_gpu (the reference's ABI entry, face SoA only) with >=
    262,144 queries finds the faces' shared corners by coordinate equality and
    runs the FP64 strip kernel: within 1e-12 of the FP64 face kernels
    (wn_winding_gpu_ex, precisions 0 and 2) and within the FP64 bar of the
    oracle; a
    repeated call (kept strip mesh) gives the same bits; a non-manifold soup
    (repeated and reversed faces, coordinate-duplicated vertices) too."""
    from igl import pack_faces
    from igl._lib import addr, lib

    h = lib()
    V, F, Q = W.config2(1 << 20)
    Q = np.ascontiguousarray(Q[:270_000])
    soup_V, soup_F = W.icosphere(2)
    soup_V = np.vstack([soup_V, soup_V[:40]])  # same coordinates, other ids
    soup_F = np.vstack([soup_F, soup_F[:30, ::-1], soup_F[5:9]])
    soup_F[::7, 0] = np.where(soup_F[::7, 0] < 40, soup_F[::7, 0] + len(soup_V) - 40,
                              soup_F[::7, 0])
    soup_Q = np.ascontiguousarray(np.random.default_rng(9).uniform(-1.5, 1.5, size=(262_144, 3)))
    for name, (tri, Qc) in {"torus": (pack_faces(V, F), Q),
                            "soup": (pack_faces(soup_V, soup_F), soup_Q)}.items():
        face = np.full(len(Qc), np.nan)
        assert h.wn_winding_gpu_ex(addr(tri), addr(Qc), addr(face), tri.shape[1], len(Qc), 0, 0,
                                   0, 0, -1.0) == 0, igl._lib.last_error()
        plain = np.full(len(Qc), np.nan)
        assert h.wn_winding_gpu_ex(addr(tri), addr(Qc), addr(plain), tri.shape[1], len(Qc), 2, 0,
                                   0, 0, -1.0) == 0, igl._lib.last_error()
        out = np.full(len(Qc), np.nan)
        assert h.mli_winding_gpu(addr(tri), addr(Qc), addr(out), tri.shape[1], len(Qc)) == 1
        again = np.full(len(Qc), np.nan)
        assert h.mli_winding_gpu(addr(tri), addr(Qc), addr(again), tri.shape[1], len(Qc)) == 1
        sub = np.arange(0, len(Qc), 331)
        ref = oracle.winding_soa(tri, Qc[sub])
        print(f"mli_winding_gpu strip route, {name}: vs face kernel "
              f"{np.abs(out - face).max():.3e}, vs oracle {np.abs(out[sub] - ref).max():.3e}",
              flush=True)
        assert np.array_equal(out, again)
        assert np.abs(out - face).max() <= 1e-12
        assert np.abs(out - plain).max() <= 1e-12
        assert np.abs(out[sub] - ref).max() <= FP64_TOL


def test_vf_entry_matches_host_packing(require_gpu, c1, c2):
    """wn_winding_gpu_vf (V and F uploaded, face SoA gathered on the device;
    the path of igl.winding_number below the strip threshold) equals the
    host-packed entries bit for bit: mli_winding_gpu for FP64,
    wn_winding_gpu_ex with default tuning for FP32c; int32 faces give the same
    bits; an empty F gives zeros; and igl.winding_number returns the same."""
    from igl._lib import addr

    h = igl.lib()
    for V, F, Q, tri, ref in (c1, c2):
        V = np.ascontiguousarray(V, dtype=np.float64)
        tri = np.ascontiguousarray(tri)
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        for precision in (0, 1):
            out = np.full(len(Q), np.nan)
            for fdtype, Fc in ((0, np.ascontiguousarray(F, dtype=np.int64)),
                               (1, np.ascontiguousarray(F, dtype=np.int32))):
                assert h.wn_winding_gpu_vf(addr(V), addr(Fc), len(V), len(Fc), fdtype, addr(Q),
                                           addr(out), len(Q), precision) == 0, igl._lib.last_error()
                if precision == 0:
                    packed = np.empty(len(Q))
                    assert h.mli_winding_gpu(addr(tri), addr(Q), addr(packed), tri.shape[1],
                                             len(Q)) == 1
                else:
                    packed = host_call(tri, Q, 1)
                assert np.array_equal(out, packed)
            bar = FP64_TOL if precision == 0 else FP32C_TOL
            assert np.abs(out - ref).max() <= bar
            w = igl.winding_number(V, F, Q, device="gpu",
                                   precision="fp64" if precision == 0 else "fp32c")
            assert np.array_equal(w, out)
    V, F, Q, tri, ref = c1
    V = np.ascontiguousarray(V, dtype=np.float64)
    F0 = np.zeros((0, 3), dtype=np.int64)
    out = np.full(len(Q), np.nan)
    assert h.wn_winding_gpu_vf(addr(V), addr(F0), len(V), 0, 0, addr(Q), addr(out), len(Q), 0) == 0
    assert np.array_equal(out, np.zeros(len(Q)))


def test_strip_route_keeps_the_last_mesh(require_gpu, monkeypatch):
    """wn_winding_gpu_mesh keeps the last strip mesh on the device: repeated
    calls (hit), the same F with new V (records regathered on the device), a
    different F, int32 faces and meshes interleaved all give the bits of a
    freshly built mesh (wn_release_host_buffers before the call), in FP64
    and FP32c."""
    from igl._lib import addr

    h = igl.lib()
    V, F, Q = W.config2(4096)
    V1, F1, _ = W.config1()
    monkeypatch.setattr(igl, "STRIP_MIN_QUERIES", {0: 1000, 1: 1000})

    def fresh(Vm, Fm, precision):
        assert h.wn_release_host_buffers() == 0
        return igl.winding_number(Vm, Fm, Q, device="gpu", precision=precision)

    cases = [(V, F), (V, F), (V * 1.25 + [0.1, 0.0, -0.2], F), (V1, F1), (V, F),
             (V, F.astype(np.int32)), (V1 + [0.05, 0.05, 0.05], F1), (V, F)]
    for precision in ("fp64", "fp32c"):
        expect = [fresh(Vm, Fm, precision) for Vm, Fm in cases]
        assert h.wn_release_host_buffers() == 0
        for (Vm, Fm), e in zip(cases, expect):
            w = igl.winding_number(Vm, Fm, Q, device="gpu", precision=precision)
            assert np.array_equal(w, e)
    # the moved mesh agrees with the plain FP64 kernel (the regathered records are right)
    Vm = V * 1.25 + [0.1, 0.0, -0.2]
    igl.winding_number(V, F, Q, device="gpu")
    w = igl.winding_number(Vm, F, Q, device="gpu")
    tri = igl.pack_faces(np.asarray(Vm, dtype=np.float64), F)
    plain = np.empty(len(Q))
    assert h.mli_winding_gpu(addr(tri), addr(Q), addr(plain), tri.shape[1], len(Q)) == 1
    assert np.abs(w - plain).max() <= 1e-12
    assert h.wn_release_host_buffers() == 0
