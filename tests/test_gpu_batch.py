"""Batched persistent forward (wn_batch_*) against the oracle and against
per-mesh launches. Bars as for the single-mesh forward: FP64 within 1e-10 of
the oracle; FP32-compensated within 1e-6 on closed and open meshes alike (the
open cylinders of the workload need the boundary correction: face plans find
their boundary edges themselves, strip plans are given them)."""

import numpy as np
import pytest

import oracle
import wn_workloads as W

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def small_batch():
    B = W.config4(n_meshes=24, seed=7, max_faces=6000, max_queries=1500)
    # edge cases: a mesh without queries and a mesh without faces
    V, F, Q = B[3]
    B[3] = (V, F, Q[:0])
    V, F, Q = B[5]
    B[5] = (V, F[:0], Q)
    return B


def _run(B, precision, use_graph=False, **cfg):
    from igl.batch import winding_number_batch

    return winding_number_batch([b[0] for b in B], [b[1] for b in B], [b[2] for b in B],
                                precision=precision, use_graph=use_graph, **cfg)


def _closed(F, nv):
    # every edge shared by exactly two faces
    e = np.sort(np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]]), axis=1)
    _, counts = np.unique(e, axis=0, return_counts=True)
    return len(F) > 0 and (counts == 2).all()


@pytest.mark.parametrize("qpl,chunk", [(1, 256), (1, 2048), (2, 1024)])
def test_batch_matches_oracle(require_gpu, small_batch, qpl, chunk):
    w64 = _run(small_batch, "fp64", qpl=qpl, chunk_faces=chunk)
    w32 = _run(small_batch, "fp32c", qpl=qpl, chunk_faces=chunk)
    worst64 = worst32 = 0.0
    for (V, F, Q), a, b in zip(small_batch, w64, w32):
        assert a.shape == (len(Q),)
        if len(Q) == 0:
            continue
        ref = oracle.winding(V.astype(np.float64), F, Q)
        e64 = np.abs(a - ref).max()
        e32 = np.abs(b - ref).max()
        worst64, worst32 = max(worst64, e64), max(worst32, e32)
        assert e64 <= 1e-10
        assert e32 <= 1e-6
    print(f"batch qpl={qpl} chunk={chunk}: worst fp64 {worst64:.2e} fp32c {worst32:.2e}", flush=True)


def test_graph_replays_match_eager(require_gpu, small_batch):
    from igl.batch import WindingBatch, pack_batch

    B = small_batch
    tri = pack_batch([b[0] for b in B], [b[1] for b in B])
    Q = torch.tensor(np.concatenate([b[2] for b in B]), device="cuda")
    plan = WindingBatch([len(b[1]) for b in B], [len(b[2]) for b in B], precision="fp32c",
                        chunk_faces=512)
    try:
        eager = plan.run(tri, Q, use_graph=False).clone()
        for i in range(4):
            # unrelated eager work and a sync between replays
            torch.randn(1000, device="cuda").sum().item()
            out = plan.run(tri, Q, use_graph=True)
            torch.cuda.synchronize()
            assert torch.equal(out, eager), f"replay {i}"
        info = plan.info()
        assert info["items"] > 0 and info["partials"] > 0
    finally:
        plan.close()


def test_batch_equals_single_mesh_launches(require_gpu, small_batch):
    import igl

    w = _run(small_batch, "fp64")
    for (V, F, Q), a in zip(small_batch, w):
        if len(Q) == 0:
            continue
        single = igl.winding_number(V, F, Q, device="gpu")
        assert np.abs(single - a).max() <= 1e-12


def test_invalid_plans_are_rejected(require_gpu):
    from igl.batch import WindingBatch

    with pytest.raises(RuntimeError):
        WindingBatch([10, 20], [5, 5], chunk_faces=100)
    with pytest.raises(RuntimeError):
        WindingBatch([10, 20], [5, 5], qpl=4)
    with pytest.raises(ValueError):
        WindingBatch([10, 20], [5])


@pytest.fixture(scope="module")
def c4_full():
    """512 meshes (6.5M faces, 5.0M queries, 7.3e10 pairs) and their FP64
    face-plan result."""
    B = W.config4()
    print("C4 full size: FP64 face plans", flush=True)
    return B, _run(B, "fp64")


def _oracle_subsample(B, results, bar_closed, bar_open, every=64, k=32):
    """Every `every`-th mesh, k queries: max error against the oracle."""
    rng = np.random.default_rng(0)
    worst = 0.0
    for i in range(0, len(B), every):
        V, F, Q = B[i]
        if len(Q) == 0:
            continue
        sub = rng.choice(len(Q), min(k, len(Q)), replace=False)
        ref = oracle.winding(V.astype(np.float64), F, Q[sub])
        for w in results:
            err = np.abs(w[i][sub] - ref).max()
            worst = max(worst, err)
            assert err <= (bar_closed if _closed(F, len(V)) else bar_open), i
    return worst


def _split_errors(B, a, b):
    closed_err = open_err = 0.0
    for (V, F, Q), x, y in zip(B, a, b):
        d = np.abs(x - y).max() if len(Q) else 0.0
        if _closed(F, len(V)):
            closed_err = max(closed_err, d)
        else:
            open_err = max(open_err, d)
    return closed_err, open_err


def test_config4_full_size(require_gpu, c4_full):
    """Face-record plans: FP32-compensated against FP64 on the GPU, FP64
    against the oracle on a subsample."""
    B, w64 = c4_full
    w32 = _run(B, "fp32c")
    closed_err, open_err = _split_errors(B, w64, w32)
    worst = _oracle_subsample(B, [w64], 1e-10, 1e-10)
    print(f"C4 face plans fp32c vs fp64: closed meshes max {closed_err:.2e}, open meshes max "
          f"{open_err:.2e}; fp64 vs oracle {worst:.2e}", flush=True)
    assert closed_err <= 1e-6
    assert open_err <= 1e-6


def test_config4_full_size_strip_plans(require_gpu, c4_full):
    """The batch headline path at full size: strip plans in FP32c and FP64
    against the FP64 face plans (FP64 strips <= 1e-12; FP32c at the face
    plans' bars) and against the oracle on a per-mesh subsample; the FP32c
    strip plan's graph replays equal its eager run."""
    from igl.batch import StripBatch, pack_batch_strips

    B, w64 = c4_full
    s64 = _run(B, "fp64", strips=True)
    worst64 = max(np.abs(a - b).max() for a, b, (_, _, Q) in zip(s64, w64, B) if len(Q))
    rec, counts, boundary = pack_batch_strips([b[0] for b in B], [b[1] for b in B],
                                              boundary=True)
    Q = torch.tensor(np.concatenate([b[2] for b in B]), device="cuda")
    qc = [len(b[2]) for b in B]
    plan = StripBatch(counts, qc, boundary=boundary)
    try:
        eager = plan.run(rec, Q, use_graph=False).clone()
        for i in range(3):
            torch.randn(1000, device="cuda").sum().item()
            out = plan.run(rec, Q, use_graph=True)
            torch.cuda.synchronize()
            assert torch.equal(out, eager), f"replay {i}"
    finally:
        plan.close()
    w = eager.cpu().numpy()
    off = np.concatenate([[0], np.cumsum(qc)])
    s32 = [w[off[k]:off[k + 1]] for k in range(len(B))]
    closed_err, open_err = _split_errors(B, w64, s32)
    worst = _oracle_subsample(B, [s64], 1e-10, 1e-10)
    worst32 = _oracle_subsample(B, [s32], 1e-6, 1e-6)
    print(f"C4 strip plans: fp64 strips vs fp64 faces {worst64:.2e}; fp32c strips vs fp64 faces: "
          f"closed max {closed_err:.2e}, open max {open_err:.2e}; vs oracle fp64 {worst:.2e}, "
          f"fp32c {worst32:.2e}", flush=True)
    assert worst64 <= 1e-12
    assert closed_err <= 1e-6
    assert open_err <= 1e-6


def test_tuned_strip_plan_reproduces_the_default(require_gpu, small_batch, tmp_path,
                                                 monkeypatch):
    """The autotuner's winner (timed on this device, cached, then read back)
    gives a plan whose result equals the default plan's to the FP32c bar,
    and a cache hit returns the same config without timing again."""
    from igl import _autotune
    from igl.batch import StripBatch, pack_batch_strips

    monkeypatch.setenv("WN_TUNE_CACHE", str(tmp_path / "tune.json"))
    B = small_batch
    rec, counts, bnd = pack_batch_strips([b[0] for b in B], [b[1] for b in B], boundary=True)
    Q = torch.tensor(np.concatenate([b[2] for b in B]), device="cuda")
    qc = [len(b[2]) for b in B]
    best = _autotune.tune_batch_strips(counts, qc, rec, Q, chunks=(1, 2, 8), boundary=bnd)
    assert best in (1, 2, 8)
    calls = []
    monkeypatch.setattr(_autotune, "_time", lambda fn, r: calls.append(1) or 0.0)
    assert _autotune.tune_batch_strips(counts, qc, rec, Q, chunks=(1, 2, 8), boundary=bnd) == best
    assert not calls
    plans = [StripBatch(counts, qc, chunk_tiles=best, boundary=bnd),
             StripBatch(counts, qc, boundary=bnd)]
    try:
        tuned, default = (p.run(rec, Q, use_graph=False).cpu().numpy() for p in plans)
    finally:
        for p in plans:
            p.close()
    assert np.abs(tuned - default).max() <= 1e-6


@pytest.mark.parametrize("chunk_tiles", [1, 8, 0])
def test_strip_batch_matches_oracle(require_gpu, small_batch, chunk_tiles):
    """Strip plans (records per mesh, FP32-compensated) against the oracle,
    with many chunks per mesh (chunk_tiles=1), a few, and the default."""
    w = _run(small_batch, "fp32c", strips=True, chunk_tiles=chunk_tiles)
    worst = 0.0
    for (V, F, Q), a in zip(small_batch, w):
        assert a.shape == (len(Q),)
        if len(Q) == 0:
            continue
        ref = oracle.winding(V.astype(np.float64), F, Q)
        err = np.abs(a - ref).max()
        worst = max(worst, err)
        assert err <= 1e-6
    print(f"strip batch chunk_tiles={chunk_tiles}: worst {worst:.2e}", flush=True)


def test_strip_batch_graph_replays_match_eager(require_gpu, small_batch):
    """Graph replays of a strip plan with its boundary correction (part of
    the captured sequence) equal the eager run; setting the boundary drops
    the captured graph, and a plan without it differs from one with it only
    on the open meshes."""
    from igl.batch import StripBatch, pack_batch_strips

    B = small_batch
    rec, counts, bnd = pack_batch_strips([b[0] for b in B], [b[1] for b in B], boundary=True)
    assert counts[5] == 0  # the mesh without faces has no records
    assert bnd.edge_counts[5] == 0 and bnd.n_edges > 0
    Q = torch.tensor(np.concatenate([b[2] for b in B]), device="cuda")
    qc = [len(b[2]) for b in B]
    plan = StripBatch(counts, qc, chunk_tiles=2)
    try:
        bare = plan.run(rec, Q, use_graph=True).clone()  # captured without edges
        plan.set_boundary(bnd)
        eager = plan.run(rec, Q, use_graph=False).clone()
        off = np.concatenate([[0], np.cumsum(qc)])
        diff = (eager - bare).abs().cpu().numpy()
        for k in range(len(B)):
            if bnd.edge_counts[k] == 0:
                assert diff[off[k]:off[k + 1]].max(initial=0.0) == 0.0, k
        assert diff.max() > 0.0
        for i in range(4):
            torch.randn(1000, device="cuda").sum().item()
            out = plan.run(rec, Q, use_graph=True)
            torch.cuda.synchronize()
            assert torch.equal(out, eager), f"replay {i}"
        info = plan.info()
        assert info["items"] > 0 and info["partials"] > 0
        assert info["faces"] == int(counts.sum()) // 255
    finally:
        plan.close()


def test_strip_batch_equals_single_mesh_strip_launches(require_gpu, small_batch):
    """Same records, same query blocks (both unsorted), same face order: a
    one-chunk strip plan reproduces igl.Mesh without splits."""
    from igl.mesh import Mesh

    w = _run(small_batch, "fp32c", strips=True, chunk_tiles=1 << 20, sort=0)
    for (V, F, Q), a in zip(small_batch, w):
        if len(Q) == 0 or len(F) == 0:
            continue
        m = Mesh(V, F)
        try:
            single = m.winding_number(Q, splits=1, sort=0)
        finally:
            m.close()
        assert np.abs(single - a).max() <= 1e-12


def test_batch_topology_device_records(require_gpu, small_batch):
    """BatchTopology gathers on the device exactly the records the host
    packer builds, and again for moved vertices (same topology)."""
    from igl.batch import BatchTopology, StripBatch, pack_batch_strips

    B = small_batch
    topo = BatchTopology([b[1] for b in B], [len(b[0]) for b in B])
    V_all = torch.tensor(np.concatenate([b[0] for b in B]), device="cuda")
    host_rec, counts = pack_batch_strips([b[0] for b in B], [b[1] for b in B])
    np.testing.assert_array_equal(topo.record_counts, counts)
    assert torch.equal(topo.records(V_all), host_rec)
    moved = [b[0] * 1.1 + 0.05 for b in B]
    host2, _ = pack_batch_strips(moved, [b[1] for b in B])
    rec2 = topo.records(torch.tensor(np.concatenate(moved), device="cuda"))
    assert torch.equal(rec2, host2)
    _, _, host_bnd = pack_batch_strips(moved, [b[1] for b in B], boundary=True)
    assert torch.equal(topo.boundary.pairs, host_bnd.pairs)
    assert torch.equal(topo.boundary.weight, host_bnd.weight)
    Q = torch.tensor(np.concatenate([b[2] for b in B]), device="cuda")
    plan = StripBatch(counts, [len(b[2]) for b in B], boundary=topo.boundary)
    try:
        assert torch.equal(plan.run(rec2, Q, use_graph=False), plan.run(host2, Q, use_graph=False))
    finally:
        plan.close()


@pytest.mark.parametrize("chunk_tiles", [1, 3, 0])
def test_strip64_batch_matches_oracle(require_gpu, small_batch, chunk_tiles):
    """FP64 strip plans against the oracle at the FP64 bar."""
    w = _run(small_batch, "fp64", strips=True, chunk_tiles=chunk_tiles)
    worst = 0.0
    for (V, F, Q), a in zip(small_batch, w):
        assert a.shape == (len(Q),)
        if len(Q):
            worst = max(worst, np.abs(a - oracle.winding(V.astype(np.float64), F, Q)).max())
    print(f"fp64 strip batch chunk_tiles={chunk_tiles}: worst {worst:.2e}", flush=True)
    assert worst <= 1e-10


def test_strip64_batch_graph_and_single_mesh(require_gpu, small_batch):
    """FP64 strip plan: graph replays equal the eager run, and a one-chunk
    unsorted plan equals single-mesh FP64 strip launches (same records, same
    query blocks, same face order) bit for bit."""
    from igl.batch import StripBatch, pack_batch_strips
    from igl.mesh import Mesh

    B = small_batch
    rec, counts = pack_batch_strips([b[0] for b in B], [b[1] for b in B])
    Q = torch.tensor(np.concatenate([b[2] for b in B]), device="cuda")
    qc = [len(b[2]) for b in B]
    plan = StripBatch(counts, qc, chunk_tiles=2, precision="fp64")
    try:
        eager = plan.run(rec, Q, use_graph=False).clone()
        for i in range(3):
            torch.randn(1000, device="cuda").sum().item()
            out = plan.run(rec, Q, use_graph=True)
            torch.cuda.synchronize()
            assert torch.equal(out, eager), f"replay {i}"
    finally:
        plan.close()
    one = StripBatch(counts, qc, chunk_tiles=1 << 20, sort=0, precision="fp64")
    try:
        w = one.run(rec, Q, use_graph=False).cpu().numpy()
    finally:
        one.close()
    off = np.concatenate([[0], np.cumsum(qc)])
    for k, (V, F, Qk) in enumerate(B):
        if len(Qk) == 0 or len(F) == 0:
            continue
        m = Mesh(V.astype(np.float64), F)
        try:
            single = m.winding_number(Qk, splits=1, sort=0, precision="fp64")
        finally:
            m.close()
        np.testing.assert_array_equal(w[off[k]:off[k + 1]], single)
