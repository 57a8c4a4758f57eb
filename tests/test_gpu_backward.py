"""Adjoint kernel (wn_backward + wn_scatter_corners) through the autograd layer.

Loss L = sum_q (w_q - t_q)^2 (the soft-occupancy loss of the differentiation
workload). Checks, all FP64:
  * against the FP64 PyTorch autograd gradient of the reference formula
    (oracle.winding_torch, src/capi.mojo:218-235,286) on an open 2,000-face
    subset: relative error <= 1e-9;
  * against FP64 central differences (h = 1e-6) of the FP64 GPU forward on
    256 random coordinates: |fd - an| <= 1e-5 * (1 + |an|);
  * on the closed 50,000-face sphere (w locally constant, true gradient 0):
    forward + backward at full size (262,144 queries), gradients finite and
    matching central differences, |fd - an| <= 1e-5 * (1 + |an|).
"""

import numpy as np
import pytest

import oracle
import wn_workloads as W

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def c5():
    return W.config5()


def _open_subset(V, F, faces=2000):
    # a connected cap: the faces nearest the +z pole (an open surface)
    cz = V[F].mean(axis=1)[:, 2]
    keep = np.argsort(-cz)[:faces]
    return F[np.sort(keep)]


def _loss_and_grads(V, F, O, t, precision="fp64"):
    import igl.autograd as A

    dev = torch.device("cuda")
    Vd = torch.tensor(V, device=dev, requires_grad=True)
    Od = torch.tensor(O, device=dev, requires_grad=True)
    Fd = torch.tensor(F, device=dev)
    td = torch.tensor(t, device=dev)
    w = A.winding_number(Vd, Fd, Od, precision)
    L = ((w - td) ** 2).sum()
    L.backward()
    return float(L), Vd.grad.cpu().numpy().astype(np.float64), Od.grad.cpu().numpy().astype(np.float64)


def test_matches_fp64_torch_reference_open_subset(require_gpu, c5):
    V, F, Q, t = c5
    V = V.astype(np.float64)
    Fs = _open_subset(V, F)
    rng = np.random.default_rng(504)
    sel = rng.choice(len(Q), 384, replace=False)
    O = Q[sel].astype(np.float64)
    ts = t[sel]
    L, gV, gO = _loss_and_grads(V, Fs, O, ts)
    Vt = torch.tensor(V, dtype=torch.float64, requires_grad=True)
    Ot = torch.tensor(O, dtype=torch.float64, requires_grad=True)
    w = oracle.winding_torch(Vt, torch.tensor(Fs), Ot)
    Lr = ((w - torch.tensor(ts)) ** 2).sum()
    Lr.backward()
    assert L == pytest.approx(float(Lr), rel=1e-12, abs=1e-12)
    for got, ref in ((gV, Vt.grad.numpy()), (gO, Ot.grad.numpy())):
        scale = np.abs(ref).max()
        err = np.abs(got - ref).max()
        print(f"open subset: max|grad| {scale:.3e}, max err {err:.3e}", flush=True)
        assert scale > 1e-3  # non-trivial gradient on the open cap
        assert err <= 1e-9 * scale


def _fd_check(V, F, O, t, coords, h=1e-6):
    import igl.autograd as A

    dev = torch.device("cuda")
    Fd = torch.tensor(F, device=dev)
    td = torch.tensor(t, device=dev, dtype=torch.float64)

    def loss(Vn, On):
        tri = A.pack_faces(torch.tensor(Vn, device=dev), Fd)
        w = A.forward(tri, torch.tensor(On, device=dev), "fp64")
        return float(((w - td) ** 2).sum())

    _, gV, gO = _loss_and_grads(V, F, O, t)
    worst = 0.0
    for which, i, d in coords:
        Vp, Vm, Op, Om = V.copy(), V.copy(), O.copy(), O.copy()
        if which == 0:
            Vp[i, d] += h
            Vm[i, d] -= h
            an = gV[i, d]
        else:
            Op[i, d] += h
            Om[i, d] -= h
            an = gO[i, d]
        fd = (loss(Vp, Op) - loss(Vm, Om)) / (2 * h)
        worst = max(worst, abs(fd - an) / (1.0 + abs(an)))
        assert abs(fd - an) <= 1e-5 * (1.0 + abs(an)), (which, i, d, fd, an)
    return worst


def test_central_differences_open_subset(require_gpu, c5):
    V, F, Q, t = c5
    V = V.astype(np.float64)
    Fs = _open_subset(V, F)
    used = np.unique(Fs)
    rng = np.random.default_rng(504)
    sel = rng.choice(len(Q), 256, replace=False)
    O = Q[sel].astype(np.float64)
    ts = t[sel]
    coords = [(0, int(rng.choice(used)), int(rng.integers(3))) for _ in range(128)]
    coords += [(1, int(rng.integers(len(O))), int(rng.integers(3))) for _ in range(128)]
    worst = _fd_check(V, Fs, O, ts, coords)
    print(f"open subset FD worst scaled error {worst:.3e}", flush=True)


def test_config5_full_size_closed_sphere(require_gpu, c5):
    """Closed sphere: w is locally constant, so the exact gradients vanish;
    the adjoint's cancellation and the FD derivatives must agree on that."""
    V, F, Q, t = c5
    L, gV, gO = _loss_and_grads(V, F, Q, t)
    assert np.isfinite(gV).all() and np.isfinite(gO).all()
    print(f"C5 full: L={L:.6e} max|dL/dV|={np.abs(gV).max():.3e} "
          f"max|dL/dO|={np.abs(gO).max():.3e}", flush=True)
    rng = np.random.default_rng(504)
    sub = rng.choice(len(Q), 4096, replace=False)
    O = Q[sub].astype(np.float64)
    coords = [(0, int(rng.integers(len(V))), int(rng.integers(3))) for _ in range(128)]
    coords += [(1, int(rng.integers(len(O))), int(rng.integers(3))) for _ in range(128)]
    worst = _fd_check(V.astype(np.float64), F, O, t[sub], coords)
    print(f"closed sphere FD worst scaled error {worst:.3e}", flush=True)


def test_split_and_unsplit_adjoint_agree(require_gpu, c5):
    import igl.autograd as A

    V, F, Q, t = c5
    dev = torch.device("cuda")
    tri = A.pack_faces(torch.tensor(V, device=dev), torch.tensor(F, device=dev))
    O = torch.tensor(Q[:2048], device=dev, dtype=torch.float64)
    g = torch.tensor(np.random.default_rng(1).normal(size=2048), device=dev)
    q1, t1 = A.backward(tri, O, g, qpl=1, splits=1)
    q2, t2 = A.backward(tri, O, g, qpl=2, splits=13)
    torch.cuda.synchronize()
    assert torch.allclose(q1, q2, rtol=1e-10, atol=1e-10)
    assert torch.allclose(t1, t2, rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("qpl,splits", [(1, 1), (2, 3), (4, 1), (4, 5), (0, 0)])
def test_lane_and_split_variants_match_fp64_torch_open_subset(require_gpu, c5, qpl, splits):
    """Every queries-per-lane / face-split variant of the adjoint (0 = the
    library default) against the FP64 torch gradient of sum_q g_q w_q on the
    open cap, with a query count that leaves partial blocks at every width."""
    import igl.autograd as A

    V, F, Q, t = c5
    V = V.astype(np.float64)
    Fs = _open_subset(V, F)
    rng = np.random.default_rng(77)
    sel = rng.choice(len(Q), 1000, replace=False)
    O = Q[sel].astype(np.float64)
    g = rng.standard_normal(len(O))
    dev = torch.device("cuda")
    Vd = torch.tensor(V, device=dev)
    Fd = torch.tensor(Fs, device=dev)
    tri = A.pack_faces(Vd, Fd)
    gq, gtri = A.backward(tri, torch.tensor(O, device=dev), torch.tensor(g, device=dev),
                          qpl=qpl, splits=splits)
    gV = A.scatter_corners(gtri, Fd, len(V)).cpu().numpy()
    gq = gq.cpu().numpy()
    Vt = torch.tensor(V, dtype=torch.float64, requires_grad=True)
    Ot = torch.tensor(O, dtype=torch.float64, requires_grad=True)
    w = oracle.winding_torch(Vt, torch.tensor(Fs), Ot)
    (w * torch.tensor(g)).sum().backward()
    for got, ref in ((gV, Vt.grad.numpy()), (gq, Ot.grad.numpy())):
        scale = np.abs(ref).max()
        err = np.abs(got - ref).max()
        print(f"qpl={qpl} splits={splits}: max|grad| {scale:.3e}, max err {err:.3e}", flush=True)
        assert scale > 1e-3
        assert err <= 1e-9 * scale


# ---- edge form of the adjoint (wn_boundary_edges + wn_backward_edges) ----

def _torch_grads(V, F, O, g):
    Vt = torch.tensor(V, dtype=torch.float64, requires_grad=True)
    Ot = torch.tensor(O, dtype=torch.float64, requires_grad=True)
    w = oracle.winding_torch(Vt, torch.tensor(F), Ot)
    (w * torch.tensor(g)).sum().backward()
    return Vt.grad.numpy(), Ot.grad.numpy()


@pytest.mark.parametrize("qpl,splits,vdtype", [(1, 1, "f64"), (2, 3, "f64"), (4, 1, "f32"),
                                                (4, 5, "f64"), (0, 0, "f32")])
def test_edge_adjoint_matches_fp64_torch_open_subset(require_gpu, c5, qpl, splits, vdtype):
    """wn_backward_edges (every lane / split variant, float32 and float64 V)
    against the FP64 torch gradient of sum_q g_q w_q on the open cap."""
    import igl.autograd as A

    V, F, Q, t = c5
    V = V.astype(np.float64 if vdtype == "f64" else np.float32)
    Fs = _open_subset(V, F)
    rng = np.random.default_rng(78)
    sel = rng.choice(len(Q), 1000, replace=False)
    O = Q[sel].astype(np.float64)
    g = rng.standard_normal(len(O))
    dev = torch.device("cuda")
    edges, weight = A.boundary_edges(Fs, len(V))
    assert 0 < len(weight) < len(Fs)
    gq, gV = A.backward_edges(torch.tensor(V, device=dev), edges, weight,
                              torch.tensor(O, device=dev), torch.tensor(g, device=dev),
                              qpl=qpl, splits=splits)
    rV, rO = _torch_grads(V.astype(np.float64), Fs, O, g)
    for got, ref in ((gV.cpu().numpy(), rV), (gq.cpu().numpy(), rO)):
        scale = np.abs(ref).max()
        err = np.abs(got - ref).max()
        print(f"edges qpl={qpl} splits={splits} {vdtype}: {len(weight)} edges, "
              f"max|grad| {scale:.3e}, max err {err:.3e}", flush=True)
        assert scale > 1e-3
        assert err <= 1e-9 * scale


def test_edge_and_face_adjoints_agree(require_gpu, c5):
    """Both autograd adjoints on the open cap and on an open cap with a
    flipped face (a weight-2 boundary), through winding_number."""
    V, F, Q, t = c5
    V = V.astype(np.float64)
    Fs = _open_subset(V, F)
    Ff = Fs.copy()
    Ff[7] = Ff[7][::-1]
    rng = np.random.default_rng(79)
    O = Q[rng.choice(len(Q), 700, replace=False)].astype(np.float64)
    ts = rng.uniform(size=len(O))
    import igl.autograd as A

    for faces in (Fs, Ff):
        out = {}
        for adjoint in ("edges", "faces"):
            Vd = torch.tensor(V, device="cuda", requires_grad=True)
            Od = torch.tensor(O, device="cuda", requires_grad=True)
            w = A.winding_number(Vd, torch.tensor(faces, device="cuda"), Od, "fp64", adjoint)
            ((w - torch.tensor(ts, device="cuda")) ** 2).sum().backward()
            out[adjoint] = (Vd.grad.cpu().numpy(), Od.grad.cpu().numpy())
        for i in range(2):
            scale = np.abs(out["faces"][i]).max()
            err = np.abs(out["edges"][i] - out["faces"][i]).max()
            print(f"edges vs faces: max|grad| {scale:.3e}, max diff {err:.3e}", flush=True)
            assert err <= 1e-9 * scale


def test_edge_adjoint_closed_sphere_is_exactly_zero(require_gpu, c5):
    """The closed sphere has no boundary edges: the edge adjoint returns exact
    zeros at full size, which the face adjoint reaches only to rounding."""
    import igl.autograd as A

    V, F, Q, t = c5
    edges, weight = A.boundary_edges(F, len(V))
    assert len(weight) == 0
    dev = torch.device("cuda")
    Vd = torch.tensor(V, device=dev)
    Od = torch.tensor(Q, device=dev, dtype=torch.float64)
    g = torch.tensor(np.random.default_rng(2).normal(size=len(Q)), device=dev)
    gq, gV = A.backward_edges(Vd, edges, weight, Od, g)
    assert not gq.any() and not gV.any()
    fq, ftri = A.backward(A.pack_faces(Vd, torch.tensor(F, device=dev)), Od, g)
    fV = A.scatter_corners(ftri, torch.tensor(F, device=dev), len(V))
    print(f"closed sphere: face adjoint max|dL/dO| {float(fq.abs().max()):.3e}, "
          f"max|dL/dV| {float(fV.abs().max()):.3e} (exact 0)", flush=True)
    assert float(fq.abs().max()) < 1e-6 and float(fV.abs().max()) < 1e-6


def test_edge_adjoint_skips_ids_outside_V(require_gpu, c5):
    """An edge whose id is outside [0, nv) contributes nothing (checked in the
    kernel); the rest of the result is unchanged."""
    import igl.autograd as A

    V, F, Q, t = c5
    Fs = _open_subset(V, F)
    edges, weight = A.boundary_edges(Fs, len(V))
    bad = torch.tensor([[0, len(V)], [len(V) + 5, 1], [-1, 2]], dtype=torch.int32, device="cuda")
    e2 = torch.cat([edges, bad])
    w2 = torch.cat([weight, torch.ones(3, dtype=torch.float64, device="cuda")])
    dev = torch.device("cuda")
    Vd = torch.tensor(V, device=dev)
    Od = torch.tensor(Q[:3000], device=dev, dtype=torch.float64)
    g = torch.tensor(np.random.default_rng(3).normal(size=3000), device=dev)
    q1, v1 = A.backward_edges(Vd, edges, weight, Od, g, splits=1)
    q2, v2 = A.backward_edges(Vd, e2, w2, Od, g, splits=1)
    assert torch.equal(q1, q2)
    # grad_V is summed with atomics (block order varies): equal to rounding
    assert float((v1 - v2).abs().max()) <= 1e-12 * float(v1.abs().max())


def test_edge_adjoint_empty_inputs(require_gpu):
    import igl.autograd as A

    dev = torch.device("cuda")
    V = torch.zeros((4, 3), dtype=torch.float64, device=dev)
    e = torch.zeros((0, 2), dtype=torch.int32, device=dev)
    m = torch.zeros(0, dtype=torch.float64, device=dev)
    O = torch.zeros((0, 3), dtype=torch.float64, device=dev)
    gq, gV = A.backward_edges(V, e, m, O, torch.zeros(0, dtype=torch.float64, device=dev))
    assert gq.shape == (0, 3) and gV.shape == (4, 3) and not gV.any()


def test_large_functional_calls_use_a_kept_strip_field(require_gpu, c5):
    """igl.autograd.winding_number at >= STRIP_MIN_QUERIES queries runs the
    strip kernels through a WindingField kept for the next call with an equal
    F; w and both gradients match the face-record route (WindingNumber) on
    an open cap x 262,144 queries: FP64 w to 1e-12, FP32c w to 1e-6;
    gradients (the same FP64 edge adjoint) to 1e-12 relative in FP64 and
    1e-6 in FP32c, whose upstream dL/dw = 2 (w - t) carries the routes'
    difference in w."""
    import igl.autograd as A

    V, F, Q, t = c5
    V = V.astype(np.float64)
    Fs = _open_subset(V, F)
    td = torch.tensor(t, device="cuda")

    def run(fn, vdtype, precision, faces):
        Vd = torch.tensor(V, device="cuda", dtype=vdtype, requires_grad=True)
        Od = torch.tensor(Q, device="cuda", requires_grad=True)
        w = fn(Vd, torch.tensor(faces, device="cuda"), Od, precision, "edges")
        ((w - td) ** 2).sum().backward()
        return w.detach().cpu().numpy(), Vd.grad.cpu().numpy(), Od.grad.cpu().numpy()

    A.release()
    for precision, wbar, gbar in (("fp64", 1e-12, 1e-12), ("fp32c", 1e-6, 1e-6)):
        for vdtype in (torch.float64, torch.float32):
            for faces in (Fs, Fs[::-1].copy(), Fs):
                got = run(A.winding_number, vdtype, precision, faces)
                field = A.kept_field()
                assert field is not None and field.precision == precision
                ref = run(A.WindingNumber.apply, vdtype, precision, faces)
                assert np.abs(got[0] - ref[0]).max() <= wbar
                for i in (1, 2):
                    scale = np.abs(ref[i]).max()
                    assert np.abs(got[i] - ref[i]).max() <= gbar * scale + 1e-300
            # a call repeating the last F reuses the kept field, with the same bits
            kept = A.kept_field()
            again = run(A.winding_number, torch.float64, precision, Fs)
            assert A.kept_field() is kept
            assert np.array_equal(again[0], run(A.winding_number, torch.float64, precision, Fs)[0])
    A.release()
    assert A.kept_field() is None


def test_kept_strip_fields_are_per_thread(require_gpu, c5):
    """Two threads calling winding_number concurrently on the same F with
    different V (each thread keeps its own field, whose records every call
    rewrites) get the values of single-threaded calls, bit for bit."""
    import threading

    import igl.autograd as A

    V, F, Q, _ = c5
    V = V.astype(np.float64)
    Fd = torch.tensor(_open_subset(V, F), device="cuda")
    Od = torch.tensor(Q, device="cuda")
    Vs = [torch.tensor(V * s, device="cuda") for s in (1.0, 1.1)]
    A.release()
    expect = [A.winding_number(v, Fd, Od, "fp32c").cpu().numpy() for v in Vs]
    A.release()
    out, errors = {}, []

    def worker(k):
        try:
            res = [A.winding_number(Vs[k], Fd, Od, "fp32c").cpu().numpy() for _ in range(4)]
            out[k] = res
            A.release()
        except Exception as e:  # reported below
            errors.append(repr(e))

    threads = [threading.Thread(target=worker, args=(k,)) for k in (0, 1)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors
    for k in (0, 1):
        for r in out[k]:
            assert np.array_equal(r, expect[k])
