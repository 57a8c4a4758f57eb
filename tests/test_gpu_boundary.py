"""The FP32-compensated boundary correction on open meshes, near their
boundary: every FP32c path (strip mesh, face kernel from the face SoA, face
kernel from V and F) against the FP64 strip kernel on all queries and
against the oracle (FP64 restatement of src/capi.mojo:205-294) on a
subsample.

Mesh: a level-6 icosphere (40,960 faces) with the cap z > 0.5 removed, so
its rim is a closed boundary curve of ~200 edges. Queries: 16,384 uniform in
the box; 16,384 at distance |d| log-uniform in [1e-7, 1e-2] from a random
point of a random boundary edge, in a random direction perpendicular to the
edge (w is not locally constant there: |grad w| ~ 1 / (2 pi |d|)); 8,192 on
both sides of the surface away from the rim. Variants: the same mesh moved to
1e6 from the origin (block-local origins), and the mesh with every face
listed twice (boundary edges of weight 2, w doubled).

Bars: uniform and surface-side queries 1e-6; rim queries 1e-6 at |d| >=
1e-5 and the near-surface bar 1e-4 closer in, for queries at least 1e-7
from the surface itself (a rim query leaving the edge almost along the
surface can lie closer than that, where the FP32 rounding of the block-local
surface, ~6e-8 x block radius, may put it on the other side; those are only
counted)."""

import numpy as np
import pytest

import igl
import oracle
import wn_workloads as W

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def open_cap():
    V, F = W.icosphere(6)
    c = V[F].mean(axis=1)
    F = np.ascontiguousarray(F[c[:, 2] <= 0.5])
    used = np.unique(F)
    remap = np.full(len(V), -1, dtype=np.int64)
    remap[used] = np.arange(len(used))
    return np.ascontiguousarray(V[used]), np.ascontiguousarray(remap[F])


def rim_queries(V, F, n, rng):
    be = np.array(list(oracle.boundary_edges(F).keys()))
    e = be[rng.integers(0, len(be), n)]
    P1, P2 = V[e[:, 0]], V[e[:, 1]]
    t = rng.uniform(0.05, 0.95, size=(n, 1))
    p = P1 + t * (P2 - P1)
    u = (P2 - P1) / np.linalg.norm(P2 - P1, axis=1, keepdims=True)
    r = rng.normal(size=(n, 3))
    r -= (r * u).sum(axis=1, keepdims=True) * u
    r /= np.linalg.norm(r, axis=1, keepdims=True)
    d = 10.0 ** rng.uniform(-7.0, -2.0, size=n)
    return p + d[:, None] * r, d


def surface_queries(V, F, n, rng):
    f = rng.integers(0, len(F), n)
    T = V[F[f]]
    p = np.einsum("nk,nkd->nd", rng.dirichlet(np.ones(3), size=n), T)
    nrm = np.cross(T[:, 1] - T[:, 0], T[:, 2] - T[:, 0])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    d = np.where(rng.random(n) < 0.5, -1.0, 1.0) * 10.0 ** rng.uniform(-4.0, -2.0, size=n)
    return p + d[:, None] * nrm


@pytest.fixture(scope="module")
def cap():
    V, F = open_cap()
    rng = np.random.default_rng(61)
    Qr, d = rim_queries(V, F, 16_384, rng)
    Q = np.concatenate([rng.uniform(V.min(0) - 0.2, V.max(0) + 0.2, size=(16_384, 3)), Qr,
                        surface_queries(V, F, 8_192, rng)])
    dist = np.concatenate([np.full(16_384, np.inf), d, np.full(8_192, np.inf)])
    return V, F, np.ascontiguousarray(Q), dist


def fp32c_paths(V, F, Q):
    from igl._lib import addr

    h = igl.lib()
    out = {}
    m = igl.Mesh(V, F)
    try:
        out["strips"] = m.winding_number(Q)
        ref = m.winding_number(Q, precision="fp64")
    finally:
        m.close()
    tri = igl.pack_faces(V, F)
    w = np.full(len(Q), np.nan)
    assert h.wn_winding_gpu_ex(addr(tri), addr(Q), addr(w), tri.shape[1], len(Q), 1, 0, 0, -1,
                               -1.0) == 0, igl._lib.last_error()
    out["face SoA"] = w
    Vc, Fc = np.ascontiguousarray(V), np.ascontiguousarray(F)
    w = np.full(len(Q), np.nan)
    assert h.wn_winding_gpu_vf(addr(Vc), addr(Fc), len(Vc), len(Fc), 0, addr(Q), addr(w), len(Q),
                               1) == 0, igl._lib.last_error()
    out["face V/F"] = w
    return out, ref


def check(name, V, F, Q, dist):
    paths, ref = fp32c_paths(V, F, Q)
    # distance to the surface itself (rim queries can leave the edge almost
    # along the surface)
    to_surface = np.sqrt(igl.point_mesh_squared_distance(Q, V, F)[0])
    on = to_surface < 1e-7
    rim = np.isfinite(dist) & ~on
    close = rim & (dist < 1e-5)
    other = ~np.isfinite(dist) & ~on
    sub = np.arange(0, len(Q), 97)
    sub = sub[~on[sub]]
    expect = oracle.winding(V, F, Q[sub])
    assert np.abs(ref[sub] - expect).max() <= 1e-10
    for path, w in paths.items():
        err = np.abs(w - ref)
        flips = int((err[on] > 0.5).sum())
        print(f"{name} {path}: uniform/surface max {err[other].max(initial=0):.2e}, rim |d| >= "
              f"1e-5 max {err[rim & ~close].max(initial=0):.2e}, rim |d| < 1e-5 max "
              f"{err[close].max(initial=0):.2e}, subsample vs oracle "
              f"{np.abs(w[sub] - expect).max():.2e}; {int(on.sum())} queries within 1e-7 of "
              f"the surface, {flips} on the other side", flush=True)
        assert err[other].max(initial=0) <= 1e-6, path
        assert err[rim & ~close].max(initial=0) <= 1e-6, path
        assert err[close].max(initial=0) <= 1e-4, path


def test_open_cap_near_its_rim(require_gpu, cap):
    V, F, Q, dist = cap
    assert len(oracle.boundary_edges(F)) > 100
    check("cap", V, F, Q, dist)


def test_open_cap_far_from_the_origin(require_gpu, cap):
    V, F, Q, dist = cap
    shift = np.array([1e6, -2e6, 5e5])
    check("cap at 1e6", V + shift, F, Q + shift, dist)


def test_open_cap_with_doubled_faces(require_gpu, cap):
    """Every face twice: boundary edges of weight 2, w twice the single
    mesh's."""
    V, F, Q, dist = cap
    F2 = np.ascontiguousarray(np.vstack([F, F]))
    edges = oracle.boundary_edges(F2)
    assert set(edges.values()) == {2}
    sel = slice(0, 8_192)
    w1 = igl.winding_number(V, F, Q[sel], device="gpu", precision="fp32c")
    w2 = igl.winding_number(V, F2, Q[sel], device="gpu", precision="fp32c")
    assert np.abs(w2 - 2.0 * w1).max() <= 2e-6
    check("doubled cap", V, F2, Q[8_192:24_576], dist[8_192:24_576])
