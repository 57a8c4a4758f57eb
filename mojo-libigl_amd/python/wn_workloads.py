"""Synthetic meshes and query sets of the winding-number workloads.

All generators are seeded with ``numpy.random.default_rng`` and produce
float64 vertices / int64 faces unless stated. Workloads:

* ``config1``  icosphere level 4 (2,562 V, 5,120 F), radial noise +-2 %
  (seed 101), 1,024 queries uniform in [-1.5, 1.5]^3 (seed 102).
* ``config2``  closed torus R = 1, r = 0.35 on a 120 x 118 grid (14,160 V,
  28,320 F); queries uniform in the bbox padded by 20 % per side
  (seed 201 for 1,024, seed 202 for 1,048,576).
* ``config3``  icosphere level 8 (655,362 V, 1,310,720 F), radius
  1 + 0.05 S / max|S| with S a sum of 8 real spherical harmonics (seed 301);
  half the queries uniform in the bbox, half at p + d n for random surface
  points p, |d| log-uniform in [1e-7, 1e-2], random sign (seed 302).
* ``config4``  512 meshes (seed 401): torus / geodesic icosphere / open
  cylinder with holes, faces log-uniform in [500, 60,000], queries
  log-uniform in [64, 65,536]; float32 V, int32 F.
* ``config5``  frequency-50 geodesic icosphere (25,002 V, 50,000 F), 5 %
  radial noise (seed 501), float32; 262,144 queries (seed 502), targets
  U[0, 1] (seed 503).

The reference's own test and bench meshes (``grid`` of tests/test_parity.py:12-26,
``bench_grid`` of bench/bench.py:33-47, ``tetrahedron`` of
tests/test_parity.py:29-34) are restated too.
"""

from __future__ import annotations

import numpy as np

# ---------------------------------------------------------------------------
# Reference test meshes
# ---------------------------------------------------------------------------


def grid(n=7, curved=True):
    """Height field z = 0.15 sin(i) cos(j) on an n x n lattice, two triangles
    per cell (reference tests/test_parity.py:12-26)."""
    V = np.array([[i, j, 0.15 * np.sin(i) * np.cos(j) if curved else 0.0]
                  for j in range(n) for i in range(n)], dtype=np.float64)
    F = []
    for j in range(n - 1):
        for i in range(n - 1):
            a = j * n + i
            F.extend(((a, a + 1, a + n + 1), (a, a + n + 1, a + n)))
    return V, np.asarray(F, dtype=np.int64).reshape(-1, 3)


def bench_grid(n: int):
    """Open height field z = 0.1 sin(x/8) cos(y/11) (reference
    bench/bench.py:33-47); n = 120 gives 14,400 V and 28,322 F."""
    y, x = np.mgrid[:n, :n]
    z = 0.1 * np.sin(x / 8.0) * np.cos(y / 11.0)
    V = np.ascontiguousarray(np.column_stack((x.ravel(), y.ravel(), z.ravel())), dtype=np.float64)
    a = (np.arange(n - 1)[:, None] * n + np.arange(n - 1)[None, :]).ravel()
    F = np.empty((2 * len(a), 3), dtype=np.int64)
    F[0::2] = np.column_stack((a, a + 1, a + n + 1))
    F[1::2] = np.column_stack((a, a + n + 1, a + n))
    return V, F


def tetrahedron():
    """Reference tests/test_parity.py:29-34 (outward oriented)."""
    V = np.array([[0.0, 0.0, 0.0], [2.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 3.0]])
    F = np.array([[0, 2, 1], [0, 1, 3], [1, 2, 3], [2, 0, 3]], dtype=np.int64)
    return V, F


# ---------------------------------------------------------------------------
# Closed and open meshes
# ---------------------------------------------------------------------------


def icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    V = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    F = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return V, F


def icosphere(level: int):
    """Recursive 4:1 subdivision of the unit icosahedron, vertices projected
    to the unit sphere: 10*4^k + 2 V, 20*4^k F, outward oriented."""
    V, F = icosahedron()
    for _ in range(level):
        nv = len(V)
        edges = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
        edges.sort(axis=1)
        uniq, inverse = np.unique(edges, axis=0, return_inverse=True)
        inverse = inverse.reshape(-1)
        mid = V[uniq[:, 0]] + V[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        V = np.concatenate([V, mid])
        m = len(F)
        ab = nv + inverse[:m]
        bc = nv + inverse[m:2 * m]
        ca = nv + inverse[2 * m:]
        a, b, c = F[:, 0], F[:, 1], F[:, 2]
        F = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1),
                            np.stack([c, ca, bc], 1), np.stack([ab, bc, ca], 1)])
    return np.ascontiguousarray(V), np.ascontiguousarray(F.astype(np.int64))


def geodesic_sphere(freq: int):
    """Frequency-n geodesic subdivision of the icosahedron (each face cut into
    n^2 triangles on a barycentric grid), projected to the unit sphere:
    10 n^2 + 2 V, 20 n^2 F, outward oriented."""
    V0, F0 = icosahedron()
    n = int(freq)
    pts, tris = [], []
    base = 0
    ij = [(i, j) for i in range(n + 1) for j in range(n + 1 - i)]
    local = {p: k for k, p in enumerate(ij)}
    loc_tris = []
    for i in range(n):
        for j in range(n - i):
            loc_tris.append((local[(i, j)], local[(i + 1, j)], local[(i, j + 1)]))
            if j < n - i - 1:
                loc_tris.append((local[(i + 1, j)], local[(i + 1, j + 1)], local[(i, j + 1)]))
    ij = np.array(ij, dtype=np.float64)
    loc_tris = np.array(loc_tris, dtype=np.int64)
    for a, b, c in F0:
        wa = (n - ij[:, 0] - ij[:, 1]) / n
        P = wa[:, None] * V0[a] + (ij[:, 0] / n)[:, None] * V0[b] + (ij[:, 1] / n)[:, None] * V0[c]
        pts.append(P)
        tris.append(loc_tris + base)
        base += len(P)
    P = np.concatenate(pts)
    T = np.concatenate(tris)
    key = np.round(P, 9)
    _, first, inverse = np.unique(key, axis=0, return_index=True, return_inverse=True)
    V = P[first]
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    F = inverse.reshape(-1)[T]
    return np.ascontiguousarray(V), np.ascontiguousarray(F.astype(np.int64))


def torus(R=1.0, r=0.35, nu=120, nv=118):
    """Closed torus on an nu (around the axis) x nv (around the tube) grid,
    outward oriented: w = 1 inside the tube."""
    u = 2.0 * np.pi * np.arange(nu) / nu
    v = 2.0 * np.pi * np.arange(nv) / nv
    uu, vv = np.meshgrid(u, v, indexing="ij")
    V = np.stack([(R + r * np.cos(vv)) * np.cos(uu), (R + r * np.cos(vv)) * np.sin(uu),
                  r * np.sin(vv)], axis=-1).reshape(-1, 3)
    i, j = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    i, j = i.ravel(), j.ravel()
    a = i * nv + j
    b = ((i + 1) % nu) * nv + j
    c = ((i + 1) % nu) * nv + (j + 1) % nv
    d = i * nv + (j + 1) % nv
    F = np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1)])
    return np.ascontiguousarray(V), np.ascontiguousarray(F.astype(np.int64))


def open_cylinder(nu: int, nv: int, holes: int, rng, radius=0.5, height=1.5):
    """Open tube (no caps) with `holes` rectangular patches of faces removed:
    an open surface, so w is non-integer."""
    u = 2.0 * np.pi * np.arange(nu) / nu
    z = np.linspace(-height / 2, height / 2, nv + 1)
    uu, zz = np.meshgrid(u, z, indexing="ij")
    V = np.stack([radius * np.cos(uu), radius * np.sin(uu), zz], axis=-1).reshape(-1, 3)
    i, j = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    i, j = i.ravel(), j.ravel()
    a = i * (nv + 1) + j
    b = ((i + 1) % nu) * (nv + 1) + j
    c = ((i + 1) % nu) * (nv + 1) + j + 1
    d = i * (nv + 1) + j + 1
    keep = np.ones(nu * nv, dtype=bool)
    for _ in range(holes):
        ci, cj = rng.integers(nu), rng.integers(nv)
        hi, hj = max(1, nu // 10), max(1, nv // 6)
        mask = (((i - ci) % nu) < hi) & (np.abs(j - cj) < hj)
        keep &= ~mask
    a, b, c, d = a[keep], b[keep], c[keep], d[keep]
    F = np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1)])
    return np.ascontiguousarray(V), np.ascontiguousarray(F.astype(np.int64))


def padded_box_queries(V, n, rng, pad=0.2):
    lo, hi = V.min(0), V.max(0)
    ext = hi - lo
    return rng.uniform(lo - pad * ext, hi + pad * ext, size=(n, 3))


def real_sph_harm(l, m, theta, phi):
    """Real spherical harmonic of degree l, order m (theta polar, phi
    azimuth), from scipy's complex one."""
    from scipy import special

    if hasattr(special, "sph_harm_y"):
        Y = special.sph_harm_y(l, abs(m), theta, phi)
    else:  # older scipy: sph_harm(m, l, azimuth, polar)
        Y = special.sph_harm(abs(m), l, phi, theta)
    if m > 0:
        return np.sqrt(2.0) * (-1) ** m * Y.real
    if m < 0:
        return np.sqrt(2.0) * (-1) ** m * Y.imag
    return Y.real


# ---------------------------------------------------------------------------
# Workloads
# ---------------------------------------------------------------------------


def config1():
    V, F = icosphere(4)
    rng = np.random.default_rng(101)
    V = V * (1.0 + rng.uniform(-0.02, 0.02, size=(len(V), 1)))
    Q = np.random.default_rng(102).uniform(-1.5, 1.5, size=(1024, 3))
    return V, F, Q


def config2(nq=1024):
    V, F = torus()
    seed = 201 if nq == 1024 else 202
    Q = padded_box_queries(V, nq, np.random.default_rng(seed))
    return V, F, Q


def config3_mesh():
    V, F = icosphere(8)
    rng = np.random.default_rng(301)
    theta = np.arccos(np.clip(V[:, 2], -1.0, 1.0))
    phi = np.arctan2(V[:, 1], V[:, 0])
    S = np.zeros(len(V))
    for _ in range(8):
        l = int(rng.integers(0, 9))
        m = int(rng.integers(-l, l + 1))
        S += rng.normal() * real_sph_harm(l, m, theta, phi)
    S /= max(np.abs(S).max(), 1e-300)
    V = V * (1.0 + 0.05 * S)[:, None]
    return np.ascontiguousarray(V), F


def config3_queries(V, F, nq=1_048_576, seed=302):
    """Half uniform in the bbox, half at signed normal offsets from random
    surface points. Returns (Q, d) with d = NaN for the uniform half."""
    rng = np.random.default_rng(seed)
    half = nq // 2
    Qu = rng.uniform(V.min(0), V.max(0), size=(half, 3))
    f = rng.integers(0, len(F), size=nq - half)
    bary = rng.dirichlet(np.ones(3), size=nq - half)
    T = V[F[f]]
    p = np.einsum("nk,nkd->nd", bary, T)
    nrm = np.cross(T[:, 1] - T[:, 0], T[:, 2] - T[:, 0])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    mag = 10.0 ** rng.uniform(-7.0, -2.0, size=nq - half)
    d = np.where(rng.random(nq - half) < 0.5, -mag, mag)
    Qn = p + d[:, None] * nrm
    Q = np.concatenate([Qu, Qn])
    dist = np.concatenate([np.full(half, np.nan), d])
    return np.ascontiguousarray(Q), dist


def config4(n_meshes=512, seed=401, max_faces=60_000, max_queries=65_536):
    """List of (V float32, F int32, Q float64) for the batched workload."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_meshes):
        target = int(np.exp(rng.uniform(np.log(500), np.log(max_faces))))
        nq = int(np.exp(rng.uniform(np.log(64), np.log(max_queries))))
        kind = int(rng.integers(0, 3))
        if kind == 0:
            nu = max(3, int(round(np.sqrt(target / 2.0 * 3.0))))
            nv = max(3, int(round(target / (2.0 * nu))))
            V, F = torus(1.0, float(rng.uniform(0.2, 0.45)), nu, nv)
        elif kind == 1:
            freq = max(1, int(round(np.sqrt(target / 20.0))))
            V, F = geodesic_sphere(freq)
        else:
            nu = max(6, int(round(np.sqrt(target / 2.0 * 2.0))))
            nv = max(2, int(round(target / (2.0 * nu))))
            V, F = open_cylinder(nu, nv, int(rng.integers(1, 4)), rng)
        V = V * rng.uniform(0.5, 2.0) + rng.uniform(-2.0, 2.0, size=3)
        Q = padded_box_queries(V, nq, rng)
        out.append((V.astype(np.float32), F.astype(np.int32), Q))
    return out


def config5():
    V, F = geodesic_sphere(50)
    rng = np.random.default_rng(501)
    V = V * (1.0 + rng.uniform(-0.05, 0.05, size=(len(V), 1)))
    V = V.astype(np.float32)
    Q = np.random.default_rng(502).uniform(V.min(0), V.max(0), size=(262_144, 3)).astype(np.float32)
    t = np.random.default_rng(503).uniform(0.0, 1.0, size=262_144)
    return V, F, Q, t
