"""Correctness oracle for the generalized winding number -- TEST INFRASTRUCTURE.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package. The product library never loads it.

Contents, each following the reference file:line it restates:

* :func:`winding` / :func:`winding_threads` -- ctypes over ``wn_oracle.c``, the
  plain-C FP64 restatement of ``src/capi.mojo:205-294`` (reference CPU path).
* :func:`pack_soa` -- the 9 x m face SoA of ``python/igl/__init__.py:248-250``.
* :func:`winding_numpy` -- the per-pair formula of ``src/capi.mojo:218-235``
  vectorised in NumPy (independent second restatement, small inputs).
* :func:`winding_mpmath` -- the same formula at 40+ digits (brute force pin).
* :func:`winding_torch` -- the formula in float64 PyTorch ops, differentiable,
  the FP64 reference gradient for the adjoint kernel.
* :func:`boundary_edges` -- net directed boundary edges by counting (pure
  Python); :func:`edge_adjoint` -- dL/dO, dL/dV of sum g w as a sum over them
  in NumPy (no reference counterpart: the reference has no derivative). Pinned
  against the autograd gradient of :func:`winding_torch` and against hand
  counts of edges in the CPU tests.
* :func:`ref_gpu_atan2` -- the reference GPU kernel's own arctangent,
  ``src/capi.mojo:326-352``.
* :func:`build_aabb` -- the median-split tree of ``python/igl/__init__.py:286-319``
  in NumPy; :func:`aabb_distance` -- the descent of ``src/capi.mojo:401-587``
  over it (``aabb_oracle.c``); :func:`brute_distance` -- the minimum over all
  faces, with no tree; :func:`closest_on_triangle` -- one point, one triangle.
* :func:`fwn_expansion`, :func:`fwn_far_omega`, :func:`exact_omega`,
  :func:`fast_winding` -- the hierarchical winding number (Barill et al.
  2018, second order; not in the reference, README.md:29) restated in NumPy
  over the same tree.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libwn_oracle.so")

_I = ctypes.c_int64
_P = ctypes.c_void_p
_D = ctypes.c_double
_handle = None


def build() -> str:
    """Compile the oracle with its Makefile (used by the build step)."""
    import subprocess

    proc = subprocess.run(["make", "-s", "-C", HERE], capture_output=True, text=True)
    if proc.returncode or not os.path.exists(LIB_PATH):
        raise RuntimeError("oracle build failed:\n" + proc.stdout + proc.stderr)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _handle
    if _handle is None:
        if not os.path.exists(LIB_PATH):
            build()
        h = ctypes.CDLL(LIB_PATH)
        h.wn_oracle_winding.argtypes = [_P, _P, _P, _I, _I, _I, _I]
        h.wn_oracle_winding.restype = None
        h.wn_oracle_winding_threads.argtypes = [_P, _P, _P, _I, _I, _I]
        h.wn_oracle_winding_threads.restype = _I
        h.wn_oracle_pair_angle.argtypes = [_P, _I, _I, _D, _D, _D]
        h.wn_oracle_pair_angle.restype = _D
        h.wn_oracle_ref_gpu_atan2.argtypes = [_D, _D]
        h.wn_oracle_ref_gpu_atan2.restype = _D
        h.wn_oracle_aabb_distance.argtypes = [_P] * 10 + [_I, _I, _I]
        h.wn_oracle_aabb_distance.restype = None
        h.wn_oracle_brute_distance.argtypes = [_P, _P, _P, _I, _P, _P, _P, _I]
        h.wn_oracle_brute_distance.restype = None
        h.wn_oracle_closest_on_triangle.argtypes = [_P, _P, _P]
        h.wn_oracle_closest_on_triangle.restype = None
        _handle = h
    return _handle


def pack_soa(V, F) -> np.ndarray:
    """``vertices[faces].transpose(1, 2, 0).reshape(9, m)``, rows
    ax,ay,az,bx,by,bz,cx,cy,cz (``python/igl/__init__.py:248-250``)."""
    V = np.ascontiguousarray(V, dtype=np.float64)
    F = np.ascontiguousarray(F, dtype=np.int64)
    return np.ascontiguousarray(V[F].transpose(1, 2, 0).reshape(9, len(F)))


def winding_soa(tri: np.ndarray, Q: np.ndarray) -> np.ndarray:
    tri = np.ascontiguousarray(tri, dtype=np.float64)
    Q = np.ascontiguousarray(Q, dtype=np.float64).reshape(-1, 3)
    out = np.empty(len(Q), dtype=np.float64)
    if len(Q):
        lib().wn_oracle_winding(tri.ctypes.data, Q.ctypes.data, out.ctypes.data,
                                tri.shape[1], len(Q), 0, len(Q))
    return out


def winding(V, F, Q) -> np.ndarray:
    """FP64 oracle: the reference CPU algorithm on (V, F, Q)."""
    return winding_soa(pack_soa(V, F), Q)


def winding_threads(tri: np.ndarray, Q: np.ndarray, threads: int) -> np.ndarray:
    tri = np.ascontiguousarray(tri, dtype=np.float64)
    Q = np.ascontiguousarray(Q, dtype=np.float64).reshape(-1, 3)
    out = np.empty(len(Q), dtype=np.float64)
    if len(Q):
        lib().wn_oracle_winding_threads(tri.ctypes.data, Q.ctypes.data, out.ctypes.data,
                                        tri.shape[1], len(Q), int(threads))
    return out


def pair_angle(tri: np.ndarray, f: int, o) -> float:
    tri = np.ascontiguousarray(tri, dtype=np.float64)
    return lib().wn_oracle_pair_angle(tri.ctypes.data, tri.shape[1], int(f),
                                      float(o[0]), float(o[1]), float(o[2]))


def ref_gpu_atan2(det: float, den: float) -> float:
    return lib().wn_oracle_ref_gpu_atan2(float(det), float(den))


def winding_numpy(V, F, Q) -> np.ndarray:
    """Vectorised FP64 restatement of ``src/capi.mojo:218-235,286``."""
    V = np.asarray(V, dtype=np.float64)
    F = np.asarray(F, dtype=np.int64)
    Q = np.asarray(Q, dtype=np.float64).reshape(-1, 3)
    T = V[F]  # (m, 3, 3)
    a = T[None, :, 0, :] - Q[:, None, :]
    b = T[None, :, 1, :] - Q[:, None, :]
    c = T[None, :, 2, :] - Q[:, None, :]
    la = np.sqrt((a * a).sum(-1))
    lb = np.sqrt((b * b).sum(-1))
    lc = np.sqrt((c * c).sum(-1))
    det = (a * np.cross(b, c)).sum(-1)
    den = la * lb * lc + (a * b).sum(-1) * lc + (b * c).sum(-1) * la + (c * a).sum(-1) * lb
    return np.arctan2(det, den).sum(-1) / (2.0 * np.pi)


def winding_mpmath(V, F, Q, dps: int = 40):
    """Brute force at ``dps`` digits; returns a list of mpmath numbers."""
    import mpmath as mp

    with mp.workdps(dps):
        Vm = [[mp.mpf(float(x)) for x in row] for row in np.asarray(V, dtype=np.float64)]
        out = []
        for o in np.asarray(Q, dtype=np.float64).reshape(-1, 3):
            om = [mp.mpf(float(x)) for x in o]
            total = mp.mpf(0)
            for f in np.asarray(F, dtype=np.int64):
                a = [Vm[f[0]][k] - om[k] for k in range(3)]
                b = [Vm[f[1]][k] - om[k] for k in range(3)]
                c = [Vm[f[2]][k] - om[k] for k in range(3)]
                la = mp.sqrt(sum(x * x for x in a))
                lb = mp.sqrt(sum(x * x for x in b))
                lc = mp.sqrt(sum(x * x for x in c))
                det = (a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0])
                       + a[2] * (b[0] * c[1] - b[1] * c[0]))
                den = (la * lb * lc + (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) * lc
                       + (b[0] * c[0] + b[1] * c[1] + b[2] * c[2]) * la
                       + (c[0] * a[0] + c[1] * a[1] + c[2] * a[2]) * lb)
                total += mp.atan2(det, den)
            out.append(total / (2 * mp.pi))
        return out


def winding_torch(V, F, Q):
    """Differentiable FP64 restatement of ``src/capi.mojo:218-235,286`` in
    PyTorch ops (CPU); the reference gradient for the adjoint kernel."""
    import torch

    T = V[F]  # (m, 3, 3)
    a = T[None, :, 0, :] - Q[:, None, :]
    b = T[None, :, 1, :] - Q[:, None, :]
    c = T[None, :, 2, :] - Q[:, None, :]
    la = a.norm(dim=-1)
    lb = b.norm(dim=-1)
    lc = c.norm(dim=-1)
    det = (a * torch.linalg.cross(b, c, dim=-1)).sum(-1)
    den = la * lb * lc + (a * b).sum(-1) * lc + (b * c).sum(-1) * la + (c * a).sum(-1) * lb
    return torch.atan2(det, den).sum(-1) / (2.0 * torch.pi)


def boundary_edges(F):
    """{(u, v): n} for every undirected edge used n > 0 more times as u -> v
    than as v -> u by the faces' directed edges (i->j, j->k, k->i)."""
    net = {}
    for f in np.asarray(F).tolist():
        for s, t in ((f[0], f[1]), (f[1], f[2]), (f[2], f[0])):
            if s == t:
                continue
            key, sign = ((s, t), 1) if s < t else ((t, s), -1)
            net[key] = net.get(key, 0) + sign
    out = {}
    for (u, v), n in net.items():
        if n > 0:
            out[(u, v)] = n
        elif n < 0:
            out[(v, u)] = -n
    return out


def edge_adjoint(V, F, Q, g):
    """(dL/dV (n, 3), dL/dO (q, 3)) of L = sum_q g_q w_q, as a sum over the
    net boundary edges: per directed edge P1 -> P2 with a = P1 - O,
    b = P2 - O, t = |a||b| + a.b, d theta/dP1 = -(a x b)/(2|a| t),
    d theta/dP2 = -(a x b)/(2|b| t), d theta/dO = -(both). Float64 NumPy."""
    V = np.asarray(V, dtype=np.float64)
    Q = np.asarray(Q, dtype=np.float64)
    g = np.asarray(g, dtype=np.float64) / (2.0 * np.pi)
    gV = np.zeros_like(V)
    gO = np.zeros_like(Q)
    E = boundary_edges(F)
    if not E:
        return gV, gO
    src = np.array([k[0] for k in E])
    dst = np.array([k[1] for k in E])
    wt = np.array(list(E.values()), dtype=np.float64)
    a = V[src][None, :, :] - Q[:, None, :]
    b = V[dst][None, :, :] - Q[:, None, :]
    la = np.linalg.norm(a, axis=-1)
    lb = np.linalg.norm(b, axis=-1)
    x = np.cross(a, b)
    t = la * lb + (a * b).sum(-1)
    k = -0.5 * wt[None, :] * g[:, None] / t
    d1 = (k / la)[..., None] * x
    d2 = (k / lb)[..., None] * x
    np.add.at(gV, src, d1.sum(0))
    np.add.at(gV, dst, d2.sum(0))
    gO = -(d1 + d2).sum(1)
    return gV, gO


# ---------------------------------------------------------------------------
# Point-mesh squared distance
# ---------------------------------------------------------------------------

def build_aabb(V, F, leaf_size: int = 8):
    """(boxes (n, 6) f64, nodes (n, 4) i64, order i64, stack size): the
    median-split tree of python/igl/__init__.py:286-319. Per node, in
    preorder: the box of its faces' corner bounds; leaves of <= leaf_size
    faces; otherwise faces sorted stably by box centre along the axis of
    largest centre spread (first axis on ties), split at len // 2."""
    V = np.ascontiguousarray(V, dtype=np.float64)
    F = np.ascontiguousarray(F, dtype=np.int64)
    tri = V[F]
    lo, hi = tri.min(axis=1), tri.max(axis=1)
    mid = 0.5 * (lo + hi)
    nodes, boxes, order = [], [], []
    deepest = 0

    def visit(ids, depth):
        nonlocal deepest
        deepest = max(deepest, depth)
        me = len(nodes)
        nodes.append(None)
        boxes.append(np.concatenate((lo[ids].min(axis=0), hi[ids].max(axis=0))))
        if len(ids) <= leaf_size:
            nodes[me] = [-1, -1, len(order), len(ids)]
            order.extend(int(i) for i in ids)
            return me
        spread = mid[ids].max(axis=0) - mid[ids].min(axis=0)
        axis = int(np.argmax(spread))
        ids = ids[np.argsort(mid[ids, axis], kind="stable")]
        half = len(ids) // 2
        a = visit(ids[:half], depth + 1)
        b = visit(ids[half:], depth + 1)
        nodes[me] = [a, b, 0, 0]
        return me

    visit(np.arange(len(F), dtype=np.int64), 1)
    return (np.ascontiguousarray(boxes, dtype=np.float64).reshape(-1, 6),
            np.ascontiguousarray(nodes, dtype=np.int64).reshape(-1, 4),
            np.asarray(order, dtype=np.int64), deepest + 1)


def aabb_distance(P, V, F, tree=None):
    """(sqrd, face index, closest point) by the reference descent
    (src/capi.mojo:523-587) over build_aabb's tree."""
    P = np.ascontiguousarray(P, dtype=np.float64).reshape(-1, 3)
    V = np.ascontiguousarray(V, dtype=np.float64)
    F = np.ascontiguousarray(F, dtype=np.int64)
    n = len(P)
    sqrd = np.full(n, np.inf)
    idx = np.full(n, -1, dtype=np.int64)
    cl = np.zeros((n, 3))
    if len(F) == 0 or n == 0:
        return sqrd, idx, cl
    boxes, nodes, order, stack_size = tree if tree is not None else build_aabb(V, F)
    stack = np.empty(stack_size, dtype=np.int64)
    lib().wn_oracle_aabb_distance(P.ctypes.data, V.ctypes.data, F.ctypes.data, boxes.ctypes.data,
                                  nodes.ctypes.data, order.ctypes.data, sqrd.ctypes.data,
                                  idx.ctypes.data, cl.ctypes.data, stack.ctypes.data, len(nodes),
                                  0, n)
    return sqrd, idx, cl


def brute_distance(P, V, F):
    """(sqrd, face index, closest point): minimum of closest_on_triangle over
    every face in index order, first face on exact ties."""
    P = np.ascontiguousarray(P, dtype=np.float64).reshape(-1, 3)
    V = np.ascontiguousarray(V, dtype=np.float64)
    F = np.ascontiguousarray(F, dtype=np.int64)
    n = len(P)
    sqrd = np.empty(n)
    idx = np.empty(n, dtype=np.int64)
    cl = np.empty((n, 3))
    if n:
        lib().wn_oracle_brute_distance(P.ctypes.data, V.ctypes.data, F.ctypes.data, len(F),
                                       sqrd.ctypes.data, idx.ctypes.data, cl.ctypes.data, n)
    return sqrd, idx, cl


def closest_on_triangle(p, tri):
    """(d2, closest point) of point p to the triangle tri (3 x 3 corners)."""
    p = np.ascontiguousarray(p, dtype=np.float64).reshape(3)
    t = np.ascontiguousarray(tri, dtype=np.float64).reshape(9)
    out = np.empty(4)
    lib().wn_oracle_closest_on_triangle(p.ctypes.data, t.ctypes.data, out.ctypes.data)
    return float(out[0]), out[1:].copy()


# ---------------------------------------------------------------------------
# Hierarchical winding number (Barill et al. 2018, second order)
# ---------------------------------------------------------------------------

def fwn_expansion(tri_corners):
    """Second-order far-field expansion of a set of triangles ((k, 3, 3)
    corners): (centre, radius, N (3,), C (3, 3), D (3, 3, 3)) with centre =
    area-weighted centroid, N = sum A n, C_ij = sum A n_i ybar_j,
    D_ijk = sum A n_i (1/12)(sum_v y_vj y_vk + s_j s_k)."""
    T = np.asarray(tri_corners, dtype=np.float64).reshape(-1, 3, 3)
    an = 0.5 * np.cross(T[:, 1] - T[:, 0], T[:, 2] - T[:, 0])
    area = np.linalg.norm(an, axis=1)
    g = T.mean(axis=1)
    c = (area[:, None] * g).sum(0) / area.sum() if area.sum() > 0 else g.mean(0)
    y = T - c
    s = y.sum(axis=1)
    N = an.sum(0)
    C = np.einsum("ti,tj->ij", an, s / 3.0)
    M = (np.einsum("tvj,tvk->tjk", y, y) + np.einsum("tj,tk->tjk", s, s)) / 12.0
    D = np.einsum("ti,tjk->ijk", an, M)
    radius = np.sqrt((y ** 2).sum(-1).max())
    return c, radius, N, C, D


def fwn_far_omega(expansion, q):
    """Solid angle of the expansion's triangles seen from q (second order):
    N.f + sum J_ij C_ij + 1/2 sum H_ijk D_ijk at r = centre - q, for
    f = r/|r|^3 and its first and second derivatives."""
    c, _, N, C, D = expansion
    r = np.asarray(c, dtype=np.float64) - np.asarray(q, dtype=np.float64)
    n = np.sqrt(r @ r)
    I = np.eye(3)
    f = r / n ** 3
    J = I / n ** 3 - 3.0 * np.outer(r, r) / n ** 5
    H = (-3.0 * (np.einsum("ij,k->ijk", I, r) + np.einsum("ik,j->ijk", I, r)
                 + np.einsum("jk,i->ijk", I, r)) / n ** 5
         + 15.0 * np.einsum("i,j,k->ijk", r, r, r) / n ** 7)
    return float(N @ f + (J * C).sum() + 0.5 * (H * D).sum())


def exact_omega(tri_corners, q):
    """Exact solid angle sum (2 atan2 per face, src/capi.mojo:218-235)."""
    T = np.asarray(tri_corners, dtype=np.float64).reshape(-1, 3, 3) - np.asarray(q, dtype=np.float64)
    a, b, c = T[:, 0], T[:, 1], T[:, 2]
    la, lb, lc = (np.linalg.norm(x, axis=1) for x in (a, b, c))
    det = np.einsum("ij,ij->i", a, np.cross(b, c))
    den = la * lb * lc + (a * b).sum(1) * lc + (b * c).sum(1) * la + (c * a).sum(1) * lb
    return float((2.0 * np.arctan2(det, den)).sum())


def fast_winding(P, V, F, beta=2.0, leaf_size=8):
    """Reference descent of the hierarchical winding number over build_aabb's
    tree, in Python (small inputs): far nodes (|q - c| > beta R) by
    expansion, near leaves exactly."""
    V = np.asarray(V, dtype=np.float64)
    F = np.asarray(F, dtype=np.int64)
    P = np.asarray(P, dtype=np.float64).reshape(-1, 3)
    if len(F) == 0:
        return np.zeros(len(P))
    boxes, nodes, order, _ = build_aabb(V, F, leaf_size)
    corners = V[F[order]]

    def span(n):
        l, r, s, k = nodes[n]
        if k > 0:
            return s, s + k
        return span(l)[0], span(r)[1]

    exps = [fwn_expansion(corners[slice(*span(n))]) for n in range(len(nodes))]
    out = np.empty(len(P))
    for i, q in enumerate(P):
        total, stack = 0.0, [0]
        while stack:
            n = stack.pop()
            e = exps[n]
            if np.sum((q - e[0]) ** 2) > beta * beta * e[1] * e[1]:
                total += fwn_far_omega(e, q)
            elif nodes[n][3] > 0:
                s, k = nodes[n][2], nodes[n][3]
                total += exact_omega(corners[s:s + k], q)
            else:
                stack.append(int(nodes[n][1]))
                stack.append(int(nodes[n][0]))
        out[i] = total / (4.0 * np.pi)
    return out
