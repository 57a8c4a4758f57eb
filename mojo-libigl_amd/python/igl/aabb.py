"""Point-mesh squared distance: ``AABB`` and ``point_mesh_squared_distance``.

Mirrors python/igl/__init__.py:322-404 of the reference (same arguments,
results and errors): ``AABB.init(V, Ele)`` builds the median-split tree
(here in the library, ``wn_aabb_build``, instead of NumPy);
``AABB.squared_distance(V, Ele, P)`` returns ``(sqrd (p,), face index (p,)
int64, closest point (p, 3))``. ``device="gpu"`` runs the same descent on the
GPU (``wn_aabb_query``) over a device copy of the tree made on first use;
its results equal the CPU path's bit for bit. With no usable HIP device a
GPU request runs the CPU path (the rule of ``igl.winding_number``, reference
python/igl/__init__.py:251,264; ``WN_GPU_STRICT=1`` raises instead).
"""

from __future__ import annotations

import threading

import numpy as np

from ._lib import addr, fan_out, last_error, lib, worker_count

LEAF_SIZE = 8


def _build_tree(vertices: np.ndarray, faces: np.ndarray, leaf_size: int = LEAF_SIZE):
    """(boxes (n, 6), nodes (n, 4), order (m,), stack size) of the reference's
    tree (python/igl/__init__.py:286-319), built by wn_aabb_build."""
    h = lib()
    args = (addr(vertices), addr(faces), len(vertices), len(faces), 0, leaf_size)
    count = h.wn_aabb_build(*args, 0, 0, 0, 0, 0)
    if count < 0:
        raise RuntimeError(f"wn_aabb_build failed: {last_error()}")
    boxes = np.empty((count, 6), dtype=np.float64)
    nodes = np.empty((count, 4), dtype=np.int64)
    order = np.empty(len(faces), dtype=np.int64)
    stack = np.zeros(1, dtype=np.int64)
    got = h.wn_aabb_build(*args, addr(boxes), addr(nodes), addr(order), count, addr(stack))
    if got != count:
        raise RuntimeError(f"wn_aabb_build failed: {last_error()}")
    return boxes, nodes, order, int(stack[0])


class AABB:
    """libigl-compatible AABB precomputation for squared-distance queries
    (python/igl/__init__.py:322-395)."""

    def __init__(self):
        self._tree = None
        self._gpu = 0

    def init(self, V, Ele):
        from . import _mesh

        vertices, faces = _mesh(V, Ele, three_dimensional=True)
        self._release_gpu()
        self._vertices = vertices
        self._faces = faces
        self._tree = _build_tree(vertices, faces) if len(faces) else ()

    def _checked(self, V, Ele, P):
        from . import _float_array, _mesh

        if self._tree is None:
            raise RuntimeError("AABB.init must be called before squared_distance")
        vertices, faces = _mesh(V, Ele, three_dimensional=True)
        if not np.array_equal(vertices, self._vertices) or not np.array_equal(faces, self._faces):
            raise ValueError("V and Ele must match the mesh passed to AABB.init")
        points = _float_array(P, "P")
        if points.ndim != 2 or points.shape[1] != 3:
            raise ValueError("P must have shape (p, 3)")
        return points

    def squared_distance(self, V, Ele, P, device="cpu"):
        points = self._checked(V, Ele, P)
        selected = str(device).lower()
        if selected not in ("cpu", "gpu"):
            raise ValueError("device must be 'cpu' or 'gpu'")
        if selected == "gpu":
            from . import _use_gpu

            if _use_gpu(True):
                return self._squared_distance_gpu(points)
        vertices, faces = self._vertices, self._faces
        count = len(points)
        if not len(faces):
            return (np.full(count, np.inf), np.full(count, -1, dtype=np.int64),
                    np.zeros((count, 3)))
        boxes, nodes, order, stack_size = self._tree
        sqrd = np.empty(count, dtype=np.float64)
        indices = np.empty(count, dtype=np.int64)
        closest = np.empty((count, 3), dtype=np.float64)
        if count:
            # one traversal stack per query span (python/igl/__init__.py:370-390)
            depth = 1 + max(stack_size.bit_length(), 1)
            workers = worker_count(count, count * len(faces) * depth, "aabb")
            stacks = np.empty((workers, stack_size), dtype=np.int64)

            def query(lo, hi, span=0):
                lib().mli_aabb_squared_distance(
                    addr(points), addr(vertices), addr(faces), addr(boxes), addr(nodes),
                    addr(order), addr(sqrd), addr(indices), addr(closest),
                    stacks[span].ctypes.data, count, len(nodes), lo, hi)

            fan_out(query, (), count, workers)
        return sqrd, indices, closest

    def _squared_distance_gpu(self, points: np.ndarray):
        import torch

        P = torch.as_tensor(points, device="cuda")
        return tuple(t.cpu().numpy() for t in self.query_device(P))

    def _ensure_gpu(self) -> int:
        if self._tree is None:
            raise RuntimeError("AABB.init must be called before squared_distance")
        if not self._gpu:
            handle = lib().wn_aabb_create(addr(self._vertices), addr(self._faces),
                                          len(self._vertices), len(self._faces), 0, LEAF_SIZE)
            if handle <= 0:
                raise RuntimeError(f"wn_aabb_create failed: {last_error()}")
            self._gpu = handle
        return self._gpu

    def query_device(self, P, sort=-1):
        """(sqrd, face index, closest point) as CUDA tensors for a (p, 3) CUDA
        tensor P, on the current stream, without host copies. The device
        tree is made on first use and kept until init() or deletion."""
        import torch

        h = lib()
        self._ensure_gpu()
        P = P.to(dtype=torch.float64).contiguous()
        n = P.shape[0]
        sqrd = torch.empty(n, dtype=torch.float64, device=P.device)
        idx = torch.empty(n, dtype=torch.int64, device=P.device)
        cl = torch.empty((n, 3), dtype=torch.float64, device=P.device)
        nbytes = h.wn_aabb_workspace_bytes(self._gpu, n, sort)
        if nbytes < 0:
            raise RuntimeError(last_error())
        ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=P.device)
        st = h.wn_aabb_query(self._gpu, P.data_ptr(), sqrd.data_ptr(), idx.data_ptr(),
                             cl.data_ptr(), n, sort, ws.data_ptr(), nbytes,
                             torch.cuda.current_stream().cuda_stream)
        if st != 0:
            raise RuntimeError(f"wn_aabb_query failed on the GPU: {last_error()}")
        return sqrd, idx, cl

    def gpu_info(self) -> dict:
        out = np.zeros(4, dtype=np.int64)
        if not self._gpu or lib().wn_aabb_info(self._gpu, out.ctypes.data) != 0:
            raise RuntimeError("no device tree")
        return dict(zip(("faces", "nodes", "stack_size", "device_bytes"), (int(v) for v in out)))

    def find(self, V, Ele, q, first=False):
        raise NotImplementedError("AABB.find is outside the covered subset")

    def intersect_ray(self, V, Ele, orig, dir):
        raise NotImplementedError("AABB ray intersection is outside the covered subset")

    def intersect_ray_first(self, V, Ele, orig, dir, min_t=np.inf):
        raise NotImplementedError("AABB ray intersection is outside the covered subset")

    def _release_gpu(self) -> None:
        if getattr(self, "_gpu", 0):
            lib().wn_aabb_destroy(self._gpu)
            self._gpu = 0

    def __del__(self):
        try:
            self._release_gpu()
        except Exception:
            pass


# per thread: the tree of the last point_mesh_squared_distance call, built
# from private copies of its V and Ele
_kept = threading.local()


def _same_bytes(a: np.ndarray, b: np.ndarray) -> bool:
    return a.shape == b.shape and a.dtype == b.dtype and np.array_equal(
        a.view(np.uint64) if a.dtype == np.float64 else a,
        b.view(np.uint64) if b.dtype == np.float64 else b)


def point_mesh_squared_distance(P, V, Ele, device="cpu"):
    """(sqrd, face index, closest point) of each row of P to the mesh
    (python/igl/__init__.py:398-401).

    The tree of the last call (per thread) is kept and reused when V and Ele
    are byte-for-byte equal (the build is deterministic, so the results are
    those of a new tree); release() drops it."""
    from . import _mesh

    vertices, faces = _mesh(V, Ele, three_dimensional=True)
    tree = getattr(_kept, "tree", None)
    if (tree is None or not _same_bytes(tree._vertices, vertices)
            or not _same_bytes(tree._faces, faces)):
        _kept.tree = None
        tree = AABB()
        tree.init(vertices.copy(), faces.copy())  # the caller may change its arrays later
        _kept.tree = tree
    return tree.squared_distance(tree._vertices, tree._faces, P, device=device)


def release() -> None:
    """Drops the tree this thread's point_mesh_squared_distance keeps."""
    _kept.tree = None


def signed_distance(P, V, F, device="cpu", winding="exact", beta=2.0):
    """(S, face index, closest point): S = sign * sqrt(squared distance) with
    sign -1 where the winding number is >= 1/2 (inside) and +1 elsewhere.

    No counterpart in the reference (it has the distance and the winding
    number separately, python/igl/__init__.py:232-283, 322-401). The
    distance is point_mesh_squared_distance's; the winding number is the
    exact one (``winding="exact"``; FP64) or the hierarchical one
    (``winding="fast"``, accuracy ``beta``), on the same device."""
    from . import _float_array, winding_number
    from .fast_winding import fast_winding_number

    points = _float_array(P, "P")
    if points.ndim != 2 or points.shape[1] != 3:
        raise ValueError("P must have shape (p, 3)")
    mode = str(winding).lower()
    if mode not in ("exact", "fast"):
        raise ValueError("winding must be 'exact' or 'fast'")
    sqrd, index, closest = point_mesh_squared_distance(points, V, F, device=device)
    if mode == "exact":
        w = winding_number(V, F, points, device=device) if len(points) else np.zeros(0)
    else:
        w = fast_winding_number(V, F, points, beta=beta, device=device)
    sign = np.where(np.asarray(w) >= 0.5, -1.0, 1.0)
    return sign * np.sqrt(sqrd), index, closest
