"""Hierarchical ("fast") winding number: ``FastWindingNumber`` and
``fast_winding_number``.

Not in the reference (its README.md:29 lists the fast winding number as not
covered). Method: Barill et al., "Fast Winding Numbers for Soups and Clouds"
(SIGGRAPH 2018), second order, over the median-split tree of ``igl.AABB``:
a node farther than ``beta`` times its radius contributes the Taylor
expansion of its faces' solid angle, nearer leaves contribute their faces
exactly. ``beta = 2`` is the published default; the error falls quickly as
``beta`` grows, and for ``beta`` beyond the mesh's extent over its smallest
leaf radius every face is summed exactly. Inputs are validated as
``igl.winding_number`` validates them.

    fw = FastWindingNumber(V, F)
    w = fw.query(Q)                     # CPU
    w = fw.query(Q, device="gpu")       # GPU (tree copied on first use)
    w = fast_winding_number(V, F, Q, beta=4.0, device="gpu")

A GPU request for an array of queries with no usable HIP device runs the CPU
walk (the rule of ``igl.winding_number``; ``WN_GPU_STRICT=1`` raises).
"""

from __future__ import annotations

import sys
import threading

import numpy as np

from ._lib import addr, fan_out, last_error, lib, worker_count

BETA = 2.0


def _is_cuda_tensor(value) -> bool:
    torch = sys.modules.get("torch")  # a CUDA tensor means torch is imported
    return torch is not None and isinstance(value, torch.Tensor) and value.is_cuda


class FastWindingNumber:
    """A mesh's tree of far-field expansions, on the host (and the GPU once
    queried there)."""

    def __init__(self, V, F, leaf_size: int = 8):
        from . import _mesh

        self._vertices, self._faces = _mesh(V, F, three_dimensional=True)
        self._leaf = int(leaf_size)
        self._host = self._create(0)
        self._gpu = 0

    def _create(self, upload: int) -> int:
        h = lib().wn_fwn_create(addr(self._vertices), addr(self._faces), len(self._vertices),
                                len(self._faces), 0, self._leaf, upload)
        if h <= 0:
            raise RuntimeError(f"wn_fwn_create failed: {last_error()}")
        return h

    def info(self) -> dict:
        out = np.zeros(4, dtype=np.int64)
        if lib().wn_fwn_info(self._gpu or self._host, out.ctypes.data) != 0:
            raise RuntimeError(last_error())
        return dict(zip(("faces", "nodes", "stack_size", "device_bytes"), (int(v) for v in out)))

    def expansions(self):
        """(expansions (n, 34), links (n, 4)) of every node, for inspection."""
        h = lib()
        n = h.wn_fwn_export(self._host, 0, 0, 0)
        e = np.empty((n, 34))
        links = np.empty((n, 4), dtype=np.int64)
        if n and h.wn_fwn_export(self._host, e.ctypes.data, links.ctypes.data, n) != n:
            raise RuntimeError(last_error())
        return e, links

    def query(self, Q, beta: float = BETA, device: str = "cpu", sort: int = -1):
        """w at the rows of Q ((q, 3) array-like -> (q,) float64), or at a
        (q, 3) CUDA tensor (-> CUDA tensor, current stream) with device="gpu".
        sort (GPU): -1 = walk queries in Morton order when q >= 4096, 1 =
        always, 0 = in the given order; the result order never changes."""
        from . import _float_array, _use_gpu

        beta = float(beta)
        if not beta > 0.0:
            raise ValueError("beta must be positive")
        selected = str(device).lower()
        if selected not in ("cpu", "gpu"):
            raise ValueError("device must be 'cpu' or 'gpu'")
        on_device = selected == "gpu" and _is_cuda_tensor(Q)
        # a CUDA tensor stays on its device; an array follows the winding
        # number's rule (no usable device: the CPU walk, same tree)
        if on_device or _use_gpu(selected == "gpu"):
            import torch

            if on_device:
                P = Q.to(torch.float64).contiguous()
            else:
                points = _float_array(Q, "Q")
                if points.ndim != 2 or points.shape[1] != 3:
                    raise ValueError("Q must have shape (q, 3)")
                P = torch.as_tensor(points, device="cuda")
            if not self._gpu:
                self._gpu = self._create(1)
            w = torch.empty(P.shape[0], dtype=torch.float64, device=P.device)
            nbytes = lib().wn_fwn_workspace_bytes(self._gpu, P.shape[0], int(sort))
            if nbytes < 0:
                raise RuntimeError(last_error())
            ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=P.device)
            st = lib().wn_fwn_query(self._gpu, P.data_ptr(), w.data_ptr(), P.shape[0], beta,
                                    int(sort), ws.data_ptr(), nbytes,
                                    torch.cuda.current_stream().cuda_stream)
            if st != 0:
                raise RuntimeError(f"wn_fwn_query failed on the GPU: {last_error()}")
            return w if on_device else w.cpu().numpy()
        points = _float_array(Q, "Q")
        if points.ndim != 2 or points.shape[1] != 3:
            raise ValueError("Q must have shape (q, 3)")
        w = np.empty(len(points), dtype=np.float64)
        if len(points):
            def span(lo, hi, span=0):
                st = lib().wn_fwn_query_cpu(self._host, addr(points), addr(w), len(points), beta,
                                            lo, hi)
                if st != 0:
                    raise RuntimeError(last_error())

            work = len(points) * max(len(self._faces).bit_length(), 1) * 64
            fan_out(span, (), len(points), worker_count(len(points), work, "winding"))
        return w

    def close(self) -> None:
        for name in ("_gpu", "_host"):
            h = getattr(self, name, 0)
            if h:
                lib().wn_fwn_destroy(h)
                setattr(self, name, 0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# per thread: the tree of the last fast_winding_number call and the bytes of
# the V and F it was built from
_kept = threading.local()


def fast_winding_number(V, F, Q, beta: float = BETA, device: str = "cpu"):
    """Hierarchical winding number of (V, F) at the rows of Q.

    The tree of the last call (per thread) is kept and reused when V and F
    are byte-for-byte equal (the build is deterministic, so the results are
    those of a new tree); release() drops it."""
    from . import _mesh

    vertices, faces = _mesh(V, F, three_dimensional=True)
    kept = getattr(_kept, "tree", None)
    if (kept is None or kept._vertices.shape != vertices.shape
            or kept._faces.shape != faces.shape
            or not np.array_equal(kept._vertices.view(np.uint64), vertices.view(np.uint64))
            or not np.array_equal(kept._faces, faces)):
        _kept.tree = None
        # private copies: the caller may change its arrays in place later
        kept = FastWindingNumber(vertices.copy(), faces.copy())
        _kept.tree = kept
    return kept.query(Q, beta=beta, device=device)


def release() -> None:
    """Drops the tree this thread's fast_winding_number keeps."""
    _kept.tree = None


def fast_winding_number_for_meshes(V, F, Q):
    """libigl's name for the mesh variant (accuracy 2), CPU."""
    return fast_winding_number(V, F, Q)
