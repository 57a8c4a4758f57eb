"""Device-resident mesh for repeated winding-number queries.

``Mesh(V, F)`` converts the mesh once into triangle-strip records (host C++,
wn_mesh_create) and keeps them in HBM; ``Mesh.winding_number(O)`` then runs
the FP32-compensated strip kernel (wn_mesh_forward), which evaluates one
vertex offset, length and square root per face instead of three. Same
accuracy contract as ``igl.winding_number(..., precision="fp32c")``.

    m = igl.Mesh(V, F)
    w = m.winding_number(O)        # numpy in -> numpy out, or a CUDA tensor
    m.close()
"""

from __future__ import annotations

import numpy as np

from ._lib import addr, last_error, lib


def strip_records(V, F) -> np.ndarray:
    """(records, 4) float64 array (x, y, z, flag) of mesh (V, F), computed on
    the host by the library's stripifier."""
    from . import _mesh

    vertices, faces = _mesh(V, F, three_dimensional=True)
    h = lib()
    # one stripification when the record count fits the first guess (about
    # 1.06 records per face on meshes of long strips), a second otherwise
    out = np.empty((len(faces) + len(faces) // 4 + 2 * 255, 4), dtype=np.float64)
    n = h.wn_mesh_records(addr(vertices), addr(faces), len(vertices), len(faces), 0, addr(out),
                          len(out))
    if n < 0:
        raise ValueError(last_error())
    if n > len(out):
        out = np.empty((n, 4), dtype=np.float64)
        if h.wn_mesh_records(addr(vertices), addr(faces), len(vertices), len(faces), 0, addr(out),
                             n) != n:
            raise RuntimeError(last_error())
    return out[:n]


def strip_topology(F, n_vertices: int):
    """(vertex id per record int32, flag per record int8) of F's strips
    (wn_strip_topology)."""
    from . import _faces

    faces = _faces(F)
    h = lib()
    cap = len(faces) + len(faces) // 4 + 2 * 255  # one call when it fits
    idx = np.empty(cap, dtype=np.int32)
    flag = np.empty(cap, dtype=np.int8)
    n = h.wn_strip_topology(addr(faces), int(n_vertices), len(faces), 0, idx.ctypes.data,
                            flag.ctypes.data, cap)
    if n < 0:
        raise (IndexError if "outside V" in last_error() else ValueError)(last_error())
    if n > cap:
        idx = np.empty(n, dtype=np.int32)
        flag = np.empty(n, dtype=np.int8)
        if h.wn_strip_topology(addr(faces), int(n_vertices), len(faces), 0, idx.ctypes.data,
                               flag.ctypes.data, n) != n:
            raise RuntimeError(last_error())
    return idx[:n], flag[:n]


def strip_boundary(F, n_vertices: int, idx) -> tuple[np.ndarray, np.ndarray]:
    """Net directed boundary edges of F as record pairs of the strip topology
    idx ((count, 2) int32) and their weights ((count,) float64)
    (wn_strip_boundary)."""
    from . import _faces

    faces = _faces(F)
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    h = lib()
    cap = max(16, len(faces) // 8)
    for _ in range(2):
        pairs = np.empty((cap, 2), dtype=np.int32)
        weight = np.empty(cap, dtype=np.float64)
        n = h.wn_strip_boundary(addr(faces), int(n_vertices), len(faces), 0, idx.ctypes.data,
                                len(idx), pairs.ctypes.data, weight.ctypes.data, cap)
        if n < 0:
            raise (IndexError if "outside V" in last_error() else ValueError)(last_error())
        if n <= cap:
            return pairs[:n], weight[:n]
        cap = n
    raise RuntimeError("wn_strip_boundary: edge count changed between calls")


def strip_records_boundary(V, F):
    """Strip records of (V, F) as strip_records gives them, with the mesh's
    boundary edges as record pairs and weights (the FP32-compensated boundary
    correction of a strip batch plan)."""
    from . import _mesh

    vertices, faces = _mesh(V, F, three_dimensional=True)
    idx, flag = strip_topology(faces, len(vertices))
    rec = np.empty((len(idx), 4), dtype=np.float64)
    rec[:, :3] = vertices[idx]
    rec[:, 3] = flag
    pairs, weight = strip_boundary(faces, len(vertices), idx)
    return rec, pairs, weight


class Mesh:
    """A mesh resident on the GPU in strip-record form."""

    def __init__(self, V, F):
        from . import _mesh

        vertices, faces = _mesh(V, F, three_dimensional=True)
        h = lib()
        handle = h.wn_mesh_create(addr(vertices), addr(faces), len(vertices), len(faces), 0)
        if handle <= 0:
            raise RuntimeError(f"wn_mesh_create failed: {last_error()}")
        self.handle = handle
        self.n_faces = len(faces)
        self.n_vertices = len(vertices)

    @classmethod
    def from_faces(cls, F, n_vertices: int):
        """A mesh from its faces alone; give it vertex positions with
        set_vertices (device tensors) before querying."""
        from . import _faces, _integer

        faces = _faces(F)
        nv = _integer(n_vertices, "n_vertices")
        if faces.size and (faces.min() < 0 or faces.max() >= nv):
            raise IndexError("face index is outside V")
        self = cls.__new__(cls)
        handle = lib().wn_mesh_create_topology(addr(faces), nv, len(faces), 0)
        if handle <= 0:
            raise RuntimeError(f"wn_mesh_create_topology failed: {last_error()}")
        self.handle = handle
        self.n_faces = len(faces)
        self.n_vertices = nv
        return self

    def set_vertices(self, V) -> None:
        """New vertex positions from a (nv, 3) float32/float64 CUDA tensor, on
        the current stream (device gather of the strip records; no
        re-stripification, no host copy)."""
        import torch

        if not (isinstance(V, torch.Tensor) and V.is_cuda):
            raise TypeError("V must be a CUDA tensor")
        if V.shape != (self.n_vertices, 3) or V.dtype not in (torch.float32, torch.float64):
            raise ValueError("V must be a (n_vertices, 3) float32 or float64 tensor")
        V = V.detach().contiguous()
        st = lib().wn_mesh_set_vertices(self.handle, V.data_ptr(), self.n_vertices,
                                        0 if V.dtype == torch.float64 else 1,
                                        torch.cuda.current_stream().cuda_stream)
        if st != 0:
            raise RuntimeError(f"wn_mesh_set_vertices failed: {last_error()}")
        self._keep = V  # the gather may still be reading it on the stream

    def info(self) -> dict:
        out = np.zeros(3, dtype=np.int64)
        if lib().wn_mesh_info(self.handle, out.ctypes.data) != 0:
            raise RuntimeError(last_error())
        return {"faces": int(out[0]), "records": int(out[1]), "tiles": int(out[2])}

    def resolve(self, nq: int, splits=0, sort=-1):
        out = np.zeros(2, dtype=np.int64)
        if lib().wn_mesh_resolve(self.handle, nq, splits, sort, out.ctypes.data) != 0:
            raise RuntimeError(last_error())
        return int(out[0]), int(out[1])

    def winding_number(self, O, splits=0, sort=-1, kappa=-1.0, out=None, precision="fp32c"):
        """w at the rows of O. O: (q, 3) array-like (result numpy float64) or
        a CUDA tensor (result a float64 CUDA tensor on the current stream).
        precision "fp32c" (FP32-compensated strip kernel) or "fp64" (FP64
        strip kernel, wn_mesh_forward64; kappa does not apply)."""
        import torch

        precision = str(precision).lower()
        if precision not in ("fp32c", "fp64"):
            raise ValueError("precision must be 'fp32c' or 'fp64'")
        host = not (isinstance(O, torch.Tensor) and O.is_cuda)
        Q = torch.as_tensor(np.asarray(O, dtype=np.float64).reshape(-1, 3) if host else O)
        Q = Q.to(device="cuda", dtype=torch.float64).contiguous()
        nq = Q.shape[0]
        splits, sort = self.resolve(nq, splits, sort)
        nbytes = lib().wn_mesh_workspace_bytes(self.handle, nq, splits, sort)
        if nbytes < 0:
            raise RuntimeError(last_error())
        ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=Q.device)
        w = out if out is not None else torch.empty(nq, dtype=torch.float64, device=Q.device)
        stream = torch.cuda.current_stream().cuda_stream
        if precision == "fp64":
            st = lib().wn_mesh_forward64(self.handle, Q.data_ptr(), w.data_ptr(), nq, splits, sort,
                                         ws.data_ptr(), nbytes, stream)
        else:
            st = lib().wn_mesh_forward(self.handle, Q.data_ptr(), w.data_ptr(), nq, splits, sort,
                                       float(kappa), ws.data_ptr(), nbytes, stream)
        if st != 0:
            raise RuntimeError(f"wn_mesh_forward failed on the GPU: {last_error()}")
        return w.cpu().numpy() if host else w

    def close(self) -> None:
        if getattr(self, "handle", 0):
            lib().wn_mesh_destroy(self.handle)
            self.handle = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
