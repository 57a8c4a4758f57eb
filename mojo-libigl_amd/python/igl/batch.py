"""Batched winding numbers over many meshes of uneven sizes.

One persistent HIP launch (optionally replayed from a hipGraph) evaluates
every (mesh, query) pair set of the batch; see wn_batch_create in
include/wn.h. Device tensors come from PyTorch; the work happens in the C ABI.

    plan = WindingBatch(face_counts, query_counts, precision="fp32c")
    tri = pack_batch([V0, V1, ...], [F0, F1, ...])      # 9 x F_total, device
    w = plan.run(tri, Q_all)                             # (Q_total,) float64
    plan.close()

``winding_number_batch`` wraps the whole sequence for host arrays.

Strip plans (``StripBatch``) take the meshes as triangle-strip records built
once on the host (``pack_batch_strips``) and run the strip kernel, which
evaluates one vertex per face instead of three:

    rec, counts, boundary = pack_batch_strips([V0, V1, ...], [F0, F1, ...], boundary=True)
    plan = StripBatch(counts, query_counts, boundary=boundary)
    w = plan.run(rec, Q_all)
"""

from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import PRECISIONS
from ._lib import addr, last_error, lib


def _mode(precision) -> int:
    return PRECISIONS[str(precision).lower()] if isinstance(precision, str) else int(precision)


class WindingBatch:
    """A planned batch: offsets, work list and device workspace."""

    def __init__(self, face_counts, query_counts, precision="fp32c", qpl=1, chunk_faces=2048,
                 sort=-1, kappa=-1.0):
        fc = np.asarray(face_counts, dtype=np.int64)
        qc = np.asarray(query_counts, dtype=np.int64)
        if fc.ndim != 1 or fc.shape != qc.shape or len(fc) == 0:
            raise ValueError("face_counts and query_counts must be equal-length 1-D sequences")
        if (fc < 0).any() or (qc < 0).any():
            raise ValueError("counts must be non-negative")
        self.face_off = np.concatenate([[0], np.cumsum(fc)]).astype(np.int64)
        self.query_off = np.concatenate([[0], np.cumsum(qc)]).astype(np.int64)
        self.precision = _mode(precision)
        self.config = dict(qpl=int(qpl), chunk_faces=int(chunk_faces), sort=int(sort), kappa=float(kappa))
        handle = lib().wn_batch_create(self.face_off.ctypes.data, self.query_off.ctypes.data,
                                       len(fc), self.precision, int(qpl), int(chunk_faces),
                                       int(sort), float(kappa))
        if handle <= 0:
            raise RuntimeError(f"wn_batch_create failed: {last_error()}")
        self.handle = handle

    @property
    def n_faces(self) -> int:
        return int(self.face_off[-1])

    @property
    def n_queries(self) -> int:
        return int(self.query_off[-1])

    def info(self) -> dict:
        out = np.zeros(6, dtype=np.int64)
        if lib().wn_batch_info(self.handle, out.ctypes.data) != 0:
            raise RuntimeError(last_error())
        keys = ("items", "partials", "workspace_bytes", "grid", "faces", "queries")
        return dict(zip(keys, (int(v) for v in out)))

    def _check_faces(self, tri) -> None:
        import torch

        if tri.shape != (9, self.n_faces) or tri.dtype != torch.float64 or not tri.is_contiguous():
            raise ValueError("tri must be a contiguous (9, F_total) float64 tensor")

    def run(self, tri, Q, out=None, use_graph=True, stream=None):
        """w for all queries. tri: (9, F_total) float64 device tensor (the
        records of a strip plan), Q: (Q_total, 3) float64 device tensor."""
        import torch

        self._check_faces(tri)
        if Q.shape != (self.n_queries, 3) or Q.dtype != torch.float64 or not Q.is_contiguous():
            raise ValueError("Q must be a contiguous (Q_total, 3) float64 tensor")
        if out is None:
            out = torch.empty(self.n_queries, dtype=torch.float64, device=Q.device)
        s = (stream or torch.cuda.current_stream()).cuda_stream
        st = lib().wn_batch_run(self.handle, tri.data_ptr(), Q.data_ptr(), out.data_ptr(),
                                1 if use_graph else 0, s)
        if st != 0:
            raise RuntimeError(f"wn_batch_run failed on the GPU: {last_error()}")
        return out

    def close(self) -> None:
        if getattr(self, "handle", 0):
            lib().wn_batch_destroy(self.handle)
            self.handle = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchBoundary:
    """Boundary edges of a strip batch's meshes on the device: record pairs
    ((E, 2) int32, indices into the batch's records), weights ((E,)
    float64) and per-mesh edge counts (host). The FP32-compensated strip plan
    adds their boundary correction (wn_batch_set_boundary)."""

    def __init__(self, pairs, weight, edge_counts):
        self.pairs = pairs
        self.weight = weight
        self.edge_counts = np.asarray(edge_counts, dtype=np.int64)
        self.edge_off = np.concatenate([[0], np.cumsum(self.edge_counts)]).astype(np.int64)

    @classmethod
    def from_host(cls, pairs_list, weight_list, record_counts, device="cuda"):
        """Per-mesh (pairs, weight) with mesh-local record indices ->
        one device list with batch record indices."""
        import torch

        off = np.concatenate([[0], np.cumsum(np.asarray(record_counts, dtype=np.int64))])
        pairs = [np.asarray(p, dtype=np.int64) + off[k] for k, p in enumerate(pairs_list)]
        if off[-1] > np.iinfo(np.int32).max:
            raise ValueError("more than 2^31-1 records in the batch")
        cat = (np.concatenate(pairs).astype(np.int32) if pairs else np.zeros((0, 2), np.int32))
        w = np.concatenate(weight_list) if weight_list else np.zeros(0)
        return cls(torch.as_tensor(np.ascontiguousarray(cat.reshape(-1, 2)), device=device),
                   torch.as_tensor(np.ascontiguousarray(w, dtype=np.float64), device=device),
                   [len(p) for p in pairs_list])

    @property
    def n_edges(self) -> int:
        return int(self.edge_off[-1])


class StripBatch(WindingBatch):
    """A planned batch over meshes in strip-record form: FP32-compensated, or
    FP64 with precision="fp64" (the FP64 strip kernel; kappa does not
    apply). ``boundary`` (a BatchBoundary, from pack_batch_strips(...,
    boundary=True) or BatchTopology.boundary) gives the FP32-compensated plan
    its open meshes' boundary correction; without it open meshes keep the
    FP32 rounding of their boundary (errors up to a few 1e-6)."""

    def __init__(self, record_counts, query_counts, chunk_tiles=0, sort=-1, kappa=-1.0,
                 precision="fp32c", boundary=None):
        rc = np.asarray(record_counts, dtype=np.int64)
        qc = np.asarray(query_counts, dtype=np.int64)
        if rc.ndim != 1 or rc.shape != qc.shape or len(rc) == 0:
            raise ValueError("record_counts and query_counts must be equal-length 1-D sequences")
        if (rc < 0).any() or (qc < 0).any() or (rc % 255).any():
            raise ValueError("record counts must be non-negative multiples of 255")
        self.face_off = np.concatenate([[0], np.cumsum(rc)]).astype(np.int64)
        self.query_off = np.concatenate([[0], np.cumsum(qc)]).astype(np.int64)
        precision = str(precision).lower()
        if precision not in ("fp32c", "fp64"):
            raise ValueError("precision must be 'fp32c' or 'fp64'")
        self.precision = PRECISIONS[precision]
        self.config = dict(chunk_tiles=int(chunk_tiles), sort=int(sort), kappa=float(kappa),
                           precision=precision)
        if precision == "fp64":
            handle = lib().wn_batch_create_strips64(self.face_off.ctypes.data,
                                                    self.query_off.ctypes.data, len(rc),
                                                    int(chunk_tiles), int(sort))
        else:
            handle = lib().wn_batch_create_strips(self.face_off.ctypes.data,
                                                  self.query_off.ctypes.data, len(rc),
                                                  int(chunk_tiles), int(sort), float(kappa))
        if handle <= 0:
            raise RuntimeError(f"wn_batch_create_strips failed: {last_error()}")
        self.handle = handle
        self.boundary = None
        if boundary is not None:
            self.set_boundary(boundary)

    def set_boundary(self, boundary) -> None:
        """Boundary edges of the meshes (a BatchBoundary with one edge count
        per mesh); the plan keeps a reference while it lives."""
        if len(boundary.edge_counts) != len(self.query_off) - 1:
            raise ValueError("boundary must have one edge count per mesh")
        st = lib().wn_batch_set_boundary(self.handle, boundary.pairs.data_ptr(),
                                         boundary.weight.data_ptr(), boundary.n_edges,
                                         boundary.edge_off.ctypes.data)
        if st != 0:
            raise RuntimeError(f"wn_batch_set_boundary failed: {last_error()}")
        self.boundary = boundary

    @property
    def n_records(self) -> int:
        return int(self.face_off[-1])

    def _check_faces(self, rec) -> None:
        import torch

        if rec.shape != (self.n_records, 4) or rec.dtype != torch.float64 or not rec.is_contiguous():
            raise ValueError("records must be a contiguous (R_total, 4) float64 tensor")


def _workers(n: int) -> int:
    return max(1, min(n, 16, os.cpu_count() or 1))


def pack_batch_strips(vertices, faces, device="cuda", boundary=False):
    """Strip records of every mesh (host stripifier), concatenated on the
    device: ((R_total, 4) float64 tensor, per-mesh record counts), and with
    boundary=True also the meshes' BatchBoundary."""
    import torch

    from .mesh import strip_records, strip_records_boundary

    # meshes stripify independently; the library calls release the GIL
    with ThreadPoolExecutor(max_workers=_workers(len(vertices))) as pool:
        if boundary:
            parts = list(pool.map(strip_records_boundary, vertices, faces))
            recs = [p[0] for p in parts]
        else:
            recs = list(pool.map(strip_records, vertices, faces))
    counts = np.array([len(r) for r in recs], dtype=np.int64)
    rec = np.concatenate(recs) if recs else np.zeros((0, 4))
    rec = torch.as_tensor(np.ascontiguousarray(rec), device=device)
    if not boundary:
        return rec, counts
    return rec, counts, BatchBoundary.from_host([p[1] for p in parts], [p[2] for p in parts],
                                                counts, device=device)


class BatchTopology:
    """Strip topology of many meshes, built once on the host and kept on the
    device: ``records(V_all)`` gathers a StripBatch's records from all
    meshes' vertices concatenated on the device (new positions, same
    topology) without re-stripifying.

        topo = BatchTopology([F0, F1, ...], [nv0, nv1, ...])
        plan = StripBatch(topo.record_counts, query_counts, boundary=topo.boundary)
        w = plan.run(topo.records(V_all), Q_all)
    """

    def __init__(self, faces, vertex_counts, device="cuda"):
        import torch

        from . import _faces
        from .mesh import strip_boundary, strip_topology

        def topology(F, nv):
            Fa = _faces(F)
            idx, flag = strip_topology(Fa, nv)
            pairs, weight = strip_boundary(Fa, nv, idx)
            return idx, flag, pairs, weight

        nvs = [int(nv) for nv in vertex_counts]
        with ThreadPoolExecutor(max_workers=_workers(len(nvs))) as pool:
            tops = list(pool.map(topology, faces, nvs))
        ids, flags, counts, off = [], [], [], 0
        for (idx, flag, _, _), nv in zip(tops, nvs):
            ids.append(idx.astype(np.int64) + off)
            flags.append(flag)
            counts.append(len(idx))
            off += nv
        if off > np.iinfo(np.int32).max:
            raise ValueError("more than 2^31-1 vertices in the batch")
        self.n_vertices = off
        self.record_counts = np.asarray(counts, dtype=np.int64)
        cat = np.concatenate(ids).astype(np.int32) if ids else np.zeros(0, np.int32)
        self.idx = torch.as_tensor(cat, device=device)
        self.flag = torch.as_tensor(np.concatenate(flags) if flags else np.zeros(0, np.int8),
                                    device=device)
        # boundary edges as record pairs: they follow the records, so moved
        # vertices need no new boundary
        self.boundary = BatchBoundary.from_host([t[2] for t in tops], [t[3] for t in tops],
                                                self.record_counts, device=device)

    def records(self, V_all, out=None):
        """(R_total, 4) float64 records for (n_vertices, 3) float32/float64
        device vertices, gathered on the current stream."""
        import torch

        if V_all.shape != (self.n_vertices, 3) or V_all.dtype not in (torch.float32, torch.float64):
            raise ValueError("V_all must be a (n_vertices, 3) float32 or float64 tensor")
        V = V_all.detach().contiguous()
        n = int(self.idx.shape[0])
        rec = out if out is not None else torch.empty((n, 4), dtype=torch.float64, device=V.device)
        st = lib().wn_strip_gather(V.data_ptr(), self.n_vertices,
                                   0 if V.dtype == torch.float64 else 1, self.idx.data_ptr(),
                                   self.flag.data_ptr(), n, rec.data_ptr(),
                                   torch.cuda.current_stream().cuda_stream)
        if st != 0:
            raise RuntimeError(f"wn_strip_gather failed: {last_error()}")
        return rec


def pack_batch(vertices, faces, device="cuda"):
    """Concatenate meshes (V_i, F_i) and gather them on the device into one
    9 x F_total float64 SoA (indices offset per mesh)."""
    import torch

    from .autograd import pack_faces

    Vs, Fs, off = [], [], 0
    for V, F in zip(vertices, faces):
        Vt = torch.as_tensor(V, device=device)
        Ft = torch.as_tensor(F, device=device).to(torch.int64)
        Vs.append(Vt.to(torch.float64))
        Fs.append(Ft + off)
        off += Vt.shape[0]
    return pack_faces(torch.cat(Vs), torch.cat(Fs))


def winding_number_batch(vertices, faces, queries, precision="fp32c", use_graph=False,
                         strips=False, **config):
    """Host convenience: lists of (V_i, F_i, O_i) -> list of w_i (numpy).
    strips=True runs a strip plan (precision "fp32c" or "fp64")."""
    import torch

    Q = torch.as_tensor(np.concatenate([np.asarray(o, dtype=np.float64).reshape(-1, 3)
                                        for o in queries]), device="cuda")
    qc = [len(np.asarray(o).reshape(-1, 3)) for o in queries]
    if strips:
        tri, counts, boundary = pack_batch_strips(vertices, faces, boundary=True)
        plan = StripBatch(counts, qc, precision=precision, boundary=boundary, **config)
    else:
        tri = pack_batch(vertices, faces)
        plan = WindingBatch([len(f) for f in faces], qc, precision=precision, **config)
    try:
        w = plan.run(tri, Q.contiguous(), use_graph=use_graph).cpu().numpy()
    finally:
        plan.close()
    return [w[a:b] for a, b in zip(plan.query_off[:-1], plan.query_off[1:])]
