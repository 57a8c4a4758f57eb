"""PyTorch autograd layer over the HIP winding-number kernels.

Tensors stay on the GPU; every computation is a C-ABI call (include/wn.h) on
PyTorch's current stream, with PyTorch only allocating device memory:

* ``pack_faces(V, F)``      -> wn_pack_soa (device gather of V[F], 9 x m f64)
* ``forward(tri, O)``       -> wn_forward (FP64 plain or FP32-compensated)
* ``WindingNumber.apply``   -> forward, then the FP64 adjoint for dL/dO and
  dL/dV: by default wn_backward_edges over the mesh's net boundary edges
  (wn_boundary_edges; interior edges cancel, a closed mesh has none), or
  with ``adjoint="faces"`` wn_backward + wn_scatter_corners over every face
  (face tiles recomputed). Both are the exact gradient; they agree to
  rounding.

``winding_number(V, F, O, precision="fp64")`` is the differentiable entry
point; ``WindingField(F, nv)`` serves repeated calls on one topology with the
FP32-compensated strip kernel forward. V may be float32 or float64, F int32 or int64, O float32 or float64;
w is float64 and gradients come back in the dtype of their input.
"""

from __future__ import annotations

import threading

import numpy as np
import torch

from . import PRECISIONS
from ._lib import addr, last_error, lib

_INT32_MAX = 2**31 - 1


def _check(status: int, what: str) -> None:
    if status != 0:
        raise RuntimeError(f"{what} failed on the GPU: {last_error()}")


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _require_cuda(*tensors) -> None:
    for t in tensors:
        if not t.is_cuda:
            raise ValueError("igl.autograd expects CUDA/HIP tensors")


def pack_faces(V: torch.Tensor, F: torch.Tensor, validate: bool = True) -> torch.Tensor:
    """9 x m float64 face SoA (rows ax..cz) gathered on the device."""
    _require_cuda(V, F)
    if V.dim() != 2 or V.shape[1] != 3:
        raise ValueError("V must have shape (n, 3)")
    if F.dim() != 2 or F.shape[1] != 3:
        raise ValueError("F must have shape (m, 3)")
    if V.dtype not in (torch.float32, torch.float64):
        raise TypeError("V must be float32 or float64")
    if F.dtype not in (torch.int32, torch.int64):
        raise TypeError("F must be int32 or int64")
    V = V.contiguous()
    F = F.contiguous()
    tri = torch.empty((9, F.shape[0]), dtype=torch.float64, device=V.device)
    bad = torch.full((1,), _INT32_MAX, dtype=torch.int32, device=V.device)
    _check(lib().wn_pack_soa(V.data_ptr(), F.data_ptr(), V.shape[0], F.shape[0],
                             0 if V.dtype == torch.float64 else 1,
                             0 if F.dtype == torch.int64 else 1,
                             tri.data_ptr(), bad.data_ptr(), _stream()), "wn_pack_soa")
    if validate and F.shape[0] and int(bad.item()) != _INT32_MAX:
        raise IndexError("face index is outside V")
    return tri


def resolve_config(nf: int, nq: int, precision: int, qpl=0, splits=0, sort=-1, kappa=-1.0):
    out = np.zeros(4, dtype=np.int64)
    _check(lib().wn_resolve_config(nf, nq, precision, qpl, splits, sort, float(kappa),
                                   out.ctypes.data), "wn_resolve_config")
    return int(out[0]), int(out[1]), int(out[2])


def forward(tri: torch.Tensor, O: torch.Tensor, precision="fp64", qpl=0, splits=0, sort=-1,
            kappa=-1.0) -> torch.Tensor:
    """w (nq,) float64 for the 9 x m SoA ``tri`` and queries O (nq, 3)."""
    mode = PRECISIONS[str(precision).lower()] if isinstance(precision, str) else int(precision)
    _require_cuda(tri, O)
    Q = O.to(torch.float64).contiguous()
    nf, nq = tri.shape[1], Q.shape[0]
    qpl, splits, sort = resolve_config(nf, nq, mode, qpl, splits, sort, kappa)
    nbytes = lib().wn_forward_workspace_bytes(nf, nq, mode, qpl, splits, sort)
    if nbytes < 0:
        raise RuntimeError(f"wn_forward_workspace_bytes failed: {last_error()}")
    ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=Q.device)
    w = torch.empty(nq, dtype=torch.float64, device=Q.device)
    _check(lib().wn_forward(tri.data_ptr(), Q.data_ptr(), w.data_ptr(), nf, nq, mode, qpl,
                            splits, sort, float(kappa), ws.data_ptr(), nbytes, _stream()),
           "wn_forward")
    return w


def backward(tri: torch.Tensor, Q: torch.Tensor, grad_w: torch.Tensor, qpl=0, splits=0):
    """(dL/dO (nq, 3), dL/dtri (9, m)) in float64 for upstream dL/dw."""
    Q = Q.to(torch.float64).contiguous()
    gw = grad_w.to(torch.float64).contiguous()
    nf, nq = tri.shape[1], Q.shape[0]
    gq = torch.empty((nq, 3), dtype=torch.float64, device=Q.device)
    gtri = torch.empty((9, nf), dtype=torch.float64, device=Q.device)
    _check(lib().wn_backward(tri.data_ptr(), Q.data_ptr(), gw.data_ptr(), gq.data_ptr(),
                             gtri.data_ptr(), nf, nq, 0, qpl, splits, _stream()), "wn_backward")
    return gq, gtri


def scatter_corners(grad_tri: torch.Tensor, F: torch.Tensor, nv: int) -> torch.Tensor:
    F = F.contiguous()
    gv = torch.empty((nv, 3), dtype=torch.float64, device=grad_tri.device)
    _check(lib().wn_scatter_corners(grad_tri.data_ptr(), F.data_ptr(),
                                    0 if F.dtype == torch.int64 else 1, F.shape[0], nv,
                                    gv.data_ptr(), _stream()), "wn_scatter_corners")
    return gv


def boundary_edges(F, n_vertices: int, device="cuda"):
    """Net directed boundary edges of F (host, wn_boundary_edges) as device
    tensors: ((ne, 2) int32 source/target ids, (ne,) float64 weights)."""
    Fh = F.detach().cpu().numpy() if isinstance(F, torch.Tensor) else np.asarray(F)
    if Fh.ndim != 2 or Fh.shape[1] != 3:
        raise ValueError("F must have shape (m, 3)")
    Fh = np.ascontiguousarray(Fh, dtype=np.int64)
    h = lib()
    cap = len(Fh) // 4 + 64  # one call when the rim fits
    e = np.zeros((cap, 2), dtype=np.int32)
    m = np.zeros(cap, dtype=np.float64)
    n = h.wn_boundary_edges(addr(Fh), int(n_vertices), len(Fh), 0, e.ctypes.data, m.ctypes.data,
                            cap)
    if n < 0:
        raise (IndexError if "outside V" in last_error() else ValueError)(last_error())
    if n > cap:
        e = np.zeros((n, 2), dtype=np.int32)
        m = np.zeros(n, dtype=np.float64)
        if h.wn_boundary_edges(addr(Fh), int(n_vertices), len(Fh), 0, e.ctypes.data,
                               m.ctypes.data, n) != n:
            raise RuntimeError(last_error())
    return (torch.as_tensor(np.ascontiguousarray(e[:n]), device=device),
            torch.as_tensor(np.ascontiguousarray(m[:n]), device=device))


def backward_edges(V: torch.Tensor, edges: torch.Tensor, weight: torch.Tensor, Q: torch.Tensor,
                   grad_w: torch.Tensor, qpl=0, splits=0):
    """(dL/dO (nq, 3), dL/dV (nv, 3)) in float64 from the boundary edges."""
    if V.dtype not in (torch.float32, torch.float64):
        raise TypeError("V must be float32 or float64")
    V = V.detach().contiguous()
    Q = Q.to(torch.float64).contiguous()
    gw = grad_w.to(torch.float64).contiguous()
    e = edges.to(torch.int32).contiguous()
    m = weight.to(torch.float64).contiguous()
    nq, nv, ne = Q.shape[0], V.shape[0], e.shape[0]
    gq = torch.empty((nq, 3), dtype=torch.float64, device=Q.device)
    gv = torch.empty((nv, 3), dtype=torch.float64, device=Q.device)
    _check(lib().wn_backward_edges(V.data_ptr(), nv, 0 if V.dtype == torch.float64 else 1,
                                   e.data_ptr(), m.data_ptr(), ne, Q.data_ptr(), gw.data_ptr(),
                                   gq.data_ptr(), gv.data_ptr(), nq, qpl, splits, _stream()),
           "wn_backward_edges")
    return gq, gv


ADJOINTS = ("edges", "faces")


class WindingNumber(torch.autograd.Function):
    """w = winding_number(V, F, O); backward gives dL/dV and dL/dO."""

    @staticmethod
    def forward(ctx, V, F, O, precision="fp64", adjoint="edges"):
        if adjoint not in ADJOINTS:
            raise ValueError(f"adjoint must be one of {ADJOINTS}")
        tri = pack_faces(V, F)
        Q = O.to(torch.float64).contiguous()
        w = forward(tri, Q, precision)
        if adjoint == "faces":
            ctx.save_for_backward(tri, Q, F)
        else:
            ctx.save_for_backward(V.detach(), Q, F)
        ctx.adjoint = adjoint
        ctx.nv = V.shape[0]
        ctx.vdtype = V.dtype
        ctx.odtype = O.dtype
        return w

    @staticmethod
    def backward(ctx, grad_w):
        need_v, _, need_o = ctx.needs_input_grad[:3]
        if ctx.adjoint == "faces":
            tri, Q, F = ctx.saved_tensors
            gq, gtri = backward(tri, Q, grad_w)
            gv = scatter_corners(gtri, F, ctx.nv) if need_v else None
        else:
            V, Q, F = ctx.saved_tensors
            edges, weight = boundary_edges(F, ctx.nv, device=Q.device)
            gq, gv = backward_edges(V, edges, weight, Q, grad_w)
        gv = gv.to(ctx.vdtype) if need_v else None
        go = gq.to(ctx.odtype) if need_o else None
        return gv, None, go, None, None


def winding_number(V: torch.Tensor, F: torch.Tensor, O: torch.Tensor, precision="fp64",
                   adjoint="edges"):
    """Differentiable generalized winding number of (V, F) at the rows of O.
    ``adjoint`` picks the backward: "edges" (boundary edges only) or "faces".

    Calls with at least igl.STRIP_MIN_QUERIES[precision] queries run on the
    strip kernels through a WindingField that is kept for the next call with
    an equal F (vertex count, precision and adjoint), as an optimisation loop
    makes them: the strip topology and boundary edges are built once, and
    each call regathers the records from V on the device. The field is kept
    per thread, and each call's stream waits for the previous call's
    forward. The route depends on the sizes only, never on earlier calls;
    release() drops the kept field."""
    if adjoint not in ADJOINTS:
        raise ValueError(f"adjoint must be one of {ADJOINTS}")
    key = str(precision).lower()
    if key not in PRECISIONS:
        raise ValueError("precision must be 'fp64' or 'fp32c'")
    _require_cuda(V, F, O)
    field = _kept_field(V, F, key, adjoint, O.shape[0] if O.dim() == 2 else 0)
    if field is None:
        return WindingNumber.apply(V, F, O, precision, adjoint)
    # the field's strip records are regathered from V by each call: this
    # call's stream first waits for the previous call's forward
    stream = torch.cuda.current_stream(V.device)
    done = getattr(_kept, "done", None)
    if done is not None:
        stream.wait_event(done)
    w = field(V, O)
    _kept.done = torch.cuda.Event()
    _kept.done.record(stream)
    return w


# per thread: the WindingField of the thread's last large call (its strip
# records are rewritten by every call, so threads do not share one)
_kept = threading.local()


def _kept_field(V, F, precision: str, adjoint: str, nq: int):
    """The WindingField for a large call (None: face-record route), reusing
    the one this thread kept from its last large call when F is equal."""
    from . import STRIP_MIN_QUERIES

    if nq < STRIP_MIN_QUERIES[PRECISIONS[precision]] or F.dim() != 2 or F.shape[0] == 0:
        return None
    if V.dim() != 2 or V.shape[1] != 3:
        raise ValueError("V must have shape (n, 3)")
    if F.shape[1] != 3:
        raise ValueError("F must have shape (m, 3)")
    if V.dtype not in (torch.float32, torch.float64):
        raise TypeError("V must be float32 or float64")
    if F.dtype not in (torch.int32, torch.int64):
        raise TypeError("F must be int32 or int64")
    Fd = F.to(device=V.device, dtype=torch.int64).contiguous()
    key = (int(V.shape[0]), precision, adjoint, int(Fd.shape[0]))
    kept = getattr(_kept, "field", None)
    if kept is not None and _kept.key == key and torch.equal(kept.F, Fd):
        return kept
    if int(Fd.min()) < 0 or int(Fd.max()) >= V.shape[0]:
        raise IndexError("face index is outside V")
    # the previous field is dropped, not closed: a pending backward may still
    # hold it, and its mesh is freed with the last reference
    _kept.field, _kept.key = None, None
    field = WindingField(Fd, V.shape[0], adjoint=adjoint, precision=precision)
    _kept.field, _kept.key = field, key
    return field


def kept_field():
    """The WindingField this thread's winding_number keeps (None if none)."""
    return getattr(_kept, "field", None)


def release() -> None:
    """Drops the WindingField that this thread's winding_number keeps between
    large calls; its device mesh is freed with the last reference (at once
    unless a pending backward still holds it)."""
    _kept.field, _kept.key, _kept.done = None, None, None


class _StripWinding(torch.autograd.Function):
    """Forward on a resident strip mesh (FP32-compensated or FP64), FP64
    adjoint."""

    @staticmethod
    def forward(ctx, V, O, field):
        field.mesh.set_vertices(V)
        Q = O.to(torch.float64).contiguous()
        w = field.mesh.winding_number(Q, precision=field.precision)
        ctx.save_for_backward(V, Q)
        ctx.field = field
        ctx.odtype = O.dtype
        return w

    @staticmethod
    def backward(ctx, grad_w):
        V, Q = ctx.saved_tensors
        field = ctx.field
        need_v, need_o = ctx.needs_input_grad[:2]
        if field.adjoint == "faces":
            tri = pack_faces(V, field.F, validate=False)
            gq, gtri = backward(tri, Q, grad_w)
            gv = scatter_corners(gtri, field.F, field.nv) if need_v else None
        else:
            gq, gv = backward_edges(V, field.edges, field.weight, Q, grad_w)
        gv = gv.to(V.dtype) if need_v else None
        go = gq.to(ctx.odtype) if need_o else None
        return gv, go, None


class WindingField:
    """Differentiable winding number of a fixed-topology mesh whose vertices
    change between calls (an optimisation loop): the strip topology and the
    boundary edges are built once on the host, every call gathers the strip
    records from V on the device and runs the strip kernel (FP32-compensated,
    or FP64 with precision="fp64"); backward is the FP64 adjoint (as
    winding_number's, "edges" by default).

        field = WindingField(F, n_vertices)
        w = field(V, O)            # V (nv, 3), O (q, 3) CUDA tensors
    """

    def __init__(self, F, n_vertices: int, adjoint="edges", precision="fp32c"):
        from .mesh import Mesh

        if adjoint not in ADJOINTS:
            raise ValueError(f"adjoint must be one of {ADJOINTS}")
        if str(precision).lower() not in ("fp32c", "fp64"):
            raise ValueError("precision must be 'fp32c' or 'fp64'")
        self.precision = str(precision).lower()
        Fh = F.detach().cpu().numpy() if isinstance(F, torch.Tensor) else np.asarray(F)
        self.mesh = Mesh.from_faces(Fh, int(n_vertices))
        self.F = torch.as_tensor(np.ascontiguousarray(Fh, dtype=np.int64), device="cuda")
        self.nv = int(n_vertices)
        self.adjoint = adjoint
        self.edges, self.weight = boundary_edges(Fh, self.nv)

    def __call__(self, V: torch.Tensor, O: torch.Tensor) -> torch.Tensor:
        _require_cuda(V, O)
        if V.shape != (self.nv, 3):
            raise ValueError("V must have shape (n_vertices, 3)")
        return _StripWinding.apply(V, O, self)

    def close(self) -> None:
        self.mesh.close()
