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
* :func:`ref_gpu_atan2` -- the reference GPU kernel's own arctangent,
  ``src/capi.mojo:326-352``.
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
