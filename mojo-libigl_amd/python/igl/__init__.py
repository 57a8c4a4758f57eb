"""MI355X-native drop-in for libigl's exact generalized winding number.

``igl.winding_number(V, F, O, device=...)`` and
``igl.winding_number_for_point(V, F, p)`` keep the reference's names,
arguments, return types and error contract (reference
python/igl/__init__.py:72-124, 232-283). ``device="cpu"`` runs the C++ FP64
sweep fanned out over query spans; ``device="gpu"`` runs the hand-written
gfx950 HIP kernels through the C ABI declared in include/wn.h.

GPU availability follows the reference (python/igl/__init__.py:251,264,
README.md:87-89): ``device="gpu"`` with no usable HIP device (none visible,
or less than 4000 MiB free) runs the CPU path and returns its result; a
kernel failure on a usable device raises ``RuntimeError`` (:262-263), now with
the runtime's reason. Setting ``WN_GPU_STRICT=1`` in the environment turns the
no-device fallback into a ``RuntimeError`` for callers who must know that the
GPU ran (this repository's GPU tests set it).

Difference from the reference, deliberate:

* keyword-only ``precision``: ``"fp64"`` (default, the reference's arithmetic)
  or ``"fp32c"`` (FP32 pairs with FP64 accumulation and FP64 re-evaluation of
  ill-conditioned pairs; GPU only).
"""

from __future__ import annotations

import operator
import os

import numpy as np

from ._lib import addr, fan_out, lib, spans, worker_count  # noqa: F401

__all__ = [
    "winding_number",
    "winding_number_for_point",
    "PRECISIONS",
]

PRECISIONS = {"fp64": 0, "fp32c": 1}

# The reference requires this much free device memory before it tries the GPU
# (python/igl/__init__.py:228).
GPU_MIN_FREE_BYTES = 4000 * 2**20


def _to_hi(function):
    """Wrap a ``(..., lo, hi)`` C entry point, ignoring ``fan_out``'s span
    (python/igl/__init__.py:16-22)."""

    def call(*arguments, span=0):
        function(*arguments)

    return call


def _float_array(value, name: str) -> np.ndarray:
    """Real dtype, integers exactly representable, finite -> C-contiguous f64
    (python/igl/__init__.py:72-83)."""
    source = np.asarray(value)
    if source.dtype.kind not in "fiu":
        raise TypeError(f"{name} must contain real numbers")
    if source.dtype.kind in "iu" and source.size:
        limit = 2**53
        if np.any(source < -limit) or np.any(source > limit):
            raise ValueError(f"{name} contains integers not exactly representable as float64")
    result = np.ascontiguousarray(source, dtype=np.float64)
    if not np.isfinite(result).all():
        raise ValueError(f"{name} must contain finite values")
    return result


def _index_array(value, name: str) -> np.ndarray:
    """Integer dtype -> C-contiguous int64 (python/igl/__init__.py:86-94)."""
    source = np.asarray(value)
    if not source.size:
        return np.ascontiguousarray(source, dtype=np.int64)
    if source.dtype.kind not in "iu" or source.dtype.kind == "b":
        raise TypeError(f"{name} must contain integers")
    if source.dtype.kind == "u" and source.size and source.max() > np.iinfo(np.int64).max:
        raise OverflowError(f"{name} contains an index outside int64 range")
    return np.ascontiguousarray(source, dtype=np.int64)


def _integer(value, name: str) -> int:
    try:
        return operator.index(value)
    except TypeError as error:
        raise TypeError(f"{name} must be an integer") from error


def _vertices(value, *, three_dimensional: bool = False) -> np.ndarray:
    result = _float_array(value, "V")
    if result.ndim != 2 or result.shape[1] not in ((3,) if three_dimensional else (2, 3)):
        expected = "(n, 3)" if three_dimensional else "(n, 2) or (n, 3)"
        raise ValueError(f"V must have shape {expected}")
    return result


def _faces(value) -> np.ndarray:
    result = _index_array(value, "F")
    if result.ndim != 2 or result.shape[1] != 3:
        raise ValueError("F must have shape (m, 3)")
    return result


def _mesh(v, f, *, three_dimensional: bool = False):
    """Validated (V, F) (python/igl/__init__.py:119-124)."""
    vertices = _vertices(v, three_dimensional=three_dimensional)
    faces = _faces(f)
    if faces.size and (faces.min() < 0 or faces.max() >= len(vertices)):
        raise IndexError("face index is outside V")
    return vertices, faces


def _precision(value) -> int:
    key = str(value).lower()
    if key not in PRECISIONS:
        raise ValueError("precision must be 'fp64' or 'fp32c'")
    return PRECISIONS[key]


def _last_error(handle) -> str:
    getter = getattr(handle, "wn_last_error", None)
    if getter is None:
        return "no reason reported"
    message = getter()
    return message.decode(errors="replace") if message else "no reason reported"


def _gpu_memory_available() -> bool:
    """A HIP device is visible with >= 4000 MiB free. Replaces the reference's
    nvidia-smi probe (python/igl/__init__.py:213-229) with two C-ABI calls."""
    handle = lib()
    if handle.wn_device_count() < 1:
        return False
    return handle.wn_device_free_bytes() >= GPU_MIN_FREE_BYTES


# Environment switch: a GPU request without a usable device raises instead of
# running the CPU path.
STRICT_ENV = "WN_GPU_STRICT"


def _use_gpu(requested: bool) -> bool:
    """Whether a call that asked for ``device="gpu"`` runs on the GPU.

    The reference tries the GPU only when ``_gpu_memory_available()`` holds
    and otherwise runs the CPU path (python/igl/__init__.py:251,264);
    ``WN_GPU_STRICT=1`` makes that case an error instead.
    """
    if not requested:
        return False
    if _gpu_memory_available():
        return True
    if os.environ.get(STRICT_ENV, "").strip() not in ("", "0"):
        raise RuntimeError(
            "device='gpu' requested but no HIP GPU with >= 4000 MiB free is available "
            f"({STRICT_ENV} is set, so the CPU path is not used)")
    return False


def pack_faces(vertices: np.ndarray, faces: np.ndarray) -> np.ndarray:
    """9 x m face SoA, rows ax,ay,az,bx,by,bz,cx,cy,cz
    (python/igl/__init__.py:248-250)."""
    return np.ascontiguousarray(vertices[faces].transpose(1, 2, 0).reshape(9, len(faces)))


# GPU calls with at least this many queries (by precision mode: 0 = FP64,
# 1 = FP32-compensated) go through the strip kernels (wn_winding_gpu_mesh
# here, a kept WindingField in igl.autograd.winding_number): stripifying on
# the host costs ~0.1-0.14 us per face and the strip kernels save ~0.93 ps
# (FP64) / ~0.36 ps (FP32c) per query-face pair over the face kernels, so a
# first call breaks even at ~100-150k / ~330k queries and calls that repeat a
# kept mesh win well below that.
STRIP_MIN_QUERIES = {0: 131072, 1: 262144}


def winding_number(V, F, O, device="cpu", *, precision="fp64"):
    """Evaluate the exact generalized winding number of a triangle soup.

    Reference: python/igl/__init__.py:232-279. ``O`` of shape ``(3,)`` returns
    a Python float, ``(q, 3)`` returns a ``(q,)`` float64 array.
    """
    vertices, faces = _mesh(V, F, three_dimensional=True)
    queries = _float_array(O, "O")
    scalar = queries.ndim == 1
    if scalar:
        if queries.shape != (3,):
            raise ValueError("o must have shape (3,)")
        queries = queries[None, :]
    if queries.ndim != 2 or queries.shape[1] != 3:
        raise ValueError("O must have shape (q, 3)")
    selected_device = str(device).lower()
    if selected_device not in ("cpu", "gpu"):
        raise ValueError("device must be 'cpu' or 'gpu'")
    mode = _precision(precision)
    if mode != 0 and selected_device != "gpu":
        raise ValueError("precision='fp32c' is a GPU mode; use device='gpu'")
    result = np.empty(len(queries), dtype=np.float64)
    if len(queries):
        # no usable device: the CPU path below, in FP64 whatever the
        # requested precision (python/igl/__init__.py:251,264)
        if _use_gpu(selected_device == "gpu"):
            handle = lib()
            if len(queries) >= STRIP_MIN_QUERIES[mode] and len(faces):
                vc = np.ascontiguousarray(vertices, dtype=np.float64)
                fc = np.ascontiguousarray(faces, dtype=np.int64)
                status = handle.wn_winding_gpu_mesh(addr(vc), addr(fc), len(vc), len(fc), 0,
                                                    addr(queries), addr(result), len(queries),
                                                    mode)
                if status != 0:
                    raise RuntimeError(
                        f"the HIP GPU winding-number kernel failed: {_last_error(handle)}")
                return float(result[0]) if scalar else result
            # V and F go up as they are and the face SoA is gathered on the
            # device (the reference packs it on the host, :248-250)
            status = handle.wn_winding_gpu_vf(addr(vertices), addr(faces), len(vertices),
                                              len(faces), 0, addr(queries), addr(result),
                                              len(queries), mode)
            if status != 0:
                raise RuntimeError(
                    f"the HIP GPU winding-number kernel failed: {_last_error(handle)}")
        else:
            triangles = pack_faces(vertices, faces)
            # Every pair costs three square roots and an arctangent, so the
            # sweep is compute bound and the query rows split across threads.
            fan_out(
                _to_hi(lib().mli_winding),
                (addr(triangles), addr(queries), addr(result), len(faces), len(queries)),
                len(queries),
                worker_count(len(queries), len(faces) * len(queries), "winding"),
            )
    return float(result[0]) if scalar else result


def winding_number_for_point(V, F, p):
    """Winding number at one point, CPU (python/igl/__init__.py:282-283)."""
    return float(winding_number(V, F, p))


from .aabb import AABB, point_mesh_squared_distance, signed_distance  # noqa: E402
from .fast_winding import (  # noqa: E402
    FastWindingNumber,
    fast_winding_number,
    fast_winding_number_for_meshes,
)


def __getattr__(name):
    # Lazy: these need PyTorch only when used.
    if name == "Mesh":
        from .mesh import Mesh

        return Mesh
    raise AttributeError(name)
