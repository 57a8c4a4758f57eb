"""ctypes bridge to the MI355X winding-number library.

Mirrors the reference's shim (python/igl/_lib.py): integer addresses cross the
ABI as ``c_int64`` (``:16-32``), the CPU kernel is fanned out over query spans
on a thread pool (``:37-95``), and ``lib()`` binds every declared symbol
(``:122-130``). Differences: the library is HIP + C++
(``mojo-libigl_amd/lib/libmojo_libigl_amd.so``, override with ``$WN_LIB``) and
is never rebuilt implicitly -- a missing library raises, so nothing silently
runs a different code path.
"""

from __future__ import annotations

import ctypes
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

PACKAGE_ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
LIB = os.environ.get("WN_LIB") or os.path.join(PACKAGE_ROOT, "lib", "libmojo_libigl_amd.so")

I = ctypes.c_int64
D = ctypes.c_double
S = ctypes.c_char_p

# name -> (argument types, result type). Pointers are passed as c_int64
# addresses, as in the reference (python/igl/_lib.py:16-32).
_SIGNATURES = {
    "mli_winding": ([I] * 7, None),
    "mli_winding_gpu": ([I] * 5, I),
    "wn_device_count": ([], I),
    "wn_device_free_bytes": ([], I),
    "wn_last_error": ([], S),
    "wn_version": ([], I),
    "wn_kernel_timer": ([I] * 2, I),
    "wn_winding_gpu_ex": ([I] * 9 + [D], I),
    "wn_winding_gpu_vf": ([I] * 9, I),
    "wn_resolve_config": ([I] * 6 + [D, I], I),
    "wn_forward_workspace_bytes": ([I] * 6, I),
    "wn_forward": ([I] * 9 + [D] + [I] * 3, I),
    "wn_pack_soa": ([I] * 9, I),
    "wn_mesh_records": ([I] * 7, I),
    "wn_mesh_create": ([I] * 5, I),
    "wn_mesh_create_topology": ([I] * 4, I),
    "wn_mesh_set_vertices": ([I] * 5, I),
    "wn_strip_topology": ([I] * 7, I),
    "wn_strip_gather": ([I] * 8, I),
    "wn_strip_boundary": ([I] * 9, I),
    "wn_batch_set_boundary": ([I] * 5, I),
    "wn_boundary_edges": ([I] * 7, I),
    "wn_backward_edges": ([I] * 14, I),
    "wn_mesh_info": ([I] * 2, I),
    "wn_mesh_resolve": ([I] * 5, I),
    "wn_mesh_workspace_bytes": ([I] * 4, I),
    "wn_mesh_forward": ([I] * 6 + [D] + [I] * 3, I),
    "wn_mesh_forward64": ([I] * 9, I),
    "wn_mesh_forward_noalloc": ([I] * 6 + [D, I], I),
    "wn_winding_gpu_mesh": ([I] * 9, I),
    "wn_release_host_buffers": ([], I),
    "wn_mesh_destroy": ([I], I),
    "wn_batch_create": ([I] * 7 + [D], I),
    "wn_batch_create_strips": ([I] * 5 + [D], I),
    "wn_batch_create_strips64": ([I] * 5, I),
    "mli_aabb_squared_distance": ([I] * 14, None),
    "wn_aabb_build": ([I] * 11, I),
    "wn_aabb_create": ([I] * 6, I),
    "wn_aabb_info": ([I] * 2, I),
    "wn_aabb_workspace_bytes": ([I] * 3, I),
    "wn_aabb_query": ([I] * 10, I),
    "wn_aabb_destroy": ([I], I),
    "wn_fwn_create": ([I] * 7, I),
    "wn_fwn_info": ([I] * 2, I),
    "wn_fwn_export": ([I] * 4, I),
    "wn_fwn_query_cpu": ([I] * 4 + [D] + [I] * 2, I),
    "wn_fwn_workspace_bytes": ([I] * 3, I),
    "wn_fwn_query": ([I] * 4 + [D] + [I] * 4, I),
    "wn_fwn_destroy": ([I], I),
    "wn_batch_run": ([I] * 6, I),
    "wn_batch_info": ([I] * 2, I),
    "wn_batch_destroy": ([I], I),
    "wn_backward": ([I] * 11, I),
    "wn_scatter_corners": ([I] * 7, I),
}

# Same policy as the reference (python/igl/_lib.py:37-50): at most four
# workers, and no split below the measured crossover of the query-face sweep.
MAX_WORKERS = 4
PARALLEL_MIN_WORK = {"winding": 120_000, "aabb": 4_000_000}


def worker_count(items: int, work: int, kernel: str) -> int:
    """Span count for a kernel with ``items`` independent outputs
    (python/igl/_lib.py:53-61)."""
    if items < 2 or work < PARALLEL_MIN_WORK[kernel]:
        return 1
    try:
        available = len(os.sched_getaffinity(0))
    except AttributeError:
        available = os.cpu_count() or 1
    return max(1, min(items, available, MAX_WORKERS))


def spans(total: int, parts: int) -> list[tuple[int, int]]:
    """``total`` items in ``parts`` contiguous near-equal spans
    (python/igl/_lib.py:64-69)."""
    step = -(-total // parts)
    return [(lo, min(lo + step, total)) for lo in range(0, total, step) if lo < total]


def fan_out(function, arguments, total: int, workers: int) -> int:
    """Call ``function(*arguments, lo, hi, span=index)`` once per span
    (python/igl/_lib.py:72-95); returns the number of spans used."""
    if workers == 1:
        function(*arguments, 0, total, span=0)
        return 1
    parts = spans(total, workers)
    with ThreadPoolExecutor(max_workers=workers) as pool:
        list(pool.map(lambda item: function(*arguments, item[1][0], item[1][1], span=item[0]),
                      enumerate(parts)))
    return len(parts)


_handle: ctypes.CDLL | None = None


def _preload_hip_runtime() -> None:
    """Make the library share PyTorch's HIP runtime.

    The library needs ``libamdhip64.so.7``; PyTorch ships its own copy of that
    runtime under ``torch/lib``. Two runtimes in one process cannot both own
    the GPU, so the one PyTorch will use is loaded first (RTLD_GLOBAL): the
    library's dependency then resolves to it, and a later ``import torch``
    maps the same file instead of a second runtime. ``WN_HIP_RUNTIME=system``
    keeps the system runtime (processes that never import torch).
    """
    if os.environ.get("WN_HIP_RUNTIME", "").lower() == "system":
        return
    import importlib.util

    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.submodule_search_locations:
        return
    for base in spec.submodule_search_locations:
        path = os.path.join(base, "lib", "libamdhip64.so")
        if os.path.exists(path):
            ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
            return


def lib() -> ctypes.CDLL:
    """Load the library and bind every symbol of ``_SIGNATURES``."""
    global _handle
    if _handle is None:
        if not os.path.exists(LIB):
            raise RuntimeError(
                f"winding-number library not found at {LIB}; build it with "
                "`python mojo-libigl_amd/build.py` (or set WN_LIB)")
        _preload_hip_runtime()
        handle = ctypes.CDLL(LIB)
        for name, (args, result) in _SIGNATURES.items():
            function = getattr(handle, name)
            function.argtypes = args
            function.restype = result
        _handle = handle
    return _handle


def last_error() -> str:
    message = lib().wn_last_error()
    return message.decode(errors="replace") if message else ""


def addr(array: np.ndarray) -> int:
    return array.ctypes.data
