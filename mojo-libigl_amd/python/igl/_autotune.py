"""Launch-shape autotuner for the batched kernels.

Times candidate configurations on the live device (PyTorch events on the
current stream) and caches the winner in a JSON file, so a shape class is
tuned once per machine and build:

  face-record batch plans: queries per lane x faces per work item (chunk)
  strip batch plans:       255-record tiles per work item (chunk_tiles)

A cache key names the device (its name and gfx architecture), the library
build (ABI version, size and modification time of the shared library), the
precision and the batch's log2 size buckets, so a winner found on another
device or with another build of the kernels is never reused.

Cache file: $WN_TUNE_CACHE, else ~/.cache/mojo-libigl-amd/tune.json.
"""

from __future__ import annotations

import json
import math
import os

import numpy as np


def cache_path() -> str:
    return os.environ.get("WN_TUNE_CACHE") or os.path.join(
        os.path.expanduser("~"), ".cache", "mojo-libigl-amd", "tune.json")


def _load() -> dict:
    try:
        with open(cache_path()) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _store(data: dict) -> None:
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as fh:
            json.dump(data, fh, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def _bucket(n: int) -> int:
    return int(math.floor(math.log2(max(1, n))))


def _time(fn, repeats: int) -> float:
    import torch

    fn()
    torch.cuda.synchronize()
    best = float("inf")
    s = torch.cuda.current_stream()
    for _ in range(repeats):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(s)
        fn()
        b.record(s)
        torch.cuda.synchronize()
        best = min(best, a.elapsed_time(b))
    return best


def device_name() -> str:
    """Name and gfx architecture of the current device."""
    try:
        import torch

        index = torch.cuda.current_device()
        props = torch.cuda.get_device_properties(index)
        arch = getattr(props, "gcnArchName", "")
        return f"{props.name}/{arch}" if arch else str(props.name)
    except Exception:
        return "unknown"


def library_build() -> str:
    """ABI version of the loaded library plus the size and modification time
    of its file: a rebuilt library is a new build."""
    from ._lib import LIB, lib

    try:
        st = os.stat(LIB)
        stamp = f"{st.st_size}.{int(st.st_mtime)}"
    except OSError:
        stamp = "nofile"
    return f"abi{int(lib().wn_version())}.{stamp}"


def cache_key(kind: str, precision: int, signature: str) -> str:
    return f"{device_name()}|{library_build()}|{kind}:p{int(precision)}:{signature}"


def _signature(counts, query_counts) -> str:
    c = np.asarray(counts, dtype=np.int64)
    q = np.asarray(query_counts, dtype=np.int64)
    return f"{len(c)}:{_bucket(int(c.sum()))}:{_bucket(int(q.sum()))}:{_bucket(int((c * q).sum()))}"


def _tuned(key: str, candidates, measure, use_cache: bool):
    """The cached winner for key, else time every candidate (measure(c) ->
    ms), cache and return the fastest."""
    if use_cache:
        hit = _load().get(key)
        if hit is not None:
            value = hit["config"]
            return tuple(value) if isinstance(value, list) else value
    results = sorted(((measure(c), i, c) for i, c in enumerate(candidates)),
                     key=lambda r: (r[0], r[1]))
    best = results[0][2]
    if use_cache:
        data = _load()
        data[key] = {"config": list(best) if isinstance(best, tuple) else best,
                     "ms": results[0][0],
                     "tried": [[list(c) if isinstance(c, tuple) else c, ms] for ms, _, c in results]}
        _store(data)
    return best


def tune_batch(face_counts, query_counts, tri, Q, precision: int, repeats: int = 2,
               use_cache: bool = True, chunks=(1024, 2048, 4096, 8192), qpls=(1, 2)):
    """Best (qpl, chunk_faces) of a face-record batch plan for this shape."""
    from .batch import WindingBatch

    fc = np.asarray(face_counts, dtype=np.int64)
    qc = np.asarray(query_counts, dtype=np.int64)

    def measure(config):
        qpl, chunk = config
        plan = WindingBatch(fc, qc, precision=precision, qpl=qpl, chunk_faces=chunk)
        try:
            return _time(lambda: plan.run(tri, Q, use_graph=False), repeats)
        finally:
            plan.close()

    key = cache_key("batch", precision, _signature(fc, qc))
    return _tuned(key, [(q, c) for q in qpls for c in chunks], measure, use_cache)


def tune_batch_strips(record_counts, query_counts, rec, Q, repeats: int = 2,
                      use_cache: bool = True, chunks=(2, 4, 8, 16, 32),
                      precision="fp32c", boundary=None) -> int:
    """Best chunk_tiles (255-record tiles per work item) of a strip batch
    plan in the given precision (with its BatchBoundary, as it will run)."""
    from .batch import StripBatch, _mode

    rc = np.asarray(record_counts, dtype=np.int64)
    qc = np.asarray(query_counts, dtype=np.int64)
    mode = _mode(precision)

    def measure(chunk):
        plan = StripBatch(rc, qc, chunk_tiles=chunk, precision=precision, boundary=boundary)
        try:
            return _time(lambda: plan.run(rec, Q, use_graph=False), repeats)
        finally:
            plan.close()

    key = cache_key("batch_strips", mode, _signature(rc, qc))
    return int(_tuned(key, list(chunks), measure, use_cache))
