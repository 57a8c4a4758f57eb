"""Launch-shape autotuner for the forward and batched kernels.

Times candidate configurations on the live device (PyTorch events on the
current stream) and caches the winner in a JSON file keyed by precision and
log2 size buckets, so a shape class is tuned once per machine:

  forward: queries per lane x face splits
  batch:   queries per lane x faces per work item (chunk)

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


def forward_candidates(nf: int, nq: int, precision: int):
    from .autograd import resolve_config

    qpls = (1, 2) if precision == 0 else (1, 2, 4)
    out = []
    for qpl in qpls:
        _, s0, _ = resolve_config(nf, nq, precision, qpl, 0)
        for splits in sorted({max(1, s0 // 2), s0, min(65535, s0 * 2)}):
            out.append((qpl, splits))
    return out


def tune_forward(tri, Q, precision: int, repeats: int = 3, use_cache: bool = True):
    """Best (qpl, splits) for this shape class; times every candidate once
    per cache miss."""
    from .autograd import forward

    nf, nq = int(tri.shape[1]), int(Q.shape[0])
    key = f"fwd:p{precision}:f{_bucket(nf)}:q{_bucket(nq)}"
    data = _load() if use_cache else {}
    if key in data:
        return tuple(data[key]["config"])
    results = []
    for qpl, splits in forward_candidates(nf, nq, precision):
        ms = _time(lambda: forward(tri, Q, precision, qpl=qpl, splits=splits), repeats)
        results.append((ms, qpl, splits))
    results.sort()
    best = (results[0][1], results[0][2])
    if use_cache:
        data = _load()
        data[key] = {"config": list(best), "ms": results[0][0],
                     "tried": [[r[1], r[2], r[0]] for r in results]}
        _store(data)
    return best


def tune_batch(face_counts, query_counts, tri, Q, precision: int, repeats: int = 2,
               use_cache: bool = True, chunks=(1024, 2048, 4096, 8192), qpls=(1, 2)):
    """Best (qpl, chunk_faces) for a batch shape."""
    from .batch import WindingBatch

    fc = np.asarray(face_counts, dtype=np.int64)
    qc = np.asarray(query_counts, dtype=np.int64)
    sig = f"{len(fc)}:{_bucket(int(fc.sum()))}:{_bucket(int(qc.sum()))}:{_bucket(int((fc * qc).sum()))}"
    key = f"batch:p{precision}:{sig}"
    data = _load() if use_cache else {}
    if key in data:
        return tuple(data[key]["config"])
    results = []
    for qpl in qpls:
        for chunk in chunks:
            plan = WindingBatch(fc, qc, precision=precision, qpl=qpl, chunk_faces=chunk)
            try:
                ms = _time(lambda: plan.run(tri, Q, use_graph=False), repeats)
            finally:
                plan.close()
            results.append((ms, qpl, chunk))
    results.sort()
    best = (results[0][1], results[0][2])
    if use_cache:
        data = _load()
        data[key] = {"config": list(best), "ms": results[0][0],
                     "tried": [[r[1], r[2], r[0]] for r in results]}
        _store(data)
    return best


def tune_batch_strips(record_counts, query_counts, rec, Q, repeats: int = 2,
                      use_cache: bool = True, chunks=(2, 4, 8, 16, 32)) -> int:
    """Best chunk_tiles (255-record tiles per work item) for a strip batch."""
    from .batch import StripBatch

    rc = np.asarray(record_counts, dtype=np.int64)
    qc = np.asarray(query_counts, dtype=np.int64)
    sig = f"{len(rc)}:{_bucket(int(rc.sum()))}:{_bucket(int(qc.sum()))}:{_bucket(int((rc * qc).sum()))}"
    key = f"batch_strips:{sig}"
    data = _load() if use_cache else {}
    if key in data:
        return int(data[key]["config"])
    results = []
    for chunk in chunks:
        plan = StripBatch(rc, qc, chunk_tiles=chunk)
        try:
            ms = _time(lambda: plan.run(rec, Q, use_graph=False), repeats)
        finally:
            plan.close()
        results.append((ms, chunk))
    results.sort()
    best = int(results[0][1])
    if use_cache:
        data = _load()
        data[key] = {"config": best, "ms": results[0][0], "tried": [[r[1], r[0]] for r in results]}
        _store(data)
    return best


def device_name() -> str:
    try:
        import torch

        return torch.cuda.get_device_name(0)
    except Exception:
        return "unknown"
