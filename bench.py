#!/usr/bin/env python
"""Throughput of the generalized-winding-number hot path on one MI355X.

One step = one forward pass of the FP32-compensated kernel (wn_forward,
precision 1) over the closed torus R = 1, r = 0.35 on a 120 x 118 grid
(28,320 faces) x 1,048,576 queries uniform in the 20 %-padded bbox, inputs
resident in HBM. The same steps are timed for the FP64-plain kernel, the
comparison point of the metric, and both outputs are compared.

Prints ONE JSON line. `value` = G query-face pairs/s of the FP32-compensated
kernel; `roofline` prices 66 algorithmic FLOP per pair (62 basic ops + 3 sqrt
+ 1 atan2, counted as one each) against MI355X's published FP32 vector peak;
`cpu_baseline` is the FP64 oracle (a C restatement of the reference CPU path)
on the host cores over a bounded query sample.

    python bench.py [--steps K] [--warmup W] [--queries N] [--gpus 1]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "mojo-libigl_amd", "python"), ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

FLOP_PER_PAIR = 66
FP32_VECTOR_PEAK_TFLOPS = 157.3
FP64_VECTOR_PEAK_TFLOPS = 78.6
METRIC = "G query-face pairs/s, FP32-comp vs FP64 plain, 1 MI355X @ max|Δw|≤1e-6"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(tri, Q, seconds: float):
    """FP64 oracle over a bounded sample of the same queries, all usable host
    cores (at most 16, the box's CPU share), contiguous query spans."""
    import numpy as np

    import oracle

    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    nf = tri.shape[1]
    probe = Q[: 64 * cores]
    t0 = time.perf_counter()
    oracle.winding_threads(tri, probe, cores)
    dt = max(time.perf_counter() - t0, 1e-6)
    rate = nf * len(probe) / dt
    nq = int(min(len(Q), max(cores, rate * seconds / nf)))
    sample = Q[:nq]
    oracle.winding_threads(tri, sample[: min(nq, 4 * cores)], cores)  # warm-up
    t0 = time.perf_counter()
    oracle.winding_threads(tri, sample, cores)
    dt = time.perf_counter() - t0
    return {
        "value": nf * nq / dt / 1e9,
        "unit": "G query-face pairs/s",
        "cores": cores,
        "kind": "port",
        "sample": f"first {nq} of the {len(Q)} queries x all {nf} faces, FP64, "
                  f"{cores} threads, one timed run of {dt:.1f} s after a warm-up",
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--queries", type=int, default=1_048_576)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64-plain comparison")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the winding-number path runs on a single MI355X here).", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import igl
    import wn_workloads as W

    h = igl.lib()
    if h.wn_device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        return 1
    log("generating workload")
    V, F, Q = W.config2(args.queries)
    tri = igl.pack_faces(V, F)
    nf, nq = len(F), len(Q)
    dev = torch.device("cuda")
    trid = torch.tensor(tri, device=dev)
    Qd = torch.tensor(Q, device=dev)
    stream = torch.cuda.current_stream()
    sptr = stream.cuda_stream

    def setup(precision):
        cfg = np.zeros(4, dtype=np.int64)
        st = h.wn_resolve_config(nf, nq, precision, 0, 0, -1, -1.0, cfg.ctypes.data)
        assert st == 0, igl._lib.last_error()
        qpl, splits, sort = int(cfg[0]), int(cfg[1]), int(cfg[2])
        nbytes = h.wn_forward_workspace_bytes(nf, nq, precision, qpl, splits, sort)
        assert nbytes >= 0, igl._lib.last_error()
        ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=dev)
        out = torch.empty(nq, dtype=torch.float64, device=dev)
        return dict(qpl=qpl, splits=splits, sort=sort, ws=ws, nbytes=nbytes, out=out)

    def run(precision, s):
        st = h.wn_forward(trid.data_ptr(), Qd.data_ptr(), s["out"].data_ptr(), nf, nq, precision,
                          s["qpl"], s["splits"], s["sort"], -1.0, s["ws"].data_ptr(), s["nbytes"],
                          sptr)
        if st != 0:
            raise RuntimeError(igl._lib.last_error())

    def timed(precision, s, steps, warmup):
        for _ in range(warmup):
            run(precision, s)
        torch.cuda.synchronize()
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        start.record(stream)
        t0 = time.perf_counter()
        for _ in range(steps):
            run(precision, s)
        end.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        return start.elapsed_time(end) / steps, wall / steps * 1e3

    s32 = setup(1)
    log(f"fp32c config qpl={s32['qpl']} splits={s32['splits']} sort={s32['sort']}; "
        f"{args.warmup} warm-up + {args.steps} timed steps")
    ms32, wall32 = timed(1, s32, args.steps, args.warmup)
    log(f"fp32c {ms32:.3f} ms/step (device), {wall32:.3f} ms/step (wall)")
    w32 = s32["out"].cpu().numpy()

    result_fp64 = None
    acc = {}
    if not args.no_fp64:
        s64 = setup(0)
        ms64, wall64 = timed(0, s64, args.steps, args.warmup)
        log(f"fp64 {ms64:.3f} ms/step (device)")
        w64 = s64["out"].cpu().numpy()
        d = np.abs(w32 - w64)
        acc = {"max_abs_dw_fp32c_vs_fp64": float(d.max()), "mean_abs_dw_fp32c_vs_fp64": float(d.mean()),
               "fp64_max_dist_from_integer": float(np.abs(w64 - np.round(w64)).max())}
        result_fp64 = {"value": nf * nq / (ms64 * 1e-3) / 1e9, "ms_per_step": ms64,
                       "qpl": s64["qpl"], "splits": s64["splits"],
                       "roofline_frac_fp64_peak": nf * nq * FLOP_PER_PAIR / (ms64 * 1e-3) / 1e12
                       / FP64_VECTOR_PEAK_TFLOPS}
        import oracle

        sub = np.arange(0, nq, max(1, nq // 512))
        ref = oracle.winding_soa(tri, Q[sub])
        acc["max_abs_dw_fp64_vs_oracle_sub512"] = float(np.abs(w64[sub] - ref).max())
        acc["max_abs_dw_fp32c_vs_oracle_sub512"] = float(np.abs(w32[sub] - ref).max())

    value = nf * nq / (ms32 * 1e-3) / 1e9
    achieved = nf * nq * FLOP_PER_PAIR / (ms32 * 1e-3) / 1e12
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "G query-face pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms32,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"closed torus R=1 r=0.35 120x118 ({nf} faces) x {nq} queries uniform in the "
                        "20%-padded bbox; FP32-compensated forward (FP32 pairs, FP64 accumulation, "
                        "FP64 re-evaluation of ill-conditioned pairs), inputs resident in HBM",
            "faces": nf,
            "queries": nq,
            "queries_per_lane": s32["qpl"],
            "splits": s32["splits"],
            "morton_sort": bool(s32["sort"]),
            "wall_ms_per_step": wall32,
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": FP32_VECTOR_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_VECTOR_PEAK_TFLOPS,
            "traffic": None,
        },
        "fp64_plain": result_fp64,
        "speedup_fp32c_vs_fp64": (result_fp64["ms_per_step"] / ms32) if result_fp64 else None,
        "accuracy": acc,
    }
    if not args.no_cpu_baseline:
        log("cpu baseline")
        line["cpu_baseline"] = cpu_baseline(tri, Q, args.cpu_seconds)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
