#!/usr/bin/env python
"""Throughput of the generalized-winding-number hot path on one MI355X.

Default workload (the metric's configuration): one step = one forward pass of
the FP32-compensated strip kernel (wn_mesh_forward on a resident igl.Mesh) over
the closed torus R = 1, r = 0.35 on a 120 x 118 grid (28,320 faces) x
1,048,576 queries uniform in the 20 %-padded bbox, inputs resident in HBM. The
same steps are timed for the FP32-compensated face-record kernel (wn_forward,
precision 1, no mesh preprocessing), for the FP64-plain kernel (precision 2:
the reference formula with the library's sqrt and atan2 for every pair), the
comparison point of the metric, and for the FP64 face and strip kernels with
the far-pair fast path; outputs are compared with each other and, on a
subsample, with the oracle.

Prints ONE JSON line. `value` = G query-face pairs/s of the FP32-compensated
strip kernel; `roofline` prices 66 algorithmic FLOP per pair (62 basic ops + 3 sqrt
+ 1 atan2, one each) against MI355X's published FP32 vector peak;
`cpu_baseline` is the FP64 oracle (a C restatement of the reference CPU path)
on the host cores over a bounded query sample.

Other workloads (one JSON line each, same timing method):
  --workload batch   512 meshes of uneven size in one persistent launch
                     (graph replay) vs 512 separate launches
  --workload stress  1.31M-face displaced sphere x 1M queries, error of the
                     FP32-compensated kernel by distance to the surface
  --workload grad    50k-face sphere x 262k queries, forward + adjoint
  --workload readme  the reference README's GPU row: 28,322-face open grid x
                     1,024 queries, end to end from host arrays
  --workload aabb    point-mesh squared distance: 243,602-face grid x 1M
                     points (and the README's 10k-point row end to end)
  --workload fast    hierarchical winding number on the 1.31M-face sphere x
                     1M queries, per accuracy parameter, vs all-pairs
  --workload copy    256 MiB device copy (calibrates the byte counters)
  --workload plain64 the plain FP64 kernel alone on the torus workload
  --workload face64  the FP64 face kernel (far-pair fast path) alone on it

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--workload NAME]
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


# Worker threads of the CPU baseline at most: the CPU share of one GPU on the
# measured boxes (their affinity mask shows more cores than the share).
CPU_THREAD_CAP = 16


def visible_cores() -> int:
    """Cores in this process's affinity mask, as the OS reports them."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def host_cores() -> int:
    return max(1, min(visible_cores(), CPU_THREAD_CAP))


def cpu_baseline(tri, Q, seconds: float):
    """FP64 oracle over a bounded sample of the same queries, all usable host
    cores up to CPU_THREAD_CAP (the affinity count itself is reported as
    host_cores_visible), contiguous query spans."""
    import oracle

    cores = host_cores()
    nf = tri.shape[1]
    oracle.winding_threads(tri, Q[: 16 * cores], cores)  # warm-up
    probe = Q[: 256 * cores]
    t0 = time.perf_counter()
    oracle.winding_threads(tri, probe, cores)
    rate = nf * len(probe) / max(time.perf_counter() - t0, 1e-6)
    # contiguous chunks of ~1 s from the start of the queries until the time
    # budget is spent (a long all-core run settles slower than a short probe)
    chunk = int(max(cores, rate / nf))
    nq, dt = 0, 0.0
    while nq < len(Q) and dt < seconds:
        t0 = time.perf_counter()
        oracle.winding_threads(tri, Q[nq: nq + chunk], cores)
        dt += time.perf_counter() - t0
        nq = min(len(Q), nq + chunk)
    # the reference's own thread policy: at most 4 workers (python/igl/_lib.py:37)
    n4 = max(4, nq // max(1, cores // 4) // 4)
    t0 = time.perf_counter()
    oracle.winding_threads(tri, Q[:n4], 4)
    dt4 = time.perf_counter() - t0
    return {
        "value": nf * nq / dt / 1e9,
        "unit": "G query-face pairs/s",
        "cores": cores,
        "host_cores_visible": visible_cores(),
        "thread_cap": CPU_THREAD_CAP,
        "kind": "port",
        "sample": f"first {nq} of the {len(Q)} queries x all {nf} faces, FP64, "
                  f"{cores} threads, {dt:.1f} s in ~1 s chunks after a warm-up",
        "four_workers": {"value": nf * n4 / dt4 / 1e9,
                         "sample": f"first {n4} queries, 4 threads, {dt4:.1f} s"},
    }


class Forward:
    """wn_forward on device tensors with a resolved configuration."""

    def __init__(self, h, trid, Qd, precision, qpl=0, splits=0, sort=-1, kappa=-1.0):
        import numpy as np
        import torch

        import igl

        self.h, self.trid, self.Qd, self.precision = h, trid, Qd, precision
        self.nf, self.nq = trid.shape[1], Qd.shape[0]
        cfg = np.zeros(4, dtype=np.int64)
        st = h.wn_resolve_config(self.nf, self.nq, precision, qpl, splits, sort, kappa, cfg.ctypes.data)
        assert st == 0, igl._lib.last_error()
        self.qpl, self.splits, self.sort = int(cfg[0]), int(cfg[1]), int(cfg[2])
        self.kappa = kappa
        self.nbytes = h.wn_forward_workspace_bytes(self.nf, self.nq, precision, self.qpl, self.splits,
                                                   self.sort)
        assert self.nbytes >= 0, igl._lib.last_error()
        self.ws = torch.empty(max(self.nbytes, 1), dtype=torch.uint8, device=Qd.device)
        self.out = torch.empty(self.nq, dtype=torch.float64, device=Qd.device)
        self.stream = torch.cuda.current_stream()

    def __call__(self):
        import igl

        st = self.h.wn_forward(self.trid.data_ptr(), self.Qd.data_ptr(), self.out.data_ptr(), self.nf,
                               self.nq, self.precision, self.qpl, self.splits, self.sort, self.kappa,
                               self.ws.data_ptr(), self.nbytes, self.stream.cuda_stream)
        if st != 0:
            raise RuntimeError(igl._lib.last_error())


def timed(fn, steps, warmup):
    """Device ms/step (HIP events on torch's current stream, which is the
    stream every call here launches on) and wall ms/step."""
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(s)
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    end.record(s)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return start.elapsed_time(end) / steps, wall / steps * 1e3


def timed_kernel(h, fn, steps, warmup):
    """timed(), plus the average device duration of the dominant kernel of
    each call over the timed steps: the library's kernel timer records HIP
    events on the stream that kernel is launched on, right before and after
    it. Returns (ms/step, wall ms/step, kernel ms/launch, launches)."""
    import numpy as np

    for _ in range(warmup):
        fn()
    out = np.zeros(2)
    h.wn_kernel_timer(1, 0)
    ms, wall = timed(fn, steps, 0)
    h.wn_kernel_timer(2, out.ctypes.data)
    h.wn_kernel_timer(0, 0)
    n = int(out[0])
    return ms, wall, (out[1] / n if n else float("nan")), n


def base_line(args, metric, value, unit, ms, dtype, workload, **extra):
    line = {
        "metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": {"workload": workload},
    }
    line.update(extra)
    return line


COUNTER_PROFILE = os.path.join(ROOT, "profiles", "fwd_torus28k_1M_counters.md")


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` on the torus workload, from the
    committed rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE, each scaled
    by the reading of a device copy of known size); None if absent."""
    try:
        text = open(COUNTER_PROFILE).read()
    except OSError:
        return None
    for row in text.splitlines():
        cells = [c.strip() for c in row.strip().strip("|").split("|")]
        if len(cells) == 5 and cells[0] == f"`{kernel}`":
            try:
                return (float(cells[3].replace(",", "")) + float(cells[4].replace(",", ""))) * 1e6
            except ValueError:
                return None
    return None


def roofline(pairs, ms, peak=FP32_VECTOR_PEAK_TFLOPS, kernel=None, launches=None,
             torus_traffic=True):
    """66 algorithmic FLOP per pair over `ms`, the kernel's average launch
    duration (kernel timer) where one is given, against `peak`; traffic from
    the torus workload's counter summary when the workload is the torus."""
    achieved = pairs * FLOP_PER_PAIR / (ms * 1e-3) / 1e12
    traffic = measured_traffic(kernel) if kernel and torus_traffic else None
    line = {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
            "frac": achieved / peak, "traffic": traffic}
    if kernel:
        line.update(kernel=kernel, kernel_ms=ms, launches=launches)
    return line


# ---------------------------------------------------------------------------


class MeshForward:
    """wn_mesh_forward (FP32c strip kernel) or wn_mesh_forward64 (FP64 strip
    kernel) on a resident mesh handle."""

    def __init__(self, h, mesh, Qd, splits=0, fp64=False):
        import torch

        import igl

        self.h, self.mesh, self.Qd = h, mesh, Qd
        self.nq = Qd.shape[0]
        self.splits, self.sort = mesh.resolve(self.nq, splits, -1)
        self.nbytes = h.wn_mesh_workspace_bytes(mesh.handle, self.nq, self.splits, self.sort)
        assert self.nbytes >= 0, igl._lib.last_error()
        self.ws = torch.empty(max(self.nbytes, 1), dtype=torch.uint8, device=Qd.device)
        self.out = torch.empty(self.nq, dtype=torch.float64, device=Qd.device)
        self.stream = torch.cuda.current_stream()
        self.fp64 = fp64

    def __call__(self):
        import igl

        if self.fp64:
            st = self.h.wn_mesh_forward64(self.mesh.handle, self.Qd.data_ptr(), self.out.data_ptr(),
                                          self.nq, self.splits, self.sort, self.ws.data_ptr(),
                                          self.nbytes, self.stream.cuda_stream)
        else:
            st = self.h.wn_mesh_forward(self.mesh.handle, self.Qd.data_ptr(), self.out.data_ptr(),
                                        self.nq, self.splits, self.sort, -1.0, self.ws.data_ptr(),
                                        self.nbytes, self.stream.cuda_stream)
        if st != 0:
            raise RuntimeError(igl._lib.last_error())


def run_torus(args, h):
    import numpy as np
    import torch

    import igl
    import oracle
    import wn_workloads as W
    from igl.mesh import Mesh

    log("generating workload")
    V, F, Q = W.config2(args.queries)
    tri = igl.pack_faces(V, F)
    nf, nq = len(F), len(Q)
    trid = torch.tensor(tri, device="cuda")
    Qd = torch.tensor(Q, device="cuda")

    t0 = time.perf_counter()
    mesh = Mesh(V, F)
    create_ms = (time.perf_counter() - t0) * 1e3
    info = mesh.info()
    strip = MeshForward(h, mesh, Qd, args.splits)
    log(f"strip path: {info['records']} records for {nf} faces (built in {create_ms:.1f} ms), "
        f"splits={strip.splits} sort={strip.sort}; {args.warmup} warm-up + {args.steps} timed steps")
    ms_strip, wall_strip, k_strip, n_strip = timed_kernel(h, strip, args.steps, args.warmup)
    log(f"fp32c strip {ms_strip:.3f} ms/step (device), {wall_strip:.3f} ms/step (wall), "
        f"fwd_strip_kernel {k_strip:.3f} ms/launch over {n_strip} launches")
    w_strip = strip.out.cpu().numpy()

    face = Forward(h, trid, Qd, 1, args.qpl, args.splits)
    ms_face, _, k_face, _ = timed_kernel(h, face, args.steps, args.warmup)
    log(f"fp32c face-record kernel {ms_face:.3f} ms/step (qpl={face.qpl} splits={face.splits}), "
        f"{k_face:.3f} ms/launch")
    w_face = face.out.cpu().numpy()

    result_fp64, acc = None, {}
    if not args.no_fp64:
        f64 = Forward(h, trid, Qd, 2)
        ms64, _, k64, _ = timed_kernel(h, f64, args.steps, args.warmup)
        log(f"fp64 plain {ms64:.3f} ms/step (device), fwd_f64_kernel {k64:.3f} ms/launch")
        w64 = f64.out.cpu().numpy()
        f64f = Forward(h, trid, Qd, 0)
        ms64f, _, k64f, _ = timed_kernel(h, f64f, args.steps, args.warmup)
        log(f"fp64 face kernel {ms64f:.3f} ms/step (device), fwd_f64f_kernel {k64f:.3f} ms/launch")
        w64f = f64f.out.cpu().numpy()
        d = np.abs(w_strip - w64)
        acc = {"max_abs_dw_fp32c_vs_fp64": float(d.max()), "mean_abs_dw_fp32c_vs_fp64": float(d.mean()),
               "max_abs_dw_face_kernel_vs_fp64": float(np.abs(w_face - w64).max()),
               "fp64_max_dist_from_integer": float(np.abs(w64 - np.round(w64)).max())}
        result_fp64 = {"value": nf * nq / (ms64 * 1e-3) / 1e9, "ms_per_step": ms64, "qpl": f64.qpl,
                       "splits": f64.splits, "kernel_ms": k64, "precision": 2,
                       "roofline_frac_fp64_peak": roofline(nf * nq, k64, FP64_VECTOR_PEAK_TFLOPS)["frac"]}
        # the FP64 face kernel of the product (precision 0): far pairs by the
        # arctangent polynomial, the rest by the plain formula
        result_fp64["face"] = {"value": nf * nq / (ms64f * 1e-3) / 1e9, "ms_per_step": ms64f,
                               "kernel_ms": k64f, "qpl": f64f.qpl, "splits": f64f.splits,
                               "speedup_vs_fp64_plain": ms64 / ms64f,
                               "roofline_frac_fp64_peak":
                                   roofline(nf * nq, k64f, FP64_VECTOR_PEAK_TFLOPS)["frac"]}
        acc["max_abs_dw_fp64_face_vs_fp64"] = float(np.abs(w64f - w64).max())
        sub = np.arange(0, nq, max(1, nq // 512))
        ref = oracle.winding_soa(tri, Q[sub])
        acc["max_abs_dw_fp64_vs_oracle_sub512"] = float(np.abs(w64[sub] - ref).max())
        acc["max_abs_dw_fp32c_vs_oracle_sub512"] = float(np.abs(w_strip[sub] - ref).max())
        acc["max_abs_dw_fp64_face_vs_oracle_sub512"] = float(np.abs(w64f[sub] - ref).max())
        s64 = MeshForward(h, mesh, Qd, args.splits, fp64=True)
        ms_s64, _, k_s64, _ = timed_kernel(h, s64, args.steps, args.warmup)
        log(f"fp64 strip {ms_s64:.3f} ms/step (device), {k_s64:.3f} ms/launch")
        w_s64 = s64.out.cpu().numpy()
        acc["max_abs_dw_fp64_strip_vs_fp64"] = float(np.abs(w_s64 - w64).max())
        acc["max_abs_dw_fp64_strip_vs_oracle_sub512"] = float(np.abs(w_s64[sub] - ref).max())
        result_fp64["strip"] = {"value": nf * nq / (ms_s64 * 1e-3) / 1e9, "ms_per_step": ms_s64,
                                "kernel_ms": k_s64, "speedup_vs_fp64_plain": ms64 / ms_s64,
                                "roofline_frac_fp64_peak":
                                    roofline(nf * nq, k_s64, FP64_VECTOR_PEAK_TFLOPS)["frac"]}
        # the drop-in call from host arrays, end to end (large calls take the
        # strip route: host stripification and upload unless the library still
        # holds this mesh from the previous call, strip kernel, download):
        # first call (mesh released before it) and repeated call
        def host_calls(precision):
            first, again = [], []
            for _ in range(2):
                igl.lib().wn_release_host_buffers()
                t0 = time.perf_counter()
                igl.winding_number(V, F, Q, device="gpu", precision=precision)
                first.append((time.perf_counter() - t0) * 1e3)
                t0 = time.perf_counter()
                w = igl.winding_number(V, F, Q, device="gpu", precision=precision)
                again.append((time.perf_counter() - t0) * 1e3)
            return min(first), min(again), w

        # the reference's own ABI entry (face SoA from the host, FP64): from
        # 262,144 queries it matches corners, stripifies and runs the FP64
        # strip kernel, keeping the strip mesh for a repeated call
        from igl._lib import addr

        w_abi = np.empty(nq)
        abi_ms = []
        for _ in range(2):
            igl.lib().wn_release_host_buffers()
            for _ in range(2):
                t0 = time.perf_counter()
                assert h.mli_winding_gpu(addr(tri), addr(Q), addr(w_abi), nf, nq) == 1
                abi_ms.append((time.perf_counter() - t0) * 1e3)
        result_fp64["reference_abi_mli_winding_gpu"] = {
            "first_call_ms": min(abi_ms[0::2]), "repeated_call_ms": min(abi_ms[1::2]),
            "repeated_value": nf * nq / (min(abi_ms[1::2]) * 1e-3) / 1e9,
            "note": "host arrays in and out (PCIe included), G pairs/s end to end"}
        acc["max_abs_dw_reference_abi_vs_fp64"] = float(np.abs(w_abi - w64).max())
        log(f"mli_winding_gpu 1M queries: first {min(abi_ms[0::2]):.2f} ms, repeated "
            f"{min(abi_ms[1::2]):.2f} ms")
        first64, e2e64, w_e2e = host_calls("fp64")
        result_fp64["host_call_end_to_end_ms"] = e2e64
        result_fp64["host_call_first_ms"] = first64
        acc["max_abs_dw_host_call_vs_fp64"] = float(np.abs(w_e2e - w64).max())
        kt = np.zeros(2)
        h.wn_kernel_timer(1, 0)
        first32, e2e32, w_e2e32 = host_calls("fp32c")
        h.wn_kernel_timer(2, kt.ctypes.data)
        h.wn_kernel_timer(0, 0)
        acc["max_abs_dw_host_call_fp32c_vs_fp64"] = float(np.abs(w_e2e32 - w64).max())
        # the same strip step again, after the FP64 runs and the host calls:
        # the kernel's duration on a device that has just run FP64 work
        _, _, k_after, n_after = timed_kernel(h, strip, max(1, args.steps // 2), 0)
        log(f"fwd_strip_kernel in the host calls {kt[1] / max(kt[0], 1):.3f} ms/launch "
            f"({int(kt[0])} launches); strip step again after them {k_after:.3f} ms/launch")
        strip_kernel_later = {"host_call_launches": int(kt[0]),
                              "host_call_kernel_ms": kt[1] / max(kt[0], 1),
                              "strip_step_again_kernel_ms": k_after,
                              "strip_step_again_launches": n_after}
    mesh.close()

    line = base_line(
        args, METRIC, nf * nq / (ms_strip * 1e-3) / 1e9, "G query-face pairs/s", ms_strip, "f32",
        f"closed torus R=1 r=0.35 120x118 ({nf} faces) x {nq} queries uniform in the 20%-padded "
        "bbox; FP32-compensated forward (FP32 pairs, FP64 accumulation, FP64 re-evaluation of "
        "ill-conditioned pairs) over the resident triangle-strip mesh, inputs resident in HBM",
        roofline=roofline(nf * nq, k_strip, kernel="wn::fwd_strip_kernel", launches=n_strip),
        fp64_plain=result_fp64,
        traffic_source="bytes per launch of wn::fwd_strip_kernel from profiles/fwd_torus28k_1M_counters.md "
                       "(rocprofv3 FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00, factors from a 256 MiB copy)",
        speedup_fp32c_vs_fp64=(result_fp64["ms_per_step"] / ms_strip) if result_fp64 else None,
        # the metric's ratio is FP32c strips over the plain FP64 kernel; the
        # same-algorithm ratios put each FP32c kernel beside its FP64 twin
        # (both with the far-pair fast path), and the face kernel beside the
        # plain one
        speedup_same_algorithm=({"face_vs_face": result_fp64["face"]["ms_per_step"] / ms_face,
                                 "strip_vs_strip": result_fp64["strip"]["ms_per_step"] / ms_strip,
                                 "face_vs_plain_face": result_fp64["ms_per_step"] / ms_face}
                                if result_fp64 else None),
        fp32c_face_kernel={"value": nf * nq / (ms_face * 1e-3) / 1e9, "ms_per_step": ms_face,
                           "kernel_ms": k_face, "qpl": face.qpl, "splits": face.splits},
        accuracy=acc)
    line["config"].update(faces=nf, queries=nq, strip_records=info["records"], splits=strip.splits,
                          morton_sort=bool(strip.sort), mesh_build_ms=create_ms,
                          wall_ms_per_step=wall_strip)
    if result_fp64 is not None:
        line["strip_kernel_after_fp64"] = strip_kernel_later
        # igl.winding_number(V, F, Q, device="gpu", precision="fp32c") from
        # host arrays, repeated on the same mesh (kernel, query upload, result
        # download) and first on it (plus stripify and mesh upload)
        line["host_call_end_to_end_ms"] = e2e32
        line["host_call_first_ms"] = first32
    if not args.no_cpu_baseline:
        log("cpu baseline")
        line["cpu_baseline"] = cpu_baseline(tri, Q, args.cpu_seconds)
    return line


def run_batch(args, h):
    import numpy as np
    import torch

    from igl import _autotune
    from igl.autograd import pack_faces
    from igl.batch import StripBatch, WindingBatch, pack_batch, pack_batch_strips

    import wn_workloads as W

    log("generating 512-mesh batch")
    B = W.config4()
    fc = np.array([len(b[1]) for b in B])
    qc = np.array([len(b[2]) for b in B])
    pairs = int((fc * qc).sum())
    tri = pack_batch([b[0] for b in B], [b[1] for b in B])
    Q = torch.tensor(np.concatenate([b[2] for b in B]), device="cuda")
    t0 = time.perf_counter()
    rec, rc, bnd = pack_batch_strips([b[0] for b in B], [b[1] for b in B], boundary=True)
    torch.cuda.synchronize()
    strip_build_s = time.perf_counter() - t0
    log(f"strip records {int(rc.sum())} for {int(fc.sum())} faces in {strip_build_s:.2f} s")
    log("autotuning strip chunk")
    t0 = time.perf_counter()
    chunk_tiles = _autotune.tune_batch_strips(rc, qc, rec, Q, use_cache=False, boundary=bnd)
    tune_strip_s = time.perf_counter() - t0
    splan = StripBatch(rc, qc, chunk_tiles=chunk_tiles, boundary=bnd)
    out = torch.empty(len(Q), dtype=torch.float64, device="cuda")
    ms_graph, wall_graph = timed(lambda: splan.run(rec, Q, out=out, use_graph=True), args.steps,
                                 args.warmup)
    log(f"batched strips fp32c (graph) {ms_graph:.3f} ms/step; chunk_tiles={chunk_tiles}")
    w_batch = out.cpu().numpy()
    sinfo = splan.info()
    splan.close()

    log("autotuning face batch (qpl x chunk)")
    t0 = time.perf_counter()
    qpl, chunk = _autotune.tune_batch(fc, qc, tri, Q, 1, use_cache=False)
    tune_s = time.perf_counter() - t0
    plan = WindingBatch(fc, qc, precision="fp32c", qpl=qpl, chunk_faces=chunk)
    ms_face, _ = timed(lambda: plan.run(tri, Q, out=out, use_graph=True), args.steps, args.warmup)
    log(f"batched face records fp32c (graph) {ms_face:.3f} ms/step; qpl={qpl} chunk={chunk}")
    w_face = out.cpu().numpy()

    # 512 separate launches of the same strip kernels (per-mesh resident
    # strip meshes, built outside the timed region), and of the face-record
    # forward (per-mesh SoA packed beforehand; each call also finds its
    # mesh's boundary edges)
    from igl.mesh import Mesh

    meshes, singles, faces_single = [], [], []
    off = 0
    for V, F, Qm in B:
        Qs = Q[off:off + len(Qm)]
        if len(F):
            meshes.append(Mesh(V, F))
            singles.append(MeshForward(h, meshes[-1], Qs))
        tri_m = pack_faces(torch.tensor(V, device="cuda"), torch.tensor(F, device="cuda"))
        faces_single.append(Forward(h, tri_m, Qs, 1))
        off += len(Qm)

    def run_all(calls):
        for f in calls:
            f()

    ms_sep, _ = timed(lambda: run_all(singles), args.steps, args.warmup)
    log(f"512 separate strip launches {ms_sep:.3f} ms/step")
    ms_sep_face, _ = timed(lambda: run_all(faces_single), args.steps, args.warmup)
    log(f"512 separate face-record launches {ms_sep_face:.3f} ms/step")
    w_sep = np.concatenate([f.out.cpu().numpy() for f in faces_single])
    for m in meshes:
        m.close()
    plan64 = WindingBatch(fc, qc, precision="fp64", qpl=1, chunk_faces=chunk)
    w64 = plan64.run(tri, Q, use_graph=False).cpu().numpy()
    ms64, _ = timed(lambda: plan64.run(tri, Q, out=out, use_graph=True), max(1, args.steps // 2), 1)
    plan64.close()
    splan64 = StripBatch(rc, qc, chunk_tiles=chunk_tiles, precision="fp64")
    ms64s, _ = timed(lambda: splan64.run(rec, Q, out=out, use_graph=True), max(1, args.steps // 2), 1)
    w64s = out.cpu().numpy()
    splan64.close()
    log(f"batched strips fp64 (graph) {ms64s:.3f} ms/step")
    info = plan.info()
    plan.close()
    d = np.abs(w_batch - w64)
    open_mesh = np.repeat(bnd.edge_counts > 0, qc)  # per query: its mesh has a boundary
    return base_line(
        args, "G query-face pairs/s, batched FP32-comp, 512 uneven meshes, 1 MI355X",
        pairs / (ms_graph * 1e-3) / 1e9, "G query-face pairs/s", ms_graph, "f32",
        f"512 meshes (torus / geodesic sphere / open cylinder with holes), {int(fc.sum())} faces, "
        f"{int(qc.sum())} queries, {pairs} pairs; triangle-strip records, one persistent launch "
        "replayed from a hipGraph",
        roofline=roofline(pairs, ms_graph),
        face_records_batched={"value": pairs / (ms_face * 1e-3) / 1e9, "ms_per_step": ms_face},
        separate_launches={"value": pairs / (ms_sep * 1e-3) / 1e9, "ms_per_step": ms_sep,
                           "kind": "wn_mesh_forward per mesh (the batch's strip kernels)"},
        separate_face_launches={"value": pairs / (ms_sep_face * 1e-3) / 1e9,
                                "ms_per_step": ms_sep_face,
                                "kind": "wn_forward FP32c per mesh (face records)"},
        speedup_batched_vs_separate=ms_sep / ms_graph,
        fp64_batched={"value": pairs / (ms64 * 1e-3) / 1e9, "ms_per_step": ms64,
                      "strips": {"value": pairs / (ms64s * 1e-3) / 1e9, "ms_per_step": ms64s,
                                 "max_abs_dw_vs_fp64_faces": float(np.abs(w64s - w64).max())}},
        strip_records={"records": int(rc.sum()), "host_build_seconds": strip_build_s,
                       "boundary_edges": bnd.n_edges,
                       "open_meshes": int((bnd.edge_counts > 0).sum())},
        tuning={"chunk_tiles": chunk_tiles, "seconds": tune_strip_s, "items": sinfo["items"],
                "grid": sinfo["grid"],
                "face_batch": {"qpl": qpl, "chunk_faces": chunk, "seconds": tune_s,
                               "items": info["items"], "grid": info["grid"]}},
        accuracy={"max_abs_dw_fp32c_vs_fp64": float(d.max()), "mean_abs_dw": float(d.mean()),
                  "max_abs_dw_open_meshes": float(d[open_mesh].max(initial=0.0)),
                  "max_abs_dw_closed_meshes": float(d[~open_mesh].max(initial=0.0)),
                  "max_abs_dw_face_batch_vs_fp64": float(np.abs(w_face - w64).max()),
                  "max_abs_dw_batched_vs_separate": float(np.abs(w_face - w_sep).max())})


def run_fp64_face(args, h, precision):
    """One FP64 face kernel alone on the torus workload (wn_forward): the
    plain kernel (precision 2), the comparison point of the metric, or the
    FP64 face kernel with the far-pair fast path (precision 0); for
    profiling."""
    import torch

    import igl
    import wn_workloads as W

    V, F, Q = W.config2(args.queries)
    tri = igl.pack_faces(V, F)
    nf, nq = len(F), len(Q)
    f64 = Forward(h, torch.tensor(tri, device="cuda"), torch.tensor(Q, device="cuda"), precision)
    ms, _, k, n = timed_kernel(h, f64, args.steps, args.warmup)
    plain = precision == 2
    kernel = "wn::fwd_f64_kernel" if plain else "wn::fwd_f64f_kernel"
    log(f"{'fp64 plain' if plain else 'fp64 face'} {ms:.3f} ms/step, {kernel} {k:.3f} ms/launch")
    what = ("plain FP64 face kernel" if plain else
            "FP64 face kernel, far pairs by the arctangent polynomial")
    return base_line(
        args, f"G query-face pairs/s, {'FP64 plain' if plain else 'FP64 face'} kernel, 1 MI355X",
        nf * nq / (ms * 1e-3) / 1e9, "G query-face pairs/s", ms, "f64",
        f"closed torus 120x118 ({nf} faces) x {nq} queries; {what} (wn_forward precision "
        f"{precision}, qpl {f64.qpl}, {f64.splits} splits)",
        roofline=roofline(nf * nq, k, FP64_VECTOR_PEAK_TFLOPS, kernel=f"{kernel}<{f64.qpl}>",
                          launches=n, torus_traffic=False))


def run_plain64(args, h):
    return run_fp64_face(args, h, 2)


def run_face64(args, h):
    return run_fp64_face(args, h, 0)


def run_stress(args, h):
    import numpy as np
    import torch

    import igl
    import wn_workloads as W
    from igl.mesh import Mesh

    log("generating 1.31M-face displaced sphere")
    V, F = W.config3_mesh()
    Q, dist = W.config3_queries(V, F, args.queries)
    tri = igl.pack_faces(V, F)
    nf, nq = len(F), len(Q)
    trid = torch.tensor(tri, device="cuda")
    Qd = torch.tensor(Q, device="cuda")
    t0 = time.perf_counter()
    mesh = Mesh(V, F)
    create_ms = (time.perf_counter() - t0) * 1e3
    strip = MeshForward(h, mesh, Qd)
    ms32, _, k32, n32 = timed_kernel(h, strip, args.steps, min(args.warmup, 1))
    w32 = strip.out.cpu().numpy()
    face = Forward(h, trid, Qd, 1)
    ms_face, _ = timed(face, max(1, args.steps // 2), min(args.warmup, 1))
    w_face = face.out.cpu().numpy()
    f64 = Forward(h, trid, Qd, 2)
    ms64, _ = timed(f64, max(1, args.steps // 2), min(args.warmup, 1))
    w64 = f64.out.cpu().numpy()
    f64f = Forward(h, trid, Qd, 0)
    ms64f, _ = timed(f64f, max(1, args.steps // 2), min(args.warmup, 1))
    w64f = f64f.out.cpu().numpy()
    mesh.close()
    uni = np.isnan(dist)
    mag = np.abs(dist)
    near = ~uni
    expect = np.where(dist[near] < 0, 1.0, 0.0)

    def buckets_of(w):
        d = np.abs(w - w64)
        out = {"uniform": {"n": int(uni.sum()), "max": float(d[uni].max()), "mean": float(d[uni].mean())}}
        for lo in range(-7, -2):
            sel = near & (mag >= 10.0 ** lo) & (mag < 10.0 ** (lo + 1))
            if sel.any():
                out[f"1e{lo}..1e{lo + 1}"] = {"n": int(sel.sum()), "max": float(d[sel].max()),
                                               "mean": float(d[sel].mean())}
        return out

    return base_line(
        args, "G query-face pairs/s, FP32-comp precision stress, 1 MI355X",
        nf * nq / (ms32 * 1e-3) / 1e9, "G query-face pairs/s", ms32, "f32",
        f"displaced icosphere level 8 ({nf} faces) x {nq} queries (half uniform, half at "
        "|d| in [1e-7, 1e-2] along the normal); FP32-compensated strip kernel",
        roofline=roofline(nf * nq, k32, kernel="wn::fwd_strip_kernel", launches=n32,
                          torus_traffic=False),
        fp64_plain={"value": nf * nq / (ms64 * 1e-3) / 1e9, "ms_per_step": ms64,
                    "face": {"value": nf * nq / (ms64f * 1e-3) / 1e9, "ms_per_step": ms64f,
                             "max_abs_dw_vs_fp64_plain": float(np.abs(w64f - w64).max())}},
        fp32c_face_kernel={"value": nf * nq / (ms_face * 1e-3) / 1e9, "ms_per_step": ms_face},
        speedup_fp32c_vs_fp64=ms64 / ms32, mesh_build_ms=create_ms,
        accuracy={"max_abs_dw_fp32c_vs_fp64": float(np.abs(w32 - w64).max()),
                  "by_distance_strip": buckets_of(w32), "by_distance_face_kernel": buckets_of(w_face),
                  "near_surface_side_errors_fp64": int((np.abs(w64[near] - expect) > 0.5).sum()),
                  "near_surface_side_errors_fp32c": int((np.abs(w32[near] - expect) > 0.5).sum())})


def run_grad(args, h):
    import numpy as np
    import torch

    import igl.autograd as A
    import wn_workloads as W

    V, F, Q, t = W.config5()
    Vd = torch.tensor(V, device="cuda", requires_grad=True)
    Fd = torch.tensor(F, device="cuda")
    Od = torch.tensor(Q, device="cuda", requires_grad=True)
    td = torch.tensor(t, device="cuda")

    def step():
        Vd.grad = None
        Od.grad = None
        w = A.winding_number(Vd, Fd, Od, "fp32c")
        ((w - td) ** 2).sum().backward()

    # the first calls of the process also load the strip, sort and adjoint
    # kernels: two more warm-up steps than the others
    ms, _ = timed(step, args.steps, args.warmup + 2)
    A.release()

    def step_face_route():
        Vd.grad = None
        Od.grad = None
        w = A.WindingNumber.apply(Vd, Fd, Od, "fp32c", "edges")
        ((w - td) ** 2).sum().backward()

    ms_face_route, _ = timed(step_face_route, args.steps, args.warmup)
    field = A.WindingField(Fd, len(V))
    field_faces = A.WindingField(Fd, len(V), adjoint="faces")

    def step_with(fld):
        def run():
            Vd.grad = None
            Od.grad = None
            w = fld(Vd, Od)
            ((w - td) ** 2).sum().backward()
        return run

    ms_faces, _ = timed(step_with(field_faces), args.steps, args.warmup)
    ms_field, _ = timed(step_with(field), args.steps, args.warmup)
    field_fwd_ms, _ = timed(lambda: field(Vd.detach(), Od.detach()), args.steps, 1)
    n_edges = int(field.weight.shape[0])
    field.close()
    field_faces.close()
    pairs = len(F) * len(Q)
    tri = A.pack_faces(Vd.detach(), Fd)
    Qd = Od.detach().double()
    fwd_ms, _ = timed(lambda: A.forward(tri, Qd, "fp32c"), args.steps, 1)
    g = torch.ones(len(Q), dtype=torch.float64, device="cuda")
    bwd_ms, _ = timed(lambda: A.backward(tri, Qd, g), args.steps, 1)
    edge_ms, _ = timed(lambda: A.backward_edges(Vd.detach(), field.edges, field.weight, Qd, g),
                       args.steps, 1)

    # the same sphere with a polar cap cut away (an open surface): the edge
    # adjoint's work is its boundary rim
    cz = V[F].mean(axis=1)[:, 2]
    Fo = F[cz < 0.6 * np.abs(V[:, 2]).max()]
    Fod = torch.tensor(Fo, device="cuda")
    eo, mo = A.boundary_edges(Fo, len(V))
    tri_o = A.pack_faces(Vd.detach(), Fod)
    gr = torch.tensor(np.random.default_rng(504).standard_normal(len(Q)), device="cuda")
    open_face_ms, _ = timed(lambda: A.backward(tri_o, Qd, gr), args.steps, 1)
    open_edge_ms, _ = timed(lambda: A.backward_edges(Vd.detach(), eo, mo, Qd, gr), args.steps, 1)
    fq, ftri = A.backward(tri_o, Qd, gr)
    fV = A.scatter_corners(ftri, Fod, len(V))
    eq, eV = A.backward_edges(Vd.detach(), eo, mo, Qd, gr)
    diff = max(float((eq - fq).abs().max() / fq.abs().max()),
               float((eV - fV).abs().max() / fV.abs().max()))
    return base_line(
        args, "G query-face pairs/s, forward + adjoint (dL/dO, dL/dV), 1 MI355X",
        pairs / (ms_field * 1e-3) / 1e9, "G query-face pairs/s", ms_field, "f32+f64",
        f"frequency-50 geodesic sphere ({len(F)} faces, float32) x {len(Q)} queries, "
        "L = sum (w - t)^2; igl.autograd.WindingField: strip records gathered from V on "
        "the device each step, FP32-compensated strip forward + FP64 edge adjoint (the "
        f"closed sphere has {n_edges} boundary edges: exact zero gradient)",
        forward_ms=field_fwd_ms, edge_adjoint_ms=edge_ms, face_adjoint_ms=bwd_ms,
        face_adjoint_step_ms=ms_faces,
        face_adjoint_pairs_per_s=pairs / (bwd_ms * 1e-3) / 1e9,
        autograd_winding_number={"ms_per_step": ms,
                                 "note": "igl.autograd.winding_number at this size (>= "
                                         "STRIP_MIN_QUERIES): strip route through a kept "
                                         "WindingField (FP32c strip forward + FP64 edge "
                                         "adjoint; topology and boundary edges built on the "
                                         "first call and reused)"},
        face_record_route={"ms_per_step": ms_face_route, "forward_ms": fwd_ms,
                           "note": "igl.autograd.WindingNumber, the route below the strip "
                                   "threshold: device face pack, FP32c face-record forward + "
                                   "FP64 edge adjoint (boundary edges from F on the host)"},
        open_cap={"faces": len(Fo), "boundary_edges": int(mo.shape[0]),
                  "face_adjoint_ms": open_face_ms, "edge_adjoint_ms": open_edge_ms,
                  "edge_pairs_per_s": int(mo.shape[0]) * len(Q) / (open_edge_ms * 1e-3) / 1e9,
                  "max_rel_diff_edges_vs_faces": diff},
        gradient_max={"dL/dV": float(Vd.grad.abs().max()), "dL/dO": float(Od.grad.abs().max())})


def run_readme(args, h):
    import numpy as np

    import igl
    import oracle
    import wn_workloads as W

    V, F = W.bench_grid(120)
    Q = np.random.default_rng(0).uniform([0, 0, -2], [119, 119, 2], size=(1024, 3))
    best = {}
    for name, kw in (("gpu_fp64", dict(device="gpu")), ("gpu_fp32c", dict(device="gpu", precision="fp32c")),
                     ("cpu", dict(device="cpu"))):
        igl.winding_number(V, F, Q, **kw)
        times = []
        for _ in range(max(3, args.steps)):
            t0 = time.perf_counter()
            w = igl.winding_number(V, F, Q, **kw)
            times.append(time.perf_counter() - t0)
        best[name] = (min(times) * 1e3, w)
    ref = oracle.winding(V, F, Q)
    pairs = len(F) * len(Q)
    ms = best["gpu_fp64"][0]
    return base_line(
        args, "G query-face pairs/s end to end, reference README GPU row shape, 1 MI355X",
        pairs / (ms * 1e-3) / 1e9, "G query-face pairs/s", ms, "f64",
        f"open height-field grid(120) ({len(F)} faces) x 1024 queries, igl.winding_number from host "
        "arrays (pack, H2D, kernel, D2H), best of runs",
        reference_published_ms={"mojo_gpu": 53.47, "mojo_cpu_4threads": 215.09, "libigl_cpu": 138.86},
        vs_reference_gpu_row=53.47 / ms,
        end_to_end_ms={k: v[0] for k, v in best.items()},
        accuracy={k: float(np.abs(v[1] - ref).max()) for k, v in best.items()})


def run_aabb(args, h):
    """Point-mesh squared distance on the reference README's AABB row shape:
    grid(350) (243,602 faces) x 10,000 points, plus 1,048,576 points for
    throughput. GPU results are checked bit for bit against the CPU path."""
    import numpy as np
    import torch

    import igl
    import wn_workloads as W

    V, F = W.bench_grid(350)
    rng = np.random.default_rng(0)
    P = rng.uniform([-10, -10, -5], [360, 360, 5], size=(10_000, 3))
    t0 = time.perf_counter()
    tree = igl.AABB()
    tree.init(V, F)
    init_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    gpu = tree.squared_distance(V, F, P, device="gpu")  # also makes the device tree
    first_call_ms = (time.perf_counter() - t0) * 1e3

    def best_of(fn, reps):
        fn()
        best = float("inf")
        for _ in range(reps):
            t = time.perf_counter()
            fn()
            best = min(best, time.perf_counter() - t)
        return best * 1e3

    e2e_ms = best_of(lambda: tree.squared_distance(V, F, P, device="gpu"), max(3, args.steps))
    cpu_ms = best_of(lambda: tree.squared_distance(V, F, P), 3)
    cpu = tree.squared_distance(V, F, P)
    same_10k = all(np.array_equal(a, b) for a, b in zip(cpu, gpu))
    Pd = torch.tensor(P, device="cuda")
    dev_ms, _ = timed(lambda: tree.query_device(Pd), args.steps, args.warmup)
    log(f"AABB 10k: device {dev_ms:.3f} ms, end to end {e2e_ms:.3f} ms, CPU {cpu_ms:.3f} ms")

    P1 = rng.uniform([-10, -10, -5], [360, 360, 5], size=(1 << 20, 3))
    P1d = torch.tensor(P1, device="cuda")
    ms_1m, _ = timed(lambda: tree.query_device(P1d), args.steps, args.warmup)
    ms_1m_sorted, _ = timed(lambda: tree.query_device(P1d, sort=1), args.steps, args.warmup)
    g1 = tuple(t.cpu().numpy() for t in tree.query_device(P1d))
    sub = np.random.default_rng(1).choice(len(P1), 65536, replace=False)
    c1 = tree.squared_distance(V, F, P1[sub])
    same_1m = all(np.array_equal(a[sub], b) for a, b in zip(g1, c1))
    log(f"AABB 1M: {ms_1m:.3f} ms in launch order, {ms_1m_sorted:.3f} ms with the Morton sort")
    info = tree.gpu_info()
    return base_line(
        args, "M point-mesh squared-distance queries/s, 1 MI355X",
        len(P1) / (ms_1m * 1e-3) / 1e6, "M queries/s", ms_1m, "f64",
        f"open height-field grid(350) ({len(F)} faces, {info['nodes']} tree nodes) x "
        f"{len(P1)} points uniform in its padded box; median-split tree, one FP64 descent per "
        "point",
        morton_sorted={"value": len(P1) / (ms_1m_sorted * 1e-3) / 1e6, "ms_per_step": ms_1m_sorted},
        readme_row={"points": len(P), "device_ms": dev_ms, "end_to_end_ms": e2e_ms,
                    "cpu_ms_host_threads": cpu_ms, "tree_init_ms": init_ms,
                    "first_gpu_call_ms": first_call_ms,
                    "reference_published_ms": {"mojo_cpu": 9.12, "libigl_cpu": 4.83},
                    "vs_reference_mojo_end_to_end": 9.12 / e2e_ms,
                    "vs_libigl_end_to_end": 4.83 / e2e_ms},
        parity={"gpu_equals_cpu_bitwise_10k": bool(same_10k),
                "gpu_equals_cpu_bitwise_1m_subsample_65536": bool(same_1m)},
        tree=info)


def run_fast(args, h):
    """Hierarchical winding number on the stress mesh (1.31M faces x 1M
    queries, half within 1e-7..1e-2 of the surface): time per beta, error
    against the exact FP64 winding number on a 65,536-query subsample, and
    the all-pairs FP32-compensated strip kernel on the same queries."""
    import numpy as np
    import torch

    import igl
    import wn_workloads as W
    from igl.mesh import Mesh

    V, F = W.config3_mesh()
    Q, dist = W.config3_queries(V, F, args.queries)
    t0 = time.perf_counter()
    fw = igl.FastWindingNumber(V, F)
    Qd = torch.tensor(Q, device="cuda")
    fw.query(Qd[:64], device="gpu")  # device copy of the tree
    build_ms = (time.perf_counter() - t0) * 1e3
    sub = np.random.default_rng(9).choice(len(Q), 65536, replace=False)
    exact = igl.winding_number(V, F, Q[sub], device="gpu")
    rows = {}
    for beta in (2.0, 4.0, 8.0):
        ms, _ = timed(lambda: fw.query(Qd, beta=beta, device="gpu"), args.steps, args.warmup)
        w = fw.query(Qd, beta=beta, device="gpu").cpu().numpy()
        err = np.abs(w[sub] - exact)
        near = ~np.isnan(dist[sub])
        rows[f"beta_{beta:g}"] = {
            "ms_per_step": ms, "queries_per_s": len(Q) / (ms * 1e-3),
            "max_abs_dw": float(err.max()), "mean_abs_dw": float(err.mean()),
            "max_abs_dw_near_surface": float(err[near].max()),
            "max_abs_dw_uniform": float(err[~near].max())}
        log(f"fast winding beta={beta}: {ms:.2f} ms, max err {err.max():.2e}")
    ms_given, _ = timed(lambda: fw.query(Qd, beta=2.0, device="gpu", sort=0), args.steps,
                        args.warmup)
    m = Mesh(V, F)
    out = torch.empty(len(Q), dtype=torch.float64, device="cuda")
    ms_strip, _ = timed(lambda: m.winding_number(Qd, out=out), max(1, args.steps // 2), 1)
    m.close()
    info = fw.info()
    fw.close()
    best = rows["beta_2"]
    return base_line(
        args, "M winding-number queries/s, hierarchical (beta 2), 1 MI355X",
        best["queries_per_s"] / 1e6, "M queries/s", best["ms_per_step"], "f64",
        f"displaced icosphere level 8 ({len(F)} faces, {info['nodes']} tree nodes) x {len(Q)} "
        "queries (half uniform, half at |d| in [1e-7, 1e-2]); second-order far-field "
        "expansions, exact FP64 leaves, queries walked in Morton order",
        by_beta=rows, tree_build_ms=build_ms,
        beta_2_queries_in_given_order_ms=ms_given,
        all_pairs_fp32c_strip_ms=ms_strip,
        speedup_beta2_vs_all_pairs_strip=ms_strip / best["ms_per_step"])


def run_copy(args, h):
    """Device-to-device copy of a known size (256 MiB read + 256 MiB written
    per launch): the reference point for scaling the FETCH_SIZE / WRITE_SIZE
    counters of the winding kernels to bytes."""
    import torch

    n = 256 * 2**20
    src = torch.ones(n // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    ms, _ = timed(lambda: dst.copy_(src), args.steps, args.warmup)
    return base_line(args, "GB/s device copy (counter calibration)", 2 * n / (ms * 1e-3) / 1e9, "GB/s",
                     ms, "f32", f"torch copy_ of {n} bytes (read {n} + write {n} bytes per launch)",
                     bytes_read_per_launch=n, bytes_written_per_launch=n)


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="torus",
                    choices=("torus", "batch", "stress", "grad", "readme", "aabb", "fast", "copy",
                             "plain64", "face64"))
    ap.add_argument("--queries", type=int, default=1_048_576)
    ap.add_argument("--qpl", type=int, default=0, help="queries per lane (0 = library default)")
    ap.add_argument("--splits", type=int, default=0, help="face splits (0 = library default)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64-plain comparison")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the winding-number path runs on a single MI355X here).", file=sys.stderr)
        return 2

    import torch  # noqa: F401  (load PyTorch's HIP runtime before the library)

    import igl

    h = igl.lib()
    if h.wn_device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        return 1
    runner = {"torus": run_torus, "batch": run_batch, "stress": run_stress, "grad": run_grad,
              "aabb": run_aabb, "fast": run_fast, "readme": run_readme,
              "copy": run_copy, "plain64": run_plain64, "face64": run_face64}[args.workload]
    print(json.dumps(runner(args, h)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
