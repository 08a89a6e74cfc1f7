#!/usr/bin/env python3
"""Benchmark of the ncarray device hot path on one MI355X.

Workload (one "step"): 3-D 7-point Laplacian with zero Dirichlet boundary on a float32
512x512x512 C-order grid (values N(0,1)), followed by the sum along axis 0 (512x512) and
the global max of the Laplacian.  The step runs as ONE fused kernel (stencil with the
axis-0 sum and the max in its epilogue).  Reported `value` is the HBM throughput of that
step: algorithmic bytes (read x once, write the Laplacian once, write the 512x512 sums)
divided by the measured step time, inputs already resident in HBM.

Also reported: the unfused pipeline (stencil, axis reduction, full reduction: three API
calls as the reference composes them), the BF16-storage / FP32-accumulate variant with
compensated sums, errors against an FP64 evaluation, the roofline fraction of the
dominant kernel, and the CPU host baseline (the oracle's restatement of the reference
host path, timed on this machine's cores).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
SHAPE = (512, 512, 512)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-errors", action="store_true", help="skip the FP64 error evaluation")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 512 MiB device copy (for PMC byte-counter calibration)")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch  # device memory and streams (torch's HIP runtime is shared with libnca_hip)
    from ncarray_amd import capi, tapes
    from ncarray_amd import torch_interop as ti

    dev = torch.device("cuda:0")
    capi.init(0)
    stream = torch.cuda.current_stream()
    sptr = ti.stream_ptr(stream)
    n = SHAPE[0] * SHAPE[1] * SHAPE[2]

    g = torch.Generator(device=dev).manual_seed(0)
    x = torch.randn(SHAPE, generator=g, device=dev, dtype=torch.float32)
    lap = torch.empty_like(x)
    col = torch.empty(SHAPE[1:], dtype=torch.float32, device=dev)
    mx = torch.empty((), dtype=torch.float32, device=dev)

    st = capi.Stencil(tapes.LAPLACIAN_OFFSETS, tapes.laplacian(), capi.FLOAT32, capi.FLOAT32, capi.FLOAT32,
                      [capi.FLOAT32])
    vx, vlap, vcol = ti.view(x), ti.view(lap), ti.view(col)
    mode = capi.ACCUM_COMPENSATED

    def fused_step():
        st.apply_reduce(vx, vlap, capi.BOUNDARY_ZERO, axis0=vcol, fuse=capi.FUSE_MAX, mode=mode,
                        max_out=mx.data_ptr(), scalars=(6.0,), stream=sptr)

    def time_loop(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        t0 = time.perf_counter()
        for a, b in evs:
            a.record(stream)
            fn()
            b.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ms = [a.elapsed_time(b) for a, b in evs]
        return ms, wall

    # ---- headline: fused FP32 step (one kernel per step) ----
    step_ms, wall = time_loop(fused_step, args.steps, args.warmup)
    ms_per_step = float(np.mean(step_ms))
    med_ms = float(np.median(step_ms))
    bytes_fused = 8 * n + 4 * SHAPE[1] * SHAPE[2]  # read x, write lap, write the axis-0 sums
    value = bytes_fused / (ms_per_step * 1e-3) / 1e9

    # per-kernel durations: the library brackets each of its launches with HIP events
    capi.profile_enable(True)
    for _ in range(args.steps):
        fused_step()
    torch.cuda.synchronize()
    prof = capi.profile_read()
    capi.profile_enable(False)
    per_kernel = {}
    for name, kms in prof:
        per_kernel.setdefault(name, []).append(kms)
    march_ms = float(np.mean(per_kernel.get("nca_stencil_march", [ms_per_step])))
    kernel_bytes = 8 * n  # dominant kernel: reads x once, writes the Laplacian once
    achieved = kernel_bytes / (march_ms * 1e-3) / 1e9

    # ---- unfused pipeline: stencil -> sum(axis 0) -> max (three API calls) ----
    col2 = torch.empty_like(col)
    mx2 = torch.empty((), dtype=torch.float32, device=dev)
    ax = (capi.ctypes.c_int64 * 1)(0)
    vcol2 = ti.view(col2)

    def unfused_step():
        st.apply(vx, vlap, capi.BOUNDARY_ZERO, scalars=(6.0,), stream=sptr)
        capi.call("nca_hip_reduce_axes", capi.ctypes.byref(vlap), ax, 1, capi.REDUCE_SUM, mode,
                  capi.ctypes.byref(vcol2), sptr)
        capi.call("nca_hip_reduce_full", capi.ctypes.byref(vlap), capi.REDUCE_MAX, 0, mx2.data_ptr(), sptr)

    un_ms, _ = time_loop(unfused_step, max(3, args.steps // 2), 2)
    un_mean = float(np.mean(un_ms))
    bytes_unfused = 8 * n + (4 * n + 4 * SHAPE[1] * SHAPE[2]) + 4 * n  # per API op, BASELINE accounting
    stencil_only_ms, _ = time_loop(lambda: st.apply(vx, vlap, capi.BOUNDARY_ZERO, scalars=(6.0,), stream=sptr),
                                   max(3, args.steps // 2), 2)

    # ---- BF16 storage, FP32 arithmetic, compensated sums ----
    xb = x.to(torch.bfloat16)
    lapb = torch.empty_like(xb)
    colb = torch.empty(SHAPE[1:], dtype=torch.float32, device=dev)
    mxb = torch.empty((), dtype=torch.bfloat16, device=dev)
    stb = capi.Stencil(tapes.LAPLACIAN_OFFSETS, tapes.laplacian(), capi.BFLOAT16, capi.FLOAT32, capi.BFLOAT16,
                       [capi.FLOAT32])
    vxb, vlapb, vcolb = ti.view(xb), ti.view(lapb), ti.view(colb)

    def bf16_step():
        stb.apply_reduce(vxb, vlapb, capi.BOUNDARY_ZERO, axis0=vcolb, fuse=capi.FUSE_MAX, mode=capi.ACCUM_COMPENSATED,
                         max_out=mxb.data_ptr(), scalars=(6.0,), stream=sptr)

    bf_ms, _ = time_loop(bf16_step, max(3, args.steps // 2), 2)
    bf_mean = float(np.mean(bf_ms))
    bytes_bf16 = 4 * n + 4 * SHAPE[1] * SHAPE[2]

    # ---- plain copy of 512 MiB (PMC calibration: known bytes) ----
    copy_info = None
    if args.calibrate_copy:
        src = torch.empty(1 << 27, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        cms, _ = time_loop(lambda: dst.copy_(src), 5, 2)
        copy_info = {"bytes_read": src.numel() * 4, "bytes_written": dst.numel() * 4, "ms": float(np.mean(cms)),
                     "GBps": 2 * src.numel() * 4 / (np.mean(cms) * 1e-3) / 1e9}
        del src, dst

    # ---- errors vs FP64 (outside the timed region) ----
    errors = None
    if not args.no_errors:
        fused_step()
        bf16_step()
        torch.cuda.synchronize()
        errors = fp64_errors(x, lap, col, mx, xb, lapb, colb, mxb)

    # ---- CPU baseline (oracle restatement of the reference host path) ----
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(x.cpu().numpy(), bytes_fused)

    out = {
        "metric": "HBM GB/s (% of 8 TB/s) fwd vs fwd+bwd, err vs FP64; CPU host baseline",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: x ~ N(0,1) float32 512^3 from torch.Generator(seed 0), resident in HBM",
        "config": {
            "workload": "3-D 7-point Laplacian, zero Dirichlet, float32 512x512x512 C-order -> sum(axis 0) "
                        "-> global max; one fused kernel per step (compensated sums)",
            "shape": list(SHAPE),
            "algorithmic_bytes_per_step": bytes_fused,
            "median_ms_per_step": round(med_ms, 4),
            "frac_of_8TBps": round(value / HBM_PEAK_GBPS, 4),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": None,
            "kernel": "nca_stencil_march (Laplacian + axis-0 sum + max epilogue)",
            "bytes_per_launch": kernel_bytes,
            "avg_launch_ms": round(march_ms, 4),
            "per_kernel_avg_ms": {k: round(float(np.mean(v)), 4) for k, v in per_kernel.items()},
        },
        "unfused_pipeline": {
            "ms_per_step": round(un_mean, 4),
            "algorithmic_bytes": bytes_unfused,
            "GBps": round(bytes_unfused / (un_mean * 1e-3) / 1e9, 2),
            "stencil_only_ms": round(float(np.mean(stencil_only_ms)), 4),
            "stencil_only_GBps": round(8 * n / (np.mean(stencil_only_ms) * 1e-3) / 1e9, 2),
        },
        "bf16_storage": {
            "ms_per_step": round(bf_mean, 4),
            "algorithmic_bytes": bytes_bf16,
            "GBps": round(bytes_bf16 / (bf_mean * 1e-3) / 1e9, 2),
            "accumulate": "float32 compensated (Neumaier) per thread, float64 across z runs",
        },
        "errors_vs_fp64": errors,
        "copy_calibration": copy_info,
        "cpu_baseline": cpu,
        "wall_s_timed_loop": round(wall, 4),
    }
    print(json.dumps(out))


def fp64_errors(x, lap, col, mx, xb, lapb, colb, mxb):
    """Max error of each output against FP64 evaluations (numpy on the host)."""
    def lap64(a):
        p = np.pad(a.astype(np.float64), 1)
        return (p[:-2, 1:-1, 1:-1] + p[2:, 1:-1, 1:-1] + p[1:-1, :-2, 1:-1] + p[1:-1, 2:, 1:-1]
                + p[1:-1, 1:-1, :-2] + p[1:-1, 1:-1, 2:] - 6.0 * p[1:-1, 1:-1, 1:-1])

    res = {}
    for tag, xs, l, c, m in (("fp32", x, lap, col, mx), ("bf16", xb, lapb, colb, mxb)):
        xn = xs.float().cpu().numpy()           # bf16 inputs exactly as stored
        L = lap64(xn)
        ln = l.float().cpu().numpy().astype(np.float64)
        absL = np.abs(L).max()
        C = ln.sum(axis=0)                      # FP64 sum of the stored Laplacian
        C_exact = L.sum(axis=0)
        absC = np.abs(ln).sum(axis=0)
        cn = c.cpu().numpy().astype(np.float64)
        res[tag] = {
            "laplacian_max_abs_err": float(np.abs(ln - L).max()),
            "laplacian_max_err_over_max_abs": float(np.abs(ln - L).max() / absL),
            "colsum_max_rel_err_vs_fp64_sum_of_stored": float((np.abs(cn - C) / np.maximum(absC, 1e-300)).max()),
            "colsum_max_err_over_sum_abs_vs_fp64_exact": float((np.abs(cn - C_exact) / np.maximum(absC, 1e-300)).max()),
            "max_equals_max_of_stored": bool(float(m.float().item()) == float(ln.max())),
            "max_abs_err_vs_fp64": float(abs(float(m.float().item()) - L.max())),
        }
        del L, ln, C, C_exact
    if "bf16" in res:
        x32 = x.cpu().numpy()
        res["bf16"]["laplacian_max_abs_err_vs_fp32_inputs"] = float(
            np.abs(lapb.float().cpu().numpy().astype(np.float64) - lap64(x32)).max())
    return res


def cpu_baseline(xn, bytes_per_step):
    """The oracle's C restatement of the host path (Laplacian -> axis-0 fold -> max),
    timed on this machine's cores: 1 thread and all threads (OpenMP over z planes)."""
    from oracle import oracle as O  # the checker/baseline, never the measured GPU path
    import multiprocessing
    ncores = min(int(os.environ.get("OMP_NUM_THREADS", multiprocessing.cpu_count())), multiprocessing.cpu_count())
    res = {}
    for threads in (1, ncores):
        reps = []
        for _ in range(2):
            t0 = time.perf_counter()
            O.laplacian_pipeline(xn, threads)
            reps.append(time.perf_counter() - t0)
        res[threads] = min(reps)
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    best_threads = ncores
    return {
        "value": round(bytes_per_step / res[best_threads] / 1e9, 3),
        "unit": "GB/s",
        "cores": best_threads,
        "kind": "port",
        "sample": "the full 512^3 float32 workload (Laplacian, axis-0 sum, max), best of 2 runs",
        "seconds_per_step": round(res[best_threads], 4),
        "single_thread_GBps": round(bytes_per_step / res[1] / 1e9, 3),
        "single_thread_seconds": round(res[1], 4),
        "cpu_model": cpu_model,
    }


if __name__ == "__main__":
    main()
