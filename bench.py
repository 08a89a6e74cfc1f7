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
import gc
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
# FP64 error evaluations of every workload (outside the timed regions); --no-errors clears it
WITH_ERRORS = True
SHAPE = (512, 512, 512)
# the fused step's kernel: float32 march stencil with the axis-0 sum and the max in its epilogue
MARCH_KERNEL = "nca_stencil_march_float32_axis0_max"
# The rocprofv3 PMC summary of that kernel (FETCH_SIZE and WRITE_SIZE per dispatch, calibrated
# against a copy of known size): its `traffic_bytes_per_launch` line is the kernel's HBM traffic,
# valid for the kernel source whose SHA-256 its `kernel_source_sha256` line records.
MARCH_PROFILE = os.path.join(ROOT, "profiles", "nca_stencil_march_laplacian_512cubed.md")


def march_traffic(st):
    """(traffic bytes per launch or None, note): the recorded PMC traffic of the headline kernel
    when the summary was measured on the kernel this build generates (same source SHA-256)."""
    import hashlib
    sha = hashlib.sha256(st.source(4).encode()).hexdigest()
    rec = {}
    try:
        with open(MARCH_PROFILE) as f:
            for line in f:
                if ":" in line and line.split(":", 1)[0].strip() in ("traffic_bytes_per_launch", "kernel_source_sha256"):
                    k, v = line.split(":", 1)
                    rec[k.strip()] = v.strip().strip("`")
    except OSError:
        return None, "no PMC summary found", sha
    if rec.get("kernel_source_sha256") != sha:
        return None, "the PMC summary was measured on another build of the kernel (source SHA-256 differs)", sha
    return int(rec["traffic_bytes_per_launch"]), os.path.relpath(MARCH_PROFILE, ROOT), sha


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-errors", action="store_true", help="skip the FP64 error evaluation")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 512 MiB device copy (for PMC byte-counter calibration)")
    ap.add_argument("--only-headline", action="store_true",
                    help="skip the other workloads (expression, binning, autograd, ragged batch)")
    ap.add_argument("--workloads", default="c0,c2,c3,c4,cp,st",
                    help="comma list of the other workloads to run: c0 (fused expression + sum), c2 (binning), "
                         "c3 (autograd pipeline), c4 (ragged batch), cp (casting, strided, permuted and "
                         "pointer-table copies, and the runtime's device copy), st (tile stencils: 2-D "
                         "Laplacian, pixel binning, pointer-axis planes)")
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
        # Python's cyclic garbage collector is off while timing (as timeit does): a full
        # collection in the middle of a host-bound loop (the eager autograd step) stalls the
        # host for tens of milliseconds and would be charged to one step
        gc_was_on = gc.isenabled()
        gc.disable()
        try:
            t0 = time.perf_counter()
            for a, b in evs:
                a.record(stream)
                fn()
                b.record(stream)
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
        finally:
            if gc_was_on:
                gc.enable()
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
    march_ms = float(np.mean(per_kernel[MARCH_KERNEL]))
    traffic, traffic_note, kernel_sha = march_traffic(st)
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
    # per API op, each input read once and each output written once: stencil, axis-0 sum, max
    bytes_unfused = 8 * n + (4 * n + 4 * SHAPE[1] * SHAPE[2]) + 4 * n
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
    global WITH_ERRORS
    WITH_ERRORS = not args.no_errors
    if not args.no_errors:
        fused_step()
        bf16_step()
        torch.cuda.synchronize()
        errors = fp64_errors(x, lap, col, mx, xb, lapb, colb, mxb)

    # ---- the other workloads: parity-tested elsewhere, measured here ----
    extra = {}
    if not args.only_headline:
        chosen = {w.strip() for w in args.workloads.split(",") if w.strip()}
        for name, fn in (("c0_fused_expression_sum", bench_c0), ("c2_binning", bench_c2),
                         ("c3_autograd_pipeline", bench_c3), ("c4_ragged_batch", bench_c4),
                         ("cp_layout_copies", bench_cp), ("st_tile_stencils", bench_st)):
            if name.split("_")[0] not in chosen:
                continue
            try:
                extra[name] = fn(torch, capi, tapes, ti, dev, sptr, time_loop)
            except Exception as e:  # report, keep the headline line
                extra[name] = {"error": f"{type(e).__name__}: {e}"}
            torch.cuda.empty_cache()

    # ---- CPU baseline (oracle restatement of the reference host path) ----
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(x.cpu().numpy(), bytes_fused, extra)

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
            "traffic": traffic,
            "traffic_over_algorithmic": round(traffic / kernel_bytes, 4) if traffic else None,
            "traffic_source": traffic_note,
            "kernel_source_sha256": kernel_sha,
            "kernel": MARCH_KERNEL + " (Laplacian + axis-0 sum + max epilogue)",
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
            "accumulate": "float32 compensated (Kahan) per thread, float64 across z runs",
        },
        "errors_vs_fp64": errors,
        "other_workloads": extra,
        "copy_calibration": copy_info,
        "cpu_baseline": cpu,
        "wall_s_timed_loop": round(wall, 4),
    }
    print(json.dumps(out))


def gbps(nbytes, ms):
    return round(nbytes / (ms * 1e-3) / 1e9, 2)


def graph_loop(torch, fn, inner, reps, between=None):
    """Capture `inner` calls of fn (which launch on torch's current stream) into one hipGraph
    and time `reps` replays with events; returns the mean ms per call.  fn is warmed up on a
    side stream first (kernels compiled, workspaces sized), so nothing allocates in capture."""
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(3):
            fn()
            if between:
                between()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    # capture on the warm-up stream: the library's per-stream workspace is already sized
    with torch.cuda.graph(graph, stream=side):
        for _ in range(inner):
            if between:
                between()
            fn()
    torch.cuda.synchronize()
    graph.replay()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    cur = torch.cuda.current_stream()
    for a, b in evs:
        a.record(cur)
        graph.replay()
        b.record(cur)
    torch.cuda.synchronize()
    return float(np.mean([a.elapsed_time(b) for a, b in evs])) / inner, graph


def cold_device_ms(torch, capi, fn, reps=10):
    """Device time of one call of fn with the Infinity Cache flushed first: a 512 MiB scratch
    buffer is read (untimed) before each call, so the cache holds only clean scratch lines (a
    write would leave dirty lines whose write-back the timed call would pay for), and the
    call's kernels are timed by the library's profiler (HIP events around each launch).
    Returns the median per call in ms."""
    scratch = torch.ones(128 << 20, dtype=torch.float32, device="cuda")
    fn()
    scratch.sum()
    torch.cuda.synchronize()
    per = []
    for _ in range(reps):
        scratch.sum()
        torch.cuda.synchronize()
        capi.profile_enable(True)
        fn()
        torch.cuda.synchronize()
        recs = capi.profile_read()
        capi.profile_enable(False)
        per.append(sum(ms for _, ms in recs))
    del scratch
    return float(np.median(per))


def f32_ulp(torch, ref64):
    """The float32 ulp at each value of a float64 tensor (spacing above |fl32(ref)|)."""
    r = ref64.float().abs()
    return (torch.nextafter(r, torch.full_like(r, float("inf"))) - r).double()


def sum_errors(torch, S, S64, abs_sum):
    """|S - S64| / |S64| and |S - S64| / sum|x| for a float32 (or float64) sum S."""
    d = abs(float(S) - float(S64))
    return {"rel_err_vs_fp64": d / abs(float(S64)) if S64 != 0 else d, "err_over_sum_abs": d / float(abs_sum)}


def bench_c0(torch, capi, tapes, ti, dev, sptr, time_loop):
    """out = a*b + exp(-a*a) - 0.5*b on 2^22 float32, then sum(out): fused in one kernel
    (out materialised), and the expression alone.  The 50 MB working set fits the 256 MB
    Infinity Cache, so these figures are cache-resident."""
    n = 1 << 22
    g = torch.Generator(device=dev).manual_seed(0)
    a = torch.rand(n, generator=g, device=dev) * 2 - 1
    b = torch.rand(n, generator=g, device=dev) * 2 - 1
    out = torch.empty_like(a)
    s = torch.empty((), dtype=torch.float32, device=dev)
    tape = capi.Tape(tapes.fused_c0(), [capi.FLOAT32] * 5, [capi.FLOAT32], capi.FLOAT32, capi.FLOAT32)
    ops = {"a": ti.view(a), "b": ti.view(b)}
    ins = [ops[k] for k in tapes.C0_OPERANDS]
    vo = ti.view(out)
    fused = lambda: tape.launch_reduce(ins, vo, capi.FUSE_SUM, capi.ACCUM_COMPENSATED, s.data_ptr(), None, None,  # noqa: E731
                                       scalars=(0.5,), stream=ti.stream_ptr())
    only = lambda: tape.launch(ins, vo, scalars=(0.5,), stream=ti.stream_ptr())  # noqa: E731
    fm, _ = time_loop(fused, 50, 5)
    om, _ = time_loop(only, 50, 5)
    torch.cuda.synchronize()
    eager_sum = float(s.item())
    gfm, _ = graph_loop(torch, fused, 20, 10)
    gom, _ = graph_loop(torch, only, 20, 10)
    # the same expression through the Python array API (the reference's call path: NCDevArray
    # operators, eager = one kernel and one allocation per operator, lazy = one fused kernel)
    import ncarray_amd as nca
    A, B = nca.from_torch(a), nca.from_torch(b)
    was_eager = nca.is_eager()

    def py_eager():
        nca.set_eager(True)
        return A * B + nca.exp(-(A * A)) - B * 0.5

    def py_lazy():
        with nca.lazy_mode():
            e = A * B + nca.exp(-(A * A)) - B * 0.5
        return nca.materialize(e)

    def py_lazy_sum():
        with nca.lazy_mode():
            e = A * B + nca.exp(-(A * A)) - B * 0.5
        return e.sum()  # a host scalar: waits for the result

    errors = None
    if WITH_ERRORS:
        # errors against an FP64 evaluation of the same inputs (the fused call's out and sum)
        fused()
        torch.cuda.synchronize()
        a64, b64 = a.double(), b.double()
        ref64 = a64 * b64 + torch.exp(-a64 * a64) - 0.5 * b64
        diff = (out.double() - ref64).abs()
        # a*b + exp(-a*a) - 0.5*b cancels to near zero for some inputs, where ulp and relative error
        # say nothing about the evaluation: errors are given relative to the terms' magnitudes (the
        # bound the tests state is 2 eps of that), and in ulp where the result keeps at least a
        # quarter of that magnitude
        scale = (a64 * b64).abs() + torch.exp(-a64 * a64) + 0.5 * b64.abs()
        keep = ref64.abs() >= 0.25 * scale
        eps = float(np.finfo(np.float32).eps)
        errors = {"expr_max_err_in_eps_of_term_magnitude": round(float((diff / scale).max()) / eps, 4),
                  "expr_max_ulp_without_cancellation": round(float((diff[keep] / f32_ulp(torch, ref64[keep])).max()), 3),
                  "expr_max_rel_err_without_cancellation": float((diff[keep] / ref64[keep].abs()).max()),
                  "sum_vs_fp64_of_exact_values": sum_errors(torch, s.item(), ref64.sum().item(), ref64.abs().sum().item()),
                  "sum_vs_fp64_of_stored_values": sum_errors(torch, s.item(), out.double().sum().item(),
                                                             out.double().abs().sum().item())}
        del a64, b64, ref64, diff, scale, keep
    cold_f = cold_device_ms(torch, capi, fused)
    cold_o = cold_device_ms(torch, capi, only)
    pe, _ = time_loop(py_eager, 30, 5)
    pl, _ = time_loop(py_lazy, 30, 5)
    ps, _ = time_loop(py_lazy_sum, 30, 5)
    nca.set_eager(was_eager)
    return {"n": n, "fused_ms": round(float(np.mean(fm)), 4), "fused_GBps": gbps(12 * n, np.mean(fm)),
            "hbm_cold_fused_ms": round(cold_f, 4), "hbm_cold_fused_GBps": gbps(12 * n, cold_f),
            "hbm_cold_expr_only_ms": round(cold_o, 4), "hbm_cold_expr_only_GBps": gbps(12 * n, cold_o),
            "python_eager_ms": round(float(np.median(pe)), 4), "python_lazy_ms": round(float(np.median(pl)), 4),
            "python_lazy_sum_ms": round(float(np.median(ps)), 4),
            "expr_only_ms": round(float(np.mean(om)), 4), "expr_only_GBps": gbps(12 * n, np.mean(om)),
            "graph_fused_ms": round(gfm, 4), "graph_fused_GBps": gbps(12 * n, gfm),
            "graph_expr_only_ms": round(gom, 4), "graph_expr_only_GBps": gbps(12 * n, gom),
            "graph_sum_equals_eager": bool(float(s.item()) == eager_sum),
            "errors_vs_fp64": errors,
            "note": "cache-resident (50 MB < 256 MB MALL); bytes = read a, b + write out; graph_* = "
                    "20 steps per hipGraph replay (no host launch overhead); hbm_cold_* = device time of one "
                    "call after reading 512 MiB flushed the Infinity Cache; python_* = the expression "
                    "written with NCDevArray operators (eager: one kernel and one output allocation per "
                    "operator; lazy: materialize() runs one fused kernel; lazy_sum: the sum fused into "
                    "the evaluation, returned to the host), median per call"}


def bench_c2(torch, capi, tapes, ti, dev, sptr, time_loop):
    """256 frames of 1024x1024 float32 (exponential, 1% zeros), 1000 radial bins."""
    import ctypes
    F, ny, nx, nb = 256, 1024, 1024, 1000
    g = torch.Generator(device=dev).manual_seed(3)
    fr = -50.0 * torch.log1p(-torch.rand((F, ny, nx), generator=g, device=dev))
    fr = torch.where(torch.rand((F, ny, nx), generator=g, device=dev) < 0.01, torch.zeros_like(fr), fr).float()
    yy, xx = torch.meshgrid(torch.arange(ny, device=dev, dtype=torch.float64),
                            torch.arange(nx, device=dev, dtype=torch.float64), indexing="ij")
    r = torch.hypot(yy - 512, xx - 512)
    bm = torch.clamp(torch.floor(r * nb / float(np.hypot(512, 512))), max=nb - 1).to(torch.int32)
    counts = torch.empty((F, nb), dtype=torch.int32, device=dev)
    sums = torch.empty((F, nb), dtype=torch.float32, device=dev)
    vf, vm, vc, vs = ti.view(fr), ti.view(bm), ti.view(counts), ti.view(sums)
    fwd = lambda: capi.call("nca_hip_bin", ctypes.byref(vf), ctypes.byref(vm), nb, capi.COUNT_NONZERO,  # noqa: E731
                            ctypes.byref(vc), ctypes.byref(vs), sptr)
    gs = torch.randn((F, nb), device=dev)
    gf = torch.empty_like(fr)
    vg, vgf = ti.view(gs), ti.view(gf)
    bwd = lambda: capi.call("nca_hip_bin_backward", ctypes.byref(vg), ctypes.byref(vm), ctypes.byref(vgf), sptr)  # noqa: E731
    fm, _ = time_loop(fwd, 10, 2)
    bm_, _ = time_loop(bwd, 10, 2)
    # per-kernel device time (HIP events the library records around each launch)
    capi.profile_enable(True)
    for _ in range(10):
        fwd()
        bwd()
    torch.cuda.synchronize()
    recs = capi.profile_read()
    capi.profile_enable(False)
    per = {}
    for name, kms in recs:
        per.setdefault(name, []).append(kms)
    n = F * ny * nx
    c2_errors = None
    if WITH_ERRORS:
        # counts and sums against a float64 scatter of the same frames by the same map
        fwd()
        torch.cuda.synchronize()
        idx = bm.view(-1).long()
        valid = (idx >= 0) & (idx < nb)
        s64 = torch.zeros((F, nb), dtype=torch.float64, device=dev)
        c64 = torch.zeros((F, nb), dtype=torch.int64, device=dev)
        for f0 in range(0, F, 32):
            blk = fr[f0:f0 + 32].reshape(fr[f0:f0 + 32].shape[0], -1)[:, valid]
            s64[f0:f0 + 32].index_add_(1, idx[valid], blk.double())
            c64[f0:f0 + 32].index_add_(1, idx[valid], (blk != 0).long())
        err = (sums.double() - s64).abs()
        c2_errors = {"counts_bit_exact_vs_fp64_scatter": bool(torch.equal(counts.long(), c64)),
                     "sums_max_rel_err_vs_fp64": float((err / s64.abs().clamp_min(1e-300)).max()),
                     "sums_max_ulp": round(float((err / f32_ulp(torch, s64)).max()), 3)}
        del s64, c64, err
    fwd_bytes = 4 * n + 4 * ny * nx + 2 * 4 * F * nb
    bwd_bytes = 4 * n + 4 * ny * nx + 4 * F * nb
    return {"fwd_ms": round(float(np.mean(fm)), 4), "fwd_GBps": gbps(fwd_bytes, np.mean(fm)),
            "bwd_ms": round(float(np.mean(bm_)), 4), "bwd_GBps": gbps(bwd_bytes, np.mean(bm_)),
            "kernel_ms": {k: round(float(np.mean(v)), 4) for k, v in per.items()},
            # the histogram kernel reads every pixel and the map once (4 B each)
            "histogram_kernel_frac_of_8TBps": round((4 * n + 4 * ny * nx) / (np.mean(per["nca_bin"]) * 1e-3) / 1e9
                                                    / HBM_PEAK_GBPS, 4),
            "fwd_bytes": fwd_bytes, "bwd_bytes": bwd_bytes,
            "counts_total_equals_nonzero_pixels": bool(int(counts.sum()) == int((fr != 0).sum())),
            "errors_vs_fp64": c2_errors}


def bench_c3(torch, capi, tapes, ti, dev, sptr, time_loop):
    """loss = sum(a*b + exp(-a*a) - 0.5*b), a = Laplacian(x), b = y on 256^3 float32 and
    bf16: fused forward (one kernel), and forward + backward (hand-written adjoints)."""
    from ncarray_amd import autograd as A
    res = {}
    n3 = 256 ** 3
    for tag, dt in (("fp32", torch.float32), ("bf16", torch.bfloat16)):
        g = torch.Generator(device=dev).manual_seed(5)
        x = torch.randn((256, 256, 256), generator=g, device=dev).to(dt).requires_grad_()
        y = torch.randn((256, 256, 256), generator=g, device=dev).to(dt).requires_grad_()
        item = x.element_size()

        def fwd():
            with torch.no_grad():
                A.laplacian_c0_loss(x, y)

        def fwdbwd():
            x.grad = None
            y.grad = None
            A.laplacian_c0_loss(x, y).backward()

        fm, _ = time_loop(fwd, 20, 3)
        bm, _ = time_loop(fwdbwd, 10, 2)
        torch.cuda.synchronize()
        # device time of one forward+backward step: the library's kernels, each bracketed by
        # HIP events on its launch stream (host launch overhead excluded)
        capi.profile_enable(True)
        for _ in range(5):
            fwdbwd()
        torch.cuda.synchronize()
        recs = capi.profile_read()
        capi.profile_enable(False)
        kms = float(sum(ms for _, ms in recs)) / 5
        # the same for one step of laplacian_c0_value_and_grad (loss and gradients in one march
        # pass, then the transposed stencil)
        xd, yd = x.detach(), y.detach()
        A.laplacian_c0_value_and_grad(xd, yd)
        torch.cuda.synchronize()
        capi.profile_enable(True)
        for _ in range(5):
            vg = A.laplacian_c0_value_and_grad(xd, yd)
        torch.cuda.synchronize()
        vrecs = capi.profile_read()
        capi.profile_enable(False)
        vms = float(sum(ms for _, ms in vrecs)) / 5
        cold_fwd = cold_device_ms(torch, capi, fwd)
        cold_vg = cold_device_ms(torch, capi, lambda: A.laplacian_c0_value_and_grad(xd, yd))
        with torch.no_grad():
            loss_eager = A.laplacian_c0_loss(x, y).clone()
        errs = c3_errors(torch, xd, yd, loss_eager, vg) if WITH_ERRORS else None
        gfm, _ = graph_loop(torch, fwd, 10, 10)  # forward only (launched from this thread)
        # forward + backward without torch's autograd (the same kernels, launched from this
        # thread), 10 steps per hipGraph replay
        gfbm, _ = graph_loop(torch, lambda: A.laplacian_c0_value_and_grad(x, y), 10, 10)
        fwd_bytes = 2 * item * n3                  # read x, y
        fb_bytes = fwd_bytes + 4 * item * n3       # + read x, y, write dx, dy
        frac = lambda nb, ms: round(nb / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)  # noqa: E731
        res[tag] = {"fwd_ms": round(float(np.mean(fm)), 4), "fwd_GBps": gbps(fwd_bytes, np.mean(fm)),
                    "fwd_bwd_ms": round(float(np.mean(bm)), 4), "fwd_bwd_GBps": gbps(fb_bytes, np.mean(bm)),
                    "fwd_frac_of_8TBps": frac(fwd_bytes, np.mean(fm)),
                    "fwd_bwd_frac_of_8TBps": frac(fb_bytes, np.mean(bm)),
                    "graph_fwd_ms": round(gfm, 4), "graph_fwd_GBps": gbps(fwd_bytes, gfm),
                    "graph_fwd_frac_of_8TBps": frac(fwd_bytes, gfm),
                    "graph_fwd_bwd_ms": round(gfbm, 4), "graph_fwd_bwd_GBps": gbps(fb_bytes, gfbm),
                    "graph_fwd_bwd_frac_of_8TBps": frac(fb_bytes, gfbm),
                    "device_fwd_bwd_ms": round(kms, 4), "device_fwd_bwd_GBps": gbps(fb_bytes, kms),
                    "device_fwd_bwd_frac_of_8TBps": frac(fb_bytes, kms),
                    "device_fwd_bwd_kernels": sorted({n for n, _ in recs}),
                    "device_value_and_grad_ms": round(vms, 4), "device_value_and_grad_GBps": gbps(fb_bytes, vms),
                    "device_value_and_grad_frac_of_8TBps": frac(fb_bytes, vms),
                    "device_value_and_grad_kernels": sorted({n for n, _ in vrecs}),
                    # the one pass's own traffic: read x, y, write dx, dy
                    "value_and_grad_pass_bytes": 4 * item * n3,
                    "device_value_and_grad_pass_frac_of_8TBps": frac(4 * item * n3, vms),
                    "graph_value_and_grad_pass_frac_of_8TBps": frac(4 * item * n3, gfbm),
                    "hbm_cold_fwd_ms": round(cold_fwd, 4), "hbm_cold_fwd_frac_of_8TBps": frac(fwd_bytes, cold_fwd),
                    "hbm_cold_value_and_grad_ms": round(cold_vg, 4),
                    "hbm_cold_value_and_grad_frac_of_8TBps": frac(fb_bytes, cold_vg),
                    "loss": float(loss_eager), "loss_value_and_grad": float(vg[0]),
                    "errors_vs_fp64": errs}
    res["note"] = ("bytes are algorithmic (each input read once, each output written once): forward = read x, "
                   "y; fwd_bwd = the forward's bytes + the backward's (read x, y, write dx, dy), also for "
                   "laplacian_c0_value_and_grad, whose one pass computes the loss and both gradients (its own "
                   "traffic, value_and_grad_pass_bytes, is 2/3 of that: the *_pass_frac_* fields); the backward "
                   "is one training-step march (the Laplacian recomputed, dx = Laplacian^T of dloss/da from LDS, "
                   "dy); fwd_bwd_ms is host-bound (torch autograd and Python per step); "
                   "graph_fwd replays a hipGraph of 10 forwards (no Python launch overhead), graph_fwd_bwd one of 10 "
                   "laplacian_c0_value_and_grad steps; device_fwd_bwd sums the kernel "
                   "durations of a torch-autograd step (forward kernel, then the backward), device_value_and_grad "
                   "those of a laplacian_c0_value_and_grad step (HIP events per launch); 134 MB working set: "
                   "cache-resident, except hbm_cold_* (device time after reading 512 MiB flushed the "
                   "Infinity Cache)")
    return res


def c3_errors(torch, x, y, loss, vg):
    """The C3 loss and gradients against FP64 on the same (stored, e.g. bf16-rounded) inputs:
    a = Laplacian(x) with zero boundary, loss = sum(a*y + exp(-a*a) - 0.5*y), dy = a - 0.5,
    dx = Laplacian(y - 2 a exp(-a*a)) (the zero-boundary Laplacian is its own adjoint)."""
    import torch.nn.functional as F

    def lap(u):
        p = F.pad(u, (1, 1, 1, 1, 1, 1))
        return (p[:-2, 1:-1, 1:-1] + p[2:, 1:-1, 1:-1] + p[1:-1, :-2, 1:-1] + p[1:-1, 2:, 1:-1]
                + p[1:-1, 1:-1, :-2] + p[1:-1, 1:-1, 2:] - 6.0 * u)

    x64, y64 = x.double(), y.double()
    a = lap(x64)
    e = torch.exp(-a * a)
    terms = a * y64 + e - 0.5 * y64
    loss64 = terms.sum().item()
    dy64 = a - 0.5
    dx64 = lap(y64 - 2.0 * a * e)
    _, dx, dy = vg
    out = {"loss": sum_errors(torch, float(loss), loss64, terms.abs().sum().item()),
           "loss_value_and_grad": sum_errors(torch, float(vg[0]), loss64, terms.abs().sum().item()),
           "dx_max_err_over_max_abs": float((dx.double() - dx64).abs().max() / dx64.abs().max()),
           "dy_max_err_over_max_abs": float((dy.double() - dy64).abs().max() / dy64.abs().max())}
    return out


def bench_cp(torch, capi, tapes, ti, dev, sptr, time_loop):
    """Copies of a 512^3 float32 array (512 MiB) into dense arrays: a casting copy to BF16, a plain
    copy, a strided slice, two permuted copies and a gather of its 512 planes through an NC table of
    plane pointers; and the two halves of a copy on their own, through the library's kernels: a
    constant written into the 512 MiB (write only) and the full sum of a 1 GiB array, four times
    the Infinity Cache (read only).  Kernel time from the HIP events the library records around
    each launch; algorithmic bytes = elements x (source + destination itemsize)."""
    n = 512
    x = torch.randn((n, n, n), device=dev)
    one = lambda sd, dd: capi.Tape([capi.pack(capi.OP_LOAD_NCARR, 0)], [sd], [], capi.FLOAT32, dd)  # noqa: E731
    f32, to_bf16 = one(capi.FLOAT32, capi.FLOAT32), one(capi.FLOAT32, capi.BFLOAT16)
    dense = torch.empty_like(x)
    half = torch.empty((n, n, n // 2), device=dev)
    ob = torch.empty((n, n, n), dtype=torch.bfloat16, device=dev)
    tbl = torch.tensor([x[i].data_ptr() for i in range(n)], dtype=torch.int64, device=dev)
    vnc = capi.make_nc_view(tbl.data_ptr(), capi.FLOAT32, (n, n, n), [1, n * 4, 4], 0)
    fill = capi.Tape([capi.pack(capi.OP_LOAD_CONST, 0)], [], [capi.FLOAT32], capi.FLOAT32, capi.FLOAT32)
    S = capi.STREAM_DEFAULT_ASYNC
    cases = [
        ("cast_f32_to_bf16", lambda: to_bf16.launch([ti.view(x)], ti.view(ob), stream=S), n ** 3 * 6),
        ("copy_f32", lambda: f32.launch([ti.view(x)], ti.view(dense), stream=S), n ** 3 * 8),
        ("slice_last_axis_step2", lambda: f32.launch([ti.view(x[:, :, ::2])], ti.view(half), stream=S), n ** 3 * 4),
        ("transpose_1_2", lambda: f32.launch([ti.view(x.transpose(1, 2))], ti.view(dense), stream=S), n ** 3 * 8),
        ("transpose_0_2", lambda: f32.launch([ti.view(x.transpose(0, 2))], ti.view(dense), stream=S), n ** 3 * 8),
        ("nc_planes_to_dense", lambda: f32.launch([vnc], ti.view(dense), stream=S), n ** 3 * 8),
        ("fill_f32_write_only", lambda: fill.launch([], ti.view(dense), scalars=(1.5,), stream=S), n ** 3 * 4),
    ]
    res = {}
    for name, fn, nbytes in cases:
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        capi.profile_enable(True)
        for _ in range(10):
            fn()
        torch.cuda.synchronize()
        recs = capi.profile_read()
        capi.profile_enable(False)
        ms = float(np.median([k for _, k in recs]))
        res[name] = {"kernel": recs[0][0], "kernel_ms": round(ms, 4), "GBps": gbps(nbytes, ms),
                     "frac_of_8TBps": round(nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)}
    # read only: the full sum of a 1 GiB float32 array (the reduction kernel and its combine, timed
    # together per call)
    big = torch.randn(1 << 28, device=dev)
    vb = ti.view(big)
    tot = torch.empty((), dtype=torch.float32, device=dev)
    red = lambda: capi.call("nca_hip_reduce_full", capi.ctypes.byref(vb), capi.REDUCE_SUM, capi.ACCUM_PLAIN,  # noqa: E731
                            tot.data_ptr(), S)
    for _ in range(3):
        red()
    torch.cuda.synchronize()
    per = []
    for _ in range(10):
        capi.profile_enable(True)
        red()
        torch.cuda.synchronize()
        recs = capi.profile_read()
        capi.profile_enable(False)
        per.append(recs)
    ms = float(np.median([sum(k for _, k in r) for r in per]))
    nb = big.numel() * 4
    res["sum_f32_1GiB_read_only"] = {"kernel": " + ".join(nm for nm, _ in per[0]), "kernel_ms": round(ms, 4),
                                     "GBps": gbps(nb, ms), "frac_of_8TBps": round(nb / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)}
    del big
    # the runtime's device-to-device copy of the same 512 MiB (hipMemcpyAsync), timed with events on
    # torch's stream: the copy ceiling the library's plain copy is compared with
    for _ in range(3):
        dense.copy_(x)
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(10)]
    for a, b in evs:
        a.record()
        dense.copy_(x)
        b.record()
    torch.cuda.synchronize()
    ms = float(np.median([a.elapsed_time(b) for a, b in evs]))
    res["runtime_d2d_copy_f32"] = {"kernel": "hipMemcpyAsync device to device (runtime)", "kernel_ms": round(ms, 4),
                                   "GBps": gbps(n ** 3 * 8, ms),
                                   "frac_of_8TBps": round(n ** 3 * 8 / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)}
    return res


def bench_st(torch, capi, tapes, ti, dev, sptr, time_loop):
    """Stencils off the dense 3-D path: a 5-point 2-D Laplacian on a 16384^2 float32 grid (zero
    boundary) and 2x2 pixel binning (step 2) of 256 frames of 1024^2 float32 into 512^2 (the tile
    kernel), and the 3-D Laplacian of a 512^3 float32 grid whose 512 planes are reached through a
    table of plane pointers (an NC pointer axis 0: the register-queue march reading each plane's
    pointer once).  Kernel time from the library's HIP events; algorithmic bytes = source read
    once + destination written once."""
    P = capi.pack
    five = [P(capi.OP_LOAD_NCARR, 1), P(capi.OP_LOAD_NCARR, 2), P(capi.OP_ADD), P(capi.OP_LOAD_NCARR, 3),
            P(capi.OP_ADD), P(capi.OP_LOAD_NCARR, 4), P(capi.OP_ADD), P(capi.OP_LOAD_CONST, 0),
            P(capi.OP_LOAD_NCARR, 0), P(capi.OP_MUL), P(capi.OP_SUB)]
    lap2 = capi.Stencil([[0, 0], [-1, 0], [1, 0], [0, -1], [0, 1]], five, capi.FLOAT32, capi.FLOAT32, capi.FLOAT32,
                        [capi.FLOAT32])
    win = [[0, a, b] for a in (0, 1) for b in (0, 1)]
    sum4 = [P(capi.OP_LOAD_NCARR, 0), P(capi.OP_LOAD_NCARR, 1), P(capi.OP_ADD), P(capi.OP_LOAD_NCARR, 2),
            P(capi.OP_ADD), P(capi.OP_LOAD_NCARR, 3), P(capi.OP_ADD)]
    binst = capi.Stencil(win, sum4, capi.FLOAT32, capi.FLOAT32, capi.FLOAT32)
    lap3 = capi.Stencil(tapes.LAPLACIAN_OFFSETS, tapes.laplacian(), capi.FLOAT32, capi.FLOAT32, capi.FLOAT32,
                        [capi.FLOAT32], is_pointer_axis=[1, 0, 0])
    g = torch.Generator(device=dev).manual_seed(0)
    S = capi.STREAM_DEFAULT_ASYNC
    res = {}

    def measure(name, fn, nbytes):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        capi.profile_enable(True)
        for _ in range(10):
            fn()
        torch.cuda.synchronize()
        recs = capi.profile_read()   # one stencil kernel per call (no fused reductions here)
        capi.profile_enable(False)
        ms = float(np.median([k for _, k in recs]))
        res[name] = {"kernel": recs[0][0], "kernel_ms": round(ms, 4), "GBps": gbps(nbytes, ms),
                     "frac_of_8TBps": round(nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)}

    x = torch.randn((16384, 16384), generator=g, device=dev)
    o = torch.empty_like(x)
    measure("laplacian_2d_16384sq_f32", lambda: lap2.apply(ti.view(x), ti.view(o), capi.BOUNDARY_ZERO, scalars=(4.0,),
                                                         stream=S), 8 * x.numel())
    del x, o
    fr = torch.rand((256, 1024, 1024), generator=g, device=dev)
    ob = torch.empty((256, 512, 512), device=dev)
    measure("pixel_binning_2x2_256x1024sq_f32", lambda: binst.apply(ti.view(fr), ti.view(ob), capi.BOUNDARY_NONE,
                                                                    stream=S), 4 * fr.numel() + 4 * ob.numel())
    del fr, ob
    x3 = torch.randn((512, 512, 512), generator=g, device=dev)
    o3 = torch.empty_like(x3)
    tbl = torch.tensor([x3[i].data_ptr() for i in range(512)], dtype=torch.int64, device=dev)
    v3 = capi.make_nc_view(tbl.data_ptr(), capi.FLOAT32, (512, 512, 512), [1, 2048, 4], 0)
    measure("laplacian_3d_512cubed_pointer_planes_f32", lambda: lap3.apply(v3, ti.view(o3), capi.BOUNDARY_ZERO,
                                                                           scalars=(6.0,), stream=S), 8 * x3.numel())
    return res


def bench_c4(torch, capi, tapes, ti, dev, sptr, time_loop):
    """4096 float32 arrays, sizes log-uniform in [1e3, 4e6] (seed 0): the C0 expression
    plus per-array sum and max, one persistent launch (and its hipGraph) vs a loop of
    per-array launches."""
    from ncarray_amd import batch
    sizes = batch.ragged_sizes(4096, 1e3, 4e6, seed=0)
    total = int(sizes.sum())
    g = torch.Generator(device=dev).manual_seed(0)
    a_all = torch.rand(total, generator=g, device=dev) * 2 - 1
    b_all = torch.rand(total, generator=g, device=dev) * 2 - 1
    o_all = torch.empty(total, device=dev)
    # items packed back to back, each starting on a 16-byte boundary
    pad, off = [], 0
    for n_ in sizes:
        pad.append(off)
        off = -(-(off + int(n_)) // 4) * 4
    span = off
    if span > total:
        a_all = torch.cat([a_all, torch.zeros(span - total, device=dev)])
        b_all = torch.cat([b_all, torch.zeros(span - total, device=dev)])
        o_all = torch.empty(span, device=dev)
    a = [a_all[p:p + int(n)] for p, n in zip(pad, sizes)]
    b = [b_all[p:p + int(n)] for p, n in zip(pad, sizes)]
    o = [o_all[p:p + int(n)] for p, n in zip(pad, sizes)]
    tape = capi.Tape(tapes.fused_c0(), [capi.FLOAT32] * 5, [capi.FLOAT32], capi.FLOAT32, capi.FLOAT32)
    rb = batch.RaggedBatch(tape, [a, b, a, a, b], o, fuse=capi.FUSE_SUM | capi.FUSE_MAX)
    cfg_default = rb.config()
    dm, _ = time_loop(lambda: rb.launch(scalars=(0.5,), stream=sptr), 5, 2)
    tuned_ms = rb.autotune(scalars=(0.5,), reps=2)
    cfg = rb.config()
    tm, _ = time_loop(lambda: rb.launch(scalars=(0.5,), stream=sptr), 5, 2)
    rb.capture(scalars=(0.5,))
    gm, _ = time_loop(lambda: rb.replay(stream=sptr), 5, 2)
    sums = torch.empty(len(sizes), device=dev)
    maxs = torch.empty(len(sizes), device=dev)
    views = [([ti.view(a[i]), ti.view(b[i]), ti.view(a[i]), ti.view(a[i]), ti.view(b[i])], ti.view(o[i]))
             for i in range(len(sizes))]
    sp, mp = sums.data_ptr(), maxs.data_ptr()

    def loop():
        for i, (ins, vo) in enumerate(views):
            tape.launch_reduce(ins, vo, capi.FUSE_SUM | capi.FUSE_MAX, capi.ACCUM_PLAIN, sp + 4 * i, mp + 4 * i, None,
                               scalars=(0.5,), stream=sptr)

    lm, _ = time_loop(loop, 2, 1)
    torch.cuda.synchronize()
    nbytes = 12 * total
    same_max = bool(torch.equal(maxs, rb.maxs))
    rel = float(((sums.double() - rb.sums.double()).abs() / (rb.sums.double().abs() + 1.0)).max())
    c4_errors = None
    if WITH_ERRORS:
        # the batch's sums against float64 sums of its stored outputs, and its max against the stored
        # outputs' max (exact), per array
        rb.launch(scalars=(0.5,), stream=sptr)
        torch.cuda.synchronize()
        worst_rel = worst_abs = 0.0
        max_exact = True
        for i in range(len(sizes)):
            oi = o[i].double()
            s64, sa = oi.sum().item(), oi.abs().sum().item()
            d = abs(float(rb.sums[i]) - s64)
            worst_rel = max(worst_rel, d / abs(s64) if s64 != 0 else d)
            worst_abs = max(worst_abs, d / sa if sa else d)
            max_exact = max_exact and float(rb.maxs[i]) == float(o[i].max())
        c4_errors = {"sum_max_rel_err_vs_fp64": worst_rel, "sum_max_err_over_sum_abs": worst_abs,
                     "max_equals_max_of_stored": bool(max_exact)}
    rb.close()
    return {"n_arrays": len(sizes), "total_elements": total, "total_GB_moved": round(nbytes / 1e9, 3),
            "batched_default_ms": round(float(np.mean(dm)), 3), "default_config": cfg_default,
            "batched_tuned_ms": round(float(np.mean(tm)), 3), "tuned_config": cfg,
            "batched_graph_ms": round(float(np.mean(gm)), 3), "batched_graph_GBps": gbps(nbytes, np.mean(gm)),
            "per_array_loop_ms": round(float(np.mean(lm)), 3), "per_array_loop_GBps": gbps(nbytes, np.mean(lm)),
            "loop_vs_batch_max_equal": same_max, "loop_vs_batch_sum_max_rel_diff": rel,
            "errors_vs_fp64": c4_errors,
            "note": ("sizes follow the stated distribution: ~1.97e9 elements (~23.6 GB moved), not 2.5 GB; "
                     "the per-array loop calls the C ABI from Python once per array (kernel + combine launch)")}


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


def cpu_info():
    """CPU model, core counts and the threads this process may use (lscpu and /proc/cpuinfo)."""
    import subprocess
    info = {"logical_cpus": os.cpu_count()}
    try:
        info["usable_cpus"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        info["usable_cpus"] = os.cpu_count()
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k = k.strip()
            if k in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU family", "Model",
                     "Stepping", "CPU max MHz"):
                info["lscpu_" + k.lower().replace(" ", "_").replace("(s)", "s")] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") and "cpu_model" not in info:
                    info["cpu_model"] = line.split(":", 1)[1].strip()
                elif line.startswith("flags") and "vector_isa" not in info:
                    flags = set(line.split(":", 1)[1].split())
                    info["vector_isa"] = [f for f in ("avx2", "fma", "avx512f", "avx512bw", "avx512_bf16") if f in flags]
                if "cpu_model" in info and "vector_isa" in info:
                    break
    except OSError:
        pass
    if info.get("lscpu_cpu_family", "").isdigit():
        info["cpu_family_hex"] = hex(int(info["lscpu_cpu_family"]))  # 0x1a: AMD Zen 5
    info["oracle_build"] = "gcc -O3 -march=x86-64-v3 (AVX2, FMA) -ffp-contract=off, OpenMP"
    return info


def cpu_baseline(xn, bytes_per_step, gpu_extra=None):
    """The oracle's C restatement of the reference host path, timed on this machine's cores at
    1 thread and at all usable threads (OpenMP, as the reference VM path splits its outer loop,
    hostengine.hh:97-118): the headline C1 pipeline on the full 512^3 grid, and bounded samples
    of the other workloads (C0 in full, 16 of C2's 256 frames, C3 at 256^3 forward and
    forward+backward, C4's first 64 arrays).  GB/s use the same algorithmic bytes as the GPU
    figures; best of 2 runs each."""
    from oracle import oracle as O  # the checker/baseline, never the measured GPU path
    from ncarray_amd import batch, tapes
    info = cpu_info()
    ncores = max(1, min(int(os.environ.get("OMP_NUM_THREADS", info["usable_cpus"])), info["usable_cpus"]))
    thread_counts = sorted({1, ncores})

    def best(fn, reps=2):
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return min(ts)

    def rates(fn, nbytes, sample):
        out = {"sample": sample, "algorithmic_bytes": int(nbytes)}
        for th in thread_counts:
            sec = best(lambda: fn(th))
            out[f"threads_{th}"] = {"seconds": round(sec, 4), "GBps": round(nbytes / sec / 1e9, 3)}
        return out

    res = {th: best(lambda: O.laplacian_pipeline(xn, th)) for th in thread_counts}
    per = {}
    rng = np.random.default_rng(0)
    # C0: the expression through the oracle's tape VM, then the float32 fold (12 B/element)
    n0 = 1 << 22
    a = rng.uniform(-1, 1, n0).astype(np.float32)
    b = rng.uniform(-1, 1, n0).astype(np.float32)
    per["c0_fused_expression_sum"] = rates(
        lambda th: O.expr_sum_mt(tapes.fused_c0(), [a, b, a, a, b], [0.5], 10, th), 12 * n0,
        "the full 2^22-element workload")
    # C2: per-frame bin counts and sums, 16 frames of 1024^2 with the radial 1000-bin map
    F2, nb = 16, 1000
    fr = rng.exponential(50.0, (F2, 1024, 1024)).astype(np.float32)
    fr[rng.random(fr.shape) < 0.01] = 0
    bm = O.radial_binmap()
    per["c2_binning"] = rates(lambda th: O.bin_f32_mt(fr, bm, nb, True, th),
                              4 * fr.size + 4 * bm.size + 2 * F2 * nb * 4, "16 of the 256 frames of 1024^2")
    del fr
    # C3: Laplacian + C0 loss at 256^3, forward and forward+backward (8 and 16 B/element)
    x3 = rng.standard_normal((256, 256, 256)).astype(np.float32)
    y3 = rng.standard_normal((256, 256, 256)).astype(np.float32)
    per["c3_forward"] = rates(lambda th: O.laplacian_c0(x3, y3, False, th), 8 * x3.size, "the full 256^3 float32 grid")
    per["c3_forward_backward"] = rates(lambda th: O.laplacian_c0(x3, y3, True, th), 16 * x3.size,
                                       "the full 256^3 float32 grid")
    del x3, y3
    # C4: the per-array loop over the first 64 arrays of the seeded ragged sizes (12 B/element)
    sizes = batch.ragged_sizes(4096, 1e3, 4e6, seed=0)[:64]
    items = [(rng.uniform(-1, 1, int(k)).astype(np.float32), rng.uniform(-1, 1, int(k)).astype(np.float32))
             for k in sizes]

    def c4(th):
        for ia, ib in items:
            O.expr_sum_mt(tapes.fused_c0(), [ia, ib, ia, ia, ib], [0.5], 10, th)

    per["c4_ragged_batch"] = rates(c4, 12 * int(sizes.sum()), f"the first 64 of the 4096 arrays ({int(sizes.sum())} elements)")
    del items
    return {
        "value": round(bytes_per_step / res[ncores] / 1e9, 3),
        "unit": "GB/s",
        "cores": ncores,
        "kind": "port",
        "sample": "the full 512^3 float32 workload (Laplacian, axis-0 sum, max), best of 2 runs",
        "seconds_per_step": round(res[ncores], 4),
        "single_thread_GBps": round(bytes_per_step / res[1] / 1e9, 3),
        "single_thread_seconds": round(res[1], 4),
        "cpu": info,
        "per_config": per,
    }


if __name__ == "__main__":
    main()
