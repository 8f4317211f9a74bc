#!/usr/bin/env python3
"""Benchmark of the interleaved block-cyclic FFT + stencil path on one MI355X.

Workload (default): 2D N=8192, P=4 (16 virtual ranks), B=32.  One step is the
round trip of the reference's example() (zmodel/main.cu:50-80): 5-point
Laplacian (h = 1/N) -> forward 2D FFT -> inverse 2D FFT scaled 1/N^2, on a
seeded smooth field (256 random Fourier modes with |k| <= N/8, amplitudes
N(0,1)/(1+|k|^2), uniform phases, plus 1e-3 uniform noise) held in HBM.

Prints one JSON line on stdout (diagnostics go to stderr):
  value       FFT GFLOP/s, 5 N^2 log2(N^2) per transform, two transforms per step
  roofline    the step's dominant kernel: its launch durations come from HIP
              events recorded after each launch on its stream, in a second
              timed pass of the same steps
  cpu_baseline the FP64 CPU oracle (a C port of the reference path) on the
              host cores, one full step of the same field
  errors      relative L2 / max error of every FP32 output against the FP64
              GPU path on the same FP32-rounded input
  library_fft the FFT round trip alone beside torch.fft.fftn + ifftn (rocFFT)
              on the natural layout of the same field and precision
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def rel_errors(a, ref, torch):
    d = (a.to(torch.complex128) - ref)
    l2 = float(torch.linalg.vector_norm(d) / torch.linalg.vector_norm(ref))
    mx = float(d.abs().max() / ref.abs().max())
    return {"rel_l2": l2, "rel_max": mx}


def run_steps(plan, x, lap, y, z, scratch, n, steps, torch, stream):
    dims = plan.shape.dims
    for _ in range(steps):
        plan.laplace(x, scale=float(n * n), out=lap, stream=stream)
        plan.forward(lap, out=y, scratch=scratch, stream=stream)
        plan.inverse(y, 1.0 / float(n) ** dims, out=z, scratch=scratch, stream=stream)


def time_path(plan, x, n, steps, warmup, torch, label):
    """Time exactly `steps` steps with no per-launch instrumentation (the step
    time and `value`), then the same `steps` again with the plan's per-launch
    HIP events, one recorded on the launch stream after every kernel (the
    stage breakdown and the dominant kernel's launch durations)."""
    lap = torch.empty_like(x)
    y = torch.empty_like(x)
    z = torch.empty_like(x)
    scratch = torch.empty(plan.scratch_bytes // x.element_size(), dtype=x.dtype, device="cuda")
    stream = torch.cuda.current_stream()
    run_steps(plan, x, lap, y, z, scratch, n, warmup, torch, stream)
    torch.cuda.synchronize()

    def timed(instrumented):
        if instrumented:
            plan.timing_reset()
            plan.timing(True)
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        start.record(stream)
        run_steps(plan, x, lap, y, z, scratch, n, steps, torch, stream)
        stop.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        if instrumented:
            plan.timing(False)
        return start.elapsed_time(stop) / steps, wall

    ms, wall = timed(False)
    ms_instr, _ = timed(True)
    stages = plan.timing_read()
    plan.timing_reset()
    log(f"{label}: {ms:.3f} ms/step (wall {1e3 * wall / steps:.3f} ms/step; with per-launch events "
        f"{ms_instr:.3f}); stages " + ", ".join(f"{k}: {c} x {t / c:.3f} ms" for k, (c, t) in stages.items()))
    time_path.last_instrumented = ms_instr
    return ms, stages, (lap, y, z)


def time_graph(plan, x, n, steps, torch):
    """One step captured in a HIP graph (torch.cuda.graph on the plan's
    stream-ordered launches), replayed `steps` times."""
    lap, y, z = torch.empty_like(x), torch.empty_like(x), torch.empty_like(x)
    scratch = torch.empty(plan.scratch_bytes // x.element_size(), dtype=x.dtype, device="cuda")
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        run_steps(plan, x, lap, y, z, scratch, n, 2, torch, None)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        run_steps(plan, x, lap, y, z, scratch, n, 1, torch, None)
    for _ in range(3):
        g.replay()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(steps):
        g.replay()
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / steps
    log(f"graph replay: {ms:.4f} ms/step")
    return ms


def library_fft_comparison(plan, x, n, steps, warmup, torch):
    """The FFT round trip alone (forward, then inverse scaled 1/N^d) on the
    packed layout against torch.fft.fftn / ifftn (hipFFT over rocFFT) on the
    natural layout of the same field and precision: the role of the
    reference's heFFTe + cuFFT example (heffte/heffte_example_cuda.cpp).
    Forward errors of both are taken against the FP64 library transform."""
    dims = plan.shape.dims
    axes = tuple(range(dims))
    nat = plan.packed_to_natural(x).reshape((n,) * dims)
    y, z = torch.empty_like(x), torch.empty_like(x)
    scratch = torch.empty(plan.scratch_bytes // x.element_size(), dtype=x.dtype, device="cuda")
    stream = torch.cuda.current_stream()
    scale = 1.0 / float(n) ** dims

    def ours(k):
        for _ in range(k):
            plan.forward(x, out=y, scratch=scratch, stream=stream)
            plan.inverse(y, scale, out=z, scratch=scratch, stream=stream)

    def lib(k):
        for _ in range(k):
            f = torch.fft.fftn(nat, dim=axes)
            torch.fft.ifftn(f, dim=axes)

    out = {}
    for name, fn in (("ours", ours), ("library", lib)):
        fn(warmup)
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        fn(steps)
        ev1.record(stream)
        torch.cuda.synchronize()
        out[name] = ev0.elapsed_time(ev1) / steps
    ref = torch.fft.fftn(nat.to(torch.complex128), dim=axes).reshape(-1)
    tab = plan.reciprocal_index_table()
    ours_nat = torch.empty_like(y)
    ours_nat[tab] = y
    del tab
    err_ours = rel_errors(ours_nat, ref, torch)
    del ours_nat
    err_lib = rel_errors(torch.fft.fftn(nat, dim=axes).reshape(-1), ref, torch)
    res = {
        "library": f"torch.fft.fftn + ifftn (hipFFT/rocFFT), natural {n}^{dims} layout, {str(x.dtype)[6:]}",
        "ms_per_round_trip": round(out["library"], 4),
        "ours_ms_per_round_trip": round(out["ours"], 4),
        "speedup_vs_library": round(out["library"] / out["ours"], 3),
        "forward_errors_vs_fp64_library": {"ours": err_ours, "library": err_lib},
    }
    log(f"library FFT comparison: {json.dumps(res)}")
    return res


def cpu_baseline(dims, n, p, b, seed):
    """FP64 CPU oracle (C port of the reference path, radix-2 DFT mode, OpenMP
    over the host cores) on one full step of the same shape, and beside it a
    stand-in for the reference's heFFTe+FFTW CPU path (absent here): scipy.fft
    (pocketfft) forward + inverse on the natural layout with one worker per
    core, plus a numpy periodic Laplacian."""
    import numpy as np

    import oracle
    oracle.build()
    threads = len(os.sched_getaffinity(0))
    s = oracle.Shape(dims, n, p, b)
    rng = np.random.default_rng(seed)
    x = (rng.uniform(-1, 1, s.total) + 1j * rng.uniform(-1, 1, s.total)).astype(np.complex128)
    field = float(n) ** dims
    oracle.set_dft_mode(1)
    try:
        t0 = time.perf_counter()
        lap = s.laplace(x)
        t1 = time.perf_counter()
        y = s.forward_fft(lap)
        t2 = time.perf_counter()
        s.inverse_fft(y, scale=1.0 / field)
        t3 = time.perf_counter()
    finally:
        oracle.set_dft_mode(0)
    del lap, y
    flops = 2 * 5.0 * field * math.log2(field)
    log(f"cpu oracle ({threads} threads): laplace {t1 - t0:.2f} s, forward {t2 - t1:.2f} s, inverse {t3 - t2:.2f} s")
    standin = None
    try:
        import scipy.fft as sfft
        xn = x.reshape((n,) * dims)
        del x
        u0 = time.perf_counter()
        lapn = -2.0 * dims * xn
        for ax in range(dims):
            lapn += np.roll(xn, 1, ax) + np.roll(xn, -1, ax)
        u1 = time.perf_counter()
        yn = sfft.fftn(lapn, workers=threads)
        u2 = time.perf_counter()
        sfft.ifftn(yn, workers=threads)
        u3 = time.perf_counter()
        log(f"cpu scipy stand-in ({threads} workers): laplace {u1 - u0:.2f} s, fftn {u2 - u1:.2f} s, "
            f"ifftn {u3 - u2:.2f} s")
        standin = {"value": flops / (u3 - u0) / 1e9, "unit": "GFLOP/s", "cores": threads,
                   "sample": "numpy periodic Laplacian + scipy.fft fftn/ifftn (pocketfft) of the same complex128 "
                             f"field on the natural layout, {u3 - u0:.2f} s, forward alone {u2 - u1:.2f} s"}
    except ImportError:
        pass
    return {
        "value": flops / (t3 - t0) / 1e9,
        "unit": "GFLOP/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"one full step (Laplacian, forward FFT, inverse FFT) of a {dims}D N={n} P={p} B={b} complex128 "
                   f"U[-1,1) field in the FP64 C oracle (radix-2 DFT mode, OpenMP); {t3 - t0:.2f} s, forward "
                   f"alone {t2 - t1:.2f} s"),
        "seconds_per_step": t3 - t0,
        "forward_seconds": t2 - t1,
        "standin_heffte_fftw": standin,
    }


def load_traffic(name):
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64", "fp32_mfma"])
    ap.add_argument("--dims", type=int, default=2, choices=[2, 3])
    ap.add_argument("--n", type=int, default=None, help="default 8192 (2D) / 512 (3D)")
    ap.add_argument("--p", type=int, default=None, help="default 4 (2D) / 2 (3D)")
    ap.add_argument("--b", type=int, default=None, help="default 32 (2D) / 16 (3D)")
    ap.add_argument("--two-stage", action="store_true", help="force the two-kernel-per-direction path")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison path")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-library-fft", action="store_true",
                    help="skip the torch.fft (rocFFT) natural-layout round-trip comparison")
    ap.add_argument("--seed", type=int, default=8192)
    ap.add_argument("--graph", action="store_true",
                    help="also time the step captured once in a HIP graph and replayed (graph_ms_per_step)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="time a plain device-to-device copy of one N^2 complex64 field (the HBM reference "
                         "kernel the rocprof FETCH_SIZE / WRITE_SIZE counters are scaled by) and exit")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "stencilfft-amd runs the whole P x P virtual rank grid on one GPU; "
                                   "--gpus must be 1 (multi-GPU sharding is a design note, not built)"}))
        sys.exit(2)

    import torch

    from stencilfft_amd import Plan
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU"}))
        sys.exit(1)

    dims = args.dims
    dflt = {2: (8192, 4, 32), 3: (512, 2, 16)}[dims]
    n = args.n or dflt[0]
    p = args.p or dflt[1]
    b = args.b or dflt[2]
    if args.copy_calibration:
        src = torch.ones(n * n, dtype=torch.complex64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 10
        ev0.record()
        for _ in range(reps):
            dst.copy_(src)
        ev1.record()
        torch.cuda.synchronize()
        ms = ev0.elapsed_time(ev1) / reps
        nbytes = 2 * src.numel() * src.element_size()
        print(json.dumps({"metric": "device-to-device copy bandwidth", "value": round(nbytes / ms / 1e6, 1),
                          "unit": "GB/s", "bytes_per_copy_read": src.numel() * src.element_size(),
                          "bytes_per_copy_write": src.numel() * src.element_size(), "ms_per_copy": round(ms, 4),
                          "frac_of_peak": round(nbytes / ms / 1e6 / HBM_PEAK_GBS, 4)}))
        return
    log(f"workload {dims}D N={n} P={p} B={b} precision={args.precision} steps={args.steps} warmup={args.warmup}")
    plan64 = Plan(dims, n, p, b, precision="fp64", two_stage=args.two_stage)
    if dims == 2:
        from stencilfft_amd.fields import mode_field_natural
        nat = mode_field_natural(n, 2, 256, n // 8, args.seed, noise=1e-3)
        x64 = plan64.natural_to_packed(nat)
        del nat
    else:
        from stencilfft_amd.fields import uniform_fields
        x64 = uniform_fields(plan64, args.seed)
    if args.precision == "fp64":
        plan = plan64
        x = x64
    else:
        plan = Plan(dims, n, p, b, precision=args.precision, two_stage=args.two_stage)
        x = x64.to(torch.complex64)
    torch.cuda.synchronize()

    ms, stages, outs = time_path(plan, x, n, args.steps, args.warmup, torch, args.precision)
    esize = x.element_size()
    field = n ** dims
    flops = 2 * 5.0 * field * math.log2(field)
    gflops = flops / (ms * 1e-3) / 1e9
    # bytes per step: 2 sweeps per direction and FFT (pack, exchange, twiddle
    # and repack fused into adjacent loads and stores) and the textbook bound
    # (1 sweep per direction), plus one sweep for the stencil
    sweep = 2 * field * esize
    bytes_2sweep = 2 * (2 * dims) * sweep + sweep
    bytes_textbook = 2 * dims * sweep + sweep
    step_frac = bytes_2sweep / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS

    # dominant kernel and its roofline (textbook bytes: one read + one write of
    # the field per launch; every stage kernel here moves exactly that)
    dom = max(stages.items(), key=lambda kv: kv[1][1])
    dom_name, (dom_count, dom_ms) = dom
    dom_avg = dom_ms / dom_count
    per_launch = 2 * field * esize
    achieved = per_launch / (dom_avg * 1e-3) / 1e9
    per_launch_2sweep = {"line": 2 * per_launch}.get(dom_name, per_launch)
    traffic = load_traffic(f"traffic_{dom_name}_n{n}_p{p}_b{b}_{args.precision}.json")
    roofline = {
        "bound": "hbm",
        "kernel": {"line": ("line_fast_kernel (fused FFT1a + exchange + twiddle + FFT1b direction pass, "
                            "forward and inverse instantiations)" if plan.fast else
                            "line_kernel (fused FFT1a + exchange + twiddle + FFT1b direction pass)"),
                   "repack": "repack_tiled_kernel", "laplace": "laplace_strip_kernel",
                   "stage_a": "stage_a_kernel", "stage_b": "stage_b_kernel"}.get(dom_name, dom_name),
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic["bytes_per_launch"] if traffic else None,
        "bytes_per_launch": per_launch,
        "avg_launch_ms": round(dom_avg, 4),
        "launches": dom_count,
        "basis": "textbook bytes: one read and one write of the N^2 field per launch",
        "achieved_2sweep": round(per_launch_2sweep / (dom_avg * 1e-3) / 1e9, 1),
        "bytes_per_launch_2sweep": per_launch_2sweep,
    }

    if args.graph:
        result_extra_graph = {"graph_ms_per_step": round(time_graph(plan, x, n, args.steps, torch), 4)}
    else:
        result_extra_graph = {}
    result_extra = {**result_extra_graph, "ms_per_step_with_launch_events": round(time_path.last_instrumented, 4),
        "stage_ms": {k: round(t / c, 4) for k, (c, t) in stages.items()},
        "stage_launches_per_step": {k: c // args.steps for k, (c, t) in stages.items()},
        "step_bytes_2sweep": bytes_2sweep,
        "step_bytes_textbook": bytes_textbook,
        "step_hbm_frac_2sweep": round(step_frac, 4),
        "step_hbm_frac_textbook": round(bytes_textbook / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
    }

    if not args.no_fp64 and args.precision != "fp64":
        x64r = x.to(torch.complex128)
        ms64, st64, outs64 = time_path(plan64, x64r, n, max(3, min(args.steps, 10)), 2, torch, "fp64")
        errs = {}
        for name, a, r in zip(("laplace", "forward", "round_trip"), outs, outs64):
            errs[name] = rel_errors(a, r, torch)
        # IFFT(FFT(lap)) against the FP64 Laplacian it started from
        errs["round_trip_vs_fp64_laplace"] = rel_errors(outs[2], outs64[0], torch)
        del outs64
        # the 4th-order derivatives (delta = 1/N) on the same rounded input,
        # outside the timed region
        derivs = ("dx4", "dy4") if dims == 2 else ("dx4", "dy4", "dz4")
        if plan.shape.b >= 2:
            for name in derivs:
                d32 = getattr(plan, name)(x, 1.0 / n)
                d64 = getattr(plan64, name)(x64r, 1.0 / n)
                errs[name] = rel_errors(d32, d64, torch)
                del d32, d64
        result_extra["errors_vs_fp64_gpu"] = errs
        result_extra["fp64_ms_per_step"] = round(ms64, 4)
        result_extra["speedup_vs_fp64_gpu"] = round(ms64 / ms, 3)
        log(f"errors vs FP64 GPU path: {json.dumps(errs)}")
        # the FP64 path with the reference's stage structure: per direction
        # FFT1a + pack + exchange, then unpack + twiddle + FFT1b, materialising
        # the post-exchange buffer (zmodel/transforms/fft.cu:199-339)
        if not args.two_stage:
            plan64s = Plan(dims, n, p, b, precision="fp64", two_stage=True)
            ms64s, _, outs64s = time_path(plan64s, x64r, n, max(3, min(args.steps, 10)), 2, torch,
                                          "fp64 two-stage")
            result_extra["fp64_two_stage_ms_per_step"] = round(ms64s, 4)
            result_extra["speedup_vs_fp64_two_stage_gpu"] = round(ms64s / ms, 3)
            del outs64s, plan64s
        del x64r

    if not args.no_library_fft:
        del outs
        torch.cuda.empty_cache()
        result_extra["library_fft"] = library_fft_comparison(plan, x, n, max(3, min(args.steps, 10)), 2, torch)
        outs = None

    cpu = None
    if not args.no_cpu_baseline:
        del outs
        torch.cuda.empty_cache()
        cpu = cpu_baseline(dims, n, p, b, args.seed)

    line = {
        "metric": "FFT GFLOP/s (5N·log2N), % HBM BW peak; FP32 vs FP64 path at stated error",
        "value": round(gflops, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"{dims}D N={n} P={p} B={b} ({p ** dims} virtual ranks): {5 if dims == 2 else 7}-point "
                         f"Laplacian (h=1/N) -> forward FFT -> inverse FFT (1/N^{dims}), complex "
                         f"{'128' if args.precision == 'fp64' else '64'}"),
            "n": n, "p": p, "b": b, "precision": args.precision,
            "path": "two-stage" if args.two_stage else "fused",
            "field": (f"256 modes |k|<=N/8, amp N(0,1)/(1+|k|^2), + 1e-3 U[-1,1) noise, seed {args.seed}"
                      if dims == 2 else f"re, im U[-1,1) (torch CUDA generator), seed {args.seed}"),
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        **result_extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
